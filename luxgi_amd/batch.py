"""Batched reflection denoise over frames of uneven sizes, graph-replayed probe-volume updates
(libluxgi.so).

ReflectionDenoiseBatch runs raytraced_reflection::denoise::system (Engine/Raytrace/
RaytracedReflection.cpp:496-520) for many frames at once through
luxgi_svgf_reflection_denoise_batched: the reprojection of every frame in one persistent launch,
then each a-trous iteration of every frame in one more. The frames' descriptors
(LuxgiReflectionDenoiseItem) live in device memory and the kernels read them at run time, so the
1 + iterations launches are captured once (capture()) and replayed every frame after the
descriptors are refreshed by one host-to-device copy from pinned memory. Each frame's outputs are
identical to luxgi_svgf_reflection_denoise on that frame.

DDGIBatch updates several probe volumes per frame by replaying a captured graph of their trace and
blend launches; the traces read the per-frame ray orientations from device memory.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .ddgi import stream_handle


def default_grid_blocks():
    """Four persistent blocks per compute unit."""
    return 4 * torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count


class ReflectionDenoiseBatch:
    """sizes: list of (width, height), one per frame. Per-frame state (accumulation targets,
    moments, a-trous pair, tile map) is allocated here; the G-buffers and noisy inputs are the
    caller's, set per frame with set_frame(). Schedule: order 2 (default) launches one block per
    16x16 unit (the grid follows the unit total); orders 0 / 1 run a persistent grid of
    grid_blocks blocks (default four per compute unit)."""

    def __init__(self, sizes, storage=abi.STORAGE_FP32, iterations=4, alpha=0.01, moments_alpha=0.2, radius=1,
                 phi_color=10.0, phi_normal=32.0, sigma_depth=1.0, grid_blocks=None, order=2):
        n = len(sizes)
        if not 0 < n <= abi.MAX_DENOISE_BATCH:
            raise ValueError(f"1 .. {abi.MAX_DENOISE_BATCH} frames per batch")
        self.n = n
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.storage = storage
        self.iterations = int(iterations)
        self.grid_blocks = int(grid_blocks or default_grid_blocks())
        self.order = int(order)
        dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
        self.items = (abi.ReflectionDenoiseItem * n)()
        self.header = abi.DenoiseBatchHeader()
        self.outputs, self.moments, self.atrous, self.tile_maps = [], [], [], []
        self._inputs = [None] * n
        for k, (w, h) in enumerate(self.sizes):
            z = lambda: torch.zeros((h, w, 4), dtype=dt, device="cuda")  # noqa: E731
            self.outputs.append([z(), z()])
            self.moments.append([z(), z()])
            self.atrous.append([torch.ones((h, w, 4), dtype=dt, device="cuda") for _ in range(2)])
            self.tile_maps.append(torch.zeros(max(lib().luxgi_svgf_tile_count(w, h), 1), dtype=torch.uint8,
                                              device="cuda"))
            it = self.items[k]
            it.width, it.height = w, h
            it.reproject.alpha = alpha
            it.reproject.momentsAlpha = moments_alpha
            it.reproject.approximateWithDDGI = 1
            it.filter = abi.ReflectionAtrousParams(radius, 0, phi_color, phi_normal, sigma_depth, 1)
            for j in range(2):
                it.outputs[j] = self.outputs[k][j].data_ptr()
                it.moments[j] = self.moments[k][j].data_ptr()
                it.atrous[j] = self.atrous[k][j].data_ptr()
            it.tileMap = self.tile_maps[k].data_ptr()
        nbytes = C.sizeof(self.items) + C.sizeof(self.header)
        self._hdr_off = C.sizeof(self.items)
        self.device = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        self._staging = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self._staged = [None, None]  # event recorded after each staging buffer's copy
        self._next = 0
        self.ping_pong = 0
        self.graph = None
        self.result_index = -1 if self.iterations == 0 else (1 if self.iterations & 1 else 0)

    def set_frame(self, k, cur, prev, noisy, view_proj_inv, prev_view_proj, camera_pos, camera_delta):
        """Frame k's inputs: G-buffer dicts of device tensors (color, normals, pbr, depth), the noisy
        signal, and the camera of ReflectionDenoiser.set_camera."""
        w, h = self.sizes[k]
        if noisy.shape[:2] != (h, w) or cur["depth"].shape != (h, w) or prev["depth"].shape != (h, w):
            raise ValueError(f"frame {k}: inputs must be {h}x{w}")
        it = self.items[k]
        it.reproject.viewProjInv = abi.mat4_colmajor(view_proj_inv)
        it.reproject.prevViewProj = abi.mat4_colmajor(prev_view_proj)
        it.reproject.cameraPos[:] = [float(camera_pos[0]), float(camera_pos[1]), float(camera_pos[2]), 1.0]
        it.reproject.cameraDelta[:] = [float(v) for v in camera_delta]
        it.cur = abi.GBuffer(cur["color"].data_ptr(), cur["normals"].data_ptr(), cur["pbr"].data_ptr(),
                             cur["depth"].data_ptr())
        it.prev = abi.GBuffer(prev["color"].data_ptr(), prev["normals"].data_ptr(), prev["pbr"].data_ptr(),
                              prev["depth"].data_ptr())
        it.noisy = noisy.data_ptr()
        self._inputs[k] = (cur, prev, noisy)  # keep the caller's tensors alive while they are referenced

    def upload(self):
        """Checks the descriptors, sets blockStart and the header (luxgi_svgf_batch_prepare) and copies
        them to device memory on the current stream, through a pinned double buffer."""
        if any(x is None for x in self._inputs):
            raise RuntimeError("set_frame() every frame before the first upload")
        for k in range(self.n):
            self.items[k].pingPong = self.ping_pong
        check(lib().luxgi_svgf_batch_prepare(C.byref(self.items), self.n, C.byref(self.header)),
              "luxgi_svgf_batch_prepare")
        j = self._next
        self._next = 1 - j
        if self._staged[j] is not None:
            self._staged[j].synchronize()  # the copy that last read this staging buffer is done
        host = self._staging[j].numpy()
        host[:self._hdr_off] = np.frombuffer(memoryview(self.items).cast("B"), dtype=np.uint8)
        host[self._hdr_off:] = np.frombuffer(memoryview(self.header).cast("B"), dtype=np.uint8)
        self.device.copy_(self._staging[j], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._staged[j] = ev

    def _grid(self):
        return max(int(self.header.totalBlocks), 1) if self.order == 2 else self.grid_blocks

    def launch(self):
        """The 1 + iterations launches over the uploaded descriptors."""
        res = C.c_int32(-2)
        base = self.device.data_ptr()
        check(lib().luxgi_svgf_reflection_denoise_batched(C.c_void_p(base + self._hdr_off), C.c_void_p(base),
                                                          self.iterations, self.storage, self._grid(), self.order,
                                                          C.byref(res), stream_handle()),
              "luxgi_svgf_reflection_denoise_batched")
        self.result_index = res.value

    def capture(self):
        """Captures launch() into a CUDA/HIP graph (after one eager warm-up run, whose effect on the
        temporal state is undone); denoise() then replays it."""
        self.upload()
        saved = self.state()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self.launch()  # warm-up outside the capture, as torch asks
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.restore(saved)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.launch()
        self.graph = g
        self._graph_grid = self._grid()

    def denoise(self):
        """One frame of every item (after set_frame): upload, launch (or graph replay), flip the
        moments ping-pong. Returns the output tensor of each frame."""
        self.upload()
        if self.graph is not None and self.order == 2 and self._grid() > self._graph_grid:
            self.capture()  # the frames grew past the captured one-unit-per-block grid
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()
        self.ping_pong = 1 - self.ping_pong
        return self.results()

    def results(self):
        r = self.result_index
        return [self.outputs[k][0] if r < 0 else self.atrous[k][r] for k in range(self.n)]

    def state(self):
        """The temporal state (for a caller that must run extra launches without disturbing it)."""
        return [t.clone() for k in range(self.n) for t in (*self.outputs[k], *self.moments[k])]

    def restore(self, saved):
        i = 0
        for k in range(self.n):
            for t in (*self.outputs[k], *self.moments[k]):
                t.copy_(saved[i])
                i += 1


def _field(st, name):
    """A ctypes struct field as a comparable value (arrays as tuples)."""
    v = getattr(st, name)
    return tuple(v) if isinstance(v, C.Array) else v


class DDGIBatch:
    """Several probe volumes of any probe and ray counts over one scene, updated each frame (trace,
    blend with the fused border, flip; DDGIRenderer.cpp:217-413 per volume) by one graph replay.

    Per frame only the volumes' ray orientations change (the ping-pong parity alternates and the
    first frame is special). So the first frame runs eagerly, capture() records one graph per
    parity, and update() copies the new orientations into a device array that the traces read
    at run time (deviceOrientation) before replaying the parity's graph.

    trace="batched" (default) traces all volumes in one launch sequence over their flat ray list
    (luxgi_ddgi_trace_sdf_batched); trace="per_volume" calls luxgi_ddgi_trace_sdf per volume.
    blend="batched" blends all volumes in one launch per map (luxgi_ddgi_probe_update_batched);
    blend="per_volume" calls luxgi_ddgi_probe_update per volume; "auto" (default) takes the batched
    form when every volume fits it. With `streams` > 0 the per-volume work is spread over that many
    side streams (default priority), forked from and joined to the current stream."""

    def __init__(self, uniforms, scene, trace_params=None, fuse_border=True, streams=4, trace="batched",
                 scene_of=None, blend="auto"):
        """scene: one DeviceScene for all volumes, or a list of them with scene_of[i] the scene of
        volume i (volumes from several scenes; trace="batched" then traces them all in one launch
        sequence with luxgi_ddgi_trace_sdf_scenes)."""
        from .ddgi import DDGIVolume, default_trace_params
        if trace not in ("batched", "per_volume"):
            raise ValueError("trace is 'batched' or 'per_volume'")
        if blend not in ("auto", "batched", "per_volume"):
            raise ValueError("blend is 'auto', 'batched' or 'per_volume'")
        self.volumes = [DDGIVolume(u) for u in uniforms]
        self.scene_list = list(scene) if isinstance(scene, (list, tuple)) else None
        if self.scene_list is not None:
            if scene_of is None or len(scene_of) != len(self.volumes):
                raise ValueError("scene_of names the scene of every volume")
            self.scene_of = [int(i) for i in scene_of]
        self.scene = scene if self.scene_list is None else None
        self.fuse_border = fuse_border
        self.trace_mode = trace
        # the volumes are independent: volume i runs on side stream i % streams (default priority),
        # forked from and joined back to the current stream, so small volumes overlap on the GPU
        # (and become parallel branches of the captured graph)
        self.streams = [torch.cuda.Stream() for _ in range(max(int(streams), 0))]
        k = len(self.volumes)
        self.orientation = torch.zeros((k, 16), dtype=torch.float32, device="cuda")
        self._staging = [torch.zeros((k, 16), dtype=torch.float32, pin_memory=True) for _ in range(2)]
        self._staged = [None, None]
        self._next = 0
        self.params = []
        for i in range(k):
            p = trace_params[i] if trace_params is not None else default_trace_params(np.eye(3))
            p.deviceOrientation = self.orientation[i].data_ptr()
            self.params.append(p)
        self.graphs = [None, None]
        self.descriptors = None
        if trace == "batched":
            if k > abi.MAX_TRACE_BATCH:
                raise ValueError(f"at most {abi.MAX_TRACE_BATCH} volumes per batched trace")
            # one launch sequence takes its march settings from one LuxgiTraceParams: the
            # volumes may differ in orientation only
            shared = [f for f, _ in abi.TraceParams._fields_ if f not in ("randomOrientation", "deviceOrientation")]
            for p in self.params[1:]:
                if any(_field(p, f) != _field(self.params[0], f) for f in shared):
                    raise ValueError("batched trace: the volumes' trace_params may differ in orientation only")
            d = (abi.TraceVolume * k)()
            for i, vol in enumerate(self.volumes):
                d[i].uniform = vol.u
                d[i].deviceOrientation = self.orientation[i].data_ptr()
                d[i].rays = vol.rays
                d[i].scene = self.scene_of[i] if self.scene_list is not None else 0
            if self.scene_list is None:
                total = C.c_int64(0)
                check(lib().luxgi_ddgi_trace_batch_prepare(C.byref(d), k, self.volumes[0].ray_storage,
                                                           C.byref(total)), "luxgi_ddgi_trace_batch_prepare")
                self.total_rays = total.value
                ws = lib().luxgi_ddgi_trace_batch_workspace_size(self.total_rays, scene.n_objects)
            else:
                n = len(self.scene_list)
                sc = (abi.TraceScene * n)()
                for j, ds in enumerate(self.scene_list):
                    sc[j].sdfData, sc[j].volume, sc[j].atlas, sc[j].skybox = ds.sdf_data, ds.volume, ds.atlas, ds.skybox
                    sc[j].skyColor[:] = list(self.params[0].skyColor)
                host = (C.c_uint8 * lib().luxgi_ddgi_trace_scenes_descriptor_bytes(n))()
                self.scene_batch = abi.TraceSceneBatch()
                check(lib().luxgi_ddgi_trace_scenes_prepare(C.byref(sc), n, C.byref(d), k, self.volumes[0].ray_storage,
                                                            C.byref(host), C.byref(self.scene_batch)),
                      "luxgi_ddgi_trace_scenes_prepare")
                self.total_rays = self.scene_batch.totalRays
                self.scene_descriptors = torch.from_numpy(np.frombuffer(host, dtype=np.uint8).copy()).cuda()
                ws = lib().luxgi_ddgi_trace_scenes_workspace_size(C.byref(self.scene_batch))
            # the descriptors do not change from frame to frame (the orientations are read through
            # deviceOrientation), so they are copied to the device once
            self.descriptors = torch.from_numpy(
                np.frombuffer(memoryview(d).cast("B"), dtype=np.uint8).copy()).cuda()
            self.trace_workspace = torch.empty(max(ws, 256), dtype=torch.uint8, device="cuda")
            self.shared_params = self.params[0]
        fits = all(v.u.irradianceProbeSideLength ** 2 <= 64 and v.u.depthProbeSideLength ** 2 <= 256
                   and v.u.maxDistance > 0 for v in self.volumes) and k <= abi.MAX_UPDATE_BATCH
        if blend == "batched" and not fits:
            raise ValueError("batched blend: every volume needs irradiance oct^2 <= 64, depth oct^2 <= 256, "
                             "maxDistance > 0")
        self.blend_mode = "batched" if blend == "batched" or (blend == "auto" and fits) else "per_volume"
        if self.blend_mode == "batched":
            self._prepare_blend()

    def _prepare_blend(self):
        """Descriptors of the batched blend for the first frame and for each ping-pong parity (the
        atlases read and written swap with the parity), on the device; one shared workspace."""
        k = len(self.volumes)
        flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | (abi.UPDATE_FUSE_BORDER if self.fuse_border else 0)
        vols = (abi.UpdateVolume * k)()
        for i, v in enumerate(self.volumes):
            vols[i].uniform = v.u
            vols[i].rays = v.rays
        ws = lib().luxgi_ddgi_update_batch_workspace_size(C.byref(vols), k)
        self.blend_workspace = torch.empty(max(ws, 256), dtype=torch.uint8, device="cuda")
        nbytes = lib().luxgi_ddgi_update_batch_descriptor_bytes(k)
        self.blend_sets = {}
        for key in ("first", 0, 1):
            parity = 0 if key == "first" else key
            for i, v in enumerate(self.volumes):
                vols[i].prev = v.atlas[parity]
                vols[i].out = v.atlas[1 - parity]
                vols[i].firstFrame = 1 if key == "first" else 0
            host = (C.c_uint8 * nbytes)()
            b = abi.UpdateBatch()
            check(lib().luxgi_ddgi_update_batch_prepare(C.byref(vols), k, flags,
                                                        C.c_void_p(self.blend_workspace.data_ptr()),
                                                        self.blend_workspace.numel(), C.byref(host), C.byref(b)),
                  "luxgi_ddgi_update_batch_prepare")
            self.blend_sets[key] = (b, torch.from_numpy(np.frombuffer(host, dtype=np.uint8).copy()).cuda())

    def _blend_all(self):
        b, d = self.blend_sets["first" if self.frames == 0 else self.volumes[0].ping_pong]
        check(lib().luxgi_ddgi_probe_update_batched(C.byref(b), C.c_void_p(d.data_ptr()), stream_handle()),
              "luxgi_ddgi_probe_update_batched")
        if not self.fuse_border:
            for vol in self.volumes:
                vol.border_update()

    @property
    def frames(self):
        return self.volumes[0].frames

    def _upload(self, rotations):
        """rotations: one 3x3 (or 4x4) orientation per volume -> the device array, column-major."""
        j = self._next
        self._next = 1 - j
        if self._staged[j] is not None:
            self._staged[j].synchronize()
        host = self._staging[j].numpy()
        for i, r in enumerate(rotations):
            m = np.eye(4)
            r = np.asarray(r, dtype=np.float64)
            m[:r.shape[0], :r.shape[1]] = r
            host[i] = m.T.reshape(-1)  # column-major
        self.orientation.copy_(self._staging[j], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        self._staged[j] = ev

    def _trace_all(self):
        if self.scene_list is not None:
            check(lib().luxgi_ddgi_trace_sdf_scenes(
                C.byref(self.scene_batch), C.c_void_p(self.descriptors.data_ptr()),
                C.c_void_p(self.scene_descriptors.data_ptr()), C.byref(self.shared_params),
                C.c_void_p(self.trace_workspace.data_ptr()), self.trace_workspace.numel(), stream_handle()),
                "luxgi_ddgi_trace_sdf_scenes")
            return
        sc = self.scene
        check(lib().luxgi_ddgi_trace_sdf_batched(
            C.c_void_p(self.descriptors.data_ptr()), len(self.volumes), self.total_rays, self.volumes[0].ray_storage,
            C.byref(sc.sdf_data), C.byref(sc.volume), C.byref(sc.atlas), C.byref(sc.skybox),
            C.byref(self.shared_params), C.c_void_p(self.trace_workspace.data_ptr()), self.trace_workspace.numel(),
            stream_handle()), "luxgi_ddgi_trace_sdf_batched")

    def _run_volume(self, vol, p, i=0):
        if self.trace_mode == "per_volume":
            vol.trace_rays(self.scene if self.scene_list is None else self.scene_list[self.scene_of[i]], p)
        if self.blend_mode == "per_volume":
            vol.probe_update(fuse_border=self.fuse_border)
            if not self.fuse_border:
                vol.border_update()

    def _run(self):
        if self.trace_mode == "batched":
            self._trace_all()
        if self.blend_mode == "batched":
            if self.trace_mode == "per_volume":
                self._run_streams()
            self._blend_all()
            return
        self._run_streams()

    def _run_streams(self):
        if not self.streams:
            for i, (vol, p) in enumerate(zip(self.volumes, self.params)):
                self._run_volume(vol, p, i)
            return
        main = torch.cuda.current_stream()
        for st in self.streams:
            st.wait_stream(main)
        for i, (vol, p) in enumerate(zip(self.volumes, self.params)):
            with torch.cuda.stream(self.streams[i % len(self.streams)]):
                self._run_volume(vol, p, i)
        for st in self.streams:
            main.wait_stream(st)

    def update(self, rotations):
        """One frame of every volume with the given orientations."""
        if len(rotations) != len(self.volumes):
            raise ValueError("one orientation per volume")
        self._upload(rotations)
        parity = self.volumes[0].ping_pong
        if self.frames > 0 and self.graphs[parity] is not None:
            self.graphs[parity].replay()
        else:
            self._run()
        for vol in self.volumes:
            vol.end_frame()

    def capture(self):
        """Records the frame for both ping-pong parities (after at least one eager frame, which also
        allocates the trace workspaces). The volumes' state is left as it was."""
        if self.frames == 0:
            raise RuntimeError("run one frame with update() before capture()")
        torch.cuda.synchronize()
        keep = [(v.ping_pong, v.frames) for v in self.volumes]
        for parity in (0, 1):
            for v in self.volumes:
                v.ping_pong = parity
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._run()
            self.graphs[parity] = g
        for v, (pp, fr) in zip(self.volumes, keep):
            v.ping_pong, v.frames = pp, fr

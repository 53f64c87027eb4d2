"""Host mirror of LuxGI's DDGI systems over the C ABI (libluxgi.so).

Mirrors Engine/DDGI/DDGIRenderer.cpp:
  DDGIPipelineInternal (:85-97)      -> DDGIVolume state: ray buffers, irradiance/depth ping-pong, frames
  init::initializeProbeGrid (:163-211) -> DDGIVolume.__init__ (atlas sizes :182-185)
  trace_rays::system (:217-330)      -> DDGIVolume.trace_rays
  probe_update::system (:377-413)    -> DDGIVolume.probe_update
  border_update::system (:441-464)   -> DDGIVolume.border_update (or fused into probe_update)
  end_frame::system (:335-342)       -> DDGIVolume.end_frame
  sample_probe::system (:467-503)    -> DDGIVolume.sample_probes
Device memory comes from torch (ROCm); every compute call goes to libluxgi.so on
the caller's current HIP stream.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib


def stream_handle():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_device(a, dtype=None):
    """numpy array / ctypes array -> contiguous cuda tensor (raw bytes for ctypes structs)."""
    if isinstance(a, C.Array):
        buf = np.frombuffer(memoryview(a).cast("B"), dtype=np.uint8).copy()
        return torch.from_numpy(buf).cuda()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


class DeviceScene:
    """SDF volume + surface atlas uploaded to the GPU (the inputs of the trace)."""

    def __init__(self, sdf_data, sdf, mip, atlas, sdf_storage=abi.STORAGE_FP32, sky_faces=None, sky_size=0,
                 layout=abi.LAYOUT_LINEAR, guard=False, atlas_layout=abi.LAYOUT_LINEAR, brick_scale=False,
                 refine=False):
        """sdf_storage FP16 with refine=True is the reduced-precision march with FP32 hits: the FP32
        volume (same layout) is kept beside the FP16 one and re-evaluated at every hit."""
        self.sdf_data = sdf_data
        self.res = int(sdf_data.resolution)
        self.mip_res = (self.res + 3) // 4
        sdt = {abi.STORAGE_FP16: torch.float16, abi.STORAGE_BF16: torch.bfloat16}.get(sdf_storage, torch.float32)
        # the FP32 volume the stored forms derive from
        src = sdf.cuda().float() if torch.is_tensor(sdf) else to_device(sdf, torch.float32)
        self.sdf = src if sdt == torch.float32 else src.to(sdt)
        self.mip = (mip.cuda() if torch.is_tensor(mip) else to_device(mip, torch.float32)).to(sdt)
        self.guard = None
        if guard:
            if sdf_storage != abi.STORAGE_FP32:
                raise ValueError("the brick guard bounds FP32 voxel values")
            C_ = int(sdf_data.cascadesCount)
            self.guard = torch.empty(((self.res // 4) ** 3) * C_, dtype=torch.float32, device="cuda")
            check(lib().luxgi_sdf_build_guard(ptr(src), self.res * C_, self.res, self.res, ptr(self.guard),
                                              stream_handle()), "luxgi_sdf_build_guard")
        C_ = int(sdf_data.cascadesCount)
        self.brick_scale = None
        self.refine = None

        def bricks(storage, dt, scale=None):
            nbytes = lib().luxgi_sdf_bricks_bytes(self.res * C_, self.res, self.res, storage)
            if nbytes == 0:
                raise ValueError("LAYOUT_BRICK4 needs dimensions that are multiples of 4")
            out = torch.empty(nbytes // (4 if storage == abi.STORAGE_FP32 else 2), dtype=dt, device="cuda")
            check(lib().luxgi_sdf_to_bricks(ptr(src), self.res * C_, self.res, self.res, storage, ptr(out), ptr(scale),
                                            stream_handle()), "luxgi_sdf_to_bricks")
            return out

        def cells(vol, W, H, D, storage, dt):
            nbytes = lib().luxgi_sdf_cells_bytes(W, H, D, storage)
            if nbytes == 0:
                raise ValueError("volume too large for LAYOUT_CELL8")
            out = torch.empty(nbytes // (4 if storage == abi.STORAGE_FP32 else 2), dtype=dt, device="cuda")
            check(lib().luxgi_sdf_to_cells(ptr(vol), W, H, D, storage, ptr(out), stream_handle()),
                  "luxgi_sdf_to_cells")
            return out

        if brick_scale and layout != abi.LAYOUT_BRICK4:
            raise ValueError("a per-brick scale needs LAYOUT_BRICK4")
        if refine:  # FP32 distances for the hits, in the same layout
            if layout == abi.LAYOUT_BRICK4:
                self.refine = bricks(abi.STORAGE_FP32, torch.float32)
            elif layout == abi.LAYOUT_CELL8:
                self.refine = cells(src, self.res * C_, self.res, self.res, abi.STORAGE_FP32, torch.float32)
            else:
                self.refine = src
        if layout == abi.LAYOUT_CELL8:
            # fine volume and coarse mip both as 8-corner cell records (one aligned record per lookup)
            mip32 = self.mip.float()
            self.sdf = cells(src, self.res * C_, self.res, self.res, sdf_storage, sdt)
            self.mip = cells(mip32, self.mip_res * C_, self.mip_res, self.mip_res, sdf_storage, sdt)
            del mip32
        if layout == abi.LAYOUT_BRICK4:
            # the engine would keep its SDF in this layout; here it is converted once
            if brick_scale:
                self.brick_scale = torch.empty((self.res // 4) ** 3 * C_, dtype=torch.float32, device="cuda")
            self.sdf = bricks(sdf_storage, sdt, self.brick_scale)
        self.volume = abi.SDFVolume(ptr(self.sdf), ptr(self.mip), self.res, self.mip_res, sdf_storage, layout,
                                    ptr(self.guard), ptr(self.brick_scale), ptr(self.refine),
                                    layout if refine else 0, 0)
        self.atlas_host = atlas
        self.objects = to_device(atlas.objects)
        self.tiles = to_device(atlas.tiles)
        self.chunks = to_device(atlas.chunks.view(np.int32))
        self.culled = to_device(atlas.culled.view(np.int32))
        if atlas.color_storage == abi.STORAGE_UNORM8:
            self.color = to_device(atlas.color)
        elif atlas.color_storage == abi.STORAGE_FP16:
            self.color = to_device(atlas.color.view(np.int16))
        else:
            self.color = to_device(atlas.color)
        self.depth = to_device(atlas.depth)
        if atlas_layout == abi.LAYOUT_BRICK4:
            # colour + depth interleaved in 4x4-texel blocks (the engine would write its atlas this way)
            res = int(atlas.data.resolution)
            packed = torch.empty(res * res * 2, dtype=torch.int32, device="cuda")
            check(lib().luxgi_atlas_pack(ptr(self.color), ptr(self.depth), res, ptr(packed), stream_handle()),
                  "luxgi_atlas_pack")
            self.color, self.depth = packed, None
        self.atlas = abi.SurfaceAtlas(atlas.data, ptr(self.chunks), ptr(self.culled), ptr(self.objects),
                                      ptr(self.tiles), ptr(self.color), ptr(self.depth), atlas.n_objects,
                                      atlas.color_storage, atlas_layout, 0)
        self.sky_faces = None if sky_faces is None else to_device(sky_faces)
        self.skybox = abi.Skybox(ptr(self.sky_faces), sky_size, 0)
        self.n_objects = atlas.n_objects


def default_trace_params(rotation, sky=(0.6, 0.7, 0.9), max_steps=250, hit_eps=0.0, frame=0):
    p = abi.TraceParams()
    p.randomOrientation = abi.mat4_colmajor(rotation)
    p.numFrames = frame
    p.infiniteBounces = 0
    p.numLights = 1
    p.intensity = 1.0
    p.maxSteps = max_steps
    p.hitEpsilon = hit_eps
    p.skyColor[:] = [sky[0], sky[1], sky[2], 1.0]
    return p


class DDGIVolume:
    """One probe volume with the reference's ping-pong frame state."""

    def __init__(self, uniform, ray_width=None, ray_storage=abi.STORAGE_FP32, atlas_storage=abi.STORAGE_FP32,
                 flags_buffer=False):
        self.u = uniform
        P = uniform.num_probes
        self.P = P
        self.ray_width = ray_width or uniform.raysPerProbe
        self.ray_storage = ray_storage
        self.atlas_storage = atlas_storage
        rdt = torch.float16 if ray_storage == abi.STORAGE_FP16 else torch.float32
        adt = torch.float16 if atlas_storage == abi.STORAGE_FP16 else torch.float32
        self.radiance = torch.zeros((P, self.ray_width, 4), dtype=rdt, device="cuda")
        self.dir_dist = torch.zeros((P, self.ray_width, 4), dtype=rdt, device="cuda")
        self.rays = abi.RayBuffers(ptr(self.radiance), ptr(self.dir_dist), self.ray_width, P, ray_storage, 0)
        Wi, Hi = uniform.irradianceTextureWidth, uniform.irradianceTextureHeight
        Wd, Hd = uniform.depthTextureWidth, uniform.depthTextureHeight
        self.irradiance = [torch.zeros((Hi, Wi, 4), dtype=adt, device="cuda") for _ in range(2)]
        self.depth = [torch.zeros((Hd, Wd, 2), dtype=adt, device="cuda") for _ in range(2)]
        self.atlas = [abi.ProbeAtlas(ptr(self.irradiance[i]), ptr(self.depth[i]), atlas_storage, 0) for i in range(2)]
        self.ray_flags = torch.zeros((P, self.ray_width), dtype=torch.uint8, device="cuda") if flags_buffer else None
        self.offsets = torch.zeros((P, 4), dtype=torch.float32, device="cuda")
        self.states = torch.zeros(P, dtype=torch.uint8, device="cuda")
        self.relocation = False
        ws = lib().luxgi_ddgi_probe_update_workspace_size(C.byref(uniform))
        self.workspace = torch.empty(max(ws, 256), dtype=torch.uint8, device="cuda")
        self.trace_workspace = None
        self.frames = 0
        self.ping_pong = 0

    def enable_relocation(self, on=True):
        """Relocation + classification in effect: the trace starts rays at the offset probe positions
        and the consumers (sample_probes, the surface-cache bounce and the reflection rays' DDGI
        branch, through the atlas descriptors) read the offsets and skip inactive probes."""
        self.relocation = bool(on)
        for a in self.atlas:
            a.probeOffsets = self.offsets.data_ptr() if on else None
            a.probeStates = self.states.data_ptr() if on else None

    # trace_rays::system (DDGIRenderer.cpp:217-330)
    def trace_rays(self, scene, params, use_offsets=None, counters=None):
        if use_offsets is None:
            use_offsets = self.relocation
        off = ptr(self.offsets) if use_offsets else None
        need = lib().luxgi_ddgi_trace_workspace_size(C.byref(self.u), scene.n_objects)
        if self.trace_workspace is None or self.trace_workspace.numel() < need:
            self.trace_workspace = torch.empty(need, dtype=torch.uint8, device="cuda")
        args = [C.byref(self.u), C.byref(scene.sdf_data), C.byref(scene.volume), C.byref(scene.atlas),
                C.byref(scene.skybox), C.byref(params), off, C.byref(self.rays), ptr(self.ray_flags),
                ptr(self.trace_workspace), self.trace_workspace.numel()]
        if counters is None:
            check(lib().luxgi_ddgi_trace_sdf(*args, stream_handle()), "luxgi_ddgi_trace_sdf")
        else:
            check(lib().luxgi_ddgi_trace_sdf_count(*args, ptr(counters), stream_handle()), "luxgi_ddgi_trace_sdf_count")

    # probe_update::system (DDGIRenderer.cpp:377-413), border fused by default
    def probe_update(self, fuse_border=True, per_probe=False, half=False, bf16=False):
        """half: the reduced-precision blend (one rounding per operand); bf16: its operands in BF16
        (needs half)."""
        flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH
        if fuse_border:
            flags |= abi.UPDATE_FUSE_BORDER
        if per_probe:
            flags |= abi.UPDATE_PER_PROBE
        if half:
            flags |= abi.UPDATE_HALF
        if bf16:
            flags |= abi.UPDATE_BF16
        write = 1 - self.ping_pong
        check(lib().luxgi_ddgi_probe_update(C.byref(self.u), C.byref(self.rays), C.byref(self.atlas[self.ping_pong]),
                                            C.byref(self.atlas[write]), 1 if self.frames == 0 else 0, flags,
                                            ptr(self.workspace), self.workspace.numel(), stream_handle()),
              "luxgi_ddgi_probe_update")

    # border_update::system (DDGIRenderer.cpp:441-464)
    def border_update(self):
        write = 1 - self.ping_pong
        check(lib().luxgi_ddgi_border_update(C.byref(self.u), C.byref(self.atlas[write]),
                                             abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH, stream_handle()),
              "luxgi_ddgi_border_update")

    def relocate_classify(self, params):
        if self.ray_flags is None:
            raise ValueError("DDGIVolume(flags_buffer=True) is required for relocation/classification")
        check(lib().luxgi_ddgi_relocate_classify(C.byref(self.u), C.byref(self.rays), ptr(self.ray_flags),
                                                 C.byref(params), ptr(self.offsets), ptr(self.states),
                                                 stream_handle()), "luxgi_ddgi_relocate_classify")

    # sample_probe::system (DDGIRenderer.cpp:467-503): per-pixel irradiance from the last update's atlases
    def sample_probes(self, depth, normals, camera_position, view_proj_inv, out=None, out_storage=abi.STORAGE_FP32):
        """depth [H][W] and normals [H][W][4] cuda float32 tensors; view_proj_inv a numpy 4x4 (row-major math)."""
        H, W = depth.shape
        if out is None:
            out = torch.empty((H, W, 4), dtype=torch.float16 if out_storage == abi.STORAGE_FP16 else torch.float32,
                              device="cuda")
        prm = abi.SampleProbeParams()
        prm.cameraPosition[:] = [float(camera_position[0]), float(camera_position[1]), float(camera_position[2]), 1.0]
        prm.viewProjInv = abi.mat4_colmajor(view_proj_inv)
        check(lib().luxgi_ddgi_sample_probes(C.byref(self.u), C.byref(self.atlas[self.ping_pong]), C.byref(prm), W, H,
                                             ptr(depth), ptr(normals), out_storage, ptr(out), stream_handle()),
              "luxgi_ddgi_sample_probes")
        return out

    # end_frame::system (DDGIRenderer.cpp:335-342)
    def end_frame(self):
        self.ping_pong = 1 - self.ping_pong
        self.frames += 1

    def update(self, scene, params, fuse_border=True):
        """One frame of the SDF branch: trace, blend (+ border), flip."""
        self.trace_rays(scene, params)
        self.probe_update(fuse_border=fuse_border)
        if not fuse_border:
            self.border_update()
        self.end_frame()

    @property
    def current_irradiance(self):
        """The atlas written by the last update (IrradianceVolume::currentIrrdance, DDGIRenderer.cpp:246)."""
        return self.irradiance[self.ping_pong]

    @property
    def current_depth(self):
        return self.depth[self.ping_pong]


def build_mip(sdf, res, max_distance=2.0, flood_passes=4, cascades=1, cascade=0):
    """GPU coarse mip of a [res]^3 FP32 SDF tensor (GlobalSDFMipmap.comp + fillFlood)."""
    mr = (res + 3) // 4
    mip = torch.ones((mr, mr, mr * cascades), dtype=torch.float32, device="cuda")
    tmp = torch.ones((mr, mr, mr), dtype=torch.float32, device="cuda")
    check(lib().luxgi_sdf_build_mip(ptr(sdf), res, cascades, cascade, float(max_distance), flood_passes, ptr(mip),
                                    ptr(tmp), stream_handle()), "luxgi_sdf_build_mip")
    return mip


def bake_sdf(prims, res, center=(0.0, 0.0, 0.0), half=1.0):
    """GPU bake of the analytic scene into a [res]^3 FP32 SDF tensor."""
    dp = to_device(np.frombuffer(prims.tobytes(), dtype=np.uint8))
    out = torch.empty((res, res, res), dtype=torch.float32, device="cuda")
    c = (C.c_float * 3)(*center)
    check(lib().luxgi_scene_bake_sdf(ptr(dp), len(prims), res, c, float(half), ptr(out), stream_handle()),
          "luxgi_scene_bake_sdf")
    torch.cuda.current_stream().synchronize()
    return out

"""GPU: the batched reflection denoise (one persistent launch per pass over frames of uneven
sizes) against the single-frame path, bit for bit; every schedule gives the same bits; a captured
graph replayed over changing inputs matches eager runs; the autotuner picks a candidate and caches
it. The batched kernels run the same block bodies as the single-frame kernels, so equality is
exact."""
import ctypes as C

import numpy as np
import pytest

from luxgi_amd import abi
from tests import _svgf_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SIZES = [(77, 45), (8, 8), (128, 64), (33, 97), (200, 120)]


def _frame(W, H, seed):
    g, prev, noisy, _, _, _, cam = _svgf_inputs.make_pair(W, H, seed)
    dev = lambda d: {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}  # noqa: E731
    return dev(g), dev(prev), torch.from_numpy(noisy).cuda(), cam


def _inputs(steps, dt=torch.float32):
    """steps x frames of (cur, prev, noisy, cam); the noisy signal changes every step."""
    out = []
    for s in range(steps):
        row = []
        for k, (w, h) in enumerate(SIZES):
            cur, prev, noisy, cam = _frame(w, h, 10 * k + 1)
            noisy = (noisy * (1.0 + 0.1 * s)).to(dt).contiguous()
            row.append((cur, prev, noisy, cam))
        out.append(row)
    return out


def _set(batch, row):
    for k, (cur, prev, noisy, cam) in enumerate(row):
        batch.set_frame(k, cur, prev, noisy, cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])


@pytest.mark.parametrize("storage,iterations,order", [(abi.STORAGE_FP32, 5, 0), (abi.STORAGE_FP16, 2, 1),
                                                      (abi.STORAGE_FP32, 0, 0), (abi.STORAGE_FP32, 5, 2)])
def test_batched_matches_single_frame_denoise(storage, iterations, order):
    from luxgi_amd.batch import ReflectionDenoiseBatch
    from luxgi_amd.svgf import ReflectionDenoiser
    dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    steps = _inputs(3, dt)
    singles = [ReflectionDenoiser(w, h, storage=storage, iterations=iterations) for w, h in SIZES]
    batch = ReflectionDenoiseBatch(SIZES, storage=storage, iterations=iterations, grid_blocks=37, order=order)
    for s, row in enumerate(steps):
        want = []
        for k, (cur, prev, noisy, cam) in enumerate(row):
            singles[k].set_camera(cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])
            want.append(singles[k].denoise(cur, prev, noisy).clone())
        _set(batch, row)
        got = batch.denoise()
        torch.cuda.synchronize()
        for k in range(len(SIZES)):
            print(f"step {s} frame {k} {SIZES[k]}", flush=True)
            assert torch.equal(got[k], want[k])
            assert torch.equal(batch.outputs[k][1], singles[k].outputs[1])
            assert torch.equal(batch.moments[k][s & 1], singles[k].moments[s & 1])
            assert torch.equal(batch.tile_maps[k][:singles[k].tile_map.numel()], singles[k].tile_map)


def test_every_schedule_gives_the_same_bits():
    from luxgi_amd.batch import ReflectionDenoiseBatch
    row = _inputs(1)[0]
    batch = ReflectionDenoiseBatch(SIZES, iterations=5)
    _set(batch, row)
    batch.upload()
    saved = batch.state()
    ref = None
    total = int(batch.header.totalBlocks)
    for grid, orders in ((1, (0, 1)), (3, (0, 1)), (64, (0, 1)), (256, (0, 1)), (5000, (0, 1)), (total, (2,))):
        for order in orders:
            batch.restore(saved)
            batch.grid_blocks, batch.order = grid, order
            batch.launch()
            torch.cuda.synchronize()
            out = [t.clone() for t in batch.results()]
            if ref is None:
                ref = out
            else:
                assert all(torch.equal(a, b) for a, b in zip(out, ref)), (grid, order)


def test_graph_replay_follows_changing_inputs():
    """Replays with new noisy tensors, cameras and the flipped ping-pong each frame, eager work and
    synchronisations between them; every replay equals the eager batch."""
    from luxgi_amd.batch import ReflectionDenoiseBatch
    steps = _inputs(4)
    eager = ReflectionDenoiseBatch(SIZES, iterations=4)
    graphed = ReflectionDenoiseBatch(SIZES, iterations=4)
    _set(graphed, steps[0])
    graphed.capture()
    for s, row in enumerate(steps):
        _set(eager, row)
        want = [t.clone() for t in eager.denoise()]
        _set(graphed, row)
        got = graphed.denoise()
        x = torch.randn(1000, device="cuda")
        (x * 2.0).sum().item()  # unrelated eager work between replays
        torch.cuda.synchronize()
        for k in range(len(SIZES)):
            assert torch.equal(got[k], want[k]), (s, k)


def test_autotuner_chooses_and_caches(tmp_path):
    from luxgi_amd.autotune import Autotuner, svgf_batch_candidates, tune_svgf_batch
    from luxgi_amd.batch import ReflectionDenoiseBatch
    row = _inputs(1)[0]
    batch = ReflectionDenoiseBatch(SIZES, iterations=4)
    _set(batch, row)
    tuner = Autotuner(path=str(tmp_path / "tune.json"))
    choice = tune_svgf_batch(batch, tuner)
    assert choice in svgf_batch_candidates(batch)
    assert [batch.grid_blocks, batch.order] == choice
    assert len(tuner.last_timings) == len(svgf_batch_candidates(batch))
    again = Autotuner(path=str(tmp_path / "tune.json"))
    assert tune_svgf_batch(batch, again) == choice and again.last_timings == {}


# ---------------------------------------------------------------------------------------------
# Several probe volumes per graph replay
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("trace,streams,blend", [("batched", 4, "auto"), ("per_volume", 4, "auto"),
                                                  ("batched", 0, "auto"), ("batched", 4, "per_volume"),
                                                  ("per_volume", 0, "batched")])
def test_ddgi_batch_graph_replay_matches_per_volume_updates(trace, streams, blend):
    """Three volumes of different probe and ray counts over the small scene, five frames with new
    random orientations each: the graph-replayed batch (orientations read from device memory) and
    the eager batch equal independent per-volume updates (orientations as host parameters), bit
    for bit, frame after frame, with the volumes traced in one launch over their flat ray list or
    one call per volume, and blended in one launch per map or one call per volume."""
    from luxgi_amd import scene
    from luxgi_amd.batch import DDGIBatch
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene, default_trace_params
    from tests import _scenes
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    ds = DeviceScene(g, sdf, mip, atlas, layout=abi.LAYOUT_BRICK4)
    unis = [_scenes.small_uniform(rays=64, counts=(5, 4, 6)), _scenes.small_uniform(rays=96, counts=(3, 3, 3)),
            _scenes.small_uniform(rays=128, counts=(6, 2, 4))]
    graphed = DDGIBatch(unis, ds, streams=streams, trace=trace, blend=blend)
    eager = DDGIBatch(unis, ds, streams=streams, trace=trace, blend=blend)
    assert graphed.blend_mode == ("per_volume" if blend == "per_volume" else "batched")
    singles = [DDGIVolume(u) for u in unis]
    rng = np.random.default_rng(5)
    for frame in range(5):
        rots = [scene.random_rotation(int(s)) for s in rng.integers(0, 2 ** 31, size=len(unis))]
        for vol, r in zip(singles, rots):
            vol.update(ds, default_trace_params(r, frame=vol.frames))
        eager.update(rots)
        graphed.update(rots)
        if frame == 0:
            graphed.capture()
        x = torch.randn(4096, device="cuda")
        (x * 3.0).sum().item()  # eager work between replays
        torch.cuda.synchronize()
        for k, vol in enumerate(singles):
            for b in (graphed, eager):
                other = b.volumes[k]
                assert other.ping_pong == vol.ping_pong and other.frames == vol.frames
                print(f"frame {frame} volume {k}", flush=True)
                assert torch.equal(other.current_irradiance, vol.current_irradiance)
                assert torch.equal(other.current_depth, vol.current_depth)
                assert torch.equal(other.radiance, vol.radiance)
    assert graphed.graphs[0] is not None and graphed.graphs[1] is not None


def test_autotuner_trace_shade_order(tmp_path):
    """tune_trace_shade_order times both shading orders of the trace (identical outputs), sets the
    fastest in the parameters and caches it."""
    from luxgi_amd.autotune import Autotuner, tune_trace_shade_order
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    from tests import _scenes
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    ds = DeviceScene(g, sdf, mip, atlas, layout=abi.LAYOUT_BRICK4)
    vol = DDGIVolume(_scenes.small_uniform(rays=64, counts=(5, 4, 6)))
    p = _scenes.trace_params(rot_seed=2)
    tuner = Autotuner(path=str(tmp_path / "tune.json"))
    choice = tune_trace_shade_order(vol, ds, p, tuner)
    assert choice in (0, 1) and p.shadeOrder == choice
    assert set(tuner.last_timings) == {"0", "1"}
    p2 = _scenes.trace_params(rot_seed=2)
    assert tune_trace_shade_order(vol, ds, p2, Autotuner(path=str(tmp_path / "tune.json"))) == choice
    assert p2.shadeOrder == choice


def _scene_set(storage=abi.STORAGE_FP32, layout=abi.LAYOUT_CELL8):
    """Three different seeded scenes (primitive sets, SDF resolutions 32, 48 and 64, atlases), all
    stored alike, for the several-scene batch."""
    from luxgi_amd import scene
    from luxgi_amd.ddgi import DeviceScene, bake_sdf, build_mip
    out = []
    for seed, res in ((11, 32), (12, 48), (13, 64)):
        prims = scene.scene_random(n=60, seed=seed, size_range=(0.1, 0.5))
        sdf = bake_sdf(prims, res)
        mip = build_mip(sdf, res, max_distance=2.0)
        g = abi.make_sdf_data(resolution=res)
        atlas = scene.build_atlas(prims, resolution=512, tile=16, seed=seed, threshold=1.05 * 2.0 / res)
        out.append(DeviceScene(g, sdf, mip, atlas, sdf_storage=storage, layout=layout))
    return out


@pytest.mark.parametrize("storage", [abi.STORAGE_FP32, abi.STORAGE_FP16])
def test_several_scene_batch_matches_per_volume_traces_on_their_scenes(storage):
    """Five volumes of different grids and ray counts over three different scenes (SDF 32^3, 48^3,
    64^3, their own atlases), traced in one launch sequence (luxgi_ddgi_trace_sdf_scenes) and
    graph-replayed over four frames: every volume's rays and atlases equal those of a
    luxgi_ddgi_trace_sdf + update on its own scene, bit for bit."""
    from luxgi_amd import scene
    from luxgi_amd.batch import DDGIBatch
    from luxgi_amd.ddgi import DDGIVolume, default_trace_params
    from tests import _scenes
    scenes = _scene_set(storage)
    unis = [_scenes.small_uniform(rays=64, counts=(5, 4, 6)), _scenes.small_uniform(rays=96, counts=(3, 3, 3)),
            _scenes.small_uniform(rays=128, counts=(6, 2, 4)), _scenes.small_uniform(rays=64, counts=(2, 2, 2)),
            _scenes.small_uniform(rays=192, counts=(4, 3, 2))]
    scene_of = [0, 1, 2, 1, 0]
    graphed = DDGIBatch(unis, scenes, scene_of=scene_of, streams=2)
    singles = [DDGIVolume(u) for u in unis]
    rng = np.random.default_rng(8)
    for frame in range(4):
        rots = [scene.random_rotation(int(s)) for s in rng.integers(0, 2 ** 31, size=len(unis))]
        for i, (vol, r) in enumerate(zip(singles, rots)):
            vol.update(scenes[scene_of[i]], default_trace_params(r, frame=vol.frames))
        graphed.update(rots)
        if frame == 0:
            graphed.capture()
        torch.cuda.synchronize()
        for k, vol in enumerate(singles):
            other = graphed.volumes[k]
            print(f"frame {frame} volume {k} scene {scene_of[k]}", flush=True)
            assert torch.equal(other.radiance, vol.radiance)
            assert torch.equal(other.dir_dist, vol.dir_dist)
            assert torch.equal(other.current_irradiance, vol.current_irradiance)
            assert torch.equal(other.current_depth, vol.current_depth)
    hits = [(v.dir_dist[..., 3] < 1000).float().mean().item() for v in singles]
    assert min(hits) > 0.03 and max(hits) < 0.95, hits  # the scenes are hit and missed


def test_several_scene_batch_rejects_mixed_storage_and_bad_scene_index():
    """All scenes of one batch store their SDF alike, and every volume names a scene of the batch:
    prepare rejects the rest on the host (no GPU call)."""
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import DDGIVolume
    from tests import _scenes
    L = _lib.lib()
    s32 = _scene_set(abi.STORAGE_FP32)
    s16 = _scene_set(abi.STORAGE_FP16)
    vol = DDGIVolume(_scenes.small_uniform(rays=64, counts=(2, 2, 2)))
    sc = (abi.TraceScene * 2)()
    for j, ds in enumerate((s32[0], s16[1])):
        sc[j].sdfData, sc[j].volume, sc[j].atlas, sc[j].skybox = ds.sdf_data, ds.volume, ds.atlas, ds.skybox
    d = (abi.TraceVolume * 1)()
    d[0].uniform, d[0].rays, d[0].scene = vol.u, vol.rays, 0
    host = (C.c_uint8 * L.luxgi_ddgi_trace_scenes_descriptor_bytes(2))()
    batch = abi.TraceSceneBatch()
    assert L.luxgi_ddgi_trace_scenes_prepare(C.byref(sc), 2, C.byref(d), 1, abi.STORAGE_FP32, C.byref(host),
                                             C.byref(batch)) == -1
    assert L.luxgi_ddgi_trace_scenes_prepare(C.byref(sc), 1, C.byref(d), 1, abi.STORAGE_FP32, C.byref(host),
                                             C.byref(batch)) == 0
    assert batch.totalRays == 8 * 64 and batch.sceneCount == 1
    d[0].scene = 1
    assert L.luxgi_ddgi_trace_scenes_prepare(C.byref(sc), 1, C.byref(d), 1, abi.STORAGE_FP32, C.byref(host),
                                             C.byref(batch)) == -1


# ---------------------------------------------------------------------------------------------
# Batched probe update (luxgi_ddgi_probe_update_batched) against per-volume updates
# ---------------------------------------------------------------------------------------------
_BLEND_VOLUMES = [  # (rays, ray buffer width, probe counts, irradiance oct, depth oct)
    (64, 64, (5, 4, 6), 6, 14), (96, 96, (3, 3, 3), 8, 16), (100, 128, (4, 2, 3), 4, 8),
    (256, 256, (2, 3, 2), 6, 14), (32, 32, (1, 1, 1), 8, 12), (160, 160, (9, 5, 7), 6, 14)]


def _fill_random_rays(vol, rng, fixup_probes=3):
    """Rays of one frame: every probe shares ray row 0's directions except `fixup_probes` probes
    with their own (the fix-up path), distances with misses beyond maxDistance and backfaces."""
    P, W, R = vol.P, vol.ray_width, int(vol.u.raysPerProbe)
    d = rng.normal(size=(R, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dd = np.zeros((P, W, 4), dtype=np.float32)
    dd[:, :R, :3] = d
    for p in rng.choice(P, size=min(fixup_probes, P - 1), replace=False) if P > 1 else []:
        e = rng.normal(size=(R, 3))
        dd[p, :R, :3] = e / np.linalg.norm(e, axis=1, keepdims=True)
    dist = rng.uniform(0.02, 1.5 * float(vol.u.maxDistance), size=(P, W))
    dist[rng.random((P, W)) < 0.1] *= -0.2
    dd[..., 3] = dist
    rad = rng.uniform(0.0, 2.0, size=(P, W, 4)).astype(np.float32)
    vol.radiance.copy_(torch.from_numpy(rad).to(vol.radiance.dtype))
    vol.dir_dist.copy_(torch.from_numpy(dd).to(vol.dir_dist.dtype))


def _blend_volume_set(ray_storage, atlas_storage):
    from luxgi_amd.ddgi import DDGIVolume
    from tests import _scenes
    out = []
    for rays, width, counts, ni, nd in _BLEND_VOLUMES:
        u = _scenes.small_uniform(rays=rays, counts=counts, ni=ni, nd=nd)
        out.append(DDGIVolume(u, ray_width=width, ray_storage=ray_storage, atlas_storage=atlas_storage))
    return out


@pytest.mark.parametrize("form", ["plain", "half", "bf16"])
@pytest.mark.parametrize("storages", [(abi.STORAGE_FP32, abi.STORAGE_FP32), (abi.STORAGE_FP16, abi.STORAGE_FP16),
                                      (abi.STORAGE_FP32, abi.STORAGE_FP16)])
@pytest.mark.parametrize("fuse", [True, False])
def test_batched_update_equals_per_volume_updates(form, storages, fuse):
    """Six volumes of different grids, ray counts (some not multiples of 32, one with a wider ray
    buffer) and octahedral sizes, three frames (the first without hysteresis), a few probes per
    frame with their own ray directions: one luxgi_ddgi_probe_update_batched per frame writes the
    atlases of a luxgi_ddgi_probe_update per volume, bit for bit, for each blend form, storage and
    border mode."""
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import stream_handle
    lib = _lib.lib()
    rs, ast = storages
    singles = _blend_volume_set(rs, ast)
    batched = _blend_volume_set(rs, ast)
    half, bf16 = form in ("half", "bf16"), form == "bf16"
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | (abi.UPDATE_FUSE_BORDER if fuse else 0)
    flags |= (abi.UPDATE_HALF if half else 0) | (abi.UPDATE_BF16 if bf16 else 0)
    k = len(batched)
    vols = (abi.UpdateVolume * k)()
    ws_bytes = lib.luxgi_ddgi_update_batch_workspace_size(C.byref(vols), 0)
    assert ws_bytes == 0  # count 0 is rejected
    for i, v in enumerate(batched):
        vols[i].uniform, vols[i].rays = v.u, v.rays
    ws_bytes = lib.luxgi_ddgi_update_batch_workspace_size(C.byref(vols), k)
    assert ws_bytes >= sum(lib.luxgi_ddgi_probe_update_workspace_size(C.byref(v.u)) for v in batched)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
    nbytes = lib.luxgi_ddgi_update_batch_descriptor_bytes(k)
    rng = np.random.default_rng(21)
    for frame in range(3):
        for a, b in zip(singles, batched):
            _fill_random_rays(a, rng)
            b.radiance.copy_(a.radiance)
            b.dir_dist.copy_(a.dir_dist)
            a.probe_update(fuse_border=fuse, half=half, bf16=bf16)
            if not fuse:
                a.border_update()
        for i, v in enumerate(batched):
            vols[i].prev, vols[i].out = v.atlas[v.ping_pong], v.atlas[1 - v.ping_pong]
            vols[i].firstFrame = 1 if frame == 0 else 0
        host = (C.c_uint8 * nbytes)()
        bt = abi.UpdateBatch()
        assert lib.luxgi_ddgi_update_batch_prepare(C.byref(vols), k, flags, C.c_void_p(ws.data_ptr()), ws_bytes,
                                                   C.byref(host), C.byref(bt)) == 0
        assert bt.count == k and bt.groups == sum((v.P + 15) // 16 for v in batched)
        desc = torch.from_numpy(np.frombuffer(host, dtype=np.uint8).copy()).cuda()
        print(f"batched update frame {frame} form {form} storages {storages} fuse {fuse} groups {bt.groups}",
              flush=True)
        assert lib.luxgi_ddgi_probe_update_batched(C.byref(bt), C.c_void_p(desc.data_ptr()), stream_handle()) == 0
        for v in batched:
            if not fuse:
                v.border_update()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(singles, batched)):
            w = 1 - a.ping_pong
            assert torch.equal(a.irradiance[w], b.irradiance[w]), (frame, i, "irradiance")
            assert torch.equal(a.depth[w], b.depth[w]), (frame, i, "depth")
            assert torch.isfinite(b.irradiance[w].float()).all()
            a.end_frame()
            b.end_frame()

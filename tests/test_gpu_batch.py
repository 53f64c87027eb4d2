"""GPU: the batched reflection denoise (one persistent launch per pass over frames of uneven
sizes) against the single-frame path, bit for bit; every schedule gives the same bits; a captured
graph replayed over changing inputs matches eager runs; the autotuner picks a candidate and caches
it. The batched kernels run the same block bodies as the single-frame kernels, so equality is
exact."""
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
@pytest.mark.parametrize("trace,streams", [("batched", 4), ("per_volume", 4), ("batched", 0)])
def test_ddgi_batch_graph_replay_matches_per_volume_updates(trace, streams):
    """Three volumes of different probe and ray counts over the small scene, five frames with new
    random orientations each: the graph-replayed batch (orientations read from device memory) and
    the eager batch equal independent per-volume updates (orientations as host parameters), bit
    for bit, frame after frame, with the volumes traced in one launch over their flat ray list or
    one call per volume."""
    from luxgi_amd import scene
    from luxgi_amd.batch import DDGIBatch
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene, default_trace_params
    from tests import _scenes
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    ds = DeviceScene(g, sdf, mip, atlas, layout=abi.LAYOUT_BRICK4)
    unis = [_scenes.small_uniform(rays=64, counts=(5, 4, 6)), _scenes.small_uniform(rays=96, counts=(3, 3, 3)),
            _scenes.small_uniform(rays=128, counts=(6, 2, 4))]
    graphed = DDGIBatch(unis, ds, streams=streams, trace=trace)
    eager = DDGIBatch(unis, ds, streams=streams, trace=trace)
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

"""GPU parity: SDF bake, coarse mip, probe-ray trace (luxgi_scene_bake_sdf,
luxgi_sdf_build_mip, luxgi_ddgi_trace_sdf[_count]) against numpy / the FP64 oracle.

Tolerances. The mip is a min of sqrt's of fp32 inputs: 2 ulp (4e-7 relative).
The trace is a chain of threshold decisions (hit test, mip/fine choice, tile
visibility); a decision taken on a value within fp32 rounding of its threshold
can flip, so parity is stated per ray population:
  * ray class (miss / hit / inside) agrees on >= 99% of rays;
  * on >= 98% of rays the stored distance agrees to 1e-4 * (1 + t) and the
    radiance to 2e-3 (absolute, radiance ~ 0..4);
  * on every ray of the same class the distance agrees to one voxel;
  * directions agree to 2e-6 + 4e-7 R on every ray (fp32 spherical-Fibonacci angle).
The counting run's totals agree with the oracle's counts to 1%."""
import ctypes as C

import numpy as np
import pytest

import oracle
from luxgi_amd import abi, scene
from tests import _scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_bake_sdf_matches_numpy():
    from luxgi_amd.ddgi import bake_sdf
    prims = scene.scene_random(n=60, seed=9)
    got = bake_sdf(prims, 48).cpu().numpy()
    want = scene.bake_sdf_numpy(prims, 48)
    np.testing.assert_allclose(got, want, atol=2e-6)


def test_mip_matches_oracle():
    from luxgi_amd.ddgi import build_mip
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    got = build_mip(torch.from_numpy(sdf).cuda(), 64, max_distance=2.0).cpu().numpy()
    np.testing.assert_allclose(got, mip, rtol=4e-7, atol=1e-7)
    # 96^3 (mip res 24, not a multiple of the 8x8x4 block)
    s2 = scene.bake_sdf_numpy(prims, 96)
    got2 = build_mip(torch.from_numpy(s2).cuda(), 96, max_distance=2.0).cpu().numpy()
    np.testing.assert_allclose(got2, oracle.sdf_mip(s2, 96, max_distance=2.0), rtol=4e-7, atol=1e-7)


def _gpu_trace(u, g, sdf, mip, atlas, params, count=False, storage=abi.STORAGE_FP32, offsets=None, sky=None,
               layout=abi.LAYOUT_LINEAR):
    from luxgi_amd.ddgi import DeviceScene, DDGIVolume
    ds = DeviceScene(g, sdf, mip, atlas, layout=layout, sky_faces=sky, sky_size=0 if sky is None else sky.shape[1])
    vol = DDGIVolume(u, ray_storage=storage, flags_buffer=True)
    if offsets is not None:
        vol.offsets.copy_(torch.from_numpy(offsets))
    counters = torch.zeros(8, dtype=torch.int64, device="cuda") if count else None
    vol.trace_rays(ds, params, use_offsets=offsets is not None, counters=counters)
    torch.cuda.synchronize()
    out = (vol.radiance.float().cpu().numpy(), vol.dir_dist.float().cpu().numpy(), vol.ray_flags.cpu().numpy())
    return out + ((counters.cpu().numpy() if count else None),)


def _compare(u, g, gpu, ref, voxel, good_fraction=0.98):
    grad, gdd, gfl = gpu[:3]
    orad, odd, ofl, _ = ref
    n = gfl.size
    same = gfl == ofl
    assert same.mean() >= 0.99, same.mean()
    # ray directions: the angle 2*pi*frac(i*(PHI-1)) is fp32 in the reference (DDGICommon.glsl:47),
    # so its rounding grows with the ray index i < R
    np.testing.assert_allclose(gdd[..., :3], odd[..., :3], atol=2e-6 + 4e-7 * u.raysPerProbe)
    t_ok = np.abs(gdd[..., 3] - odd[..., 3]) <= 1e-4 * (1 + np.abs(odd[..., 3]))
    r_ok = np.all(np.abs(grad[..., :3] - orad[..., :3]) <= 2e-3, axis=-1)
    good = same & t_ok & r_ok
    assert good.mean() >= good_fraction, (good.mean(), n)
    assert np.all(np.abs(gdd[..., 3] - odd[..., 3])[same] <= voxel + 1e-4 * (1 + np.abs(odd[..., 3][same])))
    assert np.all(grad[..., 3] == 0)


@pytest.mark.parametrize("max_steps,hit_eps", [(250, 0.0), (64, 1e-3)])
def test_trace_matches_oracle_config1(max_steps, hit_eps):
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64)
    p = _scenes.trace_params(max_steps=max_steps, hit_eps=hit_eps)
    gpu = _gpu_trace(u, g, sdf, mip, atlas, p)
    ref = _scenes.oracle_trace(u, g, sdf, mip, atlas, p)
    _compare(u, g, gpu, ref, 2.0 / 64)
    fl = gpu[2]
    assert (fl == 1).mean() > 0.3 and (fl == 0).mean() > 0.01  # the scene has hits and misses


def test_trace_unorm8_atlas_and_offsets():
    prims, sdf, mip, g, atlas = _scenes.small_scene(color_storage=abi.STORAGE_UNORM8)
    u = _scenes.small_uniform(rays=96, counts=(5, 4, 6))
    p = _scenes.trace_params(rot_seed=11)
    rng = np.random.default_rng(2)
    off = np.zeros((u.num_probes, 4), np.float32)
    off[:, :3] = rng.uniform(-0.05, 0.05, (u.num_probes, 3))
    gpu = _gpu_trace(u, g, sdf, mip, atlas, p, offsets=off)
    ref = _scenes.oracle_trace(u, g, sdf, mip, atlas, p, probe_offsets=off)
    _compare(u, g, gpu, ref, 2.0 / 64)


@pytest.mark.parametrize("order", [0, 1])
def test_trace_sky_cubemap_matches_oracle(order):
    """Misses sample the sky cube map (GISDFRays.comp:121-124); with the hit sort (order 0) the
    march stores them, in ray order the shade pass does."""
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=128, counts=(5, 4, 6))
    p = _scenes.trace_params(rot_seed=17)
    p.shadeOrder = order
    sky = np.random.default_rng(8).uniform(0.0, 3.0, size=(6, 8, 8, 4)).astype(np.float32)
    gpu = _gpu_trace(u, g, sdf, mip, atlas, p, sky=sky, layout=abi.LAYOUT_BRICK4)
    ref = _scenes.oracle_trace(u, g, sdf, mip, atlas, p, sky_faces=sky, sky_size=8)
    _compare(u, g, gpu, ref, 2.0 / 64)
    miss = (gpu[2] == 0) & (ref[2] == 0)
    assert miss.sum() > 50
    # the fp32 ray direction may differ by 2e-6 + 4e-7 R (see _compare); through the cube map
    # (size / 2 / |major axis| <= 6.9 texels per unit at size 8, texel steps <= 3) that is <= 1.5e-3
    np.testing.assert_allclose(gpu[0][miss][:, :3], ref[0][miss][:, :3], atol=1.5e-3)
    assert np.ptp(gpu[0][miss][:, 0]) > 0.5  # the misses saw different sky texels


def test_trace_counts_match_oracle():
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64)
    p = _scenes.trace_params()
    gpu = _gpu_trace(u, g, sdf, mip, atlas, p, count=True)
    ref = _scenes.oracle_trace(u, g, sdf, mip, atlas, p, want_counts=True)
    gc, oc = gpu[3].astype(np.float64), ref[3].astype(np.float64)
    assert gc[5] == oc[5] == u.num_probes * 64
    # the reference evaluates the hit normal inside the march for every hit (SDFCommon.glsl:172-184);
    # the GPU evaluates it only where the result is used (not for inside hits, GISDFRays.comp:93-96)
    oc[2] -= float((ref[2] == 2).sum())
    for k in range(8):
        assert abs(gc[k] - oc[k]) <= 0.01 * max(oc[k], 1.0), (k, gc, oc)
    # the plain trace writes the same outputs as the counting trace
    plain = _gpu_trace(u, g, sdf, mip, atlas, p)
    assert np.array_equal(plain[1], gpu[1]) and np.array_equal(plain[2], gpu[2])


def test_trace_brick_layout_is_bitwise_identical():
    """LUXGI_LAYOUT_BRICK4 and LUXGI_LAYOUT_CELL8 only change where voxels live and the brick guard
    only skips fine lookups whose value is not used: outputs equal the plain linear trace's bit for
    bit."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=128, counts=(6, 5, 4))
    p = _scenes.trace_params(rot_seed=21)
    outs = []
    for layout, guard in ((abi.LAYOUT_LINEAR, False), (abi.LAYOUT_BRICK4, False), (abi.LAYOUT_BRICK4, True),
                          (abi.LAYOUT_LINEAR, True), (abi.LAYOUT_CELL8, False), (abi.LAYOUT_CELL8, True)):
        ds = DeviceScene(g, sdf, mip, atlas, layout=layout, guard=guard)
        vol = DDGIVolume(u, flags_buffer=True)
        vol.trace_rays(ds, p)
        torch.cuda.synchronize()
        outs.append((vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy()))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("shape,storage", [((13, 7, 9), abi.STORAGE_FP32), ((24, 16, 8), abi.STORAGE_FP16),
                                           ((1, 1, 1), abi.STORAGE_FP32)])
def test_cell_records_hold_the_clamped_corners(shape, storage):
    """luxgi_sdf_to_cells: record of cell (x, y, z) = the 8 voxels (x+i, y+j, z+k) clamped to the
    volume, at brick (x/8, y/8, z/8) x-fastest, slot (z%8)*64 + (y%8)*8 + x%8 inside the brick."""
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    W, H, D = shape
    vol = np.random.default_rng(4).uniform(-1, 1, (D, H, W)).astype(np.float32)
    nbytes = _lib.lib().luxgi_sdf_cells_bytes(W, H, D, storage)
    nbx, nby, nbz = (W + 7) // 8, (H + 7) // 8, (D + 7) // 8
    assert nbytes == nbx * nby * nbz * 512 * 8 * (2 if storage == abi.STORAGE_FP16 else 4)
    dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    out = torch.full((nbytes // (2 if storage == abi.STORAGE_FP16 else 4),), 7.0, dtype=dt, device="cuda")
    src = torch.from_numpy(vol).cuda()
    assert _lib.lib().luxgi_sdf_to_cells(ptr(src), W, H, D, storage, ptr(out), stream_handle()) == 0
    torch.cuda.synchronize()
    rec = out.float().cpu().numpy().reshape(-1, 8)
    ref = vol.astype(np.float16).astype(np.float32) if storage == abi.STORAGE_FP16 else vol

    def slot(x, y, z):
        return ((z & 7) << 6) | ((y & 7) << 3) | (x & 7)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                idx = (((z >> 3) * nby + (y >> 3)) * nbx + (x >> 3)) * 512 + slot(x, y, z)
                want = [ref[min(z + k, D - 1), min(y + j, H - 1), min(x + i, W - 1)]
                        for k in (0, 1) for j in (0, 1) for i in (0, 1)]
                assert np.array_equal(rec[idx], np.array(want, np.float32)), (x, y, z)
    written = np.zeros(rec.shape[0], bool)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                written[(((z >> 3) * nby + (y >> 3)) * nbx + (x >> 3)) * 512 + slot(x, y, z)] = True
    assert np.all(rec[~written] == 0)  # padding records of partial bricks


def test_trace_fp16_rays():
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64, counts=(4, 4, 4))
    p = _scenes.trace_params(rot_seed=3)
    g32 = _gpu_trace(u, g, sdf, mip, atlas, p)
    g16 = _gpu_trace(u, g, sdf, mip, atlas, p, storage=abi.STORAGE_FP16)
    # fp16 storage is the fp32 result rounded to half
    np.testing.assert_array_equal(g16[1], g32[1].astype(np.float16).astype(np.float32))
    np.testing.assert_array_equal(g16[0], g32[0].astype(np.float16).astype(np.float32))


def test_trace_packed_atlas_is_bitwise_identical():
    """LUXGI_LAYOUT_BRICK4 surface atlas (colour + depth interleaved in 4x4-texel blocks) gives the
    same hit radiance as the two linear images, bit for bit."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene(color_storage=abi.STORAGE_UNORM8)
    u = _scenes.small_uniform(rays=96, counts=(6, 5, 4))
    p = _scenes.trace_params(rot_seed=23)
    outs = []
    for al in (abi.LAYOUT_LINEAR, abi.LAYOUT_BRICK4):
        ds = DeviceScene(g, sdf, mip, atlas, layout=abi.LAYOUT_BRICK4, atlas_layout=al)
        vol = DDGIVolume(u, flags_buffer=True)
        vol.trace_rays(ds, p)
        torch.cuda.synchronize()
        outs.append((vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy()))
    assert (outs[0][0][..., :3] > 0).any()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_trace_fp16_sdf_bricks_match_linear_and_oracle():
    """FP16 SDF storage: the brick layout is bit-identical to the linear FP16 layout, and both match
    the oracle run on the same FP16-rounded volume (same population statistics as FP32)."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64)
    p = _scenes.trace_params(rot_seed=17)
    outs = []
    for layout in (abi.LAYOUT_LINEAR, abi.LAYOUT_BRICK4, abi.LAYOUT_CELL8):
        ds = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_FP16, layout=layout)
        vol = DDGIVolume(u, flags_buffer=True)
        vol.trace_rays(ds, p)
        torch.cuda.synchronize()
        outs.append((vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy()))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
    s16 = sdf.astype(np.float16).astype(np.float32)
    m16 = mip.astype(np.float16).astype(np.float32)
    ref = _scenes.oracle_trace(u, g, s16, m16, atlas, p)
    _compare(u, g, outs[1] + (None,), ref, 2.0 / 64)


def _bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_trace_bf16_sdf_layouts_match_each_other_and_oracle():
    """BF16 SDF storage (the reduced-precision form's other 16-bit format): the linear, brick and
    cell-record layouts give the same rays bit for bit, and they match the oracle run on the same
    BF16-rounded volume (same population statistics as FP32); with the FP32 re-evaluation at hits
    they match the FP64 oracle on the FP32 volume within the population bounds that BF16 rounding
    allows."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64)
    p = _scenes.trace_params(rot_seed=29)
    outs = []
    for layout in (abi.LAYOUT_LINEAR, abi.LAYOUT_BRICK4, abi.LAYOUT_CELL8):
        ds = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_BF16, layout=layout)
        vol = DDGIVolume(u, flags_buffer=True)
        vol.trace_rays(ds, p)
        torch.cuda.synchronize()
        outs.append((vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy()))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
    ref = _scenes.oracle_trace(u, g, _bf16_round(sdf), _bf16_round(mip), atlas, p)
    _compare(u, g, outs[2] + (None,), ref, 2.0 / 64)
    ds = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_BF16, layout=abi.LAYOUT_CELL8, refine=True)
    vol = DDGIVolume(u, flags_buffer=True)
    vol.trace_rays(ds, p)
    torch.cuda.synchronize()
    got = (vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy(), None)
    # Against FP64 on the FP32 volume, BF16's 2^-9 rounding moves the step at which 3.8% of the rays
    # detect their hit (the FP64 march on the BF16-rounded volume agrees with the one on the FP32
    # volume for 96.2% of the rays, every class the same); those rays stay within one voxel.
    _compare(u, g, got, _scenes.oracle_trace(u, g, sdf, mip, atlas, p), 2.0 / 64, good_fraction=0.95)


def test_trace_fp16_scaled_bricks_and_fp32_refine_vs_oracle():
    """Reduced-precision march: FP16 bricks with the per-brick power-of-two scale, with and without
    the FP32 re-evaluation at hits, against the FP64 oracle on the FP32 volume (same population
    bounds as the FP32 trace). The re-evaluation makes the hit distances FP32-accurate, so their
    error against the oracle does not grow."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64)
    p = _scenes.trace_params(rot_seed=19)
    ref = _scenes.oracle_trace(u, g, sdf, mip, atlas, p)
    errs = []
    for refine in (False, True):
        ds = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_FP16, layout=abi.LAYOUT_BRICK4,
                         brick_scale=True, refine=refine)
        vol = DDGIVolume(u, flags_buffer=True)
        vol.trace_rays(ds, p)
        torch.cuda.synchronize()
        got = (vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy(), None)
        _compare(u, g, got, ref, 2.0 / 64)
        hit = (got[2] == 1) & (ref[2] == 1)
        errs.append(float(np.abs(got[1][..., 3] - ref[1][..., 3])[hit].mean()))
    assert errs[1] <= errs[0] * 1.01 + 1e-7, errs


@pytest.mark.parametrize("layout,guard", [(abi.LAYOUT_BRICK4, False), (abi.LAYOUT_LINEAR, True)])
def test_trace_shade_order_is_bitwise_identical(layout, guard):
    """Shading the front hits grouped by surface-atlas chunk (the default: the march writes the
    misses and inside hits, a counting sort orders the rest) or every ray in ray order gives the
    same outputs bit for bit; the counting trace (always in ray order) agrees too."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=96, counts=(7, 5, 6))
    outs = []
    for order, count in ((0, False), (1, False), (0, True)):
        p = _scenes.trace_params(rot_seed=29)
        p.shadeOrder = order
        ds = DeviceScene(g, sdf, mip, atlas, layout=layout, guard=guard)
        vol = DDGIVolume(u, flags_buffer=True)
        counters = torch.zeros(8, dtype=torch.int64, device="cuda") if count else None
        for _ in range(2):  # the second call reuses the workspace: the bins must start from zero again
            vol.trace_rays(ds, p, counters=counters)
        torch.cuda.synchronize()
        outs.append((vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy()))
    assert (outs[0][2] == 1).any() and (outs[0][2] == 0).any()
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("layout", [abi.LAYOUT_LINEAR, abi.LAYOUT_BRICK4, abi.LAYOUT_CELL8])
def test_trace_multi_cascade_matches_oracle(layout):
    """A three-cascade clipmap (half extents 0.5, 1, 2 at 32^3, packed along x) of the small scene;
    probes inside and outside cascade 0, so rays start in different cascades and move out through
    the others. Same tolerances as the single-cascade trace, with the outer cascade's voxel."""
    prims, _, _, _, atlas = _scenes.small_scene()
    sdf, mip, g = _scenes.cascaded_sdf(prims, 32, (0.5, 1.0, 2.0))
    u = _scenes.small_uniform(rays=96, counts=(5, 4, 6))
    p = _scenes.trace_params(rot_seed=23)
    gpu = _gpu_trace(u, g, sdf, mip, atlas, p, layout=layout)
    ref = _scenes.oracle_trace(u, g, sdf, mip, atlas, p)
    _compare(u, g, gpu, ref, 2.0 * 2.0 / 32)
    fl = gpu[2]
    assert (fl == 1).mean() > 0.3 and (fl == 0).mean() > 0.01


def test_mip_of_each_cascade_matches_oracle():
    from luxgi_amd.ddgi import build_mip
    prims, _, _, _, _ = _scenes.small_scene()
    halves = (0.5, 1.0, 2.0)
    sdf, mip, g = _scenes.cascaded_sdf(prims, 32, halves)
    t = torch.from_numpy(sdf).cuda()
    for c, h in enumerate(halves):
        got = build_mip(t, 32, max_distance=2.0 * h, cascades=3, cascade=c).cpu().numpy()
        np.testing.assert_allclose(got[:, :, c * 8:(c + 1) * 8], mip[:, :, c * 8:(c + 1) * 8], rtol=4e-7, atol=1e-7)


@pytest.mark.parametrize("order,ray_storage", [(0, abi.STORAGE_FP32), (1, abi.STORAGE_FP32), (0, abi.STORAGE_FP16)])
def test_batched_trace_equals_per_volume_traces(order, ray_storage):
    """luxgi_ddgi_trace_sdf_batched over three volumes (different grids and ray counts, one with
    relocation offsets, host and device orientations) writes exactly what one luxgi_ddgi_trace_sdf
    call per volume writes: radiance, direction + distance and ray class."""
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene, ptr, stream_handle
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    ds = DeviceScene(g, sdf, mip, atlas, layout=abi.LAYOUT_BRICK4)
    unis = [_scenes.small_uniform(rays=64, counts=(5, 4, 6)), _scenes.small_uniform(rays=96, counts=(3, 3, 3)),
            _scenes.small_uniform(rays=128, counts=(6, 2, 4))]
    rng = np.random.default_rng(order)
    rots = [scene.random_rotation(int(s)) for s in rng.integers(0, 2 ** 31, size=3)]
    singles, batched = [], []
    for k, u in enumerate(unis):
        for lst in (singles, batched):
            v = DDGIVolume(u, ray_storage=ray_storage, flags_buffer=True)
            if k == 1:
                v.offsets.copy_(torch.from_numpy(rng.uniform(-0.05, 0.05, (u.num_probes, 4)).astype(np.float32)))
            lst.append(v)
        singles[k].offsets.copy_(batched[k].offsets)
    def m4(r):  # the 4x4 orientation (random_rotation's upper 3x3 in an identity frame)
        m = np.eye(4)
        m[:3, :3] = np.asarray(r)[:3, :3]
        return m

    dev_rot = torch.from_numpy(np.ascontiguousarray(m4(rots[2]).T.reshape(-1), np.float32)).cuda()
    for k, v in enumerate(singles):
        p = _scenes.trace_params(rot_seed=3)
        p.randomOrientation = abi.mat4_colmajor(m4(rots[k]))
        p.shadeOrder = order
        v.trace_rays(ds, p, use_offsets=(k == 1))
    d = (abi.TraceVolume * 3)()
    for k, v in enumerate(batched):
        d[k].uniform = v.u
        if k == 2:
            d[k].deviceOrientation = dev_rot.data_ptr()
        else:
            d[k].orientation = abi.mat4_colmajor(m4(rots[k]))
        d[k].probeOffsets = v.offsets.data_ptr() if k == 1 else None
        d[k].rays = v.rays
        d[k].rayFlags = v.ray_flags.data_ptr()
    total = C.c_int64(0)
    assert lib().luxgi_ddgi_trace_batch_prepare(C.byref(d), 3, ray_storage, C.byref(total)) == 0
    assert total.value == sum(u.num_probes * u.raysPerProbe for u in unis)
    dd = torch.from_numpy(np.frombuffer(memoryview(d).cast("B"), dtype=np.uint8).copy()).cuda()
    ws = torch.empty(lib().luxgi_ddgi_trace_batch_workspace_size(total.value, ds.n_objects), dtype=torch.uint8,
                     device="cuda")
    p = _scenes.trace_params(rot_seed=3)
    p.shadeOrder = order
    assert lib().luxgi_ddgi_trace_sdf_batched(ptr(dd), 3, total.value, ray_storage, C.byref(ds.sdf_data),
                                              C.byref(ds.volume), C.byref(ds.atlas), C.byref(ds.skybox), C.byref(p),
                                              ptr(ws), ws.numel(), stream_handle()) == 0
    torch.cuda.synchronize()
    for k in range(3):
        print(f"order {order} storage {ray_storage} volume {k}", flush=True)
        assert torch.equal(batched[k].radiance, singles[k].radiance)
        assert torch.equal(batched[k].dir_dist, singles[k].dir_dist)
        assert torch.equal(batched[k].ray_flags, singles[k].ray_flags)
    assert (singles[0].ray_flags == 1).any() and (singles[0].ray_flags == 0).any()


def test_trace_through_a_global_sdf_built_from_mesh_sdfs():
    """End to end on the GPU: the small scene as instances of baked mesh SDFs, rasterized into the
    global SDF (luxgi_sdf_rasterize), its coarse mip (luxgi_sdf_build_mip), then the probe trace;
    against the oracle's rasterization, mip and trace of the same instances. Same tolerances as
    the other trace tests (the FP32 rasterization is within 2e-5 of FP64, so only rays within
    rounding of a threshold can change class)."""
    from luxgi_amd import meshsdf
    from luxgi_amd.ddgi import build_mip
    prims, _, _, _, atlas = _scenes.small_scene()
    res = 64
    msh, objs, bounds, which = meshsdf.primitive_instances(prims)
    cas = meshsdf.Cascade(0, (0.0, 0.0, 0.0), 1.0, res)
    lists = meshsdf.chunk_lists(cas, bounds)
    b = meshsdf.GlobalSDFBuilder(msh, objs, len(prims), which=which)
    vol = torch.empty((res, res, res), dtype=torch.float32, device="cuda")
    b.rasterize(cas, lists, vol)
    gmip = build_mip(vol, res, max_distance=2.0)
    torch.cuda.synchronize()
    gsdf = vol.cpu().numpy()
    osdf = oracle.sdf_rasterize(cas.params(), objs, b.pool.entries, b.pool.pool, len(prims), *lists)
    np.testing.assert_allclose(gsdf, osdf, atol=2e-5)
    omip = oracle.sdf_mip(osdf.astype(np.float32), res, max_distance=2.0)
    np.testing.assert_allclose(gmip.cpu().numpy(), omip, atol=3e-5)
    g = abi.make_sdf_data(resolution=res)
    u = _scenes.small_uniform(rays=96, counts=(5, 4, 6))
    p = _scenes.trace_params(rot_seed=31)
    gpu = _gpu_trace(u, g, gsdf, gmip.cpu().numpy(), atlas, p, layout=abi.LAYOUT_BRICK4)
    ref = _scenes.oracle_trace(u, g, osdf.astype(np.float32), omip, atlas, p)
    _compare(u, g, gpu, ref, 2.0 / res)
    assert (gpu[2] == 1).mean() > 0.3

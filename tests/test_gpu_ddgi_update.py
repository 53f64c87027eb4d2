"""GPU parity: probe blend + border (luxgi_ddgi_probe_update / luxgi_ddgi_border_update)
against the FP64 oracle on the same seeded ray buffers.

Tolerances: FP32 storage 2e-5 relative (+1e-6 absolute) on irradiance and depth;
FP16 storage: the value rounded to fp16 (2^-11 relative) plus 2e-5. Border texels
and untouched padding are compared bit for bit with their sources."""
import ctypes as C

import numpy as np
import pytest

import oracle
from luxgi_amd import abi, scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _rays(u, rot_seed=3, seed=0, perturb=False):
    P, R = u.num_probes, u.raysPerProbe
    dirs = oracle.ray_directions(abi.mat4_colmajor(scene.random_rotation(rot_seed)), R).astype(np.float32)
    rng = np.random.default_rng(seed)
    rad = np.zeros((P, R, 4), np.float32)
    dd = np.zeros((P, R, 4), np.float32)
    dd[:, :, :3] = dirs[None]
    if perturb:  # some probes get their own directions: the per-probe weight path must run for them
        for p in range(0, P, 3):
            d = dirs + rng.normal(scale=0.05, size=dirs.shape).astype(np.float32)
            dd[p, :, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rad[:, :, :3] = rng.uniform(0, 2, (P, R, 3))
    dd[:, :, 3] = rng.uniform(0.05, 3.0, (P, R))
    dd[:, ::17, 3] = 60000.0  # misses / inside hits
    return rad, dd


def _run_gpu(u, rad, dd, prev_i, prev_d, first, flags, storage=abi.STORAGE_FP32):
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    t_rad = torch.from_numpy(rad).to(dt).cuda()
    t_dd = torch.from_numpy(dd).to(dt).cuda()
    rays = abi.RayBuffers(ptr(t_rad), ptr(t_dd), rad.shape[1], rad.shape[0], storage, 0)
    p_i = torch.from_numpy(prev_i).to(dt).cuda()
    p_d = torch.from_numpy(prev_d).to(dt).cuda()
    o_i = p_i.clone()
    o_d = p_d.clone()
    prev = abi.ProbeAtlas(ptr(p_i), ptr(p_d), storage, 0)
    out = abi.ProbeAtlas(ptr(o_i), ptr(o_d), storage, 0)
    ws = torch.empty(_lib.lib().luxgi_ddgi_probe_update_workspace_size(C.byref(u)), dtype=torch.uint8, device="cuda")
    rc = _lib.lib().luxgi_ddgi_probe_update(C.byref(u), C.byref(rays), C.byref(prev), C.byref(out), int(first), flags,
                                            ptr(ws), ws.numel(), stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    return o_i.float().cpu().numpy(), o_d.float().cpu().numpy()


CASES = [
    # (probe counts, rays, irr oct, depth oct)
    ((8, 8, 8), 64, 6, 14),     # config-1 shapes
    ((4, 3, 5), 256, 8, 16),    # reference oct sizes, 256 rays
    ((3, 2, 2), 300, 8, 16),    # ragged: rays not a multiple of the 256-ray chunk or of 64
    ((1, 1, 1), 1, 4, 4),       # single probe, single ray
]


@pytest.mark.parametrize("counts,rays,ni,nd", CASES)
@pytest.mark.parametrize("per_probe", [False, True])
def test_probe_update_matches_oracle_fp32(counts, rays, ni, nd, per_probe):
    u = abi.make_uniform(counts, (-1, -1, -1), (0.3, 0.4, 0.5), rays, irr_oct=ni, depth_oct=nd, hysteresis=0.97,
                         max_distance=1.5)
    rad, dd = _rays(u, seed=rays)
    rng = np.random.default_rng(1)
    prev_i = rng.uniform(0, 1, (u.irradianceTextureHeight, u.irradianceTextureWidth, 4)).astype(np.float32)
    prev_d = rng.uniform(0, 1, (u.depthTextureHeight, u.depthTextureWidth, 2)).astype(np.float32)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    if per_probe:
        flags |= abi.UPDATE_PER_PROBE
    for first in (1, 0):
        gi, gd = _run_gpu(u, rad, dd, prev_i, prev_d, first, flags)
        oi, od = oracle.probe_update(u, rad, dd, prev_i, prev_d, first, flags)
        np.testing.assert_allclose(gi, oi, rtol=2e-5, atol=1e-6)
        np.testing.assert_allclose(gd, od, rtol=2e-5, atol=1e-6)
        # global padding untouched, bit for bit
        assert np.array_equal(gi[0], prev_i[0]) and np.array_equal(gi[:, 0], prev_i[:, 0])
        assert np.array_equal(gd[-1], prev_d[-1]) and np.array_equal(gd[:, -1], prev_d[:, -1])
        # every border texel equals its interior source bit for bit
        for atlas, side in ((gi, ni), (gd, nd)):
            ref = oracle.border_update(atlas.astype(np.float64), side, *counts).astype(np.float32)
            assert np.array_equal(ref, atlas)


def test_probe_update_unshared_directions():
    u = abi.make_uniform((4, 2, 3), (-1, -1, -1), (0.3, 0.4, 0.5), 128, irr_oct=8, depth_oct=16, max_distance=1.5)
    rad, dd = _rays(u, seed=4, perturb=True)
    z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    gi, gd = _run_gpu(u, rad, dd, z_i, z_d, 1, flags)
    oi, od = oracle.probe_update(u, rad, dd, z_i, z_d, 1, flags)
    np.testing.assert_allclose(gi, oi, rtol=2e-5, atol=1e-6)
    np.testing.assert_allclose(gd, od, rtol=2e-5, atol=1e-6)


def test_probe_update_fp16_storage():
    u = abi.make_uniform((4, 4, 4), (-1, -1, -1), (0.3, 0.4, 0.5), 256, irr_oct=8, depth_oct=16, max_distance=1.5)
    rad, dd = _rays(u, seed=8)
    rad16 = rad.astype(np.float16).astype(np.float32)
    dd16 = dd.astype(np.float16).astype(np.float32)
    z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    gi, gd = _run_gpu(u, rad16, dd16, z_i, z_d, 1, flags, storage=abi.STORAGE_FP16)
    oi, od = oracle.probe_update(u, rad16, dd16, z_i, z_d, 1, flags)
    np.testing.assert_allclose(gi, oi, rtol=2 ** -11 + 2e-5, atol=1e-6)
    np.testing.assert_allclose(gd, od, rtol=2 ** -11 + 2e-5, atol=1e-6)


def test_border_update_standalone_equals_fused():
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    u = abi.make_uniform((5, 2, 3), (-1, -1, -1), (0.3, 0.4, 0.5), 64, irr_oct=8, depth_oct=16)
    rad, dd = _rays(u, seed=2)
    z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    fused_i, fused_d = _run_gpu(u, rad, dd, z_i, z_d, 1, 7)
    plain_i, plain_d = _run_gpu(u, rad, dd, z_i, z_d, 1, 3)
    t_i = torch.from_numpy(plain_i).cuda()
    t_d = torch.from_numpy(plain_d).cuda()
    atlas = abi.ProbeAtlas(ptr(t_i), ptr(t_d), abi.STORAGE_FP32, 0)
    assert _lib.lib().luxgi_ddgi_border_update(C.byref(u), C.byref(atlas), 3, stream_handle()) == 0
    torch.cuda.synchronize()
    assert np.array_equal(t_i.cpu().numpy(), fused_i)
    assert np.array_equal(t_d.cpu().numpy(), fused_d)


@pytest.mark.parametrize("counts,rays,ni,nd", CASES)
def test_probe_update_half_mfma_error_bound(counts, rays, ni, nd):
    """Reduced-precision blend (LUXGI_UPDATE_HALF: FP16 weights and ray data on the FP16 MFMA, FP32
    accumulation) against the FP64 oracle on the FP32 rays. Each product carries at most two FP16
    roundings (2 x 2^-11 relative) and the normalisation uses the sum of the rounded weights, so
    every texel stays within 1.5e-3 relative (+1e-5) and the mean relative error within 2e-4;
    border and padding rules are the same as the FP32 form's (bit for bit)."""
    u = abi.make_uniform(counts, (-1, -1, -1), (0.3, 0.4, 0.5), rays, irr_oct=ni, depth_oct=nd, hysteresis=0.97,
                         max_distance=1.5)
    rad, dd = _rays(u, seed=rays + 7)
    rng = np.random.default_rng(2)
    prev_i = rng.uniform(0, 1, (u.irradianceTextureHeight, u.irradianceTextureWidth, 4)).astype(np.float32)
    prev_d = rng.uniform(0, 1, (u.depthTextureHeight, u.depthTextureWidth, 2)).astype(np.float32)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    for first in (1, 0):
        gi, gd = _run_gpu(u, rad, dd, prev_i, prev_d, first, flags | abi.UPDATE_HALF)
        oi, od = oracle.probe_update(u, rad, dd, prev_i, prev_d, first, flags)
        for g, o in ((gi, oi), (gd, od)):
            np.testing.assert_allclose(g, o, rtol=1.5e-3, atol=1e-5)
            rel = np.abs(g - o) / (np.abs(o) + 1e-6)
            assert float(rel.mean()) <= 2e-4, float(rel.mean())
        for atlas, side in ((gi, ni), (gd, nd)):
            ref = oracle.border_update(atlas.astype(np.float64), side, *counts).astype(np.float32)
            assert np.array_equal(ref, atlas)


def test_probe_update_half_constant_inputs_exact():
    """Constant radiance and distance: the FP16 weights cancel in the normalisation, so the texels
    equal the closed forms (L/2)^(1/gamma) and (r/2, r^2/2) to FP32 rounding."""
    u = abi.make_uniform((4, 2, 2), (-1, -1, -1), (0.3, 0.4, 0.5), 256, irr_oct=8, depth_oct=16, max_distance=1.5)
    P, R = u.num_probes, 256
    dirs = oracle.ray_directions(abi.mat4_colmajor(scene.random_rotation(5)), R).astype(np.float32)
    rad = np.zeros((P, R, 4), np.float32)
    rad[..., :3] = 0.5
    dd = np.zeros((P, R, 4), np.float32)
    dd[..., :3] = dirs[None]
    dd[..., 3] = 0.76  # r = 0.75: r / maxDistance = 0.5 and its square are exact in FP16
    z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_HALF
    gi, gd = _run_gpu(u, rad, dd, z_i, z_d, 1, flags)
    n, nd = 8, 16
    irr_int = np.stack([gi[1 + (p // 8) * 10 + 1: (p // 8) * 10 + n + 2, 1 + (p % 8) * 10 + 1: (p % 8) * 10 + n + 2]
                        for p in range(P)])
    r = np.float32(np.float32(0.76) - np.float32(0.01))
    np.testing.assert_allclose(irr_int[..., :3], (0.25) ** (1 / 5.0), rtol=2e-6)
    dep_int = np.stack([gd[1 + (p // 8) * 18 + 1: (p // 8) * 18 + nd + 2, 1 + (p % 8) * 18 + 1: (p % 8) * 18 + nd + 2]
                        for p in range(P)])
    np.testing.assert_allclose(dep_int[..., 0], r / 2, rtol=2e-6)
    np.testing.assert_allclose(dep_int[..., 1], r * r / 2, rtol=2e-6)


@pytest.mark.parametrize("counts,rays,ni,nd", CASES[:3])
def test_probe_update_bf16_mfma_error_bound(counts, rays, ni, nd):
    """Reduced-precision blend with BF16 operands (LUXGI_UPDATE_HALF | LUXGI_UPDATE_BF16, BF16 MFMA,
    FP32 accumulation) against the FP64 oracle on the FP32 rays. A BF16 rounding is at most 2^-9
    relative, so each product carries at most 2 x 2^-9; normalised by the sum of the rounded weights,
    every texel stays within 1e-2 relative (+1e-4), the mean relative error within 1.5e-3 and the
    normwise error within 2e-3. Borders and padding as in the FP32 form, bit for bit."""
    u = abi.make_uniform(counts, (-1, -1, -1), (0.3, 0.4, 0.5), rays, irr_oct=ni, depth_oct=nd, hysteresis=0.97,
                         max_distance=1.5)
    rad, dd = _rays(u, seed=rays + 11)
    rng = np.random.default_rng(3)
    prev_i = rng.uniform(0, 1, (u.irradianceTextureHeight, u.irradianceTextureWidth, 4)).astype(np.float32)
    prev_d = rng.uniform(0, 1, (u.depthTextureHeight, u.depthTextureWidth, 2)).astype(np.float32)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    for first in (1, 0):
        gi, gd = _run_gpu(u, rad, dd, prev_i, prev_d, first, flags | abi.UPDATE_HALF | abi.UPDATE_BF16)
        oi, od = oracle.probe_update(u, rad, dd, prev_i, prev_d, first, flags)
        for g, o in ((gi, oi), (gd, od)):
            np.testing.assert_allclose(g, o, rtol=1e-2, atol=1e-4)
            rel = np.abs(g - o) / (np.abs(o) + 1e-6)
            assert float(rel.mean()) <= 1.5e-3, float(rel.mean())
            assert oracle.relative_error_report(g, o)["normwise_rel"] <= 2e-3
        for atlas, side in ((gi, ni), (gd, nd)):
            ref = oracle.border_update(atlas.astype(np.float64), side, *counts).astype(np.float32)
            assert np.array_equal(ref, atlas)



@pytest.mark.parametrize("storage", [abi.STORAGE_FP32, abi.STORAGE_FP16])
def test_probe_update_padded_ray_rows(storage):
    """Ray buffers whose rows are wider than the ray count (LuxgiRayBuffers.width > raysPerProbe):
    the blend reads row 0's directions (in place for FP32 rays, widened for FP16 rays) and every
    probe's rays by the row pitch, and never the padding, which holds NaN here."""
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    u = abi.make_uniform((4, 3, 2), (-1, -1, -1), (0.3, 0.4, 0.5), 96, irr_oct=8, depth_oct=16, max_distance=1.5)
    rad, dd = _rays(u, seed=11)
    if storage == abi.STORAGE_FP16:
        rad, dd = rad.astype(np.float16).astype(np.float32), dd.astype(np.float16).astype(np.float32)
    P, R = rad.shape[:2]
    width = R + 13
    dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    t_rad = torch.full((P, width, 4), float("nan"), dtype=dt, device="cuda")
    t_dd = torch.full((P, width, 4), float("nan"), dtype=dt, device="cuda")
    t_rad[:, :R] = torch.from_numpy(rad).to(dt).cuda()
    t_dd[:, :R] = torch.from_numpy(dd).to(dt).cuda()
    rays = abi.RayBuffers(ptr(t_rad), ptr(t_dd), width, P, storage, 0)
    rng = np.random.default_rng(5)
    prev_i = rng.uniform(0, 1, (u.irradianceTextureHeight, u.irradianceTextureWidth, 4)).astype(np.float32)
    prev_d = rng.uniform(0, 1, (u.depthTextureHeight, u.depthTextureWidth, 2)).astype(np.float32)
    if storage == abi.STORAGE_FP16:
        prev_i, prev_d = prev_i.astype(np.float16).astype(np.float32), prev_d.astype(np.float16).astype(np.float32)
    p_i = torch.from_numpy(prev_i).to(dt).cuda()
    p_d = torch.from_numpy(prev_d).to(dt).cuda()
    o_i, o_d = p_i.clone(), p_d.clone()
    prev = abi.ProbeAtlas(ptr(p_i), ptr(p_d), storage, 0)
    out = abi.ProbeAtlas(ptr(o_i), ptr(o_d), storage, 0)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    ws = torch.empty(_lib.lib().luxgi_ddgi_probe_update_workspace_size(C.byref(u)), dtype=torch.uint8, device="cuda")
    for mask in (flags, abi.UPDATE_IRRADIANCE | abi.UPDATE_FUSE_BORDER):  # with and without the depth map
        rc = _lib.lib().luxgi_ddgi_probe_update(C.byref(u), C.byref(rays), C.byref(prev), C.byref(out), 0, mask,
                                                ptr(ws), ws.numel(), stream_handle())
        assert rc == 0
        torch.cuda.synchronize()
        gi, gd = o_i.float().cpu().numpy(), o_d.float().cpu().numpy()
        oi, od = oracle.probe_update(u, rad, dd, prev_i, prev_d, 0, mask)
        tol = 2 ** -11 + 2e-5 if storage == abi.STORAGE_FP16 else 2e-5
        assert np.isfinite(gi).all() and np.isfinite(gd).all()
        np.testing.assert_allclose(gi, oi, rtol=tol, atol=1e-6)
        if mask & abi.UPDATE_DEPTH:
            np.testing.assert_allclose(gd, od, rtol=tol, atol=1e-6)

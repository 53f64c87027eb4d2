"""GPU parity: SVGF reprojection, a-trous and whole-frame denoise (reflection RGB
and shadow scalar) against the FP64 oracle on the same seeded G-buffers.

Tolerances. Tile maps are compared exactly. Reprojection outputs: world
positions are rebuilt from NDC depth near 1 in fp32 (Math.glsl:35-42 as the
reference computes them), which moves reprojected coordinates by ~1e-4 px
against FP64; with the seeded (uncorrelated) histories that changes the
bilinear history by up to ~1e-4, and threshold decisions (disocclusion,
sumw >= 0.01, history clip) can flip. So: >= 99.5% of pixels agree to 1e-3
relative + 1e-3 absolute in every channel and the mean absolute colour error is
<= 1e-4. A-trous outputs (no branching on inputs other than the tile map) agree
to 1e-4 relative + 1e-5 absolute on every pixel."""
import ctypes as C

import numpy as np
import pytest

import oracle
from luxgi_amd import abi
from tests import _svgf_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev(g):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in g.items()}


def _frac_close(a, b, rtol, atol):
    ok = np.all(np.abs(a - b) <= atol + rtol * np.abs(b), axis=-1)
    return ok.mean()


def _refl_params(cam):
    p = abi.ReflectionReprojectParams()
    p.viewProjInv = abi.mat4_colmajor(cam["VPI"])
    p.prevViewProj = abi.mat4_colmajor(cam["VPp"])
    p.cameraPos[:] = [*cam["eye"], 1.0]
    p.cameraDelta[:] = list(cam["delta"])
    p.alpha = 0.01
    p.momentsAlpha = 0.2
    p.approximateWithDDGI = 1
    return p


@pytest.mark.parametrize("W,H,seed", [(77, 45, 0), (128, 64, 1), (8, 8, 2)])
def test_reflection_reproject_and_atrous(W, H, seed):
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, seed)
    p = _refl_params(cam)
    dg, dp = _dev(g), _dev(prev)
    t_noisy, t_hist, t_hmom = (torch.from_numpy(x).cuda() for x in (noisy, hist, hmom))
    oc = torch.zeros((H, W, 4), device="cuda")
    om = torch.zeros((H, W, 4), device="cuda")
    tiles = torch.zeros(_lib.lib().luxgi_svgf_tile_count(W, H), dtype=torch.uint8, device="cuda")
    rc = _lib.lib().luxgi_svgf_reflection_reproject(C.byref(p), W, H, 0, C.byref(gbuffer_struct(dg)),
                                                    C.byref(gbuffer_struct(dp)), ptr(t_noisy), ptr(t_hist),
                                                    ptr(t_hmom), ptr(oc), ptr(om), ptr(tiles), stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    ref_c, ref_m, ref_t = oracle.reflection_reproject(p, W, H, g, prev, noisy, hist, hmom)
    assert np.array_equal(tiles.cpu().numpy(), ref_t)
    assert ref_t.any() and not ref_t.all() or W * H <= 64
    f_c = _frac_close(oc.cpu().numpy(), ref_c, 1e-3, 1e-3)
    f_m = _frac_close(om.cpu().numpy(), ref_m, 1e-3, 1e-3)
    assert f_c >= 0.995 and f_m >= 0.995, (f_c, f_m)
    e_c = float(np.abs(oc.cpu().numpy() - ref_c).mean())
    assert e_c <= 1e-4, e_c
    # a-trous iterations on identical inputs
    x = oc.cpu().numpy()
    for i in range(5):
        ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
        tin = torch.from_numpy(x).cuda()
        tout = torch.ones_like(tin)
        thist = torch.zeros_like(tin)
        rc = _lib.lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, 0, C.byref(gbuffer_struct(dg)), ptr(tiles),
                                                     ptr(tin), ptr(tout), ptr(thist), stream_handle())
        assert rc == 0
        torch.cuda.synchronize()
        ref = oracle.reflection_atrous(ap, W, H, g["normals"], g["pbr"], ref_t, x)
        got = tout.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5)
        assert np.array_equal(thist.cpu().numpy(), got)
        x = got


def test_reflection_denoise_frames():
    """Three frames of denoise::system: ping-pong moments, history copy after iteration 1."""
    from luxgi_amd.svgf import ReflectionDenoiser
    W, H = 96, 56
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, 3)
    dn = ReflectionDenoiser(W, H, iterations=4)
    dn.set_camera(cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])
    dg, dp = _dev(g), _dev(prev)
    o_out = [np.zeros((H, W, 4), np.float32), np.zeros((H, W, 4), np.float32)]
    o_mom = [np.zeros((H, W, 4), np.float32), np.zeros((H, W, 4), np.float32)]
    pp = 0
    rng = np.random.default_rng(7)
    for frame in range(3):
        nz = (noisy * rng.uniform(0.8, 1.2, noisy.shape)).astype(np.float32)
        got = dn.denoise(dg, dp, torch.from_numpy(nz).cuda()).cpu().numpy()
        c, m, t = oracle.reflection_reproject(dn.rp, W, H, g, prev, nz, o_out[1], o_mom[1 - pp])
        o_out[0] = c.astype(np.float32)
        o_mom[pp] = m.astype(np.float32)
        x = o_out[0]
        for i in range(4):
            ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
            x = oracle.reflection_atrous(ap, W, H, g["normals"], g["pbr"], t, x).astype(np.float32)
            if i == 1:
                o_out[1] = x.copy()
        pp = 1 - pp
        f = _frac_close(got, x, 2e-3, 2e-3)
        assert f >= 0.995, (frame, f)
        # feed the GPU state back into the oracle so that rare threshold flips do not accumulate
        o_out[1] = dn.outputs[1].cpu().numpy()
        o_mom = [dn.moments[0].cpu().numpy(), dn.moments[1].cpu().numpy()]


@pytest.mark.parametrize("W,H,seed", [(77, 45, 0), (64, 32, 5)])
def test_shadow_reproject_and_atrous(W, H, seed):
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, seed)
    hist2 = np.ascontiguousarray(hist[..., :2])
    p = abi.ShadowReprojectParams()
    p.viewProjInv = abi.mat4_colmajor(cam["VPI"])
    p.alpha = 0.01
    p.momentsAlpha = 0.2
    dg, dp = _dev(g), _dev(prev)
    t_mask = torch.from_numpy(mask.view(np.int32)).cuda()
    t_hist, t_hmom = torch.from_numpy(hist2).cuda(), torch.from_numpy(hmom).cuda()
    oc = torch.zeros((H, W, 2), device="cuda")
    om = torch.zeros((H, W, 4), device="cuda")
    tiles = torch.zeros(_lib.lib().luxgi_svgf_tile_count(W, H), dtype=torch.uint8, device="cuda")
    rc = _lib.lib().luxgi_svgf_shadow_reproject(C.byref(p), W, H, 0, C.byref(gbuffer_struct(dg)),
                                                C.byref(gbuffer_struct(dp)), ptr(t_mask), ptr(t_hist), ptr(t_hmom),
                                                ptr(oc), ptr(om), ptr(tiles), stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    ref_c, ref_m, ref_t = oracle.shadow_reproject(p, W, H, g, prev, mask, hist2, hmom)
    assert np.array_equal(tiles.cpu().numpy(), ref_t)
    f_c = _frac_close(oc.cpu().numpy(), ref_c, 1e-3, 1e-3)
    f_m = _frac_close(om.cpu().numpy(), ref_m, 1e-3, 1e-3)
    assert f_c >= 0.995 and f_m >= 0.995, (f_c, f_m)
    x = oc.cpu().numpy()
    for i in range(4):
        power = 1.2 if i == 3 else 0.0
        ap = abi.ShadowAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, power, 0.1, 1000.0)
        tin = torch.from_numpy(x).cuda()
        tout = torch.ones_like(tin)
        rc = _lib.lib().luxgi_svgf_shadow_atrous(C.byref(ap), W, H, 0, C.byref(gbuffer_struct(dg)), ptr(tiles),
                                                 ptr(tin), ptr(tout), None, stream_handle())
        assert rc == 0
        torch.cuda.synchronize()
        ref = oracle.shadow_atrous(ap, W, H, g["normals"], g["pbr"], ref_t, x)
        got = tout.cpu().numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-5)
        x = got


def test_shadow_denoise_frames():
    """Three frames of the shadow denoise::system (RaytracedShadow.cpp:543-609) through the whole-frame
    entry point: moments ping-pong, the history colour copied after a-trous iteration 1 (:532-535) and
    power applied on the last iteration only (:510), against the oracle chain."""
    from luxgi_amd.svgf import ShadowDenoiser
    W, H = 88, 52
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, 6)
    dn = ShadowDenoiser(W, H, iterations=4, far=1000.0)
    dn.set_camera(cam["VPI"])
    dg, dp = _dev(g), _dev(prev)
    o_out = [np.zeros((H, W, 2), np.float32), np.zeros((H, W, 2), np.float32)]
    o_mom = [np.zeros((H, W, 4), np.float32), np.zeros((H, W, 4), np.float32)]
    pp = 0
    rng = np.random.default_rng(11)
    for frame in range(3):
        m = (mask ^ rng.integers(0, 2 ** 32, mask.shape, dtype=np.uint64).astype(np.uint32) & 0x01010101).astype(
            np.uint32)
        got = dn.denoise(dg, dp, torch.from_numpy(m.view(np.int32)).cuda()).cpu().numpy()
        c, mo, t = oracle.shadow_reproject(dn.rp, W, H, g, prev, m, o_out[1], o_mom[1 - pp])
        o_out[0] = c.astype(np.float32)
        o_mom[pp] = mo.astype(np.float32)
        x = o_out[0]
        for i in range(4):
            ap = abi.ShadowAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1.2 if i == 3 else 0.0, 0.1, 1000.0)
            x = oracle.shadow_atrous(ap, W, H, g["normals"], g["pbr"], t, x).astype(np.float32)
            if i == 1:
                o_out[1] = x.copy()
        assert np.array_equal(dn.tile_map.cpu().numpy(), t)
        f = _frac_close(got, x, 2e-3, 2e-3)
        assert f >= 0.995, (frame, f)
        # the history copy holds iteration 1's output, before the last pass's power
        f_h = _frac_close(dn.outputs[1].cpu().numpy(), o_out[1], 2e-3, 2e-3)
        assert f_h >= 0.995, (frame, f_h)
        pp = 1 - pp
        assert dn.ping_pong == pp
        # feed the GPU state back into the oracle so that rare threshold flips do not accumulate
        o_out[1] = dn.outputs[1].cpu().numpy()
        o_mom = [dn.moments[0].cpu().numpy(), dn.moments[1].cpu().numpy()]


def test_shadow_denoise_power_only_on_last_pass():
    """With power = 1 the whole-frame output equals the chain without power; with power 1.2 it is that
    output to the power 1.2 on filtered pixels (the last pass applies pow to the filtered value)."""
    from luxgi_amd.svgf import ShadowDenoiser
    W, H = 64, 40
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, 8)
    outs = []
    for power in (1.0, 1.2):
        dn = ShadowDenoiser(W, H, iterations=3, power=power, far=1000.0)
        dn.set_camera(cam["VPI"])
        outs.append(dn.denoise(_dev(g), _dev(prev), torch.from_numpy(mask.view(np.int32)).cuda()).cpu().numpy())
        tiles = dn.tile_map.cpu().numpy()
    filt = np.zeros((H, W), bool)
    for ty in range((H + 7) // 8):
        for tx in range((W + 7) // 8):
            if tiles[ty * ((W + 7) // 8) + tx]:
                filt[ty * 8:ty * 8 + 8, tx * 8:tx * 8 + 8] = True
    sel = filt & (outs[0][..., 0] > 0)
    assert sel.sum() > 100
    np.testing.assert_allclose(outs[1][sel, 0], outs[0][sel, 0] ** 1.2, rtol=2e-5, atol=1e-6)
    np.testing.assert_array_equal(outs[1][..., 1], outs[0][..., 1])  # variance untouched by power


def test_reflection_fp16_storage_close_to_fp32():
    from luxgi_amd.svgf import ReflectionDenoiser
    W, H = 64, 40
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, 4)
    outs = []
    for st in (abi.STORAGE_FP32, abi.STORAGE_FP16):
        dn = ReflectionDenoiser(W, H, storage=st, iterations=4)
        dn.set_camera(cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])
        dt = torch.float16 if st == abi.STORAGE_FP16 else torch.float32
        outs.append(dn.denoise(_dev(g), _dev(prev), torch.from_numpy(noisy).to(dt).cuda()).float().cpu().numpy())
    f = _frac_close(outs[1], outs[0], 4e-3, 4e-3)
    assert f >= 0.99, f


@pytest.mark.parametrize("radius,step", [(0, 1), (1, 8), (2, 1), (2, 4), (2, 8), (1, 16)])
def test_atrous_footprints_lds_and_direct(radius, step):
    """Footprints within the 8-pixel LDS halo (radius * step <= 8) and beyond it (direct form):
    reflection and shadow a-trous against the oracle, on a tile map with copy and filter tiles."""
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    W, H = 77, 45
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, 5)
    dg = _dev(g)
    rng = np.random.default_rng(step * 10 + radius)
    tiles = (rng.random(_lib.lib().luxgi_svgf_tile_count(W, H)) < 0.7).astype(np.uint8)
    ttiles = torch.from_numpy(tiles).cuda()
    x4 = rng.uniform(0, 1.5, (H, W, 4)).astype(np.float32)
    ap = abi.ReflectionAtrousParams(radius, step, 10.0, 32.0, 1.0, 1)
    tin, tout = torch.from_numpy(x4).cuda(), torch.zeros((H, W, 4), device="cuda")
    assert _lib.lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, 0, C.byref(gbuffer_struct(dg)), ptr(ttiles),
                                                   ptr(tin), ptr(tout), None, stream_handle()) == 0
    torch.cuda.synchronize()
    ref = oracle.reflection_atrous(ap, W, H, g["normals"], g["pbr"], tiles, x4)
    np.testing.assert_allclose(tout.cpu().numpy(), ref, rtol=1e-4, atol=1e-5)
    x2 = np.ascontiguousarray(x4[..., :2] / 1.5)
    sp = abi.ShadowAtrousParams(radius, step, 10.0, 32.0, 1.0, 1.2, 0.1, 1000.0)
    sin, sout = torch.from_numpy(x2).cuda(), torch.ones((H, W, 2), device="cuda")
    assert _lib.lib().luxgi_svgf_shadow_atrous(C.byref(sp), W, H, 0, C.byref(gbuffer_struct(dg)), ptr(ttiles),
                                               ptr(sin), ptr(sout), None, stream_handle()) == 0
    torch.cuda.synchronize()
    sref = oracle.shadow_atrous(sp, W, H, g["normals"], g["pbr"], tiles, x2)
    np.testing.assert_allclose(sout.cpu().numpy(), sref, rtol=1e-4, atol=1e-5)


def test_atrous_rejects_bad_steps():
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    W, H = 16, 16
    g, *_ = _svgf_inputs.make_pair(W, H, 0)
    dg = _dev(g)
    t = torch.zeros((H, W, 4), device="cuda")
    tiles = torch.ones(4, dtype=torch.uint8, device="cuda")
    for step in (0, -2, 1 << 20):
        ap = abi.ReflectionAtrousParams(1, step, 10.0, 32.0, 1.0, 1)
        assert _lib.lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, 0, C.byref(gbuffer_struct(dg)), ptr(tiles),
                                                       ptr(t), ptr(t), None, stream_handle()) != 0

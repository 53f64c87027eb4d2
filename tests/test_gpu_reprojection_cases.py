"""GPU parity of the temporal reprojections on the closed-form cases that pin the oracle
(tests/test_oracle_reprojection.py): 17x17 clip of a ramp and of a checkerboard, integer and
fractional pans, the virtual-point parallax, the 3x3 fallback, a moving camera, and the shadow
half-plane / constant / panned masks. The oracle's values there are fixed by hand arithmetic,
so these runs pin the GPU kernels to the reference's rules on the paths the seeded-scene
tests reach only by chance.

Tolerance: the inputs are exact in fp32 (power-of-two sizes, dyadic motions), so the GPU
reprojects onto the same pixels; every pixel agrees within 2e-5 absolute (fp32 rounding of
the bilinear weights, the clip and the mixes), tile maps exactly."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.test_oracle_reprojection import (_checker, _gbuf, _hmom, _mask_from_bits, _refl, _shadow)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _refl_cases():
    rng = np.random.default_rng(21)
    out = []
    # ramp + clip
    W, H = 64, 40
    g = _gbuf(W, H)
    n = np.zeros((H, W, 4), np.float32)
    n[..., 0] = 0.25 + np.arange(W) / 64.0
    n[..., 1:3] = 0.5
    h = np.zeros((H, W, 4), np.float32)
    h[..., 0] = 4.0
    h[..., 1:3] = 0.5
    out.append(("ramp", _refl(), W, H, g, g, n, h, _hmom(W, H)))
    # checkerboard + clip
    W = H = 48
    g = _gbuf(W, H)
    n = np.zeros((H, W, 4), np.float32)
    n[..., 1] = _checker(W, H)
    n[..., 0] = n[..., 2] = 0.25
    h = np.zeros((H, W, 4), np.float32)
    h[..., 1] = -3.0
    h[..., 0] = h[..., 2] = 0.25
    out.append(("checker", _refl(alpha=0.05), W, H, g, g, n, h, _hmom(W, H, length=9.0)))
    # integer and fractional pans over random / affine histories
    for f in (3.0, 5.25, -2.5):
        W, H = 64, 32
        g = _gbuf(W, H, motion=(f / W, 0.5 / H))
        prev = _gbuf(W, H)
        n = np.zeros((H, W, 4), np.float32)
        n[..., :3] = _checker(W, H)[..., None]
        h = np.zeros((H, W, 4), np.float32)
        h[..., :3] = rng.uniform(0.1, 0.9, (H, W, 3))
        hm = _hmom(W, H)
        hm[..., 2] = (np.arange(W) % 5)[None, :] + 1.0
        hm[..., 0] = rng.uniform(0.2, 0.4, (H, W))
        out.append((f"pan{f}", _refl(alpha=0.0, moments_alpha=0.0), W, H, g, prev, n, h, hm))
    # virtual point
    W = H = 64
    g = _gbuf(W, H, depth=-0.5, curvature=0.0)
    n = np.zeros((H, W, 4), np.float32)
    n[..., :3] = _checker(W, H)[..., None]
    n[..., 3] = 0.625
    h = np.zeros((H, W, 4), np.float32)
    h[..., :3] = rng.uniform(0.1, 0.9, (H, W, 3))
    out.append(("virtual", _refl(alpha=0.0, moments_alpha=0.0, cam=(0.0, 0.0, -1.0)), W, H, g, g, n, h,
                _hmom(W, H, length=1.0)))
    # 3x3 fallback
    W = H = 32
    g = _gbuf(W, H)
    prev = _gbuf(W, H)
    prev["color"][14:16, 17:19, 3] = 2.0
    n = np.zeros((H, W, 4), np.float32)
    n[..., :3] = _checker(W, H)[..., None]
    h = np.zeros((H, W, 4), np.float32)
    h[..., :3] = rng.uniform(0.3, 0.7, (H, W, 3))
    hm = _hmom(W, H, length=2.0)
    hm[14, 17, 2] = 6.0
    out.append(("fallback", _refl(alpha=0.0, moments_alpha=0.0), W, H, g, prev, n, h, hm))
    # moving camera
    W = H = 24
    g = _gbuf(W, H)
    n = np.zeros((H, W, 4), np.float32)
    n[..., :3] = 0.5
    h = np.zeros((H, W, 4), np.float32)
    h[..., :3] = 0.5
    h[..., 0] = 0.5005
    out.append(("camdelta", _refl(alpha=0.01, delta=(0.0, 0.0, 0.1)), W, H, g, g, n, h, _hmom(W, H, length=20.0)))
    return out


def _shadow_cases():
    out = []
    W, H = 64, 32
    bits = np.zeros((H, W), np.uint8)
    bits[:, :32] = 1
    g = _gbuf(W, H)
    for hv in (1.0, 0.0):
        h = np.zeros((H, W, 2), np.float32)
        h[..., 0] = hv
        out.append((f"half{hv}", _shadow(), W, H, g, g, bits, h, _hmom(W, H, length=3.0)))
    W, H = 32, 24
    g = _gbuf(W, H)
    for val in (0, 1):
        out.append((f"const{val}", _shadow(alpha=0.1), W, H, g, g, np.full((H, W), val, np.uint8),
                    np.full((H, W, 2), 0.9, np.float32), _hmom(W, H, length=1.0)))
    W, H = 64, 16
    bits = _checker(W, H).astype(np.uint8)
    h = np.zeros((H, W, 2), np.float32)
    h[..., 0] = 0.3 + np.arange(W) / 256.0
    hm = _hmom(W, H)
    hm[..., 2] = (np.arange(W) % 7)[None, :] + 1.0
    for f in (2.75, 2.25):
        out.append((f"pan{f}", _shadow(), W, H, _gbuf(W, H, motion=(f / W, 0.0)), _gbuf(W, H), bits, h, hm))
    return out


def _dev(g):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in g.items()}


@pytest.mark.parametrize("case", range(8))
def test_reflection_reproject_closed_form_cases(case):
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    name, p, W, H, g, prev, noisy, hist, hmom = _refl_cases()[case]
    print(f"[reflection case {name} {W}x{H}]", flush=True)
    dg, dp = _dev(g), _dev(prev)
    t = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (noisy, hist, hmom)]
    oc = torch.zeros((H, W, 4), device="cuda")
    om = torch.zeros((H, W, 4), device="cuda")
    tiles = torch.zeros(_lib.lib().luxgi_svgf_tile_count(W, H), dtype=torch.uint8, device="cuda")
    rc = _lib.lib().luxgi_svgf_reflection_reproject(C.byref(p), W, H, 0, C.byref(gbuffer_struct(dg)),
                                                    C.byref(gbuffer_struct(dp)), ptr(t[0]), ptr(t[1]), ptr(t[2]),
                                                    ptr(oc), ptr(om), ptr(tiles), stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    rc_, rm_, rt_ = oracle.reflection_reproject(p, W, H, g, prev, noisy, hist, hmom)
    assert np.array_equal(tiles.cpu().numpy(), rt_)
    np.testing.assert_allclose(oc.cpu().numpy(), rc_, rtol=0, atol=2e-5, err_msg=name)
    np.testing.assert_allclose(om.cpu().numpy(), rm_, rtol=0, atol=2e-5, err_msg=name)


@pytest.mark.parametrize("case", range(6))
def test_shadow_reproject_closed_form_cases(case):
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    name, p, W, H, g, prev, bits, hist, hmom = _shadow_cases()[case]
    print(f"[shadow case {name} {W}x{H}]", flush=True)
    mask = _mask_from_bits(bits)
    dg, dp = _dev(g), _dev(prev)
    t_mask = torch.from_numpy(mask.view(np.int32)).cuda()
    t_hist, t_hmom = torch.from_numpy(hist).cuda(), torch.from_numpy(hmom).cuda()
    oc = torch.zeros((H, W, 2), device="cuda")
    om = torch.zeros((H, W, 4), device="cuda")
    tiles = torch.zeros(_lib.lib().luxgi_svgf_tile_count(W, H), dtype=torch.uint8, device="cuda")
    rc = _lib.lib().luxgi_svgf_shadow_reproject(C.byref(p), W, H, 0, C.byref(gbuffer_struct(dg)),
                                                C.byref(gbuffer_struct(dp)), ptr(t_mask), ptr(t_hist), ptr(t_hmom),
                                                ptr(oc), ptr(om), ptr(tiles), stream_handle())
    assert rc == 0
    torch.cuda.synchronize()
    rc_, rm_, rt_ = oracle.shadow_reproject(p, W, H, g, prev, mask, hist, hmom)
    assert np.array_equal(tiles.cpu().numpy(), rt_)
    np.testing.assert_allclose(oc.cpu().numpy(), rc_, rtol=0, atol=2e-5, err_msg=name)
    np.testing.assert_allclose(om.cpu().numpy(), rm_, rtol=0, atol=2e-5, err_msg=name)

"""Pins of the FP64 oracle of the learned a-trous pass and its backward (CPU).

- With the hand-set parameters of luxgi_amd.learned.reference_like_weights(live_units=False) and
  normals that agree everywhere, the learned pass is the reference's reflection a-trous pass
  (oracle.reflection_atrous, Reflection/DenoiseAtrous.comp:63-134) up to the e^-16 weight floor.
- A constant weight c reduces each filtered pixel to (c_p + c sum c_q) / (1 + c n) (variance
  (v_p + c^2 sum v_q) / (1 + c n)^2), computed here with array shifts.
- Sky pixels are 0 and out-of-band roughness passes through.
- Both gradients against central finite differences of the oracle's own forward: the input
  gradient perturbs the blended colour with the guide held (the forward is linear in it), the
  parameter gradient perturbs each of the 81 parameters.
- Clamped taps (o > 8) carry no parameter gradient."""
import math

import numpy as np
import pytest

import oracle
from luxgi_amd import abi
from tests import _svgf_inputs


def _inputs(W=20, H=14, seed=0, same_normals=False):
    g, _, noisy, _, _, _, _ = _svgf_inputs.make_pair(W, H, seed)
    rng = np.random.default_rng(seed + 100)
    img = noisy.copy()
    img[..., 3] = rng.uniform(0.0, 0.3, (H, W))  # variance
    pbr = g["pbr"].copy()
    pbr[..., 1] = np.where(rng.random((H, W)) < 0.8, rng.uniform(0.1, 0.4, (H, W)), pbr[..., 1])
    sky = rng.random((H, W)) < 0.05
    pbr[sky, 3] = 1.0
    nrm = g["normals"].copy()
    if same_normals:
        nrm[..., :2] = _svgf_inputs.dir_to_oct(np.array([0.3, 0.8, -0.52]) / np.linalg.norm([0.3, 0.8, -0.52]))
    return nrm, pbr, img


def _params(radius=1, step=1, phi=10.0, sigma=1.0, level=0.0):
    p = abi.LearnedAtrousParams()
    p.radius, p.stepSize, p.phiColor, p.sigmaDepth, p.level = radius, step, phi, sigma, level
    return p


def _random_weights(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, scale, 81)
    w[60:70] = np.abs(w[60:70]) + 0.2  # b1 > 0: live units
    w[80] = -1.0
    return w


def _mode(pbr):
    sky = pbr[..., 3] == 1.0
    r = pbr[..., 1]
    passthrough = ~sky & ((r <= np.float32(0.05)) | (r >= np.float32(0.45)))
    return sky, passthrough, ~sky & ~passthrough


@pytest.mark.parametrize("step", [1, 2, 3])
def test_reduces_to_reference_atrous_when_normals_agree(step):
    from luxgi_amd.learned import reference_like_weights
    W, H = 24, 18
    nrm, pbr, img = _inputs(W, H, seed=step, same_normals=True)
    wts = reference_like_weights(live_units=False)
    got, _ = oracle.learned_atrous_forward(_params(1, step, 10.0, 0.7), wts, nrm, pbr, img)
    rp = abi.ReflectionAtrousParams(1, step, 10.0, 32.0, 0.7, 1)
    want = oracle.reflection_atrous(rp, W, H, nrm, pbr, np.ones(((H + 7) // 8) * ((W + 7) // 8), np.uint8), img)
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)


def test_constant_weight_is_a_box_filter():
    W, H = 17, 13
    nrm, pbr, img = _inputs(W, H, seed=4)
    c = 0.37
    wts = np.zeros(81)
    wts[80] = math.log(c)
    step = 2
    got, sw = oracle.learned_atrous_forward(_params(2, step), wts, nrm, pbr, img)
    acc = np.zeros((H, W, 4))
    n = np.zeros((H, W))
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            if dx == 0 and dy == 0:
                continue
            ys, xs = np.mgrid[0:H, 0:W]
            qy, qx = ys + dy * step, xs + dx * step
            ok = (qy >= 0) & (qy < H) & (qx >= 0) & (qx < W)
            q = img[np.clip(qy, 0, H - 1), np.clip(qx, 0, W - 1)].astype(np.float64)
            acc += np.where(ok[..., None], q, 0.0)
            n += ok
    sky, pt, filt = _mode(pbr)
    want = np.empty((H, W, 4))
    want[..., :3] = (img[..., :3] + c * acc[..., :3]) / (1 + c * n)[..., None]
    want[..., 3] = (img[..., 3] + c * c * acc[..., 3]) / (1 + c * n) ** 2
    np.testing.assert_allclose(got[filt], want[filt], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(sw[filt], (1 + c * n)[filt], rtol=1e-12)
    assert np.all(got[sky] == 0.0)
    np.testing.assert_array_equal(got[pt], img[pt].astype(np.float64))
    assert filt.sum() > 100 and pt.sum() > 5 and sky.sum() > 3


@pytest.mark.parametrize("radius,step,seed", [(1, 1, 0), (2, 2, 1)])
def test_input_gradient_matches_finite_differences(radius, step, seed):
    W, H = 16, 12
    nrm, pbr, img = _inputs(W, H, seed=seed)
    wts = _random_weights(seed)
    p = _params(radius, step, level=0.5)
    rng = np.random.default_rng(seed + 7)
    gout = rng.normal(0.0, 1.0, (H, W, 4))
    gout[..., 3] = 0.0
    gin, _ = oracle.learned_atrous_backward(p, wts, nrm, pbr, img, gout)
    for _ in range(40):
        y, x, k = int(rng.integers(0, H)), int(rng.integers(0, W)), int(rng.integers(0, 3))
        up, dn = img.copy(), img.copy()
        up[y, x, k] += np.float32(0.01)
        dn[y, x, k] -= np.float32(0.01)
        h = float(up[y, x, k]) - float(dn[y, x, k])
        lu = (oracle.learned_atrous_forward(p, wts, nrm, pbr, img, color=up)[0] * gout).sum()
        ld = (oracle.learned_atrous_forward(p, wts, nrm, pbr, img, color=dn)[0] * gout).sum()
        assert abs((lu - ld) / h - gin[y, x, k]) <= 1e-9 * (1 + abs(gin[y, x, k])), (y, x, k)
    assert np.all(gin[..., 3] == 0.0)


@pytest.mark.parametrize("radius,step,seed", [(1, 1, 2), (2, 1, 3)])
def test_parameter_gradient_matches_finite_differences(radius, step, seed):
    W, H = 14, 10
    nrm, pbr, img = _inputs(W, H, seed=seed)
    wts = _random_weights(seed)
    p = _params(radius, step, level=1.0)
    gout = np.random.default_rng(seed + 9).normal(0.0, 1.0, (H, W, 4))
    gout[..., 3] = 0.0
    _, gp = oracle.learned_atrous_backward(p, wts, nrm, pbr, img, gout)
    eps = 1e-6
    for i in range(81):
        up, dn = wts.copy(), wts.copy()
        up[i] += eps
        dn[i] -= eps
        lu = (oracle.learned_atrous_forward(p, up, nrm, pbr, img)[0] * gout).sum()
        ld = (oracle.learned_atrous_forward(p, dn, nrm, pbr, img)[0] * gout).sum()
        fd = (lu - ld) / (2 * eps)
        assert abs(fd - gp[i]) <= 1e-6 * (1 + abs(gp[i])), (i, fd, gp[i])
    assert np.abs(gp).max() > 1e-3


def test_clamped_taps_carry_no_parameter_gradient():
    W, H = 12, 9
    nrm, pbr, img = _inputs(W, H, seed=5)
    wts = np.zeros(81)
    wts[80] = 20.0  # o = 20 > 8 on every tap: weight e^8
    gout = np.ones((H, W, 4))
    gin, gp = oracle.learned_atrous_backward(_params(1, 1), wts, nrm, pbr, img, gout)
    assert np.all(gp == 0.0)
    out, sw = oracle.learned_atrous_forward(_params(1, 1), wts, nrm, pbr, img)
    _, _, filt = _mode(pbr)
    assert np.all(sw[filt] >= 1 + 3 * math.exp(8.0) - 1e-6)

"""Pins of the FP64 oracle of the per-pixel learned edge-stopping scales (CPU).

The network (12 -> 32 -> 32 -> 5, ReLU) maps per-pixel features to one scale per a-trous pass,
s_i = exp(clamp(o_i, -4, 4)), which multiplies the luminance term of the reference's
edge-stopping weight (EdgeStop.glsl:24-40) in the reflection a-trous pass
(Reflection/DenoiseAtrous.comp:63-134).

- s = 1 is the reference pass, bit for bit (oracle.reflection_atrous).
- A scale gathers: s -> infinity leaves only the taps of equal luminance, s -> 0 drops the
  luminance term (closed forms on a two-tap row).
- The features: closed forms on a depth ramp, a constant image, known normals, moments and
  history lengths.
- The MLP: hand-set single paths through named weights pin the packing (W1 [32][12] row-major,
  then b1, W2 [32][32], b2, W3 [5][32], b3), the ReLUs and the output clamp.
- The backward: the scale gradient against central differences of the scaled forward; the input
  gradient against central differences along luminance-preserving colour directions (along which
  the frozen weights are exact); the MLP's parameter gradient against central differences, with
  and without FP16 operand rounding."""
import numpy as np
import pytest

import oracle
from luxgi_amd import abi
from tests import _svgf_inputs

NP = 1637


def _inputs(W=20, H=14, seed=0):
    g, _, noisy, _, hmom, _, _ = _svgf_inputs.make_pair(W, H, seed)
    rng = np.random.default_rng(seed + 100)
    img = noisy.copy()
    img[..., 3] = rng.uniform(0.0, 0.3, (H, W))  # variance channel
    pbr = g["pbr"].copy()
    pbr[..., 1] = np.where(rng.random((H, W)) < 0.8, rng.uniform(0.1, 0.4, (H, W)), pbr[..., 1])
    sky = rng.random((H, W)) < 0.05
    pbr[sky, 3] = 1.0
    tiles = (rng.random(((H + 7) // 8) * ((W + 7) // 8)) < 0.8).astype(np.uint8)
    return g["normals"], pbr, img, hmom, tiles


def _ap(step=1):
    return abi.ReflectionAtrousParams(1, step, 10.0, 32.0, 1.0, 1)


def _random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, NP)
    p[384:416] = np.abs(p[384:416]) + 0.1  # b1 > 0: live units
    p[1440:1472] = np.abs(p[1440:1472]) + 0.1  # b2 > 0
    return p


def test_param_count_and_layout():
    assert oracle.learned_pixel_param_count() == NP == 32 * 12 + 32 + 32 * 32 + 32 + 5 * 32 + 5


@pytest.mark.parametrize("step", [1, 2, 4])
def test_unit_scale_is_the_reference_pass(step):
    nrm, pbr, img, _, tiles = _inputs(seed=step)
    H, W = img.shape[:2]
    ref = oracle.reflection_atrous(_ap(step), W, H, nrm, pbr, tiles, img)
    a = oracle.reflection_atrous_scaled(_ap(step), W, H, nrm, pbr, tiles, img, None)
    b = oracle.reflection_atrous_scaled(_ap(step), W, H, nrm, pbr, tiles, img, np.ones((H, W), np.float32))
    assert np.array_equal(a, ref) and np.array_equal(b, ref)


def test_scale_limits_on_a_two_tap_row():
    """A 3x1 image, one filtered centre (x = 1) whose two taps have luminance equal to and far from
    the centre's: a huge scale keeps only the equal tap, a zero scale weighs both by the depth and
    normal terms alone (exp(-exp(0)) with equal depths and normals)."""
    W, H = 3, 1
    nrm = np.zeros((H, W, 4), np.float32)
    nrm[..., :2] = _svgf_inputs.dir_to_oct(np.array([[[0.0, 0.0, 1.0]] * W]))
    pbr = np.zeros((H, W, 4), np.float32)
    pbr[..., 1], pbr[..., 3] = 0.2, 5.0
    img = np.zeros((H, W, 4), np.float32)
    img[0, 0, :3] = [0.5, 0.5, 0.5]  # luminance 0.5, as the centre
    img[0, 1, :3] = [0.5, 0.5, 0.5]
    img[0, 2, :3] = [2.0, 2.0, 2.0]
    img[..., 3] = 0.04
    tiles = np.ones(1, np.uint8)
    k = 2.0 / 3.0  # kw[1] * kw[0]
    w0 = np.exp(-np.exp(0.0))  # equal depth: wZ = exp(0); wN = 1
    for s, want_far in ((1e6, 0.0), (1e-12, w0 * k)):
        sc = np.full((H, W), s, np.float32)
        out = oracle.reflection_atrous_scaled(_ap(), W, H, nrm, pbr, tiles, img, sc)
        sw = 1.0 + w0 * k + want_far
        want = (0.5 + w0 * k * 0.5 + want_far * 2.0) / sw
        assert out[0, 1, 0] == pytest.approx(want, rel=1e-9)


def test_features_closed_forms():
    W, H = 9, 7
    color = np.zeros((H, W, 4), np.float32)
    color[..., :3] = 0.4  # constant: luminance 0.4
    color[..., 3] = 0.09  # variance
    moments = np.zeros((H, W, 4), np.float32)
    moments[..., 0], moments[..., 1], moments[..., 2] = 0.5, 0.34, 12.0  # std sqrt(0.09) = 0.3
    n = np.array([0.36, -0.48, 0.8])
    normals = np.zeros((H, W, 4), np.float32)
    normals[..., :2] = _svgf_inputs.dir_to_oct(np.broadcast_to(n, (H, W, 3)))
    pbr = np.zeros((H, W, 4), np.float32)
    ys, xs = np.mgrid[0:H, 0:W]
    pbr[..., 3] = 2.0 + 0.25 * xs + 0.5 * ys  # depth ramp
    f = oracle.learned_pixel_features(color, moments, normals, pbr)
    x, y = 4, 3
    z = 2.0 + 0.25 * x + 0.5 * y
    l = 0.299 * 0.4 + 0.587 * 0.4 + 0.114 * 0.4
    assert f[y, x, 0] == pytest.approx(0.25 / z, rel=1e-6)
    assert f[y, x, 1] == pytest.approx(0.5 / z, rel=1e-6)
    assert f[y, x, 0:2].tolist() != f[y, 0, 0:2].tolist()  # edge-clamped neighbours: half the step at x = 0
    assert f[y, 0, 0] == pytest.approx(0.125 / pbr[y, 0, 3], rel=1e-6)
    assert np.allclose(f[y, x, 2:5], n, atol=1e-6)
    assert f[y, x, 5] == pytest.approx(np.log1p(l), rel=1e-6)
    assert f[y, x, 6] == pytest.approx(np.log1p(l), rel=1e-6)
    assert abs(f[y, x, 7]) < 1e-6  # constant neighbourhood: zero std
    assert f[y, x, 8] == pytest.approx(0.3 / (l + 1e-3), rel=1e-5)
    assert f[y, x, 9] == pytest.approx(np.log1p(0.5), rel=1e-6)
    assert f[y, x, 10] == pytest.approx(0.3 / (0.5 + 1e-3), rel=1e-5)
    assert f[y, x, 11] == pytest.approx(12.0 / 32.0)
    # a 3x3 checker of two luminances: std / mean of the nine values
    color[..., :3] = np.where(((xs + ys) % 2 == 0)[..., None], 0.2, 0.8)
    f = oracle.learned_pixel_features(color, moments, normals, pbr)
    vals = np.array([0.2 if (xx + yy) % 2 == 0 else 0.8 for yy in (2, 3, 4) for xx in (3, 4, 5)])
    mu, sd = vals.mean(), vals.std()
    assert f[3, 4, 6] == pytest.approx(np.log1p(mu), rel=1e-6)
    assert f[3, 4, 7] == pytest.approx(sd / (mu + 1e-3), rel=1e-6)


def test_mlp_packing_relu_and_clamp_by_single_paths():
    """o_3 = W3[3][7] relu(W2[7][20] relu(W1[20][5] f_5 + b1[20]) + b2[7]) + b3[3] with every other
    weight 0, and the clamp at +-4."""
    f = np.zeros((2, 12))
    f[0, 5], f[1, 5] = 0.7, -0.7
    p = np.zeros(NP)
    p[20 * 12 + 5] = 2.0           # W1[20][5]
    p[384 + 20] = 0.1              # b1[20]
    p[416 + 7 * 32 + 20] = 1.5     # W2[7][20]
    p[1440 + 7] = -0.05            # b2[7]
    p[1472 + 3 * 32 + 7] = 0.8     # W3[3][7]
    p[1632 + 3] = 0.25             # b3[3]
    p[1632 + 1] = 9.0              # b3[1]: clamped to 4
    o, s = oracle.learned_pixel_mlp(p, f)
    h1 = np.maximum(2.0 * f[:, 5] + 0.1, 0.0)
    h2 = np.maximum(1.5 * h1 - 0.05, 0.0)
    assert np.allclose(o[:, 3], 0.8 * h2 + 0.25, rtol=1e-12)
    assert np.allclose(o[:, 1], 9.0) and np.allclose(s[:, 1], np.exp(4.0))
    assert np.allclose(o[:, [0, 2, 4]], 0.0) and np.allclose(s[:, [0, 2, 4]], 1.0)
    assert h1[1] == 0.0 and o[1, 3] == pytest.approx(0.25)  # the negative input dies at the first ReLU
    # the transposed W2 entry ([20][7]) carries nothing along this path
    q = p.copy()
    q[416 + 7 * 32 + 20], q[416 + 20 * 32 + 7] = 0.0, 1.5
    o2, _ = oracle.learned_pixel_mlp(q, f)
    assert np.allclose(o2[:, 3], 0.8 * max(-0.05, 0.0) + 0.25)


def test_mlp_half_rounds_the_operands():
    p = np.zeros(NP)
    p[0] = 0.1                      # W1[0][0], not representable in FP16
    p[416] = 1.0                    # W2[0][0]
    p[1472] = 1.0                   # W3[0][0]
    f = np.zeros((1, 12))
    f[0, 0] = 1.0
    o, _ = oracle.learned_pixel_mlp(p, f, half=False)
    oh, _ = oracle.learned_pixel_mlp(p, f, half=True)
    assert o[0, 0] == pytest.approx(0.1, rel=1e-12)
    assert oh[0, 0] == pytest.approx(float(np.float16(0.1)), rel=1e-12)


@pytest.mark.parametrize("step", [1, 2])
def test_scale_gradient_matches_central_differences(step):
    nrm, pbr, img, _, tiles = _inputs(seed=10 + step)
    H, W = img.shape[:2]
    rng = np.random.default_rng(step)
    sc = np.exp(rng.uniform(-1.0, 1.0, (H, W))).astype(np.float32)
    g = rng.normal(size=(H, W, 4))
    g[..., 3] = 0.0
    _, gs = oracle.learned_pixel_atrous_backward(_ap(step), W, H, nrm, pbr, tiles, img, sc, g)
    assert np.abs(gs).max() > 1e-3
    for (y, x) in [(3, 4), (7, 11), (10, 2), (5, 17)]:
        h = 1e-3 * sc[y, x]
        sp, sm = sc.astype(np.float64).copy(), sc.astype(np.float64).copy()
        sp[y, x] += h
        sm[y, x] -= h
        # the oracle reads the scale as FP32: use an FP64 step on a value that survives the cast
        lp = (oracle.reflection_atrous_scaled(_ap(step), W, H, nrm, pbr, tiles, img, sp.astype(np.float32)) * g).sum()
        lm = (oracle.reflection_atrous_scaled(_ap(step), W, H, nrm, pbr, tiles, img, sm.astype(np.float32)) * g).sum()
        dh = float(np.float32(sp[y, x])) - float(np.float32(sm[y, x]))
        assert gs[y, x] == pytest.approx((lp - lm) / dh, rel=2e-3, abs=1e-7), (y, x)


@pytest.mark.parametrize("step", [1, 3])
def test_input_gradient_matches_central_differences_along_luminance_preserving_directions(step):
    nrm, pbr, img, _, tiles = _inputs(seed=20 + step)
    H, W = img.shape[:2]
    rng = np.random.default_rng(step + 7)
    sc = np.exp(rng.uniform(-1.0, 1.0, (H, W))).astype(np.float32)
    g = rng.normal(size=(H, W, 4))
    g[..., 3] = 0.0
    gin, _ = oracle.learned_pixel_atrous_backward(_ap(step), W, H, nrm, pbr, tiles, img, sc, g)
    assert np.all(gin[..., 3] == 0.0)
    d = np.array([0.587, -0.299, 0.0])  # luminance 0.299 r + 0.587 g + 0.114 b unchanged
    d2 = np.array([0.114, 0.0, -0.299])
    for (y, x) in [(3, 4), (8, 9), (12, 15), (0, 0)]:
        for dv in (d, d2):
            h = 1e-2
            ip, im = img.astype(np.float64).copy(), img.astype(np.float64).copy()
            ip[y, x, :3] += h * dv
            im[y, x, :3] -= h * dv
            lp = (oracle.reflection_atrous_scaled(_ap(step), W, H, nrm, pbr, tiles, ip, sc) * g).sum()
            lm = (oracle.reflection_atrous_scaled(_ap(step), W, H, nrm, pbr, tiles, im, sc) * g).sum()
            # the forward reads FP32: the effective step is the cast difference
            dh = (ip[y, x, :3].astype(np.float32).astype(np.float64) - im[y, x, :3].astype(np.float32))
            want = (lp - lm)
            got = float(gin[y, x, :3] @ dh)
            assert got == pytest.approx(want, rel=1e-4, abs=1e-8), (y, x, dv)


@pytest.mark.parametrize("half", [False, True])
def test_mlp_parameter_gradient_matches_central_differences(half):
    rng = np.random.default_rng(4)
    p = _random_params(5)
    f = rng.uniform(-1.0, 1.0, (6, 12))
    gS = rng.normal(size=(6, 5))
    gp = oracle.learned_pixel_mlp_backward(p, f, gS, half=half)
    idx = list(range(0, 384, 37)) + list(range(384, 416, 7)) + list(range(416, 1440, 101)) + \
        list(range(1440, 1472, 9)) + list(range(1472, 1632, 17)) + list(range(1632, 1637))
    if half:
        # the gradient of the rounded network taken straight through the roundings: central
        # differences see it only where no rounding follows (the output biases); elsewhere it
        # stays within FP16 operand precision of the unrounded gradient
        g64 = oracle.learned_pixel_mlp_backward(p, f, gS, half=False)
        assert np.allclose(gp, g64, rtol=2e-2, atol=2e-2 * np.abs(g64).max())
        idx = list(range(1632, 1637))
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = (oracle.learned_pixel_mlp(pp, f, half=half)[1] * gS).sum()
        lm = (oracle.learned_pixel_mlp(pm, f, half=half)[1] * gS).sum()
        assert gp[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-9), i

"""Pins the oracle's SVGF passes and the probe relocation rule with closed forms.

SVGF (Reflection/DenoiseReprojection.comp, Reflection/DenoiseAtrous.comp,
Shadow/DenoiseAtrous.comp, Raytraced/EdgeStop.glsl, Raytraced/Reprojection.glsl):
  * a static camera over an unchanged G-buffer reprojects every pixel onto itself,
    so the history is clipped to the neighbourhood box (extent sd + 0.001) and mixed
    with alpha = max(alpha, 1 / histLen) -- values computed by hand below;
  * a disoccluded frame (every previous meshId differs) restarts: alpha = 1, histLen = 1;
  * on a flat plane the edge-stop weight is exp(-|dl|/phi - exp(-|dz|/sigma)) * max(0, n.n')^phi_n,
    i.e. exp(-1) between equal depths and ~1 across a large depth step (the reference's
    sign, kept as is): an impulse spreads over the 5x5 B3-spline lattice with those weights,
    and the variance channel with their squares;
  * copy tiles pass colour through (reflection) or write 0 (shadow).
Relocation: hand-built ray buffers with known classification and push."""
import math

import numpy as np
import pytest

import oracle
from luxgi_amd import abi

E1 = math.exp(-1.0)
IDENT = np.eye(4, dtype=np.float32)


def _plane_gbuffer(W, H, depth=0.5, rough=0.2, mesh=1.0):
    g = {"color": np.zeros((H, W, 4), np.float32), "normals": np.zeros((H, W, 4), np.float32),
         "pbr": np.zeros((H, W, 4), np.float32), "depth": np.full((H, W), depth, np.float32)}
    g["color"][..., 3] = mesh          # meshId
    g["pbr"][..., 1] = rough           # roughness
    g["pbr"][..., 2] = 0.1             # curvature != 0: plain motion-vector reprojection
    g["pbr"][..., 3] = depth           # depth compared by the edge-stop
    return g                           # normals: octahedral (0, 0) = +z, motion 0


def _refl_params(alpha=0.1, moments_alpha=0.2):
    p = abi.ReflectionReprojectParams()
    p.viewProjInv = abi.mat4_colmajor(IDENT)
    p.prevViewProj = abi.mat4_colmajor(IDENT)
    p.alpha = alpha
    p.momentsAlpha = moments_alpha
    return p


def test_reflection_reproject_static_camera_clips_history():
    W = H = 24
    g = _plane_gbuffer(W, H)
    g["depth"][0, 0] = 1.0  # one sky pixel
    noisy = np.zeros((H, W, 4), np.float32)
    noisy[..., :3] = 0.5    # grey: luminance 0.5 for any weights summing to 1
    hist = np.zeros((H, W, 4), np.float32)
    hist[..., :3] = (1.0, 0.5, 0.5)
    hmom = np.zeros((H, W, 4), np.float32)
    hmom[..., 0], hmom[..., 1], hmom[..., 2] = 0.2, 0.05, 3.0
    out, mom, tiles = oracle.reflection_reproject(_refl_params(), W, H, g, g, noisy, hist, hmom)
    a = 0.25  # max(alpha 0.1, 1 / histLen 4); moments: max(0.2, 0.25)
    c = out[8:16, 8:16]  # pixels whose 17x17 box lies inside the image: mean 0.5, sd 0
    np.testing.assert_allclose(c[..., 0], 0.5 + 0.001 * (1 - a), atol=1e-7)
    np.testing.assert_allclose(c[..., 1:3], 0.5, atol=1e-7)
    m0, m1 = 0.2 * (1 - a) + 0.5 * a, 0.05 * (1 - a) + 0.25 * a
    np.testing.assert_allclose(mom[1:, 1:, 0], m0, atol=1e-12)
    np.testing.assert_allclose(mom[1:, 1:, 1], m1, atol=1e-12)
    np.testing.assert_allclose(c[..., 3], m1 - m0 * m0, atol=1e-12)
    assert np.all(mom[1:, 1:, 2] == 4.0)
    assert np.all(out[0, 0] == 0) and np.all(mom[0, 0] == 0)
    assert np.all(tiles == 1)


def test_reflection_reproject_disocclusion_restarts_history():
    W, H = 20, 12
    g = _plane_gbuffer(W, H)
    prev = _plane_gbuffer(W, H, mesh=2.0)
    rng = np.random.default_rng(0)
    noisy = np.zeros((H, W, 4), np.float32)
    noisy[..., :3] = rng.uniform(0, 1, (H, W, 3))
    hist = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    hmom = np.zeros((H, W, 4), np.float32)
    hmom[..., 2] = 9.0
    out, mom, tiles = oracle.reflection_reproject(_refl_params(), W, H, g, prev, noisy, hist, hmom)
    np.testing.assert_allclose(out[..., :3], noisy[..., :3], atol=1e-12)
    assert np.all(mom[..., 2] == 1.0)
    np.testing.assert_allclose(out[..., 3], 0.0, atol=1e-12)
    # roughness outside (0.05, 0.45) votes no denoise tile
    g2 = _plane_gbuffer(W, H, rough=0.6)
    _, _, tiles2 = oracle.reflection_reproject(_refl_params(), W, H, g2, g2, noisy, hist, hmom)
    assert np.all(tiles == 1) and np.all(tiles2 == 0)


def _atrous_params(step, radius=2, phi_color=1e30, sigma_depth=1.0):
    return abi.ReflectionAtrousParams(radius, step, phi_color, 32.0, sigma_depth, 0)


B3 = np.array([1.0 / 6.0, 2.0 / 3.0, 1.0, 2.0 / 3.0, 1.0 / 6.0])


def test_reflection_atrous_impulse_response_on_plane():
    W = H = 32
    step = 2
    g = _plane_gbuffer(W, H)
    inp = np.zeros((H, W, 4), np.float32)
    inp[..., 3] = 1.0
    x0 = y0 = 15
    inp[y0, x0, :3] = 0.8
    tiles = np.ones(16, np.uint8)
    out = oracle.reflection_atrous(_atrous_params(step), W, H, g["normals"], g["pbr"], tiles, inp)
    K = np.outer(B3, B3)
    sw = 1.0 + E1 * (K.sum() - 1.0)                  # 1 + e^-1 * 55/9
    sq = 1.0 + E1 * E1 * ((K * K).sum() - 1.0)       # 1 + e^-2 * 901/324
    assert abs(K.sum() - 64.0 / 9.0) < 1e-12
    # the impulse pixel keeps a / sw; lattice neighbours receive e^-1 k a / sw of it
    want = np.zeros((H, W))
    for j in range(-2, 3):
        for i in range(-2, 3):
            y, x = y0 + j * step, x0 + i * step
            want[y, x] = 0.8 * (1.0 if i == j == 0 else E1 * K[j + 2, i + 2]) / sw
    interior = np.zeros((H, W), bool)
    interior[4:-4, 4:-4] = True  # every tap of these pixels lies inside the image
    np.testing.assert_allclose(out[..., 0][interior], want[interior], atol=1e-12)
    np.testing.assert_allclose(out[..., 1][interior], out[..., 2][interior], atol=0)
    np.testing.assert_allclose(out[..., 3][interior], sq / (sw * sw), rtol=1e-12)


def test_reflection_atrous_depth_step_weights_grow_across_edge():
    """exp(-exp(-|dz|/sigma)): equal depths weigh e^-1, a large step ~1 (EdgeStop.glsl:35-38)."""
    W = H = 16
    g = _plane_gbuffer(W, H)
    g["pbr"][:, 8:, 3] = 60.0  # right half far behind: exp(-59.5) ~ 0 -> weight exp(-0) = 1
    inp = np.zeros((H, W, 4), np.float32)
    inp[..., 3] = 1.0
    inp[:, 8:, :3] = 1.0      # right half bright
    out = oracle.reflection_atrous(_atrous_params(1), W, H, g["normals"], g["pbr"], np.ones(4, np.uint8), inp)
    # pixel (x=7, y=8): taps x+1, x+2 lie on the far plane (weight 1 each), others e^-1
    K = np.outer(B3, B3)
    far = K[:, 3:].sum()
    near = K[:, :3].sum() - 1.0
    want = far / (1.0 + far + E1 * near)
    assert abs(out[8, 7, 0] - want) < 1e-9, (out[8, 7, 0], want)
    assert want > E1 * far / (1.0 + E1 * (far + near))  # more than with no depth step


def test_reflection_atrous_copy_sky_and_rough_pixels():
    W, H = 16, 8
    g = _plane_gbuffer(W, H)
    g["pbr"][0, :8, 1] = 0.9   # rough: copied
    g["pbr"][1, :8, 3] = 1.0   # sky depth: zero
    rng = np.random.default_rng(1)
    inp = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    tiles = np.array([0, 1], np.uint8)  # left tile copy, right tile filtered
    out = oracle.reflection_atrous(_atrous_params(1, phi_color=4.0), W, H, g["normals"], g["pbr"], tiles, inp)
    np.testing.assert_array_equal(out[:, :8], inp[:, :8].astype(np.float64))
    tiles = np.ones(2, np.uint8)
    out = oracle.reflection_atrous(_atrous_params(1, phi_color=4.0), W, H, g["normals"], g["pbr"], tiles, inp)
    np.testing.assert_array_equal(out[0, :8], inp[0, :8].astype(np.float64))
    assert np.all(out[1, :8] == 0)


def test_shadow_atrous_constant_power_and_copy_tiles():
    W, H = 16, 16
    g = _plane_gbuffer(W, H)
    inp = np.zeros((H, W, 2), np.float32)
    inp[..., 0] = 0.4
    inp[..., 1] = 0.3
    prm = abi.ShadowAtrousParams(1, 1, 10.0, 32.0, 1.0, 1.2, 0.1, 1000.0)
    out = oracle.shadow_atrous(prm, W, H, g["normals"], g["pbr"], np.array([1, 0, 1, 1], np.uint8), inp)
    K = np.outer(B3[1:4], B3[1:4])
    sw = 1.0 + E1 * (K.sum() - 1.0)   # radius 1: 3x3 taps
    sq = 1.0 + E1 * E1 * ((K * K).sum() - 1.0)
    inner = (slice(1, 7), slice(1, 7))
    v, var = float(np.float32(0.4)), float(np.float32(0.3))  # inputs and power are fp32
    np.testing.assert_allclose(out[inner][..., 0], v ** float(np.float32(1.2)), rtol=1e-12)
    np.testing.assert_allclose(out[inner][..., 1], var * sq / (sw * sw), rtol=1e-12)
    assert np.all(out[:8, 8:] == 0)  # tile (1, 0) is a copy tile: the shadow denoiser writes 0


# ---------------------------------------------------------------------------
# probe relocation / classification
# ---------------------------------------------------------------------------
def _ray_buffers(P, R):
    return np.zeros((P, R, 4), np.float32), np.zeros((P, R), np.uint8)


def test_relocate_classification_and_push():
    u = abi.make_uniform((2, 2, 1), (0, 0, 0), (1.0, 1.0, 1.0), 8)
    dd, fl = _ray_buffers(4, 8)
    dd[..., 0] = 1.0
    dd[..., 3] = 5.0
    # probe 0: 3 of 8 rays start inside (> 0.25 * 8): inactive, never moved
    fl[0, :3] = 2
    fl[0, 3:] = 1
    dd[0, 3:, 3] = 0.05
    # probe 1: a near front hit at t = 0.04 along +x (closest), another at 0.07 along -y
    fl[1, 0] = 1
    dd[1, 0] = (1, 0, 0, 0.04)
    fl[1, 1] = 1
    dd[1, 1] = (0, -1, 0, 0.07)
    # probe 2: only misses and far hits: inactive (no hit within one spacing), not pushed
    fl[2, :2] = 1
    dd[2, :2, 3] = 3.0
    # probe 3: a hit along -z at 0.01 pushes past the clamp (0.45 * step)
    fl[3, 4] = 1
    dd[3, 4] = (0, 0, -1, 0.01)
    off0 = np.zeros((4, 4), np.float32)
    off0[3, 2] = 0.4
    prm = abi.RelocateParams(0.25, 0.1, 0.45, 1)
    off, st = oracle.relocate(u, dd, fl, prm, off0)
    assert st.tolist() == [1, 0, 1, 0]
    np.testing.assert_array_equal(off[0], off0[0])
    np.testing.assert_allclose(off[1], [-(0.1 - 0.04), 0, 0, 0], atol=1e-7)
    np.testing.assert_array_equal(off[2], off0[2])
    np.testing.assert_allclose(off[3], [0, 0, 0.45, 0], atol=1e-7)  # 0.4 + 0.09 clamped to 0.45
    # relocate = 0 classifies only
    off_n, st_n = oracle.relocate(u, dd, fl, abi.RelocateParams(0.25, 0.1, 0.45, 0), off0)
    assert st_n.tolist() == st.tolist()
    np.testing.assert_array_equal(off_n, off0)


def test_relocate_ties_pick_lowest_ray():
    u = abi.make_uniform((1, 1, 1), (0, 0, 0), (1.0, 1.0, 1.0), 4)
    dd, fl = _ray_buffers(1, 4)
    fl[0, 1:3] = 1
    dd[0, 1] = (0, 1, 0, 0.05)
    dd[0, 2] = (1, 0, 0, 0.05)
    off, st = oracle.relocate(u, dd, fl, abi.RelocateParams(0.25, 0.1, 0.45, 1), np.zeros((1, 4), np.float32))
    assert st[0] == 0
    np.testing.assert_allclose(off[0], [0, -0.05, 0, 0], atol=1e-7)

"""GPU: the per-pixel learned edge-stopping scales against the FP64 oracle.

- luxgi_svgf_reflection_atrous_scaled with every scale 1 is luxgi_svgf_reflection_atrous bit for
  bit (LDS-tiled and direct footprints, FP32 and FP16), and with random scales matches the
  oracle's scaled pass.
- luxgi_learned_pixel_scales: the FP32 features within 1e-5 of the FP64 oracle's; the scales
  of 99% of the pixels within 1e-5 (all within 3e-3) of the oracle's MLP run with FP16 operands
  on the GPU's features (the MFMA forward's arithmetic up to the FP32 summation order, which can
  move a hidden activation across an FP16 rounding boundary).
- luxgi_learned_pixel_pass_backward: dL/d(input) and dL/d(scale) against the oracle's backward.
- luxgi_learned_pixel_mlp_backward: dL/d(parameters) against the oracle's with FP16 operands.
- The untrained network (output layer zero) gives scales of exactly 1: its chain is the
  reference chain, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import oracle
from luxgi_amd import abi
from tests import _svgf_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _inputs(W, H, seed):
    g, _, noisy, _, hmom, _, _ = _svgf_inputs.make_pair(W, H, seed)
    rng = np.random.default_rng(seed + 100)
    img = noisy.copy()
    img[..., 3] = rng.uniform(0.0, 0.3, (H, W))
    pbr = g["pbr"].copy()
    pbr[..., 1] = np.where(rng.random((H, W)) < 0.8, rng.uniform(0.1, 0.4, (H, W)), pbr[..., 1])
    pbr[rng.random((H, W)) < 0.05, 3] = 1.0
    g = dict(g)
    g["pbr"] = pbr
    tiles = (rng.random(((H + 7) // 8) * ((W + 7) // 8)) < 0.8).astype(np.uint8)
    return g, img, hmom, tiles


def _gb(g):
    from luxgi_amd.svgf import gbuffer_struct
    d = {k: _dev(v) for k, v in g.items()}
    return d, gbuffer_struct(d)


def _random_params(seed, out_scale=0.3):
    from luxgi_amd import learned_pixel as lp
    p = lp.initial_params(seed)
    rng = np.random.default_rng(seed + 1)
    p[lp.W3:lp.B3] = rng.normal(0.0, out_scale, lp.B3 - lp.W3)
    p[lp.B3:] = rng.normal(0.0, 0.2, 5)
    return p


@pytest.mark.parametrize("storage", [abi.STORAGE_FP32, abi.STORAGE_FP16])
@pytest.mark.parametrize("step", [1, 4, 16])
def test_unit_scales_are_the_reference_pass_and_random_scales_match_the_oracle(storage, step):
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    W, H = 77, 45
    g, img, _, tiles = _inputs(W, H, step)
    gd, gb = _gb(g)
    dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    x = _dev(img).to(dt)
    tm = _dev(tiles)
    ap = abi.ReflectionAtrousParams(1, step, 10.0, 32.0, 1.0, 1)
    ref = torch.zeros_like(x)
    got = torch.zeros_like(x)
    ones = torch.ones((H, W), dtype=torch.float32, device="cuda")
    assert lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, storage, C.byref(gb), ptr(tm), ptr(x), ptr(ref), None,
                                              stream_handle()) == 0
    assert lib().luxgi_svgf_reflection_atrous_scaled(C.byref(ap), W, H, storage, C.byref(gb), ptr(tm), ptr(x), ptr(got),
                                                     None, ptr(ones), stream_handle()) == 0
    torch.cuda.synchronize()
    print(f"scaled pass step {step} storage {storage}", flush=True)
    assert torch.equal(got, ref)
    if storage != abi.STORAGE_FP32:
        return
    rng = np.random.default_rng(step)
    sc = np.exp(rng.uniform(-2.0, 2.0, (H, W))).astype(np.float32)
    assert lib().luxgi_svgf_reflection_atrous_scaled(C.byref(ap), W, H, storage, C.byref(gb), ptr(tm), ptr(x), ptr(got),
                                                     None, ptr(_dev(sc)), stream_handle()) == 0
    want = oracle.reflection_atrous_scaled(ap, W, H, g["normals"], g["pbr"], tiles, img, sc)
    a = got.cpu().numpy()
    ok = np.all(np.abs(a - want) <= 1e-4 + 1e-4 * np.abs(want), axis=-1)
    assert ok.mean() >= 0.999, ok.mean()
    assert not np.allclose(a, ref.cpu().numpy())  # the scales change the result


@pytest.mark.parametrize("W,H,seed", [(77, 45, 0), (256, 33, 1), (5, 3, 2)])
def test_features_and_scales_match_the_oracle(W, H, seed):
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    g, img, hmom, _ = _inputs(W, H, seed)
    gd, gb = _gb(g)
    p = _random_params(seed)
    pd = _dev(p.astype(np.float32))
    scales = torch.zeros((5, H, W), dtype=torch.float32, device="cuda")
    feat = torch.zeros((H, W, 12), dtype=torch.float32, device="cuda")
    print(f"scales {W}x{H}", flush=True)
    assert lib().luxgi_learned_pixel_scales(ptr(pd), W, H, C.byref(gb), ptr(_dev(img)), ptr(_dev(hmom)), ptr(scales),
                                            ptr(feat), stream_handle()) == 0
    torch.cuda.synchronize()
    f = feat.cpu().numpy()
    fo = oracle.learned_pixel_features(img, hmom, g["normals"], g["pbr"])
    assert np.allclose(f, fo, rtol=1e-5, atol=1e-5), np.abs(f - fo).max()
    o, s = oracle.learned_pixel_mlp(p.astype(np.float32).astype(np.float64), f.astype(np.float64), half=True)
    s = np.moveaxis(s, -1, 0)
    got = scales.cpu().numpy()
    # the same FP16 operands up to the FP32 summation order; where that order moves a hidden
    # activation across an FP16 rounding boundary, one operand differs by an FP16 ulp
    rel = np.abs(got - s) / s
    assert rel.max() <= 3e-3, rel.max()
    assert (rel <= 1e-5).mean() >= 0.99, (rel <= 1e-5).mean()
    assert np.ptp(got) > 0.1  # the scales vary


@pytest.mark.parametrize("step", [1, 2, 8, 16])
def test_pass_backward_matches_the_oracle(step):
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    W, H = 61, 37
    g, img, _, tiles = _inputs(W, H, 10 + step)
    gd, gb = _gb(g)
    rng = np.random.default_rng(step)
    sc = np.exp(rng.uniform(-1.0, 1.0, (H, W))).astype(np.float32)
    go = rng.normal(size=(H, W, 4)).astype(np.float32)
    go[..., 3] = 0.0
    ap = abi.ReflectionAtrousParams(1, step, 10.0, 32.0, 1.0, 1)
    gin = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    gsc = torch.zeros((H, W), dtype=torch.float32, device="cuda")
    ws = torch.empty(lib().luxgi_learned_pixel_pass_backward_workspace_size(W, H), dtype=torch.uint8, device="cuda")
    print(f"pass backward step {step}", flush=True)
    assert lib().luxgi_learned_pixel_pass_backward(C.byref(ap), W, H, C.byref(gb), ptr(_dev(tiles)), ptr(_dev(img)),
                                                   ptr(_dev(sc)), ptr(_dev(go)), ptr(gin), ptr(gsc), ptr(ws),
                                                   ws.numel(), stream_handle()) == 0
    torch.cuda.synchronize()
    wi, ws_ = oracle.learned_pixel_atrous_backward(ap, W, H, g["normals"], g["pbr"], tiles, img, sc, go)
    a, b = gin.cpu().numpy(), gsc.cpu().numpy()
    assert np.all(a[..., 3] == 0.0)
    assert np.allclose(a, wi, rtol=1e-4, atol=1e-5 * np.abs(wi).max()), np.abs(a - wi).max()
    assert np.allclose(b, ws_, rtol=1e-3, atol=1e-4 * np.abs(ws_).max()), np.abs(b - ws_).max()
    assert np.abs(ws_).max() > 1e-3


def test_mlp_backward_matches_the_oracle_and_is_deterministic():
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    W, H = 97, 41
    g, img, hmom, _ = _inputs(W, H, 5)
    gd, gb = _gb(g)
    p = _random_params(7)
    pd = _dev(p.astype(np.float32))
    scales = torch.zeros((5, H, W), dtype=torch.float32, device="cuda")
    feat = torch.zeros((H, W, 12), dtype=torch.float32, device="cuda")
    assert lib().luxgi_learned_pixel_scales(ptr(pd), W, H, C.byref(gb), ptr(_dev(img)), ptr(_dev(hmom)), ptr(scales),
                                            ptr(feat), stream_handle()) == 0
    rng = np.random.default_rng(3)
    gs = rng.normal(size=(5, H, W)).astype(np.float32)
    ws = torch.empty(lib().luxgi_learned_pixel_mlp_backward_workspace_size(), dtype=torch.uint8, device="cuda")
    grads = []
    for _ in range(2):
        gp = torch.zeros(abi.LEARNED_PIXEL_PARAMS, dtype=torch.float32, device="cuda")
        print("mlp backward", flush=True)
        assert lib().luxgi_learned_pixel_mlp_backward(ptr(pd), W, H, ptr(feat), ptr(scales), ptr(_dev(gs)), ptr(gp),
                                                      ptr(ws), ws.numel(), stream_handle()) == 0
        torch.cuda.synchronize()
        grads.append(gp.cpu().numpy())
    assert np.array_equal(grads[0], grads[1])
    f = feat.cpu().numpy().astype(np.float64)
    want = oracle.learned_pixel_mlp_backward(p.astype(np.float32).astype(np.float64), f,
                                             np.moveaxis(gs, 0, -1).astype(np.float64), half=True)
    got = grads[0]
    assert np.allclose(got, want, rtol=2e-3, atol=1e-4 * np.abs(want).max()), np.abs(got - want).max()


def test_untrained_network_chain_is_the_reference_chain():
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.learned_pixel import PixelScaleChain
    W, H = 80, 48
    g, img, hmom, tiles = _inputs(W, H, 9)
    gd, gb = _gb(g)
    chain = PixelScaleChain(W, H, iterations=5)
    x, m, tm = _dev(img), _dev(hmom), _dev(tiles)
    out = chain.forward(gd, x, m, tm).clone()
    assert torch.all(chain.scales == 1.0)
    src = x
    bufs = [torch.zeros_like(x), torch.zeros_like(x)]
    for i in range(5):
        ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
        assert lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, abi.STORAGE_FP32, C.byref(gb), ptr(tm), ptr(src),
                                                  ptr(bufs[i & 1]), None, stream_handle()) == 0
        src = bufs[i & 1]
    torch.cuda.synchronize()
    assert torch.equal(out, src)


def test_training_on_rendered_views_beats_plain_svgf_on_held_out_views():
    """Six seeded config-1-style views (192x128; reflections rendered by the method, the target the
    mean of 256 frames), three reference-denoiser frames each, 300 Adam steps; on two held-out
    scenes, with each denoiser feeding back its own filtered history over frames 4..12, the learned
    denoiser's error against the target is below plain SVGF's."""
    from luxgi_amd import learned_pixel as lp
    W, H = 192, 128
    keep = (2, 5, 11)
    views, samples = [], []
    for k in range(6):
        v = lp.ReflectionView(100 + k, W, H, keep_raw=max(keep) + 1)
        for s in lp.reference_sequence_samples(v, keep):
            samples.append((len(views), s))
        views.append(v)
    assert all(v.mask.float().mean() > 0.3 for v in views)
    chain = lp.PixelScaleChain(W, H)
    print("training", flush=True)
    losses = lp.train(chain, views, samples, 300, lr=3e-3)
    assert np.mean(losses[-50:]) < np.mean(losses[:50])
    params = chain.params_numpy()
    res = []
    for k in range(2):
        v = lp.ReflectionView(900 + k, W, H, keep_raw=13)
        res.append(lp.evaluate(params, v, (4, 6, 8, 10, 12)))
        print(f"held-out view {k}: {res[-1]}", flush=True)
    plain = np.mean([r["plain"] for r in res])
    learned = np.mean([r["learned"] for r in res])
    assert learned < plain, (learned, plain)


def test_learned_denoiser_falls_back_and_keeps_the_temporal_structure():
    """Unvalidated: every frame equals ReflectionDenoiser's bit for bit. Validated better: each frame
    is the reprojection followed by the scaled chain, with pass 1's output as the next history."""
    from luxgi_amd import learned_pixel as lp
    from luxgi_amd.svgf import ReflectionDenoiser
    W, H = 64, 48
    g, prev, noisy, _, _, _, cam = _svgf_inputs.make_pair(W, H, 3)
    gd = {k: _dev(v) for k, v in g.items()}
    pd = {k: _dev(v) for k, v in prev.items()}
    x = _dev(noisy)
    camt = (cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])
    ref = ReflectionDenoiser(W, H, iterations=5)
    ref.set_camera(*camt)
    den = lp.LearnedPixelDenoiser(W, H, params=_random_params(4), iterations=5)
    den.set_camera(*camt)
    for _ in range(3):
        assert torch.equal(den.denoise(gd, pd, x), ref.denoise(gd, pd, x))
    worse = lp.LearnedPixelDenoiser(W, H)
    worse.validate(1.0, 0.5)
    assert not worse.use_learned
    den.validate(0.5, 1.0)
    assert den.use_learned
    # learned frames against a hand-run reprojection + chain with the filtered history
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    d = den.plain
    hist_c, hist_m = d.outputs[1].clone(), d.moments[1 - d.ping_pong].clone()
    chain = lp.PixelScaleChain(W, H, params=_random_params(4), iterations=5)
    for f in range(3):
        oc, om = torch.zeros_like(x), torch.zeros_like(x)
        tiles = torch.zeros_like(d.tile_map)
        assert lib().luxgi_svgf_reflection_reproject(C.byref(d.rp), W, H, abi.STORAGE_FP32, C.byref(gbuffer_struct(gd)),
                                                     C.byref(gbuffer_struct(pd)), ptr(x), ptr(hist_c), ptr(hist_m),
                                                     ptr(oc), ptr(om), ptr(tiles), stream_handle()) == 0
        want = chain.forward(gd, oc, om, tiles).clone()
        got = den.denoise(gd, pd, x)
        torch.cuda.synchronize()
        print(f"learned denoiser frame {f}", flush=True)
        assert torch.equal(got, want)
        hist_c, hist_m = chain.images[1].clone(), om

"""GPU parity of the learned a-trous pass, its backward and Adam, and the learned chain's gain over
the reference's a-trous chain on reflections it was not trained on.

Tolerances. The forward has no threshold decisions on the inputs other than the pixel's mode
(sky / pass-through / filtered, all exact compares of G-buffer values): every pixel agrees to
1e-4 relative + 1e-5 absolute (fp32 features and __expf against FP64). The input gradient is a
weighted sum of the same weights: 1e-4 relative + 1e-6 absolute. The parameter gradient sums
~10^5 tap terms of both signs through fp32 atomics: each component agrees to 2e-3 of the
largest component + 1e-3 relative."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from luxgi_amd import abi, scene
from tests import _scenes
from tests.test_oracle_learned import _inputs, _params, _random_weights

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _gb(nrm, pbr):
    H, W = pbr.shape[:2]
    return {"color": torch.zeros((H, W, 4), device="cuda"), "normals": torch.from_numpy(nrm).cuda(),
            "pbr": torch.from_numpy(pbr).cuda(), "depth": torch.ones((H, W), device="cuda")}


def _forward(p, wts, gbd, img):
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    H, W = img.shape[:2]
    tw = torch.from_numpy(np.asarray(wts, np.float32)).cuda()
    ti = torch.from_numpy(img).cuda()
    out = torch.full((H, W, 4), -7.0, device="cuda")
    sw = torch.zeros((H, W, 4), device="cuda")
    rc = lib().luxgi_svgf_learned_atrous(C.byref(p), ptr(tw), W, H, C.byref(gbuffer_struct(gbd)), ptr(ti), ptr(out),
                                         ptr(sw), stream_handle())
    assert rc == 0
    return tw, ti, out, sw


# footprints radius * step <= 8 run the LDS-tiled kernel, larger ones the direct kernel
@pytest.mark.parametrize("W,H,radius,step,seed", [(77, 45, 1, 1, 0), (128, 64, 2, 2, 1), (40, 33, 1, 8, 2),
                                                  (9, 7, 2, 1, 3), (50, 40, 1, 16, 4), (60, 44, 2, 8, 5)])
def test_forward_matches_oracle(W, H, radius, step, seed):
    nrm, pbr, img = _inputs(W, H, seed=seed)
    wts = _random_weights(seed)
    p = _params(radius, step, level=0.25 * seed)
    _, _, out, sw = _forward(p, wts, _gb(nrm, pbr), img)
    torch.cuda.synchronize()
    want, want_sw = oracle.learned_atrous_forward(p, wts, nrm, pbr, img)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(sw.cpu().numpy()[..., 0], want_sw, rtol=1e-4)


def test_reference_like_weights_reproduce_the_reference_pass():
    """The hand-set MLP with agreeing normals on the GPU against the reference's a-trous oracle."""
    from luxgi_amd.learned import reference_like_weights
    W, H = 64, 40
    nrm, pbr, img = _inputs(W, H, seed=6, same_normals=True)
    p = _params(1, 2, 10.0, 0.7)
    _, _, out, _ = _forward(p, reference_like_weights(live_units=False), _gb(nrm, pbr), img)
    torch.cuda.synchronize()
    rp = abi.ReflectionAtrousParams(1, 2, 10.0, 32.0, 0.7, 1)
    want = oracle.reflection_atrous(rp, W, H, nrm, pbr, np.ones(((H + 7) // 8) * ((W + 7) // 8), np.uint8), img)
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-4, atol=2e-5)


@pytest.mark.parametrize("W,H,radius,step,seed", [(61, 37, 1, 1, 4), (96, 80, 2, 2, 5)])
def test_backward_matches_oracle(W, H, radius, step, seed):
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    nrm, pbr, img = _inputs(W, H, seed=seed)
    wts = _random_weights(seed)
    p = _params(radius, step, level=0.5)
    gbd = _gb(nrm, pbr)
    tw, ti, out, sw = _forward(p, wts, gbd, img)
    gout = np.random.default_rng(seed).normal(0.0, 1.0, (H, W, 4)).astype(np.float32)
    tg = torch.from_numpy(gout).cuda()
    gin = torch.full((H, W, 4), 5.0, device="cuda")
    gp = torch.zeros(81, device="cuda")
    for _ in range(2):  # accumulates
        rc = lib().luxgi_svgf_learned_atrous_backward(C.byref(p), ptr(tw), W, H, C.byref(gbuffer_struct(gbd)), ptr(ti),
                                                      ptr(out), ptr(sw), ptr(tg), ptr(gin), ptr(gp), stream_handle())
        assert rc == 0
    torch.cuda.synchronize()
    g64 = gout.astype(np.float64)
    g64[..., 3] = 0.0
    want_in, want_p = oracle.learned_atrous_backward(p, wts, nrm, pbr, img, g64)
    got_in = gin.cpu().numpy()
    np.testing.assert_allclose(got_in, want_in, rtol=1e-4, atol=1e-6 * np.abs(want_in).max())
    got_p = gp.cpu().numpy() / 2.0
    assert np.all(np.abs(got_p - want_p) <= 2e-3 * np.abs(want_p).max() + 1e-3 * np.abs(want_p)), \
        np.abs(got_p - want_p).max()


def test_adam_step_matches_the_published_update():
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    rng = np.random.default_rng(3)
    n = 300
    w = rng.normal(0, 1, n)
    m = np.zeros(n)
    v = np.zeros(n)
    tw = torch.from_numpy(w.astype(np.float32)).cuda()
    tm = torch.zeros(n, device="cuda")
    tv = torch.zeros(n, device="cuda")
    lr, b1, b2, eps = 0.01, 0.9, 0.999, 1e-8
    for t in range(1, 4):
        g = rng.normal(0, 1, n)
        tg = torch.from_numpy(g.astype(np.float32)).cuda()
        assert lib().luxgi_adam_step(ptr(tw), ptr(tg), ptr(tm), ptr(tv), n, lr, b1, b2, eps, t, stream_handle()) == 0
        torch.cuda.synchronize()
        assert torch.all(tg == 0)
        g = g.astype(np.float32).astype(np.float64)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    np.testing.assert_allclose(tw.cpu().numpy(), w, rtol=1e-5, atol=1e-6)


def test_learned_entry_points_reject_bad_arguments():
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    nrm, pbr, img = _inputs(8, 8, seed=0)
    gbd = _gb(nrm, pbr)
    t = torch.zeros((8, 8, 4), device="cuda")
    tw = torch.zeros(81, device="cuda")
    gb = gbuffer_struct(gbd)
    L = lib()
    st = stream_handle()
    o = t.clone()
    assert L.luxgi_svgf_learned_atrous(C.byref(_params(3, 1)), ptr(tw), 8, 8, C.byref(gb), ptr(t), ptr(o), None, st) < 0
    assert L.luxgi_svgf_learned_atrous(C.byref(_params(1, 1)), ptr(tw), 8, 8, C.byref(gb), ptr(t), ptr(t), None, st) < 0
    assert L.luxgi_svgf_learned_atrous(C.byref(_params(1, 0)), ptr(tw), 8, 8, C.byref(gb), ptr(t), ptr(o), None, st) < 0
    assert L.luxgi_svgf_learned_atrous(C.byref(_params(1, 1)), None, 8, 8, C.byref(gb), ptr(t), ptr(o), None, st) < 0
    assert L.luxgi_svgf_learned_atrous_backward(C.byref(_params(1, 1)), ptr(tw), 8, 8, C.byref(gb), ptr(t), ptr(o),
                                                None, ptr(t), ptr(o), ptr(tw), st) < 0
    assert L.luxgi_adam_step(ptr(tw), ptr(tw), ptr(tw), ptr(tw), 81, 0.01, 0.9, 0.999, 1e-8, 0, st) < 0


# ---------------------------------------------------------------------------------------------
# Gain over the reference chain
# ---------------------------------------------------------------------------------------------
def _reflection_view(W, H, eye, cam_seed, frames, target_frames=256):
    """G-buffer of the small scene from a static camera, SDF reflection frames (one GGX-sampled ray
    per pixel, a new blue-noise sample each frame) accumulated by the reflection reprojection,
    and the converged target (mean of `target_frames` raw frames)."""
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import DeviceScene, ptr, stream_handle, to_device
    from luxgi_amd.producers import BlueNoiseTables, sdf_reflection
    from luxgi_amd.svgf import gbuffer_struct, make_gbuffer
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    VP, _, eye, _ = scene.camera_pair(W, H, seed=cam_seed, eye=eye)
    VPI = np.linalg.inv(VP)
    prims_dev = to_device(np.frombuffer(prims.tobytes(), dtype=np.uint8))
    gb, _, _ = make_gbuffer(prims_dev, len(prims), g, torch.from_numpy(sdf).cuda(), 64, VPI, VP, W, H, seed=3,
                            with_mask=False)
    # misses: the far-plane value 1 in pbr.w (the reference's sky test); hits: roughness in the GGX band
    miss = gb["depth"] == 1.0
    gb["pbr"][..., 3] = torch.where(miss, torch.ones_like(gb["depth"]), gb["pbr"][..., 3])
    gb["pbr"][..., 1] = torch.where(miss, gb["pbr"][..., 1], gb["pbr"][..., 1].clamp(0.1, 0.4))
    ds = DeviceScene(g, sdf, mip, atlas)
    rng = np.random.default_rng(cam_seed)
    nz = BlueNoiseTables(np.stack([rng.permutation(256) for _ in range(4)], axis=1).astype(np.uint8),
                         rng.integers(0, 256, (128, 128, 4), dtype=np.uint8))
    p = abi.SdfReflectionParams()
    p.trim, p.roughDDGIIntensity, p.approximateWithDDGI = 1.0, 1.0, 0
    p.cameraPosition[:] = [float(eye[0]), float(eye[1]), float(eye[2]), 1.0]
    p.viewProjInv = abi.mat4_colmajor(VPI)
    p.skyColor[:] = [0.6, 0.7, 0.9, 1.0]
    raw = []
    target = torch.zeros((H, W, 4), device="cuda")
    for f in range(target_frames):
        p.numFrames = f
        out = torch.zeros((H, W, 4), device="cuda")
        sdf_reflection(p, nz, gb, ds, out)
        target += out
        if f < frames:
            raw.append(out)
    target /= target_frames
    rp = abi.ReflectionReprojectParams()
    rp.viewProjInv = abi.mat4_colmajor(VPI)
    rp.prevViewProj = abi.mat4_colmajor(VP)
    rp.cameraPos[:] = [float(eye[0]), float(eye[1]), float(eye[2]), 1.0]
    rp.alpha, rp.momentsAlpha, rp.approximateWithDDGI = 0.01, 0.2, 1
    hist_c = torch.zeros((H, W, 4), device="cuda")
    hist_m = torch.zeros((H, W, 4), device="cuda")
    tiles = torch.zeros(lib().luxgi_svgf_tile_count(W, H), dtype=torch.uint8, device="cuda")
    seq = []
    gbs = gbuffer_struct(gb)
    for f in range(frames):
        oc = torch.zeros((H, W, 4), device="cuda")
        om = torch.zeros((H, W, 4), device="cuda")
        assert lib().luxgi_svgf_reflection_reproject(C.byref(rp), W, H, abi.STORAGE_FP32, C.byref(gbs), C.byref(gbs),
                                                     ptr(raw[f]), ptr(hist_c), ptr(hist_m), ptr(oc), ptr(om),
                                                     ptr(tiles), stream_handle()) == 0
        seq.append((oc, tiles.clone()))
        hist_c, hist_m = oc, om
    torch.cuda.synchronize()
    return gb, seq, target


def _reference_chain(gb, img, tiles, iterations=4):
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.svgf import gbuffer_struct
    H, W = img.shape[:2]
    gbs = gbuffer_struct(gb)
    bufs = [torch.zeros_like(img), torch.zeros_like(img)]
    src = img
    for i in range(iterations):
        ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
        assert lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, abi.STORAGE_FP32, C.byref(gbs), ptr(tiles),
                                                  ptr(src), ptr(bufs[i & 1]), None, stream_handle()) == 0
        src = bufs[i & 1]
    return src


def test_learned_chain_beats_the_reference_chain_on_a_held_out_view():
    """Trained on one view's reflections (frames 1..15 of the accumulation), evaluated on another
    view (frames 4..12): the mean squared error against the 256-frame mean over the filtered
    pixels is below the reference chain's (phiColor 10, phiNormal 32, sigmaDepth 1, 4 passes)."""
    from luxgi_amd.learned import LearnedAtrous, filtered_mask
    W, H = 192, 128
    torch.manual_seed(0)
    gA, seqA, tA = _reflection_view(W, H, (0.1, 0.3, 0.75), 9, 16)
    gB, seqB, tB = _reflection_view(W, H, (-0.25, 0.4, 0.6), 11, 13)
    mA, mB = filtered_mask(gA), filtered_mask(gB)
    assert mA.float().mean() > 0.3 and mB.float().mean() > 0.3

    def mse(x, t, m):
        return float((((x[..., :3] - t[..., :3]) ** 2).sum(-1) * m).sum() / m.sum())

    la = LearnedAtrous(W, H)
    frames_b = range(4, 13)
    before = np.mean([mse(la.forward(gB, seqB[k][0]), tB, mB) for k in frames_b])
    rng = np.random.default_rng(0)
    losses = []
    for step in range(300):
        k = int(rng.integers(1, 16))
        losses.append(la.train_step(gA, seqA[k][0], tA, mA, lr=1e-2))
    torch.cuda.synchronize()
    learned = np.mean([mse(la.forward(gB, seqB[k][0]), tB, mB) for k in frames_b])
    ref = np.mean([mse(_reference_chain(gB, seqB[k][0], seqB[k][1]), tB, mB) for k in frames_b])
    raw = np.mean([mse(seqB[k][0], tB, mB) for k in frames_b])
    first = float(torch.stack(losses[:20]).mean())
    last = float(torch.stack(losses[-20:]).mean())
    print(f"\n[learned atrous] held-out MSE: unfiltered {raw:.6f}, reference chain {ref:.6f}, "
          f"learned init {before:.6f}, learned trained {learned:.6f}; train loss {first:.6f} -> {last:.6f}",
          flush=True)
    assert all(math.isfinite(x) for x in (raw, ref, learned))
    assert last < first
    assert learned < 0.9 * ref, (learned, ref)
    assert learned < raw

    # the denoiser that uses the learned chain only when it validates better than the reference's
    from luxgi_amd.learned import LearnedReflectionDenoiser
    samples = [(gB, seqB[k][0], seqB[k][1], tB, mB) for k in frames_b]
    good = LearnedReflectionDenoiser(W, H, weights=la.weights_numpy())
    assert not good.use_learned  # nothing validated yet: the reference path
    v = good.validate(samples)
    assert good.use_learned and v["learned_mse"] < v["reference_mse"], v
    blur = np.zeros(81)
    blur[80] = 8.0  # every tap weighted e^8: a box blur over each pass's footprint
    bad = LearnedReflectionDenoiser(W, H, weights=blur, radius=2)
    v = bad.validate(samples)
    assert bad.use_learned == (v["learned_mse"] < v["reference_mse"]), v


def test_learned_denoiser_falls_back_to_the_reference_path_and_keeps_its_temporal_structure():
    """Unvalidated: every frame equals ReflectionDenoiser's bit for bit. Learned: each frame equals
    the reprojection followed by the learned chain, with pass 1's output as the next history."""
    from luxgi_amd._lib import lib
    from luxgi_amd.ddgi import ptr, stream_handle
    from luxgi_amd.learned import LearnedAtrous, LearnedReflectionDenoiser
    from luxgi_amd.svgf import ReflectionDenoiser, gbuffer_struct
    from tests import _svgf_inputs
    W, H = 64, 48
    g, prev, noisy, _, _, _, cam = _svgf_inputs.make_pair(W, H, 3)
    dev = lambda d: {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in d.items()}  # noqa: E731
    g, prev = dev(g), dev(prev)
    frames = [torch.from_numpy(noisy * (1.0 + 0.2 * f)).cuda() for f in range(3)]
    ld = LearnedReflectionDenoiser(W, H)
    ref = ReflectionDenoiser(W, H, iterations=4)
    for d in (ld, ref):
        d.set_camera(cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])
    for f in frames:
        assert torch.equal(ld.denoise(g, prev, f), ref.denoise(g, prev, f))
    # learned path against a hand-run reprojection + chain
    ld2 = LearnedReflectionDenoiser(W, H)
    ld2.set_camera(cam["VPI"], cam["VPp"], cam["eye"], cam["delta"])
    ld2.use_learned = True
    la = LearnedAtrous(W, H)
    rp = ld2.plain.rp
    hist_c = torch.zeros((H, W, 4), device="cuda")
    mom = [torch.zeros((H, W, 4), device="cuda"), torch.zeros((H, W, 4), device="cuda")]
    tiles = torch.zeros(lib().luxgi_svgf_tile_count(W, H), dtype=torch.uint8, device="cuda")
    pp = 0
    for f in frames:
        got = ld2.denoise(g, prev, f).clone()
        oc = torch.zeros((H, W, 4), device="cuda")
        assert lib().luxgi_svgf_reflection_reproject(C.byref(rp), W, H, abi.STORAGE_FP32, C.byref(gbuffer_struct(g)),
                                                     C.byref(gbuffer_struct(prev)), ptr(f), ptr(hist_c),
                                                     ptr(mom[1 - pp]), ptr(oc), ptr(mom[pp]), ptr(tiles),
                                                     stream_handle()) == 0
        want = la.forward(g, oc).clone()
        hist_c = la.images[1].clone()
        pp = 1 - pp
        torch.cuda.synchronize()
        assert torch.equal(got, want)

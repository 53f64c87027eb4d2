"""Per-pixel learned edge-stopping scales for the reflection a-trous chain (libluxgi.so).

A 12 -> 32 -> 32 -> 5 ReLU MLP with FP16 weights (FP32 accumulation, on the MFMA) maps per-pixel
features (depth gradient, normal, luminance and its 3x3 mean and spread, the variance channel, the
reprojection's moments and history length) to one scale per a-trous pass. Pass i is the
reference's reflection a-trous pass (Reflection/DenoiseAtrous.comp:63-134, driven by
denoise::system, Engine/Raytrace/RaytracedReflection.cpp:419-494) with the luminance term of
every tap's edge-stopping weight (EdgeStop.glsl:24-40) multiplied by the pixel's scale. With the
output layer at zero every scale is exactly 1 and the chain is the reference chain, bit for bit.

Training runs on the GPU: luxgi_learned_pixel_scales keeps the features, the scaled passes keep
their inputs, luxgi_learned_pixel_pass_backward returns dL/d(input) and dL/d(scale) per pass,
luxgi_learned_pixel_mlp_backward adds dL/d(parameters), luxgi_adam_step updates them. The loss is
the mean squared error of the chain's output against a converged target over the filtered pixels.

Parameter layout: W1 [32][12], b1 [32], W2 [32][32], b2 [32], W3 [5][32], b3 [5].
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .ddgi import ptr, stream_handle
from .svgf import gbuffer_struct

N_PARAMS = abi.LEARNED_PIXEL_PARAMS
N_FEATURES, N_HIDDEN, N_PASSES = abi.LEARNED_PIXEL_FEATURES, 32, abi.LEARNED_PIXEL_PASSES
W1, B1 = 0, N_HIDDEN * N_FEATURES
W2 = B1 + N_HIDDEN
B2 = W2 + N_HIDDEN * N_HIDDEN
W3 = B2 + N_HIDDEN
B3 = W3 + N_PASSES * N_HIDDEN


def initial_params(seed=0, scale=0.3):
    """Hidden layers with random weights and positive biases (live units), output layer zero: every
    scale is 1, so the untrained chain is the reference chain; the output layer's gradient is
    non-zero from the first step."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    p[W1:B1] = rng.normal(0.0, scale * np.sqrt(2.0 / N_FEATURES), N_HIDDEN * N_FEATURES)
    p[B1:W2] = 0.1
    p[W2:B2] = rng.normal(0.0, scale * np.sqrt(2.0 / N_HIDDEN), N_HIDDEN * N_HIDDEN)
    p[B2:W3] = 0.1
    return p


def filtered_mask(gbuffer):
    """Pixels the a-trous pass filters: not sky (pbr.w != 1) and roughness in (0.05, 0.45)."""
    pbr = gbuffer["pbr"]
    r = pbr[..., 1]
    return (pbr[..., 3] != 1.0) & (r > 0.05) & (r < 0.45)


class PixelScaleChain:
    """`iterations` (at most 5) scaled a-trous passes over RGBA32F (rgb, variance) images on the
    current stream, the scales from the network."""

    def __init__(self, width, height, params=None, iterations=5, radius=1, phi_color=10.0, phi_normal=32.0,
                 sigma_depth=1.0):
        if not 1 <= iterations <= N_PASSES:
            raise ValueError(f"iterations must be in [1, {N_PASSES}]")
        params = initial_params() if params is None else np.asarray(params, dtype=np.float64).reshape(-1)
        if params.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {params.size}")
        self.W, self.H = width, height
        self.iterations = iterations
        self.params = torch.from_numpy(params.astype(np.float32)).cuda()
        self.grad = torch.zeros(N_PARAMS, dtype=torch.float32, device="cuda")
        self.m = torch.zeros_like(self.grad)
        self.v = torch.zeros_like(self.grad)
        self.step = 0
        self.aps = [abi.ReflectionAtrousParams(radius, 1 << i, phi_color, phi_normal, sigma_depth, 1)
                    for i in range(iterations)]
        z = lambda: torch.zeros((height, width, 4), dtype=torch.float32, device="cuda")  # noqa: E731
        self.images = [z() for _ in range(iterations)]
        self.grads = [z(), z()]
        self.scales = torch.ones((N_PASSES, height, width), dtype=torch.float32, device="cuda")
        self.features = torch.zeros((height, width, N_FEATURES), dtype=torch.float32, device="cuda")
        self.grad_scales = torch.zeros((N_PASSES, height, width), dtype=torch.float32, device="cuda")
        self.pass_ws = torch.empty(lib().luxgi_learned_pixel_pass_backward_workspace_size(width, height),
                                   dtype=torch.uint8, device="cuda")
        self.mlp_ws = torch.empty(lib().luxgi_learned_pixel_mlp_backward_workspace_size(), dtype=torch.uint8,
                                  device="cuda")
        self._inputs = [None] * iterations
        self._tiles = None

    def compute_scales(self, gbuffer, inp, moments, keep_features=True):
        """The network's scales of the passes; keep_features stores the features for backward()."""
        check(lib().luxgi_learned_pixel_scales(ptr(self.params), self.W, self.H, C.byref(gbuffer_struct(gbuffer)),
                                               ptr(inp), ptr(moments), ptr(self.scales),
                                               ptr(self.features) if keep_features else None, stream_handle()),
              "luxgi_learned_pixel_scales")
        return self.scales

    def forward(self, gbuffer, inp, moments, tile_map, history_out=None, keep_features=True):
        """The chain on `inp` ([H][W][4] FP32, the reprojection's colour + variance) with the
        reprojection's `moments` and tile map; pass 1's output is also written to `history_out`
        when given (the reference's filtered-history feedback). Returns the last pass's output (an
        internal buffer, valid until the next call). Inference passes keep_features=False."""
        if inp.shape != (self.H, self.W, 4) or inp.dtype != torch.float32:
            raise ValueError("input must be an [H][W][4] float32 tensor")
        self.compute_scales(gbuffer, inp, moments, keep_features)
        gb = gbuffer_struct(gbuffer)
        src = inp.contiguous()
        self._tiles = tile_map
        for i, ap in enumerate(self.aps):
            self._inputs[i] = src
            hist = history_out if (i == 1 and history_out is not None) else None
            check(lib().luxgi_svgf_reflection_atrous_scaled(
                C.byref(ap), self.W, self.H, abi.STORAGE_FP32, C.byref(gb), ptr(tile_map), ptr(src), ptr(self.images[i]),
                ptr(hist), ptr(self.scales[i]), stream_handle()), "luxgi_svgf_reflection_atrous_scaled")
            src = self.images[i]
        return src

    def backward(self, gbuffer, grad_out):
        """Back-propagates dL/d(output rgb) through the last forward(): adds dL/d(parameters) into
        self.grad and returns dL/d(input rgb) (an internal buffer)."""
        gb = gbuffer_struct(gbuffer)
        g = grad_out.contiguous()
        self.grad_scales.zero_()
        for i in reversed(range(self.iterations)):
            dst = self.grads[i & 1]
            if dst.data_ptr() == g.data_ptr():
                dst = self.grads[1 - (i & 1)]
            check(lib().luxgi_learned_pixel_pass_backward(
                C.byref(self.aps[i]), self.W, self.H, C.byref(gb), ptr(self._tiles), ptr(self._inputs[i]),
                ptr(self.scales[i]), ptr(g), ptr(dst), ptr(self.grad_scales[i]), ptr(self.pass_ws), self.pass_ws.numel(),
                stream_handle()), "luxgi_learned_pixel_pass_backward")
            g = dst
        check(lib().luxgi_learned_pixel_mlp_backward(
            ptr(self.params), self.W, self.H, ptr(self.features), ptr(self.scales), ptr(self.grad_scales),
            ptr(self.grad), ptr(self.mlp_ws), self.mlp_ws.numel(), stream_handle()), "luxgi_learned_pixel_mlp_backward")
        return g

    def adam(self, lr=3e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step on the accumulated gradient (cleared afterwards)."""
        self.step += 1
        check(lib().luxgi_adam_step(ptr(self.params), ptr(self.grad), ptr(self.m), ptr(self.v), N_PARAMS, lr, beta1,
                                    beta2, eps, self.step, stream_handle()), "luxgi_adam_step")

    def train_step(self, gbuffer, inp, moments, tile_map, target, mask, lr=3e-3):
        """One step on one image: forward, MSE over `mask` (bool [H][W]) and the rgb channels,
        backward, Adam. Returns the loss as a device scalar."""
        out = self.forward(gbuffer, inp, moments, tile_map)
        mf = mask.to(torch.float32)[..., None]
        n = float(max(int(mask.sum().item()), 1))
        diff = (out[..., :3] - target[..., :3]) * mf
        g = torch.zeros((self.H, self.W, 4), dtype=torch.float32, device="cuda")
        g[..., :3] = diff * (2.0 / n)
        self.backward(gbuffer, g)
        self.adam(lr)
        return (diff * diff).sum() / n

    def params_numpy(self):
        return self.params.detach().cpu().numpy().astype(np.float64)


def masked_mse(x, target, mask):
    m = mask.to(torch.float32)
    return float((((x[..., :3] - target[..., :3]) ** 2).sum(-1) * m).sum() / m.sum().clamp_min(1.0))


class LearnedPixelDenoiser:
    """The reflection denoise (denoise::system, RaytracedReflection.cpp:496-520) with the scaled
    a-trous chain, used only once validate() has shown it closer to the converged target than the
    reference chain on the caller's samples; before that, every frame is the reference path
    unchanged. Both keep the reference's temporal structure: reprojection, the passes, and pass 1's
    output fed back as the next frame's history (RaytracedReflection.cpp:488-491)."""

    def __init__(self, width, height, params=None, iterations=5, alpha=0.01, moments_alpha=0.2, radius=1,
                 phi_color=10.0, phi_normal=32.0, sigma_depth=1.0):
        from .svgf import ReflectionDenoiser
        self.W, self.H = width, height
        self.plain = ReflectionDenoiser(width, height, iterations=iterations, alpha=alpha, moments_alpha=moments_alpha,
                                        radius=radius, phi_color=phi_color, phi_normal=phi_normal,
                                        sigma_depth=sigma_depth)
        self.chain = PixelScaleChain(width, height, params=params, iterations=iterations, radius=radius,
                                     phi_color=phi_color, phi_normal=phi_normal, sigma_depth=sigma_depth)
        self.use_learned = False
        self.validation = None

    def set_camera(self, view_proj_inv, prev_view_proj, camera_pos, camera_delta):
        self.plain.set_camera(view_proj_inv, prev_view_proj, camera_pos, camera_delta)

    def validate(self, learned_mse, reference_mse):
        """Chooses the learned chain iff its error on the caller's samples is below the reference's."""
        self.validation = {"learned_mse": float(learned_mse), "reference_mse": float(reference_mse)}
        self.use_learned = self.validation["learned_mse"] < self.validation["reference_mse"]
        return self.validation

    def denoise(self, cur, prev, noisy):
        """One frame; returns the output tensor."""
        if not self.use_learned:
            return self.plain.denoise(cur, prev, noisy)
        d = self.plain
        check(lib().luxgi_svgf_reflection_reproject(
            C.byref(d.rp), self.W, self.H, abi.STORAGE_FP32, C.byref(gbuffer_struct(cur)), C.byref(gbuffer_struct(prev)),
            ptr(noisy), ptr(d.outputs[1]), ptr(d.moments[1 - d.ping_pong]), ptr(d.outputs[0]),
            ptr(d.moments[d.ping_pong]), ptr(d.tile_map), stream_handle()), "luxgi_svgf_reflection_reproject")
        out = self.chain.forward(cur, d.outputs[0], d.moments[d.ping_pong], d.tile_map,
                                 history_out=d.outputs[1] if self.chain.iterations > 1 else None, keep_features=False)
        d.ping_pong = 1 - d.ping_pong
        d.result = out
        return out


# ---------------------------------------------------------------------------------------------
# Training data rendered by the method, training, and the evaluation against plain SVGF
# ---------------------------------------------------------------------------------------------
TARGET_FRAMES = 256  # the blue-noise sampler repeats after 256 frames (BlueNoise.glsl:14-16)


class ReflectionView:
    """One static view of a seeded config-1-style scene (scene.scene_small_varied): the G-buffer from
    SDF primary rays, the SDF reflection frames (one GGX-sampled ray per pixel, a new blue-noise
    sample each frame, luxgi_sdf_reflection), and the converged target, the mean of TARGET_FRAMES
    frames. Misses are sky (pbr.w = 1); hits get roughness in the filtered band."""

    def __init__(self, seed, width, height, res=64, keep_raw=0):
        from . import scene
        from .ddgi import DeviceScene, bake_sdf, build_mip, to_device
        from .producers import BlueNoiseTables
        from .svgf import make_gbuffer
        self.seed, self.W, self.H = seed, width, height
        prims = scene.scene_small_varied(seed)
        sdf = bake_sdf(prims, res)
        mip = build_mip(sdf, res, max_distance=2.0)
        g = abi.make_sdf_data(resolution=res)
        atlas = scene.build_atlas(prims, resolution=512, tile=32, seed=seed, threshold=1.05 * 2.0 / res)
        self.scene = DeviceScene(g, sdf, mip, atlas)
        rng = np.random.default_rng(seed)
        eye = rng.uniform([-0.6, -0.3, -0.6], [0.6, 0.5, 0.7])
        while True:
            look = rng.uniform(-0.6, 0.6, size=3)
            if np.linalg.norm(look - eye) > 0.4:
                break
        VP, _, eye, _ = scene.camera_pair(width, height, seed=seed, eye=tuple(eye), target=tuple(look))
        self.eye, self.VP, self.VPI = eye, VP, np.linalg.inv(VP)
        prims_dev = to_device(np.frombuffer(prims.tobytes(), dtype=np.uint8))
        gb, _, _ = make_gbuffer(prims_dev, len(prims), g, sdf, res, self.VPI, VP, width, height, seed=seed,
                                with_mask=False)
        miss = gb["depth"] == 1.0
        gb["pbr"][..., 3] = torch.where(miss, torch.ones_like(gb["depth"]), gb["pbr"][..., 3])
        gb["pbr"][..., 1] = torch.where(miss, gb["pbr"][..., 1], gb["pbr"][..., 1].clamp(0.1, 0.4))
        self.gbuffer = gb
        self.mask = filtered_mask(gb)
        self.noise = BlueNoiseTables(np.stack([rng.permutation(256) for _ in range(4)], axis=1).astype(np.uint8),
                                     rng.integers(0, 256, (128, 128, 4), dtype=np.uint8))
        p = abi.SdfReflectionParams()
        p.trim, p.roughDDGIIntensity, p.approximateWithDDGI = 1.0, 1.0, 0
        p.cameraPosition[:] = [float(eye[0]), float(eye[1]), float(eye[2]), 1.0]
        p.viewProjInv = abi.mat4_colmajor(self.VPI)
        p.skyColor[:] = [0.6, 0.7, 0.9, 1.0]
        self.params = p
        self.raw = []
        self.target = torch.zeros((height, width, 4), device="cuda")
        for f in range(TARGET_FRAMES):
            out = self.frame(f)
            self.target += out
            if f < keep_raw:
                self.raw.append(out)
        self.target /= TARGET_FRAMES

    def frame(self, f):
        """Raw frame f (1 spp)."""
        from .producers import sdf_reflection
        self.params.numFrames = f
        out = torch.zeros((self.H, self.W, 4), device="cuda")
        sdf_reflection(self.params, self.noise, self.gbuffer, self.scene, out)
        return out

    def camera(self):
        """(viewProjInv, prevViewProj, cameraPos, cameraDelta) of the static camera."""
        return self.VPI, self.VP, self.eye, (0.0, 0.0, 0.0)


def reference_sequence_samples(view, keep, iterations=5):
    """The reference denoiser over the view's frames 0..max(keep): the a-trous input (the
    reprojection's colour + variance), the moments and the tile map of each kept frame, with the
    reference chain's filtered history fed back."""
    from .svgf import ReflectionDenoiser
    d = ReflectionDenoiser(view.W, view.H, iterations=iterations)
    d.set_camera(*view.camera())
    out = []
    for f in range(max(keep) + 1):
        d.denoise(view.gbuffer, view.gbuffer, view.raw[f] if f < len(view.raw) else view.frame(f))
        if f in keep:
            out.append((d.outputs[0].clone(), d.moments[1 - d.ping_pong].clone(), d.tile_map.clone()))
    return out


def train(chain, views, samples, steps, lr=3e-3, seed=0, log=None):
    """`steps` Adam steps, each on one (view, sample) drawn at random; returns the losses."""
    rng = np.random.default_rng(seed)
    losses = []
    for s in range(steps):
        i = int(rng.integers(0, len(samples)))
        vi, (inp, mom, tiles) = samples[i]
        v = views[vi]
        losses.append(chain.train_step(v.gbuffer, inp, mom, tiles, v.target, v.mask, lr=lr))
        if log and (s + 1) % 200 == 0:
            log(f"step {s + 1}: loss {float(torch.stack(losses[-200:]).mean()):.6f}")
    return [float(x) for x in torch.stack(losses).cpu()]


def evaluate(params, view, frames, iterations=5):
    """Plain SVGF (ReflectionDenoiser) and the learned denoiser over the same frames of `view`,
    each feeding back its own filtered history: mean squared error against the target over the
    filtered pixels, averaged over `frames`, with the reprojection's unfiltered output beside."""
    from .svgf import ReflectionDenoiser
    plain = ReflectionDenoiser(view.W, view.H, iterations=iterations)
    plain.set_camera(*view.camera())
    learned = LearnedPixelDenoiser(view.W, view.H, params=params, iterations=iterations)
    learned.set_camera(*view.camera())
    learned.use_learned = True
    e = {"plain": [], "learned": [], "unfiltered": []}
    for f in range(max(frames) + 1):
        raw = view.raw[f] if f < len(view.raw) else view.frame(f)
        a = plain.denoise(view.gbuffer, view.gbuffer, raw)
        if f in frames:
            e["plain"].append(masked_mse(a, view.target, view.mask))
            e["unfiltered"].append(masked_mse(plain.outputs[0], view.target, view.mask))
        b = learned.denoise(view.gbuffer, view.gbuffer, raw)
        if f in frames:
            e["learned"].append(masked_mse(b, view.target, view.mask))
    return {k: float(np.mean(v)) for k, v in e.items()}


def time_chains(params, view, reps=20, iterations=5):
    """1080p-style cost: the learned chain (scales + scaled passes) against the reference chain on
    one of the view's frames, HIP events on the current stream, ms per chain."""
    from .svgf import ReflectionDenoiser
    d = ReflectionDenoiser(view.W, view.H, iterations=iterations)
    d.set_camera(*view.camera())
    d.denoise(view.gbuffer, view.gbuffer, view.raw[0] if view.raw else view.frame(0))
    inp, mom, tiles = d.outputs[0].clone(), d.moments[1 - d.ping_pong].clone(), d.tile_map.clone()
    chain = PixelScaleChain(view.W, view.H, params=params, iterations=iterations)
    gb = gbuffer_struct(view.gbuffer)
    bufs = [torch.empty_like(inp), torch.empty_like(inp)]

    def reference():
        src = inp
        for i in range(iterations):
            ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
            check(lib().luxgi_svgf_reflection_atrous(C.byref(ap), view.W, view.H, abi.STORAGE_FP32, C.byref(gb),
                                                     ptr(tiles), ptr(src), ptr(bufs[i & 1]), None, stream_handle()),
                  "luxgi_svgf_reflection_atrous")
            src = bufs[i & 1]

    out = {}
    for name, fn in (("reference_chain_ms", reference),
                     ("learned_chain_ms", lambda: chain.forward(view.gbuffer, inp, mom, tiles, keep_features=False)),
                     ("scales_ms", lambda: chain.compute_scales(view.gbuffer, inp, mom, keep_features=False))):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        out[name] = ev[0].elapsed_time(ev[1]) / reps
    return out


def run(train_scenes=200, eval_scenes=8, width=960, height=540, eval_width=1920, eval_height=1080, steps=3000,
        lr=3e-3, keep=(2, 5, 11), eval_frames=(4, 6, 8, 10, 12), seed=0, log=None):
    """Renders the training views, trains a network from initial_params(seed) and evaluates it on
    held-out scenes at the evaluation size against plain SVGF; returns a report (and the params)."""
    import time
    t0 = time.perf_counter()
    views, samples = [], []
    for k in range(train_scenes):
        v = ReflectionView(1000 + k, width, height, keep_raw=max(keep) + 1)
        for s in reference_sequence_samples(v, keep):
            samples.append((len(views), s))
        v.raw = []  # the samples hold what training needs
        views.append(v)
        if log and (k + 1) % 25 == 0:
            log(f"rendered {k + 1} training views")
    t1 = time.perf_counter()
    chain = PixelScaleChain(width, height, params=initial_params(seed))
    losses = train(chain, views, samples, steps, lr=lr, seed=seed, log=log)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    params = chain.params_numpy()
    del views, samples, chain
    torch.cuda.empty_cache()
    evals = []
    timing = None
    for k in range(eval_scenes):
        v = ReflectionView(5000 + k, eval_width, eval_height, keep_raw=max(eval_frames) + 1)
        evals.append(evaluate(params, v, eval_frames))
        if timing is None:
            timing = time_chains(params, v)
        if log:
            log(f"eval view {k}: {evals[-1]}")
        del v
    t3 = time.perf_counter()
    mean = {key: float(np.mean([e[key] for e in evals])) for key in evals[0]}
    report = {"train_scenes": train_scenes, "train_size": [width, height], "train_samples": train_scenes * len(keep),
              "steps": steps, "lr": lr, "target_frames": TARGET_FRAMES, "eval_scenes": eval_scenes,
              "eval_size": [eval_width, eval_height], "eval_frames": list(eval_frames),
              "loss_first_100": float(np.mean(losses[:100])), "loss_last_100": float(np.mean(losses[-100:])),
              "eval_mse": mean, "eval_per_scene": evals,
              "learned_over_plain": mean["learned"] / mean["plain"], "timing_ms": timing,
              "render_s": t1 - t0, "train_s": t2 - t1, "eval_s": t3 - t2}
    return report, params


if __name__ == "__main__":
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="train and evaluate the per-pixel edge-stopping network")
    ap.add_argument("--train-scenes", type=int, default=200)
    ap.add_argument("--eval-scenes", type=int, default=8)
    ap.add_argument("--width", type=int, default=960)
    ap.add_argument("--height", type=int, default=540)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=None, help="JSON report path")
    ap.add_argument("--params-out", default=None, help=".npy path for the trained parameters")
    a = ap.parse_args()
    rep, prm = run(train_scenes=a.train_scenes, eval_scenes=a.eval_scenes, width=a.width, height=a.height,
                   steps=a.steps, lr=a.lr, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(rep), flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(rep, fh, indent=1)
    if a.params_out:
        np.save(a.params_out, prm)

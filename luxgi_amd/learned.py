"""Learned edge-stopping weights for the reflection a-trous chain (libluxgi.so).

The chain is the reference's reflection a-trous filter (Reflection/DenoiseAtrous.comp:63-134,
driven by denoise::system, Engine/Raytrace/RaytracedReflection.cpp:419-494: `iterations` passes
of step 1 << i) with each tap's weight exp(-|dl|/phi - wZ) * wNormal * k (EdgeStop.glsl:24-40)
replaced by exp(MLP(features)), one MLP (81 parameters) shared by the passes and told the pass by
its `level` input. Training runs on the GPU: luxgi_svgf_learned_atrous keeps each pixel's weight
sum and centre quantities, luxgi_svgf_learned_atrous_backward returns dL/d(input) and accumulates dL/d(parameters),
luxgi_adam_step updates them. The loss is the mean squared error of the chain's output against a
converged target over the pixels the filter touches (non-sky, roughness in (0.05, 0.45)).

Parameter layout: W1 [10][6] row-major, b1 [10], w2 [10], b2. Features per tap:
(min(|l_p - l_q| / phi_p, 16), exp(-|z_p - z_q| / sigmaDepth), max(0, n_p.n_q), k(dx) k(dy),
level, min(sqrt(var_p), 4)).
"""
import ctypes as C
import math

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .ddgi import ptr, stream_handle
from .svgf import gbuffer_struct

N_PARAMS = abi.LEARNED_ATROUS_PARAMS
N_IN, N_HIDDEN = 6, 10


def reference_like_weights(phi_normal=32.0, live_units=True, seed=0):
    """MLP parameters whose weight equals the reference's for radius 1 where the normals agree.

    o = -f0 - f1 + g(f3) - phi_normal * relu(1 - f2), where g is the piecewise-linear function through
    (4/9, ln 4/9), (2/3, ln 2/3), (1, 0): exp(g(k)) = k at the three kernel weights of radius 1, and
    -phi_normal (1 - n.n') is the first-order term of ln(max(0, n.n')^phi_normal) at n.n' = 1. With
    live_units the five unused hidden units get small random weights (a positive bias keeps them
    active, so training can move them); without, they are zero and the MLP is exactly the above."""
    W1 = np.zeros((N_HIDDEN, N_IN))
    b1 = np.zeros(N_HIDDEN)
    w2 = np.zeros(N_HIDDEN)
    s1 = math.log(1.5) / (2.0 / 9.0)
    s2 = 3.0 * math.log(1.5)
    W1[0, 0], w2[0] = 1.0, -1.0
    W1[1, 1], w2[1] = 1.0, -1.0
    W1[2, 3], b1[2], w2[2] = 1.0, -4.0 / 9.0, s1
    W1[3, 3], b1[3], w2[3] = 1.0, -2.0 / 3.0, s2 - s1
    W1[4, 2], b1[4], w2[4] = -1.0, 1.0, -float(phi_normal)
    if live_units:
        rng = np.random.default_rng(seed)
        W1[5:] = rng.normal(0.0, 0.1, (N_HIDDEN - 5, N_IN))
        b1[5:] = 0.1
        w2[5:] = rng.normal(0.0, 0.01, N_HIDDEN - 5)
    b2 = math.log(4.0 / 9.0)
    return np.concatenate([W1.reshape(-1), b1, w2, [b2]]).astype(np.float64)


def filtered_mask(gbuffer):
    """Pixels the a-trous pass filters: not sky (pbr.w != 1) and roughness in (0.05, 0.45)."""
    pbr = gbuffer["pbr"]
    r = pbr[..., 1]
    return (pbr[..., 3] != 1.0) & (r > 0.05) & (r < 0.45)


class LearnedAtrous:
    """`iterations` learned passes over RGBA32F (rgb, variance) images on the current stream."""

    def __init__(self, width, height, weights=None, iterations=4, radius=1, phi_color=10.0, sigma_depth=1.0):
        if weights is None:
            weights = reference_like_weights()
        weights = np.asarray(weights, dtype=np.float32).reshape(-1)
        if weights.size != N_PARAMS:
            raise ValueError(f"expected {N_PARAMS} parameters, got {weights.size}")
        if iterations < 1:
            raise ValueError("iterations must be >= 1")
        self.W, self.H = width, height
        self.iterations = iterations
        self.weights = torch.from_numpy(weights.copy()).cuda()
        self.grad = torch.zeros(N_PARAMS, dtype=torch.float32, device="cuda")
        self.m = torch.zeros_like(self.grad)
        self.v = torch.zeros_like(self.grad)
        self.step = 0
        self.params = []
        for i in range(iterations):
            p = abi.LearnedAtrousParams()
            p.radius, p.stepSize = radius, 1 << i
            p.phiColor, p.sigmaDepth = phi_color, sigma_depth
            p.level = i / (iterations - 1) if iterations > 1 else 0.0
            self.params.append(p)
        z = lambda: torch.zeros((height, width, 4), dtype=torch.float32, device="cuda")  # noqa: E731
        self.images = [z() for _ in range(iterations)]  # outputs of the passes
        # per pass and pixel: (1 + tap weights, luminance, 1 / phi, min(sqrt(var), 4)) for the backward
        self.records = [z() for _ in range(iterations)]
        self.grads = [z(), z()]
        self._inputs = [None] * iterations

    def forward(self, gbuffer, inp):
        """Runs the chain on `inp` ([H][W][4] FP32); returns the last pass's output (an internal
        buffer, valid until the next call). Keeps what backward() needs."""
        if inp.shape != (self.H, self.W, 4) or inp.dtype != torch.float32:
            raise ValueError("input must be an [H][W][4] float32 tensor")
        gb = gbuffer_struct(gbuffer)
        src = inp.contiguous()
        for i, p in enumerate(self.params):
            self._inputs[i] = src
            check(lib().luxgi_svgf_learned_atrous(C.byref(p), ptr(self.weights), self.W, self.H, C.byref(gb), ptr(src),
                                                  ptr(self.images[i]), ptr(self.records[i]), stream_handle()),
                  "luxgi_svgf_learned_atrous")
            src = self.images[i]
        return src

    def backward(self, gbuffer, grad_out):
        """Back-propagates dL/d(output rgb) through the last forward(); adds into self.grad and
        returns dL/d(input rgb) (an internal buffer)."""
        gb = gbuffer_struct(gbuffer)
        g = grad_out.contiguous()
        for i in reversed(range(self.iterations)):
            dst = self.grads[i & 1]
            if dst.data_ptr() == g.data_ptr():
                dst = self.grads[1 - (i & 1)]
            check(lib().luxgi_svgf_learned_atrous_backward(
                C.byref(self.params[i]), ptr(self.weights), self.W, self.H, C.byref(gb), ptr(self._inputs[i]),
                ptr(self.images[i]), ptr(self.records[i]), ptr(g), ptr(dst), ptr(self.grad), stream_handle()),
                "luxgi_svgf_learned_atrous_backward")
            g = dst
        return g

    def adam(self, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step on the accumulated gradient (cleared afterwards)."""
        self.step += 1
        check(lib().luxgi_adam_step(ptr(self.weights), ptr(self.grad), ptr(self.m), ptr(self.v), N_PARAMS, lr, beta1,
                                    beta2, eps, self.step, stream_handle()), "luxgi_adam_step")

    def train_step(self, gbuffer, inp, target, mask, lr=1e-2):
        """One step on one image: forward, MSE over `mask` (bool [H][W] tensor) and the rgb channels,
        backward, Adam. Returns the loss as a device scalar."""
        out = self.forward(gbuffer, inp)
        mf = mask.to(torch.float32)[..., None]
        n = float(max(int(mask.sum().item()), 1))
        diff = (out[..., :3] - target[..., :3]) * mf
        g = torch.zeros((self.H, self.W, 4), dtype=torch.float32, device="cuda")
        g[..., :3] = diff * (2.0 / n)
        self.backward(gbuffer, g)
        self.adam(lr)
        return (diff * diff).sum() / n

    def weights_numpy(self):
        return self.weights.detach().cpu().numpy().astype(np.float64)


def reference_chain(gbuffer, inp, tile_map, iterations=4, radius=1, phi_color=10.0, phi_normal=32.0,
                    sigma_depth=1.0):
    """The reference's a-trous chain on an FP32 (rgb, variance) image: `iterations` passes of
    luxgi_svgf_reflection_atrous with step 1 << i and the reprojection's tile map."""
    H, W = inp.shape[:2]
    gb = gbuffer_struct(gbuffer)
    bufs = [torch.empty_like(inp), torch.empty_like(inp)]
    src = inp.contiguous()
    for i in range(iterations):
        ap = abi.ReflectionAtrousParams(radius, 1 << i, phi_color, phi_normal, sigma_depth, 1)
        check(lib().luxgi_svgf_reflection_atrous(C.byref(ap), W, H, abi.STORAGE_FP32, C.byref(gb), ptr(tile_map),
                                                 ptr(src), ptr(bufs[i & 1]), None, stream_handle()),
              "luxgi_svgf_reflection_atrous")
        src = bufs[i & 1]
    return src


def masked_mse(x, target, mask):
    m = mask.to(torch.float32)
    return float((((x[..., :3] - target[..., :3]) ** 2).sum(-1) * m).sum() / m.sum().clamp_min(1.0))


class LearnedReflectionDenoiser:
    """The reflection denoise (denoise::system, RaytracedReflection.cpp:496-520) with the learned
    a-trous chain in place of the reference's, used only when validate() has shown it closer to the
    converged target than the reference's chain on the caller's samples; otherwise (and before any
    validation) every frame goes through the reference's path unchanged. The learned chain keeps the
    reference's temporal structure: reprojection, then the passes, with the output of pass 1 copied
    back as the next frame's history."""

    def __init__(self, width, height, weights=None, iterations=4, alpha=0.01, moments_alpha=0.2, radius=1,
                 phi_color=10.0, phi_normal=32.0, sigma_depth=1.0):
        from .svgf import ReflectionDenoiser
        self.W, self.H = width, height
        self.plain = ReflectionDenoiser(width, height, iterations=iterations, alpha=alpha,
                                        moments_alpha=moments_alpha, radius=radius, phi_color=phi_color,
                                        phi_normal=phi_normal, sigma_depth=sigma_depth)
        self.learned = LearnedAtrous(width, height, weights=weights, iterations=iterations, radius=radius,
                                     phi_color=phi_color, sigma_depth=sigma_depth)
        self.chain = dict(iterations=iterations, radius=radius, phi_color=phi_color, phi_normal=phi_normal,
                          sigma_depth=sigma_depth)
        self.use_learned = False
        self.validation = None

    def set_camera(self, view_proj_inv, prev_view_proj, camera_pos, camera_delta):
        self.plain.set_camera(view_proj_inv, prev_view_proj, camera_pos, camera_delta)

    def validate(self, samples):
        """samples: (gbuffer, reprojected input, tile map, converged target, mask) tuples. Chooses the
        learned chain iff its mean squared error over the samples is below the reference chain's."""
        lm, pm = [], []
        for gb, inp, tiles, target, mask in samples:
            lm.append(masked_mse(self.learned.forward(gb, inp), target, mask))
            pm.append(masked_mse(reference_chain(gb, inp, tiles, **self.chain), target, mask))
        self.validation = {"learned_mse": float(np.mean(lm)), "reference_mse": float(np.mean(pm))}
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
        out = self.learned.forward(cur, d.outputs[0])
        if self.learned.iterations > 1:
            d.outputs[1].copy_(self.learned.images[1])  # the history copy after iteration 1 (:488-491)
        d.ping_pong = 1 - d.ping_pong
        d.result = out
        return out

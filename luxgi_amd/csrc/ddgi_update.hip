// ddgi_update.hip -- DDGI probe blend (irradiance + depth) and border copy for gfx950.
//
// Reference passes:
//   IrradianceProbeUpdate.comp / DepthProbeUpdate.comp  Shaders/DDGI/ProbeUpdate.glsl:53-152
//   IrradianceBorderUpdate.comp / DepthBorderUpdate.comp Shaders/DDGI/BorderUpdate.glsl:136-156
//   host dispatches Engine/DDGI/DDGIRenderer.cpp:347-375 (update), :416-438 (border)
//
// Design. The reference gives every probe one workgroup and loops over its rays
// in 64-ray LDS chunks, evaluating a weight per (texel, ray). Every probe of a
// frame traces the same ray directions (GISDFRays.comp:73 depends only on the
// ray id and the frame's rotation), so the weight matrix W[texel][ray] is the
// same for all probes and the blend is a contraction W . B with B = per-ray
// (radiance | distance, distance^2). The library builds W once per call from
// ray row 0 (build_weights), and the update kernel checks every staged ray's
// direction against row 0 bit for bit: chunks that match use W, chunks that do
// not fall back to evaluating the weights from their own directions, which is
// the reference's arithmetic. The border copy (BorderUpdate.glsl) is fused into
// the epilogue: the finished interior texels of a probe are staged in LDS and
// the 4n+4 border texels are written from there.
#include "common.hpp"

#include <algorithm>
#include <mutex>
#include <type_traits>

namespace luxgi {

constexpr int kProbesPerBlock = 4;
constexpr int kRayChunk = 256;
constexpr int kMaxSide = 16;  // oct side <= 16 (reference: 8 and 16)
constexpr float kFltEps = 0.00000001f;  // ProbeUpdate.glsl:51

struct UpdateArgs {
    int px, py, pz, P, R, rayWidth;
    int ni, nd;
    int Wi, Wd;
    float maxDistance, sharpness, hysteresis, invGamma;
    int firstFrame, doIrr, doDepth, fuseBorder, forcePerProbe, half, bf16;
};

// Shared weights W[r][t] (thresholded at FLT_EPS as ProbeUpdate.glsl:93) and
// their row sums S[t], from the directions of ray row 0.
template <class SR>
__global__ void build_weights_kernel(const SR* __restrict__ dirDist, int R, int n, int depth, float sharpness,
                                     float* __restrict__ W, float* __restrict__ S) {
    const int T = n * n;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const f3 td = texel_direction(t % n, t / n, n);
    float total = 0.0f;
    for (int r = 0; r < R; ++r) {
        float4 d = Texel<SR>::ld4(dirDist, (size_t)r);
        float w = fmaxf(0.0f, dot3(td, mk(d.x, d.y, d.z)));
        if (depth) w = powf(w, sharpness);
        if (w >= kFltEps) total += w;
        else w = 0.0f;
        W[(size_t)r * T + t] = w;
    }
    S[t] = total;
}

// Rays of one chunk staged in LDS: a = (r, g, b, depth distance), b = (dx, dy, dz, distance^2)
struct RayLds {
    float4 a[kProbesPerBlock][kRayChunk];
    float4 b[kProbesPerBlock][kRayChunk];
};
struct OutLds {
    float4 irr[kProbesPerBlock][kMaxSide * kMaxSide];
    float2 dep[kProbesPerBlock][kMaxSide * kMaxSide];
};
union UpdateLds {
    RayLds rays;
    OutLds out;
};

// accumulate one texel of one map for one probe over a chunk
template <int NC>
__device__ __forceinline__ void accumulate(const UpdateLds& lds, int g, int t, int T, int r0, int nr, bool shared,
                                           const float* __restrict__ W, f3 texelDir, float sharpness, float* acc,
                                           float& total) {
    if (shared) {
        const float* w = W + (size_t)r0 * T + t;
        if (NC == 3) {
#pragma unroll 4
            for (int r = 0; r < nr; ++r) {
                float wr = w[(size_t)r * T];
                float4 a = lds.rays.a[g][r];
                acc[0] += a.x * wr;
                acc[1] += a.y * wr;
                acc[2] += a.z * wr;
                total += wr;
            }
        } else {
#pragma unroll 4
            for (int r = 0; r < nr; ++r) {
                float wr = w[(size_t)r * T];
                float dr = lds.rays.a[g][r].w, d2 = lds.rays.b[g][r].w;
                acc[0] += dr * wr;
                acc[1] += d2 * wr;
                total += wr;
            }
        }
    } else {
        // per-probe weights (ProbeUpdate.glsl:66-103)
        for (int r = 0; r < nr; ++r) {
            float4 a = lds.rays.a[g][r], b = lds.rays.b[g][r];
            float w = fmaxf(0.0f, dot3(texelDir, mk(b.x, b.y, b.z)));
            if (NC == 2) w = powf(w, sharpness);
            if (w >= kFltEps) {
                if (NC == 3) {
                    acc[0] += a.x * w;
                    acc[1] += a.y * w;
                    acc[2] += a.z * w;
                } else {
                    acc[0] += a.w * w;
                    acc[1] += b.w * w;
                }
                total += w;
            }
        }
    }
}

// one group of kProbesPerBlock probes starting at p0 (the body of probe_update_kernel)
template <class SR, class SA>
__device__ __forceinline__ void update_group(const UpdateArgs& A, int p0, const SR* __restrict__ radiance,
                                             const SR* __restrict__ dirDist, const SA* __restrict__ prevIrr,
                                             const SA* __restrict__ prevDepth, SA* __restrict__ outIrr,
                                             SA* __restrict__ outDepth, const float* __restrict__ Wi,
                                             const float* __restrict__ Wd, const uint8_t* __restrict__ onlyFlagged,
                                             UpdateLds& lds, int* sShared) {
    const int tid = threadIdx.x;
    auto wanted = [&](int p) { return p < A.P && (!onlyFlagged || onlyFlagged[p] != 0); };
    const int Ti = A.ni * A.ni, Td = A.nd * A.nd;
    const int pxy = A.px * A.py;

    // Work items. Depth: thread t < Td owns texel t of all probes of the block.
    // Irradiance: with Ti <= 64 each wave owns one probe (t = tid % 64); else like depth.
    const bool irrPerWave = Ti <= 64;
    const int ti = irrPerWave ? (tid & 63) : tid;
    const int giBegin = irrPerWave ? (tid >> 6) : 0;
    const int giEnd = irrPerWave ? giBegin + 1 : kProbesPerBlock;
    const bool irrActive = A.doIrr && ti < Ti;
    const bool depActive = A.doDepth && tid < Td;
    const f3 dirI = texel_direction(ti % A.ni, ti / A.ni, A.ni);
    const f3 dirD = texel_direction(tid % A.nd, tid / A.nd, A.nd);

    float accI[kProbesPerBlock][3], totI[kProbesPerBlock];
    float accD[kProbesPerBlock][2], totD[kProbesPerBlock];
#pragma unroll
    for (int g = 0; g < kProbesPerBlock; ++g) {
        accI[g][0] = accI[g][1] = accI[g][2] = totI[g] = 0.0f;
        accD[g][0] = accD[g][1] = totD[g] = 0.0f;
    }

    for (int r0 = 0; r0 < A.R; r0 += kRayChunk) {
        const int nr = min(kRayChunk, A.R - r0);
        if (tid < kProbesPerBlock) sShared[tid] = A.forcePerProbe ? 0 : 1;
        __syncthreads();
        // stage rays of the chunk (ProbeUpdate.glsl:53-64, 75-78) and compare directions with row 0
        for (int i = tid; i < kProbesPerBlock * nr; i += blockDim.x) {
            const int g = i / nr, r = i % nr;
            const int p = p0 + g;
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = make_float4(1.f, 0.f, 0.f, 0.f);
            if (p < A.P) {
                const size_t e = (size_t)p * A.rayWidth + r0 + r;
                float4 dd = Texel<SR>::ld4(dirDist, e);
                float4 rr = A.doIrr ? Texel<SR>::ld4(radiance, e) : make_float4(0.f, 0.f, 0.f, 0.f);
                float dist = fminf(A.maxDistance, dd.w - 0.01f);
                if (dist == -1.0f) dist = A.maxDistance;
                a = make_float4(rr.x, rr.y, rr.z, dist);
                b = make_float4(dd.x, dd.y, dd.z, dist * dist);
                float4 d0 = Texel<SR>::ld4(dirDist, (size_t)(r0 + r));
                if (d0.x != dd.x || d0.y != dd.y || d0.z != dd.z) sShared[g] = 0;
            }
            lds.rays.a[g][r] = a;
            lds.rays.b[g][r] = b;
        }
        __syncthreads();
        if (depActive) {
#pragma unroll
            for (int g = 0; g < kProbesPerBlock; ++g) {
                const bool sh = sShared[g] != 0;
                accumulate<2>(lds, g, tid, Td, r0, nr, sh, Wd, dirD, A.sharpness, accD[g], totD[g]);
            }
        }
        if (irrActive) {
            for (int g = giBegin; g < giEnd; ++g) {
                const bool sh = sShared[g] != 0;
                accumulate<3>(lds, g, ti, Ti, r0, nr, sh, Wi, dirI, A.sharpness, accI[g], totI[g]);
            }
        }
        __syncthreads();
    }

    // Epilogue: normalise, gamma, hysteresis, store interior, stage for the border.
    if (depActive) {
        const int lx = tid % A.nd, ly = tid / A.nd;
        for (int g = 0; g < kProbesPerBlock; ++g) {
            const int p = p0 + g;
            if (p >= A.P) break;
            if (!wanted(p)) continue;
            const float total = totD[g];
            float x = accD[g][0], y = accD[g][1];
            if (total > kFltEps) {
                float s = 1.0f / (2.0f * total);
                x *= s;
                y *= s;
            }
            const int cx = (p % pxy) * (A.nd + 2) + lx + 2, cy = (p / pxy) * (A.nd + 2) + ly + 2;
            const size_t e = (size_t)cy * A.Wd + cx;
            if (!A.firstFrame) {
                float2 pv = Texel<SA>::ld2(prevDepth, e);
                x = mixf(x, pv.x, A.hysteresis);
                y = mixf(y, pv.y, A.hysteresis);
            }
            Texel<SA>::st2(outDepth, e, make_float2(x, y));
            lds.out.dep[g][tid] = make_float2(x, y);
        }
    }
    if (irrActive) {
        const int lx = ti % A.ni, ly = ti / A.ni;
        for (int g = giBegin; g < giEnd; ++g) {
            const int p = p0 + g;
            if (p >= A.P) break;
            if (!wanted(p)) continue;
            const float total = totI[g];
            float r = accI[g][0], gg = accI[g][1], b = accI[g][2];
            if (total > kFltEps) {
                float s = 1.0f / (2.0f * total);
                r *= s;
                gg *= s;
                b *= s;
            }
            r = powf(r, A.invGamma);
            gg = powf(gg, A.invGamma);
            b = powf(b, A.invGamma);
            const int cx = (p % pxy) * (A.ni + 2) + lx + 2, cy = (p / pxy) * (A.ni + 2) + ly + 2;
            const size_t e = (size_t)cy * A.Wi + cx;
            if (!A.firstFrame) {
                float4 pv = Texel<SA>::ld4(prevIrr, e);
                r = mixf(r, pv.x, A.hysteresis);
                gg = mixf(gg, pv.y, A.hysteresis);
                b = mixf(b, pv.z, A.hysteresis);
            }
            float4 v = make_float4(r, gg, b, 1.0f);
            Texel<SA>::st4(outIrr, e, v);
            lds.out.irr[g][ti] = v;
        }
    }
    if (!A.fuseBorder) return;
    __syncthreads();
    // Border texels (BorderUpdate.glsl:25-133 rule): 4n+4 copies per probe from interior texels.
    for (int map = 0; map < 2; ++map) {
        if ((map == 0 && !A.doIrr) || (map == 1 && !A.doDepth)) continue;
        const int n = map == 0 ? A.ni : A.nd;
        const int nb = 4 * n + 4;
        for (int i = tid; i < kProbesPerBlock * nb; i += blockDim.x) {
            const int g = i / nb, k = i % nb;
            const int p = p0 + g;
            if (!wanted(p)) continue;
            int sx, sy, dx, dy;
            const int side = k / n, j = k % n + 1;
            if (k < 4 * n) {
                if (side == 0) { sx = n + 1 - j; sy = 1; dx = j; dy = 0; }
                else if (side == 1) { sx = n + 1 - j; sy = n; dx = j; dy = n + 1; }
                else if (side == 2) { sx = 1; sy = n + 1 - j; dx = 0; dy = j; }
                else { sx = n; sy = n + 1 - j; dx = n + 1; dy = j; }
            } else {
                const int c = k - 4 * n;
                const int cs[4][4] = {{n, n, 0, 0}, {1, n, n + 1, 0}, {n, 1, 0, n + 1}, {1, 1, n + 1, n + 1}};
                sx = cs[c][0]; sy = cs[c][1]; dx = cs[c][2]; dy = cs[c][3];
            }
            const int ox = (p % pxy) * (n + 2) + 1, oy = (p / pxy) * (n + 2) + 1;
            const int src = (sy - 1) * n + (sx - 1);
            if (map == 0) {
                Texel<SA>::st4(outIrr, (size_t)(oy + dy) * A.Wi + ox + dx, lds.out.irr[g][src]);
            } else {
                Texel<SA>::st2(outDepth, (size_t)(oy + dy) * A.Wd + ox + dx, lds.out.dep[g][src]);
            }
        }
    }
}

// Grid-stride over groups of kProbesPerBlock probes. Fix-up mode (onlyFlagged != null):
// only groups holding a flagged probe do any work, and only flagged probes are written.
template <class SR, class SA>
__global__ __launch_bounds__(256) void probe_update_kernel(UpdateArgs A, const SR* __restrict__ radiance,
                                                           const SR* __restrict__ dirDist,
                                                           const SA* __restrict__ prevIrr,
                                                           const SA* __restrict__ prevDepth, SA* __restrict__ outIrr,
                                                           SA* __restrict__ outDepth, const float* __restrict__ Wi,
                                                           const float* __restrict__ Wd,
                                                           const uint8_t* __restrict__ onlyFlagged) {
    __shared__ UpdateLds lds;
    __shared__ int sShared[kProbesPerBlock];
    const int groups = (A.P + kProbesPerBlock - 1) / kProbesPerBlock;
    for (int grp = blockIdx.x; grp < groups; grp += gridDim.x) {
        const int p0 = grp * kProbesPerBlock;
        if (onlyFlagged) {
            bool any = false;
            for (int g = 0; g < kProbesPerBlock; ++g) any |= p0 + g < A.P && onlyFlagged[p0 + g] != 0;
            if (!any) continue;  // uniform across the block
        }
        update_group<SR, SA>(A, p0, radiance, dirDist, prevIrr, prevDepth, outIrr, outDepth, Wi, Wd, onlyFlagged,
                             lds, sShared);
        __syncthreads();  // the next group reuses the LDS
    }
}

// The fix-up after the shared-direction blend: per-probe reference arithmetic for the probes flagged
// in `mismatch` (directions not those of ray row 0). A block covers 256 probes and leaves at once
// when none of them is flagged, the usual case.
template <class SR, class SA>
__global__ __launch_bounds__(256) void fixup_kernel(UpdateArgs A, const SR* __restrict__ radiance,
                                                    const SR* __restrict__ dirDist, const SA* __restrict__ prevIrr,
                                                    const SA* __restrict__ prevDepth, SA* __restrict__ outIrr,
                                                    SA* __restrict__ outDepth, const uint8_t* __restrict__ mismatch) {
    __shared__ UpdateLds lds;
    __shared__ int sShared[kProbesPerBlock];
    const int base = blockIdx.x * 256;
    const int p = base + (int)threadIdx.x;
    if (!__syncthreads_or(p < A.P && mismatch[p] != 0)) return;  // uniform across the block
    const int end = min(base + 256, A.P);
    for (int q0 = base; q0 < end; q0 += kProbesPerBlock) {
        bool any = false;
        for (int g = 0; g < kProbesPerBlock; ++g) any |= q0 + g < A.P && mismatch[q0 + g] != 0;
        if (!any) continue;  // uniform across the block
        update_group<SR, SA>(A, q0, radiance, dirDist, prevIrr, prevDepth, outIrr, outDepth, nullptr, nullptr,
                             mismatch, lds, sShared);
        __syncthreads();  // the next group reuses the LDS
    }
}

// Stand-alone border copy, in place (BorderUpdate.glsl:136-156).
template <class SA, int CH>
__global__ void border_kernel(SA* __restrict__ atlas, int W, int n, int pxy, int P) {
    const int p = blockIdx.x;
    if (p >= P) return;
    const int nb = 4 * n + 4;
    const int ox = (p % pxy) * (n + 2) + 1, oy = (p / pxy) * (n + 2) + 1;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        int sx, sy, dx, dy;
        const int side = k / n, j = k % n + 1;
        if (k < 4 * n) {
            if (side == 0) { sx = n + 1 - j; sy = 1; dx = j; dy = 0; }
            else if (side == 1) { sx = n + 1 - j; sy = n; dx = j; dy = n + 1; }
            else if (side == 2) { sx = 1; sy = n + 1 - j; dx = 0; dy = j; }
            else { sx = n; sy = n + 1 - j; dx = n + 1; dy = j; }
        } else {
            const int c = k - 4 * n;
            const int cs[4][4] = {{n, n, 0, 0}, {1, n, n + 1, 0}, {n, 1, 0, n + 1}, {1, 1, n + 1, n + 1}};
            sx = cs[c][0]; sy = cs[c][1]; dx = cs[c][2]; dy = cs[c][3];
        }
        const size_t s = (size_t)(oy + sy) * W + ox + sx, d = (size_t)(oy + dy) * W + ox + dx;
        if (CH == 4) Texel<SA>::st4(atlas, d, Texel<SA>::ld4(atlas, s));
        else Texel<SA>::st2(atlas, d, Texel<SA>::ld2(atlas, s));
    }
}

// ---------------------------------------------------------------------------
// MFMA form. With shared ray directions the blend of a block of 16 probes is
//   depth: C[Td x 16] = W_d[Td x R] . B_r[R x 16]  and  W_d . B_r2   (probe g, r and r^2)
//   irr:   C[Ti x 16] = W_i[Ti x R] . B_c[R x 16]  for c = R, G, B
// contracted on the 16x16x32 MFMA (probe_blend_map_kernel below), one map per launch. The
// epilogue normalises by 1/(2 S[t]), applies the gamma, stages the interiors in LDS and writes
// each probe's full bordered tile with the hysteresis blend. A probe whose staged directions
// differ from ray row 0 is flagged; probe_update_kernel then recomputes the flagged probes with
// their own directions.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int kMG = 16;    // probes per block

// directions of ray row 0 as FP32 (the values every probe's rays are compared with)
template <class SR>
__global__ void row0_kernel(const SR* __restrict__ dirDist, int R, float* __restrict__ dirs0) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) reinterpret_cast<float4*>(dirs0)[r] = Texel<SR>::ld4(dirDist, (size_t)r);
}

// source interior texel of bordered-tile texel (bx, by) of an n x n probe
// (BorderUpdate.glsl:25-133: edges mirror across the edge's centre, corners take the opposite corner)
__device__ __forceinline__ void border_source(int bx, int by, int n, int& sx, int& sy) {
    const bool inX = bx >= 1 && bx <= n, inY = by >= 1 && by <= n;
    if (inX && inY) { sx = bx; sy = by; }
    else if (inX) { sx = n + 1 - bx; sy = by == 0 ? 1 : n; }
    else if (inY) { sy = n + 1 - by; sx = bx == 0 ? 1 : n; }
    else { sx = bx == 0 ? n : 1; sy = by == 0 ? n : 1; }
}

// ---------------------------------------------------------------------------
// Blend on the 16x16x32 MFMA with FP32 accumulation, 16 probes per block, 32 rays per k-step,
// operands in FP16 (v_mfma_f32_16x16x32_f16) or, in the reduced-precision form on request, BF16
// (v_mfma_f32_16x16x32_bf16). Operand lanes: A[m][k] = W at texel m, ray k -- lane l holds rays
// 8(l/16) .. 8(l/16)+7 of texel row l%16, one 16-byte load from the texel-major weights;
// B[k][n] = ray data of probe n -- lane l holds the same 8 rays of probe l%16, one 16-byte LDS
// read; D[4(l/16)+v][l%16]. Depth moments enter as r/maxDistance and its square (both in [0, 1]).
//   SPLIT (the plain form): every operand x is carried as hi = fp16(x), lo = fp16(x - hi) and
//     each product as hi*hi + hi*lo + lo*hi: about 22 significant bits per operand, the dropped
//     lo*lo term and the FP16 subnormal floor sit below the FP32 rounding of the normalised sums
//     because all operands are O(1). Normalised by the sum of hi + lo.
//   !SPLIT (LUXGI_UPDATE_HALF): one FP16 (or BF16) rounding per operand, normalised by the sum
//     of the rounded weights, so a constant input stays constant.
// The ray data of each k-step is staged in LDS (double-buffered) by all 256 lanes; a probe whose
// directions differ from ray row 0 is flagged for the per-probe fix-up pass.
// ---------------------------------------------------------------------------
typedef _Float16 half8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
template <class T> struct Lp8;
template <> struct Lp8<_Float16> { typedef half8v V; };
template <> struct Lp8<__bf16> { typedef bf16x8v V; };
constexpr int kHK = 32;  // rays per k-step

// Shared weights of one map in the MFMA operand type T (hi part, and with `split` the lo part), texel
// major [tpad][Rpad], from ray row 0's directions, and their sums in kWChunk-ray chunks,
// Sp[chunk][tpad] (the blend adds the chunks in order). Block (tile, chunk): 16 texels x kWChunk rays.
constexpr int kWChunk = 64;

template <class T>
__device__ __forceinline__ void weights_block(int tile, int chunk, const float* __restrict__ dirs0, int R, int Rpad,
                                              int n, int tpad, int depth, float sharpness, int split,
                                              T* __restrict__ Whi, T* __restrict__ Wlo, float* __restrict__ Sp) {
    __shared__ float part[16][17];
    const int tl = threadIdx.x & 15, j = threadIdx.x >> 4;
    const int t = tile * 16 + tl;
    const bool valid = t < n * n;
    const f3 td = valid ? texel_direction(t % n, t / n, n) : mk(0.f, 0.f, 1.f);
    float s = 0.0f;
    const int r1 = min(Rpad, (chunk + 1) * kWChunk);
    for (int r = chunk * kWChunk + j; r < r1; r += 16) {
        float w = 0.0f;
        if (valid && r < R) {
            w = fmaxf(0.0f, dot3(td, mk(dirs0[r * 4 + 0], dirs0[r * 4 + 1], dirs0[r * 4 + 2])));
            if (depth) w = powf(w, sharpness);
            if (!(w >= kFltEps)) w = 0.0f;
        }
        const T h = (T)w;
        Whi[(size_t)t * Rpad + r] = h;
        if (split) {
            const T l = (T)(w - (float)h);
            Wlo[(size_t)t * Rpad + r] = l;
            s += (float)h + (float)l;
        } else {
            s += (float)h;
        }
    }
    part[j][tl] = s;
    __syncthreads();
    if (j == 0) {
        float tot = 0.0f;
        for (int k = 0; k < 16; ++k) tot += part[k][tl];
        Sp[(size_t)chunk * tpad + t] = tot;
    }
}

// both maps' weights in one launch: blockIdx.z 0 irradiance, 1 depth; x: 16-texel tile, y: ray chunk
template <class T>
__global__ __launch_bounds__(256) void weights_lowp_kernel(const float* __restrict__ dirs0, int R, int Rpad, int ni,
                                                           int tpI, int nd, int tpD, float sharpness, int split,
                                                           int doIrr, int doDepth, T* __restrict__ Wih,
                                                           T* __restrict__ Wil, float* __restrict__ Sih,
                                                           T* __restrict__ Wdh, T* __restrict__ Wdl,
                                                           float* __restrict__ Sdh) {
    const bool depth = blockIdx.z == 1;
    if (depth ? !doDepth : !doIrr) return;  // uniform per block
    const int tp = depth ? tpD : tpI;
    if ((int)blockIdx.x * 16 >= tp) return;
    weights_block<T>(blockIdx.x, blockIdx.y, dirs0, R, Rpad, depth ? nd : ni, tp, depth ? 1 : 0, sharpness, split,
                     depth ? Wdh : Wih, depth ? Wdl : Wil, depth ? Sdh : Sih);
}

__device__ __forceinline__ floatx4 mfma32(half8v a, half8v b, floatx4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc, 0, 0, 0);
}
__device__ __forceinline__ floatx4 mfma32(bf16x8v a, bf16x8v b, floatx4 acc) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
}
template <bool SPLIT, class V>
__device__ __forceinline__ floatx4 mfma_lp(V ah, V al, V bh, V bl, floatx4 acc) {
    acc = mfma32(ah, bh, acc);
    if (SPLIT) {
        acc = mfma32(ah, bl, acc);
        acc = mfma32(al, bh, acc);
    }
    return acc;
}

// ---------------------------------------------------------------------------
// The same contraction with one map per launch (MAP 0: depth moments r, r^2; MAP 1: irradiance
// R, G, B), as the reference dispatches DepthProbeUpdate.comp and IrradianceProbeUpdate.comp
// separately (DDGIRenderer.cpp:347-375). Each launch holds only its own map's accumulators and
// stages only its own output tile in LDS (depth 32 KB, irradiance 16 KB), so four 256-lane blocks
// fit a CU and the 1,024 blocks of a 16,384-probe volume are resident at once. The depth launch
// reads only the direction/distance rays and checks the directions against row 0; the irradiance
// launch reads only the radiance rays (and the directions too when the depth map is not
// updated in the same call, CHECK).
// ---------------------------------------------------------------------------
constexpr int kWB = 4;  // previous-atlas loads each lane keeps in flight in the tile writer

// The rows are padded so that the LDS accesses of a wave spread over the banks: a probe's ray row
// is 40 elements (80 B), so the 16-byte B-operand reads of 8 consecutive lanes (probes col..col+7,
// one ray octet) start 20 dwords apart and cover all 32 banks once; an output row is one texel
// longer than the tile, so the epilogue's stores of the 16 probes of one texel fall in different
// banks instead of all in one.
constexpr int kHKPad = kHK + 8;
template <int CH, class T>
struct MapStage {
    T b[2][CH][kMG][kHKPad];  // [hi / lo][channel][probe][ray]
};
template <int MAP>
struct MapOut;
template <>
struct MapOut<0> {
    float2 v[kMG][256 + 1];
    __device__ __forceinline__ float4 get(int g, int t) const { return make_float4(v[g][t].x, v[g][t].y, 0.f, 0.f); }
};
template <>
struct MapOut<1> {
    float4 v[kMG][64 + 1];
    __device__ __forceinline__ float4 get(int g, int t) const { return v[g][t]; }
};
template <int MAP, class T>
union MapLds {
    MapStage<MAP == 0 ? 2 : 3, T> st[2];
    MapOut<MAP> out;
};

// Bordered tiles of the kMG staged interiors of one map, blended with the previous atlas
// (ProbeUpdate.glsl:144-145) and, when fuseBorder, with the border copy of BorderUpdate.glsl.
// Consecutive lanes cover consecutive texels of an atlas row (the 16 probes' tiles are adjacent),
// and each lane keeps kWB previous-atlas loads in flight before it blends and stores.
template <class SA, int CH, class Staged>
__device__ __forceinline__ void write_staged(const UpdateArgs& A, const Staged& o, int p0, int n, int W,
                                             const SA* __restrict__ prev, SA* __restrict__ out) {
    const int pw = n + 2, pxy = A.px * A.py;
    const int total = kMG * pw * pw;
    const int tstride = CH == 4 ? 64 : 256;
    // the 16 probes' tile origins in the atlas, computed once per block
    __shared__ int sOrg[2][kMG];
    if (threadIdx.x < kMG) {
        const int p = p0 + (int)threadIdx.x;
        sOrg[0][threadIdx.x] = (p % pxy) * pw + 1;
        sOrg[1][threadIdx.x] = (p / pxy) * pw + 1;
    }
    __syncthreads();
    // i / pw as a multiply-high by m = floor(2^32 / pw) + 1: exact while i * pw < 2^32, and here
    // i < 16 pw^2 + 1024 with pw <= 66
    const uint64_t mPw = (1ull << 32) / (uint32_t)pw + 1;
    for (int i0 = threadIdx.x; i0 < total; i0 += 256 * kWB) {
        float4 pv[kWB];
        size_t dst[kWB];
        int src[kWB];
        bool ok[kWB];
#pragma unroll
        for (int u = 0; u < kWB; ++u) {
            const int i = i0 + u * 256;
            const int q = (int)(((uint64_t)(uint32_t)i * mPw) >> 32);  // i / pw
            const int bx = i - q * pw, g = q & (kMG - 1), by = q / kMG;
            const int p = p0 + g;
            const bool interior = bx >= 1 && bx <= n && by >= 1 && by <= n;
            ok[u] = i < total && p < A.P && (interior || A.fuseBorder);
            int sx = 0, sy = 0;
            border_source(bx, by, n, sx, sy);
            src[u] = g * tstride + (sy - 1) * n + (sx - 1);
            const int ox = sOrg[0][g], oy = sOrg[1][g];
            dst[u] = (size_t)(oy + by) * W + ox + bx;
            pv[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (ok[u] && !A.firstFrame) {
                const size_t ps = (size_t)(oy + sy) * W + ox + sx;
                if (CH == 4) pv[u] = Texel<SA>::ld4(prev, ps);
                else {
                    const float2 t = Texel<SA>::ld2(prev, ps);
                    pv[u] = make_float4(t.x, t.y, 0.f, 0.f);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kWB; ++u) {
            if (!ok[u]) continue;
            const int g = src[u] / tstride, t = src[u] % tstride;
            float4 v = o.get(g, t);
            if (CH == 4) {
                v.w = 1.0f;
                if (!A.firstFrame) {
                    v.x = mixf(v.x, pv[u].x, A.hysteresis);
                    v.y = mixf(v.y, pv[u].y, A.hysteresis);
                    v.z = mixf(v.z, pv[u].z, A.hysteresis);
                }
                Texel<SA>::st4(out, dst[u], v);
            } else {
                if (!A.firstFrame) {
                    v.x = mixf(v.x, pv[u].x, A.hysteresis);
                    v.y = mixf(v.y, pv[u].y, A.hysteresis);
                }
                Texel<SA>::st2(out, dst[u], make_float2(v.x, v.y));
            }
        }
    }
}

// one k-step's ray data of this map: thread t stages rays 2(t%16), 2(t%16)+1 of probe t/16
struct MapPrefetch {
    float4 a[2];   // dd (depth) or radiance (irradiance)
    float4 d[2];   // dd, irradiance with CHECK only
    float4 r0[2];  // ray row 0's directions of the same rays (depth, or CHECK), loaded with them
};

template <int MAP, bool CHECK, class SR>
__device__ __forceinline__ void map_load(const UpdateArgs& A, int k0, int p0, const SR* __restrict__ radiance,
                                         const SR* __restrict__ dirDist, const float* __restrict__ dirs0,
                                         MapPrefetch& f) {
    const int g = threadIdx.x >> 4, rp = (threadIdx.x & 15) * 2, p = p0 + g;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = k0 + rp + q;
        f.a[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (CHECK) f.d[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (MAP == 0 || CHECK) f.r0[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (p < A.P && r < A.R) {
            const size_t e = (size_t)p * A.rayWidth + r;
            f.a[q] = Texel<SR>::ld4(MAP == 0 ? dirDist : radiance, e);
            if (CHECK) f.d[q] = Texel<SR>::ld4(dirDist, e);
            if (MAP == 0 || CHECK) f.r0[q] = reinterpret_cast<const float4*>(dirs0)[r];
        }
    }
}

template <int MAP, bool CHECK, bool SPLIT, class T>
__device__ __forceinline__ void map_store(const UpdateArgs& A, int k0, int p0, const float* __restrict__ dirs0,
                                          const MapPrefetch& f, MapStage<MAP == 0 ? 2 : 3, T>& st, int* sMis) {
    constexpr int CH = MAP == 0 ? 2 : 3;
    const int g = threadIdx.x >> 4, rp = (threadIdx.x & 15) * 2, p = p0 + g;
    const float invMax = 1.0f / A.maxDistance;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = k0 + rp + q;
        float v[CH];
        if (MAP == 0) {
            v[0] = 0.0f;
            v[1] = 0.0f;
        } else {
            v[0] = f.a[q].x;
            v[1] = f.a[q].y;
            v[CH - 1] = f.a[q].z;
        }
        if (p < A.P && r < A.R) {
            if (MAP == 0) {
                float dist = fminf(A.maxDistance, f.a[q].w - 0.01f);  // ProbeUpdate.glsl:75-78
                if (dist == -1.0f) dist = A.maxDistance;
                v[0] = dist * invMax;
                v[1] = v[0] * v[0];
            }
            if (MAP == 0 || CHECK) {
                const float4 d = MAP == 0 ? f.a[q] : f.d[q];
                if (d.x != f.r0[q].x || d.y != f.r0[q].y || d.z != f.r0[q].z) sMis[g] = 1;
            }
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const T h = (T)v[c];
            st.b[0][c][g][rp + q] = h;
            if (SPLIT) st.b[1][c][g][rp + q] = (T)(v[c] - (float)h);
        }
    }
}

// one block of 16 probes starting at p0 (the body of the single-volume and batched map kernels)
template <class SR, class SA, class T, bool SPLIT, int MAP, bool CHECK>
__device__ __forceinline__ void blend_map_block(const UpdateArgs& A, int p0, const SR* __restrict__ radiance,
                                                const SR* __restrict__ dirDist, const SA* __restrict__ prev,
                                                SA* __restrict__ out, const float* __restrict__ dirs0,
                                                const T* __restrict__ Wh, const T* __restrict__ Wl,
                                                const float* __restrict__ Sh, uint8_t* __restrict__ mismatch,
                                                int tpad, int Rpad, MapLds<MAP, T>& lds, int* sMis) {
    typedef typename Lp8<T>::V V;
    constexpr int CH = MAP == 0 ? 2 : 3;
    constexpr int NJ = MAP == 0 ? 4 : 1;  // m-tiles per wave: depth up to 16 tiles, irradiance up to 4
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int col = lane & 15, kq = lane >> 4;
    const int Mt = tpad / 16;
    if (tid < kMG) sMis[tid] = 0;
    floatx4 acc[NJ][CH];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[j][c] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int nSteps = Rpad / kHK;
    MapPrefetch f;
    map_load<MAP, CHECK, SR>(A, 0, p0, radiance, dirDist, dirs0, f);
    __syncthreads();  // sMis initialised
    map_store<MAP, CHECK, SPLIT, T>(A, 0, p0, dirs0, f, lds.st[0], sMis);
    __syncthreads();
    for (int s = 0; s < nSteps; ++s) {
        const int k0 = s * kHK;
        const bool more = s + 1 < nSteps;
        if (more) map_load<MAP, CHECK, SR>(A, k0 + kHK, p0, radiance, dirDist, dirs0, f);
        const MapStage<CH, T>& st = lds.st[s & 1];
        const int ko = k0 + 8 * kq;
        V bh[CH], bl[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            bh[c] = *reinterpret_cast<const V*>(&st.b[0][c][col][8 * kq]);
            bl[c] = SPLIT ? *reinterpret_cast<const V*>(&st.b[1][c][col][8 * kq]) : bh[c];
        }
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int m = wave + 4 * j;
            if (m < Mt) {
                const size_t off = (size_t)(m * 16 + col) * Rpad + ko;
                const V ah = *reinterpret_cast<const V*>(Wh + off);
                const V al = SPLIT ? *reinterpret_cast<const V*>(Wl + off) : ah;
#pragma unroll
                for (int c = 0; c < CH; ++c) acc[j][c] = mfma_lp<SPLIT>(ah, al, bh[c], bl[c], acc[j][c]);
            }
        }
        if (more) map_store<MAP, CHECK, SPLIT, T>(A, k0 + kHK, p0, dirs0, f, lds.st[(s + 1) & 1], sMis);
        __syncthreads();
    }
    // interiors: normalise by 1/(2 S), undo the depth scaling or gamma-encode (ProbeUpdate.glsl:133-141)
    const int n = MAP == 0 ? A.nd : A.ni;
    const int nT = n * n;
    const int rowq = kq * 4;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int m = wave + 4 * j;
        if (m >= Mt) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int t = m * 16 + rowq + v;
            if (t >= nT) continue;
            float sm = 0.0f;
            for (int c = 0; c * kWChunk < Rpad; ++c) sm += Sh[(size_t)c * tpad + t];
            const float sc = sm > kFltEps ? 1.0f / (2.0f * sm) : 1.0f;
            if constexpr (MAP == 0) {
                lds.out.v[col][t] = make_float2(acc[j][0][v] * sc * A.maxDistance,
                                                acc[j][1][v] * sc * (A.maxDistance * A.maxDistance));
            } else {
                lds.out.v[col][t] = make_float4(powf(acc[j][0][v] * sc, A.invGamma),
                                                powf(acc[j][1][v] * sc, A.invGamma),
                                                powf(acc[j][CH - 1][v] * sc, A.invGamma), 1.0f);
            }
        }
    }
    __syncthreads();
    if ((MAP == 0 || CHECK) && tid < kMG && p0 + tid < A.P) mismatch[p0 + tid] = (uint8_t)sMis[tid];
    if constexpr (MAP == 0) write_staged<SA, 2>(A, lds.out, p0, A.nd, A.Wd, prev, out);
    else write_staged<SA, 4>(A, lds.out, p0, A.ni, A.Wi, prev, out);
}

template <class SR, class SA, class T, bool SPLIT, int MAP, bool CHECK>
__global__ __launch_bounds__(256, 4) void probe_blend_map_kernel(
    UpdateArgs A, const SR* __restrict__ radiance, const SR* __restrict__ dirDist, const SA* __restrict__ prev,
    SA* __restrict__ out, const float* __restrict__ dirs0, const T* __restrict__ Wh, const T* __restrict__ Wl,
    const float* __restrict__ Sh, uint8_t* __restrict__ mismatch, int tpad, int Rpad) {
    __shared__ MapLds<MAP, T> lds;
    __shared__ int sMis[kMG];
    blend_map_block<SR, SA, T, SPLIT, MAP, CHECK>(A, blockIdx.x * kMG, radiance, dirDist, prev, out, dirs0, Wh, Wl, Sh,
                                                  mismatch, tpad, Rpad, lds, sMis);
}

// ---------------------------------------------------------------------------
// Several volumes in one launch per pass (no counterpart in the reference, which dispatches the
// update per volume): each volume's arguments and workspace pointers in a device descriptor, the
// blocks of all volumes' 16-probe groups in one grid (block -> volume by the groups' prefix,
// uniform per block). Per volume the passes are those of luxgi_ddgi_probe_update.
// ---------------------------------------------------------------------------
struct UpdateVolumeDesc {
    UpdateArgs A;
    const void* rad;
    const void* dd;
    const void* prevI;
    const void* prevD;
    void* outI;
    void* outD;
    float* dirs0;
    void *Wih, *Wil, *Wdh, *Wdl;
    float *Sih, *Sdh;
    uint8_t* mismatch;
    int tpI, tpD, Rp;
    int groupStart;  // first 16-probe group of this volume in the batch grid
};

__device__ __forceinline__ int batch_update_volume(const UpdateVolumeDesc* __restrict__ d, int n, int group) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].groupStart <= group) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class SR>
__global__ void row0_batched_kernel(const UpdateVolumeDesc* __restrict__ d) {
    const UpdateVolumeDesc& D = d[blockIdx.y];
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < D.A.R)
        reinterpret_cast<float4*>(D.dirs0)[r] = Texel<SR>::ld4(static_cast<const SR*>(D.dd), (size_t)r);
}

// blockIdx.y: volume, blockIdx.z: map (0 irradiance, 1 depth), x: (16-texel tile, ray chunk)
template <class T>
__global__ __launch_bounds__(256) void weights_batched_kernel(const UpdateVolumeDesc* __restrict__ d, int split) {
    const UpdateVolumeDesc& D = d[blockIdx.y];
    const bool depth = blockIdx.z == 1;
    if (depth ? !D.A.doDepth : !D.A.doIrr) return;
    const int tp = depth ? D.tpD : D.tpI;
    const int chunks = (D.Rp + kWChunk - 1) / kWChunk;
    const int tile = blockIdx.x / chunks, chunk = blockIdx.x % chunks;
    if (tile * 16 >= tp) return;
    weights_block<T>(tile, chunk, D.dirs0, D.A.R, D.Rp, depth ? D.A.nd : D.A.ni, tp, depth ? 1 : 0, D.A.sharpness,
                     split, static_cast<T*>(depth ? D.Wdh : D.Wih), static_cast<T*>(depth ? D.Wdl : D.Wil),
                     depth ? D.Sdh : D.Sih);
}

template <class SR, class SA, class T, bool SPLIT, int MAP, bool CHECK>
__global__ __launch_bounds__(256, 4) void probe_blend_map_batched_kernel(const UpdateVolumeDesc* __restrict__ d, int n) {
    __shared__ MapLds<MAP, T> lds;
    __shared__ int sMis[kMG];
    const UpdateVolumeDesc& D = d[batch_update_volume(d, n, blockIdx.x)];
    const int p0 = (blockIdx.x - D.groupStart) * kMG;
    if (p0 >= D.A.P) return;  // uniform per block
    if constexpr (MAP == 0)
        blend_map_block<SR, SA, T, SPLIT, 0, CHECK>(D.A, p0, static_cast<const SR*>(D.rad), static_cast<const SR*>(D.dd),
                                                    static_cast<const SA*>(D.prevD), static_cast<SA*>(D.outD), D.dirs0,
                                                    static_cast<const T*>(D.Wdh), static_cast<const T*>(D.Wdl), D.Sdh,
                                                    D.mismatch, D.tpD, D.Rp, lds, sMis);
    else
        blend_map_block<SR, SA, T, SPLIT, 1, CHECK>(D.A, p0, static_cast<const SR*>(D.rad), static_cast<const SR*>(D.dd),
                                                    static_cast<const SA*>(D.prevI), static_cast<SA*>(D.outI), D.dirs0,
                                                    static_cast<const T*>(D.Wih), static_cast<const T*>(D.Wil), D.Sih,
                                                    D.mismatch, D.tpI, D.Rp, lds, sMis);
}

// the per-probe fix-up of flagged probes (directions not those of row 0), 4 probes at a time
template <class SR, class SA>
__global__ __launch_bounds__(256) void fixup_batched_kernel(const UpdateVolumeDesc* __restrict__ d, int n) {
    __shared__ UpdateLds lds;
    __shared__ int sShared[kProbesPerBlock];
    const UpdateVolumeDesc& D = d[batch_update_volume(d, n, blockIdx.x)];
    const int p0 = (blockIdx.x - D.groupStart) * kMG;
    const int tid = threadIdx.x;
    if (!__syncthreads_or(tid < kMG && p0 + tid < D.A.P && D.mismatch[p0 + tid] != 0)) return;  // uniform
    UpdateArgs F = D.A;
    F.forcePerProbe = 1;
    for (int q0 = p0; q0 < min(p0 + kMG, F.P); q0 += kProbesPerBlock) {
        bool any = false;
        for (int g = 0; g < kProbesPerBlock; ++g) any |= q0 + g < F.P && D.mismatch[q0 + g] != 0;
        if (!any) continue;  // uniform across the block
        update_group<SR, SA>(F, q0, static_cast<const SR*>(D.rad), static_cast<const SR*>(D.dd),
                             static_cast<const SA*>(D.prevI), static_cast<const SA*>(D.prevD),
                             static_cast<SA*>(D.outI), static_cast<SA*>(D.outD), nullptr, nullptr, D.mismatch, lds,
                             sShared);
        __syncthreads();  // the next group reuses the LDS
    }
}

static int check_uniform(const LuxgiDDGIUniform* u) {
    if (!u) return LUXGI_ERR_INVALID_ARGUMENT;
    const int px = u->probeCounts[0], py = u->probeCounts[1], pz = u->probeCounts[2];
    if (px <= 0 || py <= 0 || pz <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    const int ni = u->irradianceProbeSideLength, nd = u->depthProbeSideLength;
    if (ni < 1 || nd < 1) return LUXGI_ERR_INVALID_ARGUMENT;
    if (u->irradianceTextureWidth != (ni + 2) * px * py + 2 || u->irradianceTextureHeight != (ni + 2) * pz + 2)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (u->depthTextureWidth != (nd + 2) * px * py + 2 || u->depthTextureHeight != (nd + 2) * pz + 2)
        return LUXGI_ERR_INVALID_ARGUMENT;
    return LUXGI_OK;
}

static size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace luxgi

using namespace luxgi;

extern "C" int luxgi_ddgi_fill_atlas_sizes(LuxgiDDGIUniform* u) {
    if (!u) return LUXGI_ERR_INVALID_ARGUMENT;
    const int px = u->probeCounts[0], py = u->probeCounts[1], pz = u->probeCounts[2];
    u->irradianceTextureWidth = (u->irradianceProbeSideLength + 2) * px * py + 2;
    u->irradianceTextureHeight = (u->irradianceProbeSideLength + 2) * pz + 2;
    u->depthTextureWidth = (u->depthProbeSideLength + 2) * px * py + 2;
    u->depthTextureHeight = (u->depthProbeSideLength + 2) * pz + 2;
    return LUXGI_OK;
}

// workspace: row-0 directions, W and S for both maps (16-row padded), mismatch flags, and the
// texel-major FP16 weights with the sums of their rounded values (reduced-precision blend)
struct UpdateWorkspace {
    float *dirs0, *Wi, *Si, *Wd, *Sd;
    uint8_t* mismatch;
    _Float16 *Wih, *Wil, *Wdh, *Wdl;  // hi / lo parts
    float *Sih, *Sdh;
};

static int pad32(int r) { return (r + 31) & ~31; }

static int pad16(int t) { return (t + 15) & ~15; }

static size_t update_workspace_layout(const LuxgiDDGIUniform* u, char* base, UpdateWorkspace* w) {
    const size_t R = (size_t)(u->raysPerProbe > 0 ? u->raysPerProbe : 0);
    const size_t P = (size_t)u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    const size_t Ti = pad16(u->irradianceProbeSideLength * u->irradianceProbeSideLength);
    const size_t Td = pad16(u->depthProbeSideLength * u->depthProbeSideLength);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    char* d0 = take(R * 16);
    char* wi = take(R * Ti * 4);
    char* si = take(Ti * 4);
    char* wd = take(R * Td * 4);
    char* sd = take(Td * 4);
    char* mm = take(P);
    const size_t Rp = (size_t)pad32((int)R);
    const size_t chunks = (Rp + kWChunk - 1) / kWChunk;
    char* wih = take(Rp * Ti * 2);
    char* wil = take(Rp * Ti * 2);
    char* sih = take(chunks * Ti * 4);
    char* wdh = take(Rp * Td * 2);
    char* wdl = take(Rp * Td * 2);
    char* sdh = take(chunks * Td * 4);
    if (w) {
        w->Wih = reinterpret_cast<_Float16*>(wih);
        w->Wil = reinterpret_cast<_Float16*>(wil);
        w->Sih = reinterpret_cast<float*>(sih);
        w->Wdh = reinterpret_cast<_Float16*>(wdh);
        w->Wdl = reinterpret_cast<_Float16*>(wdl);
        w->Sdh = reinterpret_cast<float*>(sdh);
        w->dirs0 = reinterpret_cast<float*>(d0);
        w->Wi = reinterpret_cast<float*>(wi);
        w->Si = reinterpret_cast<float*>(si);
        w->Wd = reinterpret_cast<float*>(wd);
        w->Sd = reinterpret_cast<float*>(sd);
        w->mismatch = reinterpret_cast<uint8_t*>(mm);
    }
    return off;
}

extern "C" size_t luxgi_ddgi_probe_update_workspace_size(const LuxgiDDGIUniform* u) {
    if (!u) return 0;
    return update_workspace_layout(u, nullptr, nullptr);
}

// weights of both maps and the per-map MFMA blends with operand type T (FP16 or BF16); dirs0: ray
// row 0's directions as float4 per ray
template <class SR, class SA, class T, bool SPLIT>
static int launch_maps(const UpdateArgs& A, const SR* rad, const SR* dd, const SA* pI, const SA* pD, SA* oI, SA* oD,
                       const UpdateWorkspace& w, const float* dirs0, int tpI, int tpD, int Rp, hipStream_t st) {
    T* Wih = reinterpret_cast<T*>(w.Wih);
    T* Wil = reinterpret_cast<T*>(w.Wil);
    T* Wdh = reinterpret_cast<T*>(w.Wdh);
    T* Wdl = reinterpret_cast<T*>(w.Wdl);
    if (A.doIrr || A.doDepth) {
        const int tiles = std::max(A.doIrr ? tpI / 16 : 0, A.doDepth ? tpD / 16 : 0);
        hipLaunchKernelGGL(weights_lowp_kernel<T>, dim3(tiles, (Rp + kWChunk - 1) / kWChunk, 2), dim3(256), 0, st,
                           dirs0, A.R, Rp, A.ni, tpI, A.nd, tpD, A.sharpness, SPLIT ? 1 : 0, A.doIrr, A.doDepth, Wih,
                           Wil, w.Sih, Wdh, Wdl, w.Sdh);
        LUXGI_LAUNCH_CHECK();
    }
    const dim3 grid((A.P + kMG - 1) / kMG);
    if (A.doDepth) {
        hipLaunchKernelGGL((probe_blend_map_kernel<SR, SA, T, SPLIT, 0, false>), grid, dim3(256), 0, st, A, rad, dd, pD,
                           oD, dirs0, Wdh, Wdl, w.Sdh, w.mismatch, tpD, Rp);
        LUXGI_LAUNCH_CHECK();
    }
    if (A.doIrr) {
        // the directions were checked by the depth launch; alone, this launch checks them
        if (A.doDepth)
            hipLaunchKernelGGL((probe_blend_map_kernel<SR, SA, T, SPLIT, 1, false>), grid, dim3(256), 0, st, A, rad, dd,
                               pI, oI, dirs0, Wih, Wil, w.Sih, w.mismatch, tpI, Rp);
        else
            hipLaunchKernelGGL((probe_blend_map_kernel<SR, SA, T, SPLIT, 1, true>), grid, dim3(256), 0, st, A, rad, dd,
                               pI, oI, dirs0, Wih, Wil, w.Sih, w.mismatch, tpI, Rp);
        LUXGI_LAUNCH_CHECK();
    }
    return LUXGI_OK;
}

template <class SR, class SA>
static int launch_update(const UpdateArgs& A, const LuxgiRayBuffers* rays, const LuxgiProbeAtlas* prev,
                         const LuxgiProbeAtlas* out, const UpdateWorkspace& w, hipStream_t st) {
    const SR* rad = static_cast<const SR*>(rays->radiance);
    const SR* dd = static_cast<const SR*>(rays->directionDistance);
    const SA* pI = static_cast<const SA*>(prev ? prev->irradiance : nullptr);
    const SA* pD = static_cast<const SA*>(prev ? prev->depth : nullptr);
    SA* oI = static_cast<SA*>(out->irradiance);
    SA* oD = static_cast<SA*>(out->depth);
    const bool mfma = !A.forcePerProbe && A.ni * A.ni <= 64 && A.nd * A.nd <= 256 && A.maxDistance > 0.0f;
    if (mfma) {
        const int tpI = pad16(A.ni * A.ni), tpD = pad16(A.nd * A.nd);
        // ray row 0's directions as float4 per ray: FP32 rays already hold them so (probe 0's rays
        // come first); FP16 rays are widened into the workspace
        const float* dirs0 = reinterpret_cast<const float*>(dd);
        if constexpr (!std::is_same<SR, float>::value) {
            hipLaunchKernelGGL(row0_kernel<SR>, dim3((A.R + 255) / 256), dim3(256), 0, st, dd, A.R, w.dirs0);
            LUXGI_LAUNCH_CHECK();
            dirs0 = w.dirs0;
        }
        const int Rp = pad32(A.R);
        int rc;
        if (A.bf16) rc = launch_maps<SR, SA, __bf16, false>(A, rad, dd, pI, pD, oI, oD, w, dirs0, tpI, tpD, Rp, st);
        else if (A.half) rc = launch_maps<SR, SA, _Float16, false>(A, rad, dd, pI, pD, oI, oD, w, dirs0, tpI, tpD, Rp, st);
        else rc = launch_maps<SR, SA, _Float16, true>(A, rad, dd, pI, pD, oI, oD, w, dirs0, tpI, tpD, Rp, st);
        if (rc != LUXGI_OK) return rc;
        // probes whose directions differ from row 0: reference arithmetic with their own directions
        UpdateArgs F = A;
        F.forcePerProbe = 1;
        hipLaunchKernelGGL((fixup_kernel<SR, SA>), dim3((A.P + 255) / 256), dim3(256), 0, st, F, rad, dd, pI, pD, oI,
                           oD, w.mismatch);
        LUXGI_LAUNCH_CHECK();
        return LUXGI_OK;
    }
    if (!A.forcePerProbe) {
        if (A.doIrr) {
            const int T = A.ni * A.ni;
            hipLaunchKernelGGL(build_weights_kernel<SR>, dim3((T + 255) / 256), dim3(256), 0, st, dd, A.R, A.ni, 0,
                               A.sharpness, w.Wi, w.Si);
            LUXGI_LAUNCH_CHECK();
        }
        if (A.doDepth) {
            const int T = A.nd * A.nd;
            hipLaunchKernelGGL(build_weights_kernel<SR>, dim3((T + 255) / 256), dim3(256), 0, st, dd, A.R, A.nd, 1,
                               A.sharpness, w.Wd, w.Sd);
            LUXGI_LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL((probe_update_kernel<SR, SA>), dim3((A.P + kProbesPerBlock - 1) / kProbesPerBlock), dim3(256), 0,
                       st, A, rad, dd, pI, pD, oI, oD, w.Wi, w.Wd, nullptr);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

// Scratch of luxgi_ddgi_probe_update_noalloc: the same workspace layout as the caller-provided
// one, held in a file-scope device array. Every GPU holds its own copy at its own address, found
// once per device with hipGetSymbolAddress; host code never passes the array's name to a kernel
// (on the host it is a stand-in, not a device address). Every part of the workspace is written
// by the update's own kernels before it is read, so nothing needs clearing between calls.
static __device__ __align__(256) char scratch[LUXGI_UPDATE_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratchLock;
void* g_scratchPtr[kMaxScratchDevices] = {};

// the scratch array's device address on the current device (found on the first call there)
int scratch_address(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LUXGI_ERR_HIP;
    if (dev < 0 || dev >= kMaxScratchDevices) return LUXGI_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> g(g_scratchLock);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LUXGI_ERR_HIP;
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return LUXGI_OK;
}
}  // namespace

extern "C" int luxgi_ddgi_probe_update_noalloc(const LuxgiDDGIUniform* u, const LuxgiRayBuffers* rays,
                                               const LuxgiProbeAtlas* prev, const LuxgiProbeAtlas* out,
                                               uint32_t firstFrame, uint32_t flags, void* stream) {
    if (!u) return LUXGI_ERR_INVALID_ARGUMENT;
    const size_t need = update_workspace_layout(u, nullptr, nullptr);
    if (need > (size_t)LUXGI_UPDATE_NOALLOC_SCRATCH_BYTES) return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc) return rc;
    return luxgi_ddgi_probe_update(u, rays, prev, out, firstFrame, flags, ws, need, stream);
}

// The checked arguments of one update (luxgi_ddgi_probe_update's rules); LUXGI_OK with nothing to do
// when neither map is requested (A.doIrr = A.doDepth = 0).
static int make_update_args(const LuxgiDDGIUniform* u, const LuxgiRayBuffers* rays, const LuxgiProbeAtlas* prev,
                            const LuxgiProbeAtlas* out, uint32_t firstFrame, uint32_t flags, UpdateArgs& A) {
    int rc = check_uniform(u);
    if (rc) return rc;
    if (!rays || !out || !rays->directionDistance) return LUXGI_ERR_INVALID_ARGUMENT;
    A = UpdateArgs{};
    A.px = u->probeCounts[0];
    A.py = u->probeCounts[1];
    A.pz = u->probeCounts[2];
    A.P = A.px * A.py * A.pz;
    A.R = u->raysPerProbe;
    A.rayWidth = rays->width;
    A.ni = u->irradianceProbeSideLength;
    A.nd = u->depthProbeSideLength;
    A.Wi = u->irradianceTextureWidth;
    A.Wd = u->depthTextureWidth;
    A.maxDistance = u->maxDistance;
    A.sharpness = u->sharpness;
    A.hysteresis = u->hysteresis;
    A.invGamma = 1.0f / u->ddgiGamma;
    A.firstFrame = firstFrame != 0;
    A.doIrr = (flags & LUXGI_UPDATE_IRRADIANCE) != 0;
    A.doDepth = (flags & LUXGI_UPDATE_DEPTH) != 0;
    A.fuseBorder = (flags & LUXGI_UPDATE_FUSE_BORDER) != 0;
    A.forcePerProbe = (flags & LUXGI_UPDATE_PER_PROBE) != 0;
    A.half = (flags & LUXGI_UPDATE_HALF) != 0;
    A.bf16 = (flags & LUXGI_UPDATE_BF16) != 0;
    if (A.half && !(u->maxDistance > 0.0f)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (A.bf16 && !A.half) return LUXGI_ERR_INVALID_ARGUMENT;  // BF16 operands are a reduced-precision form
    if (A.R <= 0 || A.R > rays->width || rays->height < A.P) return LUXGI_ERR_INVALID_ARGUMENT;
    if (A.ni > kMaxSide || A.nd > kMaxSide) return LUXGI_ERR_UNSUPPORTED;
    if (!A.doIrr && !A.doDepth) return LUXGI_OK;
    if (A.doIrr && (!out->irradiance || !rays->radiance)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (A.doDepth && !out->depth) return LUXGI_ERR_INVALID_ARGUMENT;
    if (!A.firstFrame && (!prev || (A.doIrr && !prev->irradiance) || (A.doDepth && !prev->depth)))
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (prev && !A.firstFrame && prev->storage != out->storage) return LUXGI_ERR_INVALID_ARGUMENT;
    return LUXGI_OK;
}

extern "C" int luxgi_ddgi_probe_update(const LuxgiDDGIUniform* u, const LuxgiRayBuffers* rays,
                                       const LuxgiProbeAtlas* prev, const LuxgiProbeAtlas* out, uint32_t firstFrame,
                                       uint32_t flags, void* workspace, size_t workspaceBytes, void* stream) {
    UpdateArgs A;
    const int rc = make_update_args(u, rays, prev, out, firstFrame, flags, A);
    if (rc) return rc;
    if (!A.doIrr && !A.doDepth) return LUXGI_OK;
    const size_t need = update_workspace_layout(u, nullptr, nullptr);
    if (!workspace || workspaceBytes < need) return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    UpdateWorkspace w;
    update_workspace_layout(u, static_cast<char*>(workspace), &w);
    hipStream_t st = as_stream(stream);
    const int rs = rays->storage, as = out->storage;
    if (rs == LUXGI_STORAGE_FP32 && as == LUXGI_STORAGE_FP32) return launch_update<float, float>(A, rays, prev, out, w, st);
    if (rs == LUXGI_STORAGE_FP16 && as == LUXGI_STORAGE_FP16)
        return launch_update<__half, __half>(A, rays, prev, out, w, st);
    if (rs == LUXGI_STORAGE_FP32 && as == LUXGI_STORAGE_FP16)
        return launch_update<float, __half>(A, rays, prev, out, w, st);
    if (rs == LUXGI_STORAGE_FP16 && as == LUXGI_STORAGE_FP32)
        return launch_update<__half, float>(A, rays, prev, out, w, st);
    return LUXGI_ERR_UNSUPPORTED;
}

// ------------------------------------------------------------------------- several volumes

static bool batchable(const UpdateArgs& A) {  // the shared-direction MFMA path
    return !A.forcePerProbe && A.ni * A.ni <= 64 && A.nd * A.nd <= 256 && A.maxDistance > 0.0f;
}

extern "C" size_t luxgi_ddgi_update_batch_descriptor_bytes(int32_t count) {
    if (count <= 0 || count > LUXGI_MAX_UPDATE_BATCH) return 0;
    return (size_t)count * sizeof(UpdateVolumeDesc);
}

extern "C" size_t luxgi_ddgi_update_batch_workspace_size(const LuxgiUpdateVolume* volumes, int32_t count) {
    if (!volumes || count <= 0 || count > LUXGI_MAX_UPDATE_BATCH) return 0;
    size_t total = 0;
    for (int k = 0; k < count; ++k) total += align256(update_workspace_layout(&volumes[k].uniform, nullptr, nullptr));
    return total;
}

extern "C" int luxgi_ddgi_update_batch_prepare(const LuxgiUpdateVolume* volumes, int32_t count, uint32_t flags,
                                               void* workspace, size_t workspaceBytes, void* descriptors,
                                               LuxgiUpdateBatch* batch) {
    if (!volumes || !descriptors || !batch || count <= 0 || count > LUXGI_MAX_UPDATE_BATCH)
        return LUXGI_ERR_INVALID_ARGUMENT;
    const size_t need = luxgi_ddgi_update_batch_workspace_size(volumes, count);
    if (!workspace || workspaceBytes < need) return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    UpdateVolumeDesc* out = static_cast<UpdateVolumeDesc*>(descriptors);
    char* base = static_cast<char*>(workspace);  // device address: only offsets are taken here
    int groups = 0, maxR = 0, maxTiles = 0, maxChunks = 0;
    for (int k = 0; k < count; ++k) {
        const LuxgiUpdateVolume& v = volumes[k];
        UpdateArgs A;
        const int rc = make_update_args(&v.uniform, &v.rays, &v.prev, &v.out, v.firstFrame, flags, A);
        if (rc) return rc;
        if (!A.doIrr && !A.doDepth) return LUXGI_ERR_INVALID_ARGUMENT;
        if (!batchable(A)) return LUXGI_ERR_UNSUPPORTED;
        if (v.rays.storage != volumes[0].rays.storage || v.out.storage != volumes[0].out.storage)
            return LUXGI_ERR_INVALID_ARGUMENT;
        UpdateWorkspace w;
        const size_t bytes = update_workspace_layout(&v.uniform, base, &w);
        base += align256(bytes);
        UpdateVolumeDesc& D = out[k];
        D = UpdateVolumeDesc{};
        D.A = A;
        D.rad = v.rays.radiance;
        D.dd = v.rays.directionDistance;
        D.prevI = v.prev.irradiance;
        D.prevD = v.prev.depth;
        D.outI = v.out.irradiance;
        D.outD = v.out.depth;
        D.dirs0 = w.dirs0;
        D.Wih = w.Wih;
        D.Wil = w.Wil;
        D.Wdh = w.Wdh;
        D.Wdl = w.Wdl;
        D.Sih = w.Sih;
        D.Sdh = w.Sdh;
        D.mismatch = w.mismatch;
        D.tpI = pad16(A.ni * A.ni);
        D.tpD = pad16(A.nd * A.nd);
        D.Rp = pad32(A.R);
        D.groupStart = groups;
        groups += (A.P + kMG - 1) / kMG;
        maxR = std::max(maxR, A.R);
        maxTiles = std::max(maxTiles, std::max(D.tpI, D.tpD) / 16);
        maxChunks = std::max(maxChunks, (D.Rp + kWChunk - 1) / kWChunk);
    }
    batch->count = count;
    batch->groups = groups;
    batch->flags = (int32_t)flags;
    batch->rayStorage = volumes[0].rays.storage;
    batch->atlasStorage = volumes[0].out.storage;
    batch->maxRays = maxR;
    batch->maxWeightBlocks = maxTiles * maxChunks;
    batch->pad = 0;
    return LUXGI_OK;
}

template <class SR, class SA, class T, bool SPLIT>
static int launch_batched(const LuxgiUpdateBatch& b, const UpdateVolumeDesc* d, hipStream_t st) {
    const bool doIrr = (b.flags & LUXGI_UPDATE_IRRADIANCE) != 0, doDepth = (b.flags & LUXGI_UPDATE_DEPTH) != 0;
    hipLaunchKernelGGL(row0_batched_kernel<SR>, dim3((b.maxRays + 255) / 256, b.count), dim3(256), 0, st, d);
    LUXGI_LAUNCH_CHECK();
    hipLaunchKernelGGL(weights_batched_kernel<T>, dim3(b.maxWeightBlocks, b.count, 2), dim3(256), 0, st, d,
                       SPLIT ? 1 : 0);
    LUXGI_LAUNCH_CHECK();
    if (doDepth) {
        hipLaunchKernelGGL((probe_blend_map_batched_kernel<SR, SA, T, SPLIT, 0, false>), dim3(b.groups), dim3(256), 0,
                           st, d, b.count);
        LUXGI_LAUNCH_CHECK();
    }
    if (doIrr) {
        if (doDepth)
            hipLaunchKernelGGL((probe_blend_map_batched_kernel<SR, SA, T, SPLIT, 1, false>), dim3(b.groups), dim3(256),
                               0, st, d, b.count);
        else
            hipLaunchKernelGGL((probe_blend_map_batched_kernel<SR, SA, T, SPLIT, 1, true>), dim3(b.groups), dim3(256),
                               0, st, d, b.count);
        LUXGI_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((fixup_batched_kernel<SR, SA>), dim3(b.groups), dim3(256), 0, st, d, b.count);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

template <class SR, class SA>
static int launch_batched_forms(const LuxgiUpdateBatch& b, const UpdateVolumeDesc* d, hipStream_t st) {
    if (b.flags & LUXGI_UPDATE_BF16) return launch_batched<SR, SA, __bf16, false>(b, d, st);
    if (b.flags & LUXGI_UPDATE_HALF) return launch_batched<SR, SA, _Float16, false>(b, d, st);
    return launch_batched<SR, SA, _Float16, true>(b, d, st);
}

extern "C" int luxgi_ddgi_probe_update_batched(const LuxgiUpdateBatch* batch, const void* descriptors, void* stream) {
    if (!batch || !descriptors || batch->count <= 0 || batch->count > LUXGI_MAX_UPDATE_BATCH || batch->groups <= 0 ||
        batch->maxRays <= 0 || batch->maxWeightBlocks <= 0)
        return LUXGI_ERR_INVALID_ARGUMENT;
    const UpdateVolumeDesc* d = static_cast<const UpdateVolumeDesc*>(descriptors);
    hipStream_t st = as_stream(stream);
    const int rs = batch->rayStorage, as = batch->atlasStorage;
    if (rs == LUXGI_STORAGE_FP32 && as == LUXGI_STORAGE_FP32) return launch_batched_forms<float, float>(*batch, d, st);
    if (rs == LUXGI_STORAGE_FP16 && as == LUXGI_STORAGE_FP16) return launch_batched_forms<__half, __half>(*batch, d, st);
    if (rs == LUXGI_STORAGE_FP32 && as == LUXGI_STORAGE_FP16) return launch_batched_forms<float, __half>(*batch, d, st);
    if (rs == LUXGI_STORAGE_FP16 && as == LUXGI_STORAGE_FP32) return launch_batched_forms<__half, float>(*batch, d, st);
    return LUXGI_ERR_UNSUPPORTED;
}

extern "C" int luxgi_ddgi_border_update(const LuxgiDDGIUniform* u, const LuxgiProbeAtlas* atlas, uint32_t which,
                                        void* stream) {
    int rc = check_uniform(u);
    if (rc) return rc;
    if (!atlas) return LUXGI_ERR_INVALID_ARGUMENT;
    const int pxy = u->probeCounts[0] * u->probeCounts[1];
    const int P = pxy * u->probeCounts[2];
    hipStream_t st = as_stream(stream);
    if (which & LUXGI_UPDATE_IRRADIANCE) {
        if (!atlas->irradiance) return LUXGI_ERR_INVALID_ARGUMENT;
        if (atlas->storage == LUXGI_STORAGE_FP32)
            hipLaunchKernelGGL((border_kernel<float, 4>), dim3(P), dim3(64), 0, st,
                               static_cast<float*>(atlas->irradiance), u->irradianceTextureWidth,
                               u->irradianceProbeSideLength, pxy, P);
        else if (atlas->storage == LUXGI_STORAGE_FP16)
            hipLaunchKernelGGL((border_kernel<__half, 4>), dim3(P), dim3(64), 0, st,
                               static_cast<__half*>(atlas->irradiance), u->irradianceTextureWidth,
                               u->irradianceProbeSideLength, pxy, P);
        else return LUXGI_ERR_UNSUPPORTED;
        LUXGI_LAUNCH_CHECK();
    }
    if (which & LUXGI_UPDATE_DEPTH) {
        if (!atlas->depth) return LUXGI_ERR_INVALID_ARGUMENT;
        if (atlas->storage == LUXGI_STORAGE_FP32)
            hipLaunchKernelGGL((border_kernel<float, 2>), dim3(P), dim3(64), 0, st, static_cast<float*>(atlas->depth),
                               u->depthTextureWidth, u->depthProbeSideLength, pxy, P);
        else if (atlas->storage == LUXGI_STORAGE_FP16)
            hipLaunchKernelGGL((border_kernel<__half, 2>), dim3(P), dim3(64), 0, st,
                               static_cast<__half*>(atlas->depth), u->depthTextureWidth, u->depthProbeSideLength, pxy,
                               P);
        else return LUXGI_ERR_UNSUPPORTED;
        LUXGI_LAUNCH_CHECK();
    }
    return LUXGI_OK;
}

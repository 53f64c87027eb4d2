// learned_pixel.hip -- per-pixel learned edge-stopping scales for the reflection a-trous chain,
// with their training kernels, for gfx950. No counterpart in the reference.
//
// A 12 -> 32 -> 32 -> 5 ReLU MLP with FP16 weights maps per-pixel features to one scale per
// a-trous pass, s_i = exp(clamp(o_i, -4, 4)). Pass i of the reference's chain
// (Reflection/DenoiseAtrous.comp:63-134, driven by RaytracedReflection.cpp:419-494) then weighs
// each tap by the reference's edge-stopping weight (EdgeStop.glsl:24-40) with the luminance term
// scaled, wL = s_i(p) |l_p - l_q| / phi_p (luxgi_svgf_reflection_atrous_scaled). All-zero output
// weights give s = 1: the reference chain, bit for bit.
//
// Features f of pixel p (the pass-0 input: rgb + variance; the reprojection's moments m1, m2 and
// history length; the G-buffer normal and linear depth):
//   f0, f1  |z(x+1) - z(x-1)| / 2z and the same along y (edge-clamped), at most 1
//   f2..f4  the decoded normal
//   f5      log(1 + l), l the input luminance
//   f6, f7  log(1 + mean) and min(std / (mean + 1e-3), 4) of the 3x3 luminances
//   f8      min(sqrt(variance) / (l + 1e-3), 4)
//   f9, f10 log(1 + m1) and min(sqrt(max(0, m2 - m1^2)) / (m1 + 1e-3), 4)
//   f11     min(history length, 32) / 32
// Parameters (FP32 master copy; the forward rounds the weights, the features and the hidden
// activations to FP16 for the MFMA and accumulates in FP32): W1 [32][12], b1 [32], W2 [32][32],
// b2 [32], W3 [5][32], b3 [5] = 1637 floats.
//
// Inference: one kernel computes the features and runs the three layers on
// v_mfma_f32_16x16x16_f16 / v_mfma_f32_16x16x32_f16 (units x features . features x 16 pixels),
// the activations passed between layers through LDS. Training: the backward of a scaled pass in
// gather form (a centre kernel keeps each filtered pixel's dL/d(out) / sum of weights and its
// luminance factor; a gather kernel adds each pixel's share of its neighbours' taps, no atomics)
// and the MLP's backward per pixel with the parameter gradient summed per block in LDS and the
// blocks' partial sums added in block order (deterministic). The weights' dependence on the input
// colour (luminance, the variance behind phi) is held fixed in the backward.
#include "svgf_common.hpp"

namespace luxgi {

constexpr int kLpIn = 12, kLpInPad = 16, kLpHid = 32, kLpOut = 5;
constexpr int kLpW1 = 0, kLpB1 = kLpW1 + kLpHid * kLpIn, kLpW2 = kLpB1 + kLpHid, kLpB2 = kLpW2 + kLpHid * kLpHid,
              kLpW3 = kLpB2 + kLpHid, kLpB3 = kLpW3 + kLpOut * kLpHid, kLpParams = kLpB3 + kLpOut;
static_assert(kLpParams == LUXGI_LEARNED_PIXEL_PARAMS, "parameter count");
constexpr float kLpClamp = 4.0f;

typedef _Float16 lp_half4 __attribute__((ext_vector_type(4)));
typedef _Float16 lp_half8 __attribute__((ext_vector_type(8)));
typedef float lp_float4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void lp_features(int W, int H, int x, int y, const float4* __restrict__ color,
                                            const float4* __restrict__ moments, const float4* __restrict__ normals,
                                            const float4* __restrict__ pbr, float* f) {
    const int xl = max(x - 1, 0), xr = min(x + 1, W - 1), yu = max(y - 1, 0), yd = min(y + 1, H - 1);
    const size_t e = (size_t)y * W + x;
    const float z = load_linear_depth(pbr, e);
    const float iz = 1.0f / fmaxf(z, 1e-4f);
    const float zl = load_linear_depth(pbr, (size_t)y * W + xl), zr = load_linear_depth(pbr, (size_t)y * W + xr);
    const float zu = load_linear_depth(pbr, (size_t)yu * W + x), zd = load_linear_depth(pbr, (size_t)yd * W + x);
    f[0] = fminf(fabsf(zr - zl) * 0.5f * iz, 1.0f);
    f[1] = fminf(fabsf(zd - zu) * 0.5f * iz, 1.0f);
    const f3 n = load_normal(normals, e);
    f[2] = n.x;
    f[3] = n.y;
    f[4] = n.z;
    const float4 c = color[e];
    const float l = luminance(c.x, c.y, c.z);
    f[5] = log1pf(l);
    float s1 = 0.0f, s2 = 0.0f;
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) {
            const int qx = min(max(x + xx, 0), W - 1), qy = min(max(y + yy, 0), H - 1);
            const float4 q = color[(size_t)qy * W + qx];
            const float lq = luminance(q.x, q.y, q.z);
            s1 += lq;
            s2 += lq * lq;
        }
    const float mu = s1 * (1.0f / 9.0f);
    const float sd = sqrtf(fmaxf(0.0f, s2 * (1.0f / 9.0f) - mu * mu));
    f[6] = log1pf(mu);
    f[7] = fminf(sd / (mu + 1e-3f), 4.0f);
    f[8] = fminf(sqrtf(fmaxf(c.w, 0.0f)) / (l + 1e-3f), 4.0f);
    const float4 m = moments[e];
    const float m1 = fmaxf(m.x, 0.0f);
    f[9] = log1pf(m1);
    f[10] = fminf(sqrtf(fmaxf(0.0f, m.y - m1 * m1)) / (m1 + 1e-3f), 4.0f);
    f[11] = fminf(m.z, 32.0f) * (1.0f / 32.0f);
}

// ---------------------------------------------------------------------------------------------
// Forward: features + MLP (MFMA). 256 pixels per block, one per lane for the features; each wave
// then runs its 64 pixels through the layers 16 at a time. MFMA lane layouts (l = lane, g = l / 16):
//   16x16x16 f16: A[l % 16][4g + j], B[4g + j][l % 16];  16x16x32 f16: A[l % 16][8g + j],
//   B[8g + j][l % 16];  D[4g + v][l % 16].  A = weights (units x inputs), B = inputs x pixels.
// ---------------------------------------------------------------------------------------------
struct LpForwardLds {
    _Float16 f[256][kLpInPad];    // features, FP16
    _Float16 h[4][16][kLpHid];    // per wave: one layer's activations of 16 pixels
};

__global__ __launch_bounds__(256) void lp_forward_kernel(const float* __restrict__ P, int W, int H,
                                                         const float4* __restrict__ color,
                                                         const float4* __restrict__ moments,
                                                         const float4* __restrict__ normals,
                                                         const float4* __restrict__ pbr, float* __restrict__ scales,
                                                         float* __restrict__ featOut) {
    __shared__ LpForwardLds lds;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int col = lane & 15, g = lane >> 4;
    const long long N = (long long)W * H;
    const long long p0 = (long long)blockIdx.x * 256;
    {
        const long long p = p0 + tid;
        float f[kLpIn];
        if (p < N) {
            lp_features(W, H, int(p % W), int(p / W), color, moments, normals, pbr, f);
            if (featOut)
                for (int i = 0; i < kLpIn; ++i) featOut[p * kLpIn + i] = f[i];
        } else {
            for (int i = 0; i < kLpIn; ++i) f[i] = 0.0f;
        }
        for (int i = 0; i < kLpIn; ++i) lds.f[tid][i] = (_Float16)f[i];
        for (int i = kLpIn; i < kLpInPad; ++i) lds.f[tid][i] = (_Float16)0.0f;
    }
    // weight fragments (A operands), rounded to FP16; biases of the rows this lane's D holds
    lp_half4 a1[2];
    lp_half8 a2[2], a3;
    float b1v[2][4], b2v[2][4], b3v[4];
    for (int mt = 0; mt < 2; ++mt) {
        const int u = mt * 16 + col;
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * g + j;
            a1[mt][j] = (_Float16)(k < kLpIn ? P[kLpW1 + u * kLpIn + k] : 0.0f);
        }
        for (int j = 0; j < 8; ++j) a2[mt][j] = (_Float16)P[kLpW2 + u * kLpHid + 8 * g + j];
        for (int v = 0; v < 4; ++v) {
            b1v[mt][v] = P[kLpB1 + mt * 16 + 4 * g + v];
            b2v[mt][v] = P[kLpB2 + mt * 16 + 4 * g + v];
        }
    }
    for (int j = 0; j < 8; ++j) a3[j] = (_Float16)(col < kLpOut ? P[kLpW3 + col * kLpHid + 8 * g + j] : 0.0f);
    for (int v = 0; v < 4; ++v) b3v[v] = 4 * g + v < kLpOut ? P[kLpB3 + 4 * g + v] : 0.0f;
    __syncthreads();
    for (int b = 0; b < 4; ++b) {
        const int pl = wave * 64 + b * 16;  // the batch's first pixel in the block
        // layer 1
        lp_half4 bf;
        for (int j = 0; j < 4; ++j) bf[j] = lds.f[pl + col][4 * g + j];
        for (int mt = 0; mt < 2; ++mt) {
            lp_float4 d = __builtin_amdgcn_mfma_f32_16x16x16f16(a1[mt], bf, lp_float4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            for (int v = 0; v < 4; ++v) lds.h[wave][col][mt * 16 + 4 * g + v] = (_Float16)fmaxf(d[v] + b1v[mt][v], 0.f);
        }
        __syncthreads();
        // layer 2
        lp_half8 bh;
        for (int j = 0; j < 8; ++j) bh[j] = lds.h[wave][col][8 * g + j];
        __syncthreads();
        for (int mt = 0; mt < 2; ++mt) {
            lp_float4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2[mt], bh, lp_float4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
            for (int v = 0; v < 4; ++v) lds.h[wave][col][mt * 16 + 4 * g + v] = (_Float16)fmaxf(d[v] + b2v[mt][v], 0.f);
        }
        __syncthreads();
        // layer 3
        for (int j = 0; j < 8; ++j) bh[j] = lds.h[wave][col][8 * g + j];
        __syncthreads();
        const lp_float4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a3, bh, lp_float4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        const long long p = p0 + pl + col;
        if (p < N) {
            for (int v = 0; v < 4; ++v) {
                const int o = 4 * g + v;
                if (o < kLpOut) scales[(size_t)o * N + p] = __expf(clampf(d[v] + b3v[v], -kLpClamp, kLpClamp));
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Backward of one scaled reflection a-trous pass (FP32 images).
// ---------------------------------------------------------------------------------------------
struct LpPassArgs {
    int W, H, tilesX, radius, stepSize;
    float phi, phiNormal, sigmaDepth;
    int sqNormal;
};

// mode of a pixel in the pass: 0 the output is the input (copy tile, or roughness outside the
// band), 1 filtered, 2 zero (sky in a filtered tile)
__device__ __forceinline__ int lp_mode(const LpPassArgs& A, const uint8_t* tileMap, const float4* pbr, int x, int y) {
    if (!tileMap[(y >> 3) * A.tilesX + (x >> 3)]) return 0;
    const size_t e = (size_t)y * A.W + x;
    const float cd = load_linear_depth(pbr, e), rough = load_roughness(pbr, e);
    if (cd == 1.0f) return 2;
    return (rough <= 0.05f || rough >= 0.45f) ? 0 : 1;
}

// record[p] = (dL/d out_p / sw_p (rgb), s_p / phi_p) for filtered centres; mode[p]; dL/ds_p
__global__ __launch_bounds__(256) void lp_pass_centre_kernel(LpPassArgs A, const float4* __restrict__ normals,
                                                             const float4* __restrict__ pbr,
                                                             const uint8_t* __restrict__ tileMap,
                                                             const float4* __restrict__ in,
                                                             const float* __restrict__ scale,
                                                             const float4* __restrict__ gOut,
                                                             float4* __restrict__ record, uint8_t* __restrict__ mode,
                                                             float* __restrict__ gScale) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.W || y >= A.H) return;
    const size_t e = (size_t)y * A.W + x;
    const int m = lp_mode(A, tileMap, pbr, x, y);
    mode[e] = (uint8_t)m;
    if (m != 1) {
        record[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        gScale[e] = 0.0f;
        return;
    }
    const float kw[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
    const float4 c = in[e];
    float var = 0.0f;
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) var += fetch4(in, x + xx, y + yy, A.W, A.H).w * kv[abs(xx)][abs(yy)];
    const float cl = luminance(c.x, c.y, c.z);
    const float phiC = A.phi * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
    const float invPhiC = 1.0f / phiC, invPhiZ = 1.0f / A.sigmaDepth;
    const float invPhiL = invPhiC * scale[e];
    const float cd = load_linear_depth(pbr, e);
    const f3 cn = load_normal(normals, e);
    float sw = 1.0f;
    float o0 = c.x, o1 = c.y, o2 = c.z;
    for (int yy = -A.radius; yy <= A.radius; ++yy)
        for (int xx = -A.radius; xx <= A.radius; ++xx) {
            const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
            if (!inside(px, py, A.W, A.H) || (xx == 0 && yy == 0)) continue;
            const size_t q = (size_t)py * A.W + px;
            const float4 sc = in[q];
            const float w = edge_stop(cd, load_linear_depth(pbr, q), invPhiZ, cn, load_normal(normals, q),
                                      A.phiNormal, A.sqNormal, cl, luminance(sc.x, sc.y, sc.z), invPhiL) *
                            (kw[abs(xx)] * kw[abs(yy)]);
            sw += w;
            o0 += w * sc.x;
            o1 += w * sc.y;
            o2 += w * sc.z;
        }
    const float isw = 1.0f / sw;
    o0 *= isw;
    o1 *= isw;
    o2 *= isw;
    const float4 g = gOut[e];
    const float r0 = g.x * isw, r1 = g.y * isw, r2 = g.z * isw;
    // dL/ds = sum over taps of r . (x_q - out) dw/ds, dw/ds = -|l_p - l_q| / phi_p * w
    float gs = 0.0f;
    for (int yy = -A.radius; yy <= A.radius; ++yy)
        for (int xx = -A.radius; xx <= A.radius; ++xx) {
            const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
            if (!inside(px, py, A.W, A.H) || (xx == 0 && yy == 0)) continue;
            const size_t q = (size_t)py * A.W + px;
            const float4 sc = in[q];
            const float sl = luminance(sc.x, sc.y, sc.z);
            const float w = edge_stop(cd, load_linear_depth(pbr, q), invPhiZ, cn, load_normal(normals, q),
                                      A.phiNormal, A.sqNormal, cl, sl, invPhiL) *
                            (kw[abs(xx)] * kw[abs(yy)]);
            const float dw = -fabsf(cl - sl) * invPhiC * w;
            gs += dw * (r0 * (sc.x - o0) + r1 * (sc.y - o1) + r2 * (sc.z - o2));
        }
    record[e] = make_float4(r0, r1, r2, invPhiL);
    gScale[e] = gs;
}

// dL/d in_q (rgb): its own term plus r_p w_pq for every filtered neighbour p whose tap q is
__global__ __launch_bounds__(256) void lp_pass_gather_kernel(LpPassArgs A, const float4* __restrict__ normals,
                                                             const float4* __restrict__ pbr,
                                                             const float4* __restrict__ in,
                                                             const float4* __restrict__ gOut,
                                                             const float4* __restrict__ record,
                                                             const uint8_t* __restrict__ mode,
                                                             float4* __restrict__ gIn) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.W || y >= A.H) return;
    const float kw[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const size_t e = (size_t)y * A.W + x;
    const int m = mode[e];
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    if (m == 0) {
        const float4 g = gOut[e];
        a0 = g.x;
        a1 = g.y;
        a2 = g.z;
    } else if (m == 1) {
        const float4 r = record[e];
        a0 = r.x;
        a1 = r.y;
        a2 = r.z;
    }
    const float4 sc = in[e];
    const float sl = luminance(sc.x, sc.y, sc.z);
    const float d = load_linear_depth(pbr, e);
    const f3 sn = load_normal(normals, e);
    const float invPhiZ = 1.0f / A.sigmaDepth;
    for (int yy = -A.radius; yy <= A.radius; ++yy)
        for (int xx = -A.radius; xx <= A.radius; ++xx) {
            const int px = x - xx * A.stepSize, py = y - yy * A.stepSize;  // the centre whose tap (xx, yy) this is
            if (!inside(px, py, A.W, A.H) || (xx == 0 && yy == 0)) continue;
            const size_t p = (size_t)py * A.W + px;
            if (mode[p] != 1) continue;
            const float4 r = record[p];
            const float4 cp = in[p];
            const float w = edge_stop(load_linear_depth(pbr, p), d, invPhiZ, load_normal(normals, p), sn, A.phiNormal,
                                      A.sqNormal, luminance(cp.x, cp.y, cp.z), sl, r.w) *
                            (kw[abs(xx)] * kw[abs(yy)]);
            a0 += r.x * w;
            a1 += r.y * w;
            a2 += r.z * w;
        }
    gIn[e] = make_float4(a0, a1, a2, 0.0f);
}

// ---------------------------------------------------------------------------------------------
// Backward of the MLP. A block takes chunks of 64 pixels: lanes 0..63 run one pixel each (the
// forward with FP16-rounded operands, then dL/do = dL/ds * s inside the clamp, dL/dh2, dL/dh1)
// into LDS, then every lane sums its parameters' terms over the chunk. Each block writes its sums
// to partial[block]; lp_reduce_kernel adds the blocks' partials in order into the gradient.
// ---------------------------------------------------------------------------------------------
constexpr int kLpChunk = 64;
constexpr int kLpBwdBlocks = 240;

struct LpBackwardLds {
    float f[kLpChunk][kLpIn];
    float h1[kLpChunk][kLpHid + 1];
    float h2[kLpChunk][kLpHid + 1];
    float g1[kLpChunk][kLpHid + 1];
    float g2[kLpChunk][kLpHid + 1];
    float go[kLpChunk][kLpOut];
};

__device__ __forceinline__ float lp_q(float v) { return (float)(_Float16)v; }

__global__ __launch_bounds__(256) void lp_mlp_backward_kernel(const float* __restrict__ P, long long N,
                                                              const float* __restrict__ feat,
                                                              const float* __restrict__ gScales,
                                                              const float* __restrict__ scales,
                                                              float* __restrict__ partial) {
    __shared__ LpBackwardLds lds;
    const int tid = threadIdx.x;
    constexpr int kPer = (kLpParams + 255) / 256;
    float acc[kPer];
    for (int i = 0; i < kPer; ++i) acc[i] = 0.0f;
    const long long chunks = (N + kLpChunk - 1) / kLpChunk;
    for (long long c = blockIdx.x; c < chunks; c += gridDim.x) {
        if (tid < kLpChunk) {
            const long long p = c * kLpChunk + tid;
            float f[kLpIn], h1[kLpHid], h2[kLpHid], go[kLpOut];
            bool m1[kLpHid], m2[kLpHid];
            if (p < N) {
                for (int i = 0; i < kLpIn; ++i) f[i] = lp_q(feat[p * kLpIn + i]);
                for (int j = 0; j < kLpHid; ++j) {
                    float a = P[kLpB1 + j];
                    for (int i = 0; i < kLpIn; ++i) a += lp_q(P[kLpW1 + j * kLpIn + i]) * f[i];
                    m1[j] = a > 0.0f;
                    h1[j] = lp_q(fmaxf(a, 0.0f));
                }
                for (int j = 0; j < kLpHid; ++j) {
                    float a = P[kLpB2 + j];
                    for (int i = 0; i < kLpHid; ++i) a += lp_q(P[kLpW2 + j * kLpHid + i]) * h1[i];
                    m2[j] = a > 0.0f;
                    h2[j] = lp_q(fmaxf(a, 0.0f));
                }
                for (int j = 0; j < kLpOut; ++j) {
                    float a = P[kLpB3 + j];
                    for (int i = 0; i < kLpHid; ++i) a += lp_q(P[kLpW3 + j * kLpHid + i]) * h2[i];
                    const size_t e = (size_t)j * N + p;
                    go[j] = (a > -kLpClamp && a < kLpClamp) ? gScales[e] * scales[e] : 0.0f;
                }
            } else {
                for (int i = 0; i < kLpIn; ++i) f[i] = 0.0f;
                for (int j = 0; j < kLpHid; ++j) {
                    h1[j] = h2[j] = 0.0f;
                    m1[j] = m2[j] = false;
                }
                for (int j = 0; j < kLpOut; ++j) go[j] = 0.0f;
            }
            float gh2[kLpHid], gh1[kLpHid];
            for (int i = 0; i < kLpHid; ++i) {
                float s = 0.0f;
                for (int j = 0; j < kLpOut; ++j) s += go[j] * lp_q(P[kLpW3 + j * kLpHid + i]);
                gh2[i] = m2[i] ? s : 0.0f;
            }
            for (int i = 0; i < kLpHid; ++i) {
                float s = 0.0f;
                for (int j = 0; j < kLpHid; ++j) s += gh2[j] * lp_q(P[kLpW2 + j * kLpHid + i]);
                gh1[i] = m1[i] ? s : 0.0f;
            }
            for (int i = 0; i < kLpIn; ++i) lds.f[tid][i] = f[i];
            for (int i = 0; i < kLpHid; ++i) {
                lds.h1[tid][i] = h1[i];
                lds.h2[tid][i] = h2[i];
                lds.g1[tid][i] = gh1[i];
                lds.g2[tid][i] = gh2[i];
            }
            for (int j = 0; j < kLpOut; ++j) lds.go[tid][j] = go[j];
        }
        __syncthreads();
        for (int i = 0; i < kPer; ++i) {
            const int k = tid + i * 256;
            if (k >= kLpParams) break;
            float s = 0.0f;
            if (k < kLpB1) {
                const int j = (k - kLpW1) / kLpIn, ii = (k - kLpW1) % kLpIn;
                for (int p = 0; p < kLpChunk; ++p) s += lds.g1[p][j] * lds.f[p][ii];
            } else if (k < kLpW2) {
                for (int p = 0; p < kLpChunk; ++p) s += lds.g1[p][k - kLpB1];
            } else if (k < kLpB2) {
                const int j = (k - kLpW2) / kLpHid, ii = (k - kLpW2) % kLpHid;
                for (int p = 0; p < kLpChunk; ++p) s += lds.g2[p][j] * lds.h1[p][ii];
            } else if (k < kLpW3) {
                for (int p = 0; p < kLpChunk; ++p) s += lds.g2[p][k - kLpB2];
            } else if (k < kLpB3) {
                const int j = (k - kLpW3) / kLpHid, ii = (k - kLpW3) % kLpHid;
                for (int p = 0; p < kLpChunk; ++p) s += lds.go[p][j] * lds.h2[p][ii];
            } else {
                for (int p = 0; p < kLpChunk; ++p) s += lds.go[p][k - kLpB3];
            }
            acc[i] += s;
        }
        __syncthreads();
    }
    for (int i = 0; i < kPer; ++i) {
        const int k = tid + i * 256;
        if (k < kLpParams) partial[(size_t)blockIdx.x * kLpParams + k] = acc[i];
    }
}

__global__ void lp_reduce_kernel(const float* __restrict__ partial, int blocks, float* __restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kLpParams) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partial[(size_t)b * kLpParams + k];
    grad[k] += s;
}

}  // namespace luxgi

using namespace luxgi;

static bool lp_gbuf_ok(const LuxgiGBuffer* g) { return g && g->normals && g->pbr; }

extern "C" int luxgi_learned_pixel_scales(const float* params, int32_t W, int32_t H, const LuxgiGBuffer* cur,
                                          const float* color, const float* moments, float* scales, float* features,
                                          void* stream) {
    if (!params || W <= 0 || H <= 0 || !lp_gbuf_ok(cur) || !color || !moments || !scales)
        return LUXGI_ERR_INVALID_ARGUMENT;
    const long long N = (long long)W * H;
    if (N > (1ll << 30)) return LUXGI_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(lp_forward_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, as_stream(stream), params,
                       W, H, reinterpret_cast<const float4*>(color), reinterpret_cast<const float4*>(moments),
                       reinterpret_cast<const float4*>(cur->normals), reinterpret_cast<const float4*>(cur->pbr),
                       scales, features);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" size_t luxgi_learned_pixel_pass_backward_workspace_size(int32_t W, int32_t H) {
    if (W <= 0 || H <= 0) return 0;
    const size_t n = (size_t)W * H;
    return ((n * sizeof(float4) + 255) & ~size_t(255)) + n;
}

extern "C" int luxgi_learned_pixel_pass_backward(const LuxgiReflectionAtrousParams* p, int32_t W, int32_t H,
                                                 const LuxgiGBuffer* cur, const uint8_t* tileMap, const float* input,
                                                 const float* lumScale, const float* gradOut, float* gradIn,
                                                 float* gradScale, void* workspace, size_t workspaceBytes,
                                                 void* stream) {
    if (!p || W <= 0 || H <= 0 || !lp_gbuf_ok(cur) || !tileMap || !input || !lumScale || !gradOut || !gradIn ||
        !gradScale)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (p->radius < 0 || p->radius > 2) return LUXGI_ERR_UNSUPPORTED;
    if (p->stepSize < 1 || p->stepSize > (1 << 16)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (gradIn == input || gradIn == gradOut) return LUXGI_ERR_INVALID_ARGUMENT;  // the gather reads both
    const size_t need = luxgi_learned_pixel_pass_backward_workspace_size(W, H);
    if (!workspace || workspaceBytes < need) return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    const size_t n = (size_t)W * H;
    float4* record = static_cast<float4*>(workspace);
    uint8_t* mode = static_cast<uint8_t*>(workspace) + ((n * sizeof(float4) + 255) & ~size_t(255));
    LpPassArgs A{W, H, (W + 7) / 8, p->radius, p->stepSize, p->phiColor, p->phiNormal, p->sigmaDepth,
                 normal_squarings(p->phiNormal)};
    const dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipStream_t st = as_stream(stream);
    const float4* nrm = reinterpret_cast<const float4*>(cur->normals);
    const float4* pbr = reinterpret_cast<const float4*>(cur->pbr);
    hipLaunchKernelGGL(lp_pass_centre_kernel, grid, dim3(256), 0, st, A, nrm, pbr, tileMap,
                       reinterpret_cast<const float4*>(input), lumScale, reinterpret_cast<const float4*>(gradOut),
                       record, mode, gradScale);
    LUXGI_LAUNCH_CHECK();
    hipLaunchKernelGGL(lp_pass_gather_kernel, grid, dim3(256), 0, st, A, nrm, pbr,
                       reinterpret_cast<const float4*>(input), reinterpret_cast<const float4*>(gradOut), record, mode,
                       reinterpret_cast<float4*>(gradIn));
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" size_t luxgi_learned_pixel_mlp_backward_workspace_size(void) {
    return (size_t)kLpBwdBlocks * kLpParams * sizeof(float);
}

extern "C" int luxgi_learned_pixel_mlp_backward(const float* params, int32_t W, int32_t H, const float* features,
                                                const float* scales, const float* gradScales, float* gradParams,
                                                void* workspace, size_t workspaceBytes, void* stream) {
    if (!params || W <= 0 || H <= 0 || !features || !scales || !gradScales || !gradParams)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (!workspace || workspaceBytes < luxgi_learned_pixel_mlp_backward_workspace_size())
        return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    const long long N = (long long)W * H;
    hipStream_t st = as_stream(stream);
    float* partial = static_cast<float*>(workspace);
    hipLaunchKernelGGL(lp_mlp_backward_kernel, dim3(kLpBwdBlocks), dim3(256), 0, st, params, N, features, gradScales,
                       scales, partial);
    LUXGI_LAUNCH_CHECK();
    hipLaunchKernelGGL(lp_reduce_kernel, dim3((kLpParams + 255) / 256), dim3(256), 0, st, partial, kLpBwdBlocks,
                       gradParams);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

// learned_atrous.hip -- a learned edge-stopping weight for the reflection a-trous pass, with its
// training kernels, for gfx950.
//
// The pass is Reflection/DenoiseAtrous.comp:63-134 (3x3 Gaussian variance at the centre, sky pixels
// output 0, pixels with roughness outside (0.05, 0.45) pass through, the (2r+1)^2 taps at stepSize
// with the centre weighted 1 and the variance combined with the squared weights) except that the
// tap weight exp(-|dl|/phi - exp(-|dz|/sigma)) * max(0, n.n')^phiN * k (EdgeStop.glsl:24-40) is
// replaced by a small MLP of the same per-tap quantities:
//   features f = (min(|l_p - l_q| / phi_p, 16), exp(-|z_p - z_q| / sigmaDepth), max(0, n_p.n_q),
//                 k(dx) k(dy), level, min(sqrt(var_p), 4))
//   h = relu(W1 f + b1) (10 units), o = w2.h + b2, weight = exp(clamp(o, -16, 8)).
// Parameters, packed: W1 [10][6] row-major, b1 [10], w2 [10], b2 (81 floats).
// Training: the backward kernel takes dL/d(out rgb) and returns dL/d(in rgb) and accumulates
// dL/d(parameters). The features and the variance channel are treated as constants (their
// dependence on the input colour is not differentiated), so the gradient is that of the filter
// with the weights held at their forward values, composed with the MLP's own derivative.
#include "common.hpp"

namespace luxgi {

constexpr int kLaIn = 6, kLaHidden = 10;
constexpr int kLaParams = kLaHidden * kLaIn + kLaHidden + kLaHidden + 1;
static_assert(kLaParams == LUXGI_LEARNED_ATROUS_PARAMS, "parameter count");

struct LaArgs {
    int W, H, radius, stepSize;
    float phiColor, sigmaDepth, level;
};

__device__ __forceinline__ float la_luminance(float r, float g, float b) {  // ReflectionCommon.glsl:19-22
    return fmaxf(r * 0.299f + g * 0.587f + b * 0.114f, 0.0001f);
}
__device__ __forceinline__ bool la_inside(int x, int y, int W, int H) { return x >= 0 && y >= 0 && x < W && y < H; }

struct LaPixel {
    float4 c;   // input rgb + variance
    f3 n;       // decoded normal
    float z;    // linear depth (pbr.w)
    float rough;
};

__device__ __forceinline__ f3 load_normal_oct(const float4* __restrict__ normals, size_t q) {
    const float2 o = reinterpret_cast<const float2*>(normals)[2 * q];
    return oct_decode(o.x, o.y);
}

__device__ __forceinline__ LaPixel la_load(const float4* __restrict__ in, const float4* __restrict__ normals,
                                           const float4* __restrict__ pbr, size_t q) {
    LaPixel p;
    p.c = in[q];
    const float2 o = reinterpret_cast<const float2*>(normals)[2 * q];
    p.n = oct_decode(o.x, o.y);
    const float4 m = pbr[q];
    p.z = m.w;
    p.rough = m.y;
    return p;
}

// centre quantities of pixel (x, y): 3x3 Gaussian variance (computeVarianceCenter, :41-61) and phi
__device__ __forceinline__ float la_phi(const LaArgs& A, const float4* __restrict__ in, int x, int y) {
    const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
    float var = 0.0f;
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx)
            if (la_inside(x + xx, y + yy, A.W, A.H)) var += in[(size_t)(y + yy) * A.W + x + xx].w * kv[abs(xx)][abs(yy)];
    return var;
}

// 0 = filtered, 1 = passes through (roughness outside (0.05, 0.45)), 2 = sky (outputs 0)
__device__ __forceinline__ int la_mode(const LaPixel& p) {
    if (p.z == 1.0f) return 2;
    return (p.rough <= 0.05f || p.rough >= 0.45f) ? 1 : 0;
}

struct LaFeatures {
    float f[kLaIn];
};

__device__ __forceinline__ LaFeatures la_features(const LaArgs& A, const LaPixel& p, float lp, float invPhi, float sdv,
                                                  const LaPixel& q, int xx, int yy) {
    const float kw[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    LaFeatures F;
    const float lq = la_luminance(q.c.x, q.c.y, q.c.z);
    F.f[0] = fminf(fabsf(lp - lq) * invPhi, 16.0f);
    F.f[1] = __expf(-fabsf(p.z - q.z) / A.sigmaDepth);
    F.f[2] = fmaxf(0.0f, dot3(p.n, q.n));
    F.f[3] = kw[abs(xx)] * kw[abs(yy)];
    F.f[4] = A.level;
    F.f[5] = sdv;
    return F;
}

// forward of the MLP; h keeps the hidden activations, returns the weight and whether o was clamped
__device__ __forceinline__ float la_mlp(const float* __restrict__ P, const LaFeatures& F, float h[kLaHidden],
                                        bool& clamped) {
    const float* W1 = P;
    const float* b1 = P + kLaHidden * kLaIn;
    const float* w2 = b1 + kLaHidden;
    const float b2 = w2[kLaHidden];
    float o = b2;
#pragma unroll
    for (int j = 0; j < kLaHidden; ++j) {
        float a = b1[j];
#pragma unroll
        for (int i = 0; i < kLaIn; ++i) a += W1[j * kLaIn + i] * F.f[i];
        h[j] = fmaxf(a, 0.0f);
        o += w2[j] * h[j];
    }
    clamped = o < -16.0f || o > 8.0f;
    return __expf(clampf(o, -16.0f, 8.0f));
}

__device__ __forceinline__ void la_stage_params(const float* __restrict__ params, float* sP) {
    for (int i = threadIdx.x; i < kLaParams; i += blockDim.x) sP[i] = params[i];
    __syncthreads();
}

// one learned a-trous pass; `record` (nullable) keeps per pixel what the backward needs: (1 + the
// tap weights, luminance, 1 / phi, min(sqrt(var), 4)) of a filtered pixel, (1, 0, 0, 0) otherwise
__global__ __launch_bounds__(256) void learned_atrous_kernel(LaArgs A, const float* __restrict__ params,
                                                             const float4* __restrict__ normals,
                                                             const float4* __restrict__ pbr,
                                                             const float4* __restrict__ in, float4* __restrict__ out,
                                                             float4* __restrict__ record) {
    __shared__ float sP[kLaParams];
    la_stage_params(params, sP);
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.W || y >= A.H) return;
    const size_t e = (size_t)y * A.W + x;
    const LaPixel p = la_load(in, normals, pbr, e);
    const int mode = la_mode(p);
    float4 o = p.c;
    float sw = 1.0f, lp = 0.0f, invPhi = 0.0f, sdv = 0.0f;
    if (mode == 2) {
        o = make_float4(0.f, 0.f, 0.f, 0.f);
    } else if (mode == 0) {
        const float var = la_phi(A, in, x, y);
        const float phi = A.phiColor * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
        lp = la_luminance(p.c.x, p.c.y, p.c.z);
        invPhi = 1.0f / phi;
        sdv = fminf(sqrtf(fmaxf(var, 0.0f)), 4.0f);
        float4 s = p.c;
        for (int yy = -A.radius; yy <= A.radius; ++yy)
            for (int xx = -A.radius; xx <= A.radius; ++xx) {
                const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
                if (!la_inside(px, py, A.W, A.H) || (xx == 0 && yy == 0)) continue;
                const LaPixel q = la_load(in, normals, pbr, (size_t)py * A.W + px);
                const LaFeatures F = la_features(A, p, lp, invPhi, sdv, q, xx, yy);
                float h[kLaHidden];
                bool cl;
                const float w = la_mlp(sP, F, h, cl);
                sw += w;
                s.x += w * q.c.x;
                s.y += w * q.c.y;
                s.z += w * q.c.z;
                s.w += w * w * q.c.w;
            }
        o = make_float4(s.x / sw, s.y / sw, s.z / sw, s.w / (sw * sw));
    }
    out[e] = o;
    if (record) record[e] = make_float4(sw, lp, invPhi, sdv);
}

// LDS-tiled form of the same pass for footprints radius * stepSize <= 8: a 16x16 block stages its
// input and guides (decoded normal, linear depth) with a halo of max(radius * stepSize, 1) once,
// so every tap, and the 3x3 variance, is an LDS read. The arithmetic is the direct kernel's, in
// the same order (the compiler may contract a multiply-add differently, so the two agree to
// rounding, not necessarily bit for bit).
constexpr int kLaBlock = 16, kLaMaxHalo = 8, kLaMaxSide = kLaBlock + 2 * kLaMaxHalo;

__global__ __launch_bounds__(256) void learned_atrous_lds_kernel(LaArgs A, const float* __restrict__ params,
                                                                 const float4* __restrict__ normals,
                                                                 const float4* __restrict__ pbr,
                                                                 const float4* __restrict__ in,
                                                                 float4* __restrict__ out,
                                                                 float4* __restrict__ record) {
    __shared__ float sP[kLaParams];
    __shared__ float4 sIn[kLaMaxSide * kLaMaxSide];
    __shared__ float4 sGuide[kLaMaxSide * kLaMaxSide];  // normal xyz, linear depth
    const int h = max(A.radius * A.stepSize, 1), side = kLaBlock + 2 * h;
    const int ox = blockIdx.x * kLaBlock - h, oy = blockIdx.y * kLaBlock - h;
    for (int i = threadIdx.x; i < kLaParams; i += blockDim.x) sP[i] = params[i];
    for (int i = threadIdx.x; i < side * side; i += blockDim.x) {
        const int ly = i / side, lx = i - ly * side;
        const int gx = ox + lx, gy = oy + ly;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f), gd = make_float4(0.f, 0.f, 0.f, 0.f);
        if (la_inside(gx, gy, A.W, A.H)) {
            const LaPixel q = la_load(in, normals, pbr, (size_t)gy * A.W + gx);
            v = q.c;
            gd = make_float4(q.n.x, q.n.y, q.n.z, q.z);
        }
        sIn[i] = v;
        sGuide[i] = gd;
    }
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x = blockIdx.x * kLaBlock + tx, y = blockIdx.y * kLaBlock + ty;
    if (x >= A.W || y >= A.H) return;
    const size_t e = (size_t)y * A.W + x;
    const int lc = (ty + h) * side + (tx + h);
    LaPixel p;
    p.c = sIn[lc];
    const float4 gc = sGuide[lc];
    p.n = mk(gc.x, gc.y, gc.z);
    p.z = gc.w;
    p.rough = pbr[e].y;
    const int mode = la_mode(p);
    float4 o = p.c;
    float sw = 1.0f, lp = 0.0f, invPhi = 0.0f, sdv = 0.0f;
    if (mode == 2) {
        o = make_float4(0.f, 0.f, 0.f, 0.f);
    } else if (mode == 0) {
        const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
        float var = 0.0f;  // out-of-image texels are staged as 0, as la_phi skips them
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx)
                if (la_inside(x + xx, y + yy, A.W, A.H)) var += sIn[lc + yy * side + xx].w * kv[abs(xx)][abs(yy)];
        const float phi = A.phiColor * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
        lp = la_luminance(p.c.x, p.c.y, p.c.z);
        invPhi = 1.0f / phi;
        sdv = fminf(sqrtf(fmaxf(var, 0.0f)), 4.0f);
        float4 s = p.c;
        for (int yy = -A.radius; yy <= A.radius; ++yy)
            for (int xx = -A.radius; xx <= A.radius; ++xx) {
                const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
                if (!la_inside(px, py, A.W, A.H) || (xx == 0 && yy == 0)) continue;
                const int l = (py - oy) * side + (px - ox);
                LaPixel q;
                q.c = sIn[l];
                const float4 gq = sGuide[l];
                q.n = mk(gq.x, gq.y, gq.z);
                q.z = gq.w;
                const LaFeatures F = la_features(A, p, lp, invPhi, sdv, q, xx, yy);
                float hh[kLaHidden];
                bool cl;
                const float w = la_mlp(sP, F, hh, cl);
                sw += w;
                s.x += w * q.c.x;
                s.y += w * q.c.y;
                s.z += w * q.c.z;
                s.w += w * w * q.c.w;
            }
        o = make_float4(s.x / sw, s.y / sw, s.z / sw, s.w / (sw * sw));
    }
    out[e] = o;
    if (record) record[e] = make_float4(sw, lp, invPhi, sdv);
}

// Backward of one pass. Per pixel x: as a centre, the parameter gradient of its taps; as a tap of
// its neighbours (and as its own centre), dL/d(in rgb). The centres' weight sums, luminances,
// 1 / phi and deviations come from the forward's records (no variance re-evaluation).
__global__ __launch_bounds__(256) void learned_atrous_backward_kernel(
    LaArgs A, const float* __restrict__ params, const float4* __restrict__ normals, const float4* __restrict__ pbr,
    const float4* __restrict__ in, const float4* __restrict__ out, const float4* __restrict__ record,
    const float4* __restrict__ gOut, float4* __restrict__ gIn, float* __restrict__ gParams) {
    __shared__ float sP[kLaParams];
    __shared__ float sG[4][kLaParams];
    la_stage_params(params, sP);
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool valid = x < A.W && y < A.H;
    float acc[kLaParams];
#pragma unroll
    for (int i = 0; i < kLaParams; ++i) acc[i] = 0.0f;
    if (valid) {
        const size_t e = (size_t)y * A.W + x;
        const LaPixel p = la_load(in, normals, pbr, e);
        const int mode = la_mode(p);
        const float4 g = gOut[e];
        float3 gi = make_float3(0.f, 0.f, 0.f);
        if (mode == 1) gi = make_float3(g.x, g.y, g.z);  // pass-through
        if (mode == 0) {
            // centre: out = (c + sum w c_q) / W -> d out / d c_p = 1 / W (weights held fixed)
            const float4 rec = record[e];
            const float Wp = rec.x, lp = rec.y, invPhi = rec.z, sdv = rec.w;
            gi = make_float3(g.x / Wp, g.y / Wp, g.z / Wp);
            const float4 op = out[e];
            for (int yy = -A.radius; yy <= A.radius; ++yy)
                for (int xx = -A.radius; xx <= A.radius; ++xx) {
                    const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
                    if (!la_inside(px, py, A.W, A.H) || (xx == 0 && yy == 0)) continue;
                    const LaPixel q = la_load(in, normals, pbr, (size_t)py * A.W + px);
                    const LaFeatures F = la_features(A, p, lp, invPhi, sdv, q, xx, yy);
                    float h[kLaHidden];
                    bool cl;
                    const float w = la_mlp(sP, F, h, cl);
                    // dL/dw = g . (c_q - out) / W ; the variance channel carries no gradient
                    const float dw = (g.x * (q.c.x - op.x) + g.y * (q.c.y - op.y) + g.z * (q.c.z - op.z)) / Wp;
                    if (cl) continue;
                    const float dO = dw * w;  // d exp(o) / d o = exp(o)
                    float* aW1 = acc;
                    float* ab1 = acc + kLaHidden * kLaIn;
                    float* aw2 = ab1 + kLaHidden;
                    aw2[kLaHidden] += dO;
#pragma unroll
                    for (int j = 0; j < kLaHidden; ++j) {
                        aw2[j] += dO * h[j];
                        const float dh = h[j] > 0.0f ? dO * sP[kLaHidden * kLaIn + kLaHidden + j] : 0.0f;
                        ab1[j] += dh;
#pragma unroll
                        for (int i = 0; i < kLaIn; ++i) aW1[j * kLaIn + i] += dh * F.f[i];
                    }
                }
        }
        // as a tap of each neighbour centre c = x - offset: + g_c w(c, x) / W_c
        for (int yy = -A.radius; yy <= A.radius; ++yy)
            for (int xx = -A.radius; xx <= A.radius; ++xx) {
                if (xx == 0 && yy == 0) continue;
                const int cx = x - xx * A.stepSize, cy = y - yy * A.stepSize;
                if (!la_inside(cx, cy, A.W, A.H)) continue;
                const size_t ce = (size_t)cy * A.W + cx;
                LaPixel c;  // the centre's guides (its colour enters through the record's luminance)
                c.n = load_normal_oct(normals, ce);
                const float4 m = pbr[ce];
                c.z = m.w;
                c.rough = m.y;
                if (la_mode(c) != 0) continue;
                const float4 rec = record[ce];
                // x is the tap (xx, yy) of the centre c
                const LaFeatures F = la_features(A, c, rec.y, rec.z, rec.w, p, xx, yy);
                float h[kLaHidden];
                bool cl;
                const float w = la_mlp(sP, F, h, cl) / rec.x;
                const float4 gc = gOut[ce];
                gi.x += gc.x * w;
                gi.y += gc.y * w;
                gi.z += gc.z * w;
            }
        gIn[e] = make_float4(gi.x, gi.y, gi.z, 0.0f);
    }
    // block reduction of the parameter gradient: waves by shuffles, then across the 4 waves in LDS
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kLaParams; ++i) {
        float v = acc[i];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) sG[wave][i] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kLaParams; i += blockDim.x)
        atomicAdd(&gParams[i], sG[0][i] + sG[1][i] + sG[2][i] + sG[3][i]);
}

}  // namespace luxgi

using namespace luxgi;

static int la_args(const LuxgiLearnedAtrousParams* p, int32_t W, int32_t H, const LuxgiGBuffer* gb, LaArgs& A) {
    if (!p || !gb || !gb->normals || !gb->pbr || W <= 0 || H <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if ((long long)W * H >= (1ll << 31) || W > (1 << 24) || H > (1 << 24)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (p->radius < 0 || p->radius > 2 || p->stepSize < 1 || p->stepSize > 65536 || !(p->sigmaDepth > 0.0f))
        return LUXGI_ERR_INVALID_ARGUMENT;
    A.W = W;
    A.H = H;
    A.radius = p->radius;
    A.stepSize = p->stepSize;
    A.phiColor = p->phiColor;
    A.sigmaDepth = p->sigmaDepth;
    A.level = p->level;
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_learned_atrous(const LuxgiLearnedAtrousParams* p, const float* weights, int32_t W,
                                         int32_t H, const LuxgiGBuffer* gb, const float* in, float* out, float* record,
                                         void* stream) {
    LaArgs A;
    const int rc = la_args(p, W, H, gb, A);
    if (rc != LUXGI_OK) return rc;
    if (!weights || !in || !out || in == out) return LUXGI_ERR_INVALID_ARGUMENT;
    if (A.stepSize <= kLaMaxHalo && A.radius * A.stepSize <= kLaMaxHalo)
        hipLaunchKernelGGL(learned_atrous_lds_kernel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0,
                           as_stream(stream), A, weights, reinterpret_cast<const float4*>(gb->normals),
                           reinterpret_cast<const float4*>(gb->pbr), reinterpret_cast<const float4*>(in),
                           reinterpret_cast<float4*>(out), reinterpret_cast<float4*>(record));
    else
        hipLaunchKernelGGL(learned_atrous_kernel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, as_stream(stream),
                           A, weights, reinterpret_cast<const float4*>(gb->normals),
                           reinterpret_cast<const float4*>(gb->pbr), reinterpret_cast<const float4*>(in),
                           reinterpret_cast<float4*>(out), reinterpret_cast<float4*>(record));
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_learned_atrous_backward(const LuxgiLearnedAtrousParams* p, const float* weights, int32_t W,
                                                  int32_t H, const LuxgiGBuffer* gb, const float* in, const float* out,
                                                  const float* record, const float* gradOut, float* gradIn,
                                                  float* gradWeights, void* stream) {
    LaArgs A;
    const int rc = la_args(p, W, H, gb, A);
    if (rc != LUXGI_OK) return rc;
    if (!weights || !in || !out || !record || !gradOut || !gradIn || !gradWeights) return LUXGI_ERR_INVALID_ARGUMENT;
    if (gradIn == gradOut || gradIn == in) return LUXGI_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(learned_atrous_backward_kernel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0,
                       as_stream(stream), A, weights, reinterpret_cast<const float4*>(gb->normals),
                       reinterpret_cast<const float4*>(gb->pbr), reinterpret_cast<const float4*>(in),
                       reinterpret_cast<const float4*>(out), reinterpret_cast<const float4*>(record),
                       reinterpret_cast<const float4*>(gradOut),
                       reinterpret_cast<float4*>(gradIn), gradWeights);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

// Adam (Kingma & Ba, 2015) on the parameter vector: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
// w -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps); the gradient is cleared for the next step.
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                            int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    g[i] = 0.0f;
}

extern "C" int luxgi_adam_step(float* weights, float* grad, float* m, float* v, int32_t n, float lr, float beta1,
                               float beta2, float eps, int32_t step, void* stream) {
    if (!weights || !grad || !m || !v || n <= 0 || step < 1 || !(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) ||
        !(beta2 >= 0.0f && beta2 < 1.0f))
        return LUXGI_ERR_INVALID_ARGUMENT;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), weights, grad, m, v, n, lr,
                       beta1, beta2, eps, c1, c2);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

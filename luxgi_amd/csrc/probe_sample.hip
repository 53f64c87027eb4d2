// probe_sample.hip -- per-pixel DDGI irradiance from the probe atlases, for gfx950.
//
// Reference: Shaders/DDGI/SampleProbe.comp:36-60 (per pixel), sampleIrradiance
// (Shaders/DDGI/DDGICommon.glsl:163-233), textureCoordFromDirection (:141-158),
// octEncode (:65-72), baseGridCoord (:125-128); host dispatch
// Engine/DDGI/DDGIRenderer.cpp:467-503. The atlases are read with bilinear
// filtering and repeat addressing (the engine's default sampler,
// RHI/Definitions.h:152-159), with exact fp32 sub-texel weights.
//
// One lane per pixel, 16x16 pixels per 256-lane block. Neighbouring pixels read
// the same 8 probes in nearby directions, so a wave's 2x2 footprints fall on a
// few atlas lines; the G-buffer is read once per pixel (depth 4 B, normal 8 B).
#include "common.hpp"

namespace luxgi {

constexpr float kPiF = 3.1415926535897932384626433832795f;  // Common/Math.glsl:7

// GLSL normalize / pow as the hardware evaluates them: reciprocal square root (v_rsq_f32) and
// 2^(y log2 x) (v_log_f32, v_exp_f32), each within a few ulp of the correctly rounded value --
// inside the precision GLSL grants these built-ins.
__device__ __forceinline__ f3 fast_normalize(f3 a) {
    const float r = __builtin_amdgcn_rsqf(dot3(a, a));
    return mk(a.x * r, a.y * r, a.z * r);
}
__device__ __forceinline__ float fast_pow(float x, float y) {  // x >= 0
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}

__device__ __forceinline__ float2 oct_encode(f3 v) {
    const float l1 = fabsf(v.x) + fabsf(v.y) + fabsf(v.z);
    float2 r = make_float2(v.x * (1.0f / l1), v.y * (1.0f / l1));
    if (v.z < 0.0f) r = make_float2((1.0f - fabsf(r.y)) * sign_not_zero(r.x), (1.0f - fabsf(r.x)) * sign_not_zero(r.y));
    return r;
}

// textureCoordFromDirection: normalized atlas coordinate of direction `dir` in probe `probe`'s tile.
// octEncode divides by the L1 norm, so `dir` need not be normalized first.
struct AtlasGeom {
    int width, height, side, perRow;
    float invW, invH;
};
__device__ __forceinline__ float2 atlas_coord(f3 dir, int probe, const AtlasGeom& g) {
    const float2 oc = oct_encode(dir);
    const float ox = (oc.x + 1.0f) * 0.5f, oy = (oc.y + 1.0f) * 0.5f;
    const float pws = float(g.side) + 2.0f;
    const float tlx = float(probe % g.perRow) * pws + 2.0f, tly = float(probe / g.perRow) * pws + 2.0f;
    return make_float2((tlx + ox * float(g.side)) * g.invW, (tly + oy * float(g.side)) * g.invH);
}
__host__ __device__ inline AtlasGeom atlas_geom(int width, int height, int side) {
    return AtlasGeom{width, height, side, (width - 2) / (side + 2), 1.0f / float(width), 1.0f / float(height)};
}

__device__ __forceinline__ int wrap_index(int i, int n) {
    const int m = i % n;
    return m < 0 ? m + n : m;
}

// texture(sampler2D, uv): bilinear, repeat addressing
template <class S>
__device__ __forceinline__ float4 bilinear4(const S* __restrict__ img, int W, int H, float u, float v) {
    const float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrap_index(int(fx), W), x1 = wrap_index(int(fx) + 1, W);
    const int y0 = wrap_index(int(fy), H), y1 = wrap_index(int(fy) + 1, H);
    const float4 t00 = Texel<S>::ld4(img, (size_t)y0 * W + x0), t10 = Texel<S>::ld4(img, (size_t)y0 * W + x1);
    const float4 t01 = Texel<S>::ld4(img, (size_t)y1 * W + x0), t11 = Texel<S>::ld4(img, (size_t)y1 * W + x1);
    return make_float4((t00.x * (1.0f - ax) + t10.x * ax) * (1.0f - ay) + (t01.x * (1.0f - ax) + t11.x * ax) * ay,
                       (t00.y * (1.0f - ax) + t10.y * ax) * (1.0f - ay) + (t01.y * (1.0f - ax) + t11.y * ax) * ay,
                       (t00.z * (1.0f - ax) + t10.z * ax) * (1.0f - ay) + (t01.z * (1.0f - ax) + t11.z * ax) * ay,
                       0.0f);
}
template <class S>
__device__ __forceinline__ float2 bilinear2(const S* __restrict__ img, int W, int H, float u, float v) {
    const float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrap_index(int(fx), W), x1 = wrap_index(int(fx) + 1, W);
    const int y0 = wrap_index(int(fy), H), y1 = wrap_index(int(fy) + 1, H);
    const float2 t00 = Texel<S>::ld2(img, (size_t)y0 * W + x0), t10 = Texel<S>::ld2(img, (size_t)y0 * W + x1);
    const float2 t01 = Texel<S>::ld2(img, (size_t)y1 * W + x0), t11 = Texel<S>::ld2(img, (size_t)y1 * W + x1);
    return make_float2((t00.x * (1.0f - ax) + t10.x * ax) * (1.0f - ay) + (t01.x * (1.0f - ax) + t11.x * ax) * ay,
                       (t00.y * (1.0f - ax) + t10.y * ax) * (1.0f - ay) + (t01.y * (1.0f - ax) + t11.y * ax) * ay);
}

struct SampleArgs {
    LuxgiDDGIUniform u;
    LuxgiSampleProbeParams pc;
    AtlasGeom irr, dep;
    int W, H;
};

// sampleIrradiance (DDGICommon.glsl:163-233)
template <class SA>
__device__ f3 sample_irradiance(const SampleArgs& A, f3 P, f3 N, f3 Wo, const SA* __restrict__ irr,
                                const SA* __restrict__ dep) {
    const LuxgiDDGIUniform& u = A.u;
    const f3 start = mk(u.startPosition[0], u.startPosition[1], u.startPosition[2]);
    const f3 stp = mk(u.step[0], u.step[1], u.step[2]);
    const int bx = clampi(int((P.x - start.x) / stp.x), 0, u.probeCounts[0] - 1);
    const int by = clampi(int((P.y - start.y) / stp.y), 0, u.probeCounts[1] - 1);
    const int bz = clampi(int((P.z - start.z) / stp.z), 0, u.probeCounts[2] - 1);
    const f3 basePos = mk(stp.x * float(bx) + start.x, stp.y * float(by) + start.y, stp.z * float(bz) + start.z);
    const f3 alpha = mk(clampf((P.x - basePos.x) / stp.x, 0.0f, 1.0f), clampf((P.y - basePos.y) / stp.y, 0.0f, 1.0f),
                        clampf((P.z - basePos.z) / stp.z, 0.0f, 1.0f));
    const float ge = u.ddgiGamma * 0.5f;
    const f3 vBias = (N + Wo * 3.0f) * u.normalBias;
    f3 sum = mk(0.f, 0.f, 0.f);
    float sumW = 0.0f;
#pragma unroll 2
    for (int i = 0; i < 8; ++i) {
        const int ox = i & 1, oy = (i >> 1) & 1, oz = (i >> 2) & 1;
        const int gx = min(bx + ox, u.probeCounts[0] - 1), gy = min(by + oy, u.probeCounts[1] - 1),
                  gz = min(bz + oz, u.probeCounts[2] - 1);
        const f3 probePos = mk(stp.x * float(gx) + start.x, stp.y * float(gy) + start.y, stp.z * float(gz) + start.z);
        const float tri = (ox ? alpha.x : 1.0f - alpha.x) * (oy ? alpha.y : 1.0f - alpha.y) *
                          (oz ? alpha.z : 1.0f - alpha.z);
        float w = 1.0f;
        const f3 dirToProbe = fast_normalize(probePos - P);
        const float sb = fmaxf(0.0001f, (dot3(dirToProbe, N) + 1.0f) * 0.5f);
        w *= sb * sb + 0.2f;
        const int probeIdx = gx + gy * u.probeCounts[0] + gz * u.probeCounts[0] * u.probeCounts[1];
        const f3 probeToPoint = P - probePos + vBias;
        // textureCoordFromDirection(-normalize(-probeToPoint)) == the direction of probeToPoint
        float2 tc = atlas_coord(probeToPoint, probeIdx, A.dep);
        const float dist = len3(probeToPoint);
        const float2 dm = bilinear2<SA>(dep, A.dep.width, A.dep.height, tc.x, tc.y);
        const float mean = dm.x;
        const float variance = fabsf(dm.x * dm.x - dm.y);
        const float md = fmaxf(dist - mean, 0.0f);
        float cheb = variance / (variance + md * md);
        cheb = fmaxf(cheb * cheb * cheb, 0.0f);
        w *= (dist <= mean) ? 1.0f : cheb;
        w = fmaxf(0.000001f, w);
        tc = atlas_coord(N, probeIdx, A.irr);
        const float4 ir = bilinear4<SA>(irr, A.irr.width, A.irr.height, tc.x, tc.y);
        const f3 pi = mk(fast_pow(ir.x, ge), fast_pow(ir.y, ge), fast_pow(ir.z, ge));
        if (w < 0.2f) w *= w * w * (1.0f / (0.2f * 0.2f));
        w *= tri;
        sum = sum + pi * w;
        sumW += w;
    }
    f3 net = sum * (1.0f / sumW);
    net = net * net;
    return net * (2.0f * kPiF);
}

template <class SA, class SO>
__global__ __launch_bounds__(256) void sample_probes_kernel(SampleArgs A, const SA* __restrict__ irr,
                                                            const SA* __restrict__ dep, const float* __restrict__ depth,
                                                            const float4* __restrict__ normals, SO* __restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.W || y >= A.H) return;
    const size_t e = (size_t)y * A.W + x;
    const float d = depth[e];
    float4 o = make_float4(0.f, 0.f, 0.f, 0.f);
    if (d != 1.0f) {
        const float tu = (float(x) + 0.5f) / float(A.W), tv = (float(y) + 0.5f) / float(A.H);
        const f3 P = world_from_depth(tu, tv, d, A.pc.viewProjInv);
        const float2 on = reinterpret_cast<const float2*>(normals)[2 * e];
        const f3 N = oct_decode(on.x, on.y);
        const f3 Wo = fast_normalize(mk(A.pc.cameraPosition[0], A.pc.cameraPosition[1], A.pc.cameraPosition[2]) - P);
        const f3 ir = sample_irradiance<SA>(A, P, N, Wo, irr, dep);
        o = make_float4(ir.x, ir.y, ir.z, 1.0f);
    }
    Texel<SO>::st4(out, e, o);
}

template <class SA, class SO>
static void launch_sample(const SampleArgs& A, const LuxgiProbeAtlas* atlas, const float* depth, const float* normals,
                          void* out, hipStream_t st) {
    hipLaunchKernelGGL((sample_probes_kernel<SA, SO>), dim3((A.W + 15) / 16, (A.H + 15) / 16), dim3(256), 0, st, A,
                       static_cast<const SA*>(atlas->irradiance), static_cast<const SA*>(atlas->depth), depth,
                       reinterpret_cast<const float4*>(normals), static_cast<SO*>(out));
}

}  // namespace luxgi

using namespace luxgi;

extern "C" int luxgi_ddgi_sample_probes(const LuxgiDDGIUniform* u, const LuxgiProbeAtlas* atlas,
                                        const LuxgiSampleProbeParams* params, int32_t W, int32_t H, const float* depth,
                                        const float* normals, int32_t outStorage, void* out, void* stream) {
    if (!u || !atlas || !atlas->irradiance || !atlas->depth || !params || W <= 0 || H <= 0 || !depth || !normals ||
        !out)
        return LUXGI_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < 3; ++i)
        if (u->probeCounts[i] <= 0 || !(u->step[i] > 0.0f)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (u->irradianceProbeSideLength <= 0 || u->depthProbeSideLength <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    // the atlas must hold every probe's bordered tile (DDGIRenderer.cpp:182-185)
    const int px = u->probeCounts[0], py = u->probeCounts[1], pz = u->probeCounts[2];
    if (u->irradianceTextureWidth < (u->irradianceProbeSideLength + 2) * px * py + 2 ||
        u->irradianceTextureHeight < (u->irradianceProbeSideLength + 2) * pz + 2 ||
        u->depthTextureWidth < (u->depthProbeSideLength + 2) * px * py + 2 ||
        u->depthTextureHeight < (u->depthProbeSideLength + 2) * pz + 2)
        return LUXGI_ERR_INVALID_ARGUMENT;
    const bool a32 = atlas->storage == LUXGI_STORAGE_FP32, a16 = atlas->storage == LUXGI_STORAGE_FP16;
    const bool o32 = outStorage == LUXGI_STORAGE_FP32, o16 = outStorage == LUXGI_STORAGE_FP16;
    if (!(a32 || a16) || !(o32 || o16)) return LUXGI_ERR_UNSUPPORTED;
    SampleArgs A;
    A.u = *u;
    A.pc = *params;
    A.irr = atlas_geom(u->irradianceTextureWidth, u->irradianceTextureHeight, u->irradianceProbeSideLength);
    A.dep = atlas_geom(u->depthTextureWidth, u->depthTextureHeight, u->depthProbeSideLength);
    A.W = W;
    A.H = H;
    hipStream_t st = as_stream(stream);
    if (a32 && o32) launch_sample<float, float>(A, atlas, depth, normals, out, st);
    else if (a32 && o16) launch_sample<float, __half>(A, atlas, depth, normals, out, st);
    else if (a16 && o32) launch_sample<__half, float>(A, atlas, depth, normals, out, st);
    else launch_sample<__half, __half>(A, atlas, depth, normals, out, st);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

// probe_common.hpp -- DDGI probe-atlas sampling shared by the per-pixel probe sampling and the
// surface-cache bounce: sampleIrradiance (Shaders/DDGI/DDGICommon.glsl:163-233),
// textureCoordFromDirection (:141-158), octEncode (:65-72), and bilinear, repeat-addressed
// texture reads (the engine's default sampler, RHI/Definitions.h:152-159) with exact fp32
// sub-texel weights.
#pragma once
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

// Relocation and classification read by the consumers (LuxgiProbeAtlas.probeOffsets/probeStates);
// null pointers give the reference's grid positions with every probe active.
struct ProbeReloc {
    const float4* offsets;
    const uint8_t* states;
};
__host__ __device__ inline ProbeReloc probe_reloc(const LuxgiProbeAtlas* a) {
    return a ? ProbeReloc{reinterpret_cast<const float4*>(a->probeOffsets), a->probeStates} : ProbeReloc{nullptr, nullptr};
}

// sampleIrradiance (DDGICommon.glsl:163-233). With relocation (rl.offsets) each neighbour probe is
// read at its grid position + offset (the trilinear weights stay on the grid); with classification
// (rl.states) inactive neighbours are skipped, and a point with no weighted neighbour gets 0 -- the
// irradiance gather of the published rule (Majercik et al., JCGT 2021). The reference has neither.
template <class SA>
__device__ f3 sample_irradiance(const LuxgiDDGIUniform& u, const AtlasGeom& irrG, const AtlasGeom& depG, f3 P, f3 N,
                                f3 Wo, const SA* __restrict__ irr, const SA* __restrict__ dep, ProbeReloc rl = {}) {
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
        const int probeIdx = gx + gy * u.probeCounts[0] + gz * u.probeCounts[0] * u.probeCounts[1];
        if (rl.states && rl.states[probeIdx] != 0) continue;
        f3 probePos = mk(stp.x * float(gx) + start.x, stp.y * float(gy) + start.y, stp.z * float(gz) + start.z);
        if (rl.offsets) {
            const float4 o = rl.offsets[probeIdx];
            probePos = probePos + mk(o.x, o.y, o.z);
        }
        const float tri = (ox ? alpha.x : 1.0f - alpha.x) * (oy ? alpha.y : 1.0f - alpha.y) *
                          (oz ? alpha.z : 1.0f - alpha.z);
        float w = 1.0f;
        const f3 dirToProbe = fast_normalize(probePos - P);
        const float sb = fmaxf(0.0001f, (dot3(dirToProbe, N) + 1.0f) * 0.5f);
        w *= sb * sb + 0.2f;
        const f3 probeToPoint = P - probePos + vBias;
        // textureCoordFromDirection(-normalize(-probeToPoint)) == the direction of probeToPoint
        float2 tc = atlas_coord(probeToPoint, probeIdx, depG);
        const float dist = len3(probeToPoint);
        const float2 dm = bilinear2<SA>(dep, depG.width, depG.height, tc.x, tc.y);
        const float mean = dm.x;
        const float variance = fabsf(dm.x * dm.x - dm.y);
        const float md = fmaxf(dist - mean, 0.0f);
        float cheb = variance / (variance + md * md);
        cheb = fmaxf(cheb * cheb * cheb, 0.0f);
        w *= (dist <= mean) ? 1.0f : cheb;
        w = fmaxf(0.000001f, w);
        tc = atlas_coord(N, probeIdx, irrG);
        const float4 ir = bilinear4<SA>(irr, irrG.width, irrG.height, tc.x, tc.y);
        const f3 pi = mk(fast_pow(ir.x, ge), fast_pow(ir.y, ge), fast_pow(ir.z, ge));
        if (w < 0.2f) w *= w * w * (1.0f / (0.2f * 0.2f));
        w *= tri;
        sum = sum + pi * w;
        sumW += w;
    }
    if (sumW == 0.0f) return mk(0.f, 0.f, 0.f);  // every weighted neighbour inactive
    f3 net = sum * (1.0f / sumW);
    net = net * net;
    return net * (2.0f * kPiF);
}

// probe grid and atlas sizes that sample_irradiance can address: positive counts and steps, and
// atlases holding every probe's bordered tile (DDGIRenderer.cpp:182-185)
inline bool probe_atlas_geometry_valid(const LuxgiDDGIUniform& u) {
    for (int i = 0; i < 3; ++i)
        if (u.probeCounts[i] <= 0 || !(u.step[i] > 0.0f)) return false;
    if (u.irradianceProbeSideLength <= 0 || u.depthProbeSideLength <= 0) return false;
    const long long pxy = (long long)u.probeCounts[0] * u.probeCounts[1], pz = u.probeCounts[2];
    return u.irradianceTextureWidth >= (u.irradianceProbeSideLength + 2) * pxy + 2 &&
           u.irradianceTextureHeight >= (u.irradianceProbeSideLength + 2) * pz + 2 &&
           u.depthTextureWidth >= (u.depthProbeSideLength + 2) * pxy + 2 &&
           u.depthTextureHeight >= (u.depthProbeSideLength + 2) * pz + 2;
}

}  // namespace luxgi

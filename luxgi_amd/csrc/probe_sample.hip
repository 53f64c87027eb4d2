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
#include "probe_common.hpp"

namespace luxgi {

struct SampleArgs {
    LuxgiDDGIUniform u;
    LuxgiSampleProbeParams pc;
    AtlasGeom irr, dep;
    ProbeReloc rl;
    int W, H;
};

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
        const f3 ir = sample_irradiance<SA>(A.u, A.irr, A.dep, P, N, Wo, irr, dep, A.rl);
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
    if (!probe_atlas_geometry_valid(*u)) return LUXGI_ERR_INVALID_ARGUMENT;
    const bool a32 = atlas->storage == LUXGI_STORAGE_FP32, a16 = atlas->storage == LUXGI_STORAGE_FP16;
    const bool o32 = outStorage == LUXGI_STORAGE_FP32, o16 = outStorage == LUXGI_STORAGE_FP16;
    if (!(a32 || a16) || !(o32 || o16)) return LUXGI_ERR_UNSUPPORTED;
    SampleArgs A;
    A.u = *u;
    A.pc = *params;
    A.irr = atlas_geom(u->irradianceTextureWidth, u->irradianceTextureHeight, u->irradianceProbeSideLength);
    A.dep = atlas_geom(u->depthTextureWidth, u->depthTextureHeight, u->depthProbeSideLength);
    A.rl = probe_reloc(atlas);
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

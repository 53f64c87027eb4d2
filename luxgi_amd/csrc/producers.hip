// producers.hip -- the noisy signals the SVGF passes denoise, traced through the global SDF, for gfx950.
//
//   luxgi_sdf_shadow_mask   SDFShadow.comp:113-157 (dispatch RaytracedShadow.cpp:353-355): one
//                           soft-shadow ray per pixel toward a point of the light's disk picked by
//                           the blue-noise sampler, packed as one visibility bit per pixel into one
//                           R32UI word per 8x4 pixel block (bit x + 8 y).
// The block of 8x4 pixels is half a wave here: lanes 0-31 and 32-63 of a wave are two blocks, and
// the word is the OR of a 32-lane ballot, stored by the block's first lane when its pixel is not
// sky -- as the reference stores it (its barrier sits inside the non-sky branch, SDFShadow.comp:153-156).
//   luxgi_sdf_reflection    SDFReflection.comp:116-162 (dispatch RaytracedReflection.cpp:307-344):
//                           one reflection ray per pixel (mirror, blue-noise GGX lobe, or the DDGI
//                           probes for rough pixels), coloured by the surface-atlas lookup at the hit
//                           or the sky. One lane per pixel, 16x16 pixels per block.
#include "atlas_sample.hpp"
#include "probe_common.hpp"
#include "sdf_trace.hpp"

namespace luxgi {

// sampleBlueNoise (Raytraced/BlueNoise.glsl:8-18) over RGBA8 tables: sobol[256], scrambling[128][128].
// int(clamp(texel * 256, 0, 255)) of a UNORM8 texel v is v itself (v/255*256 lies in [v, v+1)
// for v < 255, and 255 clamps), so the sampler works on the bytes.
__device__ __forceinline__ float blue_noise(const uint8_t* __restrict__ sobol, const uint8_t* __restrict__ scr,
                                            int x, int y, int frame, int dim) {
    x %= 128;
    y %= 128;
    frame %= 256;
    dim %= 4;
    const uint8_t* s = scr + ((size_t)y * 128 + x) * 4;
    const int ranked = frame ^ int(s[2]);
    int value = int(sobol[(size_t)ranked * 4 + dim]);
    value ^= int(s[dim % 2]);
    return (0.5f + float(value)) / 256.0f;
}

struct ShadowArgs {
    LuxgiSdfShadowParams p;
    LuxgiGlobalSDFData g;
    int W, H, mw, mh, maxSteps;
};

// fetchLight with softShadow (SDFShadow.comp:43-109)
__device__ __forceinline__ void fetch_light_soft(const LuxgiLight& L, f3 wp, f3 n, float rx, float ry, f3& Wi,
                                                 float& tmax, float& atten) {
    f3 ld = mk(0.f, 0.f, 0.f);
    float radius = 0.0f;
    tmax = 0.0f;
    atten = 0.0f;
    if (L.type == 0.0f) {
        ld = mk(-L.direction[0], -L.direction[1], -L.direction[2]);
        tmax = kGlobalSdfWorldSize;
        radius = L.direction[3];
        atten = 1.0f;
    } else if (L.type == 2.0f) {
        const f3 dv = mk(L.position[0], L.position[1], L.position[2]) - wp;
        const float dist = len3(dv);
        ld = normalize3(dv);
        atten = 1.0f;
        tmax = dist;
        radius = L.direction[3] / dist;
    } else if (L.type == 1.0f) {
        const f3 dv = mk(L.position[0], L.position[1], L.position[2]) - wp;
        const float cutoff = 1.0f - L.angle;
        ld = normalize3(dv);
        const float dist = len3(dv);
        const float theta = dot3(ld, mk(L.direction[0], L.direction[1], L.direction[2]));
        const float e = cutoff - cutoff * 0.9f;
        atten = (theta - cutoff) / e * (L.radius / (dist * dist + 1.0f));
        atten = clampf(atten, 0.0f, 1.0f);
        tmax = dist;
        radius = L.direction[3] / dist;
    }
    // cross(lightDir, (0, 1, 0)) and cross(tangent, lightDir)
    const f3 t = normalize3(mk(-ld.z, 0.0f, ld.x));
    const f3 b = normalize3(mk(t.y * ld.z - t.z * ld.y, t.z * ld.x - t.x * ld.z, t.x * ld.y - t.y * ld.x));
    const float pr = radius * sqrtf(rx);
    const float pa = ry * 2.0f * kPiF;
    const float dx = pr * cosf(pa), dy = pr * sinf(pa);
    Wi = normalize3(ld + t * dx + b * dy);
    atten *= clampf(dot3(n, Wi), 0.0f, 1.0f);
}

template <class SS, int FL>
__global__ __launch_bounds__(256) void shadow_mask_kernel(ShadowArgs A, const float* __restrict__ depth,
                                                          const float* __restrict__ normals,
                                                          const uint8_t* __restrict__ sobol,
                                                          const uint8_t* __restrict__ scr, SdfView<SS> sdf,
                                                          uint32_t* __restrict__ mask) {
    // block = 2 x 4 pixel blocks of 8x4 (16 x 8 pixels); lanes 0-31 of a wave are one 8x4 block
    const int sub = threadIdx.x >> 5, l = threadIdx.x & 31;
    const int wx = blockIdx.x * 2 + (sub & 1), wy = blockIdx.y * 4 + (sub >> 1);
    const int x = wx * 8 + (l & 7), y = wy * 4 + (l >> 3);
    const bool inside = wx < A.mw && wy < A.mh;
    // texture() with the default sampler: repeat addressing past the image edge
    const int sx = x % A.W, sy = y % A.H;
    const size_t e = (size_t)sy * A.W + sx;
    const float d = inside ? depth[e] : 1.0f;
    uint32_t result = 0;
    if (d != 1.0f) {
        const float tu = (float(x) + 0.5f) / float(A.W), tv = (float(y) + 0.5f) / float(A.H);
        const f3 wp = world_from_depth(tu, tv, d, A.p.viewProjInv);
        const float2 on = reinterpret_cast<const float2*>(normals)[2 * e];
        const f3 n = oct_decode(on.x, on.y);
        const float rx = blue_noise(sobol, scr, x, y, int(A.p.numFrames), 0);
        const float ry = blue_noise(sobol, scr, x, y, int(A.p.numFrames), 1);
        f3 Wi;
        float tmax, atten;
        fetch_light_soft(A.p.light, wp, n, rx, ry, Wi, tmax, atten);
        const float NoL = dot3(n, Wi);
        if (NoL > 0.0f && atten > 0.0f) {
            const float sb = A.p.shadowBias;
            const float bias = 2.0f * sb * clampf(1.0f - NoL, 0.0f, 1.0f) + sb;
            const SdfHitTime h = trace_global_sdf<SS, FL>(sdf, A.g, wp + n * sb, Wi, tmax - bias, 0.0f, A.maxSteps);
            result = h.time >= 0.0f ? 0u : 1u;
        }
    }
    // OR of the 32 lanes of this half wave
    const unsigned long long bits = __ballot(result != 0u);
    const uint32_t word = (uint32_t)(bits >> (threadIdx.x & 32));
    if (l == 0 && inside && d != 1.0f) mask[(size_t)wy * A.mw + wx] = word;
}

struct ReflArgsP {
    LuxgiSdfReflectionParams p;
    LuxgiGlobalSDFData g;
    LuxgiSurfaceAtlasData ad;
    LuxgiDDGIUniform u;
    AtlasGeom irrG, depG;
    ProbeReloc rl;
    int W, H, maxSteps, skySize, outHalf, ddgi;
    float sky[4];
};

// importanceSampleGGX (Raytraced/BRDF.glsl:176-200): the half vector (the pdf is not used here)
__device__ __forceinline__ f3 importance_sample_ggx(float ex, float ey, f3 N, float roughness) {
    const float a = roughness * roughness, m2 = a * a;
    const float phi = 2.0f * kPiF * ex;
    const float cosT = sqrtf((1.0f - ey) / (1.0f + (m2 - 1.0f) * ey));
    const float sinT = sqrtf(1.0f - cosT * cosT);
    const f3 Hl = mk(cosf(phi) * sinT, sinf(phi) * sinT, cosT);
    const f3 up = fabsf(N.z) < 0.999f ? mk(0.f, 0.f, 1.f) : mk(1.f, 0.f, 0.f);
    const f3 t = normalize3(mk(up.y * N.z - up.z * N.y, up.z * N.x - up.x * N.z, up.x * N.y - up.y * N.x));
    const f3 b = mk(N.y * t.z - N.z * t.y, N.z * t.x - N.x * t.z, N.x * t.y - N.y * t.x);
    return normalize3(t * Hl.x + b * Hl.y + N * Hl.z);
}

__device__ __forceinline__ f3 reflect3(f3 I, f3 N) { return I - N * (2.0f * dot3(N, I)); }

template <class SS, int FL, class SC, class SA>
__global__ __launch_bounds__(256) void reflection_kernel(ReflArgsP A, const float* __restrict__ normals,
                                                         const float* __restrict__ pbr, const float* __restrict__ depth,
                                                         const uint8_t* __restrict__ sobol,
                                                         const uint8_t* __restrict__ scr, SdfView<SS> sdf,
                                                         const uint32_t* __restrict__ chunks,
                                                         const uint32_t* __restrict__ culled,
                                                         const LuxgiAtlasObject* __restrict__ objects,
                                                         const float* __restrict__ objInv,
                                                         const LuxgiAtlasTile* __restrict__ tiles,
                                                         const SC* __restrict__ color, const float* __restrict__ atlasDepth,
                                                         const float* __restrict__ skyFaces,
                                                         const SA* __restrict__ irr, const SA* __restrict__ dep,
                                                         void* __restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.W || y >= A.H) return;
    const size_t e = (size_t)y * A.W + x;
    const float d = depth[e];
    if (d == 1.0f) return;  // sky: the reference leaves the texel as it was
    const float tu = (float(x) + 0.5f) / float(A.W), tv = (float(y) + 0.5f) / float(A.H);
    f3 wp = world_from_depth(tu, tv, d, A.p.viewProjInv);
    const float roughness = pbr[e * 4 + 1];
    const float2 on = reinterpret_cast<const float2*>(normals)[2 * e];
    const f3 n = oct_decode(on.x, on.y);
    const f3 Wo = normalize3(mk(A.p.cameraPosition[0], A.p.cameraPosition[1], A.p.cameraPosition[2]) - wp);
    wp = wp + n * A.g.cascadeVoxelSize[0];  // self-occlusion bias
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    bool traceRay = true;
    f3 R;
    if (roughness < 0.05f) {  // MIRROR_REFLECTIONS_ROUGHNESS_THRESHOLD
        R = reflect3(mk(-Wo.x, -Wo.y, -Wo.z), n);
    } else if (roughness > 0.45f && A.p.approximateWithDDGI == 1u) {  // DDGI_REFLECTIONS_ROUGHNESS_THRESHOLD
        R = reflect3(mk(-Wo.x, -Wo.y, -Wo.z), n);
        traceRay = false;
        if (A.ddgi) {
            const f3 ir = sample_irradiance<SA>(A.u, A.irrG, A.depG, wp, R, Wo, irr, dep, A.rl);
            col = make_float4(A.p.roughDDGIIntensity * ir.x, A.p.roughDDGIIntensity * ir.y,
                              A.p.roughDDGIIntensity * ir.z, 0.0f);
        }
    } else {
        const float ex = blue_noise(sobol, scr, x, y, int(A.p.numFrames), 0) * A.p.trim;
        const float ey = blue_noise(sobol, scr, x, y, int(A.p.numFrames), 1) * A.p.trim;
        const f3 Hh = importance_sample_ggx(ex, ey, n, roughness);
        R = reflect3(mk(-Wo.x, -Wo.y, -Wo.z), Hh);
    }
    if (traceRay) {
        const SdfHitTime h = trace_global_sdf<SS, FL>(sdf, A.g, wp, R, kGlobalSdfWorldSize, 0.0f, A.maxSteps);
        if (h.time >= 0.0f) {
            Counts c{};
            const f3 hp = wp + R * h.time;
            col = sample_atlas<SC, false>(A.ad, chunks, culled, objects, objInv, tiles, color, atlasDepth, hp,
                                           mk(-R.x, -R.y, -R.z), A.g.cascadeVoxelSize[h.cascade] * 1.05f, c);
        } else {
            col = skyFaces ? sample_cube4(skyFaces, A.skySize, R) : make_float4(A.sky[0], A.sky[1], A.sky[2], A.sky[3]);
        }
    }
    if (A.outHalf) Texel<__half>::st4(static_cast<__half*>(out), e, col);
    else Texel<float>::st4(static_cast<float*>(out), e, col);
}

}  // namespace luxgi

using namespace luxgi;

template <class SS, int FL>
static void launch_shadow(const ShadowArgs& A, const float* depth, const float* normals, const LuxgiBlueNoise* nz,
                          const LuxgiSDFVolume* vol, uint32_t* mask, hipStream_t st) {
    hipLaunchKernelGGL((shadow_mask_kernel<SS, FL>), dim3((A.mw + 1) / 2, (A.mh + 3) / 4), dim3(256), 0, st, A, depth,
                       normals, nz->sobol, nz->scrambling, sdf_view<SS>(vol), mask);
}

extern "C" int luxgi_sdf_shadow_mask(const LuxgiSdfShadowParams* params, const LuxgiBlueNoise* noise, int32_t width,
                                     int32_t height, const float* depth, const float* normals,
                                     const LuxgiGlobalSDFData* sdfData, const LuxgiSDFVolume* sdf, uint32_t* mask,
                                     void* stream) {
    if (!params || !noise || !noise->sobol || !noise->scrambling || !depth || !normals || !mask)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0 || (long long)width * height >= (1ll << 31)) return LUXGI_ERR_INVALID_ARGUMENT;
    const int st = sdf_volume_status(sdfData, sdf);
    if (st != LUXGI_OK) return st;
    if (sdf->storage == LUXGI_STORAGE_BF16) return LUXGI_ERR_UNSUPPORTED;  // BF16: the probe-ray trace only
    ShadowArgs A;
    A.p = *params;
    A.g = *sdfData;
    A.W = width;
    A.H = height;
    A.mw = (width + 7) / 8;
    A.mh = (height + 3) / 4;
    A.maxSteps = params->maxSteps > 0 ? params->maxSteps : 250;
    hipStream_t s = as_stream(stream);
    const bool half = sdf->storage == LUXGI_STORAGE_FP16;
    switch (sdf->layout) {
        case LUXGI_LAYOUT_BRICK4:
            if (half) launch_shadow<__half, LUXGI_LAYOUT_BRICK4>(A, depth, normals, noise, sdf, mask, s);
            else launch_shadow<float, LUXGI_LAYOUT_BRICK4>(A, depth, normals, noise, sdf, mask, s);
            break;
        case LUXGI_LAYOUT_CELL8:
            if (half) launch_shadow<__half, LUXGI_LAYOUT_CELL8>(A, depth, normals, noise, sdf, mask, s);
            else launch_shadow<float, LUXGI_LAYOUT_CELL8>(A, depth, normals, noise, sdf, mask, s);
            break;
        default:
            if (half) launch_shadow<__half, LUXGI_LAYOUT_LINEAR>(A, depth, normals, noise, sdf, mask, s);
            else launch_shadow<float, LUXGI_LAYOUT_LINEAR>(A, depth, normals, noise, sdf, mask, s);
    }
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

// ------------------------------------------------------------------------- reflection entry point

static size_t refl_align256(size_t v) { return (v + 255) & ~size_t(255); }

extern "C" size_t luxgi_sdf_reflection_workspace_size(int32_t objectCapacity) {
    return refl_align256((size_t)(objectCapacity > 0 ? objectCapacity : 0) * 16 * sizeof(float));
}

__global__ void reflection_prepare_kernel(const LuxgiAtlasObject* __restrict__ objects, int n, float* __restrict__ inv) {
    object_inverse_kernel_body(objects, n, inv);
}

template <class SS, int FL, class SC, class SA>
static void launch_reflection(const ReflArgsP& A, const LuxgiGBuffer* gb, const LuxgiBlueNoise* nz,
                              const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at, const float* inv,
                              const float* skyFaces, const LuxgiProbeAtlas* probes, void* out, hipStream_t st) {
    hipLaunchKernelGGL((reflection_kernel<SS, FL, SC, SA>), dim3((A.W + 15) / 16, (A.H + 15) / 16), dim3(256), 0, st,
                       A, gb->normals, gb->pbr, gb->depth, nz->sobol, nz->scrambling, sdf_view<SS>(vol), at->chunks,
                       at->culledObjects, at->objects, inv, at->tiles, static_cast<const SC*>(at->color), at->depth,
                       skyFaces, probes ? static_cast<const SA*>(probes->irradiance) : nullptr,
                       probes ? static_cast<const SA*>(probes->depth) : nullptr, out);
}

template <class SS, int FL, class SC>
static void launch_reflection_sa(const ReflArgsP& A, const LuxgiGBuffer* gb, const LuxgiBlueNoise* nz,
                                 const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at, const float* inv,
                                 const float* sky, const LuxgiProbeAtlas* pr, void* out, hipStream_t st) {
    if (pr && pr->storage == LUXGI_STORAGE_FP16)
        launch_reflection<SS, FL, SC, __half>(A, gb, nz, vol, at, inv, sky, pr, out, st);
    else
        launch_reflection<SS, FL, SC, float>(A, gb, nz, vol, at, inv, sky, pr, out, st);
}

template <class SS, int FL>
static int launch_reflection_sc(const ReflArgsP& A, const LuxgiGBuffer* gb, const LuxgiBlueNoise* nz,
                                const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at, const float* inv,
                                const float* sky, const LuxgiProbeAtlas* pr, void* out, hipStream_t st) {
    if (at->layout == LUXGI_LAYOUT_BRICK4) {
        launch_reflection_sa<SS, FL, PackedAtlas>(A, gb, nz, vol, at, inv, sky, pr, out, st);
        return LUXGI_OK;
    }
    switch (at->colorStorage) {
        case LUXGI_STORAGE_FP32: launch_reflection_sa<SS, FL, float>(A, gb, nz, vol, at, inv, sky, pr, out, st); break;
        case LUXGI_STORAGE_FP16: launch_reflection_sa<SS, FL, __half>(A, gb, nz, vol, at, inv, sky, pr, out, st); break;
        case LUXGI_STORAGE_UNORM8: launch_reflection_sa<SS, FL, Unorm8>(A, gb, nz, vol, at, inv, sky, pr, out, st); break;
        default: return LUXGI_ERR_UNSUPPORTED;
    }
    return LUXGI_OK;
}

extern "C" int luxgi_sdf_reflection(const LuxgiSdfReflectionParams* params, const LuxgiBlueNoise* noise, int32_t width,
                                    int32_t height, const LuxgiGBuffer* gb, const LuxgiGlobalSDFData* sdfData,
                                    const LuxgiSDFVolume* sdf, const LuxgiSurfaceAtlas* at, const LuxgiSkybox* sky,
                                    const LuxgiDDGIUniform* ddgi, const LuxgiProbeAtlas* probes, int32_t outStorage,
                                    void* out, void* workspace, size_t workspaceBytes, void* stream) {
    if (!params || !noise || !noise->sobol || !noise->scrambling || !gb || !gb->normals || !gb->pbr || !gb->depth ||
        !at || !out)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (width <= 0 || height <= 0 || (long long)width * height >= (1ll << 31)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (outStorage != LUXGI_STORAGE_FP32 && outStorage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    const int st = sdf_volume_status(sdfData, sdf);
    if (st != LUXGI_OK) return st;
    if (sdf->storage == LUXGI_STORAGE_BF16) return LUXGI_ERR_UNSUPPORTED;  // BF16: the probe-ray trace only
    if (!at->chunks || !at->culledObjects || !at->objects || !at->tiles || !at->color || at->objectCapacity < 0)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (at->layout == LUXGI_LAYOUT_LINEAR) {
        if (!at->depth) return LUXGI_ERR_INVALID_ARGUMENT;
    } else if (at->layout == LUXGI_LAYOUT_BRICK4) {
        if (at->colorStorage != LUXGI_STORAGE_UNORM8 || (at->data.resolution & 3u)) return LUXGI_ERR_INVALID_ARGUMENT;
    } else {
        return LUXGI_ERR_INVALID_ARGUMENT;
    }
    if (at->data.resolution == 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if (sky && sky->faces && sky->size <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    const bool useDdgi = params->approximateWithDDGI == 1u;
    if (useDdgi) {
        if (!ddgi || !probes || !probes->irradiance || !probes->depth) return LUXGI_ERR_INVALID_ARGUMENT;
        if (probes->storage != LUXGI_STORAGE_FP32 && probes->storage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
        if (!probe_atlas_geometry_valid(*ddgi)) return LUXGI_ERR_INVALID_ARGUMENT;
    }
    if (workspaceBytes < luxgi_sdf_reflection_workspace_size(at->objectCapacity) ||
        (at->objectCapacity > 0 && !workspace))
        return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    ReflArgsP A{};
    A.p = *params;
    A.g = *sdfData;
    A.ad = at->data;
    if (useDdgi) {
        A.u = *ddgi;
        A.irrG = atlas_geom(ddgi->irradianceTextureWidth, ddgi->irradianceTextureHeight, ddgi->irradianceProbeSideLength);
        A.depG = atlas_geom(ddgi->depthTextureWidth, ddgi->depthTextureHeight, ddgi->depthProbeSideLength);
    }
    A.ddgi = useDdgi ? 1 : 0;
    A.W = width;
    A.H = height;
    A.maxSteps = params->maxSteps > 0 ? params->maxSteps : 250;
    A.skySize = sky ? sky->size : 0;
    A.outHalf = outStorage == LUXGI_STORAGE_FP16 ? 1 : 0;
    for (int i = 0; i < 4; ++i) A.sky[i] = params->skyColor[i];
    hipStream_t s = as_stream(stream);
    float* inv = static_cast<float*>(workspace);
    if (at->objectCapacity > 0) {
        hipLaunchKernelGGL(reflection_prepare_kernel, dim3((at->objectCapacity + 255) / 256), dim3(256), 0, s,
                           at->objects, at->objectCapacity, inv);
        LUXGI_LAUNCH_CHECK();
    }
    const float* skyFaces = sky && sky->faces ? sky->faces : nullptr;
    const LuxgiProbeAtlas* pr = useDdgi ? probes : nullptr;
    A.rl = probe_reloc(pr);
    const bool half = sdf->storage == LUXGI_STORAGE_FP16;
    int rc;
    switch (sdf->layout) {
        case LUXGI_LAYOUT_BRICK4:
            rc = half ? launch_reflection_sc<__half, LUXGI_LAYOUT_BRICK4>(A, gb, noise, sdf, at, inv, skyFaces, pr, out, s)
                      : launch_reflection_sc<float, LUXGI_LAYOUT_BRICK4>(A, gb, noise, sdf, at, inv, skyFaces, pr, out, s);
            break;
        case LUXGI_LAYOUT_CELL8:
            rc = half ? launch_reflection_sc<__half, LUXGI_LAYOUT_CELL8>(A, gb, noise, sdf, at, inv, skyFaces, pr, out, s)
                      : launch_reflection_sc<float, LUXGI_LAYOUT_CELL8>(A, gb, noise, sdf, at, inv, skyFaces, pr, out, s);
            break;
        default:
            rc = half ? launch_reflection_sc<__half, LUXGI_LAYOUT_LINEAR>(A, gb, noise, sdf, at, inv, skyFaces, pr, out, s)
                      : launch_reflection_sc<float, LUXGI_LAYOUT_LINEAR>(A, gb, noise, sdf, at, inv, skyFaces, pr, out, s);
    }
    if (rc != LUXGI_OK) return rc;
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

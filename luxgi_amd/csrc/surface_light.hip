// surface_light.hip -- surface-cache lighting of the global surface atlas, for gfx950.
//
// Reference passes, all drawn as one quad per atlas tile (GlobalSurfaceAtlas.cpp:554-574)
// into the light cache and blended there (GlobalSurfaceAtlas.cpp:886-1085):
//   CopyEmissive.frag:22-27          light = emissive                       (blend none)
//   SDFDeferredLight.frag:79-130     light += direct(light l), per light l   (blend add)
//   SDFAtlasIndirectLight.frag:44-68 light += DDGI bounce                   (blend add)
// Here one kernel evaluates the sum per texel: one block per tile draw, the
// block's lanes stride over the tile's texels. The quad covers the tile's full
// rect X..X+Wt-1 while the tile UV maps onto the rect minus its one-texel padding
// (extends.zw = (Wt - 1) / res, GlobalSurfaceAtlas.cpp:272-276), so texel p
// samples the G-buffer bilinearly at ((p + 0.5 - X) / Wt) * extends.zw + extends.xy,
// as the reference's fragments do. The direct term traces a shadow ray through
// the global SDF (tracyGlobalSDF with cascadeTraceStartBias 2,
// SDFDeferredLight.frag:104-112); the bounce is sampleIrradiance of the probe
// atlases (DDGICommon.glsl:163-233).
#include "probe_common.hpp"
#include "sdf_trace.hpp"

namespace luxgi {

constexpr int kMaxLights = 32;  // MAX_LIGHTS, Common/Light.glsl:30

struct SurfaceArgs {
    LuxgiSurfaceLightParams p;
    LuxgiGlobalSDFData g;
    LuxgiDDGIUniform u;
    AtlasGeom irrG, depG;
    ProbeReloc rl;
    int res;       // atlas resolution
    int sdfRes, mipRes, maxSteps;
    int indirect;  // probe atlases bound and the indirect pass requested
    int outHalf;
};

// bilinear, repeat addressing, all four channels (texture() with the default sampler)
template <class S>
__device__ __forceinline__ float4 bilinear_rgba(const S* __restrict__ img, int W, int H, float u, float v) {
    const float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f;
    const float fx = floorf(x), fy = floorf(y);
    const float ax = x - fx, ay = y - fy;
    const int x0 = wrap_index(int(fx), W), x1 = wrap_index(int(fx) + 1, W);
    const int y0 = wrap_index(int(fy), H), y1 = wrap_index(int(fy) + 1, H);
    const float4 a = Texel<S>::ld4(img, (size_t)y0 * W + x0), b = Texel<S>::ld4(img, (size_t)y0 * W + x1);
    const float4 c = Texel<S>::ld4(img, (size_t)y1 * W + x0), d = Texel<S>::ld4(img, (size_t)y1 * W + x1);
    auto lerp2 = [&](float p, float q, float r, float s) {
        return (p * (1.0f - ax) + q * ax) * (1.0f - ay) + (r * (1.0f - ax) + s * ax) * ay;
    };
    return make_float4(lerp2(a.x, b.x, c.x, d.x), lerp2(a.y, b.y, c.y, d.y), lerp2(a.z, b.z, c.z, d.z),
                       lerp2(a.w, b.w, c.w, d.w));
}

// BRDF (Raytraced/BRDF.glsl:12-79, the 7-argument form)
__device__ __forceinline__ f3 brdf_ggx(f3 albedo, f3 n, float rough, float metal, f3 view, f3 h, f3 l) {
    const f3 F0 = mk(mixf(0.04f, albedo.x, metal), mixf(0.04f, albedo.y, metal), mixf(0.04f, albedo.z, metal));
    const float cosLi = fmaxf(0.0f, dot3(n, l)), cosLh = fmaxf(0.0f, dot3(n, h)), NdotV = fmaxf(0.0f, dot3(n, view));
    const float c = clampf(1.0f - fmaxf(dot3(h, view), 0.0f), 0.0f, 1.0f);
    const float c5 = c * c * c * c * c;
    const f3 F = mk(F0.x + (1.0f - F0.x) * c5, F0.y + (1.0f - F0.y) * c5, F0.z + (1.0f - F0.z) * c5);
    const float alpha = rough * rough, a2 = alpha * alpha;
    const float denom = cosLh * cosLh * (a2 - 1.0f) + 1.0f;
    const float D = a2 / (kPiF * denom * denom);
    const float r = rough + 1.0f, k = r * r * 0.125f;
    const float G = cosLi / (cosLi * (1.0f - k) + k) * (NdotV / (NdotV * (1.0f - k) + k));
    const float spec = D * G / fmaxf(0.00001f, 4.0f * cosLi * NdotV);
    const float km = (1.0f - metal) * (1.0f / kPiF);
    return mk((1.0f - F.x) * km * albedo.x + F.x * spec, (1.0f - F.y) * km * albedo.y + F.y * spec,
              (1.0f - F.z) * km * albedo.z + F.z * spec);
}

// fetchLight (SDFDeferredLight.frag:41-76): types 0 directional, 1 spot, 2 point (Light.glsl:13-15)
__device__ __forceinline__ void fetch_light(const LuxgiLight& L, f3 wp, f3& Wi, float& tmax, float& atten) {
    Wi = mk(0.f, 0.f, 0.f);
    tmax = kGlobalSdfWorldSize;
    atten = 1.0f;
    const f3 lp = mk(L.position[0], L.position[1], L.position[2]);
    const f3 ld = mk(L.direction[0], L.direction[1], L.direction[2]);
    if (L.type == 0.0f) {
        Wi = mk(-ld.x, -ld.y, -ld.z);
    } else if (L.type == 2.0f) {
        const f3 dv = lp - wp;
        const float dist = len3(dv);
        Wi = normalize3(dv);
        atten = L.radius / (dist * dist + 1.0f);
        tmax = dist;
    } else if (L.type == 1.0f) {
        const f3 dv = lp - wp;
        const float cutoff = 1.0f - L.angle;
        Wi = normalize3(dv);
        const float dist = len3(dv);
        const float theta = dot3(Wi, ld);
        const float e = cutoff - cutoff * 0.9f;
        atten = (theta - cutoff) / e * (L.radius / (dist * dist + 1.0f));
        atten = clampf(atten, 0.0f, 1.0f);
        tmax = dist;
    }
}

template <class SS, int FL, class SA>
__global__ __launch_bounds__(256) void surface_light_kernel(SurfaceArgs A, const LuxgiLight* __restrict__ lights,
                                                            const LuxgiAtlasTile* __restrict__ tiles, int numTiles,
                                                            const LuxgiSurfaceDraw* __restrict__ draws,
                                                            const __half* __restrict__ gbColor,
                                                            const __half* __restrict__ gbNormal,
                                                            const Unorm8* __restrict__ gbPbr,
                                                            const __half* __restrict__ gbEmissive, SdfView<SS> sdf,
                                                            const SA* __restrict__ irr, const SA* __restrict__ dep,
                                                            void* __restrict__ out) {
    __shared__ LuxgiLight sl[kMaxLights];
    const LuxgiSurfaceDraw dr = draws[blockIdx.x];
    if (dr.tile >= (uint32_t)numTiles) return;  // uniform over the block
    const int nl = (A.p.passes & LUXGI_SURFACE_DIRECT) ? A.p.numLights : 0;
    for (int i = threadIdx.x; i < nl; i += blockDim.x) sl[i] = lights[i];
    __syncthreads();
    const LuxgiAtlasTile& t = tiles[dr.tile];
    const float res = float(A.res);
    const int X = int(rintf(t.extends[0] * res)), Y = int(rintf(t.extends[1] * res));
    const int Wt = int(rintf(t.extends[2] * res)) + 1, Ht = int(rintf(t.extends[3] * res)) + 1;
    const f3 cam = mk(A.p.cameraPosition[0], A.p.cameraPosition[1], A.p.cameraPosition[2]);
    const float sb = A.p.cameraPosition[3];
    for (int k = threadIdx.x; k < Wt * Ht; k += blockDim.x) {
        const int px = X + k % Wt, py = Y + k / Wt;
        if (px < 0 || py < 0 || px >= A.res || py >= A.res) continue;
        const float au = ((float(px) + 0.5f - float(X)) / float(Wt)) * t.extends[2] + t.extends[0];
        const float av = ((float(py) + 0.5f - float(Y)) / float(Ht)) * t.extends[3] + t.extends[1];
        f3 acc = mk(0.f, 0.f, 0.f);
        if (A.p.passes & LUXGI_SURFACE_EMISSIVE) {
            const float4 e = bilinear_rgba<__half>(gbEmissive, A.res, A.res, au, av);
            acc = mk(e.x, e.y, e.z);
        }
        if (nl > 0 || A.indirect) {
            const float4 c = bilinear_rgba<__half>(gbColor, A.res, A.res, au, av);
            const float4 nt = bilinear_rgba<__half>(gbNormal, A.res, A.res, au, av);
            const float4 pbr = bilinear_rgba<Unorm8>(gbPbr, A.res, A.res, au, av);
            const f3 n = oct_decode(nt.x, nt.y);
            const f3 wp = mk(c.w, nt.z, nt.w);
            const f3 view = normalize3(cam - wp);
            for (int l = 0; l < nl; ++l) {
                if (!((dr.lightMask >> l) & 1u)) continue;
                f3 L;
                float dist, atten;
                fetch_light(sl[l], wp, L, dist, atten);
                const float NoL = dot3(n, L);
                const float bias = 2.0f * sb * clampf(1.0f - NoL, 0.0f, 1.0f) + sb;
                float shadow = 0.0f;
                if (NoL > 0.0f) {
                    shadow = 1.0f;
                    if (atten > 0.0f) {
                                const SdfHitTime h = trace_global_sdf<SS, FL>(sdf, A.g, wp + n * sb, L, dist - bias,
                                                                         2.0f, A.maxSteps);
                        shadow = h.time >= 0.0f ? 0.0f : 1.0f;
                    }
                }
                const float inten = __powf(sl[l].intensity, 1.4f) + 0.1f;
                const f3 Lrad = mk(sl[l].color[0] * inten, sl[l].color[1] * inten, sl[l].color[2] * inten);
                const f3 Lh = normalize3(L + view);
                const float cosLi = fmaxf(0.0f, dot3(n, L));
                const f3 b = brdf_ggx(mk(c.x, c.y, c.z), n, pbr.y, pbr.x, view, Lh, L);
                const float s = cosLi * shadow * atten;
                acc = acc + mk(b.x * Lrad.x * s, b.y * Lrad.y * s, b.z * Lrad.z * s);
            }
            if (A.indirect) {
                const float m = pbr.x;
                const f3 alb = mk(fminf(c.x, 0.9f), fminf(c.y, 0.9f), fminf(c.z, 0.9f));
                const f3 diff = mk((alb.x - alb.x * m) / kPiF, (alb.y - alb.y * m) / kPiF, (alb.z - alb.z * m) / kPiF);
                const f3 ir = sample_irradiance<SA>(A.u, A.irrG, A.depG, wp, n, view, irr, dep, A.rl);
                acc = acc + mk(A.p.indirectIntensity * diff.x * ir.x, A.p.indirectIntensity * diff.y * ir.y,
                               A.p.indirectIntensity * diff.z * ir.z);
            }
        }
        const size_t e = (size_t)py * A.res + px;
        const float4 o = make_float4(acc.x, acc.y, acc.z, 1.0f);
        if (A.outHalf) Texel<__half>::st4(static_cast<__half*>(out), e, o);
        else Texel<float>::st4(static_cast<float*>(out), e, o);
    }
}

}  // namespace luxgi

using namespace luxgi;

template <class SS, int FL, class SA>
static void launch_surface(const SurfaceArgs& A, const LuxgiLight* lights, const LuxgiAtlasTile* tiles,
                           int numTiles, const LuxgiSurfaceDraw* draws, int numDraws, const LuxgiSurfaceGBuffer* gb,
                           const LuxgiSDFVolume* vol, const LuxgiProbeAtlas* probes, void* out, hipStream_t st) {
    const SdfView<SS> v = sdf_view<SS>(vol);
    hipLaunchKernelGGL((surface_light_kernel<SS, FL, SA>), dim3(numDraws), dim3(256), 0, st, A, lights, tiles,
                       numTiles, draws, static_cast<const __half*>(gb->color), static_cast<const __half*>(gb->normal),
                       static_cast<const Unorm8*>(gb->pbr), static_cast<const __half*>(gb->emissive), v,
                       probes ? static_cast<const SA*>(probes->irradiance) : nullptr,
                       probes ? static_cast<const SA*>(probes->depth) : nullptr, out);
}

template <class SS, int FL>
static void launch_surface_sa(const SurfaceArgs& A, const LuxgiLight* lights, const LuxgiAtlasTile* tiles,
                              int numTiles, const LuxgiSurfaceDraw* draws, int numDraws, const LuxgiSurfaceGBuffer* gb,
                              const LuxgiSDFVolume* vol, const LuxgiProbeAtlas* probes, void* out, hipStream_t st) {
    if (probes && probes->storage == LUXGI_STORAGE_FP16)
        launch_surface<SS, FL, __half>(A, lights, tiles, numTiles, draws, numDraws, gb, vol, probes, out, st);
    else
        launch_surface<SS, FL, float>(A, lights, tiles, numTiles, draws, numDraws, gb, vol, probes, out, st);
}

extern "C" int luxgi_surface_light(const LuxgiSurfaceLightParams* params, const LuxgiLight* lights,
                                   const LuxgiAtlasTile* tiles, int32_t numTiles, const LuxgiSurfaceDraw* draws,
                                   int32_t numDraws,
                                   const LuxgiSurfaceGBuffer* gb, const LuxgiGlobalSDFData* sdfData,
                                   const LuxgiSDFVolume* vol, const LuxgiDDGIUniform* ddgi,
                                   const LuxgiProbeAtlas* probes, int32_t outStorage, void* lightCache,
                                   void* stream) {
    if (!params || !gb || !lightCache || numDraws < 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if (params->passes & ~(LUXGI_SURFACE_EMISSIVE | LUXGI_SURFACE_DIRECT | LUXGI_SURFACE_INDIRECT))
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (outStorage != LUXGI_STORAGE_FP32 && outStorage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    if (gb->resolution <= 0 || (size_t)gb->resolution * gb->resolution >= (1ull << 31)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (numDraws == 0) return LUXGI_OK;
    if (!tiles || !draws || numTiles <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    const bool direct = (params->passes & LUXGI_SURFACE_DIRECT) && params->numLights > 0;
    const bool indirect = (params->passes & LUXGI_SURFACE_INDIRECT) != 0;
    if ((params->passes & LUXGI_SURFACE_EMISSIVE) && !gb->emissive) return LUXGI_ERR_INVALID_ARGUMENT;
    if ((direct || indirect) && (!gb->color || !gb->normal || !gb->pbr)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (params->numLights < 0 || params->numLights > LUXGI_MAX_LIGHTS) return LUXGI_ERR_INVALID_ARGUMENT;
    if (direct) {
        if (!lights) return LUXGI_ERR_INVALID_ARGUMENT;
        const int st = sdf_volume_status(sdfData, vol);
        if (st != LUXGI_OK) return st;
        if (vol->storage == LUXGI_STORAGE_BF16) return LUXGI_ERR_UNSUPPORTED;  // BF16: the probe-ray trace only
    }
    if (indirect) {
        if (!ddgi || !probes || !probes->irradiance || !probes->depth) return LUXGI_ERR_INVALID_ARGUMENT;
        if (probes->storage != LUXGI_STORAGE_FP32 && probes->storage != LUXGI_STORAGE_FP16)
            return LUXGI_ERR_UNSUPPORTED;
        if (!probe_atlas_geometry_valid(*ddgi)) return LUXGI_ERR_INVALID_ARGUMENT;
    }
    SurfaceArgs A{};
    A.p = *params;
    if (!direct) A.p.numLights = 0;
    if (direct) A.g = *sdfData;
    if (indirect) {
        A.u = *ddgi;
        A.irrG = atlas_geom(ddgi->irradianceTextureWidth, ddgi->irradianceTextureHeight, ddgi->irradianceProbeSideLength);
        A.depG = atlas_geom(ddgi->depthTextureWidth, ddgi->depthTextureHeight, ddgi->depthProbeSideLength);
    }
    A.res = gb->resolution;
    A.maxSteps = params->maxSteps > 0 ? params->maxSteps : 250;
    A.indirect = indirect ? 1 : 0;
    A.outHalf = outStorage == LUXGI_STORAGE_FP16 ? 1 : 0;
    hipStream_t st = as_stream(stream);
    const LuxgiSDFVolume* v = direct ? vol : nullptr;
    const LuxgiProbeAtlas* pr = indirect ? probes : nullptr;
    A.rl = probe_reloc(pr);
    const bool half = v && v->storage == LUXGI_STORAGE_FP16;
    const int layout = v ? v->layout : LUXGI_LAYOUT_LINEAR;
#define LUXGI_SURFACE_LAUNCH(SS_, FL_) \
    launch_surface_sa<SS_, FL_>(A, lights, tiles, numTiles, draws, numDraws, gb, v, pr, lightCache, st)
    if (layout == LUXGI_LAYOUT_BRICK4) {
        if (half) LUXGI_SURFACE_LAUNCH(__half, LUXGI_LAYOUT_BRICK4);
        else LUXGI_SURFACE_LAUNCH(float, LUXGI_LAYOUT_BRICK4);
    } else if (layout == LUXGI_LAYOUT_CELL8) {
        if (half) LUXGI_SURFACE_LAUNCH(__half, LUXGI_LAYOUT_CELL8);
        else LUXGI_SURFACE_LAUNCH(float, LUXGI_LAYOUT_CELL8);
    } else {
        if (half) LUXGI_SURFACE_LAUNCH(__half, LUXGI_LAYOUT_LINEAR);
        else LUXGI_SURFACE_LAUNCH(float, LUXGI_LAYOUT_LINEAR);
    }
#undef LUXGI_SURFACE_LAUNCH
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

// sdf_trace.hpp -- tracyGlobalSDF (Shaders/SDF/SDFCommon.glsl:98-194) as a device function, for
// the passes other than the probe-ray march that trace the global SDF for a yes/no or a hit
// distance: the surface-cache shadow ray (SDFDeferredLight.frag:104-112) and the screen-space
// shadow and reflection producers. Step rule, hit test and cascade walk are the probe-ray march's
// (the trace kernel), with the per-cascade start offset dir * voxel * cascadeTraceStartBias.
#pragma once
#include "common.hpp"

namespace luxgi {

template <class SS>
struct SdfView {
    const SS* sdf;             // fine volume, LINEAR, BRICK4 or CELL8 layout
    const SS* mip;             // coarse volume, LINEAR (CELL8 with a CELL8 fine volume)
    const float* guard;        // optional brick guard (FP32 LINEAR volumes only)
    const float* brickScale;   // optional per-brick scale (BRICK4)
    int res, mipRes;
};

constexpr float kGlobalSdfWorldSize = 60000.0f;  // GLOBAL_SDF_WORLD_SIZE, SDFCommon.glsl:6

struct SdfHitTime {
    float time;  // -1 = no hit
    float sdf;
    int cascade;
};

// fine and coarse lookups in the volume's layout (FL: LuxgiLayout; CELL8 holds both in cell records)
template <class SS, int FL>
__device__ __forceinline__ float sdf_fine(const SdfView<SS>& v, int W, float u, float y, float z) {
    if constexpr (FL == LUXGI_LAYOUT_BRICK4) return sample3d_brick4<SS>(v.sdf, W, v.res, v.res, u, y, z, v.brickScale);
    else if constexpr (FL == LUXGI_LAYOUT_CELL8) return sample3d_cell8<SS>(v.sdf, W, v.res, v.res, u, y, z);
    else return sample_volume<SS>(v.sdf, W, v.res, v.res, u, y, z);
}
template <class SS, int FL>
__device__ __forceinline__ float sdf_coarse(const SdfView<SS>& v, int W, float u, float y, float z) {
    if constexpr (FL == LUXGI_LAYOUT_CELL8) return sample3d_cell8<SS>(v.mip, W, v.mipRes, v.mipRes, u, y, z);
    else return sample_volume<SS>(v.mip, W, v.mipRes, v.mipRes, u, y, z);
}

template <class SS, int FL>
__device__ SdfHitTime trace_global_sdf(const SdfView<SS>& v, const LuxgiGlobalSDFData& g, f3 o, f3 d,
                                       float maxDistance, float startBias, int maxSteps) {
    SdfHitTime hit{-1.0f, 0.0f, 0};
    const int C = (int)g.cascadesCount;
    const float traceMax = fminf(maxDistance, g.cascadePosDistance[C - 1][3] * 2.0f);
    const float chunkSizeDistance = 32.0f / g.resolution;   // GLOBAL_SDF_RASTERIZE_CHUNK_SIZE / resolution
    const float chunkMarginDistance = 4.0f / g.resolution;  // GLOBAL_SDF_RASTERIZE_CHUNK_MARGIN / resolution
    const int texW = v.res * C, mipW = v.mipRes * C;
    const f3 traceEnd = o + d * traceMax;
    float nextStart = 0.0f;
    for (int cas = 0; cas < C && hit.time < 0.0f; ++cas) {
        const f3 cpos = mk(g.cascadePosDistance[cas][0], g.cascadePosDistance[cas][1], g.cascadePosDistance[cas][2]);
        const float half = g.cascadePosDistance[cas][3];
        const float voxel = g.cascadeVoxelSize[cas];
        const f3 wp = o + d * (voxel * startBias);
        // lineHitAABB (SDFCommon.glsl:84-95) from the biased start to the unbiased end
        const f3 dl = traceEnd - wp;
        const f3 inv = mk(1.0f / dl.x, 1.0f / dl.y, 1.0f / dl.z);
        const f3 en = (cpos - mk(half, half, half) - wp) * inv;
        const f3 ex = (cpos + mk(half, half, half) - wp) * inv;
        float ix = fmaxf(fminf(en.x, ex.x), fmaxf(fminf(en.y, ex.y), fminf(en.z, ex.z)));
        float iy = fminf(fmaxf(en.x, ex.x), fminf(fmaxf(en.y, ex.y), fmaxf(en.z, ex.z)));
        ix = clampf(ix, 0.0f, 1.0f) * traceMax;
        iy = clampf(iy, 0.0f, 1.0f) * traceMax;
        ix = fmaxf(ix, nextStart);
        float stepTime = ix;
        if (ix >= iy) stepTime = iy;
        else nextStart = iy;
        const float cmax = half * 2.0f;
        // divisions by the per-cascade constants as multiplications by their reciprocals, as in
        // the probe-ray march (exact for power-of-two divisors, within one ulp otherwise)
        const float invExtent = 1.0f / cmax, invVoxel = 1.0f / voxel, invC = 1.0f / float(C);
        for (int step = 0; step < maxSteps && stepTime < iy; ++step) {
            const f3 p = wp + d * stepTime;
            const float ux = clampf((p.x - cpos.x) * invExtent + 0.5f, 0.0f, 1.0f);
            const float uy = clampf((p.y - cpos.y) * invExtent + 0.5f, 0.0f, 1.0f);
            const float uz = clampf((p.z - cpos.z) * invExtent + 0.5f, 0.0f, 1.0f);
            const float tx = (float(cas) + ux) * invC;
            bool needFine = true;
            if (v.guard) {  // a guard >= twice the margin proves the fine value is not used
                const int gx = clampi(int(floorf(tx * float(texW) - 0.5f)), 0, texW - 1);
                const int gy = clampi(int(floorf(uy * float(v.res) - 0.5f)), 0, v.res - 1);
                const int gz = clampi(int(floorf(uz * float(v.res) - 0.5f)), 0, v.res - 1);
                const float gv = v.guard[((size_t)(gz >> 2) * (v.res >> 2) + (gy >> 2)) * (texW >> 2) + (gx >> 2)];
                needFine = !(gv >= chunkMarginDistance * 2.0f * (1.0f + 1e-5f));
            }
            // both lookups are issued together; the fine value is used only under the mip test
            float sd = sdf_coarse<SS, FL>(v, mipW, tx, uy, uz);
            const float st = needFine ? sdf_fine<SS, FL>(v, texW, tx, uy, uz) : 1.0f;
            if (sd < chunkSizeDistance) {
                if (needFine && st < chunkMarginDistance * 2.0f) sd = st;
            } else {
                sd = chunkSizeDistance;
            }
            sd *= cmax;
            const float thick = voxel * 0.5f * clampf(stepTime * invVoxel, 0.0f, 1.0f);
            if (sd < thick) {
                hit.time = fmaxf(stepTime + sd - thick, 0.0f);
                hit.sdf = sd;
                hit.cascade = cas;
                break;
            }
            stepTime += fmaxf(sd, voxel);
        }
    }
    return hit;
}

// Validation of a global SDF for the tracers of this header: LuxgiStatus (LUXGI_OK when usable).
inline int sdf_volume_status(const LuxgiGlobalSDFData* g, const LuxgiSDFVolume* vol) {
    if (!g || !vol || !vol->sdf || !vol->mip) return LUXGI_ERR_INVALID_ARGUMENT;
    if (g->cascadesCount < 1 || g->cascadesCount > 4) return LUXGI_ERR_INVALID_ARGUMENT;
    if (vol->resolution <= 0 || vol->mipResolution <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if (vol->storage != LUXGI_STORAGE_FP32 && vol->storage != LUXGI_STORAGE_FP16 && vol->storage != LUXGI_STORAGE_BF16)
        return LUXGI_ERR_UNSUPPORTED;
    if (vol->layout != LUXGI_LAYOUT_LINEAR && vol->layout != LUXGI_LAYOUT_BRICK4 && vol->layout != LUXGI_LAYOUT_CELL8)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (vol->layout == LUXGI_LAYOUT_BRICK4 && (vol->resolution & 3)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (vol->brickScale && vol->layout != LUXGI_LAYOUT_BRICK4) return LUXGI_ERR_INVALID_ARGUMENT;
    // the guard bounds FP32 voxel values; FP16 storage would round below them
    if (vol->guard && ((vol->resolution & 3) || vol->storage != LUXGI_STORAGE_FP32)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (vol->refine) {
        const int rl = vol->refineLayout;
        if (rl != LUXGI_LAYOUT_LINEAR && rl != LUXGI_LAYOUT_BRICK4 && rl != LUXGI_LAYOUT_CELL8)
            return LUXGI_ERR_INVALID_ARGUMENT;
        if (rl == LUXGI_LAYOUT_BRICK4 && (vol->resolution & 3)) return LUXGI_ERR_INVALID_ARGUMENT;
    }
    // the linear and brick samplers address the volumes with 32-bit byte offsets, the cell-record
    // sampler with 32-bit record indices
    const size_t es = vol->storage == LUXGI_STORAGE_FP32 ? 4 : 2;
    const int C = (int)g->cascadesCount, r = vol->resolution, m = vol->mipResolution;
    auto too_big = [&](int layout, size_t elem) {
        if (layout == LUXGI_LAYOUT_CELL8)
            return cell8_records(C * r, r, r) >= (1ull << 32) || cell8_records(C * m, m, m) >= (1ull << 32);
        const size_t R = (size_t)r;
        const size_t fine = layout == LUXGI_LAYOUT_BRICK4 ? C * (R / 4) * (R / 4) * (R / 4) * kBrickElems * elem
                                                          : C * R * R * R * elem;
        return fine >= (1ull << 32) || (size_t)C * m * m * m * elem >= (1ull << 32);
    };
    if (too_big(vol->layout, es) || (vol->refine && too_big(vol->refineLayout, 4))) return LUXGI_ERR_UNSUPPORTED;
    return LUXGI_OK;
}

template <class SS>
inline SdfView<SS> sdf_view(const LuxgiSDFVolume* vol) {
    SdfView<SS> v{};
    if (vol) {
        v.sdf = static_cast<const SS*>(vol->sdf);
        v.mip = static_cast<const SS*>(vol->mip);
        v.guard = vol->guard;
        v.brickScale = vol->brickScale;
        v.res = vol->resolution;
        v.mipRes = vol->mipResolution;
    }
    return v;
}

}  // namespace luxgi

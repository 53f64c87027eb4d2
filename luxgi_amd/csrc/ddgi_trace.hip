// ddgi_trace.hip -- DDGI probe rays through the global SDF with surface-atlas
// radiance at the hits, for gfx950.
//
// Reference: Shaders/DDGI/GISDFRays.comp:63-128 (ray generation and encode),
// Shaders/SDF/SDFCommon.glsl:98-194 (tracyGlobalSDF), :64-71 (cascade UV),
// :84-95 (lineHitAABB), :196-199 (surface threshold),
// Shaders/SDF/AtlasCommon.glsl:50-157 (surface-atlas lookup);
// host dispatch Engine/DDGI/DDGIRenderer.cpp:304-321.
//
// One lane per ray; the 64 lanes of a wave trace 64 rays of the same probe, so
// early march steps of a wave read neighbouring voxels. The reference inverts
// each candidate object's transform per hit (AtlasCommon.glsl:134); here the
// inverses are computed once per call into the workspace (trace_prepare_kernel).
// The march (march_kernel) and the per-hit work (shade_kernel) are separate
// launches, so the march loop carries only the registers it needs.
#include "atlas_sample.hpp"
#include "sdf_trace.hpp"

#include <algorithm>
#include <type_traits>

namespace luxgi {

constexpr float kWorldSize = 60000.0f;  // GLOBAL_SDF_WORLD_SIZE, SDFCommon.glsl:6

// Everything a trace reads about its scene: the global SDF (clipmap data, fine volume, coarse mip and
// their options), the surface atlas with its object/tile/chunk buffers, and the sky. A single call
// (or a batch over one scene) carries one in TraceArgs; a batch over several scenes indexes an
// array of them in device memory (LuxgiTraceScene, the same fields).
struct TraceScene {
    LuxgiGlobalSDFData g;
    LuxgiSurfaceAtlasData ad;
    const void* sdf;
    const void* mip;
    const float* guard;       // optional brick guard (FP32 volumes)
    const float* brickScale;  // optional per-brick scale of a BRICK4 fine volume
    const float* refine;      // optional FP32 fine volume re-evaluated at hits
    const uint32_t* chunks;
    const uint32_t* culled;
    const LuxgiAtlasObject* objects;
    const LuxgiAtlasTile* tiles;
    const void* color;
    const float* atlasDepth;
    const float* skyFaces;
    long long objectStart;    // the scene's first object in the inverse-transform workspace
    int objectCapacity;       // entries of `objects`
    float sky[3];
    float invCascades;        // 1 / cascadesCount
    int res, mipRes, skySize, refineLayout;
};

struct TraceArgs {
    TraceScene scene;              // the scene of a single call or of a one-scene batch
    const TraceScene* scenes;      // a several-scene batch: per-volume scenes (device memory)
    int maxSteps, P, R, rayWidth;
    float hitEps;
    int fineLayout;                // LuxgiLayout of the fine volume (the kernels' template argument)
    int sortHits;                  // shade the hit rays grouped by surface-atlas chunk
    LuxgiTraceVolume vol;          // the volume of a single-volume call
    const LuxgiTraceVolume* vols;  // the volumes of a batched call (device memory), their count
    int nVols;                     // and the total of their rays
    long long totalRays;
};

constexpr int kChunkBins = kChunksRes * kChunksRes * kChunksRes;
constexpr uint32_t kNoChunk = 0xffffffffu;  // misses and inside hits: no atlas lookup

// Adds the wave's active lanes to counters[key] with one atomic per distinct key and returns each
// lane's slot (the counter's old value + the lane's rank among the lanes with its key). Called by
// all 64 lanes of the wave together.
__device__ __forceinline__ uint32_t wave_aggregated_add(uint32_t* __restrict__ counters, uint32_t key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(active);
    uint32_t slot = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const uint32_t lk = (uint32_t)__shfl((int)key, leader);
        const unsigned long long peers = todo & __ballot(key == lk);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counters[lk], (uint32_t)__popcll(peers));
        base = (uint32_t)__shfl((int)base, leader);
        if ((peers >> lane) & 1ull) slot = base + (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
        todo &= ~peers;
    }
    return slot;
}

// ---------------------------------------------------------------------------
// March (tracyGlobalSDF, SDFCommon.glsl:98-194): one lane per ray, the 64 lanes
// of a wave on 64 rays of one probe. Blocks are mapped to probes XCD by XCD:
// the dispatcher deals blocks to the 8 XCDs round-robin, so block b runs on XCD
// b % 8; the mapping gives each XCD a contiguous range of rays (a slab of
// probes) that it walks in order, which keeps the SDF lines its rays touch in
// its own L2. The hit record (hitTime, texture UV of the hit step, hit SDF,
// cascade) goes to the workspace and shade_kernel turns it into the outputs.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void ray_setup(const TraceArgs& A, const LuxgiTraceVolume& v, int probe, int ray,
                                          f3& origin, f3& dir) {
    // probeLocation (DDGICommon.glsl:106-114) + GISDFRays.comp:72-73
    const LuxgiDDGIUniform& u = v.uniform;
    const int px = u.probeCounts[0], pxy = u.probeCounts[0] * u.probeCounts[1];
    const int gx = probe % px, gy = (probe % pxy) / px, gz = probe / pxy;
    origin = mk(u.step[0] * float(gx) + u.startPosition[0], u.step[1] * float(gy) + u.startPosition[1],
                u.step[2] * float(gz) + u.startPosition[2]);
    if (v.probeOffsets) {
        const float4 o = reinterpret_cast<const float4*>(v.probeOffsets)[probe];
        origin = origin + mk(o.x, o.y, o.z);
    }
    const f3 sf = spherical_fibonacci(float(ray), float(u.raysPerProbe));
    const float* rot = v.deviceOrientation ? v.deviceOrientation : v.orientation;  // column-major 4x4
    float m[9];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) m[c * 3 + r] = rot[c * 4 + r];
    dir = normalize3(mk(m[0] * sf.x + m[3] * sf.y + m[6] * sf.z, m[1] * sf.x + m[4] * sf.y + m[7] * sf.z,
                        m[2] * sf.x + m[5] * sf.y + m[8] * sf.z));
}

// the volume of flat ray `gid` in a batch: the last whose rayStart <= gid
__device__ __forceinline__ int trace_volume_of(const LuxgiTraceVolume* __restrict__ vols, int n, long long gid) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (vols[mid].rayStart <= gid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// fine-volume lookup in the volume's layout (FL: LuxgiLayout)
template <class SS, int FL>
__device__ __forceinline__ float fine_lookup(const SS* __restrict__ sdf, const float* __restrict__ scale, int W, int H,
                                             int D, float u, float v, float w) {
    if constexpr (FL == LUXGI_LAYOUT_BRICK4) return sample3d_brick4<SS>(sdf, W, H, D, u, v, w, scale);
    else if constexpr (FL == LUXGI_LAYOUT_CELL8) return sample3d_cell8<SS>(sdf, W, H, D, u, v, w);
    else return sample_volume<SS>(sdf, W, H, D, u, v, w);
}

// coarse-mip lookup: CELL8 records with a CELL8 volume, linear otherwise
template <class SS, int FL>
__device__ __forceinline__ float mip_lookup(const SS* __restrict__ mip, int W, int H, int D, float u, float v,
                                            float w) {
    if constexpr (FL == LUXGI_LAYOUT_CELL8) return sample3d_cell8<SS>(mip, W, H, D, u, v, w);
    else return sample_volume<SS>(mip, W, H, D, u, v, w);
}

// the FP32 re-evaluation volume, in its own layout
__device__ __forceinline__ float refine_lookup(const TraceArgs& A, const TraceScene& S, int W, int H, int D, float u, float v, float w) {
    if (S.refineLayout == LUXGI_LAYOUT_BRICK4) return sample3d_brick4<float>(S.refine, W, H, D, u, v, w);
    if (S.refineLayout == LUXGI_LAYOUT_CELL8) return sample3d_cell8<float>(S.refine, W, H, D, u, v, w);
    return sample_volume<float>(S.refine, W, H, D, u, v, w);
}

// Result of a ray without an atlas lookup (GISDFRays.comp:93-96,121-124): an inside hit stores
// zero radiance, a miss the sky in its direction; both store the world distance.
template <class SR>
__device__ __forceinline__ void store_nonhit(const TraceArgs& A, const TraceScene& S, const float* __restrict__ skyFaces,
                                             SR* __restrict__ outRad, SR* __restrict__ outDD,
                                             uint8_t* __restrict__ flags, size_t e, f3 dir, bool inside) {
    float4 rad = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!inside) {
        const float3 s = skyFaces ? sample_cube(skyFaces, S.skySize, dir) : make_float3(S.sky[0], S.sky[1], S.sky[2]);
        rad = make_float4(s.x, s.y, s.z, 0.f);
    }
    Texel<SR>::st4(outRad, e, rad);
    Texel<SR>::st4(outDD, e, make_float4(dir.x, dir.y, dir.z, kWorldSize));
    if (flags) flags[e] = inside ? 2 : 0;
}

// Slot of a hit in its chunk's run of the sorted order: one atomic per lane, or one per distinct
// chunk of the wave.
__device__ __forceinline__ uint32_t hist_add(uint32_t* __restrict__ bins, uint32_t key, bool active) {
#if LUXGI_HIST_PER_LANE
    return active ? atomicAdd(&bins[key], 1u) : 0u;
#else
    return wave_aggregated_add(bins, key, active);
#endif
}

// State of one ray's sphere trace (SDFCommon.glsl:98-171), advanced one step at a time.
struct MarchRay {
    f3 o, d, traceEnd;
    f3 cpos;             // the current cascade's centre
    float half, voxel;   // and its half extent and voxel size
    float invExtent, invVoxel;  // 1 / (2 half) and 1 / voxel (see march_step)
    float stepTime, iy, nextStart;
    int cas, step;
};

// lineHitAABB (SDFCommon.glsl:84-95) and the entry into cascade r.cas (:121-141)
__device__ __forceinline__ void march_enter(const TraceArgs& A, const TraceScene& S, MarchRay& r, float traceMax) {
    const int cas = r.cas;
    r.cpos = mk(S.g.cascadePosDistance[cas][0], S.g.cascadePosDistance[cas][1], S.g.cascadePosDistance[cas][2]);
    r.half = S.g.cascadePosDistance[cas][3];
    r.voxel = S.g.cascadeVoxelSize[cas];
    r.invExtent = 1.0f / (r.half * 2.0f);
    r.invVoxel = 1.0f / r.voxel;
    const float half = r.half;
    const f3 dl = r.traceEnd - r.o;
    const f3 inv = mk(1.0f / dl.x, 1.0f / dl.y, 1.0f / dl.z);
    const f3 en = (r.cpos - mk(half, half, half) - r.o) * inv;
    const f3 ex = (r.cpos + mk(half, half, half) - r.o) * inv;
    float ix = fmaxf(fminf(en.x, ex.x), fmaxf(fminf(en.y, ex.y), fminf(en.z, ex.z)));
    float iy = fminf(fmaxf(en.x, ex.x), fminf(fmaxf(en.y, ex.y), fmaxf(en.z, ex.z)));
    ix = clampf(ix, 0.0f, 1.0f) * traceMax;
    iy = clampf(iy, 0.0f, 1.0f) * traceMax;
    ix = fmaxf(ix, r.nextStart);
    r.stepTime = ix;
    if (ix >= iy) r.stepTime = iy;
    else r.nextStart = iy;
    r.iy = iy;
    r.step = 0;
}

__device__ __forceinline__ void march_begin(const TraceArgs& A, const TraceScene& S, const LuxgiTraceVolume& v, int probe, int ray,
                                            float traceMax, MarchRay& r) {
    ray_setup(A, v, probe, ray, r.o, r.d);
    r.traceEnd = r.o + r.d * traceMax;
    r.nextStart = 0.0f;
    r.cas = 0;
    march_enter(A, S, r, traceMax);
}

// One step of cascade r.cas (the body of SDFCommon.glsl:143-170). Returns true on a hit and fills
// the hit record; otherwise advances stepTime.
template <class SS, int FL, bool COUNT>
__device__ __forceinline__ bool march_step(const TraceArgs& A, const TraceScene& S, const SS* __restrict__ sdf, const SS* __restrict__ mip,
                                           const float* __restrict__ guard, MarchRay& r, float4& ra, float2& rb,
                                           unsigned long long& cMip, unsigned long long& cFine) {
    const int C = (int)S.g.cascadesCount;
    const float chunkSizeDistance = 32.0f / S.g.resolution;
    const float chunkMarginDistance = 4.0f / S.g.resolution;
    const int texW = S.res * C, mipW = S.mipRes * C;
    // The cascade UV (SDFCommon.glsl:64-71) divides by the cascade extent and the cascade count, and
    // the hit thickness (:166-167) by the voxel size; those divisors are fixed per cascade, so the
    // divisions are multiplications by their reciprocals: identical results when the divisor is a
    // power of two (extents of 2^k, one, two or four cascades), within one ulp otherwise (GLSL's
    // own division is within 2.5 ulp).
    const float cmax = r.half * 2.0f;
    const float stepTime = r.stepTime;
    const f3 p = r.o + r.d * stepTime;
    const float ux = clampf((p.x - r.cpos.x) * r.invExtent + 0.5f, 0.0f, 1.0f);
    const float uy = clampf((p.y - r.cpos.y) * r.invExtent + 0.5f, 0.0f, 1.0f);
    const float uz = clampf((p.z - r.cpos.z) * r.invExtent + 0.5f, 0.0f, 1.0f);
    const float tx = (float(r.cas) + ux) * S.invCascades;
    // The fine sample is used only where the mip distance is below the chunk size and the fine
    // distance below twice the chunk margin (SDFCommon.glsl:150-162). A brick guard >= that margin
    // proves the second test fails for every voxel of the footprint (trilinear filtering is a
    // convex combination), so the fine lookup is skipped; the factor 1 + 1e-5 covers fp32
    // rounding of the filter.
    bool needFine = true;
    if (guard) {
        const int gx = clampi(int(floorf(tx * float(texW) - 0.5f)), 0, texW - 1);
        const int gy = clampi(int(floorf(uy * float(S.res) - 0.5f)), 0, S.res - 1);
        const int gz = clampi(int(floorf(uz * float(S.res) - 0.5f)), 0, S.res - 1);
        const float gv = guard[((size_t)(gz >> 2) * (S.res >> 2) + (gy >> 2)) * (texW >> 2) + (gx >> 2)];
        needFine = !(gv >= chunkMarginDistance * 2.0f * (1.0f + 1e-5f));
    }
    const float sm = mip_lookup<SS, FL>(mip, mipW, S.mipRes, S.mipRes, tx, uy, uz);
    float sd = sm;
    float st = 1.0f;
    if (needFine) st = fine_lookup<SS, FL>(sdf, S.brickScale, texW, S.res, S.res, tx, uy, uz);
    if (COUNT) cMip++;
    if (sd < chunkSizeDistance) {
        if (COUNT) cFine++;
        if (needFine && st < chunkMarginDistance * 2.0f) sd = st;
    } else {
        sd = chunkSizeDistance;
    }
    sd *= cmax;
    const float voxel = r.voxel;
    const float thick = A.hitEps > 0.0f ? A.hitEps : voxel * 0.5f * clampf(stepTime * r.invVoxel, 0.0f, 1.0f);
    if (sd < thick) {
        if (S.refine) {  // FP32 re-evaluation of the hit step, same selection rule
            const float f32 = refine_lookup(A, S, texW, S.res, S.res, tx, uy, uz);
            const float s32 = sm < chunkSizeDistance ? (f32 < chunkMarginDistance * 2.0f ? f32 : sm)
                                                     : chunkSizeDistance;
            sd = s32 * cmax;
        }
        ra = make_float4(fmaxf(stepTime + sd - thick, 0.0f), tx, uy, uz);
        rb = make_float2(sd, float(r.cas));
        return true;
    }
    r.stepTime = stepTime + fmaxf(sd, voxel);
    return false;
}

// What a finished ray leaves: its hit record and, with the hit sort, either its final result (no
// atlas lookup: miss or inside hit) or its atlas chunk (returned; kNoChunk otherwise).
template <class SR>
__device__ __forceinline__ uint32_t march_finish(const TraceArgs& A, const TraceScene& S, const LuxgiTraceVolume& v, const MarchRay& r,
                                                 float4 ra, float2 rb, int probe, int ray, size_t rid,
                                                 float4* __restrict__ recA, float2* __restrict__ recB,
                                                 const float* __restrict__ skyFaces) {
    recA[rid] = ra;
    recB[rid] = rb;
    uint32_t key = kNoChunk;
    if (A.sortHits) {
        const bool inside = ra.x >= 0.0f && rb.x <= 0.0f && ra.x <= S.g.cascadeVoxelSize[0];
        if (ra.x >= 0.0f && !inside) {  // atlas chunk of a front hit (the lookup of sample_atlas)
            const f3 hp = r.o + r.d * ra.x;
            const float pc[3] = {hp.x, hp.y, hp.z};
            int cc[3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
                cc[i] = clampi(int(floorf(pc[i] / S.ad.chunkSize + float(kChunksRes) * 0.5f)), 0, kChunksRes - 1);
            key = (uint32_t)((cc[2] * kChunksRes + cc[1]) * kChunksRes + cc[0]);
        } else {
            store_nonhit<SR>(A, S, skyFaces, static_cast<SR*>(v.rays.radiance), static_cast<SR*>(v.rays.directionDistance),
                             v.rayFlags, (size_t)probe * v.rays.width + ray, r.d, inside);
        }
    }
    return key;
}

__device__ __forceinline__ float trace_max_distance(const TraceArgs& A, const TraceScene& S) {
    return fminf(kWorldSize, S.g.cascadePosDistance[(int)S.g.cascadesCount - 1][3] * 2.0f);
}

// The sphere trace of one ray (SDFCommon.glsl:98-171) through its cascades; returns its hit chunk
// (kNoChunk when it needs no atlas lookup).
template <class SS, class SR, bool COUNT, int FL>
__device__ __forceinline__ uint32_t march_ray(const TraceArgs& A, const TraceScene& S, const LuxgiTraceVolume& v,
                                              int probe, int ray, size_t rid, float4* __restrict__ recA,
                                              float2* __restrict__ recB, unsigned long long& cMip,
                                              unsigned long long& cFine, unsigned long long& cSteps) {
    const SS* sdf = static_cast<const SS*>(S.sdf);
    const SS* mip = static_cast<const SS*>(S.mip);
    const int C = (int)S.g.cascadesCount;
    const float traceMax = trace_max_distance(A, S);
    MarchRay r;
    march_begin(A, S, v, probe, ray, traceMax, r);
    float4 ra = make_float4(-1.0f, 0.f, 0.f, 0.f);
    float2 rb = make_float2(0.f, 0.f);
    for (;;) {
        int step = 0;
        bool hit = false;
        for (; step < A.maxSteps && r.stepTime < r.iy; ++step)
            if (march_step<SS, FL, COUNT>(A, S, sdf, mip, S.guard, r, ra, rb, cMip, cFine)) {
                hit = true;
                break;
            }
        if (COUNT) cSteps += step;
        if (hit || ++r.cas >= C) break;
        march_enter(A, S, r, traceMax);
    }
    return march_finish<SR>(A, S, v, r, ra, rb, probe, ray, rid, recA, recB, S.skyFaces);
}

// The volume of flat ray `gid` of a batch (MODE 1: one scene, MODE 2: a scene per volume). In MODE
// 2 the volumes' ray ranges start at multiples of 64 (luxgi_ddgi_trace_scenes_prepare), so a wave
// lies in one volume: the search runs on the wave's first ray and the volume and scene
// descriptors are read as wave-uniform data.
template <int MODE>
__device__ __forceinline__ int batch_volume(const TraceArgs& A, long long gid) {
    if constexpr (MODE == 2) {
        const int w0 = __builtin_amdgcn_readfirstlane((int)(gid & ~63ll));
        return __builtin_amdgcn_readfirstlane(trace_volume_of(A.vols, A.nVols, w0));
    } else {
        return trace_volume_of(A.vols, A.nVols, gid);
    }
}

__device__ __forceinline__ long long volume_rays(const LuxgiTraceVolume& v) {
    return (long long)v.uniform.probeCounts[0] * v.uniform.probeCounts[1] * v.uniform.probeCounts[2] *
           v.uniform.raysPerProbe;
}

// Sphere trace of one ray per lane. Writes the hit record; with the hit sort it also writes the
// final result of the rays without an atlas lookup and the (chunk, slot) of the others. MODE 0 (one
// volume): 64 rays of one probe per wave (rays padded to a multiple of 64 per probe). MODE 1 and 2
// (batches over one scene / a scene per volume): the flat ray list of all volumes, each ray finding
// its volume by rayStart; a ray past its volume's rays (padding, or a stale totalRays) does nothing.
template <class SS, class SR, bool COUNT, int FL, int MODE>
__global__ __launch_bounds__(256) void march_kernel(TraceArgs A, float4* __restrict__ recA, float2* __restrict__ recB,
                                                    uint2* __restrict__ hitKeys, uint32_t* __restrict__ bins,
                                                    unsigned long long* __restrict__ counters) {
    const long long gb = xcd_block(blockIdx.x, gridDim.x);
    const long long gid = gb * blockDim.x + threadIdx.x;
    unsigned long long cMip = 0, cFine = 0, cSteps = 0;
    uint32_t key = kNoChunk;
    bool valid;
    size_t rid;
    if constexpr (MODE != 0) {
        valid = gid < A.totalRays;
        rid = (size_t)gid;
        if (MODE == 2 || valid) {
            const LuxgiTraceVolume& v = A.vols[batch_volume<MODE>(A, gid)];
            const int R = v.uniform.raysPerProbe;
            const long long local = gid - v.rayStart;
            valid = valid && local >= 0 && local < volume_rays(v);
            if (valid)
                key = march_ray<SS, SR, COUNT, FL>(A, MODE == 2 ? A.scenes[v.scene] : A.scene, v, int(local / R),
                                                   int(local % R), rid, recA, recB, cMip, cFine, cSteps);
        }
    } else {
        const int Rp = (A.R + 63) & ~63;
        const int probe = int(gid / Rp), ray = int(gid % Rp);
        valid = probe < A.P && ray < A.R;
        rid = (size_t)probe * A.R + ray;
        if (valid) key = march_ray<SS, SR, COUNT, FL>(A, A.scene, A.vol, probe, ray, rid, recA, recB, cMip, cFine, cSteps);
    }
    if (A.sortHits) {  // chunk histogram; each hit keeps its slot within its chunk
        const uint32_t slot = hist_add(bins, key, key != kNoChunk);
        if (valid) hitKeys[rid] = make_uint2(key, slot);
    }
    if (COUNT) {
        __shared__ unsigned long long sc[3];
        if (threadIdx.x < 3) sc[threadIdx.x] = 0;
        __syncthreads();
        atomicAdd(&sc[0], cMip);
        atomicAdd(&sc[1], cFine);
        atomicAdd(&sc[2], cSteps);
        __syncthreads();
        if (threadIdx.x == 0) {
            atomicAdd(&counters[0], sc[0]);
            atomicAdd(&counters[1], sc[1]);
            atomicAdd(&counters[6], sc[2]);
        }
    }
}

// Normal, surface-atlas lookup and ray-result encode (SDFCommon.glsl:172-184,
// GISDFRays.comp:89-127) for one ray.
template <class SS, class SC, class SR, bool COUNT>
__device__ __forceinline__ void shade_ray(const TraceArgs& A, const TraceScene& S, const LuxgiTraceVolume& v, int probe,
                                          int ray, size_t rid, const float* __restrict__ objInvBase,
                                          const float4* __restrict__ recA, const float2* __restrict__ recB, Counts& c) {
    const SS* sdf = static_cast<const SS*>(S.sdf);
    const float* skyFaces = S.skyFaces;
    SR* outRad = static_cast<SR*>(v.rays.radiance);
    SR* outDD = static_cast<SR*>(v.rays.directionDistance);
    uint8_t* flags = v.rayFlags;
    f3 origin, dir;
    ray_setup(A, v, probe, ray, origin, dir);
    const float4 ra = recA[rid];
    const float hitTime = ra.x;
    const size_t e = (size_t)probe * v.rays.width + ray;
    if (hitTime >= 0.0f) {
        const float2 rb = recB[rid];
        const float hitSdf = rb.x;
        const int hitCascade = int(rb.y);
        if (COUNT) c.hits++;
        if (hitSdf <= 0.0f && hitTime <= S.g.cascadeVoxelSize[0]) {
            store_nonhit<SR>(A, S, skyFaces, outRad, outDD, flags, e, dir, true);
        } else {
            const int C = (int)S.g.cascadesCount;
            const int texW = S.res * C;
            const float o = 1.0f / S.g.resolution;
            const float tx = ra.y, uy = ra.z, uz = ra.w;
            auto nl = [&](float u_, float v_, float w_) {
                if (S.refine) return refine_lookup(A, S, texW, S.res, S.res, u_, v_, w_);
                if (A.fineLayout == LUXGI_LAYOUT_BRICK4)
                    return sample3d_brick4<SS>(sdf, texW, S.res, S.res, u_, v_, w_, S.brickScale);
                if (A.fineLayout == LUXGI_LAYOUT_CELL8) return sample3d_cell8<SS>(sdf, texW, S.res, S.res, u_, v_, w_);
                return sample_volume<SS>(sdf, texW, S.res, S.res, u_, v_, w_);
            };
            const float xp = nl(tx + o, uy, uz), xn = nl(tx - o, uy, uz);
            const float yp = nl(tx, uy + o, uz), yn = nl(tx, uy - o, uz);
            const float zp = nl(tx, uy, uz + o), zn = nl(tx, uy, uz - o);
            const f3 n = normalize3(mk(xp - xn, yp - yn, zp - zn));
            if (COUNT) c.normals++;
            const f3 hp = origin + dir * hitTime;
            const float thr = S.g.cascadeVoxelSize[hitCascade] * 1.05f;
            const float4 s = sample_atlas<SC, COUNT>(S.ad, S.chunks, S.culled, S.objects, objInvBase + S.objectStart * 16,
                                                     S.tiles, static_cast<const SC*>(S.color), S.atlasDepth, hp, n,
                                                     thr, c);
            const float dist = fmaxf(hitTime + S.g.cascadeVoxelSize[hitCascade] * 0.5f, 0.0f);
            Texel<SR>::st4(outRad, e, make_float4(s.x, s.y, s.z, 0.f));
            Texel<SR>::st4(outDD, e, make_float4(dir.x, dir.y, dir.z, dist));
            if (flags) flags[e] = 1;
        }
    } else {
        store_nonhit<SR>(A, S, skyFaces, outRad, outDD, flags, e, dir, false);
    }
}

// One ray per lane: the sorted front hits (`order`, grouped by atlas chunk, so the object and tile
// loops of a wave run in step), or every ray (MODE 0: 64 rays of a probe per wave; MODE 1 and 2:
// the flat ray list of a batch, as in march_kernel).
template <class SS, class SC, class SR, bool COUNT, int MODE>
__global__ __launch_bounds__(256) void shade_kernel(TraceArgs A, const float* __restrict__ objInvBase,
                                                    const float4* __restrict__ recA, const float2* __restrict__ recB,
                                                    const uint32_t* __restrict__ order,
                                                    const uint32_t* __restrict__ hitCount,
                                                    unsigned long long* __restrict__ counters) {
    const long long gid = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    Counts c = {0, 0, 0, 0, 0, 0, 0};
    bool active;
    if (MODE == 2) {  // ray order only: a wave lies in one volume
        active = gid < A.totalRays;
        const LuxgiTraceVolume& v = A.vols[batch_volume<2>(A, gid)];
        const int R = v.uniform.raysPerProbe;
        const long long local = gid - v.rayStart;
        active = active && local >= 0 && local < volume_rays(v);
        if (active)
            shade_ray<SS, SC, SR, COUNT>(A, A.scenes[v.scene], v, int(local / R), int(local % R), (size_t)gid,
                                         objInvBase, recA, recB, c);
    } else if (MODE == 1 || order) {  // a flat ray id
        long long rid = -1;
        if (order) {
            if (gid < (long long)*hitCount) rid = order[gid];
        } else if (gid < A.totalRays) {
            rid = gid;
        }
        active = rid >= 0;
        if (active) {
            if constexpr (MODE == 1) {
                const LuxgiTraceVolume& v = A.vols[trace_volume_of(A.vols, A.nVols, rid)];
                const int R = v.uniform.raysPerProbe;
                const long long local = rid - v.rayStart;
                active = local >= 0 && local < volume_rays(v);
                if (active)
                    shade_ray<SS, SC, SR, COUNT>(A, A.scene, v, int(local / R), int(local % R), (size_t)rid,
                                                 objInvBase, recA, recB, c);
            } else {
                shade_ray<SS, SC, SR, COUNT>(A, A.scene, A.vol, int(rid / A.R), int(rid % A.R), (size_t)rid,
                                             objInvBase, recA, recB, c);
            }
        }
    } else {
        const int Rp = (A.R + 63) & ~63;
        const int probe = int(gid / Rp), ray = int(gid % Rp);
        active = probe < A.P && ray < A.R;
        if (active)
            shade_ray<SS, SC, SR, COUNT>(A, A.scene, A.vol, probe, ray, (size_t)probe * A.R + ray, objInvBase, recA,
                                         recB, c);
    }
    if (COUNT) {
        __shared__ unsigned long long sc[8];
        if (threadIdx.x < 8) sc[threadIdx.x] = 0;
        __syncthreads();
        if (active) {
            atomicAdd(&sc[2], c.normals);
            atomicAdd(&sc[3], c.objects);
            atomicAdd(&sc[4], c.tiles);
            atomicAdd(&sc[5], 1ull);
            atomicAdd(&sc[7], c.hits);
        }
        __syncthreads();
        if (threadIdx.x < 8 && threadIdx.x != 0 && threadIdx.x != 1 && threadIdx.x != 6)
            atomicAdd(&counters[threadIdx.x], sc[threadIdx.x]);
    }
}

// counting sort of the rays by hit chunk: zero the bins, exclusive scan, scatter
__global__ void zero_bins_kernel(uint32_t* __restrict__ bins, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bins[i] = 0u;
}

constexpr int kScanBlock = 1024;
constexpr int kScanBlocks = (kChunkBins + kScanBlock - 1) / kScanBlock;
static_assert(kScanBlocks <= 64, "the block totals are scanned by one wave");

// order[start of the chunk + slot] = ray for every front hit; block 0 also stores the hit count
__global__ __launch_bounds__(256) void scatter_hits_kernel(const uint2* __restrict__ keys,
                                                           const uint32_t* __restrict__ starts,
                                                           const uint32_t* __restrict__ totals,
                                                           uint32_t* __restrict__ hitCount,
                                                           uint32_t* __restrict__ order, long long n) {
    __shared__ uint32_t blockStart[64];
    if (threadIdx.x < 64) {
        const uint32_t t = threadIdx.x < kScanBlocks ? totals[threadIdx.x] : 0u;
        const uint32_t x = wave_inclusive_scan(t);
        blockStart[threadIdx.x] = x - t;
        if (blockIdx.x == 0 && threadIdx.x == 63) *hitCount = x;
    }
    __syncthreads();
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint2 k = keys[i];
        if (k.x != kNoChunk) order[blockStart[k.x / kScanBlock] + starts[k.x] + k.y] = (uint32_t)i;
    }
}

// per-object inverse transforms, once per call
__global__ void trace_prepare_kernel(const LuxgiAtlasObject* __restrict__ objects, int n, float* __restrict__ inv) {
    object_inverse_kernel_body(objects, n, inv);
}

// the same for every scene of a several-scene batch: scene blockIdx.y, its objects from its objectStart
__global__ void trace_prepare_scenes_kernel(const TraceScene* __restrict__ scenes, float* __restrict__ inv) {
    const TraceScene& S = scenes[blockIdx.y];
    object_inverse_kernel_body(S.objects, S.objectCapacity, inv + S.objectStart * 16);
}

}  // namespace luxgi

using namespace luxgi;

static size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

struct TraceWorkspace {
    float* inv;
    float4* recA;
    float2* recB;
    uint2* keys;      // hit sort: (chunk, slot in the chunk) per ray
    uint32_t* bins;   // per-chunk hit counts, then their in-block exclusive scan
    uint32_t* totals; // scan block totals, then the hit count at [kScanBlocks]
    uint32_t* order;  // front hits in chunk order
};

static size_t trace_workspace_layout(size_t rays, int64_t objects, char* base, TraceWorkspace* w) {
    size_t off = 0;
    if (w) w->inv = reinterpret_cast<float*>(base + off);
    off += align256((size_t)(objects > 0 ? objects : 0) * 16 * sizeof(float));
    if (w) w->recA = reinterpret_cast<float4*>(base + off);
    off += align256(rays * sizeof(float4));
    if (w) w->recB = reinterpret_cast<float2*>(base + off);
    off += align256(rays * sizeof(float2));
    if (w) w->keys = reinterpret_cast<uint2*>(base + off);
    off += align256(rays * sizeof(uint2));
    if (w) w->bins = reinterpret_cast<uint32_t*>(base + off);
    off += align256((size_t)kChunkBins * sizeof(uint32_t));
    if (w) w->totals = reinterpret_cast<uint32_t*>(base + off);
    off += align256((size_t)(kScanBlocks + 1) * sizeof(uint32_t));
    if (w) w->order = reinterpret_cast<uint32_t*>(base + off);
    off += align256(rays * sizeof(uint32_t));
    return off;
}

extern "C" size_t luxgi_ddgi_trace_workspace_size(const LuxgiDDGIUniform* u, int32_t objectCapacity) {
    if (!u) return 0;
    const size_t P = (size_t)u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    const size_t R = (size_t)(u->raysPerProbe > 0 ? u->raysPerProbe : 0);
    return trace_workspace_layout(P * R, objectCapacity, nullptr, nullptr);
}

// The kernels of one trace: object inverses, (hit-sort bins), march, (scan + scatter), shade. MODE 0:
// one volume, 1: a batch over A.scene, 2: a batch over per-volume scenes (A.scenes, nScenes of them,
// at most maxObjects objects each).
template <class SS, class SC, class SR, bool COUNT, int MODE>
static int launch_trace(const TraceArgs& A, int nScenes, int maxObjects, const TraceWorkspace& w,
                        unsigned long long* counters, hipStream_t st) {
    const long long nRays = MODE != 0 ? A.totalRays : (long long)A.P * A.R;
    const long long lanes = MODE != 0 ? nRays : (long long)A.P * ((A.R + 63) & ~63);
    if (MODE == 2) {
        if (maxObjects > 0) {
            hipLaunchKernelGGL(trace_prepare_scenes_kernel, dim3((maxObjects + 255) / 256, nScenes), dim3(256), 0, st,
                               A.scenes, w.inv);
            LUXGI_LAUNCH_CHECK();
        }
    } else {
        hipLaunchKernelGGL(trace_prepare_kernel, dim3(std::max(1, (A.scene.objectCapacity + 255) / 256)), dim3(256), 0,
                           st, A.scene.objects, A.scene.objectCapacity, w.inv);
        LUXGI_LAUNCH_CHECK();
    }
    if (A.sortHits) {
        hipLaunchKernelGGL(zero_bins_kernel, dim3((kChunkBins + 255) / 256), dim3(256), 0, st, w.bins, kChunkBins);
        LUXGI_LAUNCH_CHECK();
    }
    const dim3 mgrid((int)((lanes + 255) / 256));
    if (A.fineLayout == LUXGI_LAYOUT_BRICK4)
        hipLaunchKernelGGL((march_kernel<SS, SR, COUNT, LUXGI_LAYOUT_BRICK4, MODE>), mgrid, dim3(256), 0, st, A, w.recA,
                           w.recB, w.keys, w.bins, counters);
    else if (A.fineLayout == LUXGI_LAYOUT_CELL8)
        hipLaunchKernelGGL((march_kernel<SS, SR, COUNT, LUXGI_LAYOUT_CELL8, MODE>), mgrid, dim3(256), 0, st, A, w.recA,
                           w.recB, w.keys, w.bins, counters);
    else
        hipLaunchKernelGGL((march_kernel<SS, SR, COUNT, LUXGI_LAYOUT_LINEAR, MODE>), mgrid, dim3(256), 0, st, A, w.recA,
                           w.recB, w.keys, w.bins, counters);
    LUXGI_LAUNCH_CHECK();
    if (A.sortHits) {
        hipLaunchKernelGGL(block_scan_kernel<kScanBlock>, dim3(kScanBlocks), dim3(kScanBlock), 0, st, w.bins, w.bins,
                           w.totals, kChunkBins);
        LUXGI_LAUNCH_CHECK();
        hipLaunchKernelGGL(scatter_hits_kernel, dim3((int)((nRays + 255) / 256)), dim3(256), 0, st, w.keys, w.bins,
                           w.totals, w.totals + kScanBlocks, w.order, nRays);
        LUXGI_LAUNCH_CHECK();
    }
    // sorted: one lane per ray covers every possible hit count; the lanes past the count exit
    const long long shadeLanes = A.sortHits ? nRays : lanes;
    hipLaunchKernelGGL((shade_kernel<SS, SC, SR, COUNT, MODE>), dim3((int)((shadeLanes + 255) / 256)), dim3(256), 0, st,
                       A, w.inv, w.recA, w.recB, A.sortHits ? w.order : nullptr, w.totals + kScanBlocks, counters);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

template <class SS, class SR, bool COUNT, int MODE>
static int dispatch_color(const TraceArgs& A, int atlasLayout, int colorStorage, int nScenes, int maxObjects,
                          const TraceWorkspace& w, unsigned long long* cnt, hipStream_t st) {
    if (atlasLayout == LUXGI_LAYOUT_BRICK4)
        return launch_trace<SS, PackedAtlas, SR, COUNT, MODE>(A, nScenes, maxObjects, w, cnt, st);
    switch (colorStorage) {
        case LUXGI_STORAGE_FP32: return launch_trace<SS, float, SR, COUNT, MODE>(A, nScenes, maxObjects, w, cnt, st);
        case LUXGI_STORAGE_FP16: return launch_trace<SS, __half, SR, COUNT, MODE>(A, nScenes, maxObjects, w, cnt, st);
        case LUXGI_STORAGE_UNORM8: return launch_trace<SS, Unorm8, SR, COUNT, MODE>(A, nScenes, maxObjects, w, cnt, st);
        default: return LUXGI_ERR_UNSUPPORTED;
    }
}

// Checks one scene (SDF + atlas + sky) and writes its kernel-side descriptor.
static int trace_scene_descriptor(const LuxgiGlobalSDFData* g, const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at,
                                  const LuxgiSkybox* sky, const float* skyColor, TraceScene& S) {
    if (!g || !vol || !at) return LUXGI_ERR_INVALID_ARGUMENT;
    const int rc = sdf_volume_status(g, vol);
    if (rc != LUXGI_OK) return rc;
    if (!at->chunks || !at->culledObjects || !at->objects || !at->tiles || !at->color) return LUXGI_ERR_INVALID_ARGUMENT;
    if (at->layout == LUXGI_LAYOUT_LINEAR) {
        if (!at->depth) return LUXGI_ERR_INVALID_ARGUMENT;
    } else if (at->layout == LUXGI_LAYOUT_BRICK4) {
        if (at->colorStorage != LUXGI_STORAGE_UNORM8 || (at->data.resolution & 3u)) return LUXGI_ERR_INVALID_ARGUMENT;
    } else {
        return LUXGI_ERR_INVALID_ARGUMENT;
    }
    if (at->data.resolution == 0 || at->objectCapacity < 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if (sky && sky->faces && sky->size <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    S = TraceScene{};
    S.g = *g;
    S.ad = at->data;
    S.sdf = vol->sdf;
    S.mip = vol->mip;
    S.guard = vol->guard;
    S.brickScale = vol->brickScale;
    S.refine = vol->refine;
    S.chunks = at->chunks;
    S.culled = at->culledObjects;
    S.objects = at->objects;
    S.tiles = at->tiles;
    S.color = at->color;
    S.atlasDepth = at->depth;
    S.skyFaces = sky && sky->faces ? sky->faces : nullptr;
    S.objectStart = 0;
    S.objectCapacity = at->objectCapacity;
    for (int i = 0; i < 3; ++i) S.sky[i] = skyColor[i];
    S.invCascades = 1.0f / float(g->cascadesCount);
    S.res = vol->resolution;
    S.mipRes = vol->mipResolution;
    S.skySize = sky ? sky->size : 0;
    S.refineLayout = vol->refineLayout;
    return LUXGI_OK;
}

// The scene arguments shared by the single and batched traces: checks them and fills A's scene
// and the call's settings.
static int trace_scene_args(const LuxgiGlobalSDFData* g, const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at,
                            const LuxgiSkybox* sky, const LuxgiTraceParams* p, bool count, TraceArgs& A) {
    if (!p) return LUXGI_ERR_INVALID_ARGUMENT;
    const int rc = trace_scene_descriptor(g, vol, at, sky, p->skyColor, A.scene);
    if (rc != LUXGI_OK) return rc;
    A.scenes = nullptr;
    A.maxSteps = p->maxSteps > 0 ? p->maxSteps : 250;
    A.hitEps = p->hitEpsilon;
    A.fineLayout = vol->layout;
    if (p->shadeOrder != 0 && p->shadeOrder != 1) return LUXGI_ERR_INVALID_ARGUMENT;
    A.sortHits = !count && p->shadeOrder == 0;  // the counting run shades in ray order
    A.vols = nullptr;
    A.nVols = 0;
    A.totalRays = 0;
    return LUXGI_OK;
}

template <bool COUNT, int MODE>
static int trace_launch(const TraceArgs& A, int sdfStorage, int atlasLayout, int colorStorage, int rayStorage,
                        size_t rays, int64_t objects, int nScenes, int maxObjects, void* workspace,
                        size_t workspaceBytes, unsigned long long* counters, hipStream_t st) {
    const size_t need = trace_workspace_layout(rays, objects, nullptr, nullptr);
    if (!workspace || workspaceBytes < need) return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    TraceWorkspace w;
    trace_workspace_layout(rays, objects, static_cast<char*>(workspace), &w);
    const bool r32 = rayStorage == LUXGI_STORAGE_FP32;
    if (rayStorage != LUXGI_STORAGE_FP32 && rayStorage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    switch (sdfStorage) {
        case LUXGI_STORAGE_FP32:
            return r32 ? dispatch_color<float, float, COUNT, MODE>(A, atlasLayout, colorStorage, nScenes, maxObjects, w,
                                                                   counters, st)
                       : dispatch_color<float, __half, COUNT, MODE>(A, atlasLayout, colorStorage, nScenes, maxObjects,
                                                                    w, counters, st);
        case LUXGI_STORAGE_FP16:
            return r32 ? dispatch_color<__half, float, COUNT, MODE>(A, atlasLayout, colorStorage, nScenes, maxObjects,
                                                                    w, counters, st)
                       : dispatch_color<__half, __half, COUNT, MODE>(A, atlasLayout, colorStorage, nScenes, maxObjects,
                                                                     w, counters, st);
        case LUXGI_STORAGE_BF16:
            return r32 ? dispatch_color<__bf16, float, COUNT, MODE>(A, atlasLayout, colorStorage, nScenes, maxObjects,
                                                                    w, counters, st)
                       : dispatch_color<__bf16, __half, COUNT, MODE>(A, atlasLayout, colorStorage, nScenes, maxObjects,
                                                                     w, counters, st);
        default: return LUXGI_ERR_UNSUPPORTED;
    }
}

// checks one volume's own arguments (grid and ray buffers)
static int trace_volume_ok(const LuxgiDDGIUniform* u, const LuxgiRayBuffers* rays) {
    if (!u || !rays || !rays->radiance || !rays->directionDistance) return LUXGI_ERR_INVALID_ARGUMENT;
    const long long P = (long long)u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    if (u->probeCounts[0] <= 0 || u->probeCounts[1] <= 0 || u->probeCounts[2] <= 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if (u->raysPerProbe <= 0 || u->raysPerProbe > rays->width || rays->height < P) return LUXGI_ERR_INVALID_ARGUMENT;
    if (P * u->raysPerProbe >= (1ll << 31)) return LUXGI_ERR_UNSUPPORTED;
    if (rays->storage != LUXGI_STORAGE_FP32 && rays->storage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    return LUXGI_OK;
}

template <bool COUNT>
static int trace_entry(const LuxgiDDGIUniform* u, const LuxgiGlobalSDFData* g, const LuxgiSDFVolume* vol,
                       const LuxgiSurfaceAtlas* at, const LuxgiSkybox* sky, const LuxgiTraceParams* p,
                       const float* probeOffsets, const LuxgiRayBuffers* rays, uint8_t* rayFlags, void* workspace,
                       size_t workspaceBytes, unsigned long long* counters, void* stream) {
    int rc = trace_volume_ok(u, rays);
    if (rc != LUXGI_OK) return rc;
    TraceArgs A;
    rc = trace_scene_args(g, vol, at, sky, p, COUNT, A);
    if (rc != LUXGI_OK) return rc;
    A.P = u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    A.R = u->raysPerProbe;
    A.rayWidth = rays->width;
    A.vol.uniform = *u;
    for (int i = 0; i < 16; ++i) A.vol.orientation[i] = p->randomOrientation[i];
    A.vol.deviceOrientation = p->deviceOrientation;
    A.vol.probeOffsets = probeOffsets;
    A.vol.rays = *rays;
    A.vol.rayFlags = rayFlags;
    A.vol.rayStart = 0;
    A.vol.scene = 0;
    return trace_launch<COUNT, 0>(A, vol->storage, at->layout, at->colorStorage, rays->storage, (size_t)A.P * A.R,
                                  at->objectCapacity, 1, at->objectCapacity, workspace, workspaceBytes, counters,
                                  as_stream(stream));
}

extern "C" int luxgi_ddgi_trace_batch_prepare(LuxgiTraceVolume* volumes, int32_t count, int32_t rayStorage,
                                              int64_t* totalRays) {
    if (!volumes || !totalRays || count <= 0 || count > LUXGI_MAX_TRACE_BATCH) return LUXGI_ERR_INVALID_ARGUMENT;
    if (rayStorage != LUXGI_STORAGE_FP32 && rayStorage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    long long total = 0;
    for (int k = 0; k < count; ++k) {
        LuxgiTraceVolume& v = volumes[k];
        const int rc = trace_volume_ok(&v.uniform, &v.rays);
        if (rc != LUXGI_OK) return rc;
        if (v.rays.storage != rayStorage) return LUXGI_ERR_INVALID_ARGUMENT;
        v.rayStart = total;
        total += (long long)v.uniform.probeCounts[0] * v.uniform.probeCounts[1] * v.uniform.probeCounts[2] *
                 v.uniform.raysPerProbe;
        if (total >= (1ll << 31)) return LUXGI_ERR_UNSUPPORTED;
    }
    *totalRays = total;
    return LUXGI_OK;
}

extern "C" size_t luxgi_ddgi_trace_batch_workspace_size(int64_t totalRays, int32_t objectCapacity) {
    if (totalRays <= 0 || totalRays >= (1ll << 31)) return 0;
    return trace_workspace_layout((size_t)totalRays, objectCapacity, nullptr, nullptr);
}

extern "C" int luxgi_ddgi_trace_sdf_batched(const LuxgiTraceVolume* volumes, int32_t count, int64_t totalRays,
                                            int32_t rayStorage, const LuxgiGlobalSDFData* g,
                                            const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at,
                                            const LuxgiSkybox* sky, const LuxgiTraceParams* p, void* workspace,
                                            size_t workspaceBytes, void* stream) {
    if (!volumes || count <= 0 || count > LUXGI_MAX_TRACE_BATCH || totalRays <= 0 || totalRays >= (1ll << 31))
        return LUXGI_ERR_INVALID_ARGUMENT;
    TraceArgs A;
    const int rc = trace_scene_args(g, vol, at, sky, p, false, A);
    if (rc != LUXGI_OK) return rc;
    A.P = A.R = A.rayWidth = 0;
    A.vols = volumes;
    A.nVols = count;
    A.totalRays = totalRays;
    return trace_launch<false, 1>(A, vol->storage, at->layout, at->colorStorage, rayStorage, (size_t)totalRays,
                                  at->objectCapacity, 1, at->objectCapacity, workspace, workspaceBytes, nullptr,
                                  as_stream(stream));
}

// ------------------------------------------------------------------------- several scenes

extern "C" size_t luxgi_ddgi_trace_scenes_descriptor_bytes(int32_t sceneCount) {
    if (sceneCount <= 0 || sceneCount > LUXGI_MAX_TRACE_BATCH) return 0;
    return (size_t)sceneCount * sizeof(TraceScene);
}

extern "C" int luxgi_ddgi_trace_scenes_prepare(const LuxgiTraceScene* scenes, int32_t sceneCount,
                                               LuxgiTraceVolume* volumes, int32_t volumeCount, int32_t rayStorage,
                                               void* descriptors, LuxgiTraceSceneBatch* batch) {
    if (!scenes || !volumes || !descriptors || !batch || sceneCount <= 0 || sceneCount > LUXGI_MAX_TRACE_BATCH ||
        volumeCount <= 0 || volumeCount > LUXGI_MAX_TRACE_BATCH)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (rayStorage != LUXGI_STORAGE_FP32 && rayStorage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    TraceScene* out = static_cast<TraceScene*>(descriptors);
    long long objects = 0;
    int maxObjects = 0;
    for (int k = 0; k < sceneCount; ++k) {
        const LuxgiTraceScene& sc = scenes[k];
        TraceScene S;
        const int rc = trace_scene_descriptor(&sc.sdfData, &sc.volume, &sc.atlas, &sc.skybox, sc.skyColor, S);
        if (rc != LUXGI_OK) return rc;
        // one kernel instantiation serves the batch: every scene stores its SDF and atlas alike
        if (sc.volume.storage != scenes[0].volume.storage || sc.volume.layout != scenes[0].volume.layout ||
            sc.atlas.layout != scenes[0].atlas.layout || sc.atlas.colorStorage != scenes[0].atlas.colorStorage)
            return LUXGI_ERR_INVALID_ARGUMENT;
        S.objectStart = objects;
        objects += sc.atlas.objectCapacity;
        maxObjects = std::max(maxObjects, (int)sc.atlas.objectCapacity);
        out[k] = S;
    }
    long long total = 0;
    for (int k = 0; k < volumeCount; ++k) {
        LuxgiTraceVolume& v = volumes[k];
        const int rc = trace_volume_ok(&v.uniform, &v.rays);
        if (rc != LUXGI_OK) return rc;
        if (v.rays.storage != rayStorage || v.scene < 0 || v.scene >= sceneCount) return LUXGI_ERR_INVALID_ARGUMENT;
        total = (total + 63) & ~63ll;  // a wave never spans two volumes
        v.rayStart = total;
        total += (long long)v.uniform.probeCounts[0] * v.uniform.probeCounts[1] * v.uniform.probeCounts[2] *
                 v.uniform.raysPerProbe;
        if (total >= (1ll << 31)) return LUXGI_ERR_UNSUPPORTED;
    }
    batch->totalRays = total;
    batch->totalObjects = objects;
    batch->sceneCount = sceneCount;
    batch->volumeCount = volumeCount;
    batch->maxObjects = maxObjects;
    batch->rayStorage = rayStorage;
    batch->sdfStorage = scenes[0].volume.storage;
    batch->sdfLayout = scenes[0].volume.layout;
    batch->atlasLayout = scenes[0].atlas.layout;
    batch->colorStorage = scenes[0].atlas.colorStorage;
    return LUXGI_OK;
}

extern "C" size_t luxgi_ddgi_trace_scenes_workspace_size(const LuxgiTraceSceneBatch* batch) {
    if (!batch || batch->totalRays <= 0 || batch->totalRays >= (1ll << 31) || batch->totalObjects < 0) return 0;
    return trace_workspace_layout((size_t)batch->totalRays, batch->totalObjects, nullptr, nullptr);
}

extern "C" int luxgi_ddgi_trace_sdf_scenes(const LuxgiTraceSceneBatch* batch, const LuxgiTraceVolume* volumes,
                                           const void* descriptors, const LuxgiTraceParams* p, void* workspace,
                                           size_t workspaceBytes, void* stream) {
    if (!batch || !volumes || !descriptors || !p) return LUXGI_ERR_INVALID_ARGUMENT;
    if (batch->sceneCount <= 0 || batch->sceneCount > LUXGI_MAX_TRACE_BATCH || batch->volumeCount <= 0 ||
        batch->volumeCount > LUXGI_MAX_TRACE_BATCH || batch->totalRays <= 0 || batch->totalRays >= (1ll << 31) ||
        batch->totalObjects < 0 || batch->maxObjects < 0)
        return LUXGI_ERR_INVALID_ARGUMENT;
    TraceArgs A{};
    A.scenes = static_cast<const TraceScene*>(descriptors);
    A.maxSteps = p->maxSteps > 0 ? p->maxSteps : 250;
    A.hitEps = p->hitEpsilon;
    A.fineLayout = batch->sdfLayout;
    A.sortHits = 0;  // shaded in ray order: a wave stays within one volume and scene
    A.vols = volumes;
    A.nVols = batch->volumeCount;
    A.totalRays = batch->totalRays;
    return trace_launch<false, 2>(A, batch->sdfStorage, batch->atlasLayout, batch->colorStorage, batch->rayStorage,
                                  (size_t)batch->totalRays, batch->totalObjects, batch->sceneCount, batch->maxObjects,
                                  workspace, workspaceBytes, nullptr, as_stream(stream));
}

extern "C" int luxgi_ddgi_trace_sdf(const LuxgiDDGIUniform* u, const LuxgiGlobalSDFData* g, const LuxgiSDFVolume* vol,
                                    const LuxgiSurfaceAtlas* at, const LuxgiSkybox* sky, const LuxgiTraceParams* p,
                                    const float* probeOffsets, const LuxgiRayBuffers* rays, uint8_t* rayFlags,
                                    void* workspace, size_t workspaceBytes, void* stream) {
    return trace_entry<false>(u, g, vol, at, sky, p, probeOffsets, rays, rayFlags, workspace, workspaceBytes, nullptr,
                              stream);
}

extern "C" int luxgi_ddgi_trace_sdf_count(const LuxgiDDGIUniform* u, const LuxgiGlobalSDFData* g,
                                          const LuxgiSDFVolume* vol, const LuxgiSurfaceAtlas* at,
                                          const LuxgiSkybox* sky, const LuxgiTraceParams* p, const float* probeOffsets,
                                          const LuxgiRayBuffers* rays, uint8_t* rayFlags, void* workspace,
                                          size_t workspaceBytes, unsigned long long* counters, void* stream) {
    if (!counters) return LUXGI_ERR_INVALID_ARGUMENT;
    return trace_entry<true>(u, g, vol, at, sky, p, probeOffsets, rays, rayFlags, workspace, workspaceBytes, counters,
                             stream);
}

namespace luxgi {
__global__ void atlas_pack_kernel(const uint32_t* __restrict__ color, const float* __restrict__ depth, int res,
                                  uint2* __restrict__ packed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= res) return;
    const size_t e = (size_t)j * res + i;
    packed[atlas_block_index(i, j, res)] = make_uint2(color[e], __float_as_uint(depth[e]));
}
}  // namespace luxgi

extern "C" int luxgi_atlas_pack(const uint32_t* color, const float* depth, int32_t res, void* packed, void* stream) {
    if (!color || !depth || !packed || res <= 0 || (res & 3) || (size_t)res * res * 8 >= (1ull << 32))
        return LUXGI_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(atlas_pack_kernel, dim3((res + 255) / 256, res), dim3(256), 0, as_stream(stream), color, depth,
                       res, static_cast<uint2*>(packed));
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

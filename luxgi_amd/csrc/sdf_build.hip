// sdf_build.hip -- global SDF build from mesh SDFs, for gfx950.
//
// SDFRasterizeModel.glsl:42-63 (+ READ_DISTANCE) and GlobalSDFClear.comp:27-33, dispatched per 32^3
// chunk of a cascade by GlobalDistanceField.cpp:700-816. One lane per voxel, 256 voxels (an 8x8x4
// brick) per block, 128 blocks per chunk. The chunk's objects are staged through LDS 28 at a time
// (the reference's per-dispatch batch, GLOBAL_SDF_RASTERIZE_MODEL_MAX_COUNT); between batches the
// running distance restarts from the stored value, as the reference's layered dispatches do.
#include "common.hpp"

namespace luxgi {

constexpr int kRasterChunk = 32;    // CONSTS_SDF_RASTERIZE_CHUNK_SIZE, GlobalDistanceField.cpp:86
constexpr int kRasterBatch = 28;    // GLOBAL_SDF_RASTERIZE_MODEL_MAX_COUNT, SDFRasterizeModel.glsl:8
constexpr int kRasterBlocksPerChunk = kRasterChunk * kRasterChunk * kRasterChunk / 256;

struct RasterArgs {
    LuxgiRasterizeParams p;
    int numObjects, numChunks, width, cascades, half;
};

// textureLod(sampler3D, uvw, level) of an R16F mesh SDF: trilinear inside the level, repeat addressing
// (the engine's default sampler for the mesh SDF textures, SDFBaker.cpp:225, RHI/Definitions.h:152-159)
__device__ __forceinline__ float mesh_sample(const uint16_t* __restrict__ pool, const LuxgiMeshSDF& m, int level,
                                             f3 uvw) {
    level = max(0, min(level, m.mipCount - 1));  // textureLod clamps the level to the chain
    const int W = m.size[level][0], H = m.size[level][1], D = m.size[level][2];
    const uint16_t* v = pool + m.offset[level];
    const float x = uvw.x * float(W) - 0.5f, y = uvw.y * float(H) - 0.5f, z = uvw.z * float(D) - 0.5f;
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const float ax = x - fx, ay = y - fy, az = z - fz;
    auto wrap = [](int i, int n) {
        const int r = i % n;
        return r < 0 ? r + n : r;
    };
    const int x0 = wrap(int(fx), W), x1 = wrap(int(fx) + 1, W);
    const int y0 = wrap(int(fy), H), y1 = wrap(int(fy) + 1, H);
    const int z0 = wrap(int(fz), D), z1 = wrap(int(fz) + 1, D);
    auto t = [&](int xi, int yi, int zi) {
        return __half2float(__ushort_as_half(v[((size_t)zi * H + yi) * W + xi]));
    };
    const float c00 = t(x0, y0, z0) * (1.0f - ax) + t(x1, y0, z0) * ax;
    const float c10 = t(x0, y1, z0) * (1.0f - ax) + t(x1, y1, z0) * ax;
    const float c01 = t(x0, y0, z1) * (1.0f - ax) + t(x1, y0, z1) * ax;
    const float c11 = t(x0, y1, z1) * (1.0f - ax) + t(x1, y1, z1) * ax;
    return (c00 * (1.0f - ay) + c10 * ay) * (1.0f - az) + (c01 * (1.0f - ay) + c11 * ay) * az;
}

// distanceToModelSDF (SDFCommon.glsl:38-62)
__device__ __forceinline__ float distance_to_model(float minDistance, const LuxgiObjectRasterizeData& o,
                                                   const LuxgiMeshSDF& m, const uint16_t* __restrict__ pool, f3 wp) {
    const f3 vp = mat4_mul(o.worldToVolume, wp, 1.0f);
    const f3 uvw = mk(vp.x * o.volumeToUVWMul[0] + o.volumeToUVWAdd[0], vp.y * o.volumeToUVWMul[1] + o.volumeToUVWAdd[1],
                      vp.z * o.volumeToUVWMul[2] + o.volumeToUVWAdd[2]);
    const f3 e = mk(o.volumeLocalBoundsExtent[0], o.volumeLocalBoundsExtent[1], o.volumeLocalBoundsExtent[2]);
    const f3 vc = mk(clampf(vp.x, -e.x, e.x), clampf(vp.y, -e.y, e.y), clampf(vp.z, -e.z, e.z));
    const f3 wc = mat4_mul(o.volumeToWorld, vc, 1.0f);
    float dv = len3(wp - wc);
    if (dv < 0.01f) dv = len3(vp - vc);
    dv = fmaxf(dv, 0.0f);
    if (minDistance <= dv) return dv;
    const float sd = (mesh_sample(pool, m, int(o.mipOffset), uvw) * 2.0f - 1.0f) * o.decodeMul;
    // combineDistanceToSDF (SDFCommon.glsl:18-36)
    float r;
    if (sd <= 0.0f && dv <= 0.0f) {
        r = sd;
    } else {
        const float ms = fmaxf(sd, 0.0f);
        r = sqrtf(ms * ms + dv * dv);
    }
    if (dv > 0.0f) r = fmaxf(dv, r);
    return r;
}

__device__ __forceinline__ float store_value(float minDistance, float maxDistance, int half) {
    const float v = clampf(minDistance / maxDistance, -1.0f, 1.0f);
    return half ? __half2float(__float2half(v)) : v;
}

__global__ __launch_bounds__(256) void rasterize_kernel(RasterArgs A, const LuxgiObjectRasterizeData* __restrict__ objects,
                                                        const LuxgiMeshSDF* __restrict__ meshes,
                                                        const uint16_t* __restrict__ pool,
                                                        const int4* __restrict__ origins,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ list, void* __restrict__ sdf) {
    __shared__ LuxgiObjectRasterizeData so[kRasterBatch];
    __shared__ LuxgiMeshSDF sm[kRasterBatch];
    __shared__ int sid[kRasterBatch];
    const int chunk = blockIdx.x / kRasterBlocksPerChunk, brick = blockIdx.x % kRasterBlocksPerChunk;
    // 8x8x4 voxel bricks: 4 x 4 x 8 of them per 32^3 chunk
    const int bx = brick & 3, by = (brick >> 2) & 3, bz = brick >> 4;
    const int4 org = origins[chunk];
    const int vx = org.x + bx * 8 + (threadIdx.x & 7), vy = org.y + by * 8 + ((threadIdx.x >> 3) & 7),
              vz = org.z + bz * 4 + (threadIdx.x >> 6);
    const int R = A.p.cascadeResolution;
    const bool inside = vx >= 0 && vy >= 0 && vz >= 0 && vx < R && vy < R && vz < R;
    const f3 wp = mk(float(vx) * A.p.cascadeCoordToPosMul[0] + A.p.cascadeCoordToPosAdd[0],
                     float(vy) * A.p.cascadeCoordToPosMul[1] + A.p.cascadeCoordToPosAdd[1],
                     float(vz) * A.p.cascadeCoordToPosMul[2] + A.p.cascadeCoordToPosAdd[2]);
    const float maxD = A.p.maxDistance;
    const uint32_t s0 = start[chunk], s1 = start[chunk + 1];
    float minDistance = maxD;
    float stored = 1.0f;  // GlobalSDFClear.comp: a chunk without objects is cleared to 1
    for (uint32_t b = s0; b < s1; b += kRasterBatch) {  // uniform over the block
        const int n = (int)min<uint32_t>(kRasterBatch, s1 - b);
        __syncthreads();
        if ((int)threadIdx.x < n) {
            const uint32_t id = list[b + threadIdx.x];
            bool ok = id < (uint32_t)A.numObjects;
            if (ok) {
                so[threadIdx.x] = objects[id];
                sm[threadIdx.x] = meshes[id];
                const LuxgiMeshSDF& m = sm[threadIdx.x];
                ok = m.mipCount >= 1 && m.mipCount <= 3;  // a malformed mesh entry is skipped
                for (int l = 0; ok && l < m.mipCount; ++l)
                    ok = m.size[l][0] > 0 && m.size[l][1] > 0 && m.size[l][2] > 0;
            }
            sid[threadIdx.x] = ok ? 1 : 0;
        }
        __syncthreads();
        if (b != s0) minDistance = maxD * stored;  // READ_DISTANCE (SDFRasterizeModel.glsl:50-52)
        for (int k = 0; k < n; ++k) {
            if (!sid[k]) continue;
            minDistance = fminf(minDistance, distance_to_model(minDistance, so[k], sm[k], pool, wp));
        }
        stored = store_value(minDistance, maxD, A.half);
    }
    if (!inside) return;
    const size_t e = ((size_t)vz * R + vy) * A.width + (size_t)A.p.cascadeIndex * R + vx;
    if (A.half) static_cast<__half*>(sdf)[e] = __float2half(stored);
    else static_cast<float*>(sdf)[e] = stored;
}

}  // namespace luxgi

using namespace luxgi;

extern "C" int luxgi_sdf_rasterize(const LuxgiRasterizeParams* params, const LuxgiObjectRasterizeData* objects,
                                   const LuxgiMeshSDF* meshes, const uint16_t* pool, int32_t numObjects,
                                   const int32_t* chunkOrigins, const uint32_t* chunkObjectStart,
                                   const uint32_t* chunkObjects, int32_t numChunks, int32_t cascades, int32_t storage,
                                   void* sdf, void* stream) {
    if (!params || !chunkOrigins || !chunkObjectStart || !sdf || numChunks < 0 || numObjects < 0 || cascades < 1)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (numObjects > 0 && (!objects || !meshes || !pool)) return LUXGI_ERR_INVALID_ARGUMENT;
    if (storage != LUXGI_STORAGE_FP32 && storage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    const int R = params->cascadeResolution;
    if (R <= 0 || params->cascadeIndex < 0 || params->cascadeIndex >= cascades || !(params->maxDistance > 0.0f))
        return LUXGI_ERR_INVALID_ARGUMENT;
    if ((long long)numChunks * kRasterBlocksPerChunk >= (1ll << 31)) return LUXGI_ERR_UNSUPPORTED;
    if (numChunks == 0) return LUXGI_OK;
    RasterArgs A;
    A.p = *params;
    A.numObjects = numObjects;
    A.numChunks = numChunks;
    A.width = R * cascades;
    A.cascades = cascades;
    A.half = storage == LUXGI_STORAGE_FP16 ? 1 : 0;
    hipLaunchKernelGGL(rasterize_kernel, dim3(numChunks * kRasterBlocksPerChunk), dim3(256), 0, as_stream(stream), A,
                       objects, meshes, pool, reinterpret_cast<const int4*>(chunkOrigins), chunkObjectStart,
                       chunkObjects, sdf);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

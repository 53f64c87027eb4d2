// atlas_cull.hip -- surface-atlas chunk culling for gfx950 (Shaders/SDF/SDFCulling.comp:36-101,
// driver GlobalSurfaceAtlas.cpp:607-641).
//
// The 40^3 chunk grid spans chunkSize * [-20, 20)^3 around the origin
// (SDFCulling.comp:42-43). A chunk lists every object whose bounding sphere
// touches its box (boxIntersectsSphere, AtlasCommon.glsl:34-38), in increasing
// object order. The lists are runs [count, id0, id1, ...] in culledObjects,
// whose entry 0 is the counter the reference initialises to 1
// (GlobalSurfaceAtlas.cpp:614-615), so offset 0 means "no list".
//
// The reference reserves each run with an atomicAdd, so where a run lands depends
// on thread timing. Here the runs are placed in chunk-address order: a count
// pass, an exclusive scan, a fill pass. A chunk's list, the set of chunks with a
// list and the final counter value (1 + the sum of count + 1 over the non-empty
// chunks) are the reference's. Deliberate deviations, in code the reference
// leaves racy or stale: an empty chunk, or one whose run would end past
// culledObjectsCapacity, gets offset 0 (the reference writes the chunk address
// into chunks[0] instead, SDFCulling.comp:60,69, and leaves the chunk's own entry
// as it was), and with runs in address order the chunks that overflow are the
// last ones.
#include "common.hpp"

namespace luxgi {

constexpr int kCullRes = 40;  // GLOBAL_SURFACE_ATLAS_CHUNKS_RESOLUTION, AtlasCommon.glsl:5
constexpr int kCullChunks = kCullRes * kCullRes * kCullRes;
constexpr int kCullBlock = 256;
constexpr int kCullScanBlock = 1024;
constexpr int kCullScanBlocks = (kCullChunks + kCullScanBlock - 1) / kCullScanBlock;
static_assert(kCullScanBlocks <= 64, "the block totals are scanned by one wave");

// boxIntersectsSphere (AtlasCommon.glsl:34-38) for one chunk, with the fp32 operations in the
// order written there and none contracted, so the oracle repeats the decision bit for bit
__device__ __forceinline__ bool chunk_touches(float cs, int cx, int cy, int cz, float4 s) {
    const float mnx = __fmul_rn(float(cx) - 20.0f, cs), mny = __fmul_rn(float(cy) - 20.0f, cs),
                mnz = __fmul_rn(float(cz) - 20.0f, cs);
    const float mxx = __fadd_rn(mnx, cs), mxy = __fadd_rn(mny, cs), mxz = __fadd_rn(mnz, cs);
    const float dx = __fsub_rn(s.x, fminf(fmaxf(s.x, mnx), mxx));
    const float dy = __fsub_rn(s.y, fminf(fmaxf(s.y, mny), mxy));
    const float dz = __fsub_rn(s.z, fminf(fmaxf(s.z, mnz), mxz));
    const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    return __fsqrt_rn(d2) <= s.w;
}

// One lane per chunk; the block stages the object spheres through LDS.
//   FILL = false: sizes[chunk] = count + 1, or 0 when empty.
//   FILL = true:  the run of the chunk at 1 + (exclusive scan of sizes), or offset 0.
template <bool FILL>
__global__ __launch_bounds__(kCullBlock) void cull_kernel(LuxgiSurfaceAtlasData d,
                                                          const LuxgiAtlasObject* __restrict__ objects,
                                                          uint32_t* __restrict__ sizes,
                                                          const uint32_t* __restrict__ prefix,
                                                          const uint32_t* __restrict__ totals,
                                                          uint32_t* __restrict__ chunks, uint32_t* __restrict__ out) {
    __shared__ float4 sph[kCullBlock];
    __shared__ uint32_t blockStart[64];
    const int addr = blockIdx.x * kCullBlock + threadIdx.x;
    const bool valid = addr < kCullChunks;
    const int cx = addr % kCullRes, cy = (addr / kCullRes) % kCullRes, cz = addr / (kCullRes * kCullRes);
    const int n = (int)d.objectsCount;
    bool test = valid;
    uint32_t start = 0, cursor = 0;
    if (FILL) {
        if (threadIdx.x < 64) {
            const uint32_t t = threadIdx.x < kCullScanBlocks ? totals[threadIdx.x] : 0u;
            const uint32_t x = wave_inclusive_scan(t);
            blockStart[threadIdx.x] = x - t;
            if (blockIdx.x == 0 && threadIdx.x == 63) out[0] = 1u + x;  // the reference's final counter
        }
        __syncthreads();
        if (valid) {
            const uint32_t size = sizes[addr];
            start = 1u + blockStart[addr / kCullScanBlock] + prefix[addr];
            const bool fits = size != 0u && (unsigned long long)start + size <= d.culledObjectsCapacity;
            chunks[addr] = fits ? start : 0u;
            if (fits) out[start] = size - 1u;
            test = fits;
            cursor = start + 1u;
        }
    }
    uint32_t count = 0;
    for (int b = 0; b < n; b += kCullBlock) {  // every lane reaches every barrier
        __syncthreads();
        if (b + (int)threadIdx.x < n) {
            const float* ob = objects[b + threadIdx.x].objectBounds;
            sph[threadIdx.x] = make_float4(ob[0], ob[1], ob[2], ob[3]);
        }
        __syncthreads();
        const int m = min(kCullBlock, n - b);
        if (test)
            for (int k = 0; k < m; ++k)
                if (chunk_touches(d.chunkSize, cx, cy, cz, sph[k])) {
                    if (FILL) out[cursor++] = (uint32_t)(b + k);
                    else ++count;
                }
    }
    if (!FILL && valid) sizes[addr] = count ? count + 1u : 0u;
}

}  // namespace luxgi

using namespace luxgi;

static size_t cull_align256(size_t v) { return (v + 255) & ~size_t(255); }

extern "C" size_t luxgi_atlas_cull_workspace_size(void) {
    return 2 * cull_align256((size_t)kCullChunks * sizeof(uint32_t)) + cull_align256(64 * sizeof(uint32_t));
}

extern "C" int luxgi_atlas_cull_chunks(const LuxgiSurfaceAtlasData* data, const LuxgiAtlasObject* objects,
                                       uint32_t* chunks, uint32_t* culledObjects, void* workspace,
                                       size_t workspaceBytes, void* stream) {
    if (!data || !chunks || !culledObjects || !workspace) return LUXGI_ERR_INVALID_ARGUMENT;
    if (data->objectsCount > 0 && !objects) return LUXGI_ERR_INVALID_ARGUMENT;
    if (!(data->chunkSize > 0.0f) || data->culledObjectsCapacity < 1) return LUXGI_ERR_INVALID_ARGUMENT;
    if (data->objectsCount > (1u << 30)) return LUXGI_ERR_UNSUPPORTED;
    if (workspaceBytes < luxgi_atlas_cull_workspace_size()) return LUXGI_ERR_WORKSPACE_TOO_SMALL;
    hipStream_t st = as_stream(stream);
    char* base = static_cast<char*>(workspace);
    uint32_t* sizes = reinterpret_cast<uint32_t*>(base);
    uint32_t* prefix = reinterpret_cast<uint32_t*>(base + cull_align256((size_t)kCullChunks * sizeof(uint32_t)));
    uint32_t* totals = reinterpret_cast<uint32_t*>(base + 2 * cull_align256((size_t)kCullChunks * sizeof(uint32_t)));
    const dim3 grid((kCullChunks + kCullBlock - 1) / kCullBlock);
    hipLaunchKernelGGL(cull_kernel<false>, grid, dim3(kCullBlock), 0, st, *data, objects, sizes, nullptr, nullptr,
                       chunks, culledObjects);
    LUXGI_LAUNCH_CHECK();
    hipLaunchKernelGGL(block_scan_kernel<kCullScanBlock>, dim3(kCullScanBlocks), dim3(kCullScanBlock), 0, st, sizes,
                       prefix, totals, kCullChunks);
    LUXGI_LAUNCH_CHECK();
    hipLaunchKernelGGL(cull_kernel<true>, grid, dim3(kCullBlock), 0, st, *data, objects, sizes, prefix, totals,
                       chunks, culledObjects);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

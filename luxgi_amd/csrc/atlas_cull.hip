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

// boxIntersectsSphere (AtlasCommon.glsl:34-38) against the box [lo, hi] of chunks lo..hi (inclusive
// chunk coordinates per axis), with the fp32 operations in the order written there and none
// contracted (contraction is off in these functions; HIP's __fmul_rn and __fadd_rn are plain
// operations that could otherwise fuse), so the oracle repeats the decision bit for bit. A box of
// several chunks uses the bounds of its outer chunks: its distance is never above that of a chunk
// inside it (each step is monotone), so it is a conservative pre-test for the chunks it holds.
__device__ __forceinline__ float chunk_lo(int c, float cs) {
#pragma clang fp contract(off)
    return __fmul_rn(float(c) - 20.0f, cs);
}
__device__ __forceinline__ bool box_touches(float cs, int3 lo, int3 hi, float4 s) {
#pragma clang fp contract(off)
    const float mnx = chunk_lo(lo.x, cs), mny = chunk_lo(lo.y, cs), mnz = chunk_lo(lo.z, cs);
    const float mxx = __fadd_rn(chunk_lo(hi.x, cs), cs), mxy = __fadd_rn(chunk_lo(hi.y, cs), cs),
                mxz = __fadd_rn(chunk_lo(hi.z, cs), cs);
    const float dx = __fsub_rn(s.x, fminf(fmaxf(s.x, mnx), mxx));
    const float dy = __fsub_rn(s.y, fminf(fmaxf(s.y, mny), mxy));
    const float dz = __fsub_rn(s.z, fminf(fmaxf(s.z, mnz), mxz));
    const float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    return __fsqrt_rn(d2) <= s.w;
}

// A block covers an 8x8x4 brick of chunks (40^3 = 5 x 5 x 10 such bricks), one lane per chunk.
// Per batch of 256 objects each lane tests one object against the brick's box; the objects that
// pass are compacted (in object order) into LDS, and every lane tests only those against its chunk.
//   FILL = false: sizes[chunk] = count + 1, or 0 when empty.
//   FILL = true:  the run of the chunk at 1 + (exclusive scan of sizes), or offset 0.
template <bool FILL>
__global__ __launch_bounds__(kCullBlock) void cull_kernel(LuxgiSurfaceAtlasData d,
                                                          const LuxgiAtlasObject* __restrict__ objects,
                                                          uint32_t* __restrict__ sizes,
                                                          const uint32_t* __restrict__ prefix,
                                                          const uint32_t* __restrict__ totals,
                                                          uint32_t* __restrict__ chunks, uint32_t* __restrict__ out) {
    static_assert(kCullBlock == 256 && kCullRes % 8 == 0 && kCullRes % 4 == 0, "8x8x4 chunk bricks");
    __shared__ float4 sph[kCullBlock];
    __shared__ uint32_t ids[kCullBlock];
    __shared__ uint32_t waveCount[kCullBlock / 64];
    __shared__ uint32_t blockStart[64];
    const int bx = blockIdx.x % (kCullRes / 8), by = (blockIdx.x / (kCullRes / 8)) % (kCullRes / 8),
              bz = blockIdx.x / ((kCullRes / 8) * (kCullRes / 8));
    const int cx = bx * 8 + (threadIdx.x & 7), cy = by * 8 + ((threadIdx.x >> 3) & 7), cz = bz * 4 + (threadIdx.x >> 6);
    const int addr = (cz * kCullRes + cy) * kCullRes + cx;
    const int3 blo = make_int3(bx * 8, by * 8, bz * 4), bhi = make_int3(bx * 8 + 7, by * 8 + 7, bz * 4 + 3);
    const int n = (int)d.objectsCount;
    bool test = true;
    uint32_t start = 0, cursor = 0;
    if (FILL) {
        if (threadIdx.x < 64) {
            const uint32_t t = threadIdx.x < kCullScanBlocks ? totals[threadIdx.x] : 0u;
            const uint32_t x = wave_inclusive_scan(t);
            blockStart[threadIdx.x] = x - t;
            if (blockIdx.x == 0 && threadIdx.x == 63) out[0] = 1u + x;  // the reference's final counter
        }
        __syncthreads();
        const uint32_t size = sizes[addr];
        start = 1u + blockStart[addr / kCullScanBlock] + prefix[addr];
        const bool fits = size != 0u && (unsigned long long)start + size <= d.culledObjectsCapacity;
        chunks[addr] = fits ? start : 0u;
        if (fits) out[start] = size - 1u;
        test = fits;
        cursor = start + 1u;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t count = 0;
    for (int b = 0; b < n; b += kCullBlock) {  // every lane reaches every barrier
        const int o = b + (int)threadIdx.x;
        float4 s = make_float4(0.f, 0.f, 0.f, -1.f);
        if (o < n) {
            const float* ob = objects[o].objectBounds;
            s = make_float4(ob[0], ob[1], ob[2], ob[3]);
        }
        const bool cand = o < n && box_touches(d.chunkSize, blo, bhi, s);
        const unsigned long long m = __ballot(cand);
        __syncthreads();  // the previous batch's readers are done
        if (lane == 0) waveCount[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t pos = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        uint32_t total = 0;
        for (int w = 0; w < kCullBlock / 64; ++w) {
            if (w < wave) pos += waveCount[w];
            total += waveCount[w];
        }
        if (cand) {
            sph[pos] = s;
            ids[pos] = (uint32_t)o;
        }
        __syncthreads();
        if (test)
            for (uint32_t k = 0; k < total; ++k)
                if (box_touches(d.chunkSize, make_int3(cx, cy, cz), make_int3(cx, cy, cz), sph[k])) {
                    if (FILL) out[cursor++] = ids[k];
                    else ++count;
                }
    }
    if (!FILL) sizes[addr] = count ? count + 1u : 0u;
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
    const dim3 grid(kCullChunks / kCullBlock);  // one block per 8x8x4 brick of chunks
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

// sdf_mip.hip -- coarse global-SDF "mip" used for empty-space skipping, gfx950.
//
// Reference: Shaders/SDF/GlobalSDFMipmap.comp:32-68 (7-tap min of
// combineDistanceToSDF over the centre and the 6 axis neighbours), driven by
// Engine/DDGI/GlobalDistanceField.cpp:823-846 (initial pass, mip factor 4) and
// fillFlood :537-573 (4 ping-pong passes at mip resolution, Mip->Tmp, Tmp->Mip).
// The distance offset |o| * maxDistance / res is in world units while the
// volume stores normalised distances; that mix is the reference's and is kept.
#include "common.hpp"

namespace luxgi {

__global__ void mip_pass_kernel(const float* __restrict__ src, int srcW, int srcH, int srcRes, int scale,
                                int texOffsetX, float* __restrict__ dst, int dstW, int dstH, int dstRes,
                                int mipOffsetX, float maxDistance) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int z = blockIdx.z * blockDim.z + threadIdx.z;
    if (x >= dstRes || y >= dstRes || z >= dstRes) return;
    const int ox[7] = {0, 1, 0, 0, -1, 0, 0}, oy[7] = {0, 0, 1, 0, 0, -1, 0}, oz[7] = {0, 0, 0, 1, 0, 0, -1};
    const float unit = maxDistance / float(srcRes);
    float best = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const int cx = clampi(x * scale + ox[k], 0, srcRes - 1) + texOffsetX;
        const int cy = clampi(y * scale + oy[k], 0, srcRes - 1);
        const int cz = clampi(z * scale + oz[k], 0, srcRes - 1);
        const float v = src[((size_t)cz * srcH + cy) * srcW + cx];
        const float d = (k == 0 ? 0.0f : 1.0f) * unit;  // length(offset) is 0 or 1
        float r;
        if (v <= 0.0f && d <= 0.0f) r = v;  // combineDistanceToSDF (SDFCommon.glsl:18-36)
        else {
            const float m = fmaxf(v, 0.0f);
            r = sqrtf(m * m + d * d);
        }
        best = k == 0 ? r : fminf(best, r);
    }
    dst[((size_t)z * dstH + y) * dstW + x + mipOffsetX] = best;
}

// one block per brick: the 5^3 apron block (clamped to the volume) + 3 zero pad elements, optionally
// divided by the brick's power-of-two scale (smallest 2^e >= max |v| over the block)
template <class D_>
__global__ void to_bricks_kernel(const float* __restrict__ src, int W, int H, int D, D_* __restrict__ dst,
                                 float* __restrict__ scaleOut) {
    __shared__ float red[kBrickElems / 64];
    const uint32_t nbx = (uint32_t)W >> 2, nby = (uint32_t)H >> 2;
    const uint32_t b = blockIdx.x;
    const int bx = int(b % nbx), by = int((b / nbx) % nby), bz = int(b / (nbx * nby));
    const uint32_t t = threadIdx.x;  // blockDim = kBrickElems
    float v = 0.0f;
    if (t < 125) {
        const int i = int(t % 5), j = int((t / 5) % 5), k = int(t / 25);
        const int x = min(4 * bx + i, W - 1), y = min(4 * by + j, H - 1), z = min(4 * bz + k, D - 1);
        v = src[((size_t)z * H + y) * W + x];
    }
    if (scaleOut) {
        float m = fabsf(v);
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if ((t & 63) == 0) red[t >> 6] = m;
        __syncthreads();
        m = fmaxf(red[0], red[1]);
        // smallest power of two >= m (frexp: m = f * 2^e with f in [0.5, 1))
        int e = 0;
        const float f = frexpf(m, &e);
        const float sc = m > 0.0f ? ldexpf(1.0f, f == 0.5f ? e - 1 : e) : 1.0f;
        v = v / sc;  // exact: sc is a power of two
        if (t == 0) scaleOut[b] = sc;
    }
    Texel<D_>::st1(dst, (size_t)b * kBrickElems + t, v);
}

// One thread per record of LUXGI_LAYOUT_CELL8: record r of brick b holds the 8 corner voxels of cell
// (x, y, z) = 8 * brick + (r % 8, r / 8 % 8, r / 64 % 8), clamped to the volume; padding records are zero.
template <class D_>
__global__ void to_cells_kernel(const float* __restrict__ src, int W, int H, int D, uint32_t nbx, uint32_t nby,
                                uint32_t records, D_* __restrict__ dst) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= records) return;
    const uint32_t b = r >> 9, m = r & 511u;
    const uint32_t bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    const int x = int(8 * bx + (m & 7u)), y = int(8 * by + ((m >> 3) & 7u)), z = int(8 * bz + (m >> 6));
    float v[8];
    const bool inside = x < W && y < H && z < D;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int xi = min(x + (k & 1), W - 1), yj = min(y + ((k >> 1) & 1), H - 1), zl = min(z + (k >> 2), D - 1);
        v[k] = inside ? src[((size_t)zl * H + yj) * W + xi] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) Texel<D_>::st1(dst, (size_t)r * 8 + k, v[k]);
}

__global__ void guard_kernel(const float* __restrict__ src, int W, int H, int D, float* __restrict__ guard) {
    const int nbx = W >> 2, nby = H >> 2, nbz = D >> 2;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbx * nby * nbz) return;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    float m = INFINITY;
    for (int z = 4 * bz; z <= min(4 * bz + 4, D - 1); ++z)
        for (int y = 4 * by; y <= min(4 * by + 4, H - 1); ++y)
            for (int x = 4 * bx; x <= min(4 * bx + 4, W - 1); ++x) m = fminf(m, src[((size_t)z * H + y) * W + x]);
    guard[b] = m;
}

}  // namespace luxgi

using namespace luxgi;

extern "C" int luxgi_sdf_build_guard(const float* sdf, int32_t W, int32_t H, int32_t D, float* guard, void* stream) {
    if (!sdf || !guard || W <= 0 || H <= 0 || D <= 0 || (W & 3) || (H & 3) || (D & 3)) return LUXGI_ERR_INVALID_ARGUMENT;
    const int n = (W >> 2) * (H >> 2) * (D >> 2);
    hipLaunchKernelGGL(guard_kernel, dim3((n + 255) / 256), dim3(256), 0, as_stream(stream), sdf, W, H, D, guard);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" size_t luxgi_sdf_bricks_bytes(int32_t W, int32_t H, int32_t D, int32_t storage) {
    if (W <= 0 || H <= 0 || D <= 0 || (W & 3) || (H & 3) || (D & 3)) return 0;
    if (storage != LUXGI_STORAGE_FP32 && storage != LUXGI_STORAGE_FP16 && storage != LUXGI_STORAGE_BF16) return 0;
    const size_t es = storage == LUXGI_STORAGE_FP32 ? 4 : 2;
    return (size_t)(W >> 2) * (H >> 2) * (D >> 2) * kBrickElems * es;
}

extern "C" int luxgi_sdf_to_bricks(const float* src, int32_t W, int32_t H, int32_t D, int32_t storage, void* dst,
                                   float* brickScale, void* stream) {
    const size_t bytes = luxgi_sdf_bricks_bytes(W, H, D, storage);
    if (!src || !dst || bytes == 0) return LUXGI_ERR_INVALID_ARGUMENT;
    if (bytes >= (1ull << 32)) return LUXGI_ERR_UNSUPPORTED;  // the samplers use 32-bit offsets
    const dim3 grid((unsigned)((size_t)(W >> 2) * (H >> 2) * (D >> 2)));
    if (storage == LUXGI_STORAGE_FP32)
        hipLaunchKernelGGL(to_bricks_kernel<float>, grid, dim3(kBrickElems), 0, as_stream(stream), src, W, H, D,
                           static_cast<float*>(dst), brickScale);
    else if (storage == LUXGI_STORAGE_FP16)
        hipLaunchKernelGGL(to_bricks_kernel<__half>, grid, dim3(kBrickElems), 0, as_stream(stream), src, W, H, D,
                           static_cast<__half*>(dst), brickScale);
    else
        hipLaunchKernelGGL(to_bricks_kernel<__bf16>, grid, dim3(kBrickElems), 0, as_stream(stream), src, W, H, D,
                           static_cast<__bf16*>(dst), brickScale);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" size_t luxgi_sdf_cells_bytes(int32_t W, int32_t H, int32_t D, int32_t storage) {
    if (W <= 0 || H <= 0 || D <= 0) return 0;
    if (storage != LUXGI_STORAGE_FP32 && storage != LUXGI_STORAGE_FP16 && storage != LUXGI_STORAGE_BF16) return 0;
    const size_t records = cell8_records(W, H, D);
    if (records >= (1ull << 32)) return 0;
    return records * 8 * (storage == LUXGI_STORAGE_FP32 ? 4 : 2);
}

extern "C" int luxgi_sdf_to_cells(const float* src, int32_t W, int32_t H, int32_t D, int32_t storage, void* dst,
                                  void* stream) {
    const size_t bytes = luxgi_sdf_cells_bytes(W, H, D, storage);
    if (!src || !dst || bytes == 0) return LUXGI_ERR_INVALID_ARGUMENT;
    const uint32_t records = (uint32_t)cell8_records(W, H, D);
    const uint32_t nbx = (uint32_t)(W + 7) / 8, nby = (uint32_t)(H + 7) / 8;
    const dim3 grid((records + 255) / 256);
    if (storage == LUXGI_STORAGE_FP32)
        hipLaunchKernelGGL(to_cells_kernel<float>, grid, dim3(256), 0, as_stream(stream), src, W, H, D, nbx, nby, records,
                           static_cast<float*>(dst));
    else if (storage == LUXGI_STORAGE_FP16)
        hipLaunchKernelGGL(to_cells_kernel<__half>, grid, dim3(256), 0, as_stream(stream), src, W, H, D, nbx, nby,
                           records, static_cast<__half*>(dst));
    else
        hipLaunchKernelGGL(to_cells_kernel<__bf16>, grid, dim3(256), 0, as_stream(stream), src, W, H, D, nbx, nby,
                           records, static_cast<__bf16*>(dst));
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_sdf_build_mip(const float* sdf, int32_t res, int32_t cascades, int32_t cascade, float maxDistance,
                                   int32_t floodPasses, float* mip, float* temp, void* stream) {
    // floodPasses alternate Mip -> Tmp -> Mip; an odd count would leave the result in `temp`
    if (!sdf || !mip || res <= 0 || cascades <= 0 || cascade < 0 || cascade >= cascades || floodPasses < 0 ||
        (floodPasses & 1))
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (floodPasses > 0 && !temp) return LUXGI_ERR_INVALID_ARGUMENT;
    const int mr = (res + 3) / 4;
    hipStream_t st = as_stream(stream);
    const dim3 blk(8, 8, 4);
    const dim3 grd((mr + 7) / 8, (mr + 7) / 8, (mr + 3) / 4);
    hipLaunchKernelGGL(mip_pass_kernel, grd, blk, 0, st, sdf, res * cascades, res, res, 4, cascade * res, mip,
                       mr * cascades, mr, mr, cascade * mr, maxDistance);
    LUXGI_LAUNCH_CHECK();
    for (int i = 1; i <= floodPasses; ++i) {
        if (i & 1)
            hipLaunchKernelGGL(mip_pass_kernel, grd, blk, 0, st, mip, mr * cascades, mr, mr, 1, cascade * mr, temp, mr,
                               mr, mr, 0, maxDistance);
        else
            hipLaunchKernelGGL(mip_pass_kernel, grd, blk, 0, st, temp, mr, mr, mr, 1, 0, mip, mr * cascades, mr, mr,
                               cascade * mr, maxDistance);
        LUXGI_LAUNCH_CHECK();
    }
    return LUXGI_OK;
}

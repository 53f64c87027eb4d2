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

__global__ void to_bricks_kernel(const float* __restrict__ src, int W, int H, int D, float* __restrict__ dst) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= W) return;
    dst[brick4_index(x, y, z, W >> 2, H >> 2)] = src[((size_t)z * H + y) * W + x];
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

extern "C" int luxgi_sdf_to_bricks(const float* src, int32_t W, int32_t H, int32_t D, float* dst, void* stream) {
    if (!src || !dst || W <= 0 || H <= 0 || D <= 0 || (W & 3) || (H & 3) || (D & 3)) return LUXGI_ERR_INVALID_ARGUMENT;
    hipLaunchKernelGGL(to_bricks_kernel, dim3((W + 63) / 64, H, D), dim3(64), 0, as_stream(stream), src, W, H, D, dst);
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

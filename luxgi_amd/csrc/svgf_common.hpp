// svgf_common.hpp -- helpers shared by the SVGF passes (svgf.hip) and the learned per-pixel
// edge-stopping scales (learned_pixel.hip): robust fetches, luminance, the edge-stopping weight and
// the a-trous arguments. Reference: Shaders/Raytraced/EdgeStop.glsl:7-40,
// Shaders/Reflection/ReflectionCommon.glsl:19-22.
#pragma once
#include "common.hpp"

namespace luxgi {

__device__ __forceinline__ bool inside(int x, int y, int W, int H) { return x >= 0 && y >= 0 && x < W && y < H; }
__device__ __forceinline__ float4 fetch4(const float4* p, int x, int y, int W, int H) {
    return inside(x, y, W, H) ? p[(size_t)y * W + x] : make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ __forceinline__ float fetch1(const float* p, int x, int y, int W, int H) {
    return inside(x, y, W, H) ? p[(size_t)y * W + x] : 0.0f;
}
template <class S>
__device__ __forceinline__ float4 fetchS4(const S* p, int x, int y, int W, int H) {
    return inside(x, y, W, H) ? Texel<S>::ld4(p, (size_t)y * W + x) : make_float4(0.f, 0.f, 0.f, 0.f);
}
template <class S>
__device__ __forceinline__ float2 fetchS2(const S* p, int x, int y, int W, int H) {
    return inside(x, y, W, H) ? Texel<S>::ld2(p, (size_t)y * W + x) : make_float2(0.f, 0.f);
}

__device__ __forceinline__ float luminance(float r, float g, float b) {  // ReflectionCommon.glsl:19-22
    return fmaxf(r * 0.299f + g * 0.587f + b * 0.114f, 0.0001f);
}

// computeEdgeStoppingWeight (EdgeStop.glsl:24-40) with the per-pixel constants folded:
// invPhiZ = 1/sigmaDepth, invPhiL = 1/phiLuma, and the normal power max(0, n.n')^phiN
// evaluated by `sqN` squarings when phiN = 2^sqN (the reference's 32), else by powf.
// exp(x) is evaluated as exp2(x * log2(e)) (v_exp_f32); the relative error stays below 1e-6
// for the arguments here, well inside the driver-ULP freedom of GLSL exp/pow.
__device__ __forceinline__ float pow_normal(float c, float phiN, int sqN) {
    if (sqN < 0) return powf(c, phiN);
    for (int i = 0; i < sqN; ++i) c *= c;
    return c;
}
__device__ __forceinline__ float edge_stop(float cd, float sd, float invPhiZ, f3 cn, f3 sn, float phiN, int sqN,
                                           float cl, float sl, float invPhiL) {
    const float wZ = __expf(-fabsf(cd - sd) * invPhiZ);
    const float wN = pow_normal(clampf(dot3(cn, sn), 0.0f, 1.0f), phiN, sqN);
    const float wL = fabsf(cl - sl) * invPhiL;
    return __expf(0.0f - fmaxf(wL, 0.0f) - fmaxf(wZ, 0.0f)) * wN;
}

struct AtrousArgs {
    int W, H, tilesX, radius, stepSize;
    float phi, phiNormal, sigmaDepth, power;
    int sqNormal;            // log2(phiNormal) when it is a power of two in [1, 64], else -1
    const float* lumScale;   // optional [H][W] scale of the luminance term (learned per pixel); null = 1
};

// squarings that evaluate x^phiNormal exactly as repeated products, or -1
__host__ __device__ inline int normal_squarings(float phiN) {
    for (int k = 0; k <= 6; ++k)
        if (phiN == float(1 << k)) return k;
    return -1;
}

__device__ __forceinline__ f3 load_normal(const float4* normals, size_t q) {
    const float2 o = reinterpret_cast<const float2*>(normals)[2 * q];  // octahedral xy only
    return oct_decode(o.x, o.y);
}
__device__ __forceinline__ float load_linear_depth(const float4* pbr, size_t q) {
    return reinterpret_cast<const float*>(pbr)[4 * q + 3];
}
__device__ __forceinline__ float load_roughness(const float4* pbr, size_t q) {
    return reinterpret_cast<const float*>(pbr)[4 * q + 1];
}

}  // namespace luxgi

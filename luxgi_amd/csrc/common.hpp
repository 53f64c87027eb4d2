// common.hpp -- device helpers shared by the gfx950 kernels of libluxgi.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../../include/luxgi.h"

namespace luxgi {

#define LUXGI_HIP_CHECK(expr)                                                                  \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            fprintf(stderr, "luxgi: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e),   \
                    __FILE__, __LINE__);                                                       \
            return LUXGI_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

#define LUXGI_LAUNCH_CHECK() LUXGI_HIP_CHECK(hipGetLastError())

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// small float3 helpers (explicit, so the operation order matches the oracle)
// ---------------------------------------------------------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 normalize3(f3 a) {
    float l = len3(a);
    return {a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float fractf(float v) { return v - floorf(v); }

// column-major mat4 * (v, w)
__device__ __forceinline__ f3 mat4_mul(const float* m, f3 v, float w) {
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * w, m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * w,
            m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * w};
}
__device__ __forceinline__ float4 mat4_mul4(const float* m, float4 v) {
    return make_float4(m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w,
                       m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w,
                       m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w,
                       m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w);
}

// worldPositionFromDepth, Common/Math.glsl:35-42
__device__ __forceinline__ f3 world_from_depth(float u, float v, float depth, const float* m) {
    const float4 w = mat4_mul4(m, make_float4(u * 2.0f - 1.0f, v * 2.0f - 1.0f, depth, 1.0f));
    return mk(w.x / w.w, w.y / w.w, w.z / w.w);
}

// ---------------------------------------------------------------------------
// texel storage: float (RGBA32F/RG32F/R32F) or __half (RGBA16F/RG16F/R16F)
// ---------------------------------------------------------------------------
template <class S> struct Texel;

template <> struct Texel<float> {
    static __device__ __forceinline__ float4 ld4(const float* p, size_t i) {
        return reinterpret_cast<const float4*>(p)[i];
    }
    static __device__ __forceinline__ void st4(float* p, size_t i, float4 v) { reinterpret_cast<float4*>(p)[i] = v; }
    static __device__ __forceinline__ float2 ld2(const float* p, size_t i) {
        return reinterpret_cast<const float2*>(p)[i];
    }
    static __device__ __forceinline__ void st2(float* p, size_t i, float2 v) { reinterpret_cast<float2*>(p)[i] = v; }
    static __device__ __forceinline__ float ld1(const float* p, size_t i) { return p[i]; }
    static __device__ __forceinline__ float cvt(float v) { return v; }
    static __device__ __forceinline__ void st1(float* p, size_t i, float v) { p[i] = v; }
};

template <> struct Texel<__half> {
    static __device__ __forceinline__ float4 ld4(const __half* p, size_t i) {
        uint2 u = reinterpret_cast<const uint2*>(p)[i];
        __half2 a = *reinterpret_cast<__half2*>(&u.x), b = *reinterpret_cast<__half2*>(&u.y);
        float2 fa = __half22float2(a), fb = __half22float2(b);
        return make_float4(fa.x, fa.y, fb.x, fb.y);
    }
    static __device__ __forceinline__ void st4(__half* p, size_t i, float4 v) {
        __half2 a = __floats2half2_rn(v.x, v.y), b = __floats2half2_rn(v.z, v.w);
        uint2 u;
        u.x = *reinterpret_cast<unsigned*>(&a);
        u.y = *reinterpret_cast<unsigned*>(&b);
        reinterpret_cast<uint2*>(p)[i] = u;
    }
    static __device__ __forceinline__ float2 ld2(const __half* p, size_t i) {
        return __half22float2(reinterpret_cast<const __half2*>(p)[i]);
    }
    static __device__ __forceinline__ void st2(__half* p, size_t i, float2 v) {
        reinterpret_cast<__half2*>(p)[i] = __floats2half2_rn(v.x, v.y);
    }
    static __device__ __forceinline__ float ld1(const __half* p, size_t i) { return __half2float(p[i]); }
    static __device__ __forceinline__ float cvt(__half v) { return __half2float(v); }
    static __device__ __forceinline__ void st1(__half* p, size_t i, float v) { p[i] = __float2half_rn(v); }
};

// BF16 scalar voxels (SDF volumes only): the upper half of an FP32, rounded to nearest even
template <> struct Texel<__bf16> {
    static __device__ __forceinline__ float ld1(const __bf16* p, size_t i) { return float(p[i]); }
    static __device__ __forceinline__ float cvt(__bf16 v) { return float(v); }
    static __device__ __forceinline__ void st1(__bf16* p, size_t i, float v) { p[i] = (__bf16)v; }
};

// RGBA8 unorm colour (surface atlas only)
struct Unorm8 {};
template <> struct Texel<Unorm8> {
    static __device__ __forceinline__ float4 ld4(const Unorm8* p, size_t i) {
        uint32_t u = reinterpret_cast<const uint32_t*>(p)[i];
        const float s = 1.0f / 255.0f;
        return make_float4(float(u & 255u) * s, float((u >> 8) & 255u) * s, float((u >> 16) & 255u) * s,
                           float(u >> 24) * s);
    }
};

// ---------------------------------------------------------------------------
// DDGI direction helpers (Shaders/DDGI/DDGICommon.glsl:42-92)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sign_not_zero(float k) { return k >= 0.0f ? 1.0f : -1.0f; }

__device__ __forceinline__ f3 oct_decode(float ox, float oy) {
    f3 v = {ox, oy, 1.0f - fabsf(ox) - fabsf(oy)};
    if (v.z < 0.0f) {
        float nx = (1.0f - fabsf(v.y)) * sign_not_zero(v.x);
        float ny = (1.0f - fabsf(v.x)) * sign_not_zero(v.y);
        v.x = nx;
        v.y = ny;
    }
    return normalize3(v);
}

// texel direction of interior texel (lx, ly) of an oct map of side n
// (normalizedOctCoord of DDGICommon.glsl:85-92 evaluated at the texel centre)
__device__ __forceinline__ f3 texel_direction(int lx, int ly, int n) {
    float s = 2.0f / float(n);
    return oct_decode((float(lx) + 0.5f) * s - 1.0f, (float(ly) + 0.5f) * s - 1.0f);
}

__device__ __forceinline__ f3 spherical_fibonacci(float i, float n) {
    const float PHI = sqrtf(5.0f) * 0.5f + 0.5f;
    const float PI_F = 3.14159265358979323846f;
    float a = i * (PHI - 1.0f);
    float phi = 2.0f * PI_F * (a - floorf(a));
    float cosTheta = 1.0f - (2.0f * i + 1.0f) * (1.0f / n);
    float sinTheta = sqrtf(clampf(1.0f - cosTheta * cosTheta, 0.0f, 1.0f));
    return {cosf(phi) * sinTheta, sinf(phi) * sinTheta, cosTheta};
}

// ---------------------------------------------------------------------------
// sampling
// ---------------------------------------------------------------------------
// Loads at a 32-bit byte offset from a wave-uniform base: the address stays one VGPR
// (global_load ... v_offset, s_base) instead of a 64-bit VGPR pair. Volumes passed to the
// samplers below are smaller than 4 GiB (checked by the callers' entry points).
template <class T>
__device__ __forceinline__ T load_at(const void* __restrict__ base, uint32_t byteOffset) {
    return *reinterpret_cast<const T*>(static_cast<const char*>(base) + byteOffset);
}

// texture(sampler3D) linear + clamp-to-edge on a [D][H][W] volume
template <class S>
__device__ __forceinline__ float sample3d(const S* __restrict__ vol, int W, int H, int D, float u, float v, float w) {
    float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f, z = w * float(D) - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    float ax = x - fx0, ay = y - fy0, az = z - fz0;
    int x0 = int(fx0), y0 = int(fy0), z0 = int(fz0);
    int x1 = clampi(x0 + 1, 0, W - 1), y1 = clampi(y0 + 1, 0, H - 1), z1 = clampi(z0 + 1, 0, D - 1);
    x0 = clampi(x0, 0, W - 1);
    y0 = clampi(y0, 0, H - 1);
    z0 = clampi(z0, 0, D - 1);
    // 32-bit element offsets: volumes are smaller than 4 GiB (checked by the entry points)
    const uint32_t uW = (uint32_t)W, uH = (uint32_t)H;
    const uint32_t r00 = ((uint32_t)z0 * uH + (uint32_t)y0) * uW, r10 = ((uint32_t)z0 * uH + (uint32_t)y1) * uW;
    const uint32_t r01 = ((uint32_t)z1 * uH + (uint32_t)y0) * uW, r11 = ((uint32_t)z1 * uH + (uint32_t)y1) * uW;
    constexpr uint32_t es = sizeof(S);
    float v000 = Texel<S>::cvt(load_at<S>(vol, (r00 + x0) * es)), v100 = Texel<S>::cvt(load_at<S>(vol, (r00 + x1) * es));
    float v010 = Texel<S>::cvt(load_at<S>(vol, (r10 + x0) * es)), v110 = Texel<S>::cvt(load_at<S>(vol, (r10 + x1) * es));
    float v001 = Texel<S>::cvt(load_at<S>(vol, (r01 + x0) * es)), v101 = Texel<S>::cvt(load_at<S>(vol, (r01 + x1) * es));
    float v011 = Texel<S>::cvt(load_at<S>(vol, (r11 + x0) * es)), v111 = Texel<S>::cvt(load_at<S>(vol, (r11 + x1) * es));
    float c00 = v000 * (1.0f - ax) + v100 * ax;
    float c10 = v010 * (1.0f - ax) + v110 * ax;
    float c01 = v001 * (1.0f - ax) + v101 * ax;
    float c11 = v011 * (1.0f - ax) + v111 * ax;
    float c0 = c00 * (1.0f - ay) + c10 * ay;
    float c1 = c01 * (1.0f - ay) + c11 * ay;
    return c0 * (1.0f - az) + c1 * az;
}

// Same filtering as sample3d<float>, with each x-adjacent voxel pair fetched by one
// 8-byte load (dword-aligned global_load_dwordx2): 4 memory requests per lookup
// instead of 8. Requires W >= 2. Clamp-to-edge is reproduced by loading the pair
// at min(max(x0, 0), W - 2) and selecting its halves.
typedef float float2_a4 __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ float sample3d_pairs(const float* __restrict__ vol, int W, int H, int D, float u, float v,
                                                float w) {
    float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f, z = w * float(D) - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    float ax = x - fx0, ay = y - fy0, az = z - fz0;
    int x0 = int(fx0), y0 = int(fy0), z0 = int(fz0);
    const int y1 = clampi(y0 + 1, 0, H - 1), z1 = clampi(z0 + 1, 0, D - 1);
    y0 = clampi(y0, 0, H - 1);
    z0 = clampi(z0, 0, D - 1);
    const int b = clampi(x0, 0, W - 2);
    const bool left = x0 < 0, right = x0 >= W - 1;
    const uint32_t uW = (uint32_t)W, uH = (uint32_t)H;
    const float2_a4 p00 = load_at<float2_a4>(vol, (((uint32_t)z0 * uH + (uint32_t)y0) * uW + (uint32_t)b) * 4u);
    const float2_a4 p10 = load_at<float2_a4>(vol, (((uint32_t)z0 * uH + (uint32_t)y1) * uW + (uint32_t)b) * 4u);
    const float2_a4 p01 = load_at<float2_a4>(vol, (((uint32_t)z1 * uH + (uint32_t)y0) * uW + (uint32_t)b) * 4u);
    const float2_a4 p11 = load_at<float2_a4>(vol, (((uint32_t)z1 * uH + (uint32_t)y1) * uW + (uint32_t)b) * 4u);
    const float v000 = right ? p00.y : p00.x, v100 = left ? p00.x : p00.y;
    const float v010 = right ? p10.y : p10.x, v110 = left ? p10.x : p10.y;
    const float v001 = right ? p01.y : p01.x, v101 = left ? p01.x : p01.y;
    const float v011 = right ? p11.y : p11.x, v111 = left ? p11.x : p11.y;
    float c00 = v000 * (1.0f - ax) + v100 * ax;
    float c10 = v010 * (1.0f - ax) + v110 * ax;
    float c01 = v001 * (1.0f - ax) + v101 * ax;
    float c11 = v011 * (1.0f - ax) + v111 * ax;
    float c0 = c00 * (1.0f - ay) + c10 * ay;
    float c1 = c01 * (1.0f - ay) + c11 * ay;
    return c0 * (1.0f - az) + c1 * az;
}

// Bricked volume (LUXGI_LAYOUT_BRICK4): W/4 x H/4 x D/4 bricks in x-fastest order, each stored
// with a +1 apron as 5x5x5 voxels (k*25 + j*5 + i for voxel (4bx+i, 4by+j, 4bz+k), clamped to the
// volume) padded to 128 elements. Every trilinear footprint whose lower corner lies in brick b
// is then inside b's 5^3 block: 4 x-adjacent voxel pairs, each one 2-element load, instead of 8
// scattered loads. W, H, D are multiples of 4. Same filtering arithmetic as sample3d.
constexpr uint32_t kBrickElems = 128;

typedef _Float16 half2_a2 __attribute__((ext_vector_type(2), aligned(2)));
template <class S> struct VoxelPair;
template <> struct VoxelPair<float> {
    typedef float2_a4 T;
    static __device__ __forceinline__ float2 get(T p) { return make_float2(p.x, p.y); }
};
template <> struct VoxelPair<__half> {
    typedef half2_a2 T;
    static __device__ __forceinline__ float2 get(T p) { return make_float2(float(p.x), float(p.y)); }
};
typedef __bf16 bf16x2_a2 __attribute__((ext_vector_type(2), aligned(2)));
template <> struct VoxelPair<__bf16> {
    typedef bf16x2_a2 T;
    static __device__ __forceinline__ float2 get(T p) { return make_float2(float(p.x), float(p.y)); }
};

__device__ __forceinline__ uint32_t brick_index(int x, int y, int z, uint32_t nbx, uint32_t nby) {
    return ((uint32_t)(z >> 2) * nby + (uint32_t)(y >> 2)) * nbx + (uint32_t)(x >> 2);
}

template <class S>
__device__ __forceinline__ float sample3d_brick4(const S* __restrict__ vol, int W, int H, int D, float u, float v,
                                                 float w, const float* __restrict__ scale = nullptr) {
    typedef typename VoxelPair<S>::T PT;
    constexpr uint32_t es = sizeof(S);
    float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f, z = w * float(D) - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    float ax = x - fx0, ay = y - fy0, az = z - fz0;
    const int x0 = int(fx0), y0 = int(fy0), z0 = int(fz0);
    // clamp-to-edge: a lower corner below 0 reads voxel 0 twice; one at the upper edge reads the
    // apron, which holds the clamped edge voxel
    const bool left = x0 < 0;
    const int xc = clampi(x0, 0, W - 1), yc = clampi(y0, 0, H - 1), zc = clampi(z0, 0, D - 1);
    const uint32_t b = brick_index(xc, yc, zc, (uint32_t)W >> 2, (uint32_t)H >> 2);
    const int jl = yc & 3, kl = zc & 3;
    const int jh = y0 < 0 ? jl : jl + 1, kh = z0 < 0 ? kl : kl + 1;
    const uint32_t base = b * kBrickElems + (uint32_t)(xc & 3);
    const float2 p00 = VoxelPair<S>::get(load_at<PT>(vol, (base + kl * 25 + jl * 5) * es));
    const float2 p10 = VoxelPair<S>::get(load_at<PT>(vol, (base + kl * 25 + jh * 5) * es));
    const float2 p01 = VoxelPair<S>::get(load_at<PT>(vol, (base + kh * 25 + jl * 5) * es));
    const float2 p11 = VoxelPair<S>::get(load_at<PT>(vol, (base + kh * 25 + jh * 5) * es));
    const float v000 = p00.x, v100 = left ? p00.x : p00.y;
    const float v010 = p10.x, v110 = left ? p10.x : p10.y;
    const float v001 = p01.x, v101 = left ? p01.x : p01.y;
    const float v011 = p11.x, v111 = left ? p11.x : p11.y;
    float c00 = v000 * (1.0f - ax) + v100 * ax;
    float c10 = v010 * (1.0f - ax) + v110 * ax;
    float c01 = v001 * (1.0f - ax) + v101 * ax;
    float c11 = v011 * (1.0f - ax) + v111 * ax;
    float c0 = c00 * (1.0f - ay) + c10 * ay;
    float c1 = c01 * (1.0f - ay) + c11 * ay;
    const float r = c0 * (1.0f - az) + c1 * az;
    return scale ? r * scale[b] : r;  // a power of two: exact
}

// Cell-record volume (LUXGI_LAYOUT_CELL8): one record of the 8 corner voxels per cell, 8x8x8-cell
// bricks in x-fastest order, cells x-fastest inside a brick.
inline size_t cell8_records(int W, int H, int D) {
    return (size_t)((W + 7) / 8) * ((H + 7) / 8) * ((D + 7) / 8) * 512;
}
__device__ __forceinline__ uint32_t cell8_index(int x, int y, int z, uint32_t nbx, uint32_t nby) {
    const uint32_t b = ((uint32_t)(z >> 3) * nby + (uint32_t)(y >> 3)) * nbx + (uint32_t)(x >> 3);
    return (b << 9) | (((uint32_t)z & 7u) << 6) | (((uint32_t)y & 7u) << 3) | ((uint32_t)x & 7u);
}

template <class S> struct CellRecord;
template <> struct CellRecord<float> {  // 32 B: two 16-byte loads of one 32-byte-aligned record
    static __device__ __forceinline__ void load(const float* __restrict__ vol, uint32_t rec, float (&v)[8]) {
        const float4* p = reinterpret_cast<const float4*>(vol) + (size_t)rec * 2;
        const float4 a = p[0], b = p[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
};
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
template <> struct CellRecord<__bf16> {  // 16 B: one load; a BF16 is the upper half of an FP32
    static __device__ __forceinline__ void load(const __bf16* __restrict__ vol, uint32_t rec, float (&v)[8]) {
        const uint4 q = reinterpret_cast<const uint4*>(vol)[rec];
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[2 * k] = __uint_as_float(w[k] << 16);
            v[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
        }
    }
};
template <> struct CellRecord<__half> {  // 16 B: one load
    static __device__ __forceinline__ void load(const __half* __restrict__ vol, uint32_t rec, float (&v)[8]) {
        const half8_t h = reinterpret_cast<const half8_t*>(vol)[rec];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = float(h[k]);
    }
};

// texture(sampler3D) linear + clamp-to-edge from cell records: the same corners and the same
// arithmetic as sample3d, so the same result bit for bit. A lower corner at -1 (the first half
// voxel) reads cell 0 and takes its lower corner for both; the upper clamp is in the record.
template <class S>
__device__ __forceinline__ float sample3d_cell8(const S* __restrict__ vol, int W, int H, int D, float u, float v,
                                                float w) {
    float x = u * float(W) - 0.5f, y = v * float(H) - 0.5f, z = w * float(D) - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    float ax = x - fx0, ay = y - fy0, az = z - fz0;
    const int x0 = int(fx0), y0 = int(fy0), z0 = int(fz0);
    const bool lx = x0 < 0, ly = y0 < 0, lz = z0 < 0;
    const uint32_t rec = cell8_index(clampi(x0, 0, W - 1), clampi(y0, 0, H - 1), clampi(z0, 0, D - 1),
                                     (uint32_t)(W + 7) >> 3, (uint32_t)(H + 7) >> 3);
    float r[8];
    CellRecord<S>::load(vol, rec, r);
    const float v000 = r[0];
    const float v100 = lx ? r[0] : r[1];
    const float v010 = ly ? r[0] : r[2];
    const float v110 = ly ? v100 : (lx ? r[2] : r[3]);
    const float v001 = lz ? r[0] : r[4];
    const float v101 = lz ? v100 : (lx ? r[4] : r[5]);
    const float v011 = lz ? v010 : (ly ? r[4] : r[6]);
    const float v111 = lz ? v110 : (ly ? (lx ? r[4] : r[5]) : (lx ? r[6] : r[7]));
    float c00 = v000 * (1.0f - ax) + v100 * ax;
    float c10 = v010 * (1.0f - ax) + v110 * ax;
    float c01 = v001 * (1.0f - ax) + v101 * ax;
    float c11 = v011 * (1.0f - ax) + v111 * ax;
    float c0 = c00 * (1.0f - ay) + c10 * ay;
    float c1 = c01 * (1.0f - ay) + c11 * ay;
    return c0 * (1.0f - az) + c1 * az;
}

// volume lookup used by the march: pair loads for FP32 volumes of width >= 2
template <class S>
__device__ __forceinline__ float sample_volume(const S* __restrict__ vol, int W, int H, int D, float u, float v,
                                               float w) {
    return sample3d<S>(vol, W, H, D, u, v, w);
}
template <>
__device__ __forceinline__ float sample_volume<float>(const float* __restrict__ vol, int W, int H, int D, float u,
                                                     float v, float w) {
    return W >= 2 ? sample3d_pairs(vol, W, H, D, u, v, w) : sample3d<float>(vol, W, H, D, u, v, w);
}

// ---------------------------------------------------------------------------
// prefix sums
// ---------------------------------------------------------------------------
// inclusive prefix sum across the 64 lanes of a wave
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, (unsigned)o);
        if (lane >= o) x += y;
    }
    return x;
}

// Exclusive scan of in[0, n) within blocks of BLOCK elements (out may alias in); the total of
// block b goes to totals[b]. Launch ceil(n / BLOCK) blocks of BLOCK lanes.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void block_scan_kernel(const uint32_t* in, uint32_t* out,
                                                           uint32_t* __restrict__ totals, int n) {
    static_assert(BLOCK % 64 == 0 && BLOCK <= 1024, "whole waves, one block");
    __shared__ uint32_t waveSum[BLOCK / 64];
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    const uint32_t v = i < n ? in[i] : 0u;
    const uint32_t x = wave_inclusive_scan(v);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) waveSum[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    for (int k = 0; k < wave; ++k) before += waveSum[k];
    if (i < n) out[i] = before + x - v;
    if (threadIdx.x == BLOCK - 1) totals[blockIdx.x] = before + x;
}

// XCD-aware block order. The dispatcher deals the workgroups of a launch to the 8 XCDs round-robin
// (workgroup b, counted x-fastest, runs on XCD b % 8); xcd_block(b, nb) renumbers them so that each
// XCD takes a contiguous 1/8 of the work, whose data then stays in that XCD's L2. The last
// nb % 8 blocks keep their numbers.
constexpr int kXcds = 8;

__device__ __forceinline__ long long xcd_block(long long b, long long nb) {
    const long long full = (nb / kXcds) * kXcds;
    if (b >= full) return b;
    return (b % kXcds) * (nb / kXcds) + b / kXcds;
}

// the (x, y) block of a 2D grid after xcd_block: each XCD takes a contiguous band of block rows
__device__ __forceinline__ int2 xcd_block_2d() {
    const long long nb = (long long)gridDim.x * gridDim.y;
    const long long t = xcd_block((long long)blockIdx.y * gridDim.x + blockIdx.x, nb);
    return make_int2(int(t % gridDim.x), int(t / gridDim.x));
}

}  // namespace luxgi

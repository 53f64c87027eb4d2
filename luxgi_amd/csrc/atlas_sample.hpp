// atlas_sample.hpp -- the global surface atlas lookup (Shaders/SDF/AtlasCommon.glsl:50-157) and the
// sky cube map read, shared by the probe-ray shading and the SDF reflection producer.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace luxgi {

constexpr int kChunksRes = 40;  // GLOBAL_SURFACE_ATLAS_CHUNKS_RESOLUTION, AtlasCommon.glsl:5

// GLSL inverse(mat4) by cofactors (column-major), evaluated once per object.
__device__ __forceinline__ void object_inverse_kernel_body(const LuxgiAtlasObject* __restrict__ objects, int n,
                                                           float* __restrict__ inv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* m = objects[i].transform;
    float o[16];
    o[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    o[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    o[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    o[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    o[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    o[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    o[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    o[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    o[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    o[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    o[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    o[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    o[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    o[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    o[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    o[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const float det = m[0] * o[0] + m[1] * o[4] + m[2] * o[8] + m[3] * o[12];
    const float r = 1.0f / det;
    for (int k = 0; k < 16; ++k) inv[(size_t)i * 16 + k] = o[k] * r;
}

struct Counts {
    unsigned long long mip, fine, normals, objects, tiles, steps, hits;
};

// Packed surface atlas (LUXGI_LAYOUT_BRICK4): per texel uint2 (RGBA8 colour, FP32 depth bits),
// 4x4-texel blocks of 128 bytes, blocks row-major.
struct PackedAtlas {};

__device__ __forceinline__ uint32_t atlas_block_index(int i, int j, int res) {
    return (((uint32_t)(j >> 2) * (uint32_t)(res >> 2) + (uint32_t)(i >> 2)) << 4) + (uint32_t)((j & 3) * 4 + (i & 3));
}

// textureGather of a repeat-addressed [res][res] image (Vulkan footprint order), as element
// indices of the atlas layout
template <class SC>
__device__ __forceinline__ void gather_footprint(int res, float u, float v, size_t idx[4]) {
    float x = u * float(res) - 0.5f, y = v * float(res) - 0.5f;
    int i0 = int(floorf(x)), j0 = int(floorf(y));
    int i1 = i0 + 1, j1 = j0 + 1;
    auto wrap = [res](int i) {
        int m = i % res;
        return m < 0 ? m + res : m;
    };
    i0 = wrap(i0); i1 = wrap(i1); j0 = wrap(j0); j1 = wrap(j1);
    if constexpr (std::is_same<SC, PackedAtlas>::value) {
        idx[0] = atlas_block_index(i0, j1, res);
        idx[1] = atlas_block_index(i1, j1, res);
        idx[2] = atlas_block_index(i1, j0, res);
        idx[3] = atlas_block_index(i0, j0, res);
    } else {
        idx[0] = (size_t)j1 * res + i0;
        idx[1] = (size_t)j1 * res + i1;
        idx[2] = (size_t)j0 * res + i1;
        idx[3] = (size_t)j0 * res + i0;
    }
}

__device__ __forceinline__ float4 unorm8x4(uint32_t u) {
    const float s = 1.0f / 255.0f;
    return make_float4(float(u & 255u) * s, float((u >> 8) & 255u) * s, float((u >> 16) & 255u) * s,
                       float(u >> 24) * s);
}

// sampleGlobalSurfaceAtlasTile (AtlasCommon.glsl:58-112)
template <class SC, bool COUNT>
__device__ float4 sample_tile(const LuxgiSurfaceAtlasData& ad, const LuxgiAtlasTile* __restrict__ tp,
                              const SC* __restrict__ color, const float* __restrict__ depth, f3 lp, f3 ln,
                              float threshold, Counts& c) {
    constexpr bool kPacked = std::is_same<SC, PackedAtlas>::value;
    float tm[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) tm[k] = tp->transform[k];
    const float4 ob = *reinterpret_cast<const float4*>(tp->objectBounds);
    const float4 ex = *reinterpret_cast<const float4*>(tp->extends);
    f3 n = normalize3(mat4_mul(tm, ln, 1.0f));
    float nw = clampf(n.z, 0.0f, 1.0f);
    nw = (nw - 0.05f) / (1.0f - 0.05f);
    if (nw <= 0.0f) return make_float4(0.f, 0.f, 0.f, 0.f);
    f3 t = mat4_mul(tm, lp, 1.0f);
    const float tileDepth = t.z / ob.z;
    const float tu = clampf(t.x / ob.x + 0.5f, 0.0f, 1.0f), tv = clampf(t.y / ob.y + 0.5f, 0.0f, 1.0f);
    const float au = tu * ex.z + ex.x, av = tv * ex.w + ex.y;
    const float bx = fractf(au * float(ad.resolution) + 0.5f), by = fractf(av * float(ad.resolution) + 0.5f);
    float bw[4] = {(1.0f - bx) * by, bx * by, bx * (1.0f - by), (1.0f - bx) * (1.0f - by)};
    size_t idx[4];
    gather_footprint<SC>((int)ad.resolution, au, av, idx);
    if (COUNT) c.tiles++;
    const float thr = 2.0f * threshold / ob.z;
    float vis[4];
    uint2 packed[4];
    if constexpr (kPacked) {
#pragma unroll
        for (int i = 0; i < 4; ++i) packed[i] = reinterpret_cast<const uint2*>(color)[idx[i]];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float tz;
        if constexpr (kPacked) tz = __uint_as_float(packed[i].y);
        else tz = depth[idx[i]];
        vis[i] = 1.0f - clampf((fabsf(tileDepth - tz) - thr) / (0.5f * thr), 0.0f, 1.0f);
        if (tz >= 1.0f) vis[i] = 0.0f;
    }
    float sw = vis[0] * bw[0] + vis[1] * bw[1] + vis[2] * bw[2] + vis[3] * bw[3];
    sw *= nw;
    if (sw <= 0.0f) return make_float4(0.f, 0.f, 0.f, 0.f);
    float r = 0.f, g = 0.f, b = 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float w = bw[i] * vis[i];
        float4 cc;
        if constexpr (kPacked) cc = unorm8x4(packed[i].x);
        else cc = Texel<SC>::ld4(color, idx[i]);
        r += cc.x * w;
        g += cc.y * w;
        b += cc.z * w;
    }
    return make_float4(r * sw, g * sw, b * sw, sw);
}

// sampleGlobalSurfaceAtlas (AtlasCommon.glsl:115-157), debug = false: (rgb / max(a, 1e-4), a) with a
// the summed tile weights, (0, 0, 0, 0) when the chunk has no list
template <class SC, bool COUNT>
__device__ float4 sample_atlas(const LuxgiSurfaceAtlasData& ad, const uint32_t* __restrict__ chunks,
                               const uint32_t* __restrict__ culled, const LuxgiAtlasObject* __restrict__ objects,
                               const float* __restrict__ objInv, const LuxgiAtlasTile* __restrict__ tiles,
                               const SC* __restrict__ color, const float* __restrict__ depth, f3 wp, f3 wn,
                               float threshold, Counts& c) {
    float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
    int cc[3];
    const float pc[3] = {wp.x, wp.y, wp.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) cc[i] = clampi(int(floorf(pc[i] / ad.chunkSize + float(kChunksRes) * 0.5f)), 0, kChunksRes - 1);
    uint32_t start = chunks[cc[2] * kChunksRes * kChunksRes + cc[1] * kChunksRes + cc[0]];
    if (start == 0) return make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t count = culled[start];
    if (count > ad.objectsCount) return make_float4(0.f, 0.f, 0.f, 0.f);
    start++;
    for (uint32_t oi = 0; oi < count; ++oi) {
        const uint32_t id = culled[start++];
        const LuxgiAtlasObject* o = objects + id;
        if (COUNT) c.objects++;
        const float4 bs = *reinterpret_cast<const float4*>(o->objectBounds);
        if (len3(mk(bs.x, bs.y, bs.z) - wp) > bs.w) continue;
        const float* inv = objInv + (size_t)id * 16;
        float m[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) m[k] = inv[k];
        const float4 l4 = mat4_mul4(m, make_float4(wp.x, wp.y, wp.z, 1.0f));
        const f3 lp = mk(l4.x, l4.y, l4.z);
        const float4 e = *reinterpret_cast<const float4*>(o->extends);
        if (fabsf(lp.x) > e.x + threshold || fabsf(lp.y) > e.y + threshold || fabsf(lp.z) > e.z + threshold) continue;
        f3 ln = mk(m[0] * wn.x + m[4] * wn.y + m[8] * wn.z, m[1] * wn.x + m[5] * wn.y + m[9] * wn.z,
                   m[2] * wn.x + m[6] * wn.y + m[10] * wn.z);
        ln = normalize3(ln);
        for (int t = 0; t < 6; ++t) {
            const uint32_t to = o->tileOffset[t];
            if (to != 0) {
                float4 s = sample_tile<SC, COUNT>(ad, tiles + to, color, depth, lp, ln, threshold, c);
                res.x += s.x;
                res.y += s.y;
                res.z += s.z;
                res.w += s.w;
            }
        }
    }
    const float a = fmaxf(res.w, 0.0001f);
    return make_float4(res.x / a, res.y / a, res.z / a, res.w);
}

// cubemap lookup (RGBA32F faces +X -X +Y -Y +Z -Z), bilinear inside the face, clamp-to-edge
__device__ float4 sample_cube4(const float* __restrict__ faces, int size, f3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) { ma = ax; face = d.x >= 0 ? 0 : 1; sc = d.x >= 0 ? -d.z : d.z; tc = -d.y; }
    else if (ay >= az) { ma = ay; face = d.y >= 0 ? 2 : 3; sc = d.x; tc = d.y >= 0 ? d.z : -d.z; }
    else { ma = az; face = d.z >= 0 ? 4 : 5; sc = d.z >= 0 ? d.x : -d.x; tc = -d.y; }
    const float s = 0.5f * (sc / ma + 1.0f), t = 0.5f * (tc / ma + 1.0f);
    const float x = s * float(size) - 0.5f, y = t * float(size) - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const float a = x - fx0, b = y - fy0;
    int x0 = int(fx0), y0 = int(fy0);
    const int x1 = clampi(x0 + 1, 0, size - 1), y1 = clampi(y0 + 1, 0, size - 1);
    x0 = clampi(x0, 0, size - 1);
    y0 = clampi(y0, 0, size - 1);
    const float4* f = reinterpret_cast<const float4*>(faces) + (size_t)face * size * size;
    const float4 t00 = f[(size_t)y0 * size + x0], t10 = f[(size_t)y0 * size + x1];
    const float4 t01 = f[(size_t)y1 * size + x0], t11 = f[(size_t)y1 * size + x1];
    return make_float4((t00.x * (1.f - a) + t10.x * a) * (1.f - b) + (t01.x * (1.f - a) + t11.x * a) * b,
                       (t00.y * (1.f - a) + t10.y * a) * (1.f - b) + (t01.y * (1.f - a) + t11.y * a) * b,
                       (t00.z * (1.f - a) + t10.z * a) * (1.f - b) + (t01.z * (1.f - a) + t11.z * a) * b,
                       (t00.w * (1.f - a) + t10.w * a) * (1.f - b) + (t01.w * (1.f - a) + t11.w * a) * b);
}
__device__ __forceinline__ float3 sample_cube(const float* __restrict__ faces, int size, f3 d) {
    const float4 c = sample_cube4(faces, size, d);
    return make_float3(c.x, c.y, c.z);
}

}  // namespace luxgi

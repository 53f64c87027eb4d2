// luxgi_oracle.hpp -- CPU restatement of LuxGI's GI hot path (TEST INFRASTRUCTURE).
//
// This header is the checker, never the product: only tests/, the smoke test and
// the CPU-baseline leg of bench.py load it (through liboracle.so). It restates
// the reference GLSL compute passes as scalar C++ templated on the arithmetic
// type T: T = double is the FP64 oracle, T = float is the FP32 CPU transcription
// timed as the CPU baseline.
//
// Texture semantics assumed for the GLSL built-ins (the Vulkan driver's sampler
// precision is not pinned by the reference):
//   texelFetch outside the image returns 0 (robustImageAccess behaviour);
//   texture() = bilinear/trilinear with exact (not 8-bit) sub-texel weights;
//   textureGather returns the 2x2 footprint in Vulkan order (i0,j1),(i1,j1),(i1,j0),(i0,j0).
//
// Every function cites the reference file:line it follows (paths relative to
// Code/Maple/src/ of the reference).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/luxgi.h"

namespace oracle {

// ---------------------------------------------------------------------------
// small vector types
// ---------------------------------------------------------------------------
template <class T> struct v2 { T x, y; };
template <class T> struct v3 { T x, y, z; };
template <class T> struct v4 { T x, y, z, w; };

template <class T> inline v3<T> mk3(T x, T y, T z) { return {x, y, z}; }
template <class T> inline v3<T> operator+(v3<T> a, v3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline v3<T> operator-(v3<T> a, v3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline v3<T> operator*(v3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline v3<T> operator*(v3<T> a, v3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline T dot(v3<T> a, v3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T length(v3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline v3<T> normalize(v3<T> a) { T l = length(a); return {a.x / l, a.y / l, a.z / l}; }
template <class T> inline T clampT(T v, T lo, T hi) { return std::min(std::max(v, lo), hi); }
template <class T> inline T fract(T v) { return v - std::floor(v); }
template <class T> inline T mixT(T a, T b, T t) { return a * (T(1) - t) + b * t; }
// GLSL min/max on NaN operands follow fmin/fmax (return the non-NaN operand).
template <class T> inline T gmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T gmax(T a, T b) { return std::fmax(a, b); }

// column-major mat4 * (v, w)
template <class T> inline v3<T> mat4_mul(const float* m, v3<T> v, T w) {
    return {T(m[0]) * v.x + T(m[4]) * v.y + T(m[8]) * v.z + T(m[12]) * w,
            T(m[1]) * v.x + T(m[5]) * v.y + T(m[9]) * v.z + T(m[13]) * w,
            T(m[2]) * v.x + T(m[6]) * v.y + T(m[10]) * v.z + T(m[14]) * w};
}
template <class T> inline v4<T> mat4_mul4(const T* m, v4<T> v) {
    return {m[0] * v.x + m[4] * v.y + m[8] * v.z + m[12] * v.w,
            m[1] * v.x + m[5] * v.y + m[9] * v.z + m[13] * v.w,
            m[2] * v.x + m[6] * v.y + m[10] * v.z + m[14] * v.w,
            m[3] * v.x + m[7] * v.y + m[11] * v.z + m[15] * v.w};
}

// GLSL inverse(mat4): adjugate / determinant, column-major in and out.
template <class T> inline void inverse4(const float* mf, T* out) {
    T m[16];
    for (int i = 0; i < 16; ++i) m[i] = T(mf[i]);
    T inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    T det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    T r = T(1) / det;
    for (int i = 0; i < 16; ++i) out[i] = inv[i] * r;
}

// ---------------------------------------------------------------------------
// DDGI indexing and direction helpers (Shaders/DDGI/DDGICommon.glsl)
// ---------------------------------------------------------------------------
constexpr double kPi = 3.1415926535897932384626433832795;  // Common/Math.glsl:7
constexpr float kGlobalSDFWorldSize = 60000.0f;            // SDF/SDFCommon.glsl:6

// DDGIRenderer.cpp:182-185
inline void atlas_sizes(int oct, int px, int py, int pz, int* w, int* h) {
    *w = (oct + 2) * px * py + 2;
    *h = (oct + 2) * pz + 2;
}

// DDGICommon.glsl:116-123
inline void probe_index_to_grid(const LuxgiDDGIUniform& u, int index, int* gx, int* gy, int* gz) {
    *gx = index % u.probeCounts[0];
    *gy = (index % (u.probeCounts[0] * u.probeCounts[1])) / u.probeCounts[0];
    *gz = index / (u.probeCounts[0] * u.probeCounts[1]);
}
// DDGICommon.glsl:136-139
inline int grid_to_probe_index(const LuxgiDDGIUniform& u, int x, int y, int z) {
    return x + y * u.probeCounts[0] + z * u.probeCounts[0] * u.probeCounts[1];
}
// DDGICommon.glsl:101-114
template <class T> inline v3<T> probe_location(const LuxgiDDGIUniform& u, int index) {
    int gx, gy, gz;
    probe_index_to_grid(u, index, &gx, &gy, &gz);
    return {T(u.step[0]) * T(gx) + T(u.startPosition[0]), T(u.step[1]) * T(gy) + T(u.startPosition[1]),
            T(u.step[2]) * T(gz) + T(u.startPosition[2])};
}
// DDGICommon.glsl:94-99 (texel integer coordinates)
inline int get_probe_id(int tx, int ty, int width, int side) {
    int pw = side + 2;
    int perSide = (width - 2) / pw;
    return tx / pw + perSide * (ty / pw);
}
// DDGICommon.glsl:85-92
template <class T> inline v2<T> normalized_oct_coord(int fx, int fy, int side) {
    int pw = side + 2;
    int ox = (fx - 2) % pw, oy = (fy - 2) % pw;
    return {(T(ox) + T(0.5)) * (T(2) / T(side)) - T(1), (T(oy) + T(0.5)) * (T(2) / T(side)) - T(1)};
}
template <class T> inline T sign_not_zero(T k) { return k >= T(0) ? T(1) : T(-1); }  // DDGICommon.glsl:55-58
// DDGICommon.glsl:65-72
template <class T> inline v2<T> oct_encode(v3<T> v) {
    T l1 = std::fabs(v.x) + std::fabs(v.y) + std::fabs(v.z);
    v2<T> r{v.x * (T(1) / l1), v.y * (T(1) / l1)};
    if (v.z < T(0)) {
        v2<T> o{(T(1) - std::fabs(r.y)) * sign_not_zero(r.x), (T(1) - std::fabs(r.x)) * sign_not_zero(r.y)};
        r = o;
    }
    return r;
}
// DDGICommon.glsl:74-82
template <class T> inline v3<T> oct_decode(v2<T> o) {
    v3<T> v{o.x, o.y, T(1) - std::fabs(o.x) - std::fabs(o.y)};
    if (v.z < T(0)) {
        T nx = (T(1) - std::fabs(v.y)) * sign_not_zero(v.x);
        T ny = (T(1) - std::fabs(v.x)) * sign_not_zero(v.y);
        v.x = nx;
        v.y = ny;
    }
    return normalize(v);
}
// Common/Math.glsl:27-33 (G-buffer decode; step(0,x)*2-1 == signNotZero)
template <class T> inline v3<T> octahedral_to_direction(T ex, T ey) { return oct_decode<T>({ex, ey}); }

// DDGICommon.glsl:42-52
template <class T> inline v3<T> spherical_fibonacci(T i, T n) {
    const T PHI = std::sqrt(T(5)) * T(0.5) + T(0.5);
    T a = i * (PHI - T(1));
    T phi = T(2) * T(kPi) * (a - std::floor(a));
    T cosTheta = T(1) - (T(2) * i + T(1)) * (T(1) / n);
    T sinTheta = std::sqrt(clampT(T(1) - cosTheta * cosTheta, T(0), T(1)));
    return {std::cos(phi) * sinTheta, std::sin(phi) * sinTheta, cosTheta};
}
// GISDFRays.comp:73: normalize(mat3(randomOrientation) * sphericalFibonacci(rayId, raysPerProbe))
template <class T> inline v3<T> ray_direction(const float* rot, int rayId, int raysPerProbe) {
    v3<T> s = spherical_fibonacci<T>(T(rayId), T(raysPerProbe));
    return normalize(mat4_mul<T>(rot, s, T(0)));
}

// Border copy table (BorderUpdate.glsl:25-133 for N = 8 / 16, generalised to any N).
// Entry k = (srcX, srcY, dstX, dstY) relative to the probe's bordered corner.
inline std::vector<int> border_table(int n) {
    std::vector<int> t;
    for (int i = 1; i <= n; ++i) { int e[4] = {n + 1 - i, 1, i, 0}; t.insert(t.end(), e, e + 4); }
    for (int i = 1; i <= n; ++i) { int e[4] = {n + 1 - i, n, i, n + 1}; t.insert(t.end(), e, e + 4); }
    for (int i = 1; i <= n; ++i) { int e[4] = {1, n + 1 - i, 0, i}; t.insert(t.end(), e, e + 4); }
    for (int i = 1; i <= n; ++i) { int e[4] = {n, n + 1 - i, n + 1, i}; t.insert(t.end(), e, e + 4); }
    int c[16] = {n, n, 0, 0, 1, n, n + 1, 0, n, 1, 0, n + 1, 1, 1, n + 1, n + 1};
    t.insert(t.end(), c, c + 16);
    return t;
}

// ---------------------------------------------------------------------------
// Probe update (Shaders/DDGI/ProbeUpdate.glsl:53-152)
// rays: radiance / dirDist as float4 [height][width]; prev/out atlases float4
// (irradiance) / float2 (depth). Output arrays are T.
// ---------------------------------------------------------------------------
template <class T>
void probe_update_probe(const LuxgiDDGIUniform& u, const float* radiance, const float* dirDist, int rayWidth,
                        const float* prevIrr, const float* prevDepth, T* outIrr, T* outDepth, bool firstFrame,
                        bool doIrr, bool doDepth, int probe) {
    const int pxy = u.probeCounts[0] * u.probeCounts[1];
    const int col = probe % pxy, row = probe / pxy;
    const T FLT_EPS = T(0.00000001f);  // ProbeUpdate.glsl:51
    for (int pass = 0; pass < 2; ++pass) {
        const bool depth = pass == 1;
        if ((depth && !doDepth) || (!depth && !doIrr)) continue;
        const int side = depth ? u.depthProbeSideLength : u.irradianceProbeSideLength;
        const int W = depth ? u.depthTextureWidth : u.irradianceTextureWidth;
        for (int ly = 0; ly < side; ++ly)
            for (int lx = 0; lx < side; ++lx) {
                // ProbeUpdate.glsl:107: gl_GlobalInvocationID + 2*gl_WorkGroupID + 2
                const int cx = col * (side + 2) + lx + 2, cy = row * (side + 2) + ly + 2;
                const int rel = get_probe_id(cx, cy, W, side);  // ProbeUpdate.glsl:109
                v3<T> texelDir = oct_decode(normalized_oct_coord<T>(cx, cy, side));
                T rx = 0, ry = 0, rz = 0, total = 0;
                for (int r = 0; r < u.raysPerProbe; ++r) {  // ProbeUpdate.glsl:66-103 over the 64-ray chunks
                    const float* dd = dirDist + ((size_t)rel * rayWidth + r) * 4;
                    v3<T> rd{T(dd[0]), T(dd[1]), T(dd[2])};
                    T d = dot(texelDir, rd);
                    T w;
                    if (depth) {
                        T dist = std::min(T(u.maxDistance), T(dd[3]) - T(0.01f));
                        if (dist == T(-1.0f)) dist = T(u.maxDistance);
                        w = std::pow(std::max(T(0), d), T(u.sharpness));
                        if (w >= FLT_EPS) { rx += dist * w; ry += dist * dist * w; total += w; }
                    } else {
                        w = std::max(T(0), d);
                        if (w >= FLT_EPS) {
                            const float* rr = radiance + ((size_t)rel * rayWidth + r) * 4;
                            rx += T(rr[0]) * w; ry += T(rr[1]) * w; rz += T(rr[2]) * w; total += w;
                        }
                    }
                }
                if (total > FLT_EPS) {  // ProbeUpdate.glsl:133-134
                    T s = T(1) / (T(2) * total);
                    rx *= s; ry *= s; rz *= s;
                }
                const size_t texel = (size_t)cy * W + cx;
                if (depth) {
                    if (!firstFrame) {
                        T h = T(u.hysteresis);
                        rx = mixT(rx, T(prevDepth[texel * 2 + 0]), h);
                        ry = mixT(ry, T(prevDepth[texel * 2 + 1]), h);
                    }
                    outDepth[texel * 2 + 0] = rx;
                    outDepth[texel * 2 + 1] = ry;
                } else {
                    T ig = T(1) / T(u.ddgiGamma);  // ProbeUpdate.glsl:141
                    rx = std::pow(rx, ig); ry = std::pow(ry, ig); rz = std::pow(rz, ig);
                    if (!firstFrame) {
                        T h = T(u.hysteresis);
                        rx = mixT(rx, T(prevIrr[texel * 4 + 0]), h);
                        ry = mixT(ry, T(prevIrr[texel * 4 + 1]), h);
                        rz = mixT(rz, T(prevIrr[texel * 4 + 2]), h);
                    }
                    outIrr[texel * 4 + 0] = rx; outIrr[texel * 4 + 1] = ry;
                    outIrr[texel * 4 + 2] = rz; outIrr[texel * 4 + 3] = T(1);
                }
            }
    }
}

// BorderUpdate.glsl:136-156, in place on a T atlas with `ch` channels.
template <class T> void border_update(T* atlas, int W, int H, int side, int ch, int px, int py, int pz) {
    std::vector<int> tab = border_table(side);
    const int n = (int)tab.size() / 4;
    for (int gz = 0; gz < pz; ++gz)
        for (int gxy = 0; gxy < px * py; ++gxy) {
            const int ox = gxy * (side + 2) + 1, oy = gz * (side + 2) + 1;  // BorderUpdate.glsl:150
            for (int k = 0; k < n; ++k) {
                int sx = ox + tab[k * 4 + 0], sy = oy + tab[k * 4 + 1];
                int dx = ox + tab[k * 4 + 2], dy = oy + tab[k * 4 + 3];
                if (sx < 0 || sy < 0 || dx < 0 || dy < 0 || sx >= W || dx >= W || sy >= H || dy >= H) continue;
                for (int c = 0; c < ch; ++c) atlas[((size_t)dy * W + dx) * ch + c] = atlas[((size_t)sy * W + sx) * ch + c];
            }
        }
}

// ---------------------------------------------------------------------------
// Texture sampling helpers
// ---------------------------------------------------------------------------
// texture(sampler3D, uvw) with linear filtering and clamp-to-edge on a [D][H][W] R32F volume.
template <class T> inline T sample3d(const float* vol, int W, int H, int D, T u, T v, T w) {
    T x = u * T(W) - T(0.5), y = v * T(H) - T(0.5), z = w * T(D) - T(0.5);
    T fx0 = std::floor(x), fy0 = std::floor(y), fz0 = std::floor(z);
    T ax = x - fx0, ay = y - fy0, az = z - fz0;
    int x0 = (int)fx0, y0 = (int)fy0, z0 = (int)fz0;
    int x1 = std::min(std::max(x0 + 1, 0), W - 1), y1 = std::min(std::max(y0 + 1, 0), H - 1), z1 = std::min(std::max(z0 + 1, 0), D - 1);
    x0 = std::min(std::max(x0, 0), W - 1); y0 = std::min(std::max(y0, 0), H - 1); z0 = std::min(std::max(z0, 0), D - 1);
    auto at = [&](int xx, int yy, int zz) { return T(vol[((size_t)zz * H + yy) * W + xx]); };
    T c00 = at(x0, y0, z0) * (T(1) - ax) + at(x1, y0, z0) * ax;
    T c10 = at(x0, y1, z0) * (T(1) - ax) + at(x1, y1, z0) * ax;
    T c01 = at(x0, y0, z1) * (T(1) - ax) + at(x1, y0, z1) * ax;
    T c11 = at(x0, y1, z1) * (T(1) - ax) + at(x1, y1, z1) * ax;
    T c0 = c00 * (T(1) - ay) + c10 * ay;
    T c1 = c01 * (T(1) - ay) + c11 * ay;
    return c0 * (T(1) - az) + c1 * az;
}

// Cubemap lookup, bilinear with clamp-to-edge inside the selected face.
template <class T> inline v3<T> sample_cube(const float* faces, int size, v3<T> d) {
    T ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int face;
    T sc, tc, ma;
    if (ax >= ay && ax >= az) { ma = ax; face = d.x >= 0 ? 0 : 1; sc = d.x >= 0 ? -d.z : d.z; tc = -d.y; }
    else if (ay >= az) { ma = ay; face = d.y >= 0 ? 2 : 3; sc = d.x; tc = d.y >= 0 ? d.z : -d.z; }
    else { ma = az; face = d.z >= 0 ? 4 : 5; sc = d.z >= 0 ? d.x : -d.x; tc = -d.y; }
    T s = T(0.5) * (sc / ma + T(1)), t = T(0.5) * (tc / ma + T(1));
    T x = s * T(size) - T(0.5), y = t * T(size) - T(0.5);
    T fx0 = std::floor(x), fy0 = std::floor(y);
    T a = x - fx0, b = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0;
    int x1 = std::min(std::max(x0 + 1, 0), size - 1), y1 = std::min(std::max(y0 + 1, 0), size - 1);
    x0 = std::min(std::max(x0, 0), size - 1); y0 = std::min(std::max(y0, 0), size - 1);
    const float* f = faces + (size_t)face * size * size * 4;
    v3<T> r;
    T* rp[3] = {&r.x, &r.y, &r.z};
    for (int c = 0; c < 3; ++c) {
        T t00 = T(f[((size_t)y0 * size + x0) * 4 + c]), t10 = T(f[((size_t)y0 * size + x1) * 4 + c]);
        T t01 = T(f[((size_t)y1 * size + x0) * 4 + c]), t11 = T(f[((size_t)y1 * size + x1) * 4 + c]);
        *rp[c] = (t00 * (T(1) - a) + t10 * a) * (T(1) - b) + (t01 * (T(1) - a) + t11 * a) * b;
    }
    return r;
}

// ---------------------------------------------------------------------------
// Global SDF trace (Shaders/SDF/SDFCommon.glsl:18-36, 64-95, 98-194)
// ---------------------------------------------------------------------------
template <class T> inline T combine_distance_to_sdf(T sdf, T d) {  // SDFCommon.glsl:18-36
    if (sdf <= T(0) && d <= T(0)) return sdf;
    T m = std::max(sdf, T(0));
    return std::sqrt(m * m + d * d);
}

template <class T> struct Hit {
    v3<T> normal{0, 0, 0};
    T time = T(-1);
    int cascade = 0;
    int steps = 0;
    T sdf = T(0);
};

struct TraceCounts {
    uint64_t mip = 0, fine = 0, normals = 0, objects = 0, tiles = 0, rays = 0, steps = 0, hits = 0;
};

// lineHitAABB (SDFCommon.glsl:84-95)
template <class T> inline v2<T> line_hit_aabb(v3<T> a, v3<T> b, v3<T> mn, v3<T> mx) {
    v3<T> inv{T(1) / (b.x - a.x), T(1) / (b.y - a.y), T(1) / (b.z - a.z)};
    v3<T> en = (mn - a) * inv, ex = (mx - a) * inv;
    v3<T> lo{gmin(en.x, ex.x), gmin(en.y, ex.y), gmin(en.z, ex.z)};
    v3<T> hi{gmax(en.x, ex.x), gmax(en.y, ex.y), gmax(en.z, ex.z)};
    T x = gmax(lo.x, gmax(lo.y, lo.z));
    T y = gmin(hi.x, gmin(hi.y, hi.z));
    return {clampT(x, T(0), T(1)), clampT(y, T(0), T(1))};
}

// tracyGlobalSDF (SDFCommon.glsl:98-194), stepScale 1; the probe rays use needsHitNormal true and
// cascadeTraceStartBias 0 (GISDFRays.comp:79-87), the surface-cache shadow rays false and 2
// (SDFDeferredLight.frag:104-111).
template <class T>
Hit<T> trace_global_sdf(const LuxgiGlobalSDFData& g, const float* sdf, const float* mip, int res, int mipRes,
                        v3<T> origin, v3<T> dir, T maxDistance, int maxSteps, T hitEps, TraceCounts* cnt,
                        T cascadeTraceStartBias = T(0), bool needsHitNormal = true) {
    Hit<T> hit;
    const int C = (int)g.cascadesCount;
    T traceMax = std::min(maxDistance, T(g.cascadePosDistance[C - 1][3]) * T(2));
    T chunkSizeDistance = T(32.0f) / T(g.resolution);   // GLOBAL_SDF_RASTERIZE_CHUNK_SIZE
    T chunkMarginDistance = T(4.0f) / T(g.resolution);  // GLOBAL_SDF_RASTERIZE_CHUNK_MARGIN
    T nextStart = 0;
    v3<T> traceEnd = origin + dir * traceMax;
    for (int c = 0; c < C && hit.time < T(0); ++c) {
        v3<T> cpos{T(g.cascadePosDistance[c][0]), T(g.cascadePosDistance[c][1]), T(g.cascadePosDistance[c][2])};
        T half = T(g.cascadePosDistance[c][3]);
        T voxel = T(g.cascadeVoxelSize[c]);
        T voxelHalf = voxel * T(0.5f);
        v3<T> wpos = origin + dir * (voxel * cascadeTraceStartBias);
        v2<T> is = line_hit_aabb(wpos, traceEnd, cpos - mk3(half, half, half), cpos + mk3(half, half, half));
        is.x *= traceMax;
        is.y *= traceMax;
        is.x = std::max(is.x, nextStart);
        T stepTime = is.x;
        if (is.x >= is.y) stepTime = is.y;
        else nextStart = is.y;
        int step = 0;
        for (; step < maxSteps && stepTime < is.y; ++step) {
            v3<T> p = wpos + dir * stepTime;
            // getGlobalSDFCascadeUV (SDFCommon.glsl:64-71)
            T cmax = half * T(2);
            v3<T> cuv{clampT((p.x - cpos.x) / cmax + T(0.5f), T(0), T(1)), clampT((p.y - cpos.y) / cmax + T(0.5f), T(0), T(1)),
                      clampT((p.z - cpos.z) / cmax + T(0.5f), T(0), T(1))};
            v3<T> tuv{(T(c) + cuv.x) / T(C), cuv.y, cuv.z};
            T sd = sample3d<T>(mip, mipRes * C, mipRes, mipRes, tuv.x, tuv.y, tuv.z);
            if (cnt) cnt->mip++;
            if (sd < chunkSizeDistance) {
                T st = sample3d<T>(sdf, res * C, res, res, tuv.x, tuv.y, tuv.z);
                if (cnt) cnt->fine++;
                if (st < chunkMarginDistance * T(2)) sd = st;
            } else {
                sd = chunkSizeDistance;
            }
            sd *= cmax;
            T thick = hitEps > T(0) ? hitEps : voxelHalf * clampT(stepTime / voxel, T(0), T(1));
            if (sd < thick) {
                hit.time = std::max(stepTime + sd - thick, T(0));
                hit.cascade = c;
                hit.sdf = sd;
                if (!needsHitNormal) break;
                T o = T(1.0f) / T(g.resolution);
                auto s = [&](T a, T b, T cc) { return sample3d<T>(sdf, res * C, res, res, a, b, cc); };
                T xp = s(tuv.x + o, tuv.y, tuv.z), xn = s(tuv.x - o, tuv.y, tuv.z);
                T yp = s(tuv.x, tuv.y + o, tuv.z), yn = s(tuv.x, tuv.y - o, tuv.z);
                T zp = s(tuv.x, tuv.y, tuv.z + o), zn = s(tuv.x, tuv.y, tuv.z - o);
                hit.normal = normalize(v3<T>{xp - xn, yp - yn, zp - zn});
                if (cnt) cnt->normals++;
                break;
            }
            stepTime += std::max(sd, voxel);  // stepScale = 1
        }
        hit.steps += step;
    }
    return hit;
}

// ---------------------------------------------------------------------------
// Surface atlas lookup (Shaders/SDF/AtlasCommon.glsl:50-157)
// colour is float4 [res][res] (already converted to float), depth float [res][res];
// gathers use repeat addressing (default sampler, RHI/Definitions.h:152-159).
// ---------------------------------------------------------------------------
template <class T>
inline void gather4(const float* img, int ch, int comp, int res, T u, T v, T out[4]) {
    T x = u * T(res) - T(0.5), y = v * T(res) - T(0.5);
    int i0 = (int)std::floor(x), j0 = (int)std::floor(y);
    auto wrap = [&](int i) { int m = i % res; return m < 0 ? m + res : m; };
    int i1 = wrap(i0 + 1), j1 = wrap(j0 + 1);
    i0 = wrap(i0); j0 = wrap(j0);
    out[0] = T(img[((size_t)j1 * res + i0) * ch + comp]);
    out[1] = T(img[((size_t)j1 * res + i1) * ch + comp]);
    out[2] = T(img[((size_t)j0 * res + i1) * ch + comp]);
    out[3] = T(img[((size_t)j0 * res + i0) * ch + comp]);
}

// sampleGlobalSurfaceAtlasTile (AtlasCommon.glsl:58-112)
template <class T>
inline v4<T> sample_atlas_tile(const LuxgiSurfaceAtlasData& d, const LuxgiAtlasTile& tile, const float* depthImg,
                               const float* colorImg, v3<T> localPos, v3<T> localNormal, T surfaceThreshold,
                               TraceCounts* cnt) {
    v3<T> n = normalize(mat4_mul<T>(tile.transform, localNormal, T(1)));
    T nw = clampT(n.z, T(0), T(1));
    nw = (nw - T(0.05f)) / (T(1) - T(0.05f));
    if (nw <= T(0)) return {0, 0, 0, 0};
    v3<T> tp = mat4_mul<T>(tile.transform, localPos, T(1));
    T tileDepth = tp.z / T(tile.objectBounds[2]);
    T tu = clampT(tp.x / T(tile.objectBounds[0]) + T(0.5), T(0), T(1));
    T tv = clampT(tp.y / T(tile.objectBounds[1]) + T(0.5), T(0), T(1));
    T au = tu * T(tile.extends[2]) + T(tile.extends[0]);
    T av = tv * T(tile.extends[3]) + T(tile.extends[1]);
    T bx = fract(au * T(d.resolution) + T(0.5f)), by = fract(av * T(d.resolution) + T(0.5f));
    T bw[4] = {(T(1) - bx) * by, bx * by, bx * (T(1) - by), (T(1) - bx) * (T(1) - by)};
    T tz[4];
    gather4<T>(depthImg, 1, 0, (int)d.resolution, au, av, tz);
    if (cnt) cnt->tiles++;
    T thr = T(2) * surfaceThreshold / T(tile.objectBounds[2]);
    T vis[4];
    for (int i = 0; i < 4; ++i) {
        vis[i] = T(1) - clampT((std::fabs(tileDepth - tz[i]) - thr) / (T(0.5f) * thr), T(0), T(1));
        if (tz[i] >= T(1)) vis[i] = T(0);
    }
    T sw = vis[0] * bw[0] + vis[1] * bw[1] + vis[2] * bw[2] + vis[3] * bw[3];
    sw *= nw;
    if (sw <= T(0)) return {0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) bw[i] *= vis[i];
    T col[3];
    for (int c = 0; c < 3; ++c) {
        T g[4];
        gather4<T>(colorImg, 4, c, (int)d.resolution, au, av, g);
        col[c] = g[0] * bw[0] + g[1] * bw[1] + g[2] * bw[2] + g[3] * bw[3];
    }
    return {col[0] * sw, col[1] * sw, col[2] * sw, sw};
}

// sampleGlobalSurfaceAtlas macro (AtlasCommon.glsl:115-157), debug = false.
template <class T>
inline v4<T> sample_surface_atlas(const LuxgiSurfaceAtlasData& d, const uint32_t* chunks, const uint32_t* culled,
                                  const LuxgiAtlasObject* objects, const LuxgiAtlasTile* tiles,
                                  const float* depthImg, const float* colorImg, v3<T> wp, v3<T> wn, T threshold,
                                  TraceCounts* cnt) {
    v4<T> res{0, 0, 0, 0};
    const int R = 40;  // GLOBAL_SURFACE_ATLAS_CHUNKS_RESOLUTION
    int cc[3];
    T pc[3] = {wp.x, wp.y, wp.z};
    for (int i = 0; i < 3; ++i) {
        T f = std::floor(pc[i] / T(d.chunkSize) + T(R * 0.5f));
        int v = (int)f;
        cc[i] = std::min(std::max(v, 0), R - 1);
    }
    uint32_t addr = (uint32_t)(cc[2] * R * R + cc[1] * R + cc[0]);
    uint32_t start = chunks[addr];
    if (start == 0) return res;
    uint32_t count = culled[start];
    if (count > d.objectsCount) return res;
    start++;
    for (uint32_t oi = 0; oi < count; ++oi) {
        const LuxgiAtlasObject& obj = objects[culled[start++]];
        if (cnt) cnt->objects++;
        v3<T> oc{T(obj.objectBounds[0]), T(obj.objectBounds[1]), T(obj.objectBounds[2])};
        if (length(oc - wp) > T(obj.objectBounds[3])) continue;
        T inv[16];
        inverse4<T>(obj.transform, inv);
        v4<T> lp4 = mat4_mul4<T>(inv, {wp.x, wp.y, wp.z, T(1)});
        v3<T> lp{lp4.x, lp4.y, lp4.z};
        v3<T> ext{T(obj.extends[0]) + threshold, T(obj.extends[1]) + threshold, T(obj.extends[2]) + threshold};
        if (std::fabs(lp.x) > ext.x || std::fabs(lp.y) > ext.y || std::fabs(lp.z) > ext.z) continue;
        v3<T> ln{inv[0] * wn.x + inv[4] * wn.y + inv[8] * wn.z, inv[1] * wn.x + inv[5] * wn.y + inv[9] * wn.z,
                 inv[2] * wn.x + inv[6] * wn.y + inv[10] * wn.z};
        ln = normalize(ln);
        for (int t = 0; t < 6; ++t) {
            uint32_t to = obj.tileOffset[t];
            if (to != 0) {
                v4<T> s = sample_atlas_tile<T>(d, tiles[to], depthImg, colorImg, lp, ln, threshold, cnt);
                res.x += s.x; res.y += s.y; res.z += s.z; res.w += s.w;
            }
        }
    }
    T a = std::max(res.w, T(0.0001f));
    res.x /= a; res.y /= a; res.z /= a;
    return res;
}

// GISDFRays.comp:63-128 for one (probe, ray). Writes radiance[4], dirDist[4], flag.
template <class T>
void trace_ray(const LuxgiDDGIUniform& u, const LuxgiGlobalSDFData& g, const float* sdf, const float* mip, int res,
               int mipRes, const LuxgiSurfaceAtlasData& ad, const uint32_t* chunks, const uint32_t* culled,
               const LuxgiAtlasObject* objects, const LuxgiAtlasTile* tiles, const float* atlasDepth,
               const float* atlasColor, const float* skyFaces, int skySize, const LuxgiTraceParams& p,
               const float* probeOffsets, int probe, int ray, T* outRad, T* outDD, uint8_t* flag, TraceCounts* cnt) {
    v3<T> origin = probe_location<T>(u, probe);
    if (probeOffsets) {
        origin.x += T(probeOffsets[probe * 4 + 0]);
        origin.y += T(probeOffsets[probe * 4 + 1]);
        origin.z += T(probeOffsets[probe * 4 + 2]);
    }
    v3<T> dir = ray_direction<T>(p.randomOrientation, ray, u.raysPerProbe);
    Hit<T> h = trace_global_sdf<T>(g, sdf, mip, res, mipRes, origin, dir, T(kGlobalSDFWorldSize),
                                   p.maxSteps > 0 ? p.maxSteps : 250, T(p.hitEpsilon), cnt);
    if (cnt) { cnt->rays++; cnt->steps += (uint64_t)h.steps; }
    T rad[4] = {0, 0, 0, 0};
    uint8_t f = 0;
    if (h.time >= T(0)) {
        if (cnt) cnt->hits++;
        if (h.sdf <= T(0) && h.time <= T(g.cascadeVoxelSize[0])) {  // GISDFRays.comp:93-96
            rad[3] = T(kGlobalSDFWorldSize);
            f = 2;
        } else {
            v3<T> hp = origin + dir * h.time;
            T thr = T(g.cascadeVoxelSize[h.cascade]) * T(1.05f);  // SDFCommon.glsl:196-199
            v4<T> sc = sample_surface_atlas<T>(ad, chunks, culled, objects, tiles, atlasDepth, atlasColor, hp,
                                               h.normal, thr, cnt);
            rad[0] = sc.x; rad[1] = sc.y; rad[2] = sc.z;
            rad[3] = std::max(h.time + T(g.cascadeVoxelSize[h.cascade]) * T(0.5f), T(0));  // GISDFRays.comp:118
            f = 1;
        }
    } else {
        v3<T> sky = skyFaces ? sample_cube<T>(skyFaces, skySize, dir)
                             : v3<T>{T(p.skyColor[0]), T(p.skyColor[1]), T(p.skyColor[2])};
        rad[0] = sky.x; rad[1] = sky.y; rad[2] = sky.z;
        rad[3] = T(kGlobalSDFWorldSize);
    }
    outRad[0] = rad[0]; outRad[1] = rad[1]; outRad[2] = rad[2]; outRad[3] = T(0);
    outDD[0] = dir.x; outDD[1] = dir.y; outDD[2] = dir.z; outDD[3] = rad[3];
    if (flag) *flag = f;
}

// ---------------------------------------------------------------------------
// Coarse SDF mip (Shaders/SDF/GlobalSDFMipmap.comp:32-68) + flood passes
// (Engine/DDGI/GlobalDistanceField.cpp:537-573, driver :823-846)
// ---------------------------------------------------------------------------
template <class T>
void mip_pass(const float* src, int srcW, int srcH, int srcD, int srcRes, int scale, int texOffsetX,
              float* dst, int dstW, int dstH, int dstRes, int mipOffsetX, T maxDistance) {
    (void)srcD; (void)dstH;
    static const int offs[7][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    for (int z = 0; z < dstRes; ++z)
        for (int y = 0; y < dstRes; ++y)
            for (int x = 0; x < dstRes; ++x) {
                T best = 0;
                for (int k = 0; k < 7; ++k) {
                    int c[3] = {x * scale + offs[k][0], y * scale + offs[k][1], z * scale + offs[k][2]};
                    for (int i = 0; i < 3; ++i) c[i] = std::min(std::max(c[i], 0), srcRes - 1);
                    c[0] += texOffsetX;
                    T v = T(src[((size_t)c[2] * srcH + c[1]) * srcW + c[0]]);
                    T len = std::sqrt(T(offs[k][0] * offs[k][0] + offs[k][1] * offs[k][1] + offs[k][2] * offs[k][2]));
                    T dv = len * (maxDistance / T(srcRes));
                    T r = combine_distance_to_sdf(v, dv);
                    best = k == 0 ? r : std::min(best, r);
                }
                dst[((size_t)z * dstH + y) * dstW + x + mipOffsetX] = (float)best;
            }
}

// ---------------------------------------------------------------------------
// SVGF (Shaders/Raytraced/Reprojection.glsl, EdgeStop.glsl,
//       Shaders/Reflection/Denoise*.comp, Shaders/Shadow/Denoise*.comp)
// Images are float [H][W][4] (RG images stored with 4 channels here for
// uniformity: channels beyond the image's count are ignored).
// ---------------------------------------------------------------------------
struct Img {
    const float* p;
    int ch;
};
template <class T> inline v4<T> fetch(Img im, int W, int H, int x, int y) {  // robust texelFetch
    if (x < 0 || y < 0 || x >= W || y >= H) return {0, 0, 0, 0};
    const float* q = im.p + ((size_t)y * W + x) * im.ch;
    v4<T> r{T(q[0]), T(0), T(0), T(0)};
    if (im.ch > 1) r.y = T(q[1]);
    if (im.ch > 2) r.z = T(q[2]);
    if (im.ch > 3) r.w = T(q[3]);
    return r;
}

// worldPositionFromDepth, Common/Math.glsl:35-42
template <class T> inline v3<T> world_from_depth(T u, T v, T depth, const float* vpi) {
    v4<T> n{u * T(2) - T(1), v * T(2) - T(1), depth, T(1)};
    T m[16];
    for (int i = 0; i < 16; ++i) m[i] = T(vpi[i]);
    v4<T> w = mat4_mul4<T>(m, n);
    return {w.x / w.w, w.y / w.w, w.z / w.w};
}

// isValid (Reprojection.glsl:12-53)
template <class T>
inline bool is_valid(int hx, int hy, v3<T> pos, v3<T> prevPos, v3<T> n, v3<T> pn, T meshId, T prevMeshId, int W, int H) {
    if (hx < 0 || hy < 0 || hx > W - 1 || hy > H - 1) return false;
    if (meshId != prevMeshId) return false;
    if (std::fabs(dot(pos - prevPos, n)) > T(5.0f)) return false;
    T d = std::fabs(dot(n, pn));
    if (d * d <= T(0.1f)) return false;
    return true;
}

// GLSL int(float): truncation toward zero
template <class T> inline int itrunc(T v) { return (int)v; }

struct GB {
    Img color, normals, pbr, depth;
};

// reproject, RGB overload (Reprojection.glsl:242-405)
template <class T>
bool reproject_rgb(int cx, int cy, T depth, const float* camPos, const float* vpi, const float* pvp, T rayLength,
                   const GB& cur, const GB& prev, Img histOut, Img histMom, int W, int H, v3<T>* hc, v2<T>* hm,
                   T* hl) {
    const T sx = T(W), sy = T(H);
    T tcx = (T(cx) + T(0.5f)) / sx, tcy = (T(cy) + T(0.5f)) / sy;
    v4<T> normal = fetch<T>(cur.normals, W, H, cx, cy);
    T mx = normal.z, my = normal.w;
    v3<T> wpos = world_from_depth<T>(tcx, tcy, depth, vpi);
    T meshId = fetch<T>(cur.color, W, H, cx, cy).w;
    T curvature = fetch<T>(cur.pbr, W, H, cx, cy).z;
    T htx = tcx + mx, hty = tcy + my;
    // computeHistoryCoord (Reprojection.glsl:78-86)
    T rx = T(cx) + mx * sx, ry = T(cy) + my * sy;
    if (rayLength > T(0) && curvature == T(0)) {  // virtualPointReprojection (Reprojection.glsl:56-75)
        T u = T(cx) / sx, v = T(cy) / sy;
        v3<T> ro = world_from_depth<T>(u, v, depth, vpi);
        v3<T> cp{T(camPos[0]), T(camPos[1]), T(camPos[2])};
        v3<T> cr = ro - cp;
        T cl = length(cr);
        cr = normalize(cr);
        v3<T> ph = cp + cr * (cl + rayLength);
        T m[16];
        for (int i = 0; i < 16; ++i) m[i] = T(pvp[i]);
        v4<T> rp = mat4_mul4<T>(m, {ph.x, ph.y, ph.z, T(1)});
        rp.x /= rp.w; rp.y /= rp.w;
        rx = (rp.x * T(0.5f) + T(0.5f)) * sx;
        ry = (rp.y * T(0.5f) + T(0.5f)) * sy;
    }
    int hx = itrunc(rx), hy = itrunc(ry);
    v3<T> n = octahedral_to_direction<T>(normal.x, normal.y);
    static const int off[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    bool v[4];
    bool valid = false;
    for (int i = 0; i < 4; ++i) {
        int lx = hx + off[i][0], ly = hy + off[i][1];
        v4<T> pn = fetch<T>(prev.normals, W, H, lx, ly);
        T pd = fetch<T>(prev.depth, W, H, lx, ly).x;
        v3<T> pp = world_from_depth<T>(htx, hty, pd, vpi);
        T pm = fetch<T>(prev.color, W, H, lx, ly).w;
        v[i] = is_valid<T>(hx, hy, wpos, pp, n, octahedral_to_direction<T>(pn.x, pn.y), meshId, pm, W, H);
        valid = valid || v[i];
    }
    v3<T> col{0, 0, 0};
    v2<T> mom{0, 0};
    if (valid) {
        T sumw = 0;
        T x = fract(rx), y = fract(ry);
        T w[4] = {(T(1) - x) * (T(1) - y), x * (T(1) - y), (T(1) - x) * y, x * y};
        for (int i = 0; i < 4; ++i) {
            int lx = hx + off[i][0], ly = hy + off[i][1];
            if (v[i]) {
                v4<T> c = fetch<T>(histOut, W, H, lx, ly);
                v4<T> m = fetch<T>(histMom, W, H, lx, ly);
                col = col + v3<T>{c.x, c.y, c.z} * w[i];
                mom.x += w[i] * m.x; mom.y += w[i] * m.y;
                sumw += w[i];
            }
        }
        valid = sumw >= T(0.01f);
        if (valid) { col = col * (T(1) / sumw); mom.x /= sumw; mom.y /= sumw; }
        else { col = {0, 0, 0}; mom = {0, 0}; }
    }
    if (!valid) {
        T count = 0;
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) {
                int px = hx + xx, py = hy + yy;
                v4<T> pn = fetch<T>(prev.normals, W, H, px, py);
                T pd = fetch<T>(prev.depth, W, H, px, py).x;
                v3<T> pp = world_from_depth<T>(htx, hty, pd, vpi);
                T pm = fetch<T>(prev.color, W, H, px, py).w;
                if (is_valid<T>(hx, hy, wpos, pp, n, octahedral_to_direction<T>(pn.x, pn.y), meshId, pm, W, H)) {
                    v4<T> c = fetch<T>(histOut, W, H, px, py);
                    v4<T> m = fetch<T>(histMom, W, H, px, py);
                    col = col + v3<T>{c.x, c.y, c.z};
                    mom.x += m.x; mom.y += m.y;
                    count += T(1);
                }
            }
        if (count > T(0)) {
            valid = true;
            col = col * (T(1) / count);
            mom.x /= count; mom.y /= count;
        }
    }
    if (valid) {
        *hl = fetch<T>(histMom, W, H, hx, hy).z;
    } else {
        col = {0, 0, 0}; mom = {0, 0}; *hl = 0;
    }
    *hc = col;
    *hm = mom;
    return valid;
}

// reproject, scalar overload (Reprojection.glsl:90-238), used by the shadow denoiser.
template <class T>
bool reproject_scalar(int cx, int cy, T depth, const float* vpi, const GB& cur, const GB& prev, Img histOut,
                      Img histMom, int W, int H, T* hc, v2<T>* hm, T* hl) {
    const T sx = T(W), sy = T(H);
    T pcx = T(cx) + T(0.5f), pcy = T(cy) + T(0.5f);
    T tcx = pcx / sx, tcy = pcy / sy;
    v4<T> normal = fetch<T>(cur.normals, W, H, cx, cy);
    T mx = normal.z, my = normal.w;
    v3<T> wpos = world_from_depth<T>(tcx, tcy, depth, vpi);
    T meshId = fetch<T>(cur.color, W, H, cx, cy).w;
    int hx = itrunc(pcx + mx * sx), hy = itrunc(pcy + my * sy);
    T fx = std::floor(T(cx)) + mx * sx, fy = std::floor(T(cy)) + my * sy;
    int fxi = itrunc(fx), fyi = itrunc(fy);
    T htx = tcx + mx, hty = tcy + my;
    v3<T> n = octahedral_to_direction<T>(normal.x, normal.y);
    static const int off[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
    bool v[4];
    bool valid = false;
    for (int i = 0; i < 4; ++i) {
        int lx = fxi + off[i][0], ly = fyi + off[i][1];
        v4<T> pn = fetch<T>(prev.normals, W, H, lx, ly);
        T pd = fetch<T>(prev.depth, W, H, lx, ly).x;
        v3<T> pp = world_from_depth<T>(htx, hty, pd, vpi);
        T pm = fetch<T>(prev.color, W, H, lx, ly).w;
        v[i] = is_valid<T>(hx, hy, wpos, pp, n, octahedral_to_direction<T>(pn.x, pn.y), meshId, pm, W, H);
        valid = valid || v[i];
    }
    T col = 0;
    v2<T> mom{0, 0};
    if (valid) {
        T sumw = 0;
        T x = fract(fx), y = fract(fy);
        T w[4] = {(T(1) - x) * (T(1) - y), x * (T(1) - y), (T(1) - x) * y, x * y};
        for (int i = 0; i < 4; ++i) {
            int lx = fxi + off[i][0], ly = fyi + off[i][1];
            if (v[i]) {
                col += w[i] * fetch<T>(histOut, W, H, lx, ly).x;
                v4<T> m = fetch<T>(histMom, W, H, lx, ly);
                mom.x += w[i] * m.x; mom.y += w[i] * m.y;
                sumw += w[i];
            }
        }
        valid = sumw >= T(0.01f);
        if (valid) { col /= sumw; mom.x /= sumw; mom.y /= sumw; }
        else { col = 0; mom = {0, 0}; }
    }
    if (!valid) {
        T count = 0;
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) {
                int px = hx + xx, py = hy + yy;
                v4<T> pn = fetch<T>(prev.normals, W, H, px, py);
                T pd = fetch<T>(prev.depth, W, H, px, py).x;
                v3<T> pp = world_from_depth<T>(htx, hty, pd, vpi);
                T pm = fetch<T>(prev.color, W, H, px, py).w;
                if (is_valid<T>(hx, hy, wpos, pp, n, octahedral_to_direction<T>(pn.x, pn.y), meshId, pm, W, H)) {
                    col += fetch<T>(histOut, W, H, px, py).x;
                    v4<T> m = fetch<T>(histMom, W, H, px, py);
                    mom.x += m.x; mom.y += m.y;
                    count += T(1);
                }
            }
        if (count > T(0)) { valid = true; col /= count; mom.x /= count; mom.y /= count; }
    }
    if (valid) *hl = fetch<T>(histMom, W, H, hx, hy).z;
    else { col = 0; mom = {0, 0}; *hl = 0; }
    *hc = col;
    *hm = mom;
    return valid;
}

// ReflectionCommon.glsl:19-22
template <class T> inline T luminance(T r, T g, T b) {
    return std::max(r * T(0.299f) + g * T(0.587f) + b * T(0.114f), T(0.0001f));
}

// Reflection/DenoiseReprojection.comp:117-218 for one pixel; returns whether
// the pixel votes "denoise" for its tile (depth != 1 and roughness inside the band).
template <class T>
bool reflection_reproject_pixel(const LuxgiReflectionReprojectParams& pc, int W, int H, int x, int y, const GB& cur,
                                const GB& prev, Img noisy, Img histOut, Img histMom, T* outColor, T* outMom) {
    T depth = fetch<T>(cur.depth, W, H, x, y).x;
    T rough = fetch<T>(cur.pbr, W, H, x, y).y;
    T oc[4] = {0, 0, 0, 0}, om[4] = {0, 0, 0, 0};
    if (depth != T(1)) {
        v4<T> cr = fetch<T>(noisy, W, H, x, y);
        v3<T> hc;
        v2<T> hm;
        T hl;
        bool ok = reproject_rgb<T>(x, y, depth, pc.cameraPos, pc.viewProjInv, pc.prevViewProj, cr.w, cur, prev,
                                   histOut, histMom, W, H, &hc, &hm, &hl);
        hl = std::min(T(32.0f), ok ? hl + T(1) : T(1));
        if (ok) {  // neighborhoodStandardDeviation (Reflection/DenoiseReprojection.comp:92-114)
            v3<T> m1{0, 0, 0}, m2{0, 0, 0};
            for (int dx = -8; dx <= 8; ++dx)
                for (int dy = -8; dy <= 8; ++dy) {
                    v4<T> s = fetch<T>(noisy, W, H, x + dx, y + dy);
                    v3<T> c{s.x, s.y, s.z};
                    m1 = m1 + c;
                    m2 = m2 + c * c;
                }
            const T wgt = T(289);
            v3<T> mean = m1 * (T(1) / wgt);
            v3<T> var = m2 * (T(1) / wgt) - mean * mean;
            v3<T> sd{std::sqrt(std::max(var.x, T(0))), std::sqrt(std::max(var.y, T(0))), std::sqrt(std::max(var.z, T(0)))};
            v3<T> mn = mean - sd, mx = mean + sd;
            // clipAABB (Reflection/DenoiseReprojection.comp:78-89)
            v3<T> ctr = (mx + mn) * T(0.5f);
            v3<T> ext = (mx - mn) * T(0.5f) + v3<T>{T(0.001f), T(0.001f), T(0.001f)};
            v3<T> cv = hc - ctr;
            T ux = std::fabs(cv.x / ext.x), uy = std::fabs(cv.y / ext.y), uz = std::fabs(cv.z / ext.z);
            T mu = std::max(std::max(ux, uy), uz);
            if (mu > T(1)) hc = ctr + cv * (T(1) / mu);
        }
        T cd = std::sqrt(T(pc.cameraDelta[0]) * T(pc.cameraDelta[0]) + T(pc.cameraDelta[1]) * T(pc.cameraDelta[1]) +
                         T(pc.cameraDelta[2]) * T(pc.cameraDelta[2]));
        T maxAcc = cd > T(0) ? T(8) : hl;
        T alpha = ok ? std::max(T(pc.alpha), T(1) / maxAcc) : T(1);
        T alphaM = ok ? std::max(T(pc.momentsAlpha), T(1) / maxAcc) : T(1);
        T l = luminance(cr.x, cr.y, cr.z);
        T m0 = mixT(hm.x, l, alphaM), m1v = mixT(hm.y, l * l, alphaM);
        T var = std::max(T(0), m1v - m0 * m0);
        oc[0] = mixT(hc.x, cr.x, alpha); oc[1] = mixT(hc.y, cr.y, alpha); oc[2] = mixT(hc.z, cr.z, alpha); oc[3] = var;
        om[0] = m0; om[1] = m1v; om[2] = hl; om[3] = 0;
    }
    for (int i = 0; i < 4; ++i) { outColor[i] = oc[i]; outMom[i] = om[i]; }
    return depth != T(1) && !(rough <= T(0.05f) || rough >= T(0.45f));
}

// computeEdgeStoppingWeight (Raytraced/EdgeStop.glsl:24-40)
template <class T>
inline T edge_stop(T cd, T sd, T phiZ, v3<T> cn, v3<T> sn, T phiN, T cl, T sl, T phiL) {
    T wZ = std::exp(-std::fabs(cd - sd) / phiZ);
    T wN = std::pow(clampT(dot(cn, sn), T(0), T(1)), phiN);
    T wL = std::fabs(cl - sl) / phiL;
    return std::exp(T(0) - std::max(wL, T(0)) - std::max(wZ, T(0))) * wN;
}

// Reflection/DenoiseAtrous.comp:63-134 for one pixel of a denoise tile.
// `lumScale` multiplies the luminance term of every tap's weight (the learned per-pixel scale; 1 is
// the reference pass).
template <class T>
void reflection_atrous_pixel(const LuxgiReflectionAtrousParams& pc, int W, int H, int x, int y, Img normals, Img pbr,
                             Img in, T* out, T lumScale = T(1)) {
    const T kw[3] = {T(1.0), T(2.0 / 3.0), T(1.0 / 6.0)};
    v4<T> c = fetch<T>(in, W, H, x, y);
    T cl = luminance(c.x, c.y, c.z);
    const T kv[2][2] = {{T(1.0 / 4.0), T(1.0 / 8.0)}, {T(1.0 / 8.0), T(1.0 / 16.0)}};
    T var = 0;  // computeVarianceCenter (:41-61)
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) var += fetch<T>(in, W, H, x + xx, y + yy).w * kv[std::abs(xx)][std::abs(yy)];
    v4<T> nrm = fetch<T>(normals, W, H, x, y);
    v4<T> pb = fetch<T>(pbr, W, H, x, y);
    T cd = pb.w, rough = pb.y;
    if (cd == T(1)) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    if (rough <= T(0.05f) || rough >= T(0.45f)) { out[0] = c.x; out[1] = c.y; out[2] = c.z; out[3] = c.w; return; }
    T phiC = T(pc.phiColor) * std::sqrt(std::max(T(0), var)) + T(1e-10f);
    T sw = 1;
    T s[4] = {c.x, c.y, c.z, c.w};
    v3<T> cn = octahedral_to_direction<T>(nrm.x, nrm.y);
    for (int yy = -pc.radius; yy <= pc.radius; ++yy)
        for (int xx = -pc.radius; xx <= pc.radius; ++xx) {
            int px = x + xx * pc.stepSize, py = y + yy * pc.stepSize;
            bool inside = px >= 0 && py >= 0 && px < W && py < H;
            T k = kw[std::abs(xx)] * kw[std::abs(yy)];
            if (inside && (xx != 0 || yy != 0)) {
                v4<T> sc = fetch<T>(in, W, H, px, py);
                T sl = luminance(sc.x, sc.y, sc.z);
                v4<T> sn = fetch<T>(normals, W, H, px, py);
                T d = fetch<T>(pbr, W, H, px, py).w;
                T w = edge_stop<T>(cd, d, T(pc.sigmaDepth), cn, octahedral_to_direction<T>(sn.x, sn.y),
                                   T(pc.phiNormal), cl, sl, phiC / lumScale);
                T wv = w * k;
                sw += wv;
                s[0] += wv * sc.x; s[1] += wv * sc.y; s[2] += wv * sc.z; s[3] += wv * wv * sc.w;
            }
        }
    out[0] = s[0] / sw; out[1] = s[1] / sw; out[2] = s[2] / sw; out[3] = s[3] / (sw * sw);
}

// Shadow mask bit (Shadow/DenoiseReprojection.comp:81-91, robust outside the mask image).
inline float shadow_bit(const uint32_t* mask, int mw, int mh, int x, int y) {
    if (x < 0 || y < 0) return 0.f;
    int wx = x / 8, wy = y / 4;
    if (wx >= mw || wy >= mh) return 0.f;
    uint32_t word = mask[(size_t)wy * mw + wx];
    int bit = (y % 4) * 8 + (x % 8);
    return float((word >> bit) & 1u);
}

// Shadow/DenoiseReprojection.comp:127-210 for one pixel; returns the tile vote.
template <class T>
bool shadow_reproject_pixel(const LuxgiShadowReprojectParams& pc, int W, int H, int x, int y, const GB& cur,
                            const GB& prev, const uint32_t* mask, Img histOut, Img histMom, T* outColor, T* outMom) {
    const int mw = (W + 7) / 8, mh = (H + 3) / 4;
    T sum = 0;  // computeMean (:103-124): 17 horizontal sums of 17 taps
    for (int dy = -8; dy <= 8; ++dy)
        for (int dx = -8; dx <= 8; ++dx) sum += T(shadow_bit(mask, mw, mh, x + dx, y + dy));
    T mean = sum / T(289);
    T depth = fetch<T>(cur.depth, W, H, x, y).x;
    T vis = 0, hl = 0;
    T ov[2] = {0, 0}, om[2] = {0, 0};
    if (depth != T(1)) {
        vis = T(shadow_bit(mask, mw, mh, x, y));
        T hv;
        v2<T> hm;
        bool ok = reproject_scalar<T>(x, y, depth, pc.viewProjInv, cur, prev, histOut, histMom, W, H, &hv, &hm, &hl);
        hl = std::min(T(32.0f), ok ? hl + T(1) : T(1));
        if (ok) {
            T sv = std::max(mean - mean * mean, T(0));
            T sd = std::sqrt(sv);
            hv = clampT(hv, mean - T(0.5f) * sd, mean + T(0.5f) * sd);
        }
        T alpha = ok ? std::max(T(pc.alpha), T(1) / hl) : T(1);
        T alphaM = ok ? std::max(T(pc.momentsAlpha), T(1) / hl) : T(1);
        om[0] = mixT(hm.x, vis, alphaM);
        om[1] = mixT(hm.y, vis * vis, alphaM);
        ov[0] = mixT(hv, vis, alpha);
        ov[1] = std::max(T(0), om[1] - om[0] * om[0]);
    }
    outMom[0] = om[0]; outMom[1] = om[1]; outMom[2] = hl; outMom[3] = 0;
    outColor[0] = ov[0]; outColor[1] = ov[1];
    return depth != T(1) && ov[0] > T(0);
}

// Shadow/DenoiseAtrous.comp:65-128 for one pixel of a denoise tile.
template <class T>
void shadow_atrous_pixel(const LuxgiShadowAtrousParams& pc, int W, int H, int x, int y, Img normals, Img pbr, Img in,
                         T* out) {
    const T kw[3] = {T(1.0), T(2.0 / 3.0), T(1.0 / 6.0)};
    v4<T> c = fetch<T>(in, W, H, x, y);
    const T kv[2][2] = {{T(1.0 / 4.0), T(1.0 / 8.0)}, {T(1.0 / 8.0), T(1.0 / 16.0)}};
    T var = 0;
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) var += fetch<T>(in, W, H, x + xx, y + yy).y * kv[std::abs(xx)][std::abs(yy)];
    v4<T> nrm = fetch<T>(normals, W, H, x, y);
    T cd = fetch<T>(pbr, W, H, x, y).w;
    if (cd < T(0)) { out[0] = c.x; out[1] = c.y; return; }
    T phiV = T(pc.phiVisibility) * std::sqrt(std::max(T(0), T(1e-10f) + var));
    T sw = 1;
    T s0 = c.x, s1 = c.y;
    v3<T> cn = octahedral_to_direction<T>(nrm.x, nrm.y);
    for (int yy = -pc.radius; yy <= pc.radius; ++yy)
        for (int xx = -pc.radius; xx <= pc.radius; ++xx) {
            int px = x + xx * pc.stepSize, py = y + yy * pc.stepSize;
            bool inside = px >= 0 && py >= 0 && px < W && py < H;
            T k = kw[std::abs(xx)] * kw[std::abs(yy)];
            if (inside && (xx != 0 || yy != 0)) {
                v4<T> sv = fetch<T>(in, W, H, px, py);
                v4<T> sn = fetch<T>(normals, W, H, px, py);
                T d = fetch<T>(pbr, W, H, px, py).w;
                T w = edge_stop<T>(cd, d, T(pc.sigmaDepth), cn, octahedral_to_direction<T>(sn.x, sn.y),
                                   T(pc.phiNormal), c.x, sv.x, phiV);
                T wv = w * k;
                sw += wv;
                s0 += wv * sv.x;
                s1 += wv * wv * sv.y;
            }
        }
    out[0] = s0 / sw;
    out[1] = s1 / (sw * sw);
    if (pc.power != 0.0f) out[0] = std::pow(out[0], T(pc.power));
}

// ---------------------------------------------------------------------------
// Probe relocation + classification. No counterpart in the reference; the rule
// (after Majercik et al., JCGT 10(2) 2021) is restated from libluxgi's
// relocate.hip header: comparisons are made on the fp32 ray values exactly as
// the GPU makes them, so states are comparable bit for bit.
// dirDist float4 [P][width], flags uint8 [P][width], offsets float4 [P] (in/out).
// ---------------------------------------------------------------------------
inline void relocate_probe(const LuxgiDDGIUniform& u, const float* dirDist, const uint8_t* flags, int width,
                           const LuxgiRelocateParams& prm, float* offsets, uint8_t* states, int p) {
    const int R = u.raysPerProbe;
    int inside = 0;
    bool near = false;
    float best = INFINITY;
    int bestRay = -1;
    for (int r = 0; r < R; ++r) {
        size_t e = (size_t)p * width + r;
        if (flags[e] == 2) inside++;
        else if (flags[e] == 1) {
            const float* d = dirDist + e * 4;
            float t = d[3];
            float ax = std::fabs(float(d[0] * t)), ay = std::fabs(float(d[1] * t)), az = std::fabs(float(d[2] * t));
            if (ax <= u.step[0] && ay <= u.step[1] && az <= u.step[2]) near = true;
            if (t < best) { best = t; bestRay = r; }
        }
    }
    const bool isInside = float(inside) > prm.backfaceThreshold * float(R);
    states[p] = (!isInside && near) ? 0 : 1;
    if (prm.relocate && !isInside && bestRay >= 0 && best < prm.minFrontfaceDistance) {
        const float* d = dirDist + ((size_t)p * width + bestRay) * 4;
        const float push = prm.minFrontfaceDistance - best;
        float* o = offsets + (size_t)p * 4;
        for (int i = 0; i < 3; ++i) {
            float v = o[i] - d[i] * push;
            float lim = prm.maxOffsetFraction * u.step[i];
            o[i] = std::min(std::max(v, -lim), lim);
        }
        o[3] = 0.f;
    }
}

// ---------------------------------------------------------------------------
// Probe sampling: SampleProbe.comp:36-60 with sampleIrradiance (DDGICommon.glsl:163-233).
// The atlases are read through the engine's default sampler (RHI/Definitions.h:152-159):
// linear filtering, repeat addressing, normalized coordinates.
// ---------------------------------------------------------------------------
// texture(sampler2D, uv): bilinear, repeat, on a [H][W][ch] float image.
template <class T> inline void sample2d_repeat(const float* img, int W, int H, int ch, T u, T v, T* out) {
    const T x = u * T(W) - T(0.5), y = v * T(H) - T(0.5);
    const T fx = std::floor(x), fy = std::floor(y);
    const T ax = x - fx, ay = y - fy;
    auto wrap = [](long long i, int n) { long long m = i % n; return int(m < 0 ? m + n : m); };
    const int x0 = wrap((long long)fx, W), x1 = wrap((long long)fx + 1, W);
    const int y0 = wrap((long long)fy, H), y1 = wrap((long long)fy + 1, H);
    for (int c = 0; c < ch; ++c) {
        const T t00 = img[((size_t)y0 * W + x0) * ch + c], t10 = img[((size_t)y0 * W + x1) * ch + c];
        const T t01 = img[((size_t)y1 * W + x0) * ch + c], t11 = img[((size_t)y1 * W + x1) * ch + c];
        out[c] = (t00 * (T(1) - ax) + t10 * ax) * (T(1) - ay) + (t01 * (T(1) - ax) + t11 * ax) * ay;
    }
}

// textureCoordFromDirection (DDGICommon.glsl:141-158)
template <class T> inline v2<T> texture_coord_from_direction(v3<T> dir, int probeIndex, int width, int height,
                                                             int side) {
    const v2<T> oc = oct_encode<T>(normalize(dir));
    const T ox = (oc.x + T(1)) * T(0.5), oy = (oc.y + T(1)) * T(0.5);
    const T pws = T(side) + T(2);
    const int perRow = (width - 2) / int(pws);
    const T tlx = T(probeIndex % perRow) * pws + T(2), tly = T(probeIndex / perRow) * pws + T(2);
    return {tlx / T(width) + (ox * T(side)) / T(width), tly / T(height) + (oy * T(side)) / T(height)};
}

// sampleIrradiance (DDGICommon.glsl:163-233); irr = float4 atlas, dep = float2 atlas.
// offsets (float4 per probe) / states (uint8 per probe, 1 = inactive), both optional, extend it with
// relocation and classification as the published irradiance gather does (Majercik et al., JCGT
// 2021; the reference has neither): neighbour probes at grid position + offset, trilinear weights
// on the grid, inactive neighbours skipped, 0 when no neighbour carries weight.
template <class T>
inline v3<T> sample_irradiance(const LuxgiDDGIUniform& u, v3<T> P, v3<T> N, v3<T> Wo, const float* irr,
                               const float* dep, const float* offsets = nullptr, const uint8_t* states = nullptr) {
    const v3<T> start{T(u.startPosition[0]), T(u.startPosition[1]), T(u.startPosition[2])};
    const v3<T> stp{T(u.step[0]), T(u.step[1]), T(u.step[2])};
    int base[3];
    const T rel[3] = {(P.x - start.x) / stp.x, (P.y - start.y) / stp.y, (P.z - start.z) / stp.z};
    for (int i = 0; i < 3; ++i) base[i] = std::min(std::max((int)rel[i], 0), u.probeCounts[i] - 1);  // baseGridCoord :125-128
    const v3<T> basePos{stp.x * T(base[0]) + start.x, stp.y * T(base[1]) + start.y, stp.z * T(base[2]) + start.z};
    const v3<T> alpha{clampT((P.x - basePos.x) / stp.x, T(0), T(1)), clampT((P.y - basePos.y) / stp.y, T(0), T(1)),
                      clampT((P.z - basePos.z) / stp.z, T(0), T(1))};
    v3<T> sum{0, 0, 0};
    T sumW = 0;
    for (int i = 0; i < 8; ++i) {
        const int off[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
        int g[3];
        for (int k = 0; k < 3; ++k) g[k] = std::min(std::max(base[k] + off[k], 0), u.probeCounts[k] - 1);
        const int probeIdx = grid_to_probe_index(u, g[0], g[1], g[2]);
        if (states && states[probeIdx] != 0) continue;
        v3<T> probePos{stp.x * T(g[0]) + start.x, stp.y * T(g[1]) + start.y, stp.z * T(g[2]) + start.z};
        if (offsets) {
            const float* o = offsets + (size_t)probeIdx * 4;
            probePos = probePos + v3<T>{T(o[0]), T(o[1]), T(o[2])};
        }
        const v3<T> tri{off[0] ? alpha.x : T(1) - alpha.x, off[1] ? alpha.y : T(1) - alpha.y,
                        off[2] ? alpha.z : T(1) - alpha.z};
        T w = 1;
        const v3<T> dirToProbe = normalize(probePos - P);
        const T sb = gmax(T(0.0001), (dot(dirToProbe, N) + T(1)) * T(0.5));
        w *= sb * sb + T(0.2);
        const v3<T> vBias = (N + Wo * T(3)) * T(u.normalBias);
        const v3<T> probeToPoint = P - probePos + vBias;
        const v3<T> dir = normalize(probeToPoint * T(-1));
        v2<T> tc = texture_coord_from_direction<T>(dir * T(-1), probeIdx, u.depthTextureWidth, u.depthTextureHeight,
                                                   u.depthProbeSideLength);
        const T dist = length(probeToPoint);
        T dm[2];
        sample2d_repeat<T>(dep, u.depthTextureWidth, u.depthTextureHeight, 2, tc.x, tc.y, dm);
        const T mean = dm[0];
        const T variance = std::fabs(dm[0] * dm[0] - dm[1]);
        const T md = gmax(dist - mean, T(0));
        T cheb = variance / (variance + md * md);
        cheb = gmax(cheb * cheb * cheb, T(0));
        w *= (dist <= mean) ? T(1) : cheb;
        w = gmax(T(0.000001), w);
        tc = texture_coord_from_direction<T>(normalize(N), probeIdx, u.irradianceTextureWidth,
                                             u.irradianceTextureHeight, u.irradianceProbeSideLength);
        T ir[4];
        sample2d_repeat<T>(irr, u.irradianceTextureWidth, u.irradianceTextureHeight, 4, tc.x, tc.y, ir);
        const T ge = T(u.ddgiGamma * 0.5f);
        const v3<T> pi{std::pow(ir[0], ge), std::pow(ir[1], ge), std::pow(ir[2], ge)};
        const T crush = T(0.2f);
        if (w < crush) w *= w * w * (T(1) / (crush * crush));
        w *= tri.x * tri.y * tri.z;
        sum = sum + pi * w;
        sumW += w;
    }
    if (sumW == T(0)) return {0, 0, 0};
    v3<T> net = sum * (T(1) / sumW);
    net = net * net;
    return net * T(2 * kPi);
}

// SampleProbe.comp:36-60 for pixel (x, y): out = (irradiance, 1), or 0 on sky (depth == 1).
template <class T>
inline void sample_probe_pixel(const LuxgiDDGIUniform& u, const LuxgiSampleProbeParams& pc, int W, int H, int x,
                               int y, const float* depth, const float* normals, const float* irr, const float* dep,
                               T* out, const float* offsets = nullptr, const uint8_t* states = nullptr) {
    const T d = T(depth[(size_t)y * W + x]);
    if (d == T(1)) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    const T tu = (T(x) + T(0.5)) / T(W), tv = (T(y) + T(0.5)) / T(H);
    const v3<T> P = world_from_depth<T>(tu, tv, d, pc.viewProjInv);
    const float* nn = normals + ((size_t)y * W + x) * 4;
    const v3<T> N = octahedral_to_direction<T>(T(nn[0]), T(nn[1]));
    const v3<T> cam{T(pc.cameraPosition[0]), T(pc.cameraPosition[1]), T(pc.cameraPosition[2])};
    const v3<T> Wo = normalize(cam - P);
    const v3<T> e = sample_irradiance<T>(u, P, N, Wo, irr, dep, offsets, states);
    out[0] = e.x; out[1] = e.y; out[2] = e.z; out[3] = T(1);
}

// ---------------------------------------------------------------------------
// threading helper: split [0, n) over `threads` std::threads
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Surface-atlas chunk culling (Shaders/SDF/SDFCulling.comp:36-101)
// ---------------------------------------------------------------------------
// boxIntersectsSphere (AtlasCommon.glsl:34-38) in fp32, operations in the written order:
// chunkMin = (vec3(chunkCoord) - 20) * chunkSize, chunkMax = chunkMin + chunkSize
// (SDFCulling.comp:42-43), distance(centre, clamp(centre, min, max)) <= radius.
inline bool chunk_touches_sphere(float cs, int cx, int cy, int cz, const float* s) {
    const int c[3] = {cx, cy, cz};
    float d2 = 0.0f;
    for (int a = 0; a < 3; ++a) {
        volatile float mn = (float(c[a]) - 20.0f) * cs;  // volatile: one rounding per operation
        volatile float mx = mn + cs;
        volatile float d = s[a] - std::min(std::max(s[a], (float)mn), (float)mx);
        volatile float sq = d * d;
        d2 = d2 + sq;
    }
    return std::sqrt(d2) <= s[3];
}

// Lists in chunk-address order: chunks[addr] = offset of [count, ids...] in culled, 0 when the
// chunk touches no object or its run would end past `capacity`; culled[0] = 1 + sum over the
// non-empty chunks of (count + 1), the counter value SDFCulling.comp:64-65 leaves.
inline void cull_chunks(float chunkSize, const float* bounds, int stride, int n, uint32_t capacity, uint32_t* chunks,
                        uint32_t* culled) {
    uint32_t cursor = 1;
    std::vector<uint32_t> ids;
    for (int addr = 0; addr < 40 * 40 * 40; ++addr) {
        const int cx = addr % 40, cy = (addr / 40) % 40, cz = addr / 1600;
        ids.clear();
        for (int i = 0; i < n; ++i)
            if (chunk_touches_sphere(chunkSize, cx, cy, cz, bounds + (size_t)i * stride)) ids.push_back((uint32_t)i);
        chunks[addr] = 0;
        if (ids.empty()) continue;
        const uint32_t start = cursor, size = (uint32_t)ids.size() + 1;
        cursor += size;
        if ((unsigned long long)start + size > capacity) continue;
        chunks[addr] = start;
        culled[start] = (uint32_t)ids.size();
        for (size_t k = 0; k < ids.size(); ++k) culled[start + 1 + k] = ids[k];
    }
    culled[0] = cursor;
}

// ---------------------------------------------------------------------------
// Surface-cache lighting: CopyEmissive.frag:22-27, SDFDeferredLight.frag:41-130,
// SDFAtlasIndirectLight.frag:44-68; one texel of one tile draw. The reference
// draws each tile as a quad over its full rect (GlobalSurfaceAtlas.cpp:554-574)
// while the tile UV maps onto the rect minus the padding (extends.zw,
// GlobalSurfaceAtlas.cpp:272-276), so the G-buffer is sampled with bilinear
// filtering at atlasUV = ((p + 0.5 - X) / Wt) * extends.zw + extends.xy.
// Passes blend: emissive copy (blend none), + direct per light, + indirect
// (GlobalSurfaceAtlas.cpp:886-1085).
// ---------------------------------------------------------------------------
// Raytraced/BRDF.glsl:12-79, the 7-argument BRDF
template <class T>
inline v3<T> brdf_ggx(v3<T> albedo, v3<T> n, T rough, T metal, v3<T> view, v3<T> h, v3<T> l) {
    const T eps = T(0.00001f);  // BRDF.glsl:9
    const v3<T> F0{mixT(T(0.04f), albedo.x, metal), mixT(T(0.04f), albedo.y, metal), mixT(T(0.04f), albedo.z, metal)};
    const T cosLi = std::max(T(0), dot(n, l)), cosLh = std::max(T(0), dot(n, h)), NdotV = std::max(T(0), dot(n, view));
    const T c = clampT(T(1) - std::max(dot(h, view), T(0)), T(0), T(1));
    const T c5 = c * c * c * c * c;
    const v3<T> F{F0.x + (T(1) - F0.x) * c5, F0.y + (T(1) - F0.y) * c5, F0.z + (T(1) - F0.z) * c5};
    const T alpha = rough * rough, a2 = alpha * alpha;
    const T denom = cosLh * cosLh * (a2 - T(1)) + T(1);
    const T D = a2 / (T(kPi) * denom * denom);
    const T r = rough + T(1), k = r * r / T(8);
    const T G = cosLi / (cosLi * (T(1) - k) + k) * (NdotV / (NdotV * (T(1) - k) + k));
    const T spec = D * G / std::max(eps, T(4) * cosLi * NdotV);
    const T km = T(1) - metal;
    return {(T(1) - F.x) * km * albedo.x / T(kPi) + F.x * spec, (T(1) - F.y) * km * albedo.y / T(kPi) + F.y * spec,
            (T(1) - F.z) * km * albedo.z / T(kPi) + F.z * spec};
}

// fetchLight (SDFDeferredLight.frag:41-76); Light.glsl:13-15 types 0 directional, 1 spot, 2 point
template <class T>
inline void fetch_light(const LuxgiLight& L, v3<T> wp, v3<T>& Wi, T& tmax, T& atten) {
    Wi = {0, 0, 0};
    tmax = T(kGlobalSDFWorldSize);
    atten = T(1);
    const v3<T> lp{T(L.position[0]), T(L.position[1]), T(L.position[2])};
    const v3<T> ld{T(L.direction[0]), T(L.direction[1]), T(L.direction[2])};
    if (L.type == 0.0f) {
        Wi = {-ld.x, -ld.y, -ld.z};
    } else if (L.type == 2.0f) {
        const v3<T> dv = lp - wp;
        const T dist = length(dv);
        Wi = normalize(dv);
        atten = T(L.radius) / (dist * dist + T(1));
        tmax = dist;
    } else if (L.type == 1.0f) {
        const v3<T> dv = lp - wp;
        const T cutoff = T(1) - T(L.angle);
        Wi = normalize(dv);
        const T dist = length(dv);
        const T theta = dot(Wi, ld);
        const T e = cutoff - cutoff * T(0.9f);
        atten = (theta - cutoff) / e;
        atten *= T(L.radius) / (dist * dist + T(1));
        atten = clampT(atten, T(0), T(1));
        tmax = dist;
    }
}

// One texel (px, py) of a tile draw. gb*: float RGBA [res][res] images (the G-buffer texels
// already converted from their storage); sdf/mip: the LINEAR FP32 global SDF; irr/dep: the probe
// atlases (float RGBA / RG), or null for no indirect pass.
template <class T>
inline void surface_light_texel(const LuxgiSurfaceLightParams& prm, const LuxgiLight* lights, uint32_t lightMask,
                                const LuxgiAtlasTile& tile, int res, int px, int py, const float* gbColor,
                                const float* gbNormal, const float* gbPbr, const float* gbEmissive,
                                const LuxgiGlobalSDFData& g, const float* sdf, const float* mip, int sdfRes,
                                int mipRes, const LuxgiDDGIUniform* u, const float* irr, const float* dep, T out[4],
                                const float* offsets = nullptr, const uint8_t* states = nullptr) {
    const int X = (int)std::lrint(tile.extends[0] * (float)res), Y = (int)std::lrint(tile.extends[1] * (float)res);
    const int Wt = (int)std::lrint(tile.extends[2] * (float)res) + 1;
    const int Ht = (int)std::lrint(tile.extends[3] * (float)res) + 1;
    const T au = ((T(px) + T(0.5) - T(X)) / T(Wt)) * T(tile.extends[2]) + T(tile.extends[0]);
    const T av = ((T(py) + T(0.5) - T(Y)) / T(Ht)) * T(tile.extends[3]) + T(tile.extends[1]);
    T acc[3] = {0, 0, 0};
    if (prm.passes & LUXGI_SURFACE_EMISSIVE) {
        T e[4];
        sample2d_repeat<T>(gbEmissive, res, res, 4, au, av, e);
        acc[0] = e[0]; acc[1] = e[1]; acc[2] = e[2];
    }
    const int nl = (prm.passes & LUXGI_SURFACE_DIRECT) ? prm.numLights : 0;
    const bool indirect = (prm.passes & LUXGI_SURFACE_INDIRECT) && u && irr && dep;
    if (nl > 0 || indirect) {
        T c[4], nt[4], pbr[4];
        sample2d_repeat<T>(gbColor, res, res, 4, au, av, c);
        sample2d_repeat<T>(gbNormal, res, res, 4, au, av, nt);
        sample2d_repeat<T>(gbPbr, res, res, 4, au, av, pbr);
        const v3<T> n = octahedral_to_direction<T>(nt[0], nt[1]);
        const v3<T> wp{c[3], nt[2], nt[3]};
        const v3<T> cam{T(prm.cameraPosition[0]), T(prm.cameraPosition[1]), T(prm.cameraPosition[2])};
        const T sb = T(prm.cameraPosition[3]);
        const v3<T> view = normalize(cam - wp);
        const int steps = prm.maxSteps > 0 ? prm.maxSteps : 250;
        for (int l = 0; l < nl; ++l) {
            if (!((lightMask >> l) & 1u)) continue;
            v3<T> L;
            T dist, atten;
            fetch_light<T>(lights[l], wp, L, dist, atten);
            const T NoL = dot(n, L);
            const T bias = T(2) * sb * clampT(T(1) - NoL, T(0), T(1)) + sb;
            T shadow = T(1);
            if (NoL > T(0)) {
                if (atten > T(0)) {
                    const Hit<T> h = trace_global_sdf<T>(g, sdf, mip, sdfRes, mipRes, wp + n * sb, L, dist - bias,
                                                         steps, T(0), nullptr, T(2), false);
                    shadow = h.time >= T(0) ? T(0) : T(1);
                }
            } else {
                shadow = T(0);
            }
            const T inten = std::pow(T(lights[l].intensity), T(1.4f)) + T(0.1f);
            const v3<T> Lh = normalize(L + view);
            const T cosLi = std::max(T(0), dot(n, L));
            const v3<T> b = brdf_ggx<T>({c[0], c[1], c[2]}, n, pbr[1], pbr[0], view, Lh, L);
            const T s = cosLi * shadow * atten;
            acc[0] += b.x * T(lights[l].color[0]) * inten * s;
            acc[1] += b.y * T(lights[l].color[1]) * inten * s;
            acc[2] += b.z * T(lights[l].color[2]) * inten * s;
        }
        if (indirect) {
            const T m = pbr[0];
            const T al[3] = {std::min(c[0], T(0.9f)), std::min(c[1], T(0.9f)), std::min(c[2], T(0.9f))};
            const v3<T> ir = sample_irradiance<T>(*u, wp, n, view, irr, dep, offsets, states);
            const T I = T(prm.indirectIntensity);
            acc[0] += I * ((al[0] - al[0] * m) / T(kPi)) * ir.x;
            acc[1] += I * ((al[1] - al[1] * m) / T(kPi)) * ir.y;
            acc[2] += I * ((al[2] - al[2] * m) / T(kPi)) * ir.z;
        }
    }
    out[0] = acc[0]; out[1] = acc[1]; out[2] = acc[2]; out[3] = T(1);
}


// ---------------------------------------------------------------------------
// Noisy-signal producers: SDF soft-shadow mask (Shaders/SDF/SDFShadow.comp:36-157)
// ---------------------------------------------------------------------------
// sampleBlueNoise (Raytraced/BlueNoise.glsl:8-18) on RGBA8 tables: sobol [256], scrambling [128][128].
// A UNORM8 texel v reads as v/255; int(clamp(v/255 * 256, 0, 255)) = v.
inline float blue_noise(const uint8_t* sobol, const uint8_t* scr, int x, int y, int frame, int dim) {
    x %= 128; y %= 128; frame %= 256; dim %= 4;
    auto byte = [](uint8_t v) { return (int)std::min(std::max((float)v / 255.0f * 256.0f, 0.0f), 255.0f); };
    const uint8_t* s = scr + ((size_t)y * 128 + x) * 4;
    const int ranked = frame ^ byte(s[2]);
    int value = byte(sobol[(size_t)ranked * 4 + dim]);
    value = value ^ byte(s[dim % 2]);
    return (0.5f + (float)value) / 256.0f;
}

// fetchLight with softShadow = true (SDFShadow.comp:43-109)
template <class T>
inline void fetch_light_soft(const LuxgiLight& L, v3<T> wp, v3<T> n, T rx, T ry, v3<T>& Wi, T& tmax, T& atten) {
    v3<T> ld{0, 0, 0};
    T radius = 0;
    tmax = 0;
    atten = 0;
    const v3<T> lp{T(L.position[0]), T(L.position[1]), T(L.position[2])};
    if (L.type == 0.0f) {
        ld = {-T(L.direction[0]), -T(L.direction[1]), -T(L.direction[2])};
        tmax = T(kGlobalSDFWorldSize);
        radius = T(L.direction[3]);
        atten = T(1);
    } else if (L.type == 2.0f) {
        const v3<T> dv = lp - wp;
        const T dist = length(dv);
        ld = normalize(dv);
        atten = T(1);
        tmax = dist;
        radius = T(L.direction[3]) / dist;
    } else if (L.type == 1.0f) {
        const v3<T> dv = lp - wp;
        const T cutoff = T(1) - T(L.angle);
        ld = normalize(dv);
        const T dist = length(dv);
        const T theta = dot(ld, v3<T>{T(L.direction[0]), T(L.direction[1]), T(L.direction[2])});
        const T e = cutoff - cutoff * T(0.9f);
        atten = (theta - cutoff) / e;
        atten *= T(L.radius) / (dist * dist + T(1));
        atten = clampT(atten, T(0), T(1));
        tmax = dist;
        radius = T(L.direction[3]) / dist;
    }
    const v3<T> t = normalize(v3<T>{-ld.z, T(0), ld.x});  // cross(lightDir, (0, 1, 0))
    const v3<T> b = normalize(v3<T>{t.y * ld.z - t.z * ld.y, t.z * ld.x - t.x * ld.z, t.x * ld.y - t.y * ld.x});
    const T pr = radius * std::sqrt(rx);
    const T pa = ry * T(2) * T(kPi);
    Wi = normalize(ld + t * (pr * std::cos(pa)) + b * (pr * std::sin(pa)));
    atten *= clampT(dot(n, Wi), T(0), T(1));
}

// Visibility bit of pixel (x, y) (x, y may pass the image edge: the G-buffer is read with repeat
// addressing at the wrapped pixel, the world position comes from the unwrapped screen position).
// depth R32F [H][W], normals RGBA32F [H][W] (octahedral normal in xy). Returns 0/1; *sky = depth == 1.
template <class T>
inline uint32_t shadow_visibility(const LuxgiSdfShadowParams& p, const uint8_t* sobol, const uint8_t* scr, int W, int H,
                                  int x, int y, const float* depth, const float* normals, const LuxgiGlobalSDFData& g,
                                  const float* sdf, const float* mip, int res, int mipRes, bool* sky) {
    const size_t e = (size_t)(y % H) * W + (x % W);
    const float d = depth[e];
    *sky = d == 1.0f;
    if (*sky) return 0;
    const T tu = (T(x) + T(0.5)) / T(W), tv = (T(y) + T(0.5)) / T(H);
    const v3<T> wp = world_from_depth<T>(tu, tv, T(d), p.viewProjInv);
    const v3<T> n = octahedral_to_direction<T>(T(normals[e * 4]), T(normals[e * 4 + 1]));
    const T rx = T(blue_noise(sobol, scr, x, y, (int)p.numFrames, 0));
    const T ry = T(blue_noise(sobol, scr, x, y, (int)p.numFrames, 1));
    v3<T> Wi;
    T tmax, atten;
    fetch_light_soft<T>(p.light, wp, n, rx, ry, Wi, tmax, atten);
    const T NoL = dot(n, Wi);
    if (!(NoL > T(0)) || !(atten > T(0))) return 0;
    const T sb = T(p.shadowBias);
    const T bias = T(2) * sb * clampT(T(1) - NoL, T(0), T(1)) + sb;
    const Hit<T> h = trace_global_sdf<T>(g, sdf, mip, res, mipRes, wp + n * sb, Wi, tmax - bias,
                                         p.maxSteps > 0 ? p.maxSteps : 250, T(0), nullptr, T(0), false);
    return h.time >= T(0) ? 0u : 1u;
}

// ---------------------------------------------------------------------------
// Noisy-signal producers: SDF reflection (Shaders/SDF/SDFReflection.comp:78-162)
// ---------------------------------------------------------------------------
// texture(samplerCube, d) with the alpha channel (the sky read of trace(), SDFReflection.comp:113)
template <class T> inline v4<T> sample_cube4(const float* faces, int size, v3<T> d) {
    T ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int face;
    T sc, tc, ma;
    if (ax >= ay && ax >= az) { ma = ax; face = d.x >= 0 ? 0 : 1; sc = d.x >= 0 ? -d.z : d.z; tc = -d.y; }
    else if (ay >= az) { ma = ay; face = d.y >= 0 ? 2 : 3; sc = d.x; tc = d.y >= 0 ? d.z : -d.z; }
    else { ma = az; face = d.z >= 0 ? 4 : 5; sc = d.z >= 0 ? d.x : -d.x; tc = -d.y; }
    T s = T(0.5) * (sc / ma + T(1)), t = T(0.5) * (tc / ma + T(1));
    T x = s * T(size) - T(0.5), y = t * T(size) - T(0.5);
    T fx0 = std::floor(x), fy0 = std::floor(y);
    T a = x - fx0, b = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0;
    int x1 = std::min(std::max(x0 + 1, 0), size - 1), y1 = std::min(std::max(y0 + 1, 0), size - 1);
    x0 = std::min(std::max(x0, 0), size - 1); y0 = std::min(std::max(y0, 0), size - 1);
    const float* f = faces + (size_t)face * size * size * 4;
    T r[4];
    for (int c = 0; c < 4; ++c) {
        T t00 = T(f[((size_t)y0 * size + x0) * 4 + c]), t10 = T(f[((size_t)y0 * size + x1) * 4 + c]);
        T t01 = T(f[((size_t)y1 * size + x0) * 4 + c]), t11 = T(f[((size_t)y1 * size + x1) * 4 + c]);
        r[c] = (t00 * (T(1) - a) + t10 * a) * (T(1) - b) + (t01 * (T(1) - a) + t11 * a) * b;
    }
    return {r[0], r[1], r[2], r[3]};
}

// importanceSampleGGX (Raytraced/BRDF.glsl:176-200), the half vector
template <class T> inline v3<T> importance_sample_ggx(T ex, T ey, v3<T> N, T roughness) {
    const T a = roughness * roughness, m2 = a * a;
    const T phi = T(2) * T(kPi) * ex;
    const T cosT = std::sqrt((T(1) - ey) / (T(1) + (m2 - T(1)) * ey));
    const T sinT = std::sqrt(T(1) - cosT * cosT);
    const v3<T> Hl{std::cos(phi) * sinT, std::sin(phi) * sinT, cosT};
    const v3<T> up = std::fabs(N.z) < T(0.999f) ? v3<T>{0, 0, 1} : v3<T>{1, 0, 0};
    const v3<T> t = normalize(v3<T>{up.y * N.z - up.z * N.y, up.z * N.x - up.x * N.z, up.x * N.y - up.y * N.x});
    const v3<T> b{N.y * t.z - N.z * t.y, N.z * t.x - N.x * t.z, N.x * t.y - N.y * t.x};
    return normalize(t * Hl.x + b * Hl.y + N * Hl.z);
}

template <class T> inline v3<T> reflect3(v3<T> I, v3<T> N) { return I - N * (T(2) * dot(N, I)); }

struct ReflectionInputs {
    const LuxgiSdfReflectionParams* p;
    const uint8_t *sobol, *scr;
    int W, H;
    const float *normals, *pbr, *depth;       // G-buffer (RGBA32F normals and PBR, R32F depth)
    const LuxgiGlobalSDFData* g;
    const float *sdf, *mip;
    int res, mipRes;
    const LuxgiSurfaceAtlasData* ad;
    const uint32_t *chunks, *culled;
    const LuxgiAtlasObject* objects;
    const LuxgiAtlasTile* tiles;
    const float *atlasDepth, *atlasColor;   // float [res][res] and RGBA [res][res][4]
    const float* skyFaces;
    int skySize;
    const LuxgiDDGIUniform* u;
    const float *irr, *dep;
    const float* offsets = nullptr;     // optional relocation / classification of the probes
    const uint8_t* states = nullptr;
};

// One pixel; returns false for sky (the texel is not written).
template <class T>
inline bool reflection_pixel(const ReflectionInputs& in, int x, int y, T out[4]) {
    const LuxgiSdfReflectionParams& p = *in.p;
    const size_t e = (size_t)y * in.W + x;
    const float d = in.depth[e];
    if (d == 1.0f) return false;
    const T tu = (T(x) + T(0.5)) / T(in.W), tv = (T(y) + T(0.5)) / T(in.H);
    v3<T> wp = world_from_depth<T>(tu, tv, T(d), p.viewProjInv);
    const T roughness = T(in.pbr[e * 4 + 1]);
    const v3<T> n = octahedral_to_direction<T>(T(in.normals[e * 4]), T(in.normals[e * 4 + 1]));
    const v3<T> cam{T(p.cameraPosition[0]), T(p.cameraPosition[1]), T(p.cameraPosition[2])};
    const v3<T> Wo = normalize(cam - wp);
    wp = wp + n * T(in.g->cascadeVoxelSize[0]);
    v4<T> col{0, 0, 0, 0};
    v3<T> R;
    bool traceRay = true;
    const v3<T> mWo{-Wo.x, -Wo.y, -Wo.z};
    if (roughness < T(0.05f)) {
        R = reflect3<T>(mWo, n);
    } else if (roughness > T(0.45f) && p.approximateWithDDGI == 1u) {
        R = reflect3<T>(mWo, n);
        traceRay = false;
        const v3<T> ir = sample_irradiance<T>(*in.u, wp, R, Wo, in.irr, in.dep, in.offsets, in.states);
        col = {T(p.roughDDGIIntensity) * ir.x, T(p.roughDDGIIntensity) * ir.y, T(p.roughDDGIIntensity) * ir.z, T(0)};
    } else {
        const T ex = T(blue_noise(in.sobol, in.scr, x, y, (int)p.numFrames, 0)) * T(p.trim);
        const T ey = T(blue_noise(in.sobol, in.scr, x, y, (int)p.numFrames, 1)) * T(p.trim);
        R = reflect3<T>(mWo, importance_sample_ggx<T>(ex, ey, n, roughness));
    }
    if (traceRay) {
        const int steps = p.maxSteps > 0 ? p.maxSteps : 250;
        const Hit<T> h = trace_global_sdf<T>(*in.g, in.sdf, in.mip, in.res, in.mipRes, wp, R, T(kGlobalSDFWorldSize),
                                             steps, T(0), nullptr, T(0), false);
        if (h.time >= T(0)) {
            const T thr = T(in.g->cascadeVoxelSize[h.cascade]) * T(1.05f);
            col = sample_surface_atlas<T>(*in.ad, in.chunks, in.culled, in.objects, in.tiles, in.atlasDepth,
                                          in.atlasColor, wp + R * h.time, v3<T>{-R.x, -R.y, -R.z}, thr, nullptr);
        } else if (in.skyFaces) {
            col = sample_cube4<T>(in.skyFaces, in.skySize, R);
        } else {
            col = {T(p.skyColor[0]), T(p.skyColor[1]), T(p.skyColor[2]), T(p.skyColor[3])};
        }
    }
    out[0] = col.x; out[1] = col.y; out[2] = col.z; out[3] = col.w;
    return true;
}


// ---------------------------------------------------------------------------
// Global SDF build from mesh SDFs (Shaders/SDF/SDFRasterizeModel.glsl:42-63, GlobalSDFClear.comp:27-33,
// distanceToModelSDF SDFCommon.glsl:38-62)
// ---------------------------------------------------------------------------
// float -> nearest half (ties to even) -> float, the rounding of an R16F imageStore
inline float round_to_half(float f) {
    if (!(std::fabs(f) < 65520.0f)) return f;  // not reached by clamped distances
    if (f == 0.0f) return f;
    int e;
    std::frexp(f, &e);                      // |f| in [2^(e-1), 2^e)
    const int ulpExp = std::max(e - 11, -24);  // 11 significant bits, subnormals below 2^-14
    return (float)std::nearbyint((double)f * std::ldexp(1.0, -ulpExp)) * (float)std::ldexp(1.0, ulpExp);
}

inline float half_bits_to_float(uint16_t h) {
    const int s = h >> 15, e = (h >> 10) & 31, m = h & 1023;
    float v;
    if (e == 0) v = std::ldexp((float)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((float)(m | 1024), e - 25);
    return s ? -v : v;
}

// textureLod of an R16F mesh SDF level: trilinear, repeat addressing, level clamped to the chain
template <class T>
inline T mesh_sample(const uint16_t* pool, const LuxgiMeshSDF& m, int level, v3<T> uvw) {
    level = std::max(0, std::min(level, m.mipCount - 1));
    const int W = m.size[level][0], H = m.size[level][1], D = m.size[level][2];
    const uint16_t* v = pool + m.offset[level];
    const T x = uvw.x * T(W) - T(0.5), y = uvw.y * T(H) - T(0.5), z = uvw.z * T(D) - T(0.5);
    const T fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    const T ax = x - fx, ay = y - fy, az = z - fz;
    auto wrap = [](long long i, int n) { long long r = i % n; return (int)(r < 0 ? r + n : r); };
    const int x0 = wrap((long long)fx, W), x1 = wrap((long long)fx + 1, W);
    const int y0 = wrap((long long)fy, H), y1 = wrap((long long)fy + 1, H);
    const int z0 = wrap((long long)fz, D), z1 = wrap((long long)fz + 1, D);
    auto t = [&](int xi, int yi, int zi) { return T(half_bits_to_float(v[((size_t)zi * H + yi) * W + xi])); };
    const T c00 = t(x0, y0, z0) * (T(1) - ax) + t(x1, y0, z0) * ax, c10 = t(x0, y1, z0) * (T(1) - ax) + t(x1, y1, z0) * ax;
    const T c01 = t(x0, y0, z1) * (T(1) - ax) + t(x1, y0, z1) * ax, c11 = t(x0, y1, z1) * (T(1) - ax) + t(x1, y1, z1) * ax;
    return (c00 * (T(1) - ay) + c10 * ay) * (T(1) - az) + (c01 * (T(1) - ay) + c11 * ay) * az;
}

template <class T>
inline T distance_to_model_sdf(T minDistance, const LuxgiObjectRasterizeData& o, const LuxgiMeshSDF& m,
                               const uint16_t* pool, v3<T> wp) {
    const v3<T> vp = mat4_mul<T>(o.worldToVolume, wp, T(1));
    const v3<T> uvw{vp.x * T(o.volumeToUVWMul[0]) + T(o.volumeToUVWAdd[0]), vp.y * T(o.volumeToUVWMul[1]) + T(o.volumeToUVWAdd[1]),
                    vp.z * T(o.volumeToUVWMul[2]) + T(o.volumeToUVWAdd[2])};
    const v3<T> e{T(o.volumeLocalBoundsExtent[0]), T(o.volumeLocalBoundsExtent[1]), T(o.volumeLocalBoundsExtent[2])};
    const v3<T> vc{clampT(vp.x, -e.x, e.x), clampT(vp.y, -e.y, e.y), clampT(vp.z, -e.z, e.z)};
    const v3<T> wc = mat4_mul<T>(o.volumeToWorld, vc, T(1));
    T dv = length(wp - wc);
    if (dv < T(0.01f)) dv = length(vp - vc);
    dv = std::max(dv, T(0));
    if (minDistance <= dv) return dv;
    const T sd = (mesh_sample<T>(pool, m, (int)o.mipOffset, uvw) * T(2) - T(1)) * T(o.decodeMul);
    T r = combine_distance_to_sdf<T>(sd, dv);
    if (dv > T(0)) r = std::max(dv, r);
    return r;
}

// One voxel of a listed chunk: objects in list order, batches of 28 restarting from the stored value.
template <class T>
inline T rasterize_voxel(const LuxgiRasterizeParams& p, const LuxgiObjectRasterizeData* objects,
                         const LuxgiMeshSDF* meshes, const uint16_t* pool, int numObjects, const uint32_t* list,
                         uint32_t n, int vx, int vy, int vz, bool half) {
    const v3<T> wp{T(vx) * T(p.cascadeCoordToPosMul[0]) + T(p.cascadeCoordToPosAdd[0]),
                   T(vy) * T(p.cascadeCoordToPosMul[1]) + T(p.cascadeCoordToPosAdd[1]),
                   T(vz) * T(p.cascadeCoordToPosMul[2]) + T(p.cascadeCoordToPosAdd[2])};
    const T maxD = T(p.maxDistance);
    T minDistance = maxD, stored = T(1);
    for (uint32_t b = 0; b < n; b += 28) {
        if (b != 0) minDistance = maxD * stored;
        for (uint32_t k = b; k < std::min<uint32_t>(n, b + 28); ++k) {
            const uint32_t id = list[k];
            if (id >= (uint32_t)numObjects) continue;
            const LuxgiMeshSDF& m = meshes[id];
            bool ok = m.mipCount >= 1 && m.mipCount <= 3;
            for (int l = 0; ok && l < m.mipCount; ++l) ok = m.size[l][0] > 0 && m.size[l][1] > 0 && m.size[l][2] > 0;
            if (!ok) continue;
            minDistance = std::min(minDistance, distance_to_model_sdf<T>(minDistance, objects[id], m, pool, wp));
        }
        stored = clampT(minDistance / maxD, T(-1), T(1));
        if (half) stored = T(round_to_half((float)stored));
    }
    return stored;
}


// ---------------------------------------------------------------------------
// Learned a-trous pass (the reflection a-trous of Reflection/DenoiseAtrous.comp:63-134 with an MLP
// tap weight; no counterpart in the reference) and its backward. `guide` gives the quantities the
// weights are computed from (luminance, variance, mode), `color` the values that are blended: the
// product path passes the same image for both; the finite-difference check perturbs `color` only.
// ---------------------------------------------------------------------------
struct LearnedAtrousImage {
    int W, H;
    const float *normals, *pbr;   // RGBA32F G-buffer images
    const float *guide, *color;   // RGBA32F (rgb, variance)
};

template <class T> inline T la_lum(T r, T g, T b) { return std::max(r * T(0.299) + g * T(0.587) + b * T(0.114), T(0.0001)); }

inline int la_mode_of(const LearnedAtrousImage& im, size_t q) {
    const float z = im.pbr[q * 4 + 3], r = im.pbr[q * 4 + 1];
    if (z == 1.0f) return 2;
    return (r <= 0.05f || r >= 0.45f) ? 1 : 0;
}

template <class T>
inline T la_center_var(const LearnedAtrousImage& im, int x, int y) {
    const T kv[2][2] = {{T(0.25), T(0.125)}, {T(0.125), T(0.0625)}};
    T v = 0;
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) {
            const int px = x + xx, py = y + yy;
            if (px >= 0 && py >= 0 && px < im.W && py < im.H) v += T(im.guide[((size_t)py * im.W + px) * 4 + 3]) * kv[std::abs(xx)][std::abs(yy)];
        }
    return v;
}

// tap (xx, yy) of centre (x, y): the MLP input
template <class T>
inline void la_tap_features(const LearnedAtrousImage& im, const LuxgiLearnedAtrousParams& p, int x, int y, int xx,
                            int yy, T f[6]) {
    const size_t c = (size_t)y * im.W + x, q = (size_t)(y + yy * p.stepSize) * im.W + (x + xx * p.stepSize);
    const T var = la_center_var<T>(im, x, y);
    const T phi = T(p.phiColor) * std::sqrt(std::max(T(0), var)) + T(1e-10);
    const T lc = la_lum<T>(T(im.guide[c * 4]), T(im.guide[c * 4 + 1]), T(im.guide[c * 4 + 2]));
    const T lq = la_lum<T>(T(im.guide[q * 4]), T(im.guide[q * 4 + 1]), T(im.guide[q * 4 + 2]));
    const v3<T> nc = octahedral_to_direction<T>(T(im.normals[c * 4]), T(im.normals[c * 4 + 1]));
    const v3<T> nq = octahedral_to_direction<T>(T(im.normals[q * 4]), T(im.normals[q * 4 + 1]));
    const T kw[3] = {T(1), T(2) / T(3), T(1) / T(6)};
    f[0] = std::min(std::fabs(lc - lq) / phi, T(16));
    f[1] = std::exp(-std::fabs(T(im.pbr[c * 4 + 3]) - T(im.pbr[q * 4 + 3])) / T(p.sigmaDepth));
    f[2] = std::max(T(0), dot(nc, nq));
    f[3] = kw[std::abs(xx)] * kw[std::abs(yy)];
    f[4] = T(p.level);
    f[5] = std::min(std::sqrt(std::max(var, T(0))), T(4));
}

// MLP forward: returns exp(clamp(o)), h the hidden activations, *clamped whether o was clamped
template <class T>
inline T la_mlp_eval(const T* P, const T f[6], T h[10], bool* clamped) {
    const T* W1 = P;
    const T* b1 = P + 60;
    const T* w2 = P + 70;
    T o = P[80];
    for (int j = 0; j < 10; ++j) {
        T a = b1[j];
        for (int i = 0; i < 6; ++i) a += W1[j * 6 + i] * f[i];
        h[j] = std::max(a, T(0));
        o += w2[j] * h[j];
    }
    *clamped = o < T(-16) || o > T(8);
    return std::exp(clampT(o, T(-16), T(8)));
}

template <class T>
inline bool la_tap_inside(const LearnedAtrousImage& im, const LuxgiLearnedAtrousParams& p, int x, int y, int xx, int yy) {
    const int px = x + xx * p.stepSize, py = y + yy * p.stepSize;
    return px >= 0 && py >= 0 && px < im.W && py < im.H && !(xx == 0 && yy == 0);
}

// forward of one pixel: out[4], *sumW
template <class T>
inline void la_forward_pixel(const LearnedAtrousImage& im, const LuxgiLearnedAtrousParams& p, const T* P, int x, int y,
                             T out[4], T* sumW) {
    const size_t e = (size_t)y * im.W + x;
    const int mode = la_mode_of(im, e);
    *sumW = T(1);
    if (mode == 2) { out[0] = out[1] = out[2] = out[3] = 0; return; }
    for (int k = 0; k < 4; ++k) out[k] = T(im.color[e * 4 + k]);
    if (mode == 1) return;
    T s[4] = {out[0], out[1], out[2], out[3]}, sw = 1;
    for (int yy = -p.radius; yy <= p.radius; ++yy)
        for (int xx = -p.radius; xx <= p.radius; ++xx) {
            if (!la_tap_inside<T>(im, p, x, y, xx, yy)) continue;
            T f[6], h[10];
            bool cl;
            la_tap_features<T>(im, p, x, y, xx, yy, f);
            const T w = la_mlp_eval<T>(P, f, h, &cl);
            const size_t q = (size_t)(y + yy * p.stepSize) * im.W + (x + xx * p.stepSize);
            sw += w;
            for (int k = 0; k < 3; ++k) s[k] += w * T(im.color[q * 4 + k]);
            s[3] += w * w * T(im.color[q * 4 + 3]);
        }
    for (int k = 0; k < 3; ++k) out[k] = s[k] / sw;
    out[3] = s[3] / (sw * sw);
    *sumW = sw;
}

// backward over the image (scatter form): gIn [H][W][4] (rgb), gP [81] accumulated
template <class T>
inline void la_backward(const LearnedAtrousImage& im, const LuxgiLearnedAtrousParams& p, const T* P, const T* gOut,
                        T* gIn, T* gP) {
    for (size_t i = 0; i < (size_t)im.W * im.H * 4; ++i) gIn[i] = 0;
    for (int y = 0; y < im.H; ++y)
        for (int x = 0; x < im.W; ++x) {
            const size_t e = (size_t)y * im.W + x;
            const int mode = la_mode_of(im, e);
            if (mode == 2) continue;
            if (mode == 1) {
                for (int k = 0; k < 3; ++k) gIn[e * 4 + k] += gOut[e * 4 + k];
                continue;
            }
            T out[4], sw;
            la_forward_pixel<T>(im, p, P, x, y, out, &sw);
            for (int k = 0; k < 3; ++k) gIn[e * 4 + k] += gOut[e * 4 + k] / sw;
            for (int yy = -p.radius; yy <= p.radius; ++yy)
                for (int xx = -p.radius; xx <= p.radius; ++xx) {
                    if (!la_tap_inside<T>(im, p, x, y, xx, yy)) continue;
                    T f[6], h[10];
                    bool cl;
                    la_tap_features<T>(im, p, x, y, xx, yy, f);
                    const T w = la_mlp_eval<T>(P, f, h, &cl);
                    const size_t q = (size_t)(y + yy * p.stepSize) * im.W + (x + xx * p.stepSize);
                    T dw = 0;
                    for (int k = 0; k < 3; ++k) {
                        gIn[q * 4 + k] += gOut[e * 4 + k] * w / sw;
                        dw += gOut[e * 4 + k] * (T(im.color[q * 4 + k]) - out[k]);
                    }
                    dw /= sw;
                    if (cl) continue;
                    const T dO = dw * w;
                    gP[80] += dO;
                    for (int j = 0; j < 10; ++j) {
                        gP[70 + j] += dO * h[j];
                        const T dh = h[j] > T(0) ? dO * P[70 + j] : T(0);
                        gP[60 + j] += dh;
                        for (int i = 0; i < 6; ++i) gP[j * 6 + i] += dh * f[i];
                    }
                }
        }
}

// ---------------------------------------------------------------------------
// Per-pixel learned edge-stopping scales (no counterpart in the reference). A 12 -> 32 -> 32 -> 5
// ReLU MLP of per-pixel features gives, for each of the five a-trous passes
// (Reflection/DenoiseAtrous.comp:63-134), a scale s_i(p) = exp(clamp(o_i, -4, 4)) of the luminance
// term of the edge-stopping weight (EdgeStop.glsl:24-40): wL = s_i(p) |l_p - l_q| / phi_p. s = 1
// is the reference pass. Parameters, packed: W1 [32][12], b1 [32], W2 [32][32], b2 [32],
// W3 [5][32], b3 [5] (1637 floats).
// ---------------------------------------------------------------------------
constexpr int kLpIn = 12, kLpHid = 32, kLpOut = 5;
constexpr int kLpW1 = 0, kLpB1 = kLpW1 + kLpHid * kLpIn, kLpW2 = kLpB1 + kLpHid, kLpB2 = kLpW2 + kLpHid * kLpHid,
              kLpW3 = kLpB2 + kLpHid, kLpB3 = kLpW3 + kLpOut * kLpHid, kLpParams = kLpB3 + kLpOut;
constexpr double kLpClamp = 4.0;

// The 12 features of pixel (x, y) from the pass-0 input (rgb + variance), the reprojection's
// moments (m1, m2, history length), the G-buffer normal (octahedral xy) and pbr (.w linear depth):
//   f0, f1  |z(x+1) - z(x-1)| / 2z and the same along y (edge-clamped neighbours), at most 1
//   f2..f4  the decoded normal
//   f5      log(1 + l), l the luminance of the input
//   f6, f7  log(1 + mean) and min(std / (mean + 1e-3), 4) of the luminance over the 3x3 neighbourhood
//   f8      min(sqrt(variance) / (l + 1e-3), 4)
//   f9, f10 log(1 + m1) and min(sqrt(max(0, m2 - m1^2)) / (m1 + 1e-3), 4)
//   f11     min(history length, 32) / 32
template <class T>
void learned_pixel_features(int W, int H, int x, int y, Img color, Img moments, Img normals, Img pbr, T* f) {
    auto cx = [&](int v) { return std::min(std::max(v, 0), W - 1); };
    auto cy = [&](int v) { return std::min(std::max(v, 0), H - 1); };
    const T z = fetch<T>(pbr, W, H, x, y).w;
    const T iz = T(1) / std::max(z, T(1e-4f));
    const T zl = fetch<T>(pbr, W, H, cx(x - 1), y).w, zr = fetch<T>(pbr, W, H, cx(x + 1), y).w;
    const T zu = fetch<T>(pbr, W, H, x, cy(y - 1)).w, zd = fetch<T>(pbr, W, H, x, cy(y + 1)).w;
    f[0] = std::min(std::fabs(zr - zl) * T(0.5f) * iz, T(1));
    f[1] = std::min(std::fabs(zd - zu) * T(0.5f) * iz, T(1));
    const v4<T> nr = fetch<T>(normals, W, H, x, y);
    const v3<T> n = octahedral_to_direction<T>(nr.x, nr.y);
    f[2] = n.x;
    f[3] = n.y;
    f[4] = n.z;
    const v4<T> c = fetch<T>(color, W, H, x, y);
    const T l = luminance(c.x, c.y, c.z);
    f[5] = std::log1p(l);
    T s1 = 0, s2 = 0;
    for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) {
            const v4<T> q = fetch<T>(color, W, H, cx(x + xx), cy(y + yy));
            const T lq = luminance(q.x, q.y, q.z);
            s1 += lq;
            s2 += lq * lq;
        }
    const T mu = s1 / T(9);
    const T sd = std::sqrt(std::max(T(0), s2 / T(9) - mu * mu));
    f[6] = std::log1p(mu);
    f[7] = std::min(sd / (mu + T(1e-3f)), T(4));
    f[8] = std::min(std::sqrt(std::max(c.w, T(0))) / (l + T(1e-3f)), T(4));
    const v4<T> m = fetch<T>(moments, W, H, x, y);
    const T m1 = std::max(m.x, T(0));
    f[9] = std::log1p(m1);
    f[10] = std::min(std::sqrt(std::max(T(0), m.y - m1 * m1)) / (m1 + T(1e-3f)), T(4));
    f[11] = std::min(m.z, T(32)) / T(32);
}

// The MLP on one feature vector: pre-activations a1, a2, hidden h1, h2, outputs o and scales s.
// `half` rounds the parameters, the features and the hidden activations to FP16 before each
// product (the MFMA operands of the GPU forward); sums are exact in T.
template <class T>
struct LpForward {
    T a1[kLpHid], h1[kLpHid], a2[kLpHid], h2[kLpHid], o[kLpOut], s[kLpOut];
};
template <class T>
void learned_pixel_mlp(const double* P, const T* f, bool half, LpForward<T>& r) {
    auto q = [&](T v) { return half ? T(round_to_half(float(v))) : v; };
    for (int j = 0; j < kLpHid; ++j) {
        T a = T(P[kLpB1 + j]);
        for (int i = 0; i < kLpIn; ++i) a += q(T(P[kLpW1 + j * kLpIn + i])) * q(f[i]);
        r.a1[j] = a;
        r.h1[j] = std::max(a, T(0));
    }
    for (int j = 0; j < kLpHid; ++j) {
        T a = T(P[kLpB2 + j]);
        for (int i = 0; i < kLpHid; ++i) a += q(T(P[kLpW2 + j * kLpHid + i])) * q(r.h1[i]);
        r.a2[j] = a;
        r.h2[j] = std::max(a, T(0));
    }
    for (int j = 0; j < kLpOut; ++j) {
        T a = T(P[kLpB3 + j]);
        for (int i = 0; i < kLpHid; ++i) a += q(T(P[kLpW3 + j * kLpHid + i])) * q(r.h2[i]);
        r.o[j] = a;
        r.s[j] = std::exp(clampT(a, T(-kLpClamp), T(kLpClamp)));
    }
}

// dL/dP += the MLP's backward for one pixel, from gS = dL/ds (the clamp passes no gradient). With
// `half` the products use the rounded operands as the forward does (the gradient of the rounded
// network, applied to the unrounded parameters).
template <class T>
void learned_pixel_mlp_backward(const double* P, const T* f, bool half, const T* gS, T* gP) {
    LpForward<T> r;
    learned_pixel_mlp<T>(P, f, half, r);
    auto q = [&](T v) { return half ? T(round_to_half(float(v))) : v; };
    T go[kLpOut], gh2[kLpHid] = {}, gh1[kLpHid] = {};
    for (int j = 0; j < kLpOut; ++j)
        go[j] = (r.o[j] > T(-kLpClamp) && r.o[j] < T(kLpClamp)) ? gS[j] * r.s[j] : T(0);
    for (int j = 0; j < kLpOut; ++j) {
        gP[kLpB3 + j] += go[j];
        for (int i = 0; i < kLpHid; ++i) {
            gP[kLpW3 + j * kLpHid + i] += go[j] * q(r.h2[i]);
            gh2[i] += go[j] * q(T(P[kLpW3 + j * kLpHid + i]));
        }
    }
    for (int j = 0; j < kLpHid; ++j) {
        const T g = r.a2[j] > T(0) ? gh2[j] : T(0);
        gP[kLpB2 + j] += g;
        for (int i = 0; i < kLpHid; ++i) {
            gP[kLpW2 + j * kLpHid + i] += g * q(r.h1[i]);
            gh1[i] += g * q(T(P[kLpW2 + j * kLpHid + i]));
        }
    }
    for (int j = 0; j < kLpHid; ++j) {
        const T g = r.a1[j] > T(0) ? gh1[j] : T(0);
        gP[kLpB1 + j] += g;
        for (int i = 0; i < kLpIn; ++i) gP[kLpW1 + j * kLpIn + i] += g * q(f[i]);
    }
}

// Backward of one scaled reflection a-trous pass over the image (the pass of
// reflection_atrous_pixel with lumScale = scale[p], copy tiles passed through): from dL/d(out rgb)
// the gradient into the input rgb and into the per-pixel scale. The weights' dependence on the
// input (luminance, the variance behind phi) is held fixed; the variance channel carries no
// gradient (the loss reads rgb).
template <class T>
void learned_pixel_atrous_backward(const LuxgiReflectionAtrousParams& pc, int W, int H, Img normals, Img pbr,
                                   const uint8_t* tileMap, Img in, const float* scale, const T* gOut, T* gIn,
                                   T* gScale) {
    const T kw[3] = {T(1.0), T(2.0 / 3.0), T(1.0 / 6.0)};
    const T kv[2][2] = {{T(1.0 / 4.0), T(1.0 / 8.0)}, {T(1.0 / 8.0), T(1.0 / 16.0)}};
    const int tw = (W + 7) / 8;
    for (size_t e = 0; e < (size_t)W * H * 4; ++e) gIn[e] = 0;
    for (size_t e = 0; e < (size_t)W * H; ++e) gScale[e] = 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t e = (size_t)y * W + x;
            const T* g = gOut + e * 4;
            const v4<T> c = fetch<T>(in, W, H, x, y);
            const v4<T> pb = fetch<T>(pbr, W, H, x, y);
            const T cd = pb.w, rough = pb.y;
            if (!tileMap[(y / 8) * tw + x / 8] || (cd != T(1) && (rough <= T(0.05f) || rough >= T(0.45f)))) {
                for (int k = 0; k < 3; ++k) gIn[e * 4 + k] += g[k];  // copied or passed through
                continue;
            }
            if (cd == T(1)) continue;  // sky: 0
            const T s = T(scale[e]);
            const T cl = luminance(c.x, c.y, c.z);
            T var = 0;
            for (int yy = -1; yy <= 1; ++yy)
                for (int xx = -1; xx <= 1; ++xx) var += fetch<T>(in, W, H, x + xx, y + yy).w * kv[std::abs(xx)][std::abs(yy)];
            const T phiC = T(pc.phiColor) * std::sqrt(std::max(T(0), var)) + T(1e-10f);
            const v4<T> nrm = fetch<T>(normals, W, H, x, y);
            const v3<T> cn = octahedral_to_direction<T>(nrm.x, nrm.y);
            // forward sums
            T sw = 1, o[3] = {c.x, c.y, c.z};
            struct Tap { int q; T wk, dl; };
            std::vector<Tap> taps;
            for (int yy = -pc.radius; yy <= pc.radius; ++yy)
                for (int xx = -pc.radius; xx <= pc.radius; ++xx) {
                    const int px = x + xx * pc.stepSize, py = y + yy * pc.stepSize;
                    if (!(px >= 0 && py >= 0 && px < W && py < H) || (xx == 0 && yy == 0)) continue;
                    const v4<T> sc = fetch<T>(in, W, H, px, py);
                    const T sl = luminance(sc.x, sc.y, sc.z);
                    const v4<T> sn = fetch<T>(normals, W, H, px, py);
                    const T d = fetch<T>(pbr, W, H, px, py).w;
                    const T w = edge_stop<T>(cd, d, T(pc.sigmaDepth), cn, octahedral_to_direction<T>(sn.x, sn.y),
                                             T(pc.phiNormal), cl, sl, phiC / s);
                    const T wk = w * kw[std::abs(xx)] * kw[std::abs(yy)];
                    sw += wk;
                    o[0] += wk * sc.x; o[1] += wk * sc.y; o[2] += wk * sc.z;
                    taps.push_back({py * W + px, wk, std::fabs(cl - sl) / phiC});
                }
            for (int k = 0; k < 3; ++k) o[k] /= sw;
            for (int k = 0; k < 3; ++k) gIn[e * 4 + k] += g[k] / sw;
            T gs = 0;
            for (const Tap& t : taps) {
                const v4<T> sc = fetch<T>(in, W, H, t.q % W, t.q / W);
                const T xq[3] = {sc.x, sc.y, sc.z};
                for (int k = 0; k < 3; ++k) {
                    gIn[(size_t)t.q * 4 + k] += g[k] * t.wk / sw;
                    gs += g[k] * (-t.dl * t.wk) * (xq[k] - o[k]) / sw;
                }
            }
            gScale[e] = gs;
        }
}

template <class F> inline void parallel_for(int n, int threads, F f) {
    if (threads <= 1 || n < 2) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    threads = std::min(threads, n);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([=, &f] {
            int b = (int)((long long)n * t / threads), e = (int)((long long)n * (t + 1) / threads);
            for (int i = b; i < e; ++i) f(i);
        });
    for (auto& th : pool) th.join();
}

}  // namespace oracle

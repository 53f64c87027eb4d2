// svgf.hip -- SVGF temporal accumulation and a-trous filtering for gfx950.
//
// Reference (paths under Code/Maple/src/):
//   Shaders/Reflection/DenoiseReprojection.comp:78-218, DenoiseAtrous.comp:41-134,
//   DenoiseCopyTiles.comp:21-25, DenoiseReset.comp:23-32 (RGB reflection)
//   Shaders/Shadow/DenoiseReprojection.comp:71-210, DenoiseAtrous.comp:43-128,
//   DenoiseCopyTiles.comp:18-22 (scalar shadow visibility)
//   Shaders/Raytraced/Reprojection.glsl:12-405, EdgeStop.glsl:7-40
//   host drivers Engine/Raytrace/RaytracedReflection.cpp:346-520, RaytracedShadow.cpp:406-560
//
// Differences in structure (not in results):
//   - The 17x17 neighbourhood statistics are separable box sums over an LDS
//     halo tile (the reference fetches 289 texels per pixel).
//   - Tile classification writes a deterministic per-8x8-tile byte map (OR of
//     the pixels' votes) instead of appending tile origins with atomics; the
//     a-trous kernel runs over all tiles and copies or filters per the map,
//     which fuses DenoiseCopyTiles into DenoiseAtrous. The indirect-args reset
//     pass (DenoiseReset) has nothing left to reset.
//   - The history copy after iteration 1 (RaytracedReflection.cpp:488-491) is a
//     second store of the same a-trous output.
// texelFetch outside the image returns 0 (robust image access), as in the oracle.
#include <algorithm>

#include "svgf_common.hpp"

namespace luxgi {

constexpr float kNormalDistance = 0.1f;  // Reprojection.glsl:7
constexpr float kPlaneDistance = 5.0f;   // Reprojection.glsl:8

struct GBuf {
    const float4* color;
    const float4* normals;
    const float4* pbr;
    const float* depth;
};


// isValid (Reprojection.glsl:42-53)
__device__ __forceinline__ bool is_valid(int hx, int hy, f3 pos, f3 prevPos, f3 n, f3 pn, float meshId,
                                         float prevMeshId, int W, int H) {
    if (hx < 0 || hy < 0 || hx > W - 1 || hy > H - 1) return false;
    if (meshId != prevMeshId) return false;
    if (fabsf(dot3(pos - prevPos, n)) > kPlaneDistance) return false;
    const float d = fabsf(dot3(n, pn));
    if (d * d <= kNormalDistance) return false;
    return true;
}


// ---------------------------------------------------------------------------
// reflection reprojection
// ---------------------------------------------------------------------------
struct ReflArgs {
    LuxgiReflectionReprojectParams pc;
    int W, H, tilesX;
};
// the same with the parameters left where they are (the batched kernels' descriptors)
struct ReflArgsRef {
    const LuxgiReflectionReprojectParams& pc;
    int W, H, tilesX;
};

// reproject, RGB overload (Reprojection.glsl:242-405)
template <class S, class RA>
__device__ bool reproject_rgb(const RA& A, int cx, int cy, float depth, float rayLength, const GBuf& cur,
                              const GBuf& prev, const S* histOut, const S* histMom, f3& hc, float2& hm, float& hl) {
    const int W = A.W, H = A.H;
    const float sx = float(W), sy = float(H);
    const float tcx = (float(cx) + 0.5f) / sx, tcy = (float(cy) + 0.5f) / sy;
    const float4 normal = fetch4(cur.normals, cx, cy, W, H);
    const float mx = normal.z, my = normal.w;
    const f3 wpos = world_from_depth(tcx, tcy, depth, A.pc.viewProjInv);
    const float meshId = fetch4(cur.color, cx, cy, W, H).w;
    const float curvature = fetch4(cur.pbr, cx, cy, W, H).z;
    const float htx = tcx + mx, hty = tcy + my;
    float rx = float(cx) + mx * sx, ry = float(cy) + my * sy;
    if (rayLength > 0.0f && curvature == 0.0f) {  // virtualPointReprojection (Reprojection.glsl:56-75)
        const f3 ro = world_from_depth(float(cx) / sx, float(cy) / sy, depth, A.pc.viewProjInv);
        const f3 cp = mk(A.pc.cameraPos[0], A.pc.cameraPos[1], A.pc.cameraPos[2]);
        f3 cr = ro - cp;
        const float cl = len3(cr);
        cr = normalize3(cr);
        const f3 ph = cp + cr * (cl + rayLength);
        float4 rp = mat4_mul4(A.pc.prevViewProj, make_float4(ph.x, ph.y, ph.z, 1.0f));
        rp.x /= rp.w;
        rp.y /= rp.w;
        rx = (rp.x * 0.5f + 0.5f) * sx;
        ry = (rp.y * 0.5f + 0.5f) * sy;
    }
    const int hx = int(rx), hy = int(ry);  // GLSL ivec2(): truncation
    const f3 n = oct_decode(normal.x, normal.y);
    bool v[4];
    bool valid = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int lx = hx + (i & 1), ly = hy + (i >> 1);
        const float4 pn = fetch4(prev.normals, lx, ly, W, H);
        const float pd = fetch1(prev.depth, lx, ly, W, H);
        const f3 pp = world_from_depth(htx, hty, pd, A.pc.viewProjInv);
        const float pm = fetch4(prev.color, lx, ly, W, H).w;
        v[i] = is_valid(hx, hy, wpos, pp, n, oct_decode(pn.x, pn.y), meshId, pm, W, H);
        valid = valid || v[i];
    }
    f3 col = mk(0.f, 0.f, 0.f);
    float2 mom = make_float2(0.f, 0.f);
    if (valid) {
        float sumw = 0.0f;
        const float x = fractf(rx), y = fractf(ry);
        const float w[4] = {(1.0f - x) * (1.0f - y), x * (1.0f - y), (1.0f - x) * y, x * y};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lx = hx + (i & 1), ly = hy + (i >> 1);
            if (v[i]) {
                const float4 c = fetchS4<S>(histOut, lx, ly, W, H);
                const float4 m = fetchS4<S>(histMom, lx, ly, W, H);
                col = col + mk(c.x, c.y, c.z) * w[i];
                mom.x += w[i] * m.x;
                mom.y += w[i] * m.y;
                sumw += w[i];
            }
        }
        valid = sumw >= 0.01f;
        if (valid) {
            col = col * (1.0f / sumw);
            mom.x /= sumw;
            mom.y /= sumw;
        } else {
            col = mk(0.f, 0.f, 0.f);
            mom = make_float2(0.f, 0.f);
        }
    }
    if (!valid) {
        float count = 0.0f;
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) {
                const int px = hx + xx, py = hy + yy;
                const float4 pn = fetch4(prev.normals, px, py, W, H);
                const float pd = fetch1(prev.depth, px, py, W, H);
                const f3 pp = world_from_depth(htx, hty, pd, A.pc.viewProjInv);
                const float pm = fetch4(prev.color, px, py, W, H).w;
                if (is_valid(hx, hy, wpos, pp, n, oct_decode(pn.x, pn.y), meshId, pm, W, H)) {
                    const float4 c = fetchS4<S>(histOut, px, py, W, H);
                    const float4 m = fetchS4<S>(histMom, px, py, W, H);
                    col = col + mk(c.x, c.y, c.z);
                    mom.x += m.x;
                    mom.y += m.y;
                    count += 1.0f;
                }
            }
        if (count > 0.0f) {
            valid = true;
            col = col * (1.0f / count);
            mom.x /= count;
            mom.y /= count;
        }
    }
    if (valid) {
        hl = fetchS4<S>(histMom, hx, hy, W, H).z;
    } else {
        col = mk(0.f, 0.f, 0.f);
        mom = make_float2(0.f, 0.f);
        hl = 0.0f;
    }
    hc = col;
    hm = mom;
    return valid;
}

// LDS of one reprojection block
struct ReprojectShared {
    // rows of 48 float3 (144 dwords, 16 banks apart): in the row sums the two rows a half-wave
    // reads then fall in disjoint banks (a 33-wide row, 99 dwords, aliases them)
    float3 tile[32][48];
    float3 h1[32][16], h2[32][16];
    int vote[4];
};

// One 16x16-pixel block (block coordinates bX, bY) of the reprojection; LDS halo of 8 for the
// 17x17 statistics. Every lane of the block calls it.
template <class S, class RA>
__device__ __forceinline__ void reproject_block(const RA& A, const GBuf& cur, const GBuf& prev,
                                                const S* __restrict__ noisy, const S* __restrict__ histOut,
                                                const S* __restrict__ histMom, S* __restrict__ outColor,
                                                S* __restrict__ outMom, uint8_t* __restrict__ tileMap, int bX, int bY,
                                                ReprojectShared& sm) {
    auto& tile = sm.tile;
    auto& h1 = sm.h1;
    auto& h2 = sm.h2;
    auto& vote = sm.vote;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int bx = bX * 16, by = bY * 16;
    const int W = A.W, H = A.H;
    if (threadIdx.x < 4) vote[threadIdx.x] = 0;
    for (int i = threadIdx.x; i < 32 * 32; i += 256) {
        const int lx = i & 31, ly = i >> 5;
        const float4 c = fetchS4<S>(noisy, bx - 8 + lx, by - 8 + ly, W, H);
        tile[ly][lx] = make_float3(c.x, c.y, c.z);
    }
    __syncthreads();
    for (int row = ty; row < 32; row += 16) {
        float3 s1 = make_float3(0.f, 0.f, 0.f), s2 = make_float3(0.f, 0.f, 0.f);
        for (int k = 0; k < 17; ++k) {
            const float3 c = tile[row][tx + k];
            s1.x += c.x; s1.y += c.y; s1.z += c.z;
            s2.x += c.x * c.x; s2.y += c.y * c.y; s2.z += c.z * c.z;
        }
        h1[row][tx] = s1;
        h2[row][tx] = s2;
    }
    __syncthreads();
    float3 m1 = make_float3(0.f, 0.f, 0.f), m2 = make_float3(0.f, 0.f, 0.f);
    for (int k = 0; k < 17; ++k) {
        const float3 a = h1[ty + k][tx], b = h2[ty + k][tx];
        m1.x += a.x; m1.y += a.y; m1.z += a.z;
        m2.x += b.x; m2.y += b.y; m2.z += b.z;
    }
    const int x = bx + tx, y = by + ty;
    if (x < W && y < H) {
        const float depth = cur.depth[(size_t)y * W + x];
        const float rough = cur.pbr[(size_t)y * W + x].y;
        float4 oc = make_float4(0.f, 0.f, 0.f, 0.f), om = make_float4(0.f, 0.f, 0.f, 0.f);
        if (depth != 1.0f) {
            const float4 cr = Texel<S>::ld4(noisy, (size_t)y * W + x);
            f3 hc;
            float2 hm;
            float hl;
            const bool ok = reproject_rgb<S, RA>(A, x, y, depth, cr.w, cur, prev, histOut, histMom, hc, hm, hl);
            hl = fminf(32.0f, ok ? hl + 1.0f : 1.0f);
            if (ok) {  // neighborhoodStandardDeviation + clipAABB (Reflection/DenoiseReprojection.comp:78-114)
                const float inv = 1.0f / 289.0f;
                const f3 mean = mk(m1.x * inv, m1.y * inv, m1.z * inv);
                const f3 var = mk(m2.x * inv, m2.y * inv, m2.z * inv) - mean * mean;
                const f3 sd = mk(sqrtf(fmaxf(var.x, 0.f)), sqrtf(fmaxf(var.y, 0.f)), sqrtf(fmaxf(var.z, 0.f)));
                const f3 mn = mean - sd, mx = mean + sd;
                const f3 ctr = (mx + mn) * 0.5f;
                const f3 ext = (mx - mn) * 0.5f + mk(0.001f, 0.001f, 0.001f);
                const f3 cv = hc - ctr;
                const float mu = fmaxf(fmaxf(fabsf(cv.x / ext.x), fabsf(cv.y / ext.y)), fabsf(cv.z / ext.z));
                if (mu > 1.0f) hc = ctr + cv * (1.0f / mu);
            }
            const float cd = sqrtf(A.pc.cameraDelta[0] * A.pc.cameraDelta[0] + A.pc.cameraDelta[1] * A.pc.cameraDelta[1] +
                                   A.pc.cameraDelta[2] * A.pc.cameraDelta[2]);
            const float maxAcc = cd > 0.0f ? 8.0f : hl;
            const float alpha = ok ? fmaxf(A.pc.alpha, 1.0f / maxAcc) : 1.0f;
            const float alphaM = ok ? fmaxf(A.pc.momentsAlpha, 1.0f / maxAcc) : 1.0f;
            const float l = luminance(cr.x, cr.y, cr.z);
            const float mm0 = mixf(hm.x, l, alphaM), mm1 = mixf(hm.y, l * l, alphaM);
            // the square rounded on its own (contraction into a fused multiply-subtract is off
            // here): a single sample's moments (m2 = l*l rounded, m1 = l) then give exactly 0, as
            // in exact arithmetic, instead of the rounding residual of l*l, which would open the
            // luminance edge-stopping of freshly disoccluded pixels to ~1e-3
            float var;
            {
#pragma clang fp contract(off)
                var = fmaxf(0.0f, mm1 - mm0 * mm0);
            }
            oc = make_float4(mixf(hc.x, cr.x, alpha), mixf(hc.y, cr.y, alpha), mixf(hc.z, cr.z, alpha), var);
            om = make_float4(mm0, mm1, hl, 0.0f);
            if (!(rough <= 0.05f || rough >= 0.45f)) vote[(ty >> 3) * 2 + (tx >> 3)] = 1;
        }
        Texel<S>::st4(outMom, (size_t)y * W + x, om);
        Texel<S>::st4(outColor, (size_t)y * W + x, oc);
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int tX = bX * 2 + (threadIdx.x & 1), tY = bY * 2 + (threadIdx.x >> 1);
        if (tX * 8 < W && tY * 8 < H) tileMap[tY * A.tilesX + tX] = (uint8_t)vote[threadIdx.x];
    }
}

template <class S>
__global__ __launch_bounds__(256) void reflection_reproject_kernel(ReflArgs A, GBuf cur, GBuf prev,
                                                                   const S* __restrict__ noisy,
                                                                   const S* __restrict__ histOut,
                                                                   const S* __restrict__ histMom, S* __restrict__ outColor,
                                                                   S* __restrict__ outMom, uint8_t* __restrict__ tileMap) {
    __shared__ ReprojectShared sm;
    const int2 b = xcd_block_2d();
    reproject_block<S>(A, cur, prev, noisy, histOut, histMom, outColor, outMom, tileMap, b.x, b.y, sm);
}


// Reflection/DenoiseAtrous.comp:63-134 for one filtered pixel. tap(px, py, sc, sn, d) gives the
// input, decoded normal and linear depth of an in-image pixel. The sums run in one fixed order,
// so the LDS-tiled and the direct kernel produce identical results.
template <class Tap>
__device__ __forceinline__ float4 reflection_filter(const AtrousArgs& A, int x, int y, float4 c, float var, f3 cn,
                                                    float cd, Tap tap) {
    const float kw[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const float cl = luminance(c.x, c.y, c.z);
    const float phiC = A.phi * sqrtf(fmaxf(0.0f, var)) + 1e-10f;
    const float invPhiC = 1.0f / phiC, invPhiZ = 1.0f / A.sigmaDepth;
    const float invPhiL = A.lumScale ? invPhiC * A.lumScale[(size_t)y * A.W + x] : invPhiC;
    float sw = 1.0f;
    float4 s = c;
    for (int yy = -A.radius; yy <= A.radius; ++yy)
        for (int xx = -A.radius; xx <= A.radius; ++xx) {
            const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
            if (inside(px, py, A.W, A.H) && (xx != 0 || yy != 0)) {
                float4 sc;
                f3 sn;
                float d;
                tap(px, py, sc, sn, d);
                const float k = kw[abs(xx)] * kw[abs(yy)];
                const float sl = luminance(sc.x, sc.y, sc.z);
                const float w = edge_stop(cd, d, invPhiZ, cn, sn, A.phiNormal, A.sqNormal, cl, sl, invPhiL);
                const float wv = w * k;
                sw += wv;
                s.x += wv * sc.x;
                s.y += wv * sc.y;
                s.z += wv * sc.z;
                s.w += wv * wv * sc.w;
            }
        }
    return make_float4(s.x / sw, s.y / sw, s.z / sw, s.w / (sw * sw));
}


// Direct form, one 8x8 tile per 64-lane block: every tap read from global memory. Used when the
// footprint radius * stepSize exceeds the LDS halo.
template <class S>
__device__ __forceinline__ void atrous_direct_pixel(const AtrousArgs& A, const float4* __restrict__ normals,
                                                    const float4* __restrict__ pbr, const uint8_t* __restrict__ tileMap,
                                                    const S* __restrict__ in, S* __restrict__ out, S* __restrict__ hist,
                                                    int x, int y) {
    const int W = A.W, H = A.H;
    const size_t e = (size_t)y * W + x;
    const float4 c = Texel<S>::ld4(in, e);
    float4 o = c;
    if (tileMap[(y >> 3) * A.tilesX + (x >> 3)]) {
        const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
        float var = 0.0f;  // computeVarianceCenter (:41-61)
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) var += fetchS4<S>(in, x + xx, y + yy, W, H).w * kv[abs(xx)][abs(yy)];
        const float cd = load_linear_depth(pbr, e), rough = load_roughness(pbr, e);
        if (cd == 1.0f) {
            o = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (!(rough <= 0.05f || rough >= 0.45f)) {
            auto tap = [&](int px, int py, float4& sc, f3& sn, float& d) {
                const size_t q = (size_t)py * W + px;
                sc = Texel<S>::ld4(in, q);
                sn = load_normal(normals, q);
                d = load_linear_depth(pbr, q);
            };
            o = reflection_filter(A, x, y, c, var, load_normal(normals, e), cd, tap);
        }
    }
    Texel<S>::st4(out, e, o);
    if (hist) Texel<S>::st4(hist, e, o);
}

template <class S>
__global__ __launch_bounds__(64) void reflection_atrous_kernel(AtrousArgs A, const float4* __restrict__ normals,
                                                               const float4* __restrict__ pbr,
                                                               const uint8_t* __restrict__ tileMap,
                                                               const S* __restrict__ in, S* __restrict__ out,
                                                               S* __restrict__ hist) {
    const int2 b = xcd_block_2d();
    const int x = b.x * 8 + (threadIdx.x & 7), y = b.y * 8 + (threadIdx.x >> 3);
    if (x >= A.W || y >= A.H) return;
    atrous_direct_pixel<S>(A, normals, pbr, tileMap, in, out, hist, x, y);
}

// LDS-tiled form: a 16x16-pixel block (2x2 tiles) stages its input and guides (decoded normal,
// linear depth) with a halo of max(radius * stepSize, 1) pixels once, so each tap is an LDS read
// and each normal is decoded once per block instead of once per tap.
constexpr int kAtrousBlock = 16;
constexpr int kAtrousMaxHalo = 8;
constexpr int kAtrousMaxSide = kAtrousBlock + 2 * kAtrousMaxHalo;

struct AtrousRegion {
    int h, side, ox, oy;
};

// Block-uniform: does any of the block's (up to) 2x2 tiles need filtering?
__device__ __forceinline__ bool block_has_filter_tile(const AtrousArgs& A, const uint8_t* tileMap, int bX, int bY) {
    const int tilesY = (A.H + 7) / 8;
    bool any = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int tX = bX * 2 + (i & 1), tY = bY * 2 + (i >> 1);
        if (tX < A.tilesX && tY < tilesY && tileMap[tY * A.tilesX + tX]) any = true;
    }
    return any;
}

template <class S, class V, class Load>
__device__ __forceinline__ AtrousRegion stage_region(const AtrousArgs& A, const float4* normals, const float4* pbr,
                                                     V* sIn, float4* sGuide, Load load, int bX, int bY) {
    AtrousRegion r;
    r.h = max(A.radius * A.stepSize, 1);
    r.side = kAtrousBlock + 2 * r.h;
    r.ox = bX * kAtrousBlock - r.h;
    r.oy = bY * kAtrousBlock - r.h;
    for (int i = threadIdx.x; i < r.side * r.side; i += 256) {
        const int ly = i / r.side, lx = i - ly * r.side;
        const int gx = r.ox + lx, gy = r.oy + ly;
        V v{};
        float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
        if (inside(gx, gy, A.W, A.H)) {
            const size_t q = (size_t)gy * A.W + gx;
            v = load(q);
            const f3 n = load_normal(normals, q);
            g = make_float4(n.x, n.y, n.z, load_linear_depth(pbr, q));
        }
        sIn[i] = v;
        sGuide[i] = g;
    }
    __syncthreads();
    return r;
}

// One 16x16-pixel block (block coordinates bX, bY) of the LDS-tiled form; every lane of the block
// calls it (the staging has a barrier).
template <class S>
__device__ __forceinline__ void atrous_lds_block(const AtrousArgs& A, const float4* __restrict__ normals,
                                                 const float4* __restrict__ pbr, const uint8_t* __restrict__ tileMap,
                                                 const S* __restrict__ in, S* __restrict__ out, S* __restrict__ hist,
                                                 int bX, int bY, float4* sIn, float4* sGuide) {
    const int W = A.W, H = A.H;
    const int tx = threadIdx.x & (kAtrousBlock - 1), ty = threadIdx.x / kAtrousBlock;
    const int x = bX * kAtrousBlock + tx, y = bY * kAtrousBlock + ty;
    const bool inImage = x < W && y < H;
    const size_t e = (size_t)y * W + x;
    if (!block_has_filter_tile(A, tileMap, bX, bY)) {  // copy tiles only (DenoiseCopyTiles.comp:21-25)
        if (inImage) {
            const float4 c = Texel<S>::ld4(in, e);
            Texel<S>::st4(out, e, c);
            if (hist) Texel<S>::st4(hist, e, c);
        }
        return;
    }
    const AtrousRegion r = stage_region<S>(A, normals, pbr, sIn, sGuide,
                                           [&](size_t q) { return Texel<S>::ld4(in, q); }, bX, bY);
    if (!inImage) return;
    const int lc = (ty + r.h) * r.side + (tx + r.h);
    const float4 c = sIn[lc];
    float4 o = c;
    if (tileMap[(y >> 3) * A.tilesX + (x >> 3)]) {
        const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
        float var = 0.0f;  // computeVarianceCenter (:41-61); out-of-image texels are staged as 0
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) var += sIn[lc + yy * r.side + xx].w * kv[abs(xx)][abs(yy)];
        const float4 gc = sGuide[lc];
        const float cd = gc.w, rough = load_roughness(pbr, e);
        if (cd == 1.0f) {
            o = make_float4(0.f, 0.f, 0.f, 0.f);
        } else if (!(rough <= 0.05f || rough >= 0.45f)) {
            auto tap = [&](int px, int py, float4& sc, f3& sn, float& d) {
                const int l = (py - r.oy) * r.side + (px - r.ox);
                sc = sIn[l];
                const float4 g = sGuide[l];
                sn = mk(g.x, g.y, g.z);
                d = g.w;
            };
            o = reflection_filter(A, x, y, c, var, mk(gc.x, gc.y, gc.z), cd, tap);
        }
    }
    Texel<S>::st4(out, e, o);
    if (hist) Texel<S>::st4(hist, e, o);
}

template <class S>
__global__ __launch_bounds__(256) void reflection_atrous_lds_kernel(AtrousArgs A, const float4* __restrict__ normals,
                                                                    const float4* __restrict__ pbr,
                                                                    const uint8_t* __restrict__ tileMap,
                                                                    const S* __restrict__ in, S* __restrict__ out,
                                                                    S* __restrict__ hist) {
    // LDS sized by the launch to the region this step stages ((16 + 2 h)^2 texels of input and
    // guide): small steps keep more blocks resident
    extern __shared__ float4 atrousLds[];
    const int side = kAtrousBlock + 2 * max(A.radius * A.stepSize, 1);
    float4* sIn = atrousLds;
    float4* sGuide = atrousLds + side * side;
    const int2 b = xcd_block_2d();
    atrous_lds_block<S>(A, normals, pbr, tileMap, in, out, hist, b.x, b.y, sIn, sGuide);
}

// ---------------------------------------------------------------------------
// shadow
// ---------------------------------------------------------------------------
struct ShadowArgs {
    LuxgiShadowReprojectParams pc;
    int W, H, tilesX, maskW, maskH;
};

__device__ __forceinline__ float shadow_bit(const uint32_t* mask, int mw, int mh, int x, int y) {
    if (x < 0 || y < 0) return 0.0f;
    const int wx = x >> 3, wy = y >> 2;
    if (wx >= mw || wy >= mh) return 0.0f;
    return float((mask[(size_t)wy * mw + wx] >> ((y & 3) * 8 + (x & 7))) & 1u);
}

// reproject, scalar overload (Reprojection.glsl:90-238)
template <class S>
__device__ bool reproject_scalar(const ShadowArgs& A, int cx, int cy, float depth, const GBuf& cur, const GBuf& prev,
                                 const S* histOut, const S* histMom, float& hc, float2& hm, float& hl) {
    const int W = A.W, H = A.H;
    const float sx = float(W), sy = float(H);
    const float pcx = float(cx) + 0.5f, pcy = float(cy) + 0.5f;
    const float tcx = pcx / sx, tcy = pcy / sy;
    const float4 normal = fetch4(cur.normals, cx, cy, W, H);
    const float mx = normal.z, my = normal.w;
    const f3 wpos = world_from_depth(tcx, tcy, depth, A.pc.viewProjInv);
    const float meshId = fetch4(cur.color, cx, cy, W, H).w;
    const int hx = int(pcx + mx * sx), hy = int(pcy + my * sy);
    const float fx = floorf(float(cx)) + mx * sx, fy = floorf(float(cy)) + my * sy;
    const int fxi = int(fx), fyi = int(fy);
    const float htx = tcx + mx, hty = tcy + my;
    const f3 n = oct_decode(normal.x, normal.y);
    bool v[4];
    bool valid = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int lx = fxi + (i & 1), ly = fyi + (i >> 1);
        const float4 pn = fetch4(prev.normals, lx, ly, W, H);
        const float pd = fetch1(prev.depth, lx, ly, W, H);
        const f3 pp = world_from_depth(htx, hty, pd, A.pc.viewProjInv);
        const float pm = fetch4(prev.color, lx, ly, W, H).w;
        v[i] = is_valid(hx, hy, wpos, pp, n, oct_decode(pn.x, pn.y), meshId, pm, W, H);
        valid = valid || v[i];
    }
    float col = 0.0f;
    float2 mom = make_float2(0.f, 0.f);
    if (valid) {
        float sumw = 0.0f;
        const float x = fractf(fx), y = fractf(fy);
        const float w[4] = {(1.0f - x) * (1.0f - y), x * (1.0f - y), (1.0f - x) * y, x * y};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int lx = fxi + (i & 1), ly = fyi + (i >> 1);
            if (v[i]) {
                col += w[i] * fetchS2<S>(histOut, lx, ly, W, H).x;
                const float4 m = fetchS4<S>(histMom, lx, ly, W, H);
                mom.x += w[i] * m.x;
                mom.y += w[i] * m.y;
                sumw += w[i];
            }
        }
        valid = sumw >= 0.01f;
        if (valid) {
            col /= sumw;
            mom.x /= sumw;
            mom.y /= sumw;
        } else {
            col = 0.0f;
            mom = make_float2(0.f, 0.f);
        }
    }
    if (!valid) {
        float count = 0.0f;
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) {
                const int px = hx + xx, py = hy + yy;
                const float4 pn = fetch4(prev.normals, px, py, W, H);
                const float pd = fetch1(prev.depth, px, py, W, H);
                const f3 pp = world_from_depth(htx, hty, pd, A.pc.viewProjInv);
                const float pm = fetch4(prev.color, px, py, W, H).w;
                if (is_valid(hx, hy, wpos, pp, n, oct_decode(pn.x, pn.y), meshId, pm, W, H)) {
                    col += fetchS2<S>(histOut, px, py, W, H).x;
                    const float4 m = fetchS4<S>(histMom, px, py, W, H);
                    mom.x += m.x;
                    mom.y += m.y;
                    count += 1.0f;
                }
            }
        if (count > 0.0f) {
            valid = true;
            col /= count;
            mom.x /= count;
            mom.y /= count;
        }
    }
    if (valid) hl = fetchS4<S>(histMom, hx, hy, W, H).z;
    else {
        col = 0.0f;
        mom = make_float2(0.f, 0.f);
        hl = 0.0f;
    }
    hc = col;
    hm = mom;
    return valid;
}

template <class S>
__global__ __launch_bounds__(256) void shadow_reproject_kernel(ShadowArgs A, GBuf cur, GBuf prev,
                                                               const uint32_t* __restrict__ mask,
                                                               const S* __restrict__ histOut,
                                                               const S* __restrict__ histMom, S* __restrict__ outColor,
                                                               S* __restrict__ outMom, uint8_t* __restrict__ tileMap) {
    __shared__ float bits[32][33];
    __shared__ float hs[32][16];
    __shared__ int vote[4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int2 blk = xcd_block_2d();
    const int bx = blk.x * 16, by = blk.y * 16;
    const int W = A.W, H = A.H;
    if (threadIdx.x < 4) vote[threadIdx.x] = 0;
    for (int i = threadIdx.x; i < 32 * 32; i += 256) {
        const int lx = i & 31, ly = i >> 5;
        bits[ly][lx] = shadow_bit(mask, A.maskW, A.maskH, bx - 8 + lx, by - 8 + ly);
    }
    __syncthreads();
    for (int row = ty; row < 32; row += 16) {
        float s = 0.0f;
        for (int k = 0; k < 17; ++k) s += bits[row][tx + k];
        hs[row][tx] = s;
    }
    __syncthreads();
    float sum = 0.0f;
    for (int k = 0; k < 17; ++k) sum += hs[ty + k][tx];
    const float mean = sum / 289.0f;  // computeMean (Shadow/DenoiseReprojection.comp:103-124)
    const int x = bx + tx, y = by + ty;
    if (x < W && y < H) {
        const float depth = cur.depth[(size_t)y * W + x];
        float2 ov = make_float2(0.f, 0.f), om = make_float2(0.f, 0.f);
        float hl = 0.0f;
        if (depth != 1.0f) {
            const float vis = bits[ty + 8][tx + 8];
            float hv;
            float2 hm;
            const bool ok = reproject_scalar<S>(A, x, y, depth, cur, prev, histOut, histMom, hv, hm, hl);
            hl = fminf(32.0f, ok ? hl + 1.0f : 1.0f);
            if (ok) {
                const float sv = fmaxf(mean - mean * mean, 0.0f);
                const float sd = sqrtf(sv);
                hv = clampf(hv, mean - 0.5f * sd, mean + 0.5f * sd);
            }
            const float alpha = ok ? fmaxf(A.pc.alpha, 1.0f / hl) : 1.0f;
            const float alphaM = ok ? fmaxf(A.pc.momentsAlpha, 1.0f / hl) : 1.0f;
            om.x = mixf(hm.x, vis, alphaM);
            om.y = mixf(hm.y, vis * vis, alphaM);
            ov.x = mixf(hv, vis, alpha);
            {
#pragma clang fp contract(off)
                ov.y = fmaxf(0.0f, om.y - om.x * om.x);  // as in the reflection reprojection
            }
            if (ov.x > 0.0f) vote[(ty >> 3) * 2 + (tx >> 3)] = 1;
        }
        Texel<S>::st4(outMom, (size_t)y * W + x, make_float4(om.x, om.y, hl, 0.0f));
        Texel<S>::st2(outColor, (size_t)y * W + x, ov);
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int tX = blk.x * 2 + (threadIdx.x & 1), tY = blk.y * 2 + (threadIdx.x >> 1);
        if (tX * 8 < W && tY * 8 < H) tileMap[tY * A.tilesX + tX] = (uint8_t)vote[threadIdx.x];
    }
}

// Shadow/DenoiseAtrous.comp:65-128 for one filtered pixel (same tap contract and fixed sum order
// as reflection_filter).
template <class Tap>
__device__ __forceinline__ float2 shadow_filter(const AtrousArgs& A, int x, int y, float2 c, float var, f3 cn, float cd,
                                                Tap tap) {
    const float kw[3] = {1.0f, 2.0f / 3.0f, 1.0f / 6.0f};
    const float phiV = A.phi * sqrtf(fmaxf(0.0f, 1e-10f + var));
    const float invPhiV = 1.0f / phiV, invPhiZ = 1.0f / A.sigmaDepth;
    float sw = 1.0f, s0 = c.x, s1 = c.y;
    for (int yy = -A.radius; yy <= A.radius; ++yy)
        for (int xx = -A.radius; xx <= A.radius; ++xx) {
            const int px = x + xx * A.stepSize, py = y + yy * A.stepSize;
            if (inside(px, py, A.W, A.H) && (xx != 0 || yy != 0)) {
                float2 sv;
                f3 sn;
                float d;
                tap(px, py, sv, sn, d);
                const float k = kw[abs(xx)] * kw[abs(yy)];
                const float w = edge_stop(cd, d, invPhiZ, cn, sn, A.phiNormal, A.sqNormal, c.x, sv.x, invPhiV);
                const float wv = w * k;
                sw += wv;
                s0 += wv * sv.x;
                s1 += wv * wv * sv.y;
            }
        }
    float2 o = make_float2(s0 / sw, s1 / (sw * sw));
    if (A.power != 0.0f) o.x = powf(o.x, A.power);
    return o;
}

// Direct form, one 8x8 tile per 64-lane block; copy tiles write 0 (Shadow/DenoiseCopyTiles.comp).
template <class S>
__global__ __launch_bounds__(64) void shadow_atrous_kernel(AtrousArgs A, const float4* __restrict__ normals,
                                                           const float4* __restrict__ pbr,
                                                           const uint8_t* __restrict__ tileMap, const S* __restrict__ in,
                                                           S* __restrict__ out, S* __restrict__ hist) {
    const int2 blk = xcd_block_2d();
    const int x = blk.x * 8 + (threadIdx.x & 7), y = blk.y * 8 + (threadIdx.x >> 3);
    const int W = A.W, H = A.H;
    if (x >= W || y >= H) return;
    const size_t e = (size_t)y * W + x;
    float2 o = make_float2(0.f, 0.f);
    if (tileMap[blk.y * A.tilesX + blk.x]) {
        const float2 c = Texel<S>::ld2(in, e);
        const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
        float var = 0.0f;
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) var += fetchS2<S>(in, x + xx, y + yy, W, H).y * kv[abs(xx)][abs(yy)];
        const float cd = load_linear_depth(pbr, e);
        if (cd < 0.0f) {
            o = c;
        } else {
            auto tap = [&](int px, int py, float2& sv, f3& sn, float& d) {
                const size_t q = (size_t)py * W + px;
                sv = Texel<S>::ld2(in, q);
                sn = load_normal(normals, q);
                d = load_linear_depth(pbr, q);
            };
            o = shadow_filter(A, x, y, c, var, load_normal(normals, e), cd, tap);
        }
    }
    Texel<S>::st2(out, e, o);
    if (hist) Texel<S>::st2(hist, e, o);
}

// LDS-tiled form (see reflection_atrous_lds_kernel).
template <class S>
__global__ __launch_bounds__(256) void shadow_atrous_lds_kernel(AtrousArgs A, const float4* __restrict__ normals,
                                                                const float4* __restrict__ pbr,
                                                                const uint8_t* __restrict__ tileMap,
                                                                const S* __restrict__ in, S* __restrict__ out,
                                                                S* __restrict__ hist) {
    extern __shared__ float4 atrousLds[];  // sized by the launch, as in reflection_atrous_lds_kernel
    const int side = kAtrousBlock + 2 * max(A.radius * A.stepSize, 1);
    float4* sGuide = atrousLds;
    float2* sIn = reinterpret_cast<float2*>(atrousLds + side * side);
    const int W = A.W, H = A.H;
    const int tx = threadIdx.x & (kAtrousBlock - 1), ty = threadIdx.x / kAtrousBlock;
    const int2 blk = xcd_block_2d();
    const int x = blk.x * kAtrousBlock + tx, y = blk.y * kAtrousBlock + ty;
    const bool inImage = x < W && y < H;
    const size_t e = (size_t)y * W + x;
    if (!block_has_filter_tile(A, tileMap, blk.x, blk.y)) {
        if (inImage) {
            Texel<S>::st2(out, e, make_float2(0.f, 0.f));
            if (hist) Texel<S>::st2(hist, e, make_float2(0.f, 0.f));
        }
        return;
    }
    const AtrousRegion r = stage_region<S>(A, normals, pbr, sIn, sGuide,
                                           [&](size_t q) { return Texel<S>::ld2(in, q); }, blk.x, blk.y);
    if (!inImage) return;
    float2 o = make_float2(0.f, 0.f);
    if (tileMap[(y >> 3) * A.tilesX + (x >> 3)]) {
        const int lc = (ty + r.h) * r.side + (tx + r.h);
        const float2 c = sIn[lc];
        const float kv[2][2] = {{1.0f / 4.0f, 1.0f / 8.0f}, {1.0f / 8.0f, 1.0f / 16.0f}};
        float var = 0.0f;
        for (int yy = -1; yy <= 1; ++yy)
            for (int xx = -1; xx <= 1; ++xx) var += sIn[lc + yy * r.side + xx].y * kv[abs(xx)][abs(yy)];
        const float4 gc = sGuide[lc];
        const float cd = gc.w;
        if (cd < 0.0f) {
            o = c;
        } else {
            auto tap = [&](int px, int py, float2& sv, f3& sn, float& d) {
                const int l = (py - r.oy) * r.side + (px - r.ox);
                sv = sIn[l];
                const float4 g = sGuide[l];
                sn = mk(g.x, g.y, g.z);
                d = g.w;
            };
            o = shadow_filter(A, x, y, c, var, mk(gc.x, gc.y, gc.z), cd, tap);
        }
    }
    Texel<S>::st2(out, e, o);
    if (hist) Texel<S>::st2(hist, e, o);
}

__host__ __device__ inline GBuf to_gbuf(const LuxgiGBuffer* g) {
    return GBuf{static_cast<const float4*>(static_cast<const void*>(g->color)),
                static_cast<const float4*>(static_cast<const void*>(g->normals)),
                static_cast<const float4*>(static_cast<const void*>(g->pbr)), g->depth};
}

// ---------------------------------------------------------------------------
// Batched reflection denoise. The work of a pass over all items is the list of their 16x16-pixel
// units (item k owns [blockStart_k, blockStart_k + units_k)); a persistent grid walks it, each
// block finds its unit's item by binary search over blockStart and runs the same block body as the
// single-frame kernels, so every item's output is identical to the single-frame path. Items, count
// and unit total are read from device memory at run time, so a captured graph follows them.
// ---------------------------------------------------------------------------
using Item = LuxgiReflectionDenoiseItem;

__device__ __forceinline__ int batch_item_of(const Item* items, int count, int w) {
    int lo = 0, hi = count - 1;  // the last item whose blockStart <= w
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].blockStart <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct BatchWalk {
    int begin, end, stride;
};

// order 0: block b takes units b, b + grid, ...; order 1: a contiguous run of ceil(total / grid)
__device__ __forceinline__ BatchWalk batch_walk(int total, int order) {
    BatchWalk w;
    if (order == 1) {
        const int per = (total + (int)gridDim.x - 1) / (int)gridDim.x;
        w.begin = (int)blockIdx.x * per;
        w.end = min(total, w.begin + per);
        w.stride = 1;
    } else {
        w.begin = blockIdx.x;
        w.end = total;
        w.stride = gridDim.x;
    }
    return w;
}

// the unit's item and its block coordinates; false when the unit lies outside the item (a header
// that does not match the items), which the whole block skips together
__device__ __forceinline__ bool batch_unit(const Item* items, int count, int w, const Item*& it, int& bX, int& bY) {
    // the item index is block-uniform: say so, so that its fields are read with scalar loads
    it = &items[__builtin_amdgcn_readfirstlane(batch_item_of(items, count, w))];
    const int W = it->width, H = it->height;
    if (W <= 0 || H <= 0) return false;
    const int nbx = (W + 15) / 16, local = w - it->blockStart;
    if (local < 0 || local >= nbx * ((H + 15) / 16)) return false;
    bX = local % nbx;
    bY = local / nbx;
    return true;
}

// one unit of the batched reprojection (every lane of the block calls it)
template <class S>
__device__ __forceinline__ void batch_reproject_unit(const Item* __restrict__ items, int count, int w,
                                                     ReprojectShared& sm) {
    const Item* it;
    int bX, bY;
    if (!batch_unit(items, count, w, it, bX, bY)) return;
    const int pp = it->pingPong & 1;
    // the parameters stay in memory (scalar loads where used) rather than copied into registers
    const ReflArgsRef A{it->reproject, it->width, it->height, (it->width + 7) / 8};
    __syncthreads();  // the previous unit's LDS reads are done
    reproject_block<S>(A, to_gbuf(&it->cur), to_gbuf(&it->prev), static_cast<const S*>(it->noisy),
                       static_cast<const S*>(it->outputs[1]), static_cast<const S*>(it->moments[1 - pp]),
                       static_cast<S*>(it->outputs[0]), static_cast<S*>(it->moments[pp]), it->tileMap, bX, bY, sm);
}

// ONE: one unit per block (unit = blockIdx.x; the grid covers the units, order 2), without the
// persistent loop, which costs the body ~50 registers
template <class S, bool ONE>
__global__ __launch_bounds__(256) void batch_reproject_kernel(const LuxgiDenoiseBatchHeader* __restrict__ hdr,
                                                              const Item* __restrict__ items, int order) {
    __shared__ ReprojectShared sm;
    const int count = hdr->count, total = hdr->totalBlocks;
    if (count <= 0 || count > LUXGI_MAX_DENOISE_BATCH) return;
    if (ONE) {  // units in XCD order: each XCD takes a contiguous run of units
        const long long u = xcd_block(blockIdx.x, gridDim.x);
        if (u < total) batch_reproject_unit<S>(items, count, (int)u, sm);
        return;
    }
    const BatchWalk wk = batch_walk(total, order);
    for (int w = wk.begin; w < wk.end; w += wk.stride) batch_reproject_unit<S>(items, count, w, sm);
}

// LDS side of the batched a-trous at iteration i: the LDS-tiled items stage (16 + 2h)^2 texels with
// h = max(radius * 2^i, 1) <= 8 and radius <= 2
__host__ __device__ inline int batch_atrous_side(int iteration) {
    const int step = 1 << iteration;
    const int h = 2 * step <= kAtrousMaxHalo ? 2 * step : kAtrousMaxHalo;
    return kAtrousBlock + 2 * (h > 1 ? h : 1);
}

// iteration i of every item: input outputs[0] (i = 0) or atrous[readIdx], output atrous[writeIdx],
// the history copy into outputs[1] after iteration 1 (RaytracedReflection.cpp:419-494)
template <class S>
__device__ __forceinline__ void batch_atrous_unit(const Item* __restrict__ items, int count, int w, int iteration,
                                                  int readIdx, int writeIdx, float4* sIn, float4* sGuide) {
    {
        const Item* it;
        int bX, bY;
        if (!batch_unit(items, count, w, it, bX, bY)) return;
        const LuxgiReflectionAtrousParams& f = it->filter;
        const int radius = min(max(f.radius, 0), 2);
        const AtrousArgs A{it->width, it->height, (it->width + 7) / 8, radius, 1 << iteration, f.phiColor, f.phiNormal,
                           f.sigmaDepth, 0.0f, normal_squarings(f.phiNormal)};
        const S* in = static_cast<const S*>(iteration == 0 ? it->outputs[0] : it->atrous[readIdx]);
        S* out = static_cast<S*>(it->atrous[writeIdx]);
        S* hist = iteration == 1 ? static_cast<S*>(it->outputs[1]) : nullptr;
        const float4* nrm = static_cast<const float4*>(static_cast<const void*>(it->cur.normals));
        const float4* pbr = static_cast<const float4*>(static_cast<const void*>(it->cur.pbr));
        if (A.stepSize <= kAtrousMaxHalo && A.radius * A.stepSize <= kAtrousMaxHalo) {
            __syncthreads();  // the previous unit's LDS reads are done
            atrous_lds_block<S>(A, nrm, pbr, it->tileMap, in, out, hist, bX, bY, sIn, sGuide);
        } else {
            const int x = bX * 16 + (threadIdx.x & 15), y = bY * 16 + (threadIdx.x >> 4);
            if (x < A.W && y < A.H) atrous_direct_pixel<S>(A, nrm, pbr, it->tileMap, in, out, hist, x, y);
        }
    }
}

template <class S, bool ONE>
__global__ __launch_bounds__(256) void batch_atrous_kernel(const LuxgiDenoiseBatchHeader* __restrict__ hdr,
                                                           const Item* __restrict__ items, int iteration, int readIdx,
                                                           int writeIdx, int order) {
    // LDS sized by the launch for the largest radius (2) at this iteration's step (batch_atrous_side)
    extern __shared__ float4 batchAtrousLds[];
    const int side = batch_atrous_side(iteration);
    float4* sIn = batchAtrousLds;
    float4* sGuide = batchAtrousLds + side * side;
    const int count = hdr->count, total = hdr->totalBlocks;
    if (count <= 0 || count > LUXGI_MAX_DENOISE_BATCH) return;
    if (ONE) {  // units in XCD order, as above
        const long long u = xcd_block(blockIdx.x, gridDim.x);
        if (u < total) batch_atrous_unit<S>(items, count, (int)u, iteration, readIdx, writeIdx, sIn, sGuide);
        return;
    }
    const BatchWalk wk = batch_walk(total, order);
    for (int w = wk.begin; w < wk.end; w += wk.stride)
        batch_atrous_unit<S>(items, count, w, iteration, readIdx, writeIdx, sIn, sGuide);
}
static bool gbuf_ok(const LuxgiGBuffer* g) { return g && g->color && g->normals && g->pbr && g->depth; }

}  // namespace luxgi

using namespace luxgi;

// Footprints within the LDS halo go through the tiled kernel (16x16 pixels, 256 lanes); larger
// steps through the direct kernel (8x8 pixels, 64 lanes).
template <class S, class K>
static void launch_atrous(const AtrousArgs& A, K ldsKernel, K directKernel, hipStream_t st, const float4* nrm,
                          const float4* pbr, const uint8_t* tileMap, const void* input, void* output, void* hist,
                          size_t ldsBytesPerTexel) {
    const S* in = static_cast<const S*>(input);
    S* out = static_cast<S*>(output);
    S* h = static_cast<S*>(hist);
    if (A.stepSize <= kAtrousMaxHalo && A.radius * A.stepSize <= kAtrousMaxHalo) {
        const int side = kAtrousBlock + 2 * std::max(A.radius * A.stepSize, 1);
        hipLaunchKernelGGL(ldsKernel, dim3((A.W + kAtrousBlock - 1) / kAtrousBlock, (A.H + kAtrousBlock - 1) / kAtrousBlock),
                           dim3(256), (size_t)side * side * ldsBytesPerTexel, st, A, nrm, pbr, tileMap, in, out, h);
    } else {
        hipLaunchKernelGGL(directKernel, dim3((A.W + 7) / 8, (A.H + 7) / 8), dim3(64), 0, st, A, nrm, pbr, tileMap, in,
                           out, h);
    }
}

extern "C" int32_t luxgi_svgf_tile_count(int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return 0;
    return ((width + 7) / 8) * ((height + 7) / 8);
}

template <class S>
static int refl_reproject(const LuxgiReflectionReprojectParams* p, int W, int H, const LuxgiGBuffer* cur,
                          const LuxgiGBuffer* prev, const void* noisy, const void* hc, const void* hm, void* oc,
                          void* om, uint8_t* tiles, hipStream_t st) {
    ReflArgs A{*p, W, H, (W + 7) / 8};
    hipLaunchKernelGGL(reflection_reproject_kernel<S>, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, st, A,
                       to_gbuf(cur), to_gbuf(prev), static_cast<const S*>(noisy), static_cast<const S*>(hc),
                       static_cast<const S*>(hm), static_cast<S*>(oc), static_cast<S*>(om), tiles);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_reflection_reproject(const LuxgiReflectionReprojectParams* p, int32_t W, int32_t H,
                                               int32_t storage, const LuxgiGBuffer* cur, const LuxgiGBuffer* prev,
                                               const void* noisy, const void* historyColor, const void* historyMoments,
                                               void* outColor, void* outMoments, uint8_t* tileMap, void* stream) {
    if (!p || W <= 0 || H <= 0 || !gbuf_ok(cur) || !gbuf_ok(prev) || !noisy || !historyColor || !historyMoments ||
        !outColor || !outMoments || !tileMap)
        return LUXGI_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    if (storage == LUXGI_STORAGE_FP32)
        return refl_reproject<float>(p, W, H, cur, prev, noisy, historyColor, historyMoments, outColor, outMoments,
                                     tileMap, st);
    if (storage == LUXGI_STORAGE_FP16)
        return refl_reproject<__half>(p, W, H, cur, prev, noisy, historyColor, historyMoments, outColor, outMoments,
                                      tileMap, st);
    return LUXGI_ERR_UNSUPPORTED;
}

static int refl_atrous(const LuxgiReflectionAtrousParams* p, int32_t W, int32_t H, int32_t storage,
                       const LuxgiGBuffer* cur, const uint8_t* tileMap, const void* input, void* output,
                       void* historyOut, const float* lumScale, void* stream) {
    if (!p || W <= 0 || H <= 0 || !cur || !cur->normals || !cur->pbr || !tileMap || !input || !output)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (p->radius < 0 || p->radius > 2) return LUXGI_ERR_UNSUPPORTED;  // kernelWeights has 3 entries
    if (p->stepSize < 1 || p->stepSize > (1 << 16)) return LUXGI_ERR_INVALID_ARGUMENT;
    AtrousArgs A{W, H, (W + 7) / 8, p->radius, p->stepSize, p->phiColor, p->phiNormal, p->sigmaDepth, 0.0f,
                 normal_squarings(p->phiNormal), lumScale};
    if (storage != LUXGI_STORAGE_FP32 && storage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    hipStream_t st = as_stream(stream);
    const float4* nrm = static_cast<const float4*>(static_cast<const void*>(cur->normals));
    const float4* pbr = static_cast<const float4*>(static_cast<const void*>(cur->pbr));
    if (storage == LUXGI_STORAGE_FP32)
        launch_atrous<float>(A, reflection_atrous_lds_kernel<float>, reflection_atrous_kernel<float>, st, nrm, pbr,
                             tileMap, input, output, historyOut, 2 * sizeof(float4));
    else
        launch_atrous<__half>(A, reflection_atrous_lds_kernel<__half>, reflection_atrous_kernel<__half>, st, nrm, pbr,
                              tileMap, input, output, historyOut, 2 * sizeof(float4));
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_reflection_atrous(const LuxgiReflectionAtrousParams* p, int32_t W, int32_t H,
                                            int32_t storage, const LuxgiGBuffer* cur, const uint8_t* tileMap,
                                            const void* input, void* output, void* historyOut, void* stream) {
    return refl_atrous(p, W, H, storage, cur, tileMap, input, output, historyOut, nullptr, stream);
}

extern "C" int luxgi_svgf_reflection_atrous_scaled(const LuxgiReflectionAtrousParams* p, int32_t W, int32_t H,
                                                   int32_t storage, const LuxgiGBuffer* cur, const uint8_t* tileMap,
                                                   const void* input, void* output, void* historyOut,
                                                   const float* lumScale, void* stream) {
    if (!lumScale) return LUXGI_ERR_INVALID_ARGUMENT;
    return refl_atrous(p, W, H, storage, cur, tileMap, input, output, historyOut, lumScale, stream);
}

extern "C" int luxgi_svgf_reflection_denoise(const LuxgiReflectionReprojectParams* rp,
                                             const LuxgiReflectionAtrousParams* ap, int32_t iterations, int32_t W,
                                             int32_t H, int32_t storage, const LuxgiGBuffer* cur,
                                             const LuxgiGBuffer* prev, const void* noisy, void* output0, void* output1,
                                             void* moments0, void* moments1, int32_t pingPong, void* atrous0,
                                             void* atrous1, uint8_t* tileMap, int32_t* result, void* stream) {
    if (iterations < 0 || !atrous0 || !atrous1 || !moments0 || !moments1 || !output0 || !output1)
        return LUXGI_ERR_INVALID_ARGUMENT;
    void* mom[2] = {moments0, moments1};
    void* atr[2] = {atrous0, atrous1};
    // accumulation (RaytracedReflection.cpp:362-417)
    int rc = luxgi_svgf_reflection_reproject(rp, W, H, storage, cur, prev, noisy, output1, mom[1 - (pingPong & 1)],
                                             output0, mom[pingPong & 1], tileMap, stream);
    if (rc) return rc;
    // atrousFilter (RaytracedReflection.cpp:419-494)
    bool pp = false;
    int writeIdx = 1;
    for (int i = 0; i < iterations; ++i) {
        const int readIdx = pp ? 1 : 0;
        writeIdx = pp ? 0 : 1;
        LuxgiReflectionAtrousParams a = *ap;
        a.stepSize = 1 << i;
        rc = luxgi_svgf_reflection_atrous(&a, W, H, storage, cur, tileMap, i == 0 ? output0 : atr[readIdx],
                                          atr[writeIdx], i == 1 ? output1 : nullptr, stream);
        if (rc) return rc;
        pp = !pp;
    }
    if (result) *result = iterations > 0 ? writeIdx : -1;
    return LUXGI_OK;
}

template <class S>
static int shadow_reproject(const LuxgiShadowReprojectParams* p, int W, int H, const LuxgiGBuffer* cur,
                            const LuxgiGBuffer* prev, const uint32_t* mask, const void* hc, const void* hm, void* oc,
                            void* om, uint8_t* tiles, hipStream_t st) {
    ShadowArgs A{*p, W, H, (W + 7) / 8, (W + 7) / 8, (H + 3) / 4};
    hipLaunchKernelGGL(shadow_reproject_kernel<S>, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, st, A,
                       to_gbuf(cur), to_gbuf(prev), mask, static_cast<const S*>(hc), static_cast<const S*>(hm),
                       static_cast<S*>(oc), static_cast<S*>(om), tiles);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_shadow_reproject(const LuxgiShadowReprojectParams* p, int32_t W, int32_t H, int32_t storage,
                                           const LuxgiGBuffer* cur, const LuxgiGBuffer* prev, const uint32_t* mask,
                                           const void* historyColor, const void* historyMoments, void* outColor,
                                           void* outMoments, uint8_t* tileMap, void* stream) {
    if (!p || W <= 0 || H <= 0 || !gbuf_ok(cur) || !gbuf_ok(prev) || !mask || !historyColor || !historyMoments ||
        !outColor || !outMoments || !tileMap)
        return LUXGI_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    if (storage == LUXGI_STORAGE_FP32)
        return shadow_reproject<float>(p, W, H, cur, prev, mask, historyColor, historyMoments, outColor, outMoments,
                                       tileMap, st);
    if (storage == LUXGI_STORAGE_FP16)
        return shadow_reproject<__half>(p, W, H, cur, prev, mask, historyColor, historyMoments, outColor, outMoments,
                                        tileMap, st);
    return LUXGI_ERR_UNSUPPORTED;
}

extern "C" int luxgi_svgf_shadow_atrous(const LuxgiShadowAtrousParams* p, int32_t W, int32_t H, int32_t storage,
                                        const LuxgiGBuffer* cur, const uint8_t* tileMap, const void* input,
                                        void* output, void* historyOut, void* stream) {
    if (!p || W <= 0 || H <= 0 || !cur || !cur->normals || !cur->pbr || !tileMap || !input || !output)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if (p->radius < 0 || p->radius > 2) return LUXGI_ERR_UNSUPPORTED;
    if (p->stepSize < 1 || p->stepSize > (1 << 16)) return LUXGI_ERR_INVALID_ARGUMENT;
    AtrousArgs A{W, H, (W + 7) / 8, p->radius, p->stepSize, p->phiVisibility, p->phiNormal, p->sigmaDepth, p->power,
                 normal_squarings(p->phiNormal)};
    if (storage != LUXGI_STORAGE_FP32 && storage != LUXGI_STORAGE_FP16) return LUXGI_ERR_UNSUPPORTED;
    hipStream_t st = as_stream(stream);
    const float4* nrm = static_cast<const float4*>(static_cast<const void*>(cur->normals));
    const float4* pbr = static_cast<const float4*>(static_cast<const void*>(cur->pbr));
    if (storage == LUXGI_STORAGE_FP32)
        launch_atrous<float>(A, shadow_atrous_lds_kernel<float>, shadow_atrous_kernel<float>, st, nrm, pbr, tileMap,
                             input, output, historyOut, sizeof(float4) + sizeof(float2));
    else
        launch_atrous<__half>(A, shadow_atrous_lds_kernel<__half>, shadow_atrous_kernel<__half>, st, nrm, pbr, tileMap,
                              input, output, historyOut, sizeof(float4) + sizeof(float2));
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_shadow_denoise(const LuxgiShadowReprojectParams* rp, const LuxgiShadowAtrousParams* ap,
                                         int32_t iterations, int32_t W, int32_t H, int32_t storage,
                                         const LuxgiGBuffer* cur, const LuxgiGBuffer* prev, const uint32_t* mask,
                                         void* output0, void* output1, void* moments0, void* moments1,
                                         int32_t pingPong, void* atrous0, void* atrous1, uint8_t* tileMap,
                                         int32_t* result, void* stream) {
    if (!ap || iterations < 0 || !atrous0 || !atrous1 || !moments0 || !moments1 || !output0 || !output1)
        return LUXGI_ERR_INVALID_ARGUMENT;
    void* mom[2] = {moments0, moments1};
    void* atr[2] = {atrous0, atrous1};
    // accumulation (RaytracedShadow.cpp:418-462): outColor = outputs[0], history = outputs[1],
    // moments written to currentMoments[pingPong], read from currentMoments[1 - pingPong]
    int rc = luxgi_svgf_shadow_reproject(rp, W, H, storage, cur, prev, mask, output1, mom[1 - (pingPong & 1)], output0,
                                         mom[pingPong & 1], tileMap, stream);
    if (rc) return rc;
    // atrousFilter (RaytracedShadow.cpp:464-541): power only on the last iteration (:510),
    // the output of iteration 1 copied into the history colour (:532-535)
    bool pp = false;
    int writeIdx = 1;
    for (int i = 0; i < iterations; ++i) {
        const int readIdx = pp ? 1 : 0;
        writeIdx = pp ? 0 : 1;
        LuxgiShadowAtrousParams a = *ap;
        a.stepSize = 1 << i;
        a.power = i == iterations - 1 ? ap->power : 0.0f;
        rc = luxgi_svgf_shadow_atrous(&a, W, H, storage, cur, tileMap, i == 0 ? output0 : atr[readIdx], atr[writeIdx],
                                      i == 1 ? output1 : nullptr, stream);
        if (rc) return rc;
        pp = !pp;
    }
    if (result) *result = iterations > 0 ? writeIdx : -1;
    return LUXGI_OK;
}

extern "C" int luxgi_svgf_batch_prepare(LuxgiReflectionDenoiseItem* items, int32_t count,
                                        LuxgiDenoiseBatchHeader* header) {
    if (!items || !header || count <= 0 || count > LUXGI_MAX_DENOISE_BATCH) return LUXGI_ERR_INVALID_ARGUMENT;
    long long total = 0;
    for (int k = 0; k < count; ++k) {
        LuxgiReflectionDenoiseItem& it = items[k];
        const int W = it.width, H = it.height;
        if (W <= 0 || H <= 0 || (long long)W * H >= (1ll << 31)) return LUXGI_ERR_INVALID_ARGUMENT;
        if (it.filter.radius < 0 || it.filter.radius > 2) return LUXGI_ERR_UNSUPPORTED;
        if (!gbuf_ok(&it.cur) || !gbuf_ok(&it.prev) || !it.noisy || !it.tileMap) return LUXGI_ERR_INVALID_ARGUMENT;
        for (int j = 0; j < 2; ++j)
            if (!it.outputs[j] || !it.moments[j] || !it.atrous[j]) return LUXGI_ERR_INVALID_ARGUMENT;
        it.blockStart = (int32_t)total;
        total += (long long)((W + 15) / 16) * ((H + 15) / 16);
        if (total >= (1ll << 31)) return LUXGI_ERR_INVALID_ARGUMENT;
    }
    header->count = count;
    header->totalBlocks = (int32_t)total;
    header->pad[0] = header->pad[1] = 0;
    return LUXGI_OK;
}

template <class S>
static void launch_batched(const LuxgiDenoiseBatchHeader* hdr, const LuxgiReflectionDenoiseItem* items, int iterations,
                           int grid, int order, hipStream_t st) {
    if (order == 2)
        hipLaunchKernelGGL((batch_reproject_kernel<S, true>), dim3(grid), dim3(256), 0, st, hdr, items, order);
    else
        hipLaunchKernelGGL((batch_reproject_kernel<S, false>), dim3(grid), dim3(256), 0, st, hdr, items, order);
    bool pp = false;
    for (int i = 0; i < iterations; ++i) {
        const int readIdx = pp ? 1 : 0, writeIdx = pp ? 0 : 1;
        const int side = batch_atrous_side(i);
        const size_t lds = (size_t)side * side * 2 * sizeof(float4);
        if (order == 2)
            hipLaunchKernelGGL((batch_atrous_kernel<S, true>), dim3(grid), dim3(256), lds, st, hdr, items, i, readIdx,
                               writeIdx, order);
        else
            hipLaunchKernelGGL((batch_atrous_kernel<S, false>), dim3(grid), dim3(256), lds, st, hdr, items, i, readIdx,
                               writeIdx, order);
        pp = !pp;
    }
}

extern "C" int luxgi_svgf_reflection_denoise_batched(const LuxgiDenoiseBatchHeader* header,
                                                     const LuxgiReflectionDenoiseItem* items, int32_t iterations,
                                                     int32_t storage, int32_t gridBlocks, int32_t order, int32_t* result,
                                                     void* stream) {
    if (!header || !items || iterations < 0 || iterations > 16 || gridBlocks <= 0 || gridBlocks > (1 << 20) ||
        order < 0 || order > 2)
        return LUXGI_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    if (storage == LUXGI_STORAGE_FP32)
        launch_batched<float>(header, items, iterations, gridBlocks, order, st);
    else if (storage == LUXGI_STORAGE_FP16)
        launch_batched<__half>(header, items, iterations, gridBlocks, order, st);
    else
        return LUXGI_ERR_UNSUPPORTED;
    LUXGI_LAUNCH_CHECK();
    if (result) *result = iterations > 0 ? ((iterations & 1) ? 1 : 0) : -1;
    return LUXGI_OK;
}

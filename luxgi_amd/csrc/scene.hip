// scene.hip -- synthetic inputs for tests and benchmarks (not on the per-frame path):
// analytic-primitive SDF baking and a G-buffer + noisy signal from SDF primary rays.
#include "common.hpp"
#include "../../include/luxgi_scene.h"

namespace luxgi {

__device__ __forceinline__ float prim_distance(const LuxgiPrimitive& pr, f3 p) {
    const f3 q0 = p - mk(pr.center[0], pr.center[1], pr.center[2]);
    if (pr.type == LUXGI_PRIM_BOX) {
        const f3 q = mk(fabsf(q0.x) - pr.size[0], fabsf(q0.y) - pr.size[1], fabsf(q0.z) - pr.size[2]);
        const f3 m = mk(fmaxf(q.x, 0.f), fmaxf(q.y, 0.f), fmaxf(q.z, 0.f));
        return len3(m) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
    }
    if (pr.type == LUXGI_PRIM_SPHERE) return len3(q0) - pr.size[0];
    // capsule along axis size[2] with half length size[1]
    const int ax = int(pr.size[2]);
    f3 c = mk(0.f, 0.f, 0.f);  // closest point of the segment on the axis
    if (ax == 0) c.x = clampf(q0.x, -pr.size[1], pr.size[1]);
    else if (ax == 1) c.y = clampf(q0.y, -pr.size[1], pr.size[1]);
    else c.z = clampf(q0.z, -pr.size[1], pr.size[1]);
    return len3(q0 - c) - pr.size[0];
}

__global__ void bake_sdf_kernel(const LuxgiPrimitive* __restrict__ prims, int n, int res, float cx, float cy, float cz,
                                float half, float* __restrict__ out) {
    __shared__ LuxgiPrimitive sp[128];
    const int x = blockIdx.x * 8 + (threadIdx.x & 7);
    const int y = blockIdx.y * 8 + ((threadIdx.x >> 3) & 7);
    const int z = blockIdx.z * 4 + (threadIdx.x >> 6);
    const float voxel = 2.0f * half / float(res);
    const f3 p = mk(cx - half + (float(x) + 0.5f) * voxel, cy - half + (float(y) + 0.5f) * voxel,
                    cz - half + (float(z) + 0.5f) * voxel);
    float d = 2.0f * half;  // clamp value 1 after normalisation
    for (int base = 0; base < n; base += 128) {
        __syncthreads();
        for (int i = threadIdx.x; i < 128 && base + i < n; i += blockDim.x) sp[i] = prims[base + i];
        __syncthreads();
        const int cnt = min(128, n - base);
        for (int i = 0; i < cnt; ++i) d = fminf(d, prim_distance(sp[i], p));
    }
    if (x < res && y < res && z < res)
        out[((size_t)z * res + y) * res + x] = clampf(d / (2.0f * half), -1.0f, 1.0f);
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ float hash01(uint32_t x) { return float(hash32(x) >> 8) * (1.0f / 16777216.0f); }

// directionToOctohedral (Common/Math.glsl:20-25)
__device__ __forceinline__ float2 dir_to_oct(f3 n) {
    const float l1 = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float2 p = make_float2(n.x / l1, n.y / l1);
    if (n.z > 0.0f) return p;
    return make_float2((1.0f - fabsf(p.y)) * (p.x >= 0.0f ? 1.0f : -1.0f),
                       (1.0f - fabsf(p.x)) * (p.y >= 0.0f ? 1.0f : -1.0f));
}

struct GArgs {
    LuxgiGlobalSDFData g;
    float vpi[16], pvp[16];
    int res, W, H, n;
    uint32_t seed;
};

__global__ __launch_bounds__(256) void gbuffer_kernel(GArgs A, const LuxgiPrimitive* __restrict__ prims,
                                                      const float* __restrict__ sdf, float4* __restrict__ color,
                                                      float4* __restrict__ normals, float4* __restrict__ pbr,
                                                      float* __restrict__ depthOut, float4* __restrict__ noisy) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= A.W || y >= A.H) return;
    const size_t e = (size_t)y * A.W + x;
    const float nx = (float(x) + 0.5f) / float(A.W) * 2.0f - 1.0f, ny = (float(y) + 0.5f) / float(A.H) * 2.0f - 1.0f;
    const float4 a = mat4_mul4(A.vpi, make_float4(nx, ny, 0.0f, 1.0f));
    const float4 b = mat4_mul4(A.vpi, make_float4(nx, ny, 1.0f, 1.0f));
    const f3 o = mk(a.x / a.w, a.y / a.w, a.z / a.w);
    const f3 far = mk(b.x / b.w, b.y / b.w, b.z / b.w);
    const f3 dir = normalize3(far - o);
    const float tMax = len3(far - o);
    const f3 c = mk(A.g.cascadePosDistance[0][0], A.g.cascadePosDistance[0][1], A.g.cascadePosDistance[0][2]);
    const float half = A.g.cascadePosDistance[0][3];
    const float voxel = A.g.cascadeVoxelSize[0];
    // clip the ray against the volume
    const f3 inv = mk(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    const f3 t0 = (c - mk(half, half, half) - o) * inv, t1 = (c + mk(half, half, half) - o) * inv;
    float tn = fmaxf(fmaxf(fminf(t0.x, t1.x), fminf(t0.y, t1.y)), fminf(t0.z, t1.z));
    float tf = fminf(fminf(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y)), fmaxf(t0.z, t1.z));
    tn = fmaxf(tn, 0.0f);
    tf = fminf(tf, tMax);
    float t = tn;
    bool hit = false;
    for (int s = 0; s < 512 && t < tf; ++s) {
        const f3 p = o + dir * t;
        const float u = clampf((p.x - c.x) / (2.0f * half) + 0.5f, 0.f, 1.f);
        const float v = clampf((p.y - c.y) / (2.0f * half) + 0.5f, 0.f, 1.f);
        const float w = clampf((p.z - c.z) / (2.0f * half) + 0.5f, 0.f, 1.f);
        const float d = sample3d<float>(sdf, A.res, A.res, A.res, u, v, w) * 2.0f * half;
        if (d < 0.25f * voxel) {
            hit = true;
            break;
        }
        t += fmaxf(d, 0.25f * voxel);
    }
    if (!hit) {
        color[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        normals[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        pbr[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        depthOut[e] = 1.0f;
        noisy[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const f3 p = o + dir * t;
    // nearest primitive -> meshId, albedo, roughness
    int best = 0;
    float bd = 1e30f;
    for (int i = 0; i < A.n; ++i) {
        const float d = fabsf(prim_distance(prims[i], p));
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    const LuxgiPrimitive pr = prims[best];
    const float h = 0.5f * voxel;
    const f3 nrm = normalize3(mk(prim_distance(pr, p + mk(h, 0, 0)) - prim_distance(pr, p - mk(h, 0, 0)),
                                 prim_distance(pr, p + mk(0, h, 0)) - prim_distance(pr, p - mk(0, h, 0)),
                                 prim_distance(pr, p + mk(0, 0, h)) - prim_distance(pr, p - mk(0, 0, h))));
    // NDC depth of the hit: unproject(nx, ny, d) = (a + d (b - a)).xyz / (a + d (b - a)).w
    // is solved for d on the axis where the ray moves most.
    float ndcDepth;
    {
        const float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
        const int k = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
        const float pk = k == 0 ? p.x : (k == 1 ? p.y : p.z);
        const float ak = k == 0 ? a.x : (k == 1 ? a.y : a.z);
        const float bk = k == 0 ? b.x : (k == 1 ? b.y : b.z);
        ndcDepth = clampf((pk * a.w - ak) / ((bk - ak) - pk * (b.w - a.w)), 0.0f, 0.99999f);
    }
    const float4 pc = mat4_mul4(A.pvp, make_float4(p.x, p.y, p.z, 1.0f));
    const float pu = (pc.x / pc.w) * 0.5f + 0.5f, pv = (pc.y / pc.w) * 0.5f + 0.5f;
    const float cu = (float(x) + 0.5f) / float(A.W), cv = (float(y) + 0.5f) / float(A.H);
    const float2 on = dir_to_oct(nrm);
    color[e] = make_float4(pr.albedo[0], pr.albedo[1], pr.albedo[2], float(best));
    normals[e] = make_float4(on.x, on.y, pu - cu, pv - cv);
    pbr[e] = make_float4(0.0f, pr.roughness, 0.0f, t);
    depthOut[e] = ndcDepth;
    // 1-spp noisy radiance: albedo * (lighting with a random visibility term) + emissive
    const uint32_t hs = A.seed * 0x9e3779b9u + (uint32_t)e * 0x85ebca6bu;
    const float ndl = fmaxf(0.0f, nrm.x * 0.3f + nrm.y * 0.8f + nrm.z * 0.52f);
    const float vis = hash01(hs) < 0.75f ? 1.0f : 0.0f;
    const float jitter = 0.5f + hash01(hs ^ 0x68bc21ebu);
    const float L = (ndl * vis + 0.1f) * jitter + pr.emissive;
    noisy[e] = make_float4(pr.albedo[0] * L, pr.albedo[1] * L, pr.albedo[2] * L, 0.05f + 0.5f * hash01(hs ^ 0x2c1b3c6du));
}

// 1-bit-per-pixel shadow visibility mask, one 8x4 word per thread
__global__ void shadow_mask_kernel(int W, int H, uint32_t seed, const float* __restrict__ depth,
                                   const float4* __restrict__ normals, uint32_t* __restrict__ mask) {
    const int mw = (W + 7) / 8, mh = (H + 3) / 4;
    const int wx = blockIdx.x * blockDim.x + threadIdx.x, wy = blockIdx.y;
    if (wx >= mw || wy >= mh) return;
    uint32_t word = 0;
    for (int b = 0; b < 32; ++b) {
        const int x = wx * 8 + (b & 7), y = wy * 4 + (b >> 3);
        if (x >= W || y >= H) continue;
        const size_t e = (size_t)y * W + x;
        if (depth[e] == 1.0f) continue;
        const float4 n = normals[e];
        const float p = 0.35f + 0.5f * fabsf(n.x);
        if (hash01(seed * 0x27d4eb2fu + (uint32_t)e * 0x165667b1u) < p) word |= 1u << b;
    }
    mask[(size_t)wy * mw + wx] = word;
}

}  // namespace luxgi

using namespace luxgi;

extern "C" int luxgi_scene_bake_sdf(const LuxgiPrimitive* prims, int32_t n, int32_t res, const float center[3],
                                    float halfExtent, float* out, void* stream) {
    if (!prims || n <= 0 || res <= 0 || !center || !out || halfExtent <= 0.f) return LUXGI_ERR_INVALID_ARGUMENT;
    const dim3 grid((res + 7) / 8, (res + 7) / 8, (res + 3) / 4);
    hipLaunchKernelGGL(bake_sdf_kernel, grid, dim3(256), 0, as_stream(stream), prims, n, res, center[0], center[1],
                       center[2], halfExtent, out);
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

extern "C" int luxgi_scene_gbuffer(const LuxgiPrimitive* prims, int32_t n, const LuxgiGlobalSDFData* g,
                                   const float* sdf, int32_t res, const float viewProjInv[16],
                                   const float prevViewProj[16], int32_t W, int32_t H, uint32_t seed, float* color,
                                   float* normals, float* pbr, float* depth, float* noisy, uint32_t* shadowMask,
                                   void* stream) {
    if (!prims || n <= 0 || !g || !sdf || res <= 0 || !viewProjInv || !prevViewProj || W <= 0 || H <= 0 || !color ||
        !normals || !pbr || !depth || !noisy)
        return LUXGI_ERR_INVALID_ARGUMENT;
    if ((size_t)res * res * res * sizeof(float) >= (1ull << 32)) return LUXGI_ERR_UNSUPPORTED;  // 32-bit sampler offsets
    GArgs A;
    A.g = *g;
    for (int i = 0; i < 16; ++i) {
        A.vpi[i] = viewProjInv[i];
        A.pvp[i] = prevViewProj[i];
    }
    A.res = res;
    A.W = W;
    A.H = H;
    A.n = n;
    A.seed = seed;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(gbuffer_kernel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, st, A, prims, sdf,
                       reinterpret_cast<float4*>(color), reinterpret_cast<float4*>(normals),
                       reinterpret_cast<float4*>(pbr), depth, reinterpret_cast<float4*>(noisy));
    LUXGI_LAUNCH_CHECK();
    if (shadowMask) {
        const int mw = (W + 7) / 8, mh = (H + 3) / 4;
        hipLaunchKernelGGL(shadow_mask_kernel, dim3((mw + 63) / 64, mh), dim3(64), 0, st, W, H, seed, depth,
                           reinterpret_cast<const float4*>(normals), shadowMask);
        LUXGI_LAUNCH_CHECK();
    }
    return LUXGI_OK;
}

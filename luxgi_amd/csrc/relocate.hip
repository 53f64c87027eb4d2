// relocate.hip -- DDGI probe relocation and classification for gfx950.
//
// The reference has no probe relocation or classification (no code under
// Code/Maple/src; the only "inside geometry" signal is the inside-hit encode of
// Shaders/DDGI/GISDFRays.comp:93-96). The rules here follow Majercik et al.,
// "Scaling Probe-Based Real-Time Dynamic Global Illumination for Production",
// JCGT 10(2), 2021, restated for an SDF tracer whose probe rays report
//   flag 0 = miss, 1 = front-face hit at distance t, 2 = origin inside geometry:
//   inside fraction  f = #flag2 / R
//   classification   inactive if f > backfaceThreshold, or if no front-face hit
//                    point lies within one probe spacing of the probe on every
//                    axis (|d.x t| <= step.x && |d.y t| <= step.y && |d.z t| <= step.z);
//   relocation       when active and the closest front-face hit t_min is closer
//                    than minFrontfaceDistance, move the probe away from it:
//                    offset -= d_min * (minFrontfaceDistance - t_min), then clamp
//                    every axis to +-maxOffsetFraction * step.
// One wave per probe; the closest hit is a deterministic (distance, ray index) min.
#include "common.hpp"

namespace luxgi {

struct RelocArgs {
    float step[3];
    int P, R, rayWidth;
    LuxgiRelocateParams prm;
};

template <class SR>
__global__ __launch_bounds__(64) void relocate_kernel(RelocArgs A, const SR* __restrict__ dirDist,
                                                      const uint8_t* __restrict__ flags, float* __restrict__ offsets,
                                                      uint8_t* __restrict__ states) {
    const int p = blockIdx.x;
    if (p >= A.P) return;
    const int lane = threadIdx.x;
    int inside = 0, near = 0;
    float best = INFINITY;
    int bestRay = 0x7fffffff;
    for (int r = lane; r < A.R; r += 64) {
        const size_t e = (size_t)p * A.rayWidth + r;
        const uint8_t f = flags[e];
        if (f == 2) inside++;
        else if (f == 1) {
            const float4 d = Texel<SR>::ld4(dirDist, e);
            const float t = d.w;
            const float ax = fabsf(d.x * t), ay = fabsf(d.y * t), az = fabsf(d.z * t);
            if (ax <= A.step[0] && ay <= A.step[1] && az <= A.step[2]) near = 1;
            if (t < best || (t == best && r < bestRay)) {
                best = t;
                bestRay = r;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        inside += __shfl_xor(inside, o);
        near |= __shfl_xor(near, o);
        const float ob = __shfl_xor(best, o);
        const int orr = __shfl_xor(bestRay, o);
        if (ob < best || (ob == best && orr < bestRay)) {
            best = ob;
            bestRay = orr;
        }
    }
    if (lane != 0) return;
    const bool isInside = float(inside) > A.prm.backfaceThreshold * float(A.R);
    const bool active = !isInside && near;
    states[p] = active ? 0 : 1;
    if (A.prm.relocate && !isInside && bestRay != 0x7fffffff && best < A.prm.minFrontfaceDistance) {
        const float4 d = Texel<SR>::ld4(dirDist, (size_t)p * A.rayWidth + bestRay);
        const float push = A.prm.minFrontfaceDistance - best;
        float4 o = reinterpret_cast<float4*>(offsets)[p];
        o.x -= d.x * push;
        o.y -= d.y * push;
        o.z -= d.z * push;
        const float lim = A.prm.maxOffsetFraction;
        o.x = clampf(o.x, -lim * A.step[0], lim * A.step[0]);
        o.y = clampf(o.y, -lim * A.step[1], lim * A.step[1]);
        o.z = clampf(o.z, -lim * A.step[2], lim * A.step[2]);
        o.w = 0.0f;
        reinterpret_cast<float4*>(offsets)[p] = o;
    }
}

}  // namespace luxgi

using namespace luxgi;

extern "C" int luxgi_ddgi_relocate_classify(const LuxgiDDGIUniform* u, const LuxgiRayBuffers* rays,
                                            const uint8_t* rayFlags, const LuxgiRelocateParams* params, float* offsets,
                                            uint8_t* states, void* stream) {
    if (!u || !rays || !rays->directionDistance || !rayFlags || !params || !offsets || !states)
        return LUXGI_ERR_INVALID_ARGUMENT;
    RelocArgs A;
    for (int i = 0; i < 3; ++i) A.step[i] = u->step[i];
    A.P = u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    A.R = u->raysPerProbe;
    A.rayWidth = rays->width;
    A.prm = *params;
    if (A.P <= 0 || A.R <= 0 || A.R > rays->width || rays->height < A.P) return LUXGI_ERR_INVALID_ARGUMENT;
    hipStream_t st = as_stream(stream);
    if (rays->storage == LUXGI_STORAGE_FP32)
        hipLaunchKernelGGL(relocate_kernel<float>, dim3(A.P), dim3(64), 0, st, A,
                           static_cast<const float*>(rays->directionDistance), rayFlags, offsets, states);
    else if (rays->storage == LUXGI_STORAGE_FP16)
        hipLaunchKernelGGL(relocate_kernel<__half>, dim3(A.P), dim3(64), 0, st, A,
                           static_cast<const __half*>(rays->directionDistance), rayFlags, offsets, states);
    else return LUXGI_ERR_UNSUPPORTED;
    LUXGI_LAUNCH_CHECK();
    return LUXGI_OK;
}

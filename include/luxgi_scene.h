/*
 * luxgi_scene.h -- synthetic-input generators shipped with libluxgi.
 *
 * These entry points build the inputs the hot path consumes (a global SDF of
 * analytic primitives, a G-buffer from SDF primary rays with a noisy 1-spp
 * signal) for tests and benchmarks. They have no counterpart on the reference's
 * per-frame path: the reference rasterises baked mesh SDFs into the clipmap
 * (Shaders/SDF/SDFRasterizeModel.glsl:42-63) and renders its G-buffer with the
 * raster pipeline (Shaders/DeferredColor.frag:60-101).
 */
#ifndef LUXGI_SCENE_H
#define LUXGI_SCENE_H

#include "luxgi.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LUXGI_PRIM_BOX 0
#define LUXGI_PRIM_SPHERE 1
#define LUXGI_PRIM_CAPSULE 2

/* Analytic primitive, 48 bytes. size: box half extents; sphere (r, -, -);
 * capsule (r, half length, axis index 0/1/2). */
typedef struct LuxgiPrimitive {
    float   center[3];
    int32_t type;
    float   size[3];
    float   roughness;
    float   albedo[3];
    float   emissive;
} LuxgiPrimitive;

/* Global SDF of `n` primitives on a res^3 grid over [center-half, center+half]^3,
 * voxel centres at min + (i + 0.5) * voxel, value clamp(d / (2*half), -1, 1)
 * (the normalisation of Shaders/SDF/SDFRasterizeModel.glsl:61). out: [res][res][res] R32F. */
int luxgi_scene_bake_sdf(const LuxgiPrimitive* prims, int32_t n, int32_t res, const float center[3], float halfExtent,
                         float* out, void* stream);

/* G-buffer from primary rays through the global SDF (single cascade).
 * Pixel (x, y) unprojects ndc ((x+0.5)/W*2-1, (y+0.5)/H*2-1) at depths 0 and 1
 * with viewProjInv; on a hit P: depth = NDC z of P, motion = prevUV - curUV from
 * prevViewProj, normals.xy = octahedral SDF-gradient normal, color = (albedo, meshId),
 * pbr = (0, roughness, 0, clip w). noisy (RGBA32F) = 1-spp radiance (rgb, reflected ray
 * length), hashed from `seed`. Misses: depth 1 and zero channels. */
int luxgi_scene_gbuffer(const LuxgiPrimitive* prims, int32_t n, const LuxgiGlobalSDFData* sdfData, const float* sdf,
                        int32_t res, const float viewProjInv[16], const float prevViewProj[16], int32_t width,
                        int32_t height, uint32_t seed, float* color, float* normals, float* pbr, float* depth,
                        float* noisy, uint32_t* shadowMask, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LUXGI_SCENE_H */

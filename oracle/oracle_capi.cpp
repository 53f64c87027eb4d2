// oracle_capi.cpp -- extern "C" entry points of the CPU restatement (TEST INFRASTRUCTURE).
//
// Loaded only by tests/, the smoke test and bench.py's CPU-baseline leg
// (through oracle/__init__.py). Each *_f64 function is the FP64 oracle, each
// *_f32 function the FP32 CPU transcription used as the CPU baseline.
// All buffers are host memory; images are float arrays, row-major.
#include "luxgi_oracle.hpp"

using namespace oracle;

namespace {

template <class T>
void probe_update_impl(const LuxgiDDGIUniform* u, const float* radiance, const float* dirDist, int rayWidth,
                       const float* prevIrr, const float* prevDepth, T* outIrr, T* outDepth, int firstFrame,
                       unsigned flags, int threads) {
    const int P = u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    const bool doIrr = flags & LUXGI_UPDATE_IRRADIANCE, doDepth = flags & LUXGI_UPDATE_DEPTH;
    parallel_for(P, threads, [&](int p) {
        probe_update_probe<T>(*u, radiance, dirDist, rayWidth, prevIrr, prevDepth, outIrr, outDepth, firstFrame != 0,
                              doIrr, doDepth, p);
    });
    if (flags & LUXGI_UPDATE_FUSE_BORDER) {
        if (doIrr)
            border_update<T>(outIrr, u->irradianceTextureWidth, u->irradianceTextureHeight,
                             u->irradianceProbeSideLength, 4, u->probeCounts[0], u->probeCounts[1], u->probeCounts[2]);
        if (doDepth)
            border_update<T>(outDepth, u->depthTextureWidth, u->depthTextureHeight, u->depthProbeSideLength, 2,
                             u->probeCounts[0], u->probeCounts[1], u->probeCounts[2]);
    }
}

struct TraceArgs {
    const LuxgiDDGIUniform* u;
    const LuxgiGlobalSDFData* g;
    const float* sdf;
    const float* mip;
    int res, mipRes;
    const LuxgiSurfaceAtlasData* ad;
    const uint32_t* chunks;
    const uint32_t* culled;
    const LuxgiAtlasObject* objects;
    const LuxgiAtlasTile* tiles;
    const float* atlasDepth;
    const float* atlasColor;  // float4
    const float* skyFaces;
    int skySize;
    const LuxgiTraceParams* params;
    const float* probeOffsets;
};

template <class T>
void trace_impl(const TraceArgs& a, int probeBegin, int probeEnd, int rayWidth, T* outRad, T* outDD, uint8_t* flags,
                unsigned long long* counts, int threads) {
    const int R = a.u->raysPerProbe;
    const int nP = probeEnd - probeBegin;
    std::vector<TraceCounts> tc(counts ? nP : 0);
    parallel_for(nP, threads, [&](int i) {
        int p = probeBegin + i;
        TraceCounts* c = counts ? &tc[i] : nullptr;
        for (int r = 0; r < R; ++r) {
            size_t e = (size_t)i * rayWidth + r;
            trace_ray<T>(*a.u, *a.g, a.sdf, a.mip, a.res, a.mipRes, *a.ad, a.chunks, a.culled, a.objects, a.tiles,
                         a.atlasDepth, a.atlasColor, a.skyFaces, a.skySize, *a.params, a.probeOffsets, p, r,
                         outRad + e * 4, outDD + e * 4, flags ? flags + e : nullptr, c);
        }
    });
    if (counts) {
        for (auto& c : tc) {
            counts[0] += c.mip; counts[1] += c.fine; counts[2] += c.normals; counts[3] += c.objects;
            counts[4] += c.tiles; counts[5] += c.rays; counts[6] += c.steps; counts[7] += c.hits;
        }
    }
}

GB make_gb(const float* const* g) { return GB{{g[0], 4}, {g[1], 4}, {g[2], 4}, {g[3], 1}}; }

template <class T>
void refl_reproject_impl(const LuxgiReflectionReprojectParams* pc, int W, int H, const float* const* cur,
                         const float* const* prev, const float* noisy, const float* histOut, const float* histMom,
                         T* outColor, T* outMom, uint8_t* tileMap, int threads) {
    GB c = make_gb(cur), p = make_gb(prev);
    const int tw = (W + 7) / 8, th = (H + 7) / 8;
    std::vector<uint8_t> vote((size_t)W * H);
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            size_t e = (size_t)y * W + x;
            vote[e] = reflection_reproject_pixel<T>(*pc, W, H, x, y, c, p, {noisy, 4}, {histOut, 4}, {histMom, 4},
                                                     outColor + e * 4, outMom + e * 4);
        }
    });
    for (int t = 0; t < tw * th; ++t) tileMap[t] = 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            if (vote[(size_t)y * W + x]) tileMap[(y / 8) * tw + x / 8] = 1;
}

template <class T>
void refl_atrous_impl(const LuxgiReflectionAtrousParams* pc, int W, int H, const float* normals, const float* pbr,
                      const uint8_t* tileMap, const float* in, T* out, int threads) {
    const int tw = (W + 7) / 8;
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            size_t e = (size_t)y * W + x;
            if (tileMap[(y / 8) * tw + x / 8]) {
                reflection_atrous_pixel<T>(*pc, W, H, x, y, {normals, 4}, {pbr, 4}, {in, 4}, out + e * 4);
            } else {
                for (int c = 0; c < 4; ++c) out[e * 4 + c] = T(in[e * 4 + c]);
            }
        }
    });
}

template <class T>
void shadow_reproject_impl(const LuxgiShadowReprojectParams* pc, int W, int H, const float* const* cur,
                           const float* const* prev, const uint32_t* mask, const float* histOut,
                           const float* histMom, T* outColor, T* outMom, uint8_t* tileMap, int threads) {
    GB c = make_gb(cur), p = make_gb(prev);
    const int tw = (W + 7) / 8, th = (H + 7) / 8;
    std::vector<uint8_t> vote((size_t)W * H);
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            size_t e = (size_t)y * W + x;
            vote[e] = shadow_reproject_pixel<T>(*pc, W, H, x, y, c, p, mask, {histOut, 2}, {histMom, 4},
                                                 outColor + e * 2, outMom + e * 4);
        }
    });
    for (int t = 0; t < tw * th; ++t) tileMap[t] = 0;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x)
            if (vote[(size_t)y * W + x]) tileMap[(y / 8) * tw + x / 8] = 1;
}

template <class T>
void shadow_atrous_impl(const LuxgiShadowAtrousParams* pc, int W, int H, const float* normals, const float* pbr,
                        const uint8_t* tileMap, const float* in, T* out, int threads) {
    const int tw = (W + 7) / 8;
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            size_t e = (size_t)y * W + x;
            if (tileMap[(y / 8) * tw + x / 8]) {
                shadow_atrous_pixel<T>(*pc, W, H, x, y, {normals, 4}, {pbr, 4}, {in, 2}, out + e * 2);
            } else {
                out[e * 2 + 0] = T(0);
                out[e * 2 + 1] = T(0);
            }
        }
    });
}

}  // namespace

extern "C" {

int oracle_border_table(int n, int* out, int cap) {
    std::vector<int> t = border_table(n);
    int cnt = (int)t.size() / 4;
    if (cap < cnt) return -cnt;
    std::memcpy(out, t.data(), t.size() * sizeof(int));
    return cnt;
}

void oracle_atlas_sizes(int oct, int px, int py, int pz, int* w, int* h) { atlas_sizes(oct, px, py, pz, w, h); }

int oracle_probe_id(int tx, int ty, int width, int side) { return get_probe_id(tx, ty, width, side); }

void oracle_probe_index_to_grid(const LuxgiDDGIUniform* u, int index, int* g) {
    probe_index_to_grid(*u, index, g, g + 1, g + 2);
}

int oracle_grid_to_probe_index(const LuxgiDDGIUniform* u, int x, int y, int z) { return grid_to_probe_index(*u, x, y, z); }

void oracle_probe_location(const LuxgiDDGIUniform* u, int index, double* out) {
    v3<double> p = probe_location<double>(*u, index);
    out[0] = p.x; out[1] = p.y; out[2] = p.z;
}

void oracle_oct_decode(double x, double y, double* out) {
    v3<double> v = oct_decode<double>({x, y});
    out[0] = v.x; out[1] = v.y; out[2] = v.z;
}

void oracle_oct_encode(const double* v, double* out) {
    v2<double> o = oct_encode<double>({v[0], v[1], v[2]});
    out[0] = o.x; out[1] = o.y;
}

void oracle_normalized_oct_coord(int fx, int fy, int side, double* out) {
    v2<double> o = normalized_oct_coord<double>(fx, fy, side);
    out[0] = o.x; out[1] = o.y;
}

void oracle_ray_directions(const float* rot, int rays, double* out) {
    for (int r = 0; r < rays; ++r) {
        v3<double> d = ray_direction<double>(rot, r, rays);
        out[r * 3 + 0] = d.x; out[r * 3 + 1] = d.y; out[r * 3 + 2] = d.z;
    }
}

void oracle_spherical_fibonacci(int i, int n, double* out) {
    v3<double> d = spherical_fibonacci<double>(double(i), double(n));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

void oracle_probe_update_f64(const LuxgiDDGIUniform* u, const float* radiance, const float* dirDist, int rayWidth,
                             const float* prevIrr, const float* prevDepth, double* outIrr, double* outDepth,
                             int firstFrame, unsigned flags, int threads) {
    probe_update_impl<double>(u, radiance, dirDist, rayWidth, prevIrr, prevDepth, outIrr, outDepth, firstFrame, flags,
                              threads);
}

void oracle_probe_update_f32(const LuxgiDDGIUniform* u, const float* radiance, const float* dirDist, int rayWidth,
                             const float* prevIrr, const float* prevDepth, float* outIrr, float* outDepth,
                             int firstFrame, unsigned flags, int threads) {
    probe_update_impl<float>(u, radiance, dirDist, rayWidth, prevIrr, prevDepth, outIrr, outDepth, firstFrame, flags,
                             threads);
}

void oracle_border_update_f64(double* atlas, int W, int H, int side, int ch, int px, int py, int pz) {
    border_update<double>(atlas, W, H, side, ch, px, py, pz);
}
void oracle_border_update_f32(float* atlas, int W, int H, int side, int ch, int px, int py, int pz) {
    border_update<float>(atlas, W, H, side, ch, px, py, pz);
}

#define TRACE_ARGS                                                                                                    \
    const LuxgiDDGIUniform *u, const LuxgiGlobalSDFData *g, const float *sdf, const float *mip, int res, int mipRes, \
        const LuxgiSurfaceAtlasData *ad, const uint32_t *chunks, const uint32_t *culled,                            \
        const LuxgiAtlasObject *objects, const LuxgiAtlasTile *tiles, const float *atlasDepth,                      \
        const float *atlasColor, const float *skyFaces, int skySize, const LuxgiTraceParams *params,                \
        const float *probeOffsets, int probeBegin, int probeEnd, int rayWidth
#define TRACE_PACK                                                                                                    \
    TraceArgs a{u, g, sdf, mip, res, mipRes, ad, chunks, culled, objects, tiles, atlasDepth, atlasColor, skyFaces,  \
                skySize, params, probeOffsets}

void oracle_trace_f64(TRACE_ARGS, double* outRad, double* outDD, uint8_t* flags, unsigned long long* counts,
                      int threads) {
    TRACE_PACK;
    trace_impl<double>(a, probeBegin, probeEnd, rayWidth, outRad, outDD, flags, counts, threads);
}

void oracle_trace_f32(TRACE_ARGS, float* outRad, float* outDD, uint8_t* flags, unsigned long long* counts,
                      int threads) {
    TRACE_PACK;
    trace_impl<float>(a, probeBegin, probeEnd, rayWidth, outRad, outDD, flags, counts, threads);
}

// Coarse mip of one cascade: initial 7-tap pass then `floodPasses` ping-pong passes
// (GlobalDistanceField.cpp:823-846 and :537-573). Arithmetic in double, storage float.
void oracle_sdf_mip(const float* sdf, int res, int cascades, int cascade, float maxDistance, int floodPasses,
                    float* mip, float* temp) {
    const int mr = (res + 3) / 4;
    mip_pass<double>(sdf, res * cascades, res, res, res, 4, cascade * res, mip, mr * cascades, mr, mr, cascade * mr,
                     maxDistance);
    for (int i = 1; i <= floodPasses; ++i) {
        if (i & 1)
            mip_pass<double>(mip, mr * cascades, mr, mr, mr, 1, cascade * mr, temp, mr, mr, mr, 0, maxDistance);
        else
            mip_pass<double>(temp, mr, mr, mr, mr, 1, 0, mip, mr * cascades, mr, mr, cascade * mr, maxDistance);
    }
}

// cur/prev: pointers {color, normals, pbr, depth}
void oracle_reflection_reproject_f64(const LuxgiReflectionReprojectParams* pc, int W, int H, const float* const* cur,
                                     const float* const* prev, const float* noisy, const float* histOut,
                                     const float* histMom, double* outColor, double* outMom, uint8_t* tileMap,
                                     int threads) {
    refl_reproject_impl<double>(pc, W, H, cur, prev, noisy, histOut, histMom, outColor, outMom, tileMap, threads);
}
void oracle_reflection_reproject_f32(const LuxgiReflectionReprojectParams* pc, int W, int H, const float* const* cur,
                                     const float* const* prev, const float* noisy, const float* histOut,
                                     const float* histMom, float* outColor, float* outMom, uint8_t* tileMap,
                                     int threads) {
    refl_reproject_impl<float>(pc, W, H, cur, prev, noisy, histOut, histMom, outColor, outMom, tileMap, threads);
}
void oracle_reflection_atrous_f64(const LuxgiReflectionAtrousParams* pc, int W, int H, const float* normals,
                                  const float* pbr, const uint8_t* tileMap, const float* in, double* out, int threads) {
    refl_atrous_impl<double>(pc, W, H, normals, pbr, tileMap, in, out, threads);
}
void oracle_reflection_atrous_f32(const LuxgiReflectionAtrousParams* pc, int W, int H, const float* normals,
                                  const float* pbr, const uint8_t* tileMap, const float* in, float* out, int threads) {
    refl_atrous_impl<float>(pc, W, H, normals, pbr, tileMap, in, out, threads);
}
void oracle_shadow_reproject_f64(const LuxgiShadowReprojectParams* pc, int W, int H, const float* const* cur,
                                 const float* const* prev, const uint32_t* mask, const float* histOut,
                                 const float* histMom, double* outColor, double* outMom, uint8_t* tileMap,
                                 int threads) {
    shadow_reproject_impl<double>(pc, W, H, cur, prev, mask, histOut, histMom, outColor, outMom, tileMap, threads);
}
void oracle_shadow_atrous_f64(const LuxgiShadowAtrousParams* pc, int W, int H, const float* normals,
                              const float* pbr, const uint8_t* tileMap, const float* in, double* out, int threads) {
    shadow_atrous_impl<double>(pc, W, H, normals, pbr, tileMap, in, out, threads);
}

int oracle_hardware_threads(void) { return (int)std::thread::hardware_concurrency(); }

double oracle_sample3d(const float* vol, int W, int H, int D, double u, double v, double w) {
    return sample3d<double>(vol, W, H, D, u, v, w);
}

}  // extern "C"

extern "C" void oracle_relocate(const LuxgiDDGIUniform* u, const float* dirDist, const uint8_t* flags, int width,
                                const LuxgiRelocateParams* prm, float* offsets, uint8_t* states) {
    const int P = u->probeCounts[0] * u->probeCounts[1] * u->probeCounts[2];
    for (int p = 0; p < P; ++p) relocate_probe(*u, dirDist, flags, width, *prm, offsets, states, p);
}

template <class T>
static void sample_probes_impl(const LuxgiDDGIUniform* u, const LuxgiSampleProbeParams* pc, int W, int H,
                               const float* depth, const float* normals, const float* irr, const float* dep,
                               const float* offsets, const uint8_t* states, T* out, int threads) {
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x)
            sample_probe_pixel<T>(*u, *pc, W, H, x, y, depth, normals, irr, dep, out + ((size_t)y * W + x) * 4,
                                  offsets, states);
    });
}
extern "C" void oracle_sample_probes_f64(const LuxgiDDGIUniform* u, const LuxgiSampleProbeParams* pc, int W, int H,
                                         const float* depth, const float* normals, const float* irr, const float* dep,
                                         const float* offsets, const uint8_t* states, double* out, int threads) {
    sample_probes_impl<double>(u, pc, W, H, depth, normals, irr, dep, offsets, states, out, threads);
}
extern "C" void oracle_sample_probes_f32(const LuxgiDDGIUniform* u, const LuxgiSampleProbeParams* pc, int W, int H,
                                         const float* depth, const float* normals, const float* irr, const float* dep,
                                         const float* offsets, const uint8_t* states, float* out, int threads) {
    sample_probes_impl<float>(u, pc, W, H, depth, normals, irr, dep, offsets, states, out, threads);
}
extern "C" void oracle_texture_coord_from_direction(const double* dir, int probeIndex, int width, int height, int side,
                                                    double* out) {
    v2<double> r = texture_coord_from_direction<double>({dir[0], dir[1], dir[2]}, probeIndex, width, height, side);
    out[0] = r.x;
    out[1] = r.y;
}

// SDFCulling.comp:36-101 over `objects` (LuxgiAtlasObject, objectBounds at float 0 of each 32-float record)
extern "C" void oracle_cull_chunks(float chunkSize, const LuxgiAtlasObject* objects, int n, uint32_t capacity,
                                   uint32_t* chunks, uint32_t* culled) {
    cull_chunks(chunkSize, n > 0 ? objects[0].objectBounds : nullptr, (int)(sizeof(LuxgiAtlasObject) / sizeof(float)),
                n, capacity, chunks, culled);
}

// Surface-cache lighting of the drawn tiles into out ([res][res][4], only drawn texels written).
template <class T>
static void surface_light_impl(const LuxgiSurfaceLightParams* prm, const LuxgiLight* lights, const LuxgiAtlasTile* tiles,
                               int numTiles, const LuxgiSurfaceDraw* draws, int numDraws, int res, const float* gbColor,
                               const float* gbNormal, const float* gbPbr, const float* gbEmissive,
                               const LuxgiGlobalSDFData* g, const float* sdf, const float* mip, int sdfRes, int mipRes,
                               const LuxgiDDGIUniform* u, const float* irr, const float* dep, const float* offsets,
                               const uint8_t* states, T* out, int threads) {
    parallel_for(numDraws, threads, [&](int d) {
        const LuxgiSurfaceDraw dr = draws[d];
        if (dr.tile >= (uint32_t)numTiles) return;
        const LuxgiAtlasTile& t = tiles[dr.tile];
        const int X = (int)std::lrint(t.extends[0] * (float)res), Y = (int)std::lrint(t.extends[1] * (float)res);
        const int Wt = (int)std::lrint(t.extends[2] * (float)res) + 1, Ht = (int)std::lrint(t.extends[3] * (float)res) + 1;
        for (int py = Y; py < Y + Ht; ++py)
            for (int px = X; px < X + Wt; ++px) {
                if (px < 0 || py < 0 || px >= res || py >= res) continue;
                surface_light_texel<T>(*prm, lights, dr.lightMask, t, res, px, py, gbColor, gbNormal, gbPbr, gbEmissive,
                                       *g, sdf, mip, sdfRes, mipRes, u, irr, dep, out + ((size_t)py * res + px) * 4,
                                       offsets, states);
            }
    });
}
extern "C" void oracle_surface_light_f64(const LuxgiSurfaceLightParams* prm, const LuxgiLight* lights,
                                         const LuxgiAtlasTile* tiles, int numTiles, const LuxgiSurfaceDraw* draws,
                                         int numDraws, int res, const float* gbColor, const float* gbNormal,
                                         const float* gbPbr, const float* gbEmissive, const LuxgiGlobalSDFData* g,
                                         const float* sdf, const float* mip, int sdfRes, int mipRes,
                                         const LuxgiDDGIUniform* u, const float* irr, const float* dep,
                                         const float* offsets, const uint8_t* states, double* out, int threads) {
    surface_light_impl<double>(prm, lights, tiles, numTiles, draws, numDraws, res, gbColor, gbNormal, gbPbr, gbEmissive,
                            g, sdf, mip, sdfRes, mipRes, u, irr, dep, offsets, states, out, threads);
}
extern "C" void oracle_surface_light_f32(const LuxgiSurfaceLightParams* prm, const LuxgiLight* lights,
                                         const LuxgiAtlasTile* tiles, int numTiles, const LuxgiSurfaceDraw* draws,
                                         int numDraws, int res, const float* gbColor, const float* gbNormal,
                                         const float* gbPbr, const float* gbEmissive, const LuxgiGlobalSDFData* g,
                                         const float* sdf, const float* mip, int sdfRes, int mipRes,
                                         const LuxgiDDGIUniform* u, const float* irr, const float* dep,
                                         const float* offsets, const uint8_t* states, float* out, int threads) {
    surface_light_impl<float>(prm, lights, tiles, numTiles, draws, numDraws, res, gbColor, gbNormal, gbPbr, gbEmissive,
                            g, sdf, mip, sdfRes, mipRes, u, irr, dep, offsets, states, out, threads);
}
// Raytraced/BRDF.glsl:65-79 at one set of inputs (unit test hook)
extern "C" void oracle_brdf(const double* albedo, const double* n, double rough, double metal, const double* view,
                            const double* h, const double* l, double* out) {
    v3<double> r = brdf_ggx<double>({albedo[0], albedo[1], albedo[2]}, {n[0], n[1], n[2]}, rough, metal,
                                    {view[0], view[1], view[2]}, {h[0], h[1], h[2]}, {l[0], l[1], l[2]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

// SDFShadow.comp:113-157 over the image: mask [ceil(H/4)][ceil(W/8)]; a word is written only when its
// block's first pixel is not sky (other words keep their value).
template <class T>
static void shadow_mask_impl(const LuxgiSdfShadowParams* p, const uint8_t* sobol, const uint8_t* scr, int W, int H,
                             const float* depth, const float* normals, const LuxgiGlobalSDFData* g, const float* sdf,
                             const float* mip, int res, int mipRes, uint32_t* mask, int threads) {
    const int mw = (W + 7) / 8, mh = (H + 3) / 4;
    parallel_for(mh, threads, [&](int wy) {
        for (int wx = 0; wx < mw; ++wx) {
            uint32_t word = 0;
            bool firstSky = true;
            for (int k = 0; k < 32; ++k) {
                bool sky;
                const uint32_t v = shadow_visibility<T>(*p, sobol, scr, W, H, wx * 8 + (k & 7), wy * 4 + (k >> 3), depth,
                                                        normals, *g, sdf, mip, res, mipRes, &sky);
                if (k == 0) firstSky = sky;
                word |= v << k;
            }
            if (!firstSky) mask[(size_t)wy * mw + wx] = word;
        }
    });
}
extern "C" void oracle_shadow_mask_f64(const LuxgiSdfShadowParams* p, const uint8_t* sobol, const uint8_t* scr, int W,
                                       int H, const float* depth, const float* normals, const LuxgiGlobalSDFData* g,
                                       const float* sdf, const float* mip, int res, int mipRes, uint32_t* mask,
                                       int threads) {
    shadow_mask_impl<double>(p, sobol, scr, W, H, depth, normals, g, sdf, mip, res, mipRes, mask, threads);
}
extern "C" float oracle_blue_noise(const uint8_t* sobol, const uint8_t* scr, int x, int y, int frame, int dim) {
    return blue_noise(sobol, scr, x, y, frame, dim);
}

// SDFReflection.comp:116-162 over the image: out [H][W][4] (sky pixels keep their value).
extern "C" void oracle_sdf_reflection_f64(const LuxgiSdfReflectionParams* p, const uint8_t* sobol, const uint8_t* scr,
                                          int W, int H, const float* normals, const float* pbr, const float* depth,
                                          const LuxgiGlobalSDFData* g, const float* sdf, const float* mip, int res,
                                          int mipRes, const LuxgiSurfaceAtlasData* ad, const uint32_t* chunks,
                                          const uint32_t* culled, const LuxgiAtlasObject* objects,
                                          const LuxgiAtlasTile* tiles, const float* atlasDepth,
                                          const float* atlasColor, const float* skyFaces, int skySize,
                                          const LuxgiDDGIUniform* u, const float* irr, const float* dep,
                                          const float* offsets, const uint8_t* states, double* out, int threads) {
    ReflectionInputs in{p,   sobol,  scr,   W,     H,       normals,    pbr,        depth, g,       sdf,  mip,
                        res, mipRes, ad,    chunks, culled, objects,    tiles,      atlasDepth,    atlasColor,
                        skyFaces, skySize, u, irr, dep, offsets, states};
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) reflection_pixel<double>(in, x, y, out + ((size_t)y * W + x) * 4);
    });
}

// SDFRasterizeModel over the listed chunks of one cascade into a float volume [R][R][cascades*R]
// (values as stored: FP16-rounded when `half`).
template <class T>
static void sdf_rasterize_impl(const LuxgiRasterizeParams* p, const LuxgiObjectRasterizeData* objects,
                               const LuxgiMeshSDF* meshes, const uint16_t* pool, int numObjects, const int32_t* origins,
                               const uint32_t* start, const uint32_t* list, int numChunks, int cascades, int half,
                               float* out, int threads) {
    const int R = p->cascadeResolution, W = R * cascades;
    parallel_for(numChunks, threads, [&](int c) {
        const int* o = origins + (size_t)c * 4;
        for (int z = 0; z < 32; ++z)
            for (int y = 0; y < 32; ++y)
                for (int x = 0; x < 32; ++x) {
                    const int vx = o[0] + x, vy = o[1] + y, vz = o[2] + z;
                    if (vx < 0 || vy < 0 || vz < 0 || vx >= R || vy >= R || vz >= R) continue;
                    const T v = rasterize_voxel<T>(*p, objects, meshes, pool, numObjects, list + start[c],
                                                   start[c + 1] - start[c], vx, vy, vz, half != 0);
                    out[((size_t)vz * R + vy) * W + (size_t)p->cascadeIndex * R + vx] = (float)v;
                }
    });
}
extern "C" void oracle_sdf_rasterize_f64(const LuxgiRasterizeParams* p, const LuxgiObjectRasterizeData* objects,
                                         const LuxgiMeshSDF* meshes, const uint16_t* pool, int numObjects,
                                         const int32_t* origins, const uint32_t* start, const uint32_t* list,
                                         int numChunks, int cascades, int half, float* out, int threads) {
    sdf_rasterize_impl<double>(p, objects, meshes, pool, numObjects, origins, start, list, numChunks, cascades, half,
                               out, threads);
}
extern "C" void oracle_sdf_rasterize_f32(const LuxgiRasterizeParams* p, const LuxgiObjectRasterizeData* objects,
                                         const LuxgiMeshSDF* meshes, const uint16_t* pool, int numObjects,
                                         const int32_t* origins, const uint32_t* start, const uint32_t* list,
                                         int numChunks, int cascades, int half, float* out, int threads) {
    sdf_rasterize_impl<float>(p, objects, meshes, pool, numObjects, origins, start, list, numChunks, cascades, half,
                              out, threads);
}
extern "C" float oracle_round_to_half(float f) { return round_to_half(f); }

// Learned a-trous pass (FP64): forward over the image and the backward.
extern "C" void oracle_learned_atrous_forward(const LuxgiLearnedAtrousParams* p, const double* P, int W, int H,
                                              const float* normals, const float* pbr, const float* guide,
                                              const float* color, double* out, double* sumW) {
    LearnedAtrousImage im{W, H, normals, pbr, guide, color};
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t e = (size_t)y * W + x;
            la_forward_pixel<double>(im, *p, P, x, y, out + e * 4, sumW + e);
        }
}
extern "C" void oracle_learned_atrous_backward(const LuxgiLearnedAtrousParams* p, const double* P, int W, int H,
                                               const float* normals, const float* pbr, const float* image,
                                               const double* gOut, double* gIn, double* gP) {
    LearnedAtrousImage im{W, H, normals, pbr, image, image};
    la_backward<double>(im, *p, P, gOut, gIn, gP);
}

// Per-pixel learned edge-stopping scales (FP64).
extern "C" void oracle_reflection_atrous_scaled_f64(const LuxgiReflectionAtrousParams* pc, int W, int H,
                                                    const float* normals, const float* pbr, const uint8_t* tileMap,
                                                    const float* in, const float* scale, double* out, int threads) {
    const int tw = (W + 7) / 8;
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) {
            const size_t e = (size_t)y * W + x;
            if (tileMap[(y / 8) * tw + x / 8]) {
                reflection_atrous_pixel<double>(*pc, W, H, x, y, {normals, 4}, {pbr, 4}, {in, 4}, out + e * 4,
                                                scale ? double(scale[e]) : 1.0);
            } else {
                for (int c = 0; c < 4; ++c) out[e * 4 + c] = double(in[e * 4 + c]);
            }
        }
    });
}
extern "C" void oracle_learned_pixel_features(int W, int H, const float* color, const float* moments,
                                              const float* normals, const float* pbr, double* out, int threads) {
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x)
            learned_pixel_features<double>(W, H, x, y, {color, 4}, {moments, 4}, {normals, 4}, {pbr, 4},
                                           out + ((size_t)y * W + x) * kLpIn);
    });
}
extern "C" void oracle_learned_pixel_mlp(const double* P, int n, const double* features, int half, double* o,
                                         double* s, int threads) {
    parallel_for(n, threads, [&](int i) {
        LpForward<double> r;
        learned_pixel_mlp<double>(P, features + (size_t)i * kLpIn, half != 0, r);
        for (int j = 0; j < kLpOut; ++j) {
            o[(size_t)i * kLpOut + j] = r.o[j];
            s[(size_t)i * kLpOut + j] = r.s[j];
        }
    });
}
extern "C" void oracle_learned_pixel_mlp_backward(const double* P, int n, const double* features, int half,
                                                  const double* gS, double* gP) {
    for (int k = 0; k < kLpParams; ++k) gP[k] = 0;
    for (int i = 0; i < n; ++i)
        learned_pixel_mlp_backward<double>(P, features + (size_t)i * kLpIn, half != 0, gS + (size_t)i * kLpOut, gP);
}
extern "C" void oracle_learned_pixel_atrous_backward(const LuxgiReflectionAtrousParams* pc, int W, int H,
                                                     const float* normals, const float* pbr, const uint8_t* tileMap,
                                                     const float* in, const float* scale, const double* gOut,
                                                     double* gIn, double* gScale) {
    learned_pixel_atrous_backward<double>(*pc, W, H, {normals, 4}, {pbr, 4}, tileMap, {in, 4}, scale, gOut, gIn,
                                          gScale);
}
extern "C" int oracle_learned_pixel_param_count() { return kLpParams; }

/*
 * luxgi.h -- C ABI of the MI355X (gfx950) global-illumination hot path.
 *
 * The entry points replace the GPU dispatches that LuxGI's ECS systems record
 * through Renderer::dispatch / Pipeline::dispatchIndirect for:
 *   - the DDGI probe update, SDF ("soft trace") branch
 *       trace_rays::system   Code/Maple/src/Engine/DDGI/DDGIRenderer.cpp:217-330
 *       probe_update::system Code/Maple/src/Engine/DDGI/DDGIRenderer.cpp:377-413
 *       border_update::system Code/Maple/src/Engine/DDGI/DDGIRenderer.cpp:441-464
 *   - the global-SDF coarse mip + flood passes
 *       merge_sdf::system / fillFlood Code/Maple/src/Engine/DDGI/GlobalDistanceField.cpp:537-573,823-846
 *   - the SVGF denoisers
 *       raytraced_reflection::denoise::system Code/Maple/src/Engine/Raytrace/RaytracedReflection.cpp:346-561
 *       raytraced_shadow::denoise::system     Code/Maple/src/Engine/Raytrace/RaytracedShadow.cpp:406-610
 *
 * Conventions
 *   - All buffer pointers are DEVICE pointers owned by the caller.
 *   - Images are linear, row-major [height][width] arrays of texels with the
 *     reference's texel addressing and channel order (Vulkan images are opaque
 *     tiled; a caller sharing memory with Vulkan must use linear images or copy).
 *   - Matrices are column-major float[16] (GLSL / glm layout).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Every call is stream-ordered; none synchronises the host or allocates.
 *   - Return value: LUXGI_OK (0) or a negative LuxgiStatus.
 */
#ifndef LUXGI_H
#define LUXGI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Status codes and storage formats                                           */
/* ------------------------------------------------------------------------- */
typedef enum LuxgiStatus {
    LUXGI_OK = 0,
    LUXGI_ERR_INVALID_ARGUMENT = -1,
    LUXGI_ERR_HIP = -2,
    LUXGI_ERR_UNSUPPORTED = -3,
    LUXGI_ERR_WORKSPACE_TOO_SMALL = -4
} LuxgiStatus;

/* Texel storage of a buffer. Channel count is fixed by the buffer's role. */
typedef enum LuxgiStorage {
    LUXGI_STORAGE_FP32 = 0,   /* RGBA32F / RG32F / R32F: the plain FP32 form      */
    LUXGI_STORAGE_FP16 = 1,   /* RGBA16F / RG16F / R16F: the reference's formats  */
    LUXGI_STORAGE_UNORM8 = 2, /* RGBA8 unorm (surface-atlas colour only)          */
    LUXGI_STORAGE_BF16 = 3    /* R16 bfloat16 (SDF volume and mip only)           */
} LuxgiStorage;

/* ------------------------------------------------------------------------- */
/* POD mirrors of the reference's GPU structs (byte-identical layouts)         */
/* ------------------------------------------------------------------------- */

/* DDGIUniform: Shaders/DDGI/DDGICommon.glsl:11-31, host Engine/DDGI/DDGIRenderer.h:22-42.
 * 96 bytes, scalar layout. */
typedef struct LuxgiDDGIUniform {
    float   startPosition[4];
    float   step[4];
    int32_t probeCounts[4];
    float   maxDistance;
    float   sharpness;
    float   hysteresis;
    float   normalBias;
    float   ddgiGamma;
    int32_t irradianceProbeSideLength;
    int32_t irradianceTextureWidth;
    int32_t irradianceTextureHeight;
    int32_t depthProbeSideLength;
    int32_t depthTextureWidth;
    int32_t depthTextureHeight;
    int32_t raysPerProbe;
} LuxgiDDGIUniform;

/* GlobalSDFData: Shaders/SDF/GlobalSDFData.glsl:4-12. 96 bytes. */
typedef struct LuxgiGlobalSDFData {
    float    cascadePosDistance[4][4]; /* xyz = cascade centre, w = half extent */
    float    cascadeVoxelSize[4];
    uint32_t cascadesCount;
    float    resolution;
    float    nearPlane;
    float    farPlane;
} LuxgiGlobalSDFData;

/* GlobalSurfaceAtlasData: Shaders/SDF/AtlasCommon.glsl:8-16. 32 bytes. */
typedef struct LuxgiSurfaceAtlasData {
    float    cameraPos[3];
    float    chunkSize;
    uint32_t culledObjectsCapacity;
    uint32_t resolution;
    uint32_t objectsCount;
    uint32_t padding;
} LuxgiSurfaceAtlasData;

/* ObjectBuffer: Shaders/SDF/AtlasCommon.glsl:18-25 (std430). 128 bytes. */
typedef struct LuxgiAtlasObject {
    float    objectBounds[4]; /* bounding sphere: centre xyz, radius w */
    uint32_t tileOffset[6];   /* 0 = no tile */
    int32_t  padding[2];
    float    transform[16];   /* local -> world, column-major */
    float    extends[4];      /* local half extents xyz */
} LuxgiAtlasObject;

/* TileBuffer: Shaders/SDF/AtlasCommon.glsl:27-32 (std430). 96 bytes. */
typedef struct LuxgiAtlasTile {
    float extends[4];      /* atlas rect (x, y, w, h) in UV */
    float transform[16];   /* object-local -> tile view, column-major */
    float objectBounds[4]; /* tile view box size xyz */
} LuxgiAtlasTile;

/* GISDFRays.comp push constants (Shaders/DDGI/GISDFRays.comp:53-60, 80 bytes)
 * followed by options that extend the reference. */
typedef struct LuxgiTraceParams {
    float    randomOrientation[16]; /* column-major; only the upper 3x3 is used */
    uint32_t numFrames;
    uint32_t infiniteBounces;
    int32_t  numLights;
    float    intensity;
    /* --- extensions (defaults reproduce the reference) --- */
    int32_t  maxSteps;      /* per-cascade step cap; reference = 250 (SDFCommon.glsl:143) */
    float    hitEpsilon;    /* 0 = reference thickness test (SDFCommon.glsl:166-167); >0 = fixed eps */
    int32_t  shadeOrder;    /* 0 = shade hit rays grouped by surface-atlas chunk (counting sort in the
                               workspace), 1 = in ray order; the outputs are identical */
    int32_t  pad1;
    float    skyColor[4];   /* sky radiance when no cubemap is bound (reference binds a unit cube) */
    const float* deviceOrientation; /* optional device array of 16 floats (column-major): read in place of
                                       randomOrientation when the kernels run, so a captured graph follows
                                       the orientation copied there before each replay */
} LuxgiTraceParams;

/* ------------------------------------------------------------------------- */
/* Resource descriptors                                                       */
/* ------------------------------------------------------------------------- */

/* Global SDF clipmap + coarse mip. Layout [res][res][cascades*res] (z, y, x),
 * x fastest, cascades packed along x (GlobalDistanceField.cpp:611-622).
 * Values are normalised distances (world distance / (2*halfExtent)). */
typedef struct LuxgiSDFVolume {
    const void* sdf;      /* R32F or R16F */
    const void* mip;      /* R32F or R16F, [mipRes][mipRes][cascades*mipRes] */
    int32_t     resolution;
    int32_t     mipResolution;
    int32_t     storage;  /* LuxgiStorage: FP32, FP16 or BF16 */
    int32_t     layout;   /* LuxgiLayout of `sdf`; the mip is LINEAR, except with LUXGI_LAYOUT_CELL8, where
                             the mip is in CELL8 records as well */
    const float* guard;   /* optional [D/4][H/4][W/4] lower bounds of `sdf` per 4x4x4 brick and its
                             +1 apron (luxgi_sdf_build_guard); lets the march skip fine lookups whose
                             value the reference would not use. NULL = no skipping. */
    const float* brickScale; /* optional, LUXGI_LAYOUT_BRICK4 only: distance = stored value * brickScale[brick]
                                (luxgi_sdf_to_bricks with a scale output). NULL = scale 1. */
    const float* refine;  /* optional FP32 fine volume (layout `refineLayout`) of the same distances: the
                             hit step's distance, the hit time and the hit normal are re-evaluated from it,
                             so a reduced-precision march returns FP32 hits. NULL = none. */
    int32_t     refineLayout;
    int32_t     pad;
} LuxgiSDFVolume;

/* Memory layout of the fine SDF volume and of the surface atlas. */
typedef enum LuxgiLayout {
    LUXGI_LAYOUT_LINEAR = 0, /* SDF: [z][y][x], the reference texture's texel order.
                                Atlas: separate colour [res][res] and depth [res][res] images. */
    LUXGI_LAYOUT_BRICK4 = 1, /* SDF: 4x4x4 bricks (x-fastest brick order), each stored with a +1 apron
                                as 5x5x5 voxels (element k*25 + j*5 + i = voxel (4bx+i, 4by+j, 4bz+k),
                                clamped to the volume) padded to 128 elements; FP32 or FP16, every
                                dimension a multiple of 4 (luxgi_sdf_to_bricks, luxgi_sdf_bricks_bytes).
                                Atlas: one 8-byte texel (RGBA8 colour, FP32 depth) per texel in 4x4-texel
                                blocks of 128 bytes, blocks row-major; UNORM8 colour, res a multiple of 4
                                (luxgi_atlas_pack); `color` points at it and `depth` is unused. */
    LUXGI_LAYOUT_CELL8 = 2   /* SDF only, for the fine volume AND the coarse mip of an LuxgiSDFVolume:
                                one record per voxel cell (x, y, z) holding the 8 corner voxels of the
                                trilinear footprint whose lower corner it is, corner (i, j, k) at element
                                i + 2j + 4k = voxel (x+i, y+j, z+k) clamped to the volume (32 B FP32 or
                                16 B FP16 per record), so a lookup is one aligned record. Records are
                                grouped in bricks of 8x8x8 cells (x-fastest brick order, 512 records per
                                brick, record (z%8)*64 + (y%8)*8 + x%8 inside its brick). Any dimensions
                                (luxgi_sdf_to_cells, luxgi_sdf_cells_bytes). Lookups return what the
                                LINEAR layout returns, bit for bit. */
} LuxgiLayout;

/* Surface atlas + object/tile/chunk buffers read by the hit lookup
 * (Shaders/SDF/AtlasCommon.glsl:115-157). */
typedef struct LuxgiSurfaceAtlas {
    LuxgiSurfaceAtlasData   data;
    const uint32_t*         chunks;         /* [40^3] start index into culledObjects, 0 = empty */
    const uint32_t*         culledObjects;  /* [count, id0, id1, ...] runs */
    const LuxgiAtlasObject* objects;
    const LuxgiAtlasTile*   tiles;
    const void*             color;          /* light cache [res][res] RGBA (storage below) */
    const float*            depth;          /* tile depth [res][res] R32F, >=1 = empty */
    int32_t                 objectCapacity; /* number of entries in `objects` */
    int32_t                 colorStorage;   /* FP32, FP16 or UNORM8 */
    int32_t                 layout;         /* LuxgiLayout of color/depth (LINEAR or BRICK4) */
    int32_t                 pad;
} LuxgiSurfaceAtlas;

/* Optional sky cubemap: 6 faces (+X,-X,+Y,-Y,+Z,-Z) of size^2 RGBA32F texels. */
typedef struct LuxgiSkybox {
    const float* faces;
    int32_t      size;
    int32_t      pad;
} LuxgiSkybox;

/* DDGI ray buffers: radiance (rgb, 0) and direction/distance (xyz, t),
 * texel (rayId, probeId) -> element probeId*width + rayId
 * (DDGIRenderer.cpp:171-177). */
typedef struct LuxgiRayBuffers {
    void*   radiance;
    void*   directionDistance;
    int32_t width;    /* texels per row (>= uniform.raysPerProbe) */
    int32_t height;   /* rows (>= probe count) */
    int32_t storage;  /* FP32 (RGBA32F) or FP16 (RGBA16F) */
    int32_t pad;
} LuxgiRayBuffers;

/* One ping-pong side of the probe atlases (DDGIRenderer.cpp:179-209):
 * irradiance RGBA, depth RG, sizes taken from the uniform.
 * probeOffsets / probeStates are the output of luxgi_ddgi_relocate_classify and are read by the
 * atlas consumers (luxgi_ddgi_sample_probes, the bounce of luxgi_surface_light, the DDGI branch
 * of luxgi_sdf_reflection): a probe sits at its grid position + offset, and an inactive probe
 * (state 1) takes no weight; a point whose contributing probes are all inactive gets 0, as in
 * the published classification rule (Majercik et al., JCGT 2021). NULL = the reference's
 * sampleIrradiance (DDGICommon.glsl:163-233): grid positions, every probe active. The update and
 * border passes ignore both. */
typedef struct LuxgiProbeAtlas {
    void*   irradiance;
    void*   depth;
    int32_t storage;  /* FP32 or FP16 */
    int32_t pad;
    const float*   probeOffsets; /* optional float4[P] (xyz used), as written by relocation */
    const uint8_t* probeStates;  /* optional uint8[P]: 0 active, 1 inactive */
} LuxgiProbeAtlas;

/* Probe-update flags */
#define LUXGI_UPDATE_IRRADIANCE   0x1u  /* run IrradianceProbeUpdate */
#define LUXGI_UPDATE_DEPTH        0x2u  /* run DepthProbeUpdate */
#define LUXGI_UPDATE_FUSE_BORDER  0x4u  /* also write the border texels (BorderUpdate) */
#define LUXGI_UPDATE_PER_PROBE    0x8u  /* force the per-probe weight path (no shared-direction GEMM) */
#define LUXGI_UPDATE_HALF        0x10u  /* reduced precision: weights and ray data rounded to FP16 for the
                                           FP16 MFMA contraction, FP32 accumulation and epilogue */
#define LUXGI_UPDATE_BF16        0x20u  /* with LUXGI_UPDATE_HALF: the operands rounded to BF16 for the BF16 MFMA
                                           (v_mfma_f32_16x16x32_bf16) instead of FP16 */

/* ------------------------------------------------------------------------- */
/* DDGI                                                                        */
/* ------------------------------------------------------------------------- */

/* Atlas sizes for a probe grid and octahedral side lengths
 * (DDGIRenderer.cpp:182-185): width = (oct+2)*px*py+2, height = (oct+2)*pz+2.
 * Fills the *TextureWidth/Height fields of `uniform`. Host-only. */
int luxgi_ddgi_fill_atlas_sizes(LuxgiDDGIUniform* uniform);

/* Workspace bytes needed by luxgi_ddgi_trace_sdf for the probe grid of `uniform`
 * and `objectCapacity` atlas objects (per-object inverse transforms and the
 * 24-byte per-ray march records that the shading kernel consumes). */
size_t luxgi_ddgi_trace_workspace_size(const LuxgiDDGIUniform* uniform, int32_t objectCapacity);

/* GISDFRays.comp (Shaders/DDGI/GISDFRays.comp:63-128): probe ray generation,
 * global-SDF sphere trace (SDFCommon.glsl:98-194), hit normal, surface-atlas
 * lookup (AtlasCommon.glsl:115-157) and ray-result encode.
 * Replaces the dispatch at DDGIRenderer.cpp:318-321.
 *   probeOffsets  optional float4[P] relocation offsets added to probe positions (NULL = none)
 *   rayFlags      optional uint8[P*width] per-ray class: 0 miss, 1 front hit, 2 inside/backface (NULL = none)
 *   workspace     device scratch of luxgi_ddgi_trace_workspace_size() bytes; calls that overlap on
 *                 one workspace race */
int luxgi_ddgi_trace_sdf(const LuxgiDDGIUniform* uniform,
                         const LuxgiGlobalSDFData* sdfData,
                         const LuxgiSDFVolume* volume,
                         const LuxgiSurfaceAtlas* atlas,
                         const LuxgiSkybox* skybox,
                         const LuxgiTraceParams* params,
                         const float* probeOffsets,
                         const LuxgiRayBuffers* rays,
                         uint8_t* rayFlags,
                         void* workspace, size_t workspaceBytes,
                         void* stream);

/* Counting run of the same trace: accumulates into counters[8] (uint64, device):
 *   [0] mip lookups, [1] fine lookups, [2] front hits that evaluate a normal (6 lookups each),
 *   [3] atlas objects tested, [4] tile samples (4-texel gathers), [5] rays,
 *   [6] total march steps, [7] rays that hit.
 * The caller zeroes counters first. Outputs are written exactly as by luxgi_ddgi_trace_sdf. */
int luxgi_ddgi_trace_sdf_count(const LuxgiDDGIUniform* uniform,
                               const LuxgiGlobalSDFData* sdfData,
                               const LuxgiSDFVolume* volume,
                               const LuxgiSurfaceAtlas* atlas,
                               const LuxgiSkybox* skybox,
                               const LuxgiTraceParams* params,
                               const float* probeOffsets,
                               const LuxgiRayBuffers* rays,
                               uint8_t* rayFlags,
                               void* workspace, size_t workspaceBytes,
                               unsigned long long* counters,
                               void* stream);

/* Workspace bytes needed by luxgi_ddgi_probe_update. */
size_t luxgi_ddgi_probe_update_workspace_size(const LuxgiDDGIUniform* uniform);

/* IrradianceProbeUpdate.comp + DepthProbeUpdate.comp (Shaders/DDGI/ProbeUpdate.glsl:53-152),
 * dispatched at DDGIRenderer.cpp:347-375. Reads `prev` (the pingPong side),
 * writes `out` (the 1-pingPong side). firstFrame != 0 disables hysteresis.
 * With LUXGI_UPDATE_FUSE_BORDER the border texels are written as by
 * luxgi_ddgi_border_update in the same pass. */
int luxgi_ddgi_probe_update(const LuxgiDDGIUniform* uniform,
                            const LuxgiRayBuffers* rays,
                            const LuxgiProbeAtlas* prev,
                            const LuxgiProbeAtlas* out,
                            uint32_t firstFrame,
                            uint32_t flags,
                            void* workspace, size_t workspaceBytes,
                            void* stream);

/* luxgi_ddgi_probe_update without a workspace argument and without any allocation. No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The workspace lives in
 * a static device array of LUXGI_UPDATE_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends; volumes whose luxgi_ddgi_probe_update_workspace_size() exceeds it are rejected with
 * LUXGI_ERR_WORKSPACE_TOO_SMALL (256 rays, 16/16 oct and up to 2M probes fit). All calls on one
 * GPU share that array, so calls that overlap on one GPU (different streams, different host
 * threads) race; calls ordered on one stream do not. Results equal luxgi_ddgi_probe_update's. */
#define LUXGI_UPDATE_NOALLOC_SCRATCH_BYTES (4u << 20)
int luxgi_ddgi_probe_update_noalloc(const LuxgiDDGIUniform* uniform,
                                    const LuxgiRayBuffers* rays,
                                    const LuxgiProbeAtlas* prev,
                                    const LuxgiProbeAtlas* out,
                                    uint32_t firstFrame,
                                    uint32_t flags,
                                    void* stream);

/* Probe update of several volumes (each its own grid, ray count and atlases) in one launch per
 * map: the per-volume sequence of luxgi_ddgi_probe_update (DDGIRenderer.cpp:347-375 per volume)
 * with the volumes' 16-probe groups laid end to end in one grid. Results equal a
 * luxgi_ddgi_probe_update call per volume with the same flags, bit for bit. Every volume must take
 * the shared-direction MFMA path (irradiance oct^2 <= 64, depth oct^2 <= 256, maxDistance > 0,
 * no LUXGI_UPDATE_PER_PROBE) and all volumes store rays alike and atlases alike. */
#define LUXGI_MAX_UPDATE_BATCH 4096
typedef struct LuxgiUpdateVolume {
    LuxgiDDGIUniform uniform;
    LuxgiRayBuffers  rays;
    LuxgiProbeAtlas  prev;       /* read unless firstFrame */
    LuxgiProbeAtlas  out;
    uint32_t         firstFrame;
    int32_t          pad;
} LuxgiUpdateVolume;

/* Launch shape of a prepared batch (filled by luxgi_ddgi_update_batch_prepare). */
typedef struct LuxgiUpdateBatch {
    int32_t count;            /* volumes */
    int32_t groups;           /* 16-probe groups over all volumes */
    int32_t flags;            /* LUXGI_UPDATE_* shared by the volumes */
    int32_t rayStorage;
    int32_t atlasStorage;
    int32_t maxRays;          /* largest raysPerProbe */
    int32_t maxWeightBlocks;  /* largest (texel tiles x ray chunks) of the weight pass */
    int32_t pad;
} LuxgiUpdateBatch;

/* Bytes of the device descriptor array of a batch of `count` volumes (0 if count is out of range). */
size_t luxgi_ddgi_update_batch_descriptor_bytes(int32_t count);
/* Device workspace bytes of the batch: the volumes' update workspaces end to end, 256-aligned. */
size_t luxgi_ddgi_update_batch_workspace_size(const LuxgiUpdateVolume* volumes, int32_t count);
/* Host-only: checks every volume as luxgi_ddgi_probe_update does, carves the volumes' workspaces
 * out of `workspace` (a device address, not dereferenced here) and writes the descriptors into
 * `descriptors` (HOST memory of luxgi_ddgi_update_batch_descriptor_bytes(count) bytes), which the
 * caller copies to device memory; fills `batch`. A batch whose workspace or descriptors are shared
 * with another call races with it. */
int luxgi_ddgi_update_batch_prepare(const LuxgiUpdateVolume* volumes, int32_t count, uint32_t flags,
                                    void* workspace, size_t workspaceBytes,
                                    void* descriptors, LuxgiUpdateBatch* batch);
/* Runs a prepared batch; `descriptors` is the DEVICE copy of what prepare wrote. */
int luxgi_ddgi_probe_update_batched(const LuxgiUpdateBatch* batch, const void* descriptors, void* stream);

/* IrradianceBorderUpdate.comp / DepthBorderUpdate.comp (Shaders/DDGI/BorderUpdate.glsl:136-156),
 * dispatched at DDGIRenderer.cpp:420-438. In-place on `atlas`.
 * which: LUXGI_UPDATE_IRRADIANCE and/or LUXGI_UPDATE_DEPTH. */
int luxgi_ddgi_border_update(const LuxgiDDGIUniform* uniform,
                             const LuxgiProbeAtlas* atlas,
                             uint32_t which,
                             void* stream);

/* Probe relocation + classification (no counterpart in the reference; rules
 * after Majercik et al., JCGT 2021). Reads the ray distances and rayFlags of a
 * trace, writes offsets (float4[P], xyz used, w = 0) and states (uint8[P]:
 * 0 active, 1 inactive). `offsets` is read and updated in place. */
typedef struct LuxgiRelocateParams {
    float backfaceThreshold;      /* fraction of inside/backface rays that marks a probe as inside (0.25) */
    float minFrontfaceDistance;   /* push away from frontfaces closer than this (world units) */
    float maxOffsetFraction;      /* offsets are clamped to +-this * step per axis (0.45) */
    int32_t relocate;             /* 0 = classification only */
} LuxgiRelocateParams;

int luxgi_ddgi_relocate_classify(const LuxgiDDGIUniform* uniform,
                                 const LuxgiRayBuffers* rays,
                                 const uint8_t* rayFlags,
                                 const LuxgiRelocateParams* params,
                                 float* offsets,
                                 uint8_t* states,
                                 void* stream);

/* SampleProbe.comp UniformBufferObject (Shaders/DDGI/SampleProbe.comp:29-33). */
typedef struct LuxgiSampleProbeParams {
    float cameraPosition[4];
    float viewProjInv[16];   /* column-major inverse(projView) */
} LuxgiSampleProbeParams;

/* SampleProbe.comp:36-60 with sampleIrradiance (Shaders/DDGI/DDGICommon.glsl:163-233):
 * per-pixel indirect irradiance from the 8 surrounding probes (Chebyshev depth
 * visibility, smooth backface and trilinear weights), read from `atlas` with
 * bilinear filtering and repeat addressing (the default sampler, Definitions.h:152-159).
 * Dispatched at DDGIRenderer.cpp:467-503.
 *   depth    R32F [height][width] NDC depth (1 = sky: output 0)
 *   normals  RGBA32F [height][width], octahedral normal in .xy
 *   out      RGBA [height][width] (irradiance, 1); outStorage FP32 or FP16 */
int luxgi_ddgi_sample_probes(const LuxgiDDGIUniform* uniform,
                             const LuxgiProbeAtlas* atlas,
                             const LuxgiSampleProbeParams* params,
                             int32_t width, int32_t height,
                             const float* depth, const float* normals,
                             int32_t outStorage, void* out,
                             void* stream);

/* ------------------------------------------------------------------------- */
/* Global SDF coarse mip (GlobalSDFMipmap.comp:32-68 + fillFlood GlobalDistanceField.cpp:537-573) */
/* ------------------------------------------------------------------------- */

/* Builds `mip` ([mipRes][mipRes][cascades*mipRes], FP32) for one cascade from
 * `sdf` ([res][res][cascades*res], FP32): the 7-tap min pass followed by
 * floodPasses (reference: 4) ping-pong passes through `temp`
 * ([mipRes]^3 FP32 scratch). maxDistance = cascade full extent (2*halfExtent). */
int luxgi_sdf_build_mip(const float* sdf, int32_t resolution, int32_t cascades, int32_t cascade,
                        float maxDistance, int32_t floodPasses,
                        float* mip, float* temp, void* stream);

/* ObjectRasterizeData: Shaders/SDF/ObjectRasterizeData.glsl (std430; host GlobalDistanceField.cpp:102-112,
 * filled by chunkCalculate :486-504). 176 bytes. */
typedef struct LuxgiObjectRasterizeData {
    float worldToVolume[16];        /* column-major */
    float volumeToWorld[16];
    float volumeToUVWMul[3];
    float mipOffset;                /* mesh-SDF mip level read (min(cascade, 2)) */
    float volumeToUVWAdd[3];
    float decodeMul;                /* distance = (texel * 2 - 1) * decodeMul */
    float volumeLocalBoundsExtent[3];
    float decodeAdd;
} LuxgiObjectRasterizeData;

/* ModelsRasterizeData UBO (SDFRasterizeModel.glsl:16-24; host GlobalDistanceField.cpp:756-761). 48 bytes. */
typedef struct LuxgiRasterizeParams {
    float   cascadeCoordToPosMul[3];
    float   maxDistance;            /* the cascade's full extent */
    float   cascadeCoordToPosAdd[3];
    int32_t cascadeResolution;
    int32_t cascadeIndex;
    int32_t pad[3];
} LuxgiRasterizeParams;

/* One object's mesh SDF (SDFBaker.cpp:158-200): RGBA-less R16F texels [z][y][x] per mip level,
 * stored as FP16 bits in a shared pool; the engine's Texture3D (repeat addressing, linear filter).
 * 72 bytes. */
typedef struct LuxgiMeshSDF {
    uint64_t offset[3];   /* element offset of mip levels 0 .. mipCount-1 in the pool */
    int32_t  size[3][3];  /* x, y, z of each level */
    int32_t  mipCount;    /* 1 .. 3 */
    int32_t  pad[2];
} LuxgiMeshSDF;

/* SDFRasterizeModel.glsl:42-63 with GlobalSDFClear.comp:27-33 (dispatch GlobalDistanceField.cpp:700-816):
 * for each listed 32^3 chunk of one cascade, every voxel gets clamp(min over the chunk's objects of
 * distanceToModelSDF (SDFCommon.glsl:38-62) / maxDistance, -1, 1), or 1 for a chunk without objects.
 * The objects of a chunk are taken in list order in batches of 28 (GLOBAL_SDF_RASTERIZE_MODEL_MAX_COUNT):
 * from the second batch on the running distance restarts from the stored value, as the reference's
 * READ_DISTANCE layers do (so FP16 storage rounds between batches). chunkOrigins [numChunks][4] holds
 * the voxel coordinates of each chunk's first voxel within the cascade; chunkObjectStart
 * [numChunks + 1] and chunkObjects form the per-chunk object lists (indices into objects and meshes;
 * out-of-range entries are skipped). sdf is the [res][res][cascades*res] global SDF (FP32 or FP16);
 * voxels outside the cascade are not written. The reference drops a full chunk's first 28 objects
 * when it overflows (the reference assigns the next layer into the full one, GlobalDistanceField.cpp:521-525);
 * here every listed object is used. */
int luxgi_sdf_rasterize(const LuxgiRasterizeParams* params, const LuxgiObjectRasterizeData* objects,
                        const LuxgiMeshSDF* meshes, const uint16_t* pool, int32_t numObjects,
                        const int32_t* chunkOrigins, const uint32_t* chunkObjectStart, const uint32_t* chunkObjects,
                        int32_t numChunks, int32_t cascades, int32_t storage, void* sdf, void* stream);

/* Bytes of a LUXGI_LAYOUT_BRICK4 volume of width x height x depth voxels (multiples of 4) in
 * `storage` (FP32 or FP16): (W/4)(H/4)(D/4) bricks of 128 elements. 0 for invalid arguments. */
size_t luxgi_sdf_bricks_bytes(int32_t width, int32_t height, int32_t depth, int32_t storage);

/* Re-lays a linear FP32 volume [depth][height][width] as LUXGI_LAYOUT_BRICK4 in `storage` (FP32,
 * or FP16 rounded to nearest); `dst` holds luxgi_sdf_bricks_bytes() bytes. With `brickScale`
 * (nullable, one float per brick) each brick is stored divided by the smallest power of two not
 * below its largest magnitude (1 for an all-zero brick), which keeps FP16 bricks of any distance
 * range in [-1, 1] and out of the subnormals; the division and the sampler's multiplication are
 * exact. No counterpart in the reference, whose Vulkan 3D image is tiled by the driver. */
int luxgi_sdf_to_bricks(const float* src, int32_t width, int32_t height, int32_t depth, int32_t storage, void* dst,
                        float* brickScale, void* stream);

/* Bytes of a width x height x depth volume in LUXGI_LAYOUT_CELL8 (storage FP32 or FP16; 0 on bad
 * arguments or when the records would not be addressable with 32-bit record indices). */
size_t luxgi_sdf_cells_bytes(int32_t width, int32_t height, int32_t depth, int32_t storage);

/* Converts a LINEAR FP32 volume ([depth][height][width]) into LUXGI_LAYOUT_CELL8 records, rounded
 * to FP16 (nearest) when storage is FP16. `dst` holds luxgi_sdf_cells_bytes() bytes; the padding
 * records of partial bricks are written as zeros. Used for both the fine volume and the mip. */
int luxgi_sdf_to_cells(const float* src, int32_t width, int32_t height, int32_t depth, int32_t storage, void* dst,
                       void* stream);

/* Packs a UNORM8 colour atlas (RGBA8 [res][res]) and its FP32 depth ([res][res]) into the
 * LUXGI_LAYOUT_BRICK4 atlas layout (res a multiple of 4; `packed` holds res*res*8 bytes). No
 * counterpart in the reference (its atlas is two driver-tiled images); the bilinear footprints of
 * the hit lookup then touch one 128-byte line in 9/16 of the cases instead of four lines. */
int luxgi_atlas_pack(const uint32_t* color, const float* depth, int32_t resolution, void* packed, void* stream);

/* Per-brick lower bounds for LuxgiSDFVolume.guard: guard[bz][by][bx] = min of the linear FP32
 * volume `sdf` ([depth][height][width], dims multiples of 4) over voxels [4b, 4b+4] on each axis
 * (clamped), i.e. every trilinear footprint whose lower corner lies in brick b. No counterpart in
 * the reference; the trace results are identical with and without it. */
int luxgi_sdf_build_guard(const float* sdf, int32_t width, int32_t height, int32_t depth, float* guard,
                          void* stream);

/* ------------------------------------------------------------------------- */
/* Surface atlas: chunk culling and surface-cache lighting                    */
/* ------------------------------------------------------------------------- */

/* Workspace bytes of luxgi_atlas_cull_chunks (fixed: two 40^3 uint arrays and the scan totals). */
size_t luxgi_atlas_cull_workspace_size(void);

/* SDFCulling.comp:36-101 (dispatch GlobalSurfaceAtlas.cpp:607-641): for each of the 40^3 chunks
 * of chunkSize * [-20, 20)^3, the objects (data->objectsCount entries of `objects`) whose bounding
 * sphere objectBounds touches the chunk box, in increasing index order. Writes chunks[40^3]
 * (offset of the chunk's run [count, id0, id1, ...] in culledObjects, 0 = none) and
 * culledObjects (data->culledObjectsCapacity entries; [0] = 1 + the sum of count + 1 over the
 * non-empty chunks, the counter value the reference leaves). Runs are placed in chunk-address
 * order (the reference's atomicAdd order is timing-dependent); an empty chunk, or one whose run
 * would end past the capacity, gets offset 0 (the reference writes the chunk address into
 * chunks[0] instead, SDFCulling.comp:60,69). */
int luxgi_atlas_cull_chunks(const LuxgiSurfaceAtlasData* data, const LuxgiAtlasObject* objects, uint32_t* chunks,
                            uint32_t* culledObjects, void* workspace, size_t workspaceBytes, void* stream);

/* Light: Shaders/Common/Light.glsl:19-28 (host Scene/Component/Light.h:21-31), 64 bytes.
 * type: 0 directional, 1 spot, 2 point (Light.glsl:13-15). */
typedef struct LuxgiLight {
    float color[4];
    float position[4];
    float direction[4];
    float intensity;
    float radius;
    float type;
    float angle;
} LuxgiLight;

#define LUXGI_MAX_LIGHTS 32         /* MAX_LIGHTS, Light.glsl:30 */
#define LUXGI_SURFACE_EMISSIVE 0x1u /* CopyEmissive.frag:22-27: light = emissive        */
#define LUXGI_SURFACE_DIRECT   0x2u /* SDFDeferredLight.frag:79-130: += per light        */
#define LUXGI_SURFACE_INDIRECT 0x4u /* SDFAtlasIndirectLight.frag:44-68: += DDGI bounce  */

/* Surface-cache G-buffer (GlobalSurfaceAtlas.cpp:407-410), [res][res] texels each:
 * color    RGBA16F (albedo rgb, world x)               SDFDeferredColor.frag
 * normal   RGBA16F (octahedral normal xy, world y, z)
 * pbr      RGBA8 UNORM (metallic, roughness, ...)
 * emissive RGBA16F (rgb; R11G11B10 in the reference) */
typedef struct LuxgiSurfaceGBuffer {
    const void* color;
    const void* normal;
    const void* pbr;
    const void* emissive;
    int32_t     resolution;
    int32_t     pad;
} LuxgiSurfaceGBuffer;

/* SDFDeferredLight.frag:31-36 / SDFAtlasIndirectLight.frag:37-40 uniforms and the passes to run. */
typedef struct LuxgiSurfaceLightParams {
    float    cameraPosition[4]; /* xyz; w = shadow bias (GlobalSurfaceAtlas.cpp:99,946) */
    float    indirectIntensity; /* IrradianceVolume intensity (GlobalSurfaceAtlas.cpp:1043) */
    uint32_t passes;            /* LUXGI_SURFACE_* */
    int32_t  numLights;         /* entries of `lights`, <= LUXGI_MAX_LIGHTS */
    int32_t  maxSteps;          /* shadow-ray steps per cascade; 0 = 250 (SDFCommon.glsl:143) */
} LuxgiSurfaceLightParams;

/* One tile quad of the lighting draws (GlobalSurfaceAtlas.cpp:554-574). */
typedef struct LuxgiSurfaceDraw {
    uint32_t tile;      /* index into the tile buffer; each tile at most once per call */
    uint32_t lightMask; /* bit l set: light l's draw list holds this tile (the host culling of
                           GlobalSurfaceAtlas.cpp:913-940) */
} LuxgiSurfaceDraw;

/* Surface-cache lighting of the drawn tiles into lightCache ([res][res] RGBA, FP32 or FP16,
 * alpha 1): emissive + sum of direct(light l) over the lights of the tile's mask + DDGI bounce,
 * per LUXGI_SURFACE_* pass. Each draw covers its tile's full rect; the G-buffer is sampled
 * bilinearly (repeat) at ((p + 0.5 - X) / Wt) * extends.zw + extends.xy, as the reference's tile
 * quads do. The direct term traces a shadow ray through the global SDF (`sdf`, any layout and
 * storage of the trace) from world + normal * bias with cascadeTraceStartBias 2
 * (SDFDeferredLight.frag:104-112). The bounce needs `ddgi` and `probes` (nullable otherwise).
 * `lights`, `tiles`, `draws` and every image are device memory; a draw naming a tile >= numTiles
 * is skipped. Texels outside the drawn tiles are left untouched. */
int luxgi_surface_light(const LuxgiSurfaceLightParams* params, const LuxgiLight* lights,
                        const LuxgiAtlasTile* tiles, int32_t numTiles, const LuxgiSurfaceDraw* draws, int32_t numDraws,
                        const LuxgiSurfaceGBuffer* gbuffer, const LuxgiGlobalSDFData* sdfData,
                        const LuxgiSDFVolume* sdf, const LuxgiDDGIUniform* ddgi, const LuxgiProbeAtlas* probes,
                        int32_t outStorage, void* lightCache, void* stream);

/* ------------------------------------------------------------------------- */
/* SVGF (reflection = RGB, shadow = scalar)                                   */
/* ------------------------------------------------------------------------- */

/* G-buffer channel contract (DeferredColor.frag:98-101, GBuffer.cpp:15-24):
 * color   = (albedo.rgb, meshId)            RGBA32F
 * normals = (octNormal.xy, prevUV - curUV)  RGBA32F
 * pbr     = (metallic, roughness, curvature, linearZ) RGBA32F
 * depth   = NDC depth, 1 = cleared / sky    R32F   */
typedef struct LuxgiGBuffer {
    const float* color;
    const float* normals;
    const float* pbr;
    const float* depth;
} LuxgiGBuffer;

/* Reflection/DenoiseReprojection.comp UBO (:23-28) + push constants (:67-73). */
typedef struct LuxgiReflectionReprojectParams {
    float   viewProjInv[16];
    float   prevViewProj[16];
    float   cameraPos[4];
    float   cameraDelta[3];
    float   alpha;
    float   momentsAlpha;
    int32_t approximateWithDDGI;
} LuxgiReflectionReprojectParams;

/* Reflection/DenoiseAtrous.comp push constants (:31-39). */
typedef struct LuxgiReflectionAtrousParams {
    int32_t radius;
    int32_t stepSize;
    float   phiColor;
    float   phiNormal;
    float   sigmaDepth;
    int32_t approximateWithDDGI;
} LuxgiReflectionAtrousParams;

/* Shadow/DenoiseReprojection.comp UBO (:24-27) + push constants (:61-65). */
typedef struct LuxgiShadowReprojectParams {
    float viewProjInv[16];
    float alpha;
    float momentsAlpha;
    float pad[2];
} LuxgiShadowReprojectParams;

/* Shadow/DenoiseAtrous.comp push constants (:29-39). */
typedef struct LuxgiShadowAtrousParams {
    int32_t radius;
    int32_t stepSize;
    float   phiVisibility;
    float   phiNormal;
    float   sigmaDepth;
    float   power;
    float   near_;
    float   far_;
} LuxgiShadowAtrousParams;

/* Number of 8x8 tiles for a width x height image (the reference's tile lists,
 * RaytracedReflection.cpp:658-668); the tile map is one byte per tile. */
int32_t luxgi_svgf_tile_count(int32_t width, int32_t height);

/* Reflection/DenoiseReprojection.comp:117-218 (temporal accumulation, moments,
 * 17x17 history clip, tile classification). outColor = (rgb, variance),
 * outMoments = (m1, m2, historyLength, 0). tileMap[t] = 1 denoise / 0 copy
 * (deterministic OR over the tile; the reference's flag is last-writer-wins).
 * Images: `storage` FP32 (RGBA32F) or FP16 (RGBA16F), width x height. */
int luxgi_svgf_reflection_reproject(const LuxgiReflectionReprojectParams* params,
                                    int32_t width, int32_t height, int32_t storage,
                                    const LuxgiGBuffer* cur, const LuxgiGBuffer* prev,
                                    const void* noisy, const void* historyColor, const void* historyMoments,
                                    void* outColor, void* outMoments, uint8_t* tileMap,
                                    void* stream);

/* One a-trous iteration (Reflection/DenoiseAtrous.comp:63-134) fused with the
 * copy-tiles pass (Reflection/DenoiseCopyTiles.comp:21-25): denoise tiles are
 * filtered, copy tiles copied. historyOut (nullable) receives the same output
 * (the history copy after iteration 1, RaytracedReflection.cpp:488-491). */
int luxgi_svgf_reflection_atrous(const LuxgiReflectionAtrousParams* params,
                                 int32_t width, int32_t height, int32_t storage,
                                 const LuxgiGBuffer* cur, const uint8_t* tileMap,
                                 const void* input, void* output, void* historyOut,
                                 void* stream);

/* Whole reflection denoise of one frame (denoise::system, RaytracedReflection.cpp:496-520):
 * reprojection then `iterations` a-trous passes (step 1<<i), history copy after
 * iteration 1. outputs[0] = accumulation target, outputs[1] = history colour;
 * moments[pingPong] is written, moments[1-pingPong] read; pingPong[0/1] buffers
 * of the a-trous chain. *result receives which atrous buffer holds the output. */
int luxgi_svgf_reflection_denoise(const LuxgiReflectionReprojectParams* reprojectParams,
                                  const LuxgiReflectionAtrousParams* atrousParams,
                                  int32_t iterations,
                                  int32_t width, int32_t height, int32_t storage,
                                  const LuxgiGBuffer* cur, const LuxgiGBuffer* prev,
                                  const void* noisy,
                                  void* output0, void* output1,
                                  void* moments0, void* moments1, int32_t pingPong,
                                  void* atrous0, void* atrous1,
                                  uint8_t* tileMap, int32_t* result,
                                  void* stream);

/* Shadow/DenoiseReprojection.comp:71-210. mask = R32UI [ceil(H/4)][ceil(W/8)],
 * bit (y%4)*8+(x%8) of word (y/4, x/8) = visibility of pixel (x, y).
 * outColor = (visibility, variance) RG, outMoments RGBA (m1, m2, historyLength, 0). */
int luxgi_svgf_shadow_reproject(const LuxgiShadowReprojectParams* params,
                                int32_t width, int32_t height, int32_t storage,
                                const LuxgiGBuffer* cur, const LuxgiGBuffer* prev,
                                const uint32_t* mask, const void* historyColor, const void* historyMoments,
                                void* outColor, void* outMoments, uint8_t* tileMap,
                                void* stream);

/* Shadow/DenoiseAtrous.comp:65-128 fused with Shadow/DenoiseCopyTiles.comp:18-22
 * (copy tiles write 0). */
int luxgi_svgf_shadow_atrous(const LuxgiShadowAtrousParams* params,
                             int32_t width, int32_t height, int32_t storage,
                             const LuxgiGBuffer* cur, const uint8_t* tileMap,
                             const void* input, void* output, void* historyOut,
                             void* stream);

/* raytraced_shadow::denoise::system (RaytracedShadow.cpp:543-609) for one frame: the reprojection
 * (accumulation, :418-462) followed by `iterations` a-trous passes (atrousFilter, :464-541).
 *   output0/output1   RG outputs[0] (reprojection result) and outputs[1] (history colour), [H][W]
 *   moments0/1        RGBA currentMoments[2]; the pass writes moments[pingPong] and reads
 *                     moments[1 - pingPong] (the caller flips pingPong after the call, :565)
 *   atrous0/1         RG a-trous ping-pong images; iteration i reads atrous[i % 2 == 1 ? 1 : 0]
 *                     (output0 for i = 0) and writes the other
 *   ap->power         applied on the last iteration only (:510); stepSize is set per pass
 * The output of iteration 1 is copied into output1 (:532-535). *result receives the index of
 * the atrous image holding the final output (-1 with iterations = 0: output0). */
int luxgi_svgf_shadow_denoise(const LuxgiShadowReprojectParams* reprojectParams,
                              const LuxgiShadowAtrousParams* atrousParams, int32_t iterations,
                              int32_t width, int32_t height, int32_t storage,
                              const LuxgiGBuffer* cur, const LuxgiGBuffer* prev, const uint32_t* mask,
                              void* output0, void* output1, void* moments0, void* moments1, int32_t pingPong,
                              void* atrous0, void* atrous1, uint8_t* tileMap, int32_t* result,
                              void* stream);

/* Library version string. */
const char* luxgi_version(void);

/* ------------------------------------------------------------------------- */
/* Noisy-signal producers traced through the global SDF                       */
/* ------------------------------------------------------------------------- */

/* Blue-noise tables of sampleBlueNoise (Raytraced/BlueNoise.glsl:8-18), RGBA8 texels as the
 * engine loads them (Engine/Noise/BlueNoise.h:15-26): sobol_256_4d ([256]) and
 * scrambling_ranking_128x128_2d_1spp ([128][128]). Device memory. */
typedef struct LuxgiBlueNoise {
    const uint8_t* sobol;
    const uint8_t* scrambling;
} LuxgiBlueNoise;

/* SDFShadow.comp UBO (:24-31) plus the march step cap. 144 bytes. */
typedef struct LuxgiSdfShadowParams {
    LuxgiLight light;           /* light.direction.w = light source radius (soft shadows) */
    float      viewProjInv[16];
    uint32_t   numFrames;       /* blue-noise sample index (the frame counter) */
    float      shadowBias;
    int32_t    maxSteps;        /* 0 = 250 (SDFCommon.glsl:143) */
    int32_t    pad;
} LuxgiSdfShadowParams;

/* SDFShadow.comp:113-157 (dispatch RaytracedShadow.cpp:340-356): one soft-shadow ray per pixel of
 * the G-buffer (depth R32F, normals RGBA32F with the octahedral normal in xy) toward the light's
 * disk point picked by the blue-noise sampler, traced through the global SDF. mask holds one
 * R32UI word per 8x4 pixel block ([ceil(H/4)][ceil(W/8)]), bit x + 8y set where the ray reaches
 * the light; a word is written only when its block's first pixel is not sky (depth 1), as in
 * the reference. Pixels past the image edge read the G-buffer with repeat addressing (the
 * default sampler) at their unwrapped screen position, as the reference's invocations do. */
int luxgi_sdf_shadow_mask(const LuxgiSdfShadowParams* params, const LuxgiBlueNoise* noise, int32_t width,
                          int32_t height, const float* depth, const float* normals,
                          const LuxgiGlobalSDFData* sdfData, const LuxgiSDFVolume* sdf, uint32_t* mask,
                          void* stream);

/* SDFReflection.comp push constants (:64-76, 112 bytes) followed by options that extend the reference. 144 bytes. */
typedef struct LuxgiSdfReflectionParams {
    float    bias;                /* unused by the reference's shader, kept for the layout */
    float    trim;                /* scales the blue-noise sample of the GGX lobe */
    float    intensity;           /* unused by the reference's shader */
    float    roughDDGIIntensity;  /* scale of the DDGI estimate for rough pixels */
    uint32_t numLights;           /* unused by the reference's shader */
    uint32_t numFrames;           /* blue-noise sample index */
    uint32_t sampleGI;            /* unused by the reference's shader */
    uint32_t approximateWithDDGI; /* 1: roughness > 0.45 reads the probes instead of tracing */
    float    cameraPosition[4];
    float    viewProjInv[16];
    /* --- extensions --- */
    int32_t  maxSteps;            /* 0 = 250 (SDFCommon.glsl:143) */
    int32_t  pad[3];
    float    skyColor[4];         /* sky when no cube map is bound (the reference binds a unit cube) */
} LuxgiSdfReflectionParams;

/* Workspace bytes of luxgi_sdf_reflection for an atlas of `objectCapacity` objects. */
size_t luxgi_sdf_reflection_workspace_size(int32_t objectCapacity);

/* SDFReflection.comp:116-162 (dispatch RaytracedReflection.cpp:307-344): per non-sky pixel of the
 * G-buffer (normals and pbr RGBA32F, depth R32F; the colour image is not read), a reflection ray
 * through the global SDF from world + normal * voxel: the mirror direction below roughness 0.05,
 * a GGX half vector from the blue-noise sampler up to 0.45 (or beyond, unless
 * approximateWithDDGI), and for approximateWithDDGI above 0.45 the DDGI irradiance in the mirror
 * direction instead of a ray. A ray's colour is the surface-atlas lookup at the hit (normal -R,
 * as the trace's `sampleGlobalSurfaceAtlas`: rgb / max(a, 1e-4), a), or the sky. Writes
 * out [H][W] RGBA (FP32 or FP16) at non-sky pixels only; the SVGF reprojection reads its alpha as
 * the ray length. `ddgi` / `probes` may be null unless approximateWithDDGI is 1. */
int luxgi_sdf_reflection(const LuxgiSdfReflectionParams* params, const LuxgiBlueNoise* noise, int32_t width,
                         int32_t height, const LuxgiGBuffer* gbuffer, const LuxgiGlobalSDFData* sdfData,
                         const LuxgiSDFVolume* sdf, const LuxgiSurfaceAtlas* atlas, const LuxgiSkybox* sky,
                         const LuxgiDDGIUniform* ddgi, const LuxgiProbeAtlas* probes, int32_t outStorage, void* out,
                         void* workspace, size_t workspaceBytes, void* stream);

/* ------------------------------------------------------------------------- */
/* Several probe volumes in one trace (no counterpart in the reference)        */
/* ------------------------------------------------------------------------- */

#define LUXGI_MAX_TRACE_BATCH 256

/* One probe volume of a batched trace: the per-volume arguments of luxgi_ddgi_trace_sdf. */
typedef struct LuxgiTraceVolume {
    LuxgiDDGIUniform uniform;
    float    orientation[16];          /* randomOrientation of this volume (column-major) */
    const float* deviceOrientation;    /* optional, as LuxgiTraceParams.deviceOrientation */
    const float* probeOffsets;         /* optional float4[P] */
    LuxgiRayBuffers rays;
    uint8_t* rayFlags;                 /* optional uint8[P*width] */
    int64_t  rayStart;                 /* first ray in the batch's flat ray list; set by
                                          luxgi_ddgi_trace_batch_prepare / luxgi_ddgi_trace_scenes_prepare */
    int32_t  scene;                    /* several-scene batches: index of the volume's scene */
    int32_t  pad;
} LuxgiTraceVolume;                    /* 232 B */

/* Host-side: checks `count` host descriptors (grid, ray counts against the ray buffers, one ray
 * storage `rayStorage` for all), sets each rayStart and returns the total ray count. No GPU call. */
int luxgi_ddgi_trace_batch_prepare(LuxgiTraceVolume* volumes, int32_t count, int32_t rayStorage,
                                   int64_t* totalRays);

/* Workspace bytes of luxgi_ddgi_trace_sdf_batched for `totalRays` rays. */
size_t luxgi_ddgi_trace_batch_workspace_size(int64_t totalRays, int32_t objectCapacity);

/* luxgi_ddgi_trace_sdf for several volumes over one scene in one launch sequence: the march, the
 * hit sort and the shade each run once over the flat list of all volumes' rays, each ray finding
 * its volume by binary search over rayStart. Every volume's outputs are identical to a
 * luxgi_ddgi_trace_sdf call on it. `volumes` is a device array of `count` prepared descriptors
 * (read when the kernels run); `params` gives the shared settings (maxSteps, hitEpsilon,
 * shadeOrder, skyColor; its orientations are ignored). */
int luxgi_ddgi_trace_sdf_batched(const LuxgiTraceVolume* volumes, int32_t count, int64_t totalRays,
                                 int32_t rayStorage, const LuxgiGlobalSDFData* sdfData, const LuxgiSDFVolume* volume,
                                 const LuxgiSurfaceAtlas* atlas, const LuxgiSkybox* sky,
                                 const LuxgiTraceParams* params, void* workspace, size_t workspaceBytes,
                                 void* stream);

/* Several probe volumes from several scenes in one launch sequence (no counterpart in the
 * reference, which traces one volume per dispatch, DDGIRenderer.cpp:318-321): each volume names its
 * scene (LuxgiTraceVolume.scene), each scene carries the scene arguments of luxgi_ddgi_trace_sdf.
 * All scenes share one SDF storage and layout and one atlas layout and colour storage; the call's
 * LuxgiTraceParams gives maxSteps and hitEpsilon (orientations are per volume); hits are shaded in
 * ray order. Per volume the results are those of luxgi_ddgi_trace_sdf on its scene. */
typedef struct LuxgiTraceScene {
    LuxgiGlobalSDFData sdfData;
    LuxgiSDFVolume     volume;
    LuxgiSurfaceAtlas  atlas;
    LuxgiSkybox        skybox;
    float              skyColor[4];    /* as LuxgiTraceParams.skyColor */
} LuxgiTraceScene;

/* Filled by luxgi_ddgi_trace_scenes_prepare; passed (host memory) to the trace. */
typedef struct LuxgiTraceSceneBatch {
    int64_t totalRays;      /* flat rays including the padding that starts each volume on a multiple of 64 */
    int64_t totalObjects;   /* atlas objects of all scenes */
    int32_t sceneCount, volumeCount, maxObjects, rayStorage;
    int32_t sdfStorage, sdfLayout, atlasLayout, colorStorage;
} LuxgiTraceSceneBatch;

/* Bytes of the device descriptors of `sceneCount` scenes. */
size_t luxgi_ddgi_trace_scenes_descriptor_bytes(int32_t sceneCount);

/* Host-side: checks the scenes and the volumes (scene indices, grids, ray buffers of storage
 * `rayStorage`), sets each volume's rayStart (multiples of 64), writes the scenes' device descriptors
 * into `descriptors` (host memory of luxgi_ddgi_trace_scenes_descriptor_bytes; the caller copies it
 * to device memory) and fills `batch`. No GPU call. */
int luxgi_ddgi_trace_scenes_prepare(const LuxgiTraceScene* scenes, int32_t sceneCount,
                                    LuxgiTraceVolume* volumes, int32_t volumeCount, int32_t rayStorage,
                                    void* descriptors, LuxgiTraceSceneBatch* batch);

/* Workspace bytes of luxgi_ddgi_trace_sdf_scenes. */
size_t luxgi_ddgi_trace_scenes_workspace_size(const LuxgiTraceSceneBatch* batch);

/* The trace: `volumes` and `descriptors` in device memory (as prepared), `batch` and `params` on
 * the host. Calls that overlap on one workspace race. */
int luxgi_ddgi_trace_sdf_scenes(const LuxgiTraceSceneBatch* batch, const LuxgiTraceVolume* volumes,
                                const void* descriptors, const LuxgiTraceParams* params,
                                void* workspace, size_t workspaceBytes, void* stream);

/* ------------------------------------------------------------------------- */
/* Batched reflection denoise (no counterpart in the reference)               */
/* ------------------------------------------------------------------------- */

/* Largest batch: frames per luxgi_svgf_reflection_denoise_batched call. */
#define LUXGI_MAX_DENOISE_BATCH 1024

/* One frame of a batch: the arguments of luxgi_svgf_reflection_denoise for it. The array of items
 * lives in device memory, so a captured hipGraph replays with whatever the caller copied there
 * before the replay (new cameras, noisy inputs, the flipped pingPong, even other sizes). */
typedef struct LuxgiReflectionDenoiseItem {
    LuxgiReflectionReprojectParams reproject;  /* 168 B */
    LuxgiReflectionAtrousParams filter;        /* a-trous parameters; stepSize is set per iteration */
    int32_t width, height;
    int32_t pingPong;                          /* moments[pingPong] written, moments[1 - pingPong] read */
    int32_t blockStart;                        /* first 16x16 work unit; set by luxgi_svgf_batch_prepare */
    LuxgiGBuffer cur, prev;
    const void* noisy;
    void* outputs[2];                          /* accumulation target, history colour */
    void* moments[2];
    void* atrous[2];                           /* the a-trous ping-pong pair */
    uint8_t* tileMap;                          /* luxgi_svgf_tile_count(width, height) bytes */
} LuxgiReflectionDenoiseItem;                  /* 336 B */

typedef struct LuxgiDenoiseBatchHeader {
    int32_t count;        /* items */
    int32_t totalBlocks;  /* 16x16 work units over all items */
    int32_t pad[2];
} LuxgiDenoiseBatchHeader;

/* Host-side: checks `count` host items (sizes, radius, pointers), sets each blockStart and fills
 * the header. The caller then copies items and header to device memory. No GPU call. */
int luxgi_svgf_batch_prepare(LuxgiReflectionDenoiseItem* items, int32_t count, LuxgiDenoiseBatchHeader* header);

/* denoise::system (RaytracedReflection.cpp:496-520) for every item of a batch of frames of any
 * sizes: the reprojection of all frames in one persistent launch, then each a-trous iteration of
 * all frames in one persistent launch (1 + iterations launches in all, whatever the count). Each
 * item's outputs are identical to luxgi_svgf_reflection_denoise on it. header and items are device
 * pointers (prepared as above); gridBlocks (> 0) is the grid and `order` how blocks walk the work
 * units: 0 persistent and strided (block b takes b, b + grid, ...), 1 persistent with contiguous
 * runs, 2 one unit per block (no loop; gridBlocks must be at least the header's totalBlocks when
 * the launch runs, units beyond the grid are not processed). *result as for
 * luxgi_svgf_reflection_denoise (the same for every item). */
int luxgi_svgf_reflection_denoise_batched(const LuxgiDenoiseBatchHeader* header, const LuxgiReflectionDenoiseItem* items,
                                          int32_t iterations, int32_t storage, int32_t gridBlocks, int32_t order,
                                          int32_t* result, void* stream);

/* ------------------------------------------------------------------------- */
/* Learned a-trous edge-stopping weights (no counterpart in the reference)     */
/* ------------------------------------------------------------------------- */

#define LUXGI_LEARNED_ATROUS_PARAMS 81 /* W1 [10][6], b1 [10], w2 [10], b2 */

/* One pass of the learned filter: the parameters of Reflection/DenoiseAtrous.comp plus the level. */
typedef struct LuxgiLearnedAtrousParams {
    int32_t radius;      /* 0 .. 2 */
    int32_t stepSize;    /* 1 << iteration */
    float   phiColor;
    float   sigmaDepth;
    float   level;       /* network input: iteration / (iterations - 1) */
    int32_t pad[3];
} LuxgiLearnedAtrousParams;

/* The reflection a-trous pass (Reflection/DenoiseAtrous.comp:63-134: 3x3 variance at the centre,
 * sky outputs 0, roughness outside (0.05, 0.45) passes through) with each tap's edge-stopping x
 * kernel weight replaced by exp(MLP(features)), the MLP of the 81 `weights` (one hidden layer of 10
 * ReLU units over the tap's luminance difference / phi, depth term, normal term, kernel weight, level
 * and the centre's standard deviation). in / out [H][W] RGBA32F (rgb, variance); gbuffer normals and
 * pbr as for luxgi_svgf_reflection_atrous. record (nullable, [H][W][4] floats) keeps what
 * luxgi_svgf_learned_atrous_backward needs per pixel: (1 + the tap weights, luminance, 1 / phi,
 * min(sqrt(variance), 4)) of a filtered pixel, (1, 0, 0, 0) of the others. Footprints
 * radius * stepSize <= 8 are staged in LDS. */
int luxgi_svgf_learned_atrous(const LuxgiLearnedAtrousParams* params, const float* weights, int32_t width,
                              int32_t height, const LuxgiGBuffer* gbuffer, const float* in, float* out, float* record,
                              void* stream);

/* Backward of one learned pass: from dL/d(out rgb) (gradOut [H][W][4], alpha ignored) writes
 * dL/d(in rgb) into gradIn ([H][W][4], alpha 0) and adds dL/d(weights) into gradWeights [81]. The
 * weights' dependence on the input colour (through luminance and variance) is held fixed; `record`
 * is the forward's. */
int luxgi_svgf_learned_atrous_backward(const LuxgiLearnedAtrousParams* params, const float* weights, int32_t width,
                                       int32_t height, const LuxgiGBuffer* gbuffer, const float* in, const float* out,
                                       const float* record, const float* gradOut, float* gradIn, float* gradWeights,
                                       void* stream);

/* One Adam step on n parameters (step counts from 1); clears `grad`. */
int luxgi_adam_step(float* weights, float* grad, float* m, float* v, int32_t n, float lr, float beta1, float beta2,
                    float eps, int32_t step, void* stream);

/* ------------------------------------------------------------------------- */
/* Per-pixel learned edge-stopping scales (no counterpart in the reference)   */
/* ------------------------------------------------------------------------- */
/* A 12 -> 32 -> 32 -> 5 ReLU MLP (FP16 weights, FP32 accumulation) maps per-pixel features of the
 * pass-0 input, the reprojection's moments and the G-buffer to one scale per a-trous pass,
 * s_i = exp(clamp(o_i, -4, 4)); pass i is the reflection a-trous pass (Reflection/DenoiseAtrous.comp
 * :63-134) with the luminance term of every tap's weight (EdgeStop.glsl:24-40) multiplied by s_i(p)
 * (luxgi_svgf_reflection_atrous_scaled). Zero output weights and biases give s = 1: the reference
 * chain, bit for bit. Parameters (FP32 master copy, rounded to FP16 by the forward): W1 [32][12],
 * b1 [32], W2 [32][32], b2 [32], W3 [5][32], b3 [5]. */
#define LUXGI_LEARNED_PIXEL_PARAMS 1637
#define LUXGI_LEARNED_PIXEL_FEATURES 12
#define LUXGI_LEARNED_PIXEL_PASSES 5

/* luxgi_svgf_reflection_atrous with lumScale[y*W+x] multiplying the luminance term of pixel
 * (x, y)'s taps (lumScale of 1 everywhere = luxgi_svgf_reflection_atrous, bit for bit). */
int luxgi_svgf_reflection_atrous_scaled(const LuxgiReflectionAtrousParams* params, int32_t width, int32_t height,
                                        int32_t storage, const LuxgiGBuffer* cur, const uint8_t* tileMap,
                                        const void* input, void* output, void* historyOut, const float* lumScale,
                                        void* stream);

/* The scales of the five passes, planar [5][H][W] FP32, from the pass-0 input `color` (RGBA32F:
 * rgb + variance) and the reprojection's `moments` (RGBA32F: m1, m2, history length). `features`
 * (optional, [H][W][12] FP32) receives the features for the backward. */
int luxgi_learned_pixel_scales(const float* params, int32_t width, int32_t height, const LuxgiGBuffer* cur,
                               const float* color, const float* moments, float* scales, float* features,
                               void* stream);

/* Training: the backward of one scaled pass (FP32 images) from dL/d(output rgb): dL/d(input rgb)
 * (variance channel 0) and dL/d(lumScale). The weights' dependence on the input (luminance, the
 * variance behind phi) is held fixed. gradIn must not alias input or gradOut. */
size_t luxgi_learned_pixel_pass_backward_workspace_size(int32_t width, int32_t height);
int luxgi_learned_pixel_pass_backward(const LuxgiReflectionAtrousParams* params, int32_t width, int32_t height,
                                      const LuxgiGBuffer* cur, const uint8_t* tileMap, const float* input,
                                      const float* lumScale, const float* gradOut, float* gradIn, float* gradScale,
                                      void* workspace, size_t workspaceBytes, void* stream);

/* Training: adds dL/d(parameters) into gradParams from dL/d(scales) ([5][H][W]), the forward's
 * scales and features. Deterministic (per-block partial sums added in block order). */
size_t luxgi_learned_pixel_mlp_backward_workspace_size(void);
int luxgi_learned_pixel_mlp_backward(const float* params, int32_t width, int32_t height, const float* features,
                                     const float* scales, const float* gradScales, float* gradParams,
                                     void* workspace, size_t workspaceBytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LUXGI_H */

"""ctypes mirrors of the POD structs declared in include/luxgi.h.

The layouts are byte-identical to the reference's GPU structs (sizes are
asserted at import time):
  DDGIUniform            Shaders/DDGI/DDGICommon.glsl:11-31        96 B
  GlobalSDFData          Shaders/SDF/GlobalSDFData.glsl:4-12       96 B
  GlobalSurfaceAtlasData Shaders/SDF/AtlasCommon.glsl:8-16         32 B
  ObjectBuffer           Shaders/SDF/AtlasCommon.glsl:18-25       128 B
  TileBuffer             Shaders/SDF/AtlasCommon.glsl:27-32        96 B
  GISDFRays push consts  Shaders/DDGI/GISDFRays.comp:53-60         80 B (+ extensions)
"""
import ctypes as C

import numpy as np

LUXGI_OK = 0
STORAGE_FP32 = 0
STORAGE_FP16 = 1
STORAGE_UNORM8 = 2
STORAGE_BF16 = 3

UPDATE_IRRADIANCE = 0x1
UPDATE_DEPTH = 0x2
UPDATE_FUSE_BORDER = 0x4
UPDATE_PER_PROBE = 0x8
UPDATE_HALF = 0x10
UPDATE_BF16 = 0x20

GLOBAL_SDF_WORLD_SIZE = 60000.0  # Shaders/SDF/SDFCommon.glsl:6
ATLAS_CHUNKS_RESOLUTION = 40     # Shaders/SDF/AtlasCommon.glsl:5

f4 = C.c_float * 4
f16 = C.c_float * 16
i4 = C.c_int32 * 4


class DDGIUniform(C.Structure):
    _fields_ = [
        ("startPosition", f4), ("step", f4), ("probeCounts", i4),
        ("maxDistance", C.c_float), ("sharpness", C.c_float), ("hysteresis", C.c_float), ("normalBias", C.c_float),
        ("ddgiGamma", C.c_float), ("irradianceProbeSideLength", C.c_int32),
        ("irradianceTextureWidth", C.c_int32), ("irradianceTextureHeight", C.c_int32),
        ("depthProbeSideLength", C.c_int32), ("depthTextureWidth", C.c_int32),
        ("depthTextureHeight", C.c_int32), ("raysPerProbe", C.c_int32),
    ]

    @property
    def num_probes(self):
        return self.probeCounts[0] * self.probeCounts[1] * self.probeCounts[2]


class GlobalSDFData(C.Structure):
    _fields_ = [
        ("cascadePosDistance", f4 * 4), ("cascadeVoxelSize", f4), ("cascadesCount", C.c_uint32),
        ("resolution", C.c_float), ("nearPlane", C.c_float), ("farPlane", C.c_float),
    ]


class SurfaceAtlasData(C.Structure):
    _fields_ = [
        ("cameraPos", C.c_float * 3), ("chunkSize", C.c_float), ("culledObjectsCapacity", C.c_uint32),
        ("resolution", C.c_uint32), ("objectsCount", C.c_uint32), ("padding", C.c_uint32),
    ]


class AtlasObject(C.Structure):
    _fields_ = [
        ("objectBounds", f4), ("tileOffset", C.c_uint32 * 6), ("padding", C.c_int32 * 2),
        ("transform", f16), ("extends", f4),
    ]


class AtlasTile(C.Structure):
    _fields_ = [("extends", f4), ("transform", f16), ("objectBounds", f4)]


class TraceParams(C.Structure):
    _fields_ = [
        ("randomOrientation", f16), ("numFrames", C.c_uint32), ("infiniteBounces", C.c_uint32),
        ("numLights", C.c_int32), ("intensity", C.c_float),
        ("maxSteps", C.c_int32), ("hitEpsilon", C.c_float), ("shadeOrder", C.c_int32), ("pad1", C.c_int32),
        ("skyColor", f4), ("deviceOrientation", C.c_void_p),
    ]


LAYOUT_LINEAR = 0
LAYOUT_BRICK4 = 1
LAYOUT_CELL8 = 2


class SDFVolume(C.Structure):
    _fields_ = [("sdf", C.c_void_p), ("mip", C.c_void_p), ("resolution", C.c_int32),
                ("mipResolution", C.c_int32), ("storage", C.c_int32), ("layout", C.c_int32), ("guard", C.c_void_p),
                ("brickScale", C.c_void_p), ("refine", C.c_void_p), ("refineLayout", C.c_int32), ("pad", C.c_int32)]


class SurfaceAtlas(C.Structure):
    _fields_ = [
        ("data", SurfaceAtlasData), ("chunks", C.c_void_p), ("culledObjects", C.c_void_p),
        ("objects", C.c_void_p), ("tiles", C.c_void_p), ("color", C.c_void_p), ("depth", C.c_void_p),
        ("objectCapacity", C.c_int32), ("colorStorage", C.c_int32), ("layout", C.c_int32), ("pad", C.c_int32),
    ]


class Skybox(C.Structure):
    _fields_ = [("faces", C.c_void_p), ("size", C.c_int32), ("pad", C.c_int32)]


class RayBuffers(C.Structure):
    _fields_ = [("radiance", C.c_void_p), ("directionDistance", C.c_void_p), ("width", C.c_int32),
                ("height", C.c_int32), ("storage", C.c_int32), ("pad", C.c_int32)]


class ProbeAtlas(C.Structure):
    _fields_ = [("irradiance", C.c_void_p), ("depth", C.c_void_p), ("storage", C.c_int32), ("pad", C.c_int32),
                ("probeOffsets", C.c_void_p), ("probeStates", C.c_void_p)]


class RelocateParams(C.Structure):
    _fields_ = [("backfaceThreshold", C.c_float), ("minFrontfaceDistance", C.c_float),
                ("maxOffsetFraction", C.c_float), ("relocate", C.c_int32)]


class SampleProbeParams(C.Structure):
    _fields_ = [("cameraPosition", f4), ("viewProjInv", f16)]


class GBuffer(C.Structure):
    _fields_ = [("color", C.c_void_p), ("normals", C.c_void_p), ("pbr", C.c_void_p), ("depth", C.c_void_p)]


class ReflectionReprojectParams(C.Structure):
    _fields_ = [("viewProjInv", f16), ("prevViewProj", f16), ("cameraPos", f4), ("cameraDelta", C.c_float * 3),
                ("alpha", C.c_float), ("momentsAlpha", C.c_float), ("approximateWithDDGI", C.c_int32)]


class ReflectionAtrousParams(C.Structure):
    _fields_ = [("radius", C.c_int32), ("stepSize", C.c_int32), ("phiColor", C.c_float), ("phiNormal", C.c_float),
                ("sigmaDepth", C.c_float), ("approximateWithDDGI", C.c_int32)]


class ShadowReprojectParams(C.Structure):
    _fields_ = [("viewProjInv", f16), ("alpha", C.c_float), ("momentsAlpha", C.c_float), ("pad", C.c_float * 2)]


class ShadowAtrousParams(C.Structure):
    _fields_ = [("radius", C.c_int32), ("stepSize", C.c_int32), ("phiVisibility", C.c_float),
                ("phiNormal", C.c_float), ("sigmaDepth", C.c_float), ("power", C.c_float),
                ("near_", C.c_float), ("far_", C.c_float)]


class Light(C.Structure):  # Shaders/Common/Light.glsl:19-28
    _fields_ = [("color", f4), ("position", f4), ("direction", f4), ("intensity", C.c_float),
                ("radius", C.c_float), ("type", C.c_float), ("angle", C.c_float)]


LIGHT_DIRECTIONAL, LIGHT_SPOT, LIGHT_POINT = 0.0, 1.0, 2.0  # Light.glsl:13-15
MAX_LIGHTS = 32
SURFACE_EMISSIVE = 0x1
SURFACE_DIRECT = 0x2
SURFACE_INDIRECT = 0x4


class SurfaceGBuffer(C.Structure):
    _fields_ = [("color", C.c_void_p), ("normal", C.c_void_p), ("pbr", C.c_void_p), ("emissive", C.c_void_p),
                ("resolution", C.c_int32), ("pad", C.c_int32)]


class SurfaceLightParams(C.Structure):
    _fields_ = [("cameraPosition", f4), ("indirectIntensity", C.c_float), ("passes", C.c_uint32),
                ("numLights", C.c_int32), ("maxSteps", C.c_int32)]


class SurfaceDraw(C.Structure):
    _fields_ = [("tile", C.c_uint32), ("lightMask", C.c_uint32)]


class BlueNoise(C.Structure):  # Raytraced/BlueNoise.glsl tables (RGBA8)
    _fields_ = [("sobol", C.c_void_p), ("scrambling", C.c_void_p)]


class SdfShadowParams(C.Structure):  # Shaders/SDF/SDFShadow.comp:24-31
    _fields_ = [("light", Light), ("viewProjInv", f16), ("numFrames", C.c_uint32), ("shadowBias", C.c_float),
                ("maxSteps", C.c_int32), ("pad", C.c_int32)]


class SdfReflectionParams(C.Structure):  # Shaders/SDF/SDFReflection.comp:64-76 + extensions
    _fields_ = [("bias", C.c_float), ("trim", C.c_float), ("intensity", C.c_float), ("roughDDGIIntensity", C.c_float),
                ("numLights", C.c_uint32), ("numFrames", C.c_uint32), ("sampleGI", C.c_uint32),
                ("approximateWithDDGI", C.c_uint32), ("cameraPosition", f4), ("viewProjInv", f16),
                ("maxSteps", C.c_int32), ("pad", C.c_int32 * 3), ("skyColor", f4)]


class ObjectRasterizeData(C.Structure):  # Shaders/SDF/ObjectRasterizeData.glsl
    _fields_ = [("worldToVolume", f16), ("volumeToWorld", f16), ("volumeToUVWMul", C.c_float * 3),
                ("mipOffset", C.c_float), ("volumeToUVWAdd", C.c_float * 3), ("decodeMul", C.c_float),
                ("volumeLocalBoundsExtent", C.c_float * 3), ("decodeAdd", C.c_float)]


class RasterizeParams(C.Structure):  # SDFRasterizeModel.glsl:16-24
    _fields_ = [("cascadeCoordToPosMul", C.c_float * 3), ("maxDistance", C.c_float),
                ("cascadeCoordToPosAdd", C.c_float * 3), ("cascadeResolution", C.c_int32),
                ("cascadeIndex", C.c_int32), ("pad", C.c_int32 * 3)]


class MeshSDF(C.Structure):
    _fields_ = [("offset", C.c_uint64 * 3), ("size", (C.c_int32 * 3) * 3), ("mipCount", C.c_int32),
                ("pad", C.c_int32 * 2)]


class LearnedAtrousParams(C.Structure):  # include/luxgi.h LuxgiLearnedAtrousParams
    _fields_ = [("radius", C.c_int32), ("stepSize", C.c_int32), ("phiColor", C.c_float), ("sigmaDepth", C.c_float),
                ("level", C.c_float), ("pad", C.c_int32 * 3)]


LEARNED_ATROUS_PARAMS = 81
LEARNED_PIXEL_PARAMS = 1637
LEARNED_PIXEL_FEATURES = 12
LEARNED_PIXEL_PASSES = 5

MAX_TRACE_BATCH = 256


class TraceVolume(C.Structure):  # include/luxgi.h LuxgiTraceVolume
    _fields_ = [("uniform", DDGIUniform), ("orientation", C.c_float * 16), ("deviceOrientation", C.c_void_p),
                ("probeOffsets", C.c_void_p), ("rays", RayBuffers), ("rayFlags", C.c_void_p),
                ("rayStart", C.c_int64), ("scene", C.c_int32), ("pad", C.c_int32)]


class TraceScene(C.Structure):  # include/luxgi.h LuxgiTraceScene
    _fields_ = [("sdfData", GlobalSDFData), ("volume", SDFVolume), ("atlas", SurfaceAtlas), ("skybox", Skybox),
                ("skyColor", C.c_float * 4)]


class TraceSceneBatch(C.Structure):  # include/luxgi.h LuxgiTraceSceneBatch
    _fields_ = [("totalRays", C.c_int64), ("totalObjects", C.c_int64), ("sceneCount", C.c_int32),
                ("volumeCount", C.c_int32), ("maxObjects", C.c_int32), ("rayStorage", C.c_int32),
                ("sdfStorage", C.c_int32), ("sdfLayout", C.c_int32), ("atlasLayout", C.c_int32),
                ("colorStorage", C.c_int32)]


MAX_UPDATE_BATCH = 4096


class UpdateVolume(C.Structure):  # include/luxgi.h LuxgiUpdateVolume
    _fields_ = [("uniform", DDGIUniform), ("rays", RayBuffers), ("prev", ProbeAtlas), ("out", ProbeAtlas),
                ("firstFrame", C.c_uint32), ("pad", C.c_int32)]


class UpdateBatch(C.Structure):  # include/luxgi.h LuxgiUpdateBatch
    _fields_ = [("count", C.c_int32), ("groups", C.c_int32), ("flags", C.c_int32), ("rayStorage", C.c_int32),
                ("atlasStorage", C.c_int32), ("maxRays", C.c_int32), ("maxWeightBlocks", C.c_int32),
                ("pad", C.c_int32)]


MAX_DENOISE_BATCH = 1024


class ReflectionDenoiseItem(C.Structure):  # include/luxgi.h LuxgiReflectionDenoiseItem
    _fields_ = [("reproject", ReflectionReprojectParams), ("filter", ReflectionAtrousParams), ("width", C.c_int32),
                ("height", C.c_int32), ("pingPong", C.c_int32), ("blockStart", C.c_int32), ("cur", GBuffer),
                ("prev", GBuffer), ("noisy", C.c_void_p), ("outputs", C.c_void_p * 2), ("moments", C.c_void_p * 2),
                ("atrous", C.c_void_p * 2), ("tileMap", C.c_void_p)]


class DenoiseBatchHeader(C.Structure):
    _fields_ = [("count", C.c_int32), ("totalBlocks", C.c_int32), ("pad", C.c_int32 * 2)]

_SIZES = {DDGIUniform: 96, GlobalSDFData: 96, SurfaceAtlasData: 32, AtlasObject: 128, AtlasTile: 96,
          TraceParams: 120, ReflectionReprojectParams: 168, ReflectionAtrousParams: 24,
          ShadowReprojectParams: 80, ShadowAtrousParams: 32, SampleProbeParams: 80, Light: 64,
          SurfaceGBuffer: 40, SurfaceLightParams: 32, SurfaceDraw: 8, BlueNoise: 16, SdfShadowParams: 144,
          SdfReflectionParams: 144, ObjectRasterizeData: 176, RasterizeParams: 48, MeshSDF: 72,
          LearnedAtrousParams: 32, ReflectionDenoiseItem: 336, DenoiseBatchHeader: 16,
          TraceVolume: 232, TraceScene: 288, TraceSceneBatch: 48, UpdateVolume: 216, UpdateBatch: 32}
for _t, _n in _SIZES.items():
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t), _n)


def mat4_colmajor(m):
    """numpy 4x4 (row-major math convention, m[row, col]) -> column-major float[16]."""
    m = np.asarray(m, dtype=np.float32)
    return (C.c_float * 16)(*m.T.reshape(-1).tolist())


def make_uniform(probe_counts, start, step, rays_per_probe, irr_oct=8, depth_oct=16, max_distance=None,
                 sharpness=50.0, hysteresis=0.98, normal_bias=0.1, gamma=5.0):
    """DDGIUniform with the reference defaults (DDGIRenderer.h:28-41, on_game_start DDGIRenderer.cpp:666-682).

    max_distance defaults to 1.5 * max(step) (DDGIRenderer.cpp:674: 1.5 * probeDistance)."""
    u = DDGIUniform()
    for i in range(3):
        u.startPosition[i] = float(start[i])
        u.step[i] = float(step[i])
        u.probeCounts[i] = int(probe_counts[i])
    u.startPosition[3] = 0.0
    u.step[3] = 0.0
    u.probeCounts[3] = 0
    u.maxDistance = float(1.5 * max(step) if max_distance is None else max_distance)
    u.sharpness = float(sharpness)
    u.hysteresis = float(hysteresis)
    u.normalBias = float(normal_bias)
    u.ddgiGamma = float(gamma)
    u.irradianceProbeSideLength = int(irr_oct)
    u.depthProbeSideLength = int(depth_oct)
    px, py, pz = (int(c) for c in probe_counts)
    u.irradianceTextureWidth = (irr_oct + 2) * px * py + 2
    u.irradianceTextureHeight = (irr_oct + 2) * pz + 2
    u.depthTextureWidth = (depth_oct + 2) * px * py + 2
    u.depthTextureHeight = (depth_oct + 2) * pz + 2
    u.raysPerProbe = int(rays_per_probe)
    return u


def make_sdf_data(center=(0.0, 0.0, 0.0), half_extent=1.0, resolution=64):
    """Single-cascade GlobalSDFData over [center - half, center + half]^3."""
    g = GlobalSDFData()
    g.cascadePosDistance[0][0] = center[0]
    g.cascadePosDistance[0][1] = center[1]
    g.cascadePosDistance[0][2] = center[2]
    g.cascadePosDistance[0][3] = half_extent
    g.cascadeVoxelSize[0] = 2.0 * half_extent / resolution
    g.cascadesCount = 1
    g.resolution = float(resolution)
    g.nearPlane = 0.1
    g.farPlane = 1000.0
    return g


def addr(x):
    """Device or host address of a torch tensor / numpy array / None."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return int(x)

"""Shared seeded inputs for the tests (small sizes the oracle finishes in seconds)."""
import functools

import numpy as np

import oracle
from luxgi_amd import abi, scene


@functools.lru_cache(maxsize=None)
def small_scene(res=64, atlas_res=512, tile=32, color_storage=abi.STORAGE_FP32):
    """Config-1 scene: 8 boxes + 3 spheres baked to a res^3 SDF over [-1, 1]^3."""
    prims = scene.scene_small(seed=1)
    sdf = scene.bake_sdf_numpy(prims, res)
    g = abi.make_sdf_data(resolution=res)
    mip = oracle.sdf_mip(sdf, res, max_distance=2.0, flood_passes=4)
    atlas = scene.build_atlas(prims, resolution=atlas_res, tile=tile, color_storage=color_storage, seed=3,
                              threshold=1.05 * 2.0 / res)
    return prims, sdf, mip, g, atlas


def small_uniform(rays=64, counts=(8, 8, 8), ni=6, nd=14):
    step = [2.0 / (c - 1) if c > 1 else 1.0 for c in counts]
    return abi.make_uniform(counts, (-1.0, -1.0, -1.0), step, rays, irr_oct=ni, depth_oct=nd, hysteresis=0.97)


def trace_params(rot_seed=5, max_steps=250, hit_eps=0.0, sky=(0.6, 0.7, 0.9)):
    p = abi.TraceParams()
    p.randomOrientation = abi.mat4_colmajor(scene.random_rotation(rot_seed))
    p.numFrames = 0
    p.numLights = 1
    p.intensity = 1.0
    p.maxSteps = max_steps
    p.hitEpsilon = hit_eps
    p.skyColor[:] = [sky[0], sky[1], sky[2], 1.0]
    return p


def oracle_trace(u, g, sdf, mip, atlas, params, **kw):
    res = int(g.resolution)
    return oracle.trace(u, g, sdf, mip, res, (res + 3) // 4, atlas.data, atlas.chunks, atlas.culled, atlas.objects,
                        atlas.tiles, atlas.depth, atlas.color_f32, params, **kw)


def np_f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def cascaded_sdf(prims, res, halves, center=(0.0, 0.0, 0.0)):
    """A global SDF clipmap of len(halves) cascades around `center` (GlobalDistanceField.cpp:656-666):
    cascade c covers center +- halves[c] at res^3, normalised by its own extent 2 halves[c], packed
    along x ([res][res][C res]); the coarse mip of each cascade with maxDistance 2 halves[c]
    (:823-846). Returns (sdf, mip, GlobalSDFData)."""
    C = len(halves)
    sdf = np.concatenate([scene.bake_sdf_numpy(prims, res, center=center, half=h) for h in halves], axis=2)
    mip = None
    for c, h in enumerate(halves):
        mip = oracle.sdf_mip(sdf, res, cascades=C, cascade=c, max_distance=2.0 * h, mip=mip)
    g = abi.make_sdf_data(center=center, half_extent=halves[0], resolution=res)
    for c, h in enumerate(halves):
        g.cascadePosDistance[c][0], g.cascadePosDistance[c][1], g.cascadePosDistance[c][2] = center
        g.cascadePosDistance[c][3] = h
        g.cascadeVoxelSize[c] = 2.0 * h / res
    g.cascadesCount = C
    return np.ascontiguousarray(sdf), mip, g

"""Host mirror of the global surface atlas passes over the C ABI (libluxgi.so).

Mirrors Engine/DDGI/GlobalSurfaceAtlas.cpp:
  on_render::culling (:607-641)            -> SurfaceCache.cull (SDFCulling.comp)
  on_render::deferredLighting (:813-967)   -> draw_list (the per-light tile lists, :913-940)
                                              + SurfaceCache.light (CopyEmissive + SDFDeferredLight)
  indirect_light::system (:995-1085)       -> SurfaceCache.light with LUXGI_SURFACE_INDIRECT
The per-frame choice of objects to relight (updatedObjects, :813-860) is the caller's: pass
the tiles of those objects to draw_list.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .ddgi import ptr, stream_handle, to_device


def _tile_axis(face):
    """tile->viewDirection of tile `face` of an object (GlobalSurfaceAtlas.cpp:249-263): the object-local
    axis e_{face/2} with sign - for even faces, + for odd."""
    z = np.zeros(3)
    z[face // 2] = 1.0 if (face & 1) else -1.0
    return z


def draw_list(objects, n_objects, lights, object_radius=None):
    """The tile quads of every object's tiles with their light masks, as
    GlobalSurfaceAtlas.cpp:913-940 builds one draw list per light: a directional light skips tiles
    whose view direction is within 1e-6 of perpendicular or facing away from its direction (a
    dot product of the object-local axis with the world direction, as the reference does); a
    point or spot light skips objects whose centre is farther than radius + light radius.
    objects: ctypes LuxgiAtlasObject array; lights: ctypes LuxgiLight array."""
    draws = []
    for i in range(n_objects):
        o = objects[i]
        c = np.array(o.objectBounds[:3], dtype=np.float64)
        rad = float(o.objectBounds[3]) if object_radius is None else float(object_radius[i])
        for face in range(6):
            tid = int(o.tileOffset[face])
            if tid == 0:
                continue
            mask = 0
            for l in range(len(lights)):
                L = lights[l]
                if L.type in (abi.LIGHT_POINT, abi.LIGHT_SPOT):
                    if np.linalg.norm(c - np.array(L.position[:3])) > rad + L.radius:
                        continue
                elif L.type == abi.LIGHT_DIRECTIONAL:
                    if float(np.dot(_tile_axis(face), np.array(L.direction[:3]))) < 1e-6:
                        continue
                mask |= 1 << l
            draws.append((tid, mask))
    arr = (abi.SurfaceDraw * max(len(draws), 1))()
    for k, (t, m) in enumerate(draws):
        arr[k].tile, arr[k].lightMask = t, m
    return arr, len(draws)


class SurfaceCache:
    """Surface-cache G-buffer and light cache of a scene.Atlas built with surface_gbuffer=True."""

    def __init__(self, atlas, out_storage=abi.STORAGE_FP16):
        self.atlas = atlas
        self.res = int(atlas.data.resolution)
        self.color = to_device(atlas.gb_color.view(np.int16))
        self.normal = to_device(atlas.gb_normal.view(np.int16))
        self.pbr = to_device(atlas.gb_pbr)
        self.emissive = to_device(atlas.gb_emissive.view(np.int16))
        self.gbuffer = abi.SurfaceGBuffer(ptr(self.color), ptr(self.normal), ptr(self.pbr), ptr(self.emissive),
                                          self.res, 0)
        self.tiles = to_device(atlas.tiles)
        self.n_tiles = len(atlas.tiles)
        self.out_storage = out_storage
        dt = torch.float16 if out_storage == abi.STORAGE_FP16 else torch.float32
        self.light_cache = torch.zeros((self.res, self.res, 4), dtype=dt, device="cuda")
        self.lights = None
        self.draws = None
        self.n_draws = 0

    def set_lights(self, lights, objects=None, n_objects=None):
        """Upload the lights and the draw list of every tile of the atlas's objects."""
        objects = self.atlas.objects if objects is None else objects
        n_objects = self.atlas.n_objects if n_objects is None else n_objects
        self.lights_host = lights
        self.lights = to_device(lights) if len(lights) else None
        self.n_lights = len(lights)
        draws, n = draw_list(objects, n_objects, lights)
        self.draws_host = draws
        self.draws = to_device(draws)
        self.n_draws = n

    def light(self, scene, passes=abi.SURFACE_EMISSIVE | abi.SURFACE_DIRECT, camera=(0.0, 0.0, 0.0),
              shadow_bias=0.02, volume=None, intensity=1.0, max_steps=0):
        """Relight the drawn tiles into self.light_cache. scene: ddgi.DeviceScene (global SDF);
        volume: ddgi.DDGIVolume whose current probe atlases feed the bounce."""
        p = abi.SurfaceLightParams()
        p.cameraPosition[:] = [camera[0], camera[1], camera[2], shadow_bias]
        p.indirectIntensity = intensity
        p.passes = passes
        p.numLights = self.n_lights if (passes & abi.SURFACE_DIRECT) else 0
        p.maxSteps = max_steps
        u = C.byref(volume.u) if volume is not None else None
        probes = C.byref(volume.atlas[volume.ping_pong]) if volume is not None else None
        check(lib().luxgi_surface_light(C.byref(p), ptr(self.lights), ptr(self.tiles), self.n_tiles, ptr(self.draws),
                                        self.n_draws, C.byref(self.gbuffer), C.byref(scene.sdf_data),
                                        C.byref(scene.volume), u, probes, self.out_storage, ptr(self.light_cache),
                                        stream_handle()), "luxgi_surface_light")
        return p


def cull_chunks(data, objects_dev, n_objects, capacity):
    """luxgi_atlas_cull_chunks into new device buffers; returns (chunks, culled) tensors."""
    d = abi.SurfaceAtlasData()
    C.memmove(C.byref(d), C.byref(data), C.sizeof(d))
    d.objectsCount = n_objects
    d.culledObjectsCapacity = capacity
    chunks = torch.empty(abi.ATLAS_CHUNKS_RESOLUTION ** 3, dtype=torch.int32, device="cuda")
    culled = torch.zeros(max(capacity, 1), dtype=torch.int32, device="cuda")
    ws = torch.empty(lib().luxgi_atlas_cull_workspace_size(), dtype=torch.uint8, device="cuda")
    check(lib().luxgi_atlas_cull_chunks(C.byref(d), ptr(objects_dev), ptr(chunks), ptr(culled), ptr(ws), ws.numel(),
                                        stream_handle()), "luxgi_atlas_cull_chunks")
    return chunks, culled

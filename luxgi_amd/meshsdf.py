"""Mesh SDFs and the global SDF build (host side).

* The reference's baked mesh-SDF asset format (Engine/DDGI/SDFBaker.cpp:158-200, read back at
  :216-232): a cereal binary archive of (uvec3 size, int32 mipCount, vector<uint8> level 0) followed by
  one vector<uint8> per further level; each vector is a uint64 byte count and R16F texels [z][y][x].
  Each texel holds (distance / maxDistance + 1) / 2 (SDFBaker.cpp:146-147); level l + 1 is the mean of
  2x2x2 texels of level l at half the size (at least 1), up to 3 levels (:158, :166-200).
* A stand-in baker for the synthetic primitives with the same encoding, grid and mip rule (the
  reference bakes triangle meshes with a BVH; its distances are replaced by the analytic ones).
* ObjectRasterizeData and the per-chunk object lists of one cascade (chunkCalculate / getChunkId,
  GlobalDistanceField.cpp:193-210,468-532), and GlobalSDFBuilder, which runs luxgi_sdf_rasterize.
"""
import ctypes as C
import struct

import numpy as np

from . import abi, scene

CHUNK = 32          # CONSTS_SDF_RASTERIZE_CHUNK_SIZE, GlobalDistanceField.cpp:86
CHUNK_MARGIN = 4    # CONSTS_SDF_RASTERIZE_CHUNK_MARGIN, :89


class MeshSDF:
    """One baked mesh SDF: levels (float16 arrays [z][y][x]) and its field metadata
    (MeshDistanceField.h:17-25: aabb, localToUVWMul/Add, maxDistance)."""

    def __init__(self, levels, aabb_min=None, aabb_max=None):
        self.levels = [np.ascontiguousarray(l, dtype=np.float16) for l in levels]
        if aabb_min is not None:
            self.set_bounds(aabb_min, aabb_max)

    @property
    def size(self):
        z, y, x = self.levels[0].shape
        return x, y, z

    def set_bounds(self, aabb_min, aabb_max):
        """Field metadata of SDFBaker.cpp:74-77 for the (padded) local bounds."""
        self.aabb_min = np.asarray(aabb_min, dtype=np.float64)
        self.aabb_max = np.asarray(aabb_max, dtype=np.float64)
        size = self.aabb_max - self.aabb_min
        self.local_to_uvw_mul = 1.0 / size
        self.local_to_uvw_add = -self.aabb_min / size
        self.max_distance = float(size.max())


def read_sdf(data):
    """Parse a baked .sdf asset (bytes or a path); returns MeshSDF (without field metadata, which the
    engine keeps in the scene file)."""
    if not isinstance(data, (bytes, bytearray)):
        data = open(data, "rb").read()
    sx, sy, sz, mips = struct.unpack_from("<3Ii", data, 0)
    off = 16
    levels = []
    dims = np.array([sx, sy, sz])
    for level in range(mips):
        n = struct.unpack_from("<Q", data, off)[0]
        off += 8
        if n != int(np.prod(dims)) * 2:
            raise ValueError(f"level {level}: {n} bytes for {tuple(dims)} texels")
        levels.append(np.frombuffer(data, dtype="<f2", count=n // 2, offset=off).reshape(dims[2], dims[1], dims[0]))
        off += n
        dims = np.maximum(dims // 2, 1)
    if off != len(data):
        raise ValueError(f"{len(data) - off} trailing bytes")
    return MeshSDF(levels)


def write_sdf(mesh):
    """Serialize a MeshSDF in the asset format (inverse of read_sdf)."""
    x, y, z = mesh.size
    out = [struct.pack("<3Ii", x, y, z, len(mesh.levels))]
    for l in mesh.levels:
        b = np.ascontiguousarray(l, dtype="<f2").tobytes()
        out.append(struct.pack("<Q", len(b)) + b)
    return b"".join(out)


def to_half(values):
    """float32 -> float16 as the baker stores texels (glm::packHalf, glm 0.9.9.8 vendored under
    Maple/lib/glm): round to nearest with ties away from zero, where numpy rounds ties to even."""
    f = np.asarray(values, dtype=np.float32)
    h = f.astype(np.float16)
    other = np.nextafter(h, np.where(f > h.astype(np.float32), np.float16(np.inf), np.float16(-np.inf)).astype(np.float16))
    tie = (h.astype(np.float64) + other.astype(np.float64)) * 0.5 == f.astype(np.float64)
    tie &= h != f
    bigger = np.abs(other.astype(np.float32)) > np.abs(h.astype(np.float32))
    return np.where(tie & bigger, other, h).astype(np.float16)


def mip_chain(level0, count=3):
    """The baker's mip rule (SDFBaker.cpp:166-200): each level is the mean of 2x2x2 texels of the
    previous one (in float, from the stored half values), stored as half; odd sizes drop their last
    layer; at most min(mip levels of the size, 3) levels."""
    z, y, x = level0.shape
    full = 1
    w, h, d = x, y, z
    while w > 1 or h > 1 or d > 1:
        w, h, d = max(w >> 1, 1) if w > 1 else w, max(h >> 1, 1) if h > 1 else h, max(d >> 1, 1) if d > 1 else d
        full += 1
    levels = [np.ascontiguousarray(level0, dtype=np.float16)]
    for _ in range(1, min(full, count)):
        prev = levels[-1].astype(np.float32)
        pz, py, px = prev.shape
        nz, ny, nx = max(pz // 2, 1), max(py // 2, 1), max(px // 2, 1)
        acc = np.zeros((nz, ny, nx), np.float32)
        for dz in range(2):
            for dy in range(2):
                for dx in range(2):
                    zi = np.minimum(np.arange(nz) * 2 + dz, pz - 1)
                    yi = np.minimum(np.arange(ny) * 2 + dy, py - 1)
                    xi = np.minimum(np.arange(nx) * 2 + dx, px - 1)
                    acc += prev[np.ix_(zi, yi, xi)]
        levels.append(to_half(acc * np.float32(1.0 / 8.0)))
    return levels


def padding_box(bmin, bmax):
    """paddingSDFBox (SDFBaker.h:39-45): 5% of the size, at least 0.1, on every side."""
    bmin, bmax = np.asarray(bmin, np.float64), np.asarray(bmax, np.float64)
    pad = np.maximum(0.05 * (bmax - bmin), 0.1)
    return bmin - pad, bmax + pad


def bake_primitive(prim, min_res=32, max_res=128, texels_per_meter=3.0):
    """Mesh SDF of one synthetic primitive in its local frame (centred at the origin), with the
    baker's grid over the padded box (texel i at min + i * size / (n - 1), SDFBaker.cpp:108-110),
    resolution rule on the unpadded mesh box (its extent / targetTexelPerMeter clamped to
    [32, 128] and truncated, :55-62), encoding and mips."""
    p = prim.copy()
    p["center"] = (0.0, 0.0, 0.0)
    ext = scene.prim_half_extents(np.array([p]))[0].astype(np.float64)
    bmin, bmax = padding_box(-ext, ext)
    size = bmax - bmin
    dims = np.clip((2.0 * ext) / texels_per_meter, min_res, max_res).astype(np.int64)
    axes = [np.arange(n) * (size[a] / (n - 1)) + bmin[a] for a, n in enumerate(dims)]
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    d = scene.prim_distance_np(np.array([p]), np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)).reshape(Z.shape)
    m = MeshSDF([], bmin, bmax)
    enc = (d / m.max_distance + 1.0) / 2.0
    m.levels = mip_chain(to_half(enc.astype(np.float32)))
    return m


class Cascade:
    """One cascade of the global SDF (GlobalDistanceField.cpp:656-666,756-761)."""

    def __init__(self, index, center, half_extent, resolution):
        self.index = index
        self.center = np.asarray(center, np.float64)
        self.half = float(half_extent)
        self.res = int(resolution)
        self.max_distance = 2.0 * self.half
        self.voxel = self.max_distance / self.res
        self.bmin = self.center - self.half
        self.bmax = self.center + self.half

    def params(self):
        p = abi.RasterizeParams()
        mul = (self.bmax - self.bmin) / self.res
        add = self.bmin + self.voxel * 0.5
        p.cascadeCoordToPosMul[:] = [float(v) for v in mul]
        p.maxDistance = self.max_distance
        p.cascadeCoordToPosAdd[:] = [float(v) for v in add]
        p.cascadeResolution = self.res
        p.cascadeIndex = self.index
        return p


def object_rasterize_data(mesh, world, cascade_index):
    """chunkCalculate's ObjectRasterizeData (GlobalDistanceField.cpp:482-504): `world` is the
    object's local-to-world 4x4 (row-major math convention)."""
    world = np.asarray(world, np.float64)
    center = (mesh.aabb_min + mesh.aabb_max) / 2.0
    t = np.eye(4)
    t[:3, 3] = -center
    w2v = np.linalg.inv(world) @ t
    v2w = np.linalg.inv(w2v)
    o = abi.ObjectRasterizeData()
    o.worldToVolume = abi.mat4_colmajor(w2v)
    o.volumeToWorld = abi.mat4_colmajor(v2w)
    o.volumeToUVWMul[:] = [float(v) for v in mesh.local_to_uvw_mul]
    o.volumeToUVWAdd[:] = [float(v) for v in mesh.local_to_uvw_add + center * mesh.local_to_uvw_mul]
    o.mipOffset = float(min(cascade_index, 2))
    o.decodeMul = mesh.max_distance
    o.decodeAdd = -mesh.max_distance
    o.volumeLocalBoundsExtent[:] = [float(v) for v in (mesh.aabb_max - mesh.aabb_min) / 2.0]
    return o


def world_bounds(mesh, world):
    """sdf.aabb.transform(world) (GlobalDistanceField.cpp:476): the box of the 8 transformed corners."""
    c = np.array([[x, y, z] for x in (mesh.aabb_min[0], mesh.aabb_max[0]) for y in (mesh.aabb_min[1], mesh.aabb_max[1])
                  for z in (mesh.aabb_min[2], mesh.aabb_max[2])])
    w = (np.asarray(world, np.float64) @ np.concatenate([c, np.ones((8, 1))], 1).T).T[:, :3]
    return w.min(0), w.max(0)


def chunk_range(cascade, bmin, bmax):
    """getChunkId (GlobalDistanceField.cpp:193-210): chunk coordinates (inclusive) of an object's world
    bounds plus a 4-voxel margin, relative to the cascade bounds shrunk by 0.1, truncated toward zero
    (ivec3 of a float). The reference's clamps there discard their results, so none is applied."""
    margin = cascade.voxel * CHUNK_MARGIN
    bias_min = cascade.bmin + 0.1
    lo = (np.asarray(bmin) - margin) - bias_min
    hi = (np.asarray(bmax) + margin) - bias_min
    size = cascade.voxel * CHUNK
    return np.trunc(lo / size).astype(np.int64), np.trunc(hi / size).astype(np.int64)


def chunk_lists(cascade, bounds):
    """Per-chunk object lists of one cascade, every chunk of the cascade listed (empty lists clear
    the chunk to 1). bounds: list of (world min, world max) per object. Chunks outside the cascade
    are dropped (the reference's writes there fall outside the image). Returns (origins int32
    [n][4], start uint32 [n + 1], objects uint32)."""
    nch = -(-cascade.res // CHUNK)
    lists = {}
    for i, (bmin, bmax) in enumerate(bounds):
        lo, hi = chunk_range(cascade, bmin, bmax)
        for z in range(max(lo[2], 0), min(hi[2], nch - 1) + 1):
            for y in range(max(lo[1], 0), min(hi[1], nch - 1) + 1):
                for x in range(max(lo[0], 0), min(hi[0], nch - 1) + 1):
                    lists.setdefault((x, y, z), []).append(i)
    origins, start, objs = [], [0], []
    for z in range(nch):
        for y in range(nch):
            for x in range(nch):
                origins.append([x * CHUNK, y * CHUNK, z * CHUNK, 0])
                objs.extend(lists.get((x, y, z), []))
                start.append(len(objs))
    return (np.asarray(origins, np.int32), np.asarray(start, np.uint32),
            np.asarray(objs if objs else [0], np.uint32)[:len(objs) if objs else 0])


class MeshPool:
    """The mesh SDF levels in one FP16 pool, with a LuxgiMeshSDF entry per object: object i uses
    mesh which[i] (default: mesh i), so instances share their mesh's texels."""

    def __init__(self, meshes, which=None):
        parts, off = [], 0
        per_mesh = []
        for m in meshes:
            e = abi.MeshSDF()
            e.mipCount = len(m.levels)
            for l, lv in enumerate(m.levels):
                e.offset[l] = off
                z, y, x = lv.shape
                e.size[l][0], e.size[l][1], e.size[l][2] = x, y, z
                parts.append(lv.reshape(-1).view(np.uint16))
                off += lv.size
            per_mesh.append(e)
        which = list(range(len(meshes))) if which is None else list(which)
        self.entries = (abi.MeshSDF * max(len(which), 1))()
        for i, k in enumerate(which):
            self.entries[i] = per_mesh[k]
        self.pool = np.concatenate(parts) if parts else np.zeros(1, np.uint16)
        self.count = len(which)


def primitive_objects(prims, cascade_index=0):
    """Mesh SDFs (one per primitive) placed at the primitives' centres: (meshes, ObjectRasterizeData
    array, world bounds)."""
    meshes, objs, bounds = [], (abi.ObjectRasterizeData * max(len(prims), 1))(), []
    for i, p in enumerate(prims):
        m = bake_primitive(p)
        world = np.eye(4)
        world[:3, 3] = p["center"]
        meshes.append(m)
        objs[i] = object_rasterize_data(m, world, cascade_index)
        bounds.append(world_bounds(m, world))
    return meshes, objs, bounds


class GlobalSDFBuilder:
    """Device copies of the objects' mesh SDFs and rasterize data; luxgi_sdf_rasterize per cascade."""

    def __init__(self, meshes, objects, n_objects, which=None):
        import torch
        from .ddgi import to_device
        self.pool = MeshPool(meshes, which)
        self.n = int(n_objects)
        self.pool_dev = to_device(self.pool.pool.view(np.int16))
        self.meshes_dev = to_device(self.pool.entries)
        self.objects_dev = to_device(objects)
        self._torch = torch

    def rasterize(self, cascade, lists, sdf, cascades=1):
        """lists = chunk_lists(cascade, ...); sdf = [R][R][cascades R] float32 or float16 tensor."""
        from ._lib import check, lib
        from .ddgi import ptr, stream_handle, to_device
        origins, start, objs = lists
        o = to_device(np.ascontiguousarray(origins, np.int32))
        s = to_device(np.ascontiguousarray(start, np.uint32).view(np.int32))
        l = to_device(np.ascontiguousarray(objs if len(objs) else np.zeros(1, np.uint32), np.uint32).view(np.int32))
        storage = abi.STORAGE_FP16 if sdf.dtype == self._torch.float16 else abi.STORAGE_FP32
        check(lib().luxgi_sdf_rasterize(C.byref(cascade.params()), ptr(self.objects_dev), ptr(self.meshes_dev),
                                        ptr(self.pool_dev), self.n, ptr(o), ptr(s), ptr(l), len(start) - 1, cascades,
                                        storage, ptr(sdf), stream_handle()), "luxgi_sdf_rasterize")
        return sdf


def primitive_instances(prims, cascade_index=0):
    """The primitives as instances of three baked unit meshes (box, sphere, capsule along x), placed
    by a world matrix with the primitive's scale and axis (the engine instances a few mesh SDFs many
    times). Returns (meshes, object rasterize data, world bounds, mesh index per object)."""
    unit = np.zeros(3, dtype=scene.PRIM_DTYPE)
    unit["type"] = (scene.PRIM_BOX, scene.PRIM_SPHERE, scene.PRIM_CAPSULE)
    unit["size"][0] = (0.5, 0.5, 0.5)
    unit["size"][1] = (0.5, 0.0, 0.0)
    unit["size"][2] = (0.25, 0.25, 0.0)  # radius 0.25, half length 0.25, axis x
    meshes = [bake_primitive(u) for u in unit]
    objs = (abi.ObjectRasterizeData * max(len(prims), 1))()
    bounds, which = [], []
    for i, p in enumerate(prims):
        t = int(p["type"])
        sz = p["size"].astype(np.float64)
        world = np.eye(4)
        if t == scene.PRIM_BOX:
            S = np.diag(sz / 0.5)
        elif t == scene.PRIM_SPHERE:
            S = np.eye(3) * (sz[0] / 0.5)
        else:  # capsule of radius r and half length h along axis a: scale the unit capsule
            r, h, a = sz[0], sz[1], int(sz[2])
            S = np.eye(3) * (r / 0.25)
            S[0, 0] = (r + h) / 0.5
            perm = np.roll(np.eye(3), a, axis=0)  # x -> axis a
            S = perm @ S
        world[:3, :3] = S
        world[:3, 3] = p["center"]
        k = min(t, 2)
        objs[i] = object_rasterize_data(meshes[k], world, cascade_index)
        bounds.append(world_bounds(meshes[k], world))
        which.append(k)
    return meshes, objs, bounds, which

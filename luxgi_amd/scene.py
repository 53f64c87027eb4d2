"""Synthetic scenes for the DDGI and SVGF passes (inputs only, not the hot path).

Builds, from a seed:
  * analytic primitives (boxes, spheres, capsules) inside [-1, 1]^3;
  * the surface-atlas buffers that Shaders/SDF/AtlasCommon.glsl:115-157 reads:
    one ObjectBuffer per primitive and six TileBuffers laid out as
    Engine/DDGI/GlobalSurfaceAtlas.cpp:231-279 lays them out (tile axis
    z = +-e[t/2], view matrix = lookAt with its translation cleared,
    objectBounds = 2 * |view extents|, extends = atlas rect in UV), the atlas
    colour and tile-depth images, the 40^3 chunk table and the culled-object
    runs [count, ids...] (slot 0 of the tile and culled buffers is reserved,
    because offset 0 means "none" in the lookup);
  * camera matrices (Vulkan clip space, depth in [0, 1]).
The SDF bake has a numpy form (float64, used by the tests on small grids) and a
GPU form (luxgi_scene_bake_sdf) for the full-size scene.
"""
import ctypes as C

import numpy as np

from . import abi

PRIM_BOX, PRIM_SPHERE, PRIM_CAPSULE = 0, 1, 2

PRIM_DTYPE = np.dtype([("center", "<f4", 3), ("type", "<i4"), ("size", "<f4", 3), ("roughness", "<f4"),
                       ("albedo", "<f4", 3), ("emissive", "<f4")])
assert PRIM_DTYPE.itemsize == 48


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
def scene_small(seed=1):
    """8 axis-aligned boxes + 3 spheres in [-1, 1]^3 with albedo U(0.1, 0.9):
    a room (floor, ceiling, 4 walls), two blocks, three spheres."""
    rng = np.random.default_rng(seed)
    prims = np.zeros(11, dtype=PRIM_DTYPE)
    boxes = [((0, -0.92, 0), (0.92, 0.04, 0.92)), ((0, 0.92, 0), (0.92, 0.04, 0.92)),
             ((-0.92, 0, 0), (0.04, 0.92, 0.92)), ((0.92, 0, 0), (0.04, 0.92, 0.92)),
             ((0, 0, -0.92), (0.92, 0.92, 0.04)), ((0, 0, 0.92), (0.92, 0.92, 0.04)),
             ((-0.35, -0.6, 0.2), (0.22, 0.28, 0.22)), ((0.4, -0.45, -0.3), (0.2, 0.43, 0.2))]
    spheres = [((0.35, -0.62, 0.45), 0.22), ((-0.4, 0.1, -0.35), 0.18), ((0.1, 0.35, 0.1), 0.15)]
    for i, (c, e) in enumerate(boxes):
        prims[i]["center"] = c
        prims[i]["type"] = PRIM_BOX
        prims[i]["size"] = e
    for j, (c, r) in enumerate(spheres):
        k = 8 + j
        prims[k]["center"] = c
        prims[k]["type"] = PRIM_SPHERE
        prims[k]["size"] = (r, 0, 0)
    prims["albedo"] = rng.uniform(0.1, 0.9, size=(11, 3))
    prims["roughness"] = rng.uniform(0.06, 0.44, size=11)
    prims["emissive"] = 0.0
    return prims


def scene_small_varied(seed):
    """scene_small(seed) (seeded albedo and roughness) with the two blocks and three spheres moved by
    up to 0.15 along each axis and scaled by 0.8-1.2, seeded: config-1-style scenes that differ in
    layout as well as in material."""
    prims = scene_small(seed)
    rng = np.random.default_rng(seed + 7919)
    for k in range(6, 11):
        prims[k]["center"] = prims[k]["center"] + rng.uniform(-0.15, 0.15, size=3)
        prims[k]["size"] = prims[k]["size"] * rng.uniform(0.8, 1.2)
    return prims


def scene_random(n=2000, seed=2, size_range=(0.01, 0.2)):
    """n primitives (box / sphere / capsule) at uniform positions in [-1, 1]^3 with
    log-uniform sizes in size_range (0.005-0.1 of the extent 2)."""
    rng = np.random.default_rng(seed)
    prims = np.zeros(n, dtype=PRIM_DTYPE)
    prims["center"] = rng.uniform(-1.0, 1.0, size=(n, 3))
    types = rng.integers(0, 3, size=n)
    prims["type"] = types
    lo, hi = np.log(size_range[0]), np.log(size_range[1])
    s = np.exp(rng.uniform(lo, hi, size=(n, 3))) * 0.5
    size = np.zeros((n, 3), dtype=np.float32)
    box = types == PRIM_BOX
    size[box] = s[box]
    sph = types == PRIM_SPHERE
    size[sph, 0] = s[sph, 0]
    cap = types == PRIM_CAPSULE
    size[cap, 0] = s[cap, 0] * 0.5
    size[cap, 1] = s[cap, 1]
    size[cap, 2] = rng.integers(0, 3, size=int(cap.sum()))
    prims["size"] = size
    prims["albedo"] = rng.uniform(0.1, 0.9, size=(n, 3))
    prims["roughness"] = rng.uniform(0.06, 0.44, size=n)
    return prims


def prim_half_extents(prims):
    e = np.zeros((len(prims), 3), dtype=np.float64)
    for i, p in enumerate(prims):
        t = int(p["type"])
        sz = p["size"].astype(np.float64)
        if t == PRIM_BOX:
            e[i] = sz
        elif t == PRIM_SPHERE:
            e[i] = sz[0]
        else:
            e[i] = sz[0]
            e[i, int(sz[2])] = sz[1] + sz[0]
    return e


def prim_distance_np(prims, pts):
    """Signed distance of every point (N, 3) to the union of primitives, float64."""
    pts = np.asarray(pts, dtype=np.float64)
    d = np.full(pts.shape[0], np.inf)
    for p in prims:
        q = pts - p["center"].astype(np.float64)
        t = int(p["type"])
        sz = p["size"].astype(np.float64)
        if t == PRIM_BOX:
            qq = np.abs(q) - sz
            di = np.linalg.norm(np.maximum(qq, 0.0), axis=1) + np.minimum(qq.max(axis=1), 0.0)
        elif t == PRIM_SPHERE:
            di = np.linalg.norm(q, axis=1) - sz[0]
        else:
            ax = int(sz[2])
            c = np.zeros_like(q)  # closest point of the segment on the axis
            c[:, ax] = np.clip(q[:, ax], -sz[1], sz[1])
            di = np.linalg.norm(q - c, axis=1) - sz[0]
        d = np.minimum(d, di)
    return d


def bake_sdf_numpy(prims, res, center=(0.0, 0.0, 0.0), half=1.0):
    """[res][res][res] (z, y, x) normalised SDF, clamp(d / (2 half), -1, 1)."""
    voxel = 2.0 * half / res
    ax = (np.arange(res) + 0.5) * voxel - half
    z, y, x = np.meshgrid(ax + center[2], ax + center[1], ax + center[0], indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    d = prim_distance_np(prims, pts)
    return np.clip(d / (2.0 * half), -1.0, 1.0).reshape(res, res, res).astype(np.float32)


# ---------------------------------------------------------------------------
# surface atlas
# ---------------------------------------------------------------------------
def _tile_frame(t):
    """Rows (s, u, -f) of the tile view rotation (GlobalSurfaceAtlas.cpp:249-277, glm::lookAt)."""
    z = np.zeros(3)
    z[t // 2] = 1.0 if (t & 1) else -1.0
    up = np.array([1.0, 0.0, 0.0]) if t in (2, 3) else np.array([0.0, 1.0, 0.0])
    f = z
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    return np.stack([s, u, -f])


def _visible_depth(prim, e, R, X, Y):
    """Tile depth (tp.z / objectBounds.z) of the surface seen through tile coords (X, Y); 1 = empty."""
    t = int(prim["type"])
    sz = prim["size"].astype(np.float64)
    ob_z = 2.0 * abs(R[2] @ e)
    out = np.ones_like(X)
    if t == PRIM_BOX:
        ex = abs(R[0] @ e)
        ey = abs(R[1] @ e)
        m = (np.abs(X) <= ex) & (np.abs(Y) <= ey)
        out[m] = 0.5
        return out
    if t == PRIM_SPHERE:
        r = sz[0]
        h2 = r * r - X * X - Y * Y
        m = h2 >= 0
        out[m] = np.sqrt(h2[m]) / ob_z
        return out
    r, hl, ax = sz[0], sz[1], int(sz[2])
    axis = np.zeros(3)
    axis[ax] = 1.0
    a_s, a_u, a_f = R[0] @ axis, R[1] @ axis, R[2] @ axis
    if abs(a_f) > 0.5:  # axis along the view direction: spherical cap on the near end
        h2 = r * r - X * X - Y * Y
        m = h2 >= 0
        out[m] = (hl + np.sqrt(h2[m])) / ob_z
        return out
    if abs(a_s) > 0.5:
        dx = X - np.clip(X, -hl, hl)
        dy = Y
    else:
        dx = X
        dy = Y - np.clip(Y, -hl, hl)
    h2 = r * r - dx * dx - dy * dy
    m = h2 >= 0
    out[m] = np.sqrt(h2[m]) / ob_z
    return out


class Atlas:
    """Host-side surface atlas (numpy / ctypes arrays)."""


def cull_chunks_np(bounds, chunk_size, capacity=None):
    """Chunk culling of Shaders/SDF/SDFCulling.comp:36-101 with the runs in chunk-address order (the
    layout luxgi_atlas_cull_chunks writes): every fp32 operation of boxIntersectsSphere
    (AtlasCommon.glsl:34-38) in the written order, so the decisions match the GPU pass bit for bit.
    bounds: (n, 4) sphere centre + radius. Returns (chunks[40^3], culled) with culled[0] the
    reference's final counter (1 + sum of count + 1 over the non-empty chunks); a run that would
    end past `capacity` (default: all fit) leaves its chunk at offset 0."""
    R40 = abi.ATLAS_CHUNKS_RESOLUTION
    b = np.asarray(bounds, dtype=np.float32).reshape(-1, 4)
    cs = np.float32(chunk_size)
    mn = (np.arange(R40, dtype=np.float32) - np.float32(R40 * 0.5)) * cs
    mx = mn + cs
    lists = {}
    for i in range(b.shape[0]):
        c, r = b[i, :3], b[i, 3]
        ax = []
        for a in range(3):  # candidate chunks per axis (a superset), then the exact fp32 test
            ax.append(np.nonzero((mx >= c[a] - r - cs) & (mn <= c[a] + r + cs))[0])
        if any(len(v) == 0 for v in ax):
            continue
        Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
        d = [c[0] - np.minimum(np.maximum(c[0], mn[X]), mx[X]), c[1] - np.minimum(np.maximum(c[1], mn[Y]), mx[Y]),
             c[2] - np.minimum(np.maximum(c[2], mn[Z]), mx[Z])]
        d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        hit = np.sqrt(d2) <= r
        for addr in ((Z * R40 + Y) * R40 + X)[hit].tolist():
            lists.setdefault(addr, []).append(i)
    chunks = np.zeros(R40 ** 3, dtype=np.uint32)
    culled = [0]
    cursor = 1
    for addr in sorted(lists):
        ids = lists[addr]
        start, size = cursor, len(ids) + 1
        cursor += size
        if capacity is not None and start + size > capacity:
            continue
        chunks[addr] = start
        culled.append(len(ids))
        culled.extend(ids)
    culled[0] = cursor
    return chunks, np.asarray(culled, dtype=np.uint32)


def _surface_normals(prim, pts):
    """Outward unit normal (the distance gradient) of one primitive at points on its surface."""
    q = pts - prim["center"].astype(np.float64)
    t = int(prim["type"])
    sz = prim["size"].astype(np.float64)
    if t == PRIM_BOX:  # the face whose slab holds the point: the largest |q_a| - s_a
        rows = np.arange(len(q))
        a = np.argmax(np.abs(q) - sz[None, :], axis=1)
        g = np.zeros_like(q)
        g[rows, a] = np.where(q[rows, a] >= 0.0, 1.0, -1.0)
    elif t == PRIM_SPHERE:
        g = q
    else:
        c = np.zeros_like(q)  # closest point of the capsule's segment
        ax = int(sz[2])
        c[:, ax] = np.clip(q[:, ax], -sz[1], sz[1])
        g = q - c
    return g / np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-30)


def _oct_encode_np(n):
    """octEncode (DDGICommon.glsl:65-72): unit vectors (m, 3) -> (m, 2) in [-1, 1]."""
    l1 = np.abs(n).sum(axis=1, keepdims=True)
    r = n[:, :2] / l1
    neg = n[:, 2] < 0
    sx = np.where(r[neg, 0] >= 0, 1.0, -1.0)
    sy = np.where(r[neg, 1] >= 0, 1.0, -1.0)
    r[neg] = np.stack([(1.0 - np.abs(r[neg, 1])) * sx, (1.0 - np.abs(r[neg, 0])) * sy], axis=1)
    return r


def build_atlas(prims, resolution=512, tile=32, color_storage=abi.STORAGE_FP32, seed=3, threshold=0.05,
                gi_distance=2.0, emissive_fraction=0.02, surface_gbuffer=False):
    """Synthetic global surface atlas of the primitives: 6 tiles per object (the view frames of
    GlobalSurfaceAtlas.cpp:249-277), tile depth, light-cache colour (albedo with emissive texels),
    the object/tile buffers and the chunk lists. With surface_gbuffer, also the surface-cache
    G-buffer of the lit tiles (GlobalSurfaceAtlas.cpp:407-410): world position, octahedral normal,
    albedo, PBR (metallic, roughness) and emissive per texel, zero where a tile sees nothing."""
    n = len(prims)
    ext = prim_half_extents(prims)
    slot = tile + 1
    per_row = resolution // slot
    n_tiles = n * 6 + 1
    if per_row * per_row < n_tiles:
        raise ValueError(f"atlas {resolution}^2 too small for {n_tiles} tiles of {tile}")
    rng = np.random.default_rng(seed)
    color = np.zeros((resolution, resolution, 4), dtype=np.float32)
    depth = np.ones((resolution, resolution), dtype=np.float32)
    objects = (abi.AtlasObject * n)()
    tiles = (abi.AtlasTile * n_tiles)()
    k = (np.arange(tile) + 0.5) / (tile - 1)
    TU, TV = np.meshgrid(k, k, indexing="xy")  # [ky][kx]
    emissive_value = 1.0 if color_storage == abi.STORAGE_UNORM8 else 4.0
    if surface_gbuffer:
        gb = {k: np.zeros((resolution, resolution, 4), dtype=np.float32) for k in ("color", "normal", "pbr", "emissive")}
        metallic = np.random.default_rng(seed + 1).uniform(0.0, 0.5, size=n)
    for i in range(n):
        p = prims[i]
        c = p["center"].astype(np.float64)
        e = ext[i]
        o = objects[i]
        o.objectBounds[:] = [c[0], c[1], c[2], float(np.linalg.norm(e))]
        o.transform[:] = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, c[0], c[1], c[2], 1]
        o.extends[:] = [e[0], e[1], e[2], 1.0]
        for t in range(6):
            tid = 1 + i * 6 + t
            o.tileOffset[t] = tid
            R = _tile_frame(t)
            ob = 2.0 * np.abs(R @ e)
            tx, ty = (tid % per_row) * slot + 1, (tid // per_row) * slot + 1
            T = tiles[tid]
            T.objectBounds[:] = [ob[0], ob[1], ob[2], 0.0]
            M = np.eye(4)
            M[:3, :3] = R
            T.transform[:] = M.T.reshape(-1).tolist()  # column-major
            T.extends[:] = [tx / resolution, ty / resolution, (tile - 1) / resolution, (tile - 1) / resolution]
            X = (TU - 0.5) * ob[0]
            Y = (TV - 0.5) * ob[1]
            dep = _visible_depth(p, e, R, X, Y)
            depth[ty:ty + tile, tx:tx + tile] = dep
            alb = p["albedo"].astype(np.float32)
            tex = alb[None, None, :] * rng.uniform(0.85, 1.15, size=(tile, tile, 1)).astype(np.float32)
            em = rng.random((tile, tile)) < emissive_fraction
            tex[em] = emissive_value
            if color_storage == abi.STORAGE_UNORM8:
                tex = np.clip(tex, 0.0, 1.0)
            color[ty:ty + tile, tx:tx + tile, :3] = tex
            color[ty:ty + tile, tx:tx + tile, 3] = 1.0
            if surface_gbuffer:
                seen = dep < 1.0
                if seen.any():
                    tp = np.stack([X[seen], Y[seen], dep[seen] * ob[2]], axis=1)
                    world = c[None, :] + tp @ R  # local = R^T tp; objects are translated, not rotated
                    nrm = _surface_normals(p, world)
                    oc = _oct_encode_np(nrm)
                    sl = (slice(ty, ty + tile), slice(tx, tx + tile))
                    blk = {k: v[sl] for k, v in gb.items()}
                    albedo = alb[None, None, :] * np.ones((tile, tile, 1), np.float32)
                    blk["color"][seen] = np.concatenate([albedo[seen], world[:, :1]], axis=1)
                    blk["normal"][seen] = np.concatenate([oc, world[:, 1:3]], axis=1)
                    blk["pbr"][seen] = [metallic[i], float(p["roughness"]), 0.0, 0.0]
                    emv = np.zeros((tile, tile, 4), np.float32)
                    emv[em] = [emissive_value, emissive_value, emissive_value, 0.0]
                    blk["emissive"][seen] = emv[seen]
    # chunk table + culled runs (Shaders/SDF/SDFCulling.comp, here on the host)
    cs = gi_distance / abi.ATLAS_CHUNKS_RESOLUTION
    bounds = np.array([list(objects[i].objectBounds) for i in range(n)], dtype=np.float32).reshape(-1, 4)
    chunks, culled = cull_chunks_np(bounds, cs)
    data = abi.SurfaceAtlasData()
    data.cameraPos[:] = [0.0, 0.0, 0.0]
    data.chunkSize = cs
    data.culledObjectsCapacity = len(culled)
    data.resolution = resolution
    data.objectsCount = n
    a = Atlas()
    a.data, a.objects, a.tiles, a.chunks, a.culled = data, objects, tiles, chunks, culled
    a.depth = depth
    a.color_f32 = color
    a.color_storage = color_storage
    if color_storage == abi.STORAGE_UNORM8:
        a.color = np.round(np.clip(color, 0, 1) * 255.0).astype(np.uint8)
        a.color_f32 = a.color.astype(np.float32) / 255.0
    elif color_storage == abi.STORAGE_FP16:
        a.color = color.astype(np.float16)
        a.color_f32 = a.color.astype(np.float32)
    else:
        a.color = color
    a.n_objects = n
    if surface_gbuffer:
        # stored as the reference's formats: RGBA16F colour/normal/emissive, RGBA8 PBR
        a.gb_color = gb["color"].astype(np.float16)
        a.gb_normal = gb["normal"].astype(np.float16)
        a.gb_emissive = gb["emissive"].astype(np.float16)
        a.gb_pbr = np.round(np.clip(gb["pbr"], 0.0, 1.0) * 255.0).astype(np.uint8)
    return a


# ---------------------------------------------------------------------------
# cameras and orientation
# ---------------------------------------------------------------------------
def perspective_zo(fovy, aspect, near, far):
    """Right-handed perspective with depth in [0, 1] (glm::perspectiveRH_ZO)."""
    f = 1.0 / np.tan(fovy / 2.0)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = far / (near - far)
    m[2, 3] = near * far / (near - far)
    m[3, 2] = -1.0
    return m


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    eye, target, up = (np.asarray(v, dtype=np.float64) for v in (eye, target, up))
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[0, 3], m[1, 3], m[2, 3] = -s @ eye, -u @ eye, f @ eye
    return m


def camera_pair(width, height, seed=4, eye=(0.0, 0.2, 2.6), target=(0.0, -0.1, 0.0), motion=0.01):
    """(viewProj, prevViewProj, cameraPos) with a small seeded camera motion."""
    rng = np.random.default_rng(seed)
    P = perspective_zo(np.radians(55.0), width / height, 0.05, 20.0)
    eye = np.asarray(eye, dtype=np.float64)
    prev_eye = eye + rng.uniform(-motion, motion, size=3)
    V = look_at(eye, target)
    Vp = look_at(prev_eye, target)
    return P @ V, P @ Vp, eye, eye - prev_eye


def random_rotation(seed):
    """Axis-angle rotation as trace_rays::system draws it (DDGIRenderer.cpp:267-272), seeded."""
    rng = np.random.default_rng(seed)
    axis = rng.uniform(-1.0, 1.0, size=3)
    axis /= np.linalg.norm(axis)
    ang = rng.uniform(0.0, 1.0) * np.pi * 2.0
    x, y, z = axis
    c, s = np.cos(ang), np.sin(ang)
    R = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    M = np.eye(4)
    M[:3, :3] = R
    return M

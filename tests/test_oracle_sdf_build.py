"""Pins for the global SDF build from mesh SDFs (f3): the reference's baked asset format and mip rule,
distanceToModelSDF (SDFCommon.glsl:38-62) and the chunk rasterization (SDFRasterizeModel.glsl).

Known answers held by the reference: two of its baked assets (tests/golden/mesh_sdf_*.sdf) parse,
re-serialise byte for byte, and their mip levels 1 and 2 are reproduced bit for bit from level 0
by the baker's rule (2x2x2 mean, glm::packHalf rounding: ties away from zero). Closed forms: a
point inside a constant-texel volume reads (2c - 1) maxDistance; outside, sqrt(max(sd, 0)^2 + d^2)
with d the distance to the volume box; rotation invariance of that box distance. Invariants: a
rasterized sphere agrees with the analytic SDF near its surface; outside all objects the rasterized
union equals the min of the single-object rasterizations (every listed object is used: no loss past
28 per chunk), inside it keeps the sign;
empty chunks clear to 1; chunk ranges of getChunkId."""
import os

import numpy as np

import oracle
from luxgi_amd import abi, meshsdf, scene

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ASSETS = ("mesh_sdf_box_32x32x35.sdf", "mesh_sdf_box_51x32x32.sdf")


def test_assets_parse_roundtrip_and_mip_rule():
    for name, size in zip(ASSETS, ((32, 32, 35), (51, 32, 32))):
        raw = open(os.path.join(GOLDEN, name), "rb").read()
        m = meshsdf.read_sdf(raw)
        assert m.size == size and len(m.levels) == 3
        assert [l.shape for l in m.levels][1] == (size[2] // 2, size[1] // 2, size[0] // 2)
        assert meshsdf.write_sdf(m) == raw
        chain = meshsdf.mip_chain(m.levels[0])
        for a, b in zip(chain, m.levels):
            assert np.array_equal(a.view(np.uint16), b.view(np.uint16))
        # encoded distances (d / maxDistance + 1) / 2 lie in [0, 1]
        assert 0.0 <= m.levels[0].min() and m.levels[0].max() <= 1.0


def test_to_half_rounds_ties_away_from_zero():
    # 1 + 2^-11 is halfway between halves 1 and 1 + 2^-10: ties to even gives 1, the baker 1 + 2^-10
    t = np.float32(1.0 + 2.0 ** -11)
    assert meshsdf.to_half(np.array([t]))[0] == np.float16(1.0 + 2.0 ** -10)
    assert np.array([t]).astype(np.float16)[0] == np.float16(1.0)
    x = np.random.default_rng(0).uniform(0, 1, 20000).astype(np.float32)
    h = meshsdf.to_half(x).astype(np.float32)
    assert np.all(np.abs(h - x) <= np.abs(x.astype(np.float16).astype(np.float32) - x) + 1e-12)


def _one_object(levels, bmin, bmax, world=np.eye(4), mip=0):
    m = meshsdf.MeshSDF(levels, bmin, bmax)
    objs = (abi.ObjectRasterizeData * 1)()
    objs[0] = meshsdf.object_rasterize_data(m, world, mip)
    pool = meshsdf.MeshPool([m])
    return m, objs, pool


def _rasterize(objs, pool, n, cascade, lists=None, half=False, precision="f64"):
    if lists is None:
        bounds = [((-10, -10, -10), (10, 10, 10))] * n
        lists = meshsdf.chunk_lists(cascade, bounds)
    o, s, l = lists
    return oracle.sdf_rasterize(cascade.params(), objs, pool.entries, pool.pool, n, o, s, l, half=half,
                                precision=precision)


def test_distance_to_model_closed_forms():
    """A constant volume c over [-0.3, 0.3]^3: inside, (2c - 1) maxD; outside, sqrt(max(sd, 0)^2 + d^2)
    with d the distance to the box; the same under a rotation of the object."""
    res = 32
    cas = meshsdf.Cascade(0, (0, 0, 0), 1.0, res)
    c = 0.55
    lv = [np.full((8, 8, 8), c, np.float16)]
    m, objs, pool = _one_object(lv, (-0.3, -0.3, -0.3), (0.3, 0.3, 0.3))
    vol = _rasterize(objs, pool, 1, cas)
    sd = (2 * float(np.float16(c)) - 1) * m.max_distance
    ax = (np.arange(res) + 0.5) * (2.0 / res) - 1.0
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    q = np.stack([X, Y, Z], -1)
    dbox = np.linalg.norm(np.maximum(np.abs(q) - 0.3, 0.0), axis=-1)
    want = np.where(dbox > 0, np.sqrt(max(sd, 0) ** 2 + dbox ** 2), sd)
    # the 0.01 rule: within 0.01 of the box the local-space distance is used (here identical, no scaling)
    np.testing.assert_allclose(vol, np.clip(want / 2.0, -1, 1), atol=2e-6)
    # rotated by 90 degrees about z: the cube maps onto itself
    rot = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1.0]])
    m2, objs2, pool2 = _one_object(lv, (-0.3, -0.3, -0.3), (0.3, 0.3, 0.3), world=rot)
    np.testing.assert_allclose(_rasterize(objs2, pool2, 1, cas), vol, atol=2e-6)


def test_rasterized_sphere_matches_analytic_near_surface():
    res = 64
    cas = meshsdf.Cascade(0, (0, 0, 0), 1.0, res)
    prims = np.zeros(1, dtype=scene.PRIM_DTYPE)
    prims["type"] = scene.PRIM_SPHERE
    prims["center"] = (0.1, -0.05, 0.2)
    prims["size"] = (0.35, 0, 0)
    meshes, objs, bounds = meshsdf.primitive_objects(prims)
    pool = meshsdf.MeshPool(meshes)
    vol = _rasterize(objs, pool, 1, cas, lists=meshsdf.chunk_lists(cas, bounds))
    ana = scene.bake_sdf_numpy(prims, res)
    near = np.abs(ana) < 0.05
    # a mesh texel is maxD / 32 wide; the baker's grid is offset from the sampler's texel centres
    # by half a texel, and the stored distance is normalised by the cascade extent 2
    texel = meshes[0].max_distance / 32
    assert np.abs(vol - ana)[near].max() < texel / 2.0
    assert np.all(vol[np.abs(ana) > 0.2] * ana[np.abs(ana) > 0.2] > 0)  # sign away from the surface


def test_union_is_min_of_single_objects_and_empty_chunks_clear():
    """60 overlapping objects in the same chunks: FP32 storage gives the min of the single-object
    rasterizations (to a few ulp of the batch restarts), every object counted; FP16 storage rounds
    after every batch of 28, so it stays within one half ulp of the rounded min per batch boundary."""
    res = 64  # 2 x 2 x 2 chunks; the objects lie in chunk (0, 0, 0) only
    cas = meshsdf.Cascade(0, (0, 0, 0), 1.0, res)
    rng = np.random.default_rng(3)
    prims = np.zeros(60, dtype=scene.PRIM_DTYPE)
    prims["type"] = scene.PRIM_SPHERE
    prims["center"] = -0.5 + rng.uniform(-0.1, 0.1, (60, 3))
    prims["size"][:, 0] = rng.uniform(0.05, 0.12, 60)
    meshes, objs, bounds = meshsdf.primitive_objects(prims)
    pool = meshsdf.MeshPool(meshes)
    lists = meshsdf.chunk_lists(cas, bounds)
    full = _rasterize(objs, pool, 60, cas, lists=lists)
    singles = []
    for i in range(60):
        o, s, l = lists
        sel = np.concatenate([l[s[c]:s[c + 1]][l[s[c]:s[c + 1]] == i] for c in range(len(s) - 1)])
        cnt = np.array([np.sum(l[s[c]:s[c + 1]] == i) for c in range(len(s) - 1)])
        singles.append(_rasterize(objs, pool, 60, cas, lists=(o, np.concatenate([[0], np.cumsum(cnt)]).astype(np.uint32),
                                                              sel.astype(np.uint32))))
    want = np.min(np.stack(singles), axis=0)
    # outside every object the union is the min; inside, distanceToModelSDF's early-out
    # (minDistance <= distance to the volume, SDFCommon.glsl:52) skips every later object whose
    # volume holds the voxel once the running distance is negative, so only the sign is fixed there
    out = want >= 0
    assert out.mean() > 0.5 and (~out).sum() > 100
    np.testing.assert_allclose(full[out], want[out], atol=1e-6)
    assert np.all(full[~out] <= 0)
    counts = np.diff(lists[1])
    assert counts.max() > 28  # chunks past one batch
    half = _rasterize(objs, pool, 60, cas, lists=lists, half=True)
    assert np.abs(half - want)[out].max() < 3 * 2.0 ** -11
    assert np.all(half[~out] <= 0)
    # a chunk whose list is empty is cleared to 1
    o, s, l = lists
    empty = np.nonzero(counts == 0)[0]
    assert len(empty) and np.all(half[o[empty[0], 2]:o[empty[0], 2] + 32, o[empty[0], 1]:o[empty[0], 1] + 32,
                                      o[empty[0], 0]:o[empty[0], 0] + 32] == 1.0)


def test_chunk_range():
    cas = meshsdf.Cascade(0, (0, 0, 0), 1.0, 256)  # voxel 2/256, chunk 0.25, bounds shrunk by 0.1
    lo, hi = meshsdf.chunk_range(cas, (-0.6, 0.2, -0.05), (-0.55, 0.25, 0.05))
    # (x - 4 voxels + 0.9) / 0.25 truncated: x: (0.30 - 0.03125) / 0.25 = 1.075 -> 1, (0.35 + 0.03125) / 0.25 = 1.525 -> 1
    assert list(lo) == [1, 4, 3] and list(hi) == [1, 4, 3]
    lo, hi = meshsdf.chunk_range(cas, (-1.0, -1.0, -1.0), (-0.95, -0.95, -0.95))
    assert list(lo) == [0, 0, 0]  # -0.23 truncates toward zero


def test_baked_resolution_follows_the_unpadded_mesh_box():
    """SDFBaker.cpp:55-62 takes the resolution from the mesh's own bounding box (extent / texels
    per metre, clamped to [32, 128], truncated); the 5% / 0.1 m padding only spaces the texels. A
    120 m long box gives 40 texels along x (the padded 132 m would give 44), and thin axes clamp
    to 32."""
    p = np.zeros(1, dtype=scene.PRIM_DTYPE)[0]
    p["type"] = scene.PRIM_BOX
    p["size"] = (60.0, 0.5, 0.5)
    m = meshsdf.bake_primitive(p)
    assert m.levels[0].shape == (32, 32, 40)
    bmin, bmax = meshsdf.padding_box((-60.0, -0.5, -0.5), (60.0, 0.5, 0.5))
    assert np.allclose(m.aabb_min, bmin) and np.allclose(m.aabb_max, bmax)

"""Closed-form checks of the synthetic-scene builders (CPU)."""
import numpy as np

from luxgi_amd import abi, scene


def _one(t, center, size):
    p = np.zeros(1, dtype=scene.PRIM_DTYPE)
    p["type"] = t
    p["center"] = center
    p["size"] = size
    return p


def test_primitive_distances_closed_form():
    box = _one(scene.PRIM_BOX, (0.1, 0.2, 0.3), (0.2, 0.1, 0.05))
    pts = np.array([[0.1, 0.2, 0.3], [0.6, 0.2, 0.3], [0.4, 0.4, 0.3], [0.1, 0.2, 0.34]])
    want = [-0.05, 0.3, np.hypot(0.1, 0.1), -0.01]
    np.testing.assert_allclose(scene.prim_distance_np(box, pts), want, atol=1e-7)
    sph = _one(scene.PRIM_SPHERE, (0, 0, 0), (0.25, 0, 0))
    np.testing.assert_allclose(scene.prim_distance_np(sph, [[1, 0, 0], [0, 0.1, 0]]), [0.75, -0.15], atol=1e-7)
    cap = _one(scene.PRIM_CAPSULE, (0, 0, 0), (0.1, 0.3, 2))  # radius 0.1, half length 0.3 along z
    pts = np.array([[0, 0, 1.0], [0.5, 0, 0.2], [0.9, 0.9, 0.9], [0, 0, 0]])
    want = [0.6, 0.4, np.sqrt(0.81 + 0.81 + 0.36) - 0.1, -0.1]
    np.testing.assert_allclose(scene.prim_distance_np(cap, pts), want, atol=1e-7)


def test_random_scene_is_mostly_empty():
    prims = scene.scene_random(n=2000, seed=2)
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1, 1, (4000, 3))
    inside = (scene.prim_distance_np(prims, pts) < 0).mean()
    assert 0.01 < inside < 0.3, inside


def test_atlas_tables():
    prims = scene.scene_small(seed=1)
    a = scene.build_atlas(prims, resolution=512, tile=32, threshold=0.05)
    assert a.data.objectsCount == 11
    # chunk runs are [count, ids...]; slot 0 holds the culling counter (1 + all run sizes)
    assert a.culled[0] == len(a.culled)
    starts = a.chunks[a.chunks > 0]
    for s in np.unique(starts)[:50]:
        n = a.culled[s]
        ids = a.culled[s + 1:s + 1 + n]
        assert 1 <= n <= 11 and np.all(ids < 11) and np.all(np.diff(ids.astype(int)) > 0)
    # every tile offset is non-zero and box tiles see depth 0.5 at their centre texel
    for i in range(8):
        for t in range(6):
            tid = a.objects[i].tileOffset[t]
            assert tid == 1 + 6 * i + t
            ex = a.tiles[tid].extends
            x = int(round(ex[0] * 512)) + 16
            y = int(round(ex[1] * 512)) + 16
            assert a.depth[y, x] == 0.5


def test_surface_gbuffer_positions_and_normals():
    """The synthetic surface G-buffer holds points on the primitive's surface (distance ~ 0) and the
    distance gradient there as the octahedral normal."""
    from luxgi_amd.scene import prim_distance_np
    prims = scene.scene_small(seed=1)
    a = scene.build_atlas(prims, resolution=512, tile=32, threshold=0.05, surface_gbuffer=True)
    col = a.gb_color.astype(np.float64)
    nrm = a.gb_normal.astype(np.float64)
    seen = a.gb_pbr[..., 1] > 0
    assert seen.sum() > 20000
    world = np.stack([col[..., 3], nrm[..., 2], nrm[..., 3]], axis=-1)[seen]
    d = prim_distance_np(prims, world)
    # on its primitive's surface (fp16 positions); a point may lie inside an overlapping primitive
    assert d.max() < 2e-3 and (np.abs(d) < 2e-3).mean() > 0.75  # the room's walls overlap at the corners
    # oct-decoded normals against the numerical gradient of the scene distance
    ox, oy = nrm[..., 0][seen], nrm[..., 1][seen]
    v = np.stack([ox, oy, 1 - np.abs(ox) - np.abs(oy)], axis=1)
    neg = v[:, 2] < 0
    v[neg, :2] = (1 - np.abs(v[neg][:, [1, 0]])) * np.where(v[neg, :2] >= 0, 1.0, -1.0)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rng = np.random.default_rng(0)
    pick = rng.choice(np.nonzero(np.abs(d) < 2e-3)[0], 500, replace=False)  # points on the union's surface
    h = 1e-4
    g = np.stack([(prim_distance_np(prims, world[pick] + e) - prim_distance_np(prims, world[pick] - e)) / (2 * h)
                  for e in np.eye(3) * h], axis=1)
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    agree = np.sum(g * v[pick], axis=1) > 0.99
    assert agree.mean() > 0.95  # edges and contacts between primitives excepted

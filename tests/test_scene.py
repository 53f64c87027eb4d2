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

"""GPU parity of the global SDF build (luxgi_sdf_rasterize) against the FP64 oracle.

Tolerances: FP32 storage within 2e-5 (fp32 transforms and trilinear weights; the early-out
comparison is continuous in its result); FP16 storage within one half ulp (at most 2^-10 relative:
a value near a rounding boundary may round either way)."""
import os

import numpy as np
import pytest

import oracle
from luxgi_amd import abi, meshsdf, scene

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _objects():
    """100 primitives plus the two baked assets of the reference, rotated and scaled non-uniformly."""
    prims = scene.scene_random(n=100, seed=4)
    meshes, objs, bounds = meshsdf.primitive_objects(prims)
    rng = np.random.default_rng(7)
    for name in ("mesh_sdf_box_32x32x35.sdf", "mesh_sdf_box_51x32x32.sdf"):
        m = meshsdf.read_sdf(os.path.join(GOLDEN, name))
        x, y, z = m.size
        m.set_bounds(-0.002 * np.array([x, y, z]), 0.002 * np.array([x, y, z]))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, a, b, c = q
        R = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
                      [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
                      [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)]])
        world = np.eye(4)
        world[:3, :3] = R @ np.diag(rng.uniform(0.7, 1.4, 3))
        world[:3, 3] = rng.uniform(-0.5, 0.5, 3)
        meshes.append(m)
        bounds.append(meshsdf.world_bounds(m, world))
        objs_list = list(objs) + [meshsdf.object_rasterize_data(m, world, 0)]
        objs = (abi.ObjectRasterizeData * len(objs_list))(*objs_list)
    return meshes, objs, bounds


@pytest.mark.parametrize("half,mip", [(False, 0), (True, 0), (False, 1)])
def test_rasterize_matches_oracle(half, mip):
    meshes, objs, bounds = _objects()
    for i in range(len(objs)):
        objs[i].mipOffset = float(mip)
    n = len(meshes)
    res = 64
    cas = meshsdf.Cascade(0, (0, 0, 0), 1.0, res)
    lists = meshsdf.chunk_lists(cas, bounds)
    b = meshsdf.GlobalSDFBuilder(meshes, objs, n)
    vol = torch.full((res, res, res), 7.0, dtype=torch.float16 if half else torch.float32, device="cuda")
    b.rasterize(cas, lists, vol)
    torch.cuda.synchronize()
    got = vol.float().cpu().numpy()
    want = oracle.sdf_rasterize(cas.params(), objs, b.pool.entries, b.pool.pool, n, *lists, half=half)
    # one half ulp is at most 2^-10 |x| (2^-24 below the normal range)
    tol = np.maximum(2.0 ** -10 * np.abs(want), 2.0 ** -24) if half else 2e-5
    assert np.all(np.abs(got - want) <= tol), np.abs(got - want).max()
    assert (want < 0).sum() > 300 and want.max() > 0.2  # voxels inside objects and far from them


def test_rasterize_second_cascade_and_bad_arguments():
    import ctypes as C
    from luxgi_amd._lib import lib
    meshes, objs, bounds = _objects()
    n, res = len(meshes), 32
    cas = meshsdf.Cascade(1, (0, 0, 0), 2.5, res)  # cascade 1 of 2: x offset res, mip level 1
    lists = meshsdf.chunk_lists(cas, bounds)
    b = meshsdf.GlobalSDFBuilder(meshes, objs, n)
    vol = torch.full((res, res, 2 * res), 7.0, dtype=torch.float32, device="cuda")
    b.rasterize(cas, lists, vol, cascades=2)
    torch.cuda.synchronize()
    got = vol.cpu().numpy()
    assert np.all(got[:, :, :res] == 7.0)  # cascade 0 untouched
    want = oracle.sdf_rasterize(cas.params(), objs, b.pool.entries, b.pool.pool, n, *lists, cascades=2,
                                out=np.full((res, res, 2 * res), 7.0, np.float32))
    assert np.abs(got - want).max() <= 2e-5
    p = cas.params()
    p.cascadeIndex = 2  # beyond `cascades`
    z = C.c_void_p(vol.data_ptr())
    assert lib().luxgi_sdf_rasterize(C.byref(p), None, None, None, 0, z, z, z, 1, 2, abi.STORAGE_FP32, z, None) < 0

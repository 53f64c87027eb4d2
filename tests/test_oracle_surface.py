"""Oracle pins for the surface-atlas passes: chunk culling (SDFCulling.comp:36-101) and the
surface-cache lighting (CopyEmissive.frag, SDFDeferredLight.frag, SDFAtlasIndirectLight.frag).

Closed forms and invariants: single-sphere chunk sets at a chunk centre / face / corner, a sphere
covering the grid, object order within a run, the final counter and capacity overflow; an
independent float64 culling away from ties; GGX normalisation of the BRDF; the sample position of
the tile quads (bilinear of an affine emissive ramp); unshadowed direct light at n = l = v; point
light attenuation; a box occluder's shadow against an analytic ray-box test; the bounce with
constant probe atlases."""
import ctypes as C

import numpy as np
import pytest

import oracle
from luxgi_amd import abi, scene

R40 = 40


def _objects(bounds):
    objs = (abi.AtlasObject * max(len(bounds), 1))()
    for i, b in enumerate(bounds):
        objs[i].objectBounds[:] = [float(v) for v in b]
    return objs


def _runs(chunks, culled):
    out = {}
    for addr in np.nonzero(chunks)[0]:
        s = int(chunks[addr])
        out[int(addr)] = culled[s + 1:s + 1 + int(culled[s])].tolist()
    return out


def _addr(x, y, z):
    return (z * R40 + y) * R40 + x


def test_cull_single_sphere_closed_forms():
    cs = 0.05
    # centre of chunk (25, 10, 31), tiny radius: that chunk only
    c = [(25 - 20 + 0.5) * cs, (10 - 20 + 0.5) * cs, (31 - 20 + 0.5) * cs, 1e-4]
    ch, cu = oracle.cull_chunks(cs, _objects([c]), 1, 1000)
    assert _runs(ch, cu) == {_addr(25, 10, 31): [0]} and cu[0] == 3
    # on the face x = 0 between chunks 19 and 20: two chunks
    ch, cu = oracle.cull_chunks(cs, _objects([[0.0, 0.025, 0.025, 1e-4]]), 1, 1000)
    assert sorted(_runs(ch, cu)) == [_addr(19, 20, 20), _addr(20, 20, 20)]
    # on the corner at the origin: the 8 chunks around it
    ch, cu = oracle.cull_chunks(cs, _objects([[0.0, 0.0, 0.0, 1e-4]]), 1, 1000)
    want = sorted(_addr(x, y, z) for x in (19, 20) for y in (19, 20) for z in (19, 20))
    assert sorted(_runs(ch, cu)) == want and cu[0] == 1 + 8 * 2
    # a sphere covering the whole grid: every chunk lists it
    ch, cu = oracle.cull_chunks(cs, _objects([[0.0, 0.0, 0.0, 10.0]]), 1, 1 + 2 * R40 ** 3)
    assert np.all(ch > 0) and cu[0] == 1 + 2 * R40 ** 3


def test_cull_order_counter_and_capacity():
    cs = 0.05
    same = [[0.01, 0.02, 0.03, 0.2]] * 3
    ch, cu = oracle.cull_chunks(cs, _objects(same), 3, 100000)
    runs = _runs(ch, cu)
    assert runs and all(v == [0, 1, 2] for v in runs.values())
    total = 1 + 4 * len(runs)
    assert cu[0] == total
    # runs in address order: a capacity for k runs keeps the first k chunks and drops the rest
    k = 5
    ch2, cu2 = oracle.cull_chunks(cs, _objects(same), 3, 1 + 4 * k)
    kept = sorted(_runs(ch2, cu2))
    assert kept == sorted(runs)[:k] and cu2[0] == total


def test_cull_matches_float64_bruteforce():
    rng = np.random.default_rng(4)
    n = 60
    b = np.concatenate([rng.uniform(-1.1, 1.1, (n, 3)), rng.uniform(0.01, 0.3, (n, 1))], axis=1).astype(np.float32)
    cs = 2.0 / R40
    ch, cu = oracle.cull_chunks(cs, _objects(b), n, 10 ** 7)
    runs = _runs(ch, cu)
    lo = (np.arange(R40) - 20.0) * cs
    hi = lo + cs
    flips = 0
    for i in range(n):
        c, r = b[i, :3].astype(np.float64), float(b[i, 3])
        d = [np.maximum(np.maximum(lo - c[a], 0.0), c[a] - hi) for a in range(3)]
        dist = np.sqrt(d[2][:, None, None] ** 2 + d[1][None, :, None] ** 2 + d[0][None, None, :] ** 2)
        for z, y, x in zip(*np.nonzero(dist <= r + 1e-5)):
            inside = i in runs.get(_addr(x, y, z), [])
            if inside != (dist[z, y, x] <= r):
                assert abs(dist[z, y, x] - r) < 1e-5  # only a tie may differ
                flips += 1
        assert all(i not in v or True for v in runs.values())
    listed = sum(len(v) for v in runs.values())
    want = sum(int(np.sum(np.sqrt(
        np.maximum(np.maximum(lo - b[i, 2], 0.0), b[i, 2] - hi)[:, None, None] ** 2 +
        np.maximum(np.maximum(lo - b[i, 1], 0.0), b[i, 1] - hi)[None, :, None] ** 2 +
        np.maximum(np.maximum(lo - b[i, 0], 0.0), b[i, 0] - hi)[None, None, :] ** 2) <= b[i, 3])) for i in range(n))
    assert abs(listed - want) <= flips
    assert flips <= 2


def test_cull_host_restatement_matches_oracle():
    rng = np.random.default_rng(5)
    n = 80
    b = np.concatenate([rng.uniform(-1.0, 1.0, (n, 3)), rng.uniform(0.02, 0.25, (n, 1))], axis=1).astype(np.float32)
    for cap in (None, 3000):
        want_ch, want_cu = oracle.cull_chunks(0.05, _objects(b), n, cap if cap else 10 ** 7)
        ch, cu = scene.cull_chunks_np(b, 0.05, capacity=cap)
        assert np.array_equal(ch, want_ch)
        assert cu[0] == want_cu[0]
        assert np.array_equal(cu[1:], want_cu[1:len(cu)])


# ---------------------------------------------------------------------------
# BRDF
# ---------------------------------------------------------------------------
def test_brdf_at_normal_incidence():
    """n = l = v = h: F = F0, G = 1, D = 1 / (pi a^2) with a = rough^2; diffuse (1 - F0)(1 - m) albedo / pi."""
    n = [0.0, 0.0, 1.0]
    for rough, metal in ((1.0, 0.0), (0.5, 0.3), (0.8, 1.0)):
        alb = np.array([0.5, 0.25, 0.9])
        got = oracle.brdf(alb, n, rough, metal, n, n, n)
        f0 = 0.04 * (1 - metal) + alb * metal
        a2 = rough ** 4
        D = a2 / (np.pi * a2 * a2)
        want = (1 - f0) * (1 - metal) * alb / np.pi + f0 * D / 4.0
        np.testing.assert_allclose(got, want, rtol=1e-7)  # Fdielectric is the fp32 literal 0.04f


def test_brdf_ggx_normalisation():
    """With l = v = h at angle t from n, spec = F0 D G / (4 cos^2 t) with F = F0, so D follows from
    the BRDF; the GGX distribution must satisfy integral D(h) (n.h) dw_h = 1 over the hemisphere."""
    for rough in (0.3, 0.6, 0.9):
        k = (rough + 1) ** 2 / 8
        ts = np.linspace(0.0, np.pi / 2, 4001)[:-1] + (np.pi / 2) / 8000
        D = np.empty_like(ts)
        for j, t in enumerate(ts):
            h = [np.sin(t), 0.0, np.cos(t)]
            spec = oracle.brdf([0, 0, 0], [0, 0, 1], rough, 0.0, h, h, h)[0]
            c = np.cos(t)
            G = (c / (c * (1 - k) + k)) ** 2
            D[j] = spec * 4 * c * c / (0.04 * G)
        integral = 2 * np.pi * np.sum(D * np.cos(ts) * np.sin(ts)) * (np.pi / 2 / 4000)
        assert abs(integral - 1.0) < 2e-3, (rough, integral)


# ---------------------------------------------------------------------------
# surface-cache lighting
# ---------------------------------------------------------------------------
def _one_tile(res, X, Y, W, H):
    tiles = (abi.AtlasTile * 1)()
    tiles[0].extends[:] = [X / res, Y / res, (W - 1) / res, (H - 1) / res]
    return tiles


def _draws(mask=1):
    d = (abi.SurfaceDraw * 1)()
    d[0].tile, d[0].lightMask = 0, mask
    return d


def _params(passes, cam=(0.0, 0.0, 5.0), bias=0.01, n_lights=0, intensity=1.0):
    p = abi.SurfaceLightParams()
    p.cameraPosition[:] = [cam[0], cam[1], cam[2], bias]
    p.passes = passes
    p.numLights = n_lights
    p.indirectIntensity = intensity
    return p


def _empty_sdf(res=32):
    sdf = np.ones((res, res, res), np.float32)
    return abi.make_sdf_data(resolution=res), sdf, oracle.sdf_mip(sdf, res, max_distance=2.0), res


def test_emissive_pass_samples_the_contracted_tile_position():
    """The tile quad spans the full rect X..X+W while the tile UV maps onto W - 1 texels, so texel
    p reads the G-buffer at x = X + (p + 0.5 - X)(W - 1)/W (in texel centres: minus 0.5). On an
    affine ramp the bilinear sample is the ramp's value there."""
    res, X, Y, W, H = 64, 10, 20, 16, 12
    em = np.zeros((res, res, 4), np.float32)
    xs = np.arange(res, dtype=np.float64)
    em[..., 0] = (0.25 * xs)[None, :]
    em[..., 1] = (0.5 * xs)[:, None].astype(np.float32)
    z = np.zeros_like(em)
    out = oracle.surface_light(_params(abi.SURFACE_EMISSIVE), None, _one_tile(res, X, Y, W, H), _draws(), res,
                               z, z, z, em)
    px = np.arange(X, X + W)
    py = np.arange(Y, Y + H)
    sx = X + (px + 0.5 - X) * (W - 1) / W - 0.5
    sy = Y + (py + 0.5 - Y) * (H - 1) / H - 0.5
    np.testing.assert_allclose(out[Y:Y + H, X:X + W, 0], np.broadcast_to(0.25 * sx, (H, W)), rtol=1e-6)
    np.testing.assert_allclose(out[Y:Y + H, X:X + W, 1], np.broadcast_to((0.5 * sy)[:, None], (H, W)), rtol=1e-6)
    assert np.all(out[Y:Y + H, X:X + W, 3] == 1)
    mask = np.ones((res, res), bool)
    mask[Y:Y + H, X:X + W] = False
    assert np.all(out[mask] == 0)  # texels outside the draw are untouched


def _flat_gbuffer(res, world, normal_oct, albedo, metal, rough):
    col = np.zeros((res, res, 4), np.float32)
    nrm = np.zeros((res, res, 4), np.float32)
    pbr = np.zeros((res, res, 4), np.float32)
    col[..., :3] = albedo
    col[..., 3] = world[0]
    nrm[..., :2] = normal_oct
    nrm[..., 2:] = world[1:]
    pbr[..., 0], pbr[..., 1] = metal, rough
    return col, nrm, pbr, np.zeros_like(col)


def _lights(*ls):
    arr = (abi.Light * len(ls))()
    for i, d in enumerate(ls):
        for k, v in d.items():
            if isinstance(v, (list, tuple)):
                getattr(arr[i], k)[:len(v)] = v
            else:
                setattr(arr[i], k, v)
    return arr


def test_direct_light_unshadowed_closed_form():
    """Flat +z surface, empty SDF, camera and a directional light straight above: n = l = v = h.
    light = brdf(n=l=v) * colour * (I^1.4 + 0.1)."""
    res = 32
    g, sdf, mip, sres = _empty_sdf()
    alb, metal, rough = np.array([0.6, 0.4, 0.2], np.float32), 0.25, 0.7
    gbuf = _flat_gbuffer(res, [0.1, -0.2, 0.0], [0.0, 0.0], alb, metal, rough)
    L = _lights(dict(color=[1.0, 0.5, 2.0, 1.0], direction=[0.0, 0.0, -1.0, 0.0], intensity=3.0, type=0.0))
    p = _params(abi.SURFACE_DIRECT, cam=(0.1, -0.2, 4.0), n_lights=1)
    out = oracle.surface_light(p, L, _one_tile(res, 4, 4, 8, 8), _draws(), res, *gbuf, sdf_data=g, sdf=sdf, mip=mip,
                               sdf_res=sres, mip_res=sres // 4)
    f0 = 0.04 * (1 - metal) + alb.astype(np.float64) * metal
    a2 = rough ** 4
    brdf = (1 - f0) * (1 - metal) * alb / np.pi + f0 * (1 / (np.pi * a2)) / 4
    want = brdf * np.array([1.0, 0.5, 2.0]) * (3.0 ** 1.4 + 0.1)
    np.testing.assert_allclose(out[4:12, 4:12, :3], np.broadcast_to(want, (8, 8, 3)), rtol=2e-6)


def test_point_and_spot_attenuation():
    """Point light at distance d straight above: attenuation radius / (d^2 + 1). A spot light whose
    cone misses the point contributes nothing (attenuation clamped to 0)."""
    res = 32
    g, sdf, mip, sres = _empty_sdf()
    gbuf = _flat_gbuffer(res, [0.0, 0.0, 0.0], [0.0, 0.0], np.array([0.5, 0.5, 0.5], np.float32), 0.0, 1.0)
    d = 0.6
    L = _lights(dict(color=[1.0, 1.0, 1.0, 1.0], position=[0.0, 0.0, d, 1.0], intensity=1.0, radius=2.0, type=2.0),
                dict(color=[1.0, 1.0, 1.0, 1.0], position=[0.0, 0.0, d, 1.0], direction=[1.0, 0.0, 0.0, 0.0],
                     intensity=1.0, radius=2.0, type=1.0, angle=0.1))
    base = dict(sdf_data=g, sdf=sdf, mip=mip, sdf_res=sres, mip_res=sres // 4)
    p = _params(abi.SURFACE_DIRECT, cam=(0.0, 0.0, 3.0), n_lights=2)
    pt = oracle.surface_light(p, L, _one_tile(res, 0, 0, 4, 4), _draws(0b01), res, *gbuf, **base)
    both = oracle.surface_light(p, L, _one_tile(res, 0, 0, 4, 4), _draws(0b11), res, *gbuf, **base)
    brdf = 0.96 * 0.5 / np.pi + 0.04 / np.pi / 4
    np.testing.assert_allclose(pt[0, 0, :3], brdf * 1.1 * 2.0 / (d * d + 1), rtol=1e-6)
    np.testing.assert_allclose(both, pt, rtol=1e-12)


def test_shadow_of_a_box_occluder():
    """Surface z = -0.5 lit from +z, a box blocker above part of it: a texel is dark exactly when
    its shadow ray (from world + n * bias, started 2 voxels along the ray) meets the box, up to the
    hit test's half-voxel thickness and the trilinear reconstruction near the box's edges."""
    res, sres = 64, 64
    prims = np.zeros(1, dtype=scene.PRIM_DTYPE)
    prims["type"] = scene.PRIM_BOX
    prims["center"] = (0.1, 0.0, 0.2)
    prims["size"] = (0.25, 0.3, 0.05)
    sdf = scene.bake_sdf_numpy(prims, sres)
    g = abi.make_sdf_data(resolution=sres)
    mip = oracle.sdf_mip(sdf, sres, max_distance=2.0)
    # a 64x64 tile mapping world x, y in [-0.8, 0.8] at z = -0.5
    col = np.zeros((res, res, 4), np.float32)
    nrm = np.zeros((res, res, 4), np.float32)
    pbr = np.zeros((res, res, 4), np.float32)
    W = 64
    xs = (np.arange(res) + 0.5) / res * 1.6 - 0.8
    col[..., :3] = 0.5
    col[..., 3] = xs[None, :]
    nrm[..., 2] = xs[:, None]
    nrm[..., 3] = -0.5
    pbr[..., 1] = 1.0
    L = _lights(dict(color=[1.0, 1.0, 1.0, 1.0], direction=[0.0, 0.0, -1.0, 0.0], intensity=1.0, type=0.0))
    p = _params(abi.SURFACE_DIRECT, cam=(0.0, 0.0, 3.0), bias=0.02, n_lights=1)
    tiles = _one_tile(res, 0, 0, W, W)
    out = oracle.surface_light(p, L, tiles, _draws(), res, col, nrm, pbr, np.zeros_like(col), sdf_data=g, sdf=sdf,
                               mip=mip, sdf_res=sres, mip_res=sres // 4)
    # the G-buffer is sampled at the contracted position: recover each texel's world x, y
    sx = (np.arange(W) + 0.5) * (W - 1) / W - 0.5
    wx = np.interp(sx, np.arange(res), xs)
    dark = out[..., 0] == 0
    lit = out[..., 0] > 0
    assert np.all(dark | lit)
    voxel = 2.0 / sres
    bx = np.abs(wx - 0.1)  # distance to the box footprint in x, y (the ray is vertical)
    inside = (bx[None, :] < 0.25 - 1.5 * voxel) & (np.abs(wx)[:, None] < 0.3 - 1.5 * voxel)
    outside = (bx[None, :] > 0.25 + 1.5 * voxel) | (np.abs(wx)[:, None] > 0.3 + 1.5 * voxel)
    assert inside.sum() > 300 and outside.sum() > 1000
    assert np.all(dark[inside]) and np.all(lit[outside])


def test_bounce_with_constant_probe_atlases():
    """Constant irradiance texels v and a far depth: sampleIrradiance returns (v^(gamma/2))^2 2 pi for
    any point and normal, so the bounce is intensity * (albedo' - albedo' m) / pi * that, with
    albedo' = min(albedo, 0.9) (SDFAtlasIndirectLight.frag:44-48,60)."""
    res = 16
    u = abi.make_uniform((3, 3, 3), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 64, irr_oct=6, depth_oct=14)
    irr = np.full((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), 0.7, np.float32)
    dep = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    dep[..., 0], dep[..., 1] = 10.0, 100.0
    alb = np.array([0.95, 0.5, 0.1], np.float32)
    gbuf = _flat_gbuffer(res, [0.13, -0.42, 0.31], [0.3, -0.2], alb, 0.2, 0.5)
    p = _params(abi.SURFACE_INDIRECT, intensity=1.7)
    out = oracle.surface_light(p, None, _one_tile(res, 2, 2, 8, 8), _draws(0), res, *gbuf, u=u, irradiance=irr,
                               depth_atlas=dep)
    e = (0.7 ** (u.ddgiGamma / 2)) ** 2 * 2 * np.pi
    a = np.minimum(alb.astype(np.float64), 0.9)
    want = 1.7 * (a - a * 0.2) / np.pi * e
    np.testing.assert_allclose(out[2:10, 2:10, :3], np.broadcast_to(want, (8, 8, 3)), rtol=1e-5)

"""Oracle pins for the noisy-signal producers: the blue-noise sampler (Raytraced/BlueNoise.glsl)
and the SDF soft-shadow mask (SDF/SDFShadow.comp).

Pins: hand-evaluated sampler values on small explicit tables; which table channel drives which
part of the sampler; 256-frame stratification of every pixel and dimension when the Sobol
columns are permutations (the property the scrambling-ranking design provides); wrapping of
pixel coordinates at 128 and frames at 256. The mask: a box occluder above a plane lit
vertically (a zero-radius light) against an analytic ray-box test; a soft light keeps the
umbra dark and the far field lit; the bit layout x + 8y of 8x4 blocks; a block whose first pixel
is sky is not written."""
import numpy as np

import oracle
from luxgi_amd import abi, scene


def _tables(seed=0, perm=True):
    rng = np.random.default_rng(seed)
    sobol = np.stack([rng.permutation(256) for _ in range(4)], axis=1).astype(np.uint8) if perm else \
        rng.integers(0, 256, (256, 4), dtype=np.uint8)
    scr = rng.integers(0, 256, (128, 128, 4), dtype=np.uint8)
    return sobol, scr


def test_blue_noise_hand_values():
    sobol = np.zeros((256, 4), np.uint8)
    k = np.arange(256)
    sobol[:, 0], sobol[:, 1], sobol[:, 2], sobol[:, 3] = k, 255 - k, (3 * k) % 256, 7
    scr = np.zeros((128, 128, 4), np.uint8)
    scr[7, 5] = (9, 200, 5, 77)  # (r, g, b, a) at x = 5, y = 7
    # dim 0: sobol[f ^ b][0] ^ r = (3 ^ 5) ^ 9 = 15
    assert oracle.blue_noise(sobol, scr, 5, 7, 3, 0) == (0.5 + 15) / 256
    # dim 1: sobol[3 ^ 5][1] ^ g = (255 - 6) ^ 200 = 249 ^ 200 = 49
    assert oracle.blue_noise(sobol, scr, 5, 7, 3, 1) == (0.5 + 49) / 256
    # dim 2 uses channel r again (dimension % 2): (3 * 6) ^ 9 = 18 ^ 9 = 27
    assert oracle.blue_noise(sobol, scr, 5, 7, 3, 2) == (0.5 + 27) / 256
    # dim 3: 7 ^ 200 = 207; coordinates wrap at 128, frames at 256, dimensions at 4
    assert oracle.blue_noise(sobol, scr, 5, 7, 3, 3) == (0.5 + 207) / 256
    assert oracle.blue_noise(sobol, scr, 133, 7 + 256, 3 + 512, 4) == oracle.blue_noise(sobol, scr, 5, 7, 3, 0)


def test_blue_noise_channels():
    sobol, scr = _tables(1)
    base = [[oracle.blue_noise(sobol, scr, 3, 4, f, d) for d in range(4)] for f in range(16)]
    s2 = scr.copy()
    s2[4, 3, 3] ^= 0xFF  # alpha: unused
    assert base == [[oracle.blue_noise(sobol, s2, 3, 4, f, d) for d in range(4)] for f in range(16)]
    s3 = scr.copy()
    s3[4, 3, 0] ^= 0x5A  # red: dimensions 0 and 2 only
    got = [[oracle.blue_noise(sobol, s3, 3, 4, f, d) for d in range(4)] for f in range(16)]
    for f in range(16):
        assert got[f][1] == base[f][1] and got[f][3] == base[f][3]
        assert got[f][0] != base[f][0] and got[f][2] != base[f][2]


def test_blue_noise_stratifies_256_frames():
    sobol, scr = _tables(2, perm=True)
    strata = np.arange(256)
    for (x, y) in ((0, 0), (17, 99), (127, 127)):
        for d in range(4):
            v = np.array([oracle.blue_noise(sobol, scr, x, y, f, d) for f in range(256)])
            assert np.array_equal(np.sort(np.round(v * 256 - 0.5).astype(int)), strata)


def _flat_view(W, H):
    """viewProjInv mapping pixel (u, v) to world (0.8 (2u - 1), 0.8 (2v - 1), -0.5) for any depth."""
    m = np.array([[0.8, 0, 0, 0], [0, 0.8, 0, 0], [0, 0, 0, -0.5], [0, 0, 0, 1.0]])
    return abi.mat4_colmajor(m)


def _box_sdf(res=64):
    prims = np.zeros(1, dtype=scene.PRIM_DTYPE)
    prims["type"] = scene.PRIM_BOX
    prims["center"] = (0.1, 0.0, 0.2)
    prims["size"] = (0.25, 0.3, 0.05)
    sdf = scene.bake_sdf_numpy(prims, res)
    return abi.make_sdf_data(resolution=res), sdf, oracle.sdf_mip(sdf, res, max_distance=2.0)


def _params(radius, W, H):
    p = abi.SdfShadowParams()
    p.light.color[:] = [1, 1, 1, 1]
    p.light.direction[:] = [0.0, 0.0, -1.0, radius]
    p.light.intensity, p.light.type = 1.0, abi.LIGHT_DIRECTIONAL
    p.viewProjInv = _flat_view(W, H)
    p.numFrames, p.shadowBias = 5, 0.02
    return p


def _bits(mask, W, H):
    mh, mw = mask.shape
    out = np.zeros((mh * 4, mw * 8), np.uint8)
    for k in range(32):
        out[k >> 3::4, k & 7::8] = (mask >> np.uint32(k)) & 1
    return out[:H, :W]


def test_shadow_mask_box_occluder_hard_and_soft():
    W, H, res = 64, 64, 64
    g, sdf, mip = _box_sdf(res)
    depth = np.full((H, W), 0.5, np.float32)
    normals = np.zeros((H, W, 4), np.float32)  # octahedral (0, 0) = +z
    sobol, scr = _tables(3)
    xs = 0.8 * (2 * (np.arange(W) + 0.5) / W - 1)
    voxel = 2.0 / res
    fx = np.abs(xs - 0.1)[None, :]
    fy = np.abs(xs)[:, None]
    hard = _bits(oracle.shadow_mask(_params(0.0, W, H), sobol, scr, W, H, depth, normals, g, sdf, mip, res, res // 4), W, H)
    inside = (fx < 0.25 - 1.5 * voxel) & (fy < 0.3 - 1.5 * voxel)
    outside = (fx > 0.25 + 1.5 * voxel) | (fy > 0.3 + 1.5 * voxel)
    assert inside.sum() > 300 and outside.sum() > 1000
    assert np.all(hard[inside] == 0) and np.all(hard[outside] == 1)
    # a light of angular radius r: the disk's rays from height 0.7 - 0.05 - bias above the plane
    # wander by at most r * 0.7 sideways before they pass the box's top
    r = 0.05
    soft = _bits(oracle.shadow_mask(_params(r, W, H), sobol, scr, W, H, depth, normals, g, sdf, mip, res, res // 4), W, H)
    m = r * 0.8 + 1.5 * voxel
    assert np.all(soft[(fx < 0.25 - m) & (fy < 0.3 - m)] == 0)
    assert np.all(soft[(fx > 0.25 + m) | (fy > 0.3 + m)] == 1)
    pen = ~((fx < 0.25 - m) & (fy < 0.3 - m)) & ~((fx > 0.25 + m) | (fy > 0.3 + m))
    assert 0.05 < soft[pen].mean() < 0.95  # the penumbra is mixed


def test_shadow_mask_sky_blocks_are_not_written():
    W, H, res = 16, 8, 32
    g = abi.make_sdf_data(resolution=res)
    sdf = np.ones((res, res, res), np.float32)
    mip = oracle.sdf_mip(sdf, res, max_distance=2.0)
    depth = np.full((H, W), 0.5, np.float32)
    depth[0, 0] = 1.0  # first pixel of block (0, 0) is sky
    depth[4, 9] = 1.0  # a sky pixel that is not a block's first
    normals = np.zeros((H, W, 4), np.float32)
    sobol, scr = _tables(4)
    prior = np.full((2, 2), 0xDEADBEEF, np.uint32)
    mask = oracle.shadow_mask(_params(0.0, W, H), sobol, scr, W, H, depth, normals, g, sdf, mip, res, res // 4, mask=prior)
    assert mask[0, 0] == 0xDEADBEEF
    full = np.uint32(0xFFFFFFFF)
    assert mask[0, 1] == full and mask[1, 0] == full
    assert mask[1, 1] == full & ~np.uint32(1 << (0 * 8 + 1))  # pixel (9, 4): x % 8 = 1, y % 4 = 0


# ---------------------------------------------------------------------------
# SDF reflection (SDF/SDFReflection.comp)
# ---------------------------------------------------------------------------
def _refl_params(W, H, cam=(0.0, 0.0, 2.0), trim=1.0, ddgi=0):
    p = abi.SdfReflectionParams()
    p.trim, p.roughDDGIIntensity, p.approximateWithDDGI = trim, 1.5, ddgi
    p.cameraPosition[:] = [cam[0], cam[1], cam[2], 1.0]
    p.viewProjInv = _flat_view(W, H)
    p.numFrames = 3
    p.skyColor[:] = [0.2, 0.3, 0.4, 1.0]
    return p


def _floor_gbuffer(W, H, rough):
    depth = np.full((H, W), 0.5, np.float32)
    normals = np.zeros((H, W, 4), np.float32)  # +z
    pbr = np.zeros((H, W, 4), np.float32)
    pbr[..., 1] = rough
    return depth, normals, pbr


def _empty_scene(res=32):
    g = abi.make_sdf_data(resolution=res)
    sdf = np.ones((res, res, res), np.float32)
    prims = scene.scene_small(seed=1)
    atlas = scene.build_atlas(prims, resolution=512, tile=32, threshold=0.05)
    return g, sdf, oracle.sdf_mip(sdf, res, max_distance=2.0), atlas


def _pixel_world(W, H):
    xs = 0.8 * (2 * (np.arange(W) + 0.5) / W - 1)
    ys = 0.8 * (2 * (np.arange(H) + 0.5) / H - 1)
    X, Y = np.meshgrid(xs, ys)
    return np.stack([X, Y, np.full_like(X, -0.5)], axis=-1)


def _mirror(W, H, cam):
    wo = np.asarray(cam)[None, None, :] - _pixel_world(W, H)
    wo /= np.linalg.norm(wo, axis=-1, keepdims=True)
    return np.stack([-wo[..., 0], -wo[..., 1], wo[..., 2]], axis=-1)  # reflect(-Wo, +z)


def test_reflection_mirror_reads_the_sky_in_the_mirror_direction():
    """Roughness below 0.05: R = reflect(-Wo, n). With an empty SDF every ray misses, and a cube map
    holding its texels' own directions returns R (to the cube's interpolation error)."""
    from tests.test_oracle_trace import _cube_texel_directions
    W, H, size = 32, 24, 32
    g, sdf, mip, atlas = _empty_scene()
    depth, normals, pbr = _floor_gbuffer(W, H, 0.01)
    sobol, scr = _tables(5)
    cam = (0.1, -0.2, 2.0)
    out = oracle.sdf_reflection(_refl_params(W, H, cam), sobol, scr, W, H, normals, pbr, depth, g, sdf, mip, 32, 8,
                                atlas, sky_faces=_cube_texel_directions(size), sky_size=size)
    err = np.linalg.norm(out[..., :3] - _mirror(W, H, cam), axis=-1)
    assert err.max() < 0.75 / size, err.max()
    # without a cube map the constant sky colour
    out2 = oracle.sdf_reflection(_refl_params(W, H, cam), sobol, scr, W, H, normals, pbr, depth, g, sdf, mip, 32, 8,
                                 atlas)
    np.testing.assert_allclose(out2.reshape(-1, 4), np.broadcast_to([0.2, 0.3, 0.4, 1.0], (W * H, 4)), rtol=1e-7)


def test_reflection_ggx_lobe():
    """GGX branch (0.05 <= roughness <= 0.45): with trim 0 the sample is E = (0, 0), whose half
    vector is n, so the ray is the mirror ray. With trim 1, over 256 frames of one pixel (a
    stratified sample set), the half vectors' cos(theta) follow the GGX distribution
    D(h) cos(theta) (D normalised as in the BRDF)."""
    from tests.test_oracle_trace import _cube_texel_directions
    W, H, size = 8, 8, 64
    g, sdf, mip, atlas = _empty_scene()
    rough = 0.4
    depth, normals, pbr = _floor_gbuffer(W, H, rough)
    sobol, scr = _tables(6)
    faces = _cube_texel_directions(size)
    cam = (0.0, 0.0, 2.0)
    zero = oracle.sdf_reflection(_refl_params(W, H, cam, trim=0.0), sobol, scr, W, H, normals, pbr, depth, g, sdf, mip,
                                 32, 8, atlas, sky_faces=faces, sky_size=size)
    assert np.linalg.norm(zero[..., :3] - _mirror(W, H, cam), axis=-1).max() < 0.75 / size
    cos_h = []
    px, py = 4, 4
    wo = np.asarray(cam) - _pixel_world(W, H)[py, px]
    wo /= np.linalg.norm(wo)
    for f in range(256):
        p = _refl_params(W, H, cam)
        p.numFrames = f
        out = oracle.sdf_reflection(p, sobol, scr, W, H, normals, pbr, depth, g, sdf, mip, 32, 8, atlas,
                                    sky_faces=faces, sky_size=size)
        r = out[py, px, :3] / np.linalg.norm(out[py, px, :3])
        h = r + wo
        cos_h.append(h[2] / np.linalg.norm(h))
    cos_h = np.sort(np.array(cos_h))
    a2 = rough ** 4
    mu = np.linspace(0.0, 1.0, 20001)
    D = a2 / (np.pi * (mu * mu * (a2 - 1) + 1) ** 2)
    dens = 2 * np.pi * D * mu  # over d(mu)
    cdf_above = np.cumsum(dens[::-1])[::-1] * (mu[1] - mu[0])  # P(cos >= mu)
    for c in (0.8, 0.9, 0.95, 0.98):
        emp = (cos_h >= c).mean()
        ana = np.interp(c, mu, cdf_above)
        assert abs(emp - ana) < 0.04, (c, emp, ana)


def test_reflection_ddgi_branch_constant_atlases():
    """Roughness above 0.45 with approximateWithDDGI: roughDDGIIntensity * sampleIrradiance in the
    mirror direction; constant irradiance v gives (v^(gamma/2))^2 2 pi, alpha 0."""
    W, H = 16, 8
    g, sdf, mip, atlas = _empty_scene()
    depth, normals, pbr = _floor_gbuffer(W, H, 0.7)
    sobol, scr = _tables(7)
    u = abi.make_uniform((3, 3, 3), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 64, irr_oct=6, depth_oct=14)
    irr = np.full((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), 0.6, np.float32)
    dep = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    dep[..., 0], dep[..., 1] = 10.0, 100.0
    out = oracle.sdf_reflection(_refl_params(W, H, ddgi=1), sobol, scr, W, H, normals, pbr, depth, g, sdf, mip, 32, 8,
                                atlas, u=u, irradiance=irr, depth_atlas=dep)
    e = 1.5 * (0.6 ** (u.ddgiGamma / 2)) ** 2 * 2 * np.pi
    np.testing.assert_allclose(out[..., :3], e, rtol=1e-6)
    assert np.all(out[..., 3] == 0)
    # without approximateWithDDGI the same pixels trace (and read the sky here)
    out2 = oracle.sdf_reflection(_refl_params(W, H, ddgi=0), sobol, scr, W, H, normals, pbr, depth, g, sdf, mip, 32, 8,
                                 atlas, u=u, irradiance=irr, depth_atlas=dep)
    np.testing.assert_allclose(out2[..., :3], np.broadcast_to([0.2, 0.3, 0.4], (H, W, 3)), rtol=1e-7)


def test_reflection_hit_reads_the_atlas_colour_and_skips_sky():
    """Mirror rays from a floor up into a box of constant atlas colour c return c (the lookup
    normalises by its weight sum) with alpha = that sum > 0; sky pixels are not written."""
    W, H, res = 32, 32, 64
    prims = np.zeros(1, dtype=scene.PRIM_DTYPE)
    prims["type"] = scene.PRIM_BOX
    prims["center"] = (0.0, 0.0, 0.3)
    prims["size"] = (0.5, 0.5, 0.1)
    prims["albedo"] = (0.3, 0.6, 0.9)
    sdf = scene.bake_sdf_numpy(prims, res)
    g = abi.make_sdf_data(resolution=res)
    mip = oracle.sdf_mip(sdf, res, max_distance=2.0)
    atlas = scene.build_atlas(prims, resolution=256, tile=32, threshold=1.05 * 2 / res, emissive_fraction=0.0)
    colour = np.array([0.3, 0.6, 0.9], np.float32)
    atlas.color_f32[..., :3] = colour
    depth, normals, pbr = _floor_gbuffer(W, H, 0.01)
    depth[5, 6] = 1.0
    sobol, scr = _tables(8)
    prior = np.full((H, W, 4), -3.0)
    out = oracle.sdf_reflection(_refl_params(W, H, cam=(0.0, 0.0, 5.0)), sobol, scr, W, H, normals, pbr, depth, g,
                                sdf, mip, res, res // 4, atlas, out=prior)
    assert np.all(out[5, 6] == -3.0)
    # nearly vertical mirror rays: floor points well under the box hit it, far ones see the sky
    wp = _pixel_world(W, H)
    under = (np.abs(wp[..., 0]) < 0.35) & (np.abs(wp[..., 1]) < 0.35)
    away = (np.abs(wp[..., 0]) > 0.7) | (np.abs(wp[..., 1]) > 0.7)
    under[5, 6] = away[5, 6] = False
    assert under.sum() > 150 and away.sum() > 50
    assert np.all(out[under][:, 3] > 0)
    np.testing.assert_allclose(out[under][:, :3], np.broadcast_to(colour, (under.sum(), 3)), rtol=1e-5)
    np.testing.assert_allclose(out[away], np.broadcast_to([0.2, 0.3, 0.4, 1.0], (away.sum(), 4)), rtol=1e-7)

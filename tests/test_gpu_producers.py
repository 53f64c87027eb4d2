"""GPU parity of the noisy-signal producers: luxgi_sdf_shadow_mask against the FP64 oracle.

Tolerance. Each bit is a chain of threshold decisions (the march's hit test) on a ray whose
direction comes through fp32 trigonometry of the blue-noise disk sample, so a bit may flip
where the ray grazes geometry: >= 99.5% of the bits of the written words agree, and the set of
written words (the blocks whose first pixel is not sky) is identical."""
import numpy as np
import pytest

import oracle
from luxgi_amd import abi, scene
from tests import _scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _noise(seed=0):
    rng = np.random.default_rng(seed)
    sobol = np.stack([rng.permutation(256) for _ in range(4)], axis=1).astype(np.uint8)
    return sobol, rng.integers(0, 256, (128, 128, 4), dtype=np.uint8)


def _light(kind):
    p = abi.SdfShadowParams()
    p.light.color[:] = [1, 1, 1, 1]
    if kind == "sun":
        d = np.array([0.3, -0.5, -0.81])
        d /= np.linalg.norm(d)
        p.light.direction[:] = [d[0], d[1], d[2], 0.03]
        p.light.type = abi.LIGHT_DIRECTIONAL
    elif kind == "point":
        p.light.position[:] = [0.1, 0.55, 0.2, 1.0]
        p.light.direction[:] = [0.0, 0.0, 0.0, 0.05]
        p.light.radius, p.light.type = 2.0, abi.LIGHT_POINT
    else:
        p.light.position[:] = [-0.4, 0.6, 0.5, 1.0]
        # the reference's spot test is dot(surface-to-light, direction) > 1 - angle (SDFShadow.comp:84-86),
        # so `direction` points from the lit surface toward the light
        p.light.direction[:] = [-0.4, 0.8, 0.45, 0.04]
        p.light.radius, p.light.type, p.light.angle = 2.0, abi.LIGHT_SPOT, 0.4
    p.light.intensity = 1.0
    p.shadowBias = 0.02
    return p


# the room of the small scene is closed: the sun lights its outer walls (camera outside), the
# point and spot lights its interior (camera inside)
@pytest.mark.parametrize("kind,W,H,layout,eye", [("sun", 256, 160, abi.LAYOUT_BRICK4, (0.1, 0.3, 2.2)),
                                                  ("point", 250, 158, abi.LAYOUT_LINEAR, (0.1, 0.3, 0.75)),
                                                  ("spot", 200, 120, abi.LAYOUT_BRICK4, (0.2, 0.4, 0.7)),
                                                  ("sun", 160, 96, abi.LAYOUT_CELL8, (0.1, 0.3, 2.2))])
def test_shadow_mask_matches_oracle(kind, W, H, layout, eye):
    from luxgi_amd.ddgi import DeviceScene, to_device
    from luxgi_amd.producers import BlueNoiseTables, sdf_shadow_mask
    from luxgi_amd.svgf import make_gbuffer
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    VP, VPp, _, _ = scene.camera_pair(W, H, seed=7, eye=eye)
    VPI = np.linalg.inv(VP)
    prims_dev = to_device(np.frombuffer(prims.tobytes(), dtype=np.uint8))
    gb, _, _ = make_gbuffer(prims_dev, len(prims), g, torch.from_numpy(sdf).cuda(), 64, VPI, VPp, W, H, seed=3,
                            with_mask=False)
    ds = DeviceScene(g, sdf, mip, atlas, layout=layout)
    sobol, scr = _noise(1)
    nz = BlueNoiseTables(sobol, scr)
    p = _light(kind)
    p.viewProjInv = abi.mat4_colmajor(VPI)
    p.numFrames = 11
    prior = torch.full(((H + 3) // 4, (W + 7) // 8), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    got = sdf_shadow_mask(p, nz, gb["depth"], gb["normals"], ds, mask=prior.clone())
    torch.cuda.synchronize()
    got = got.cpu().numpy().view(np.uint32)
    depth = gb["depth"].cpu().numpy()
    want = oracle.shadow_mask(p, sobol, scr, W, H, depth, gb["normals"].cpu().numpy(), g, sdf, mip, 64, 16,
                              mask=np.full(got.shape, 0x5A5A5A5A, np.uint32))
    written = depth[0::4, 0::8] != 1.0
    assert written.shape == got.shape
    assert np.all(got[~written] == 0x5A5A5A5A) and np.all(want[~written] == 0x5A5A5A5A)
    bits_g = np.unpackbits(got[written].view(np.uint8))
    bits_w = np.unpackbits(want[written].view(np.uint8))
    agree = (bits_g == bits_w).mean()
    assert agree >= 0.995, agree
    assert 0.005 < bits_w.mean() < 0.995  # both lit and shadowed pixels


def test_shadow_mask_rejects_bad_arguments():
    from luxgi_amd._lib import lib
    import ctypes as C
    p = abi.SdfShadowParams()
    assert lib().luxgi_sdf_shadow_mask(C.byref(p), None, 8, 8, None, None, None, None, None, None) < 0


@pytest.mark.parametrize("layout,atlas_layout", [(abi.LAYOUT_BRICK4, abi.LAYOUT_LINEAR), (abi.LAYOUT_LINEAR, abi.LAYOUT_BRICK4),
                                                 (abi.LAYOUT_CELL8, abi.LAYOUT_BRICK4)])
def test_sdf_reflection_matches_oracle(layout, atlas_layout):
    """All three branches (mirror, GGX lobe, DDGI) against the FP64 oracle: >= 99% of non-sky pixels
    agree to 2e-3 (1 + |v|) on every channel (a GGX ray through fp32 trigonometry may graze past
    a different surface), and sky pixels keep the buffer's previous value."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene, to_device
    from luxgi_amd.producers import BlueNoiseTables, sdf_reflection
    from luxgi_amd.svgf import make_gbuffer
    W, H = 192, 128
    color_storage = abi.STORAGE_UNORM8 if atlas_layout == abi.LAYOUT_BRICK4 else abi.STORAGE_FP32
    prims, sdf, mip, g, atlas = _scenes.small_scene(color_storage=color_storage)
    VP, VPp, eye, _ = scene.camera_pair(W, H, seed=9, eye=(0.1, 0.3, 0.75))
    VPI = np.linalg.inv(VP)
    prims_dev = to_device(np.frombuffer(prims.tobytes(), dtype=np.uint8))
    gb, _, _ = make_gbuffer(prims_dev, len(prims), g, torch.from_numpy(sdf).cuda(), 64, VPI, VPp, W, H, seed=3,
                            with_mask=False)
    gb["pbr"][:, : W // 3, 1] = 0.01  # mirror
    gb["pbr"][:, 2 * W // 3:, 1] = 0.7  # DDGI
    sky = np.random.default_rng(8).uniform(0.0, 2.0, size=(6, 8, 8, 4)).astype(np.float32)
    ds = DeviceScene(g, sdf, mip, atlas, layout=layout, atlas_layout=atlas_layout, sky_faces=sky, sky_size=8)
    u = _scenes.small_uniform(rays=64, counts=(5, 4, 6))
    vol = DDGIVolume(u)
    rng = np.random.default_rng(3)
    irr = rng.uniform(0.2, 0.9, (u.irradianceTextureHeight, u.irradianceTextureWidth, 4)).astype(np.float32)
    dep = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    dep[..., 0] = rng.uniform(0.05, 0.6, dep.shape[:2])
    dep[..., 1] = dep[..., 0] ** 2 + rng.uniform(0.0, 0.02, dep.shape[:2])
    vol.irradiance[vol.ping_pong].copy_(torch.from_numpy(irr))
    vol.depth[vol.ping_pong].copy_(torch.from_numpy(dep))
    sobol, scr = _noise(2)
    nz = BlueNoiseTables(sobol, scr)
    p = abi.SdfReflectionParams()
    p.trim, p.roughDDGIIntensity, p.approximateWithDDGI, p.numFrames = 0.9, 1.2, 1, 17
    p.cameraPosition[:] = [eye[0], eye[1], eye[2], 1.0]
    p.viewProjInv = abi.mat4_colmajor(VPI)
    out = torch.full((H, W, 4), -5.0, dtype=torch.float32, device="cuda")
    sdf_reflection(p, nz, gb, ds, out, volume=vol)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    host = {k: v.cpu().numpy() for k, v in gb.items()}
    want = oracle.sdf_reflection(p, sobol, scr, W, H, host["normals"], host["pbr"], host["depth"], g, sdf, mip, 64, 16,
                                 atlas, sky_faces=sky, sky_size=8, u=u, irradiance=irr, depth_atlas=dep,
                                 out=np.full((H, W, 4), -5.0))
    sky_px = host["depth"] == 1.0
    assert np.all(got[sky_px] == -5.0)
    ok = np.all(np.abs(got - want) <= 2e-3 * (1 + np.abs(want)), axis=-1)[~sky_px]
    assert ok.mean() >= 0.99, ok.mean()
    hit = (want[..., 3] > 0) & (want[..., 3] != 1.0)
    assert hit[~sky_px].mean() > 0.2  # rays that read the atlas

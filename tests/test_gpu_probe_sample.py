"""GPU parity: per-pixel probe sampling (luxgi_ddgi_sample_probes, SampleProbe.comp:36-60 +
sampleIrradiance DDGICommon.glsl:163-233) against the FP64 oracle on the same atlases and G-buffer.

Tolerance. The world position is rebuilt from NDC depth in fp32 (Math.glsl:35-42), which is
ill-conditioned near the far plane; the oracle's own FP32 instantiation differs from FP64 by
<= 2.5e-3 relative on these inputs (99.9% of channels within 1e-3). So: >= 99.5% of channels
agree to 1e-3 relative + 1e-5 absolute, every channel to 1e-2 relative + 1e-4 absolute, and the
mean relative error is <= 2e-4. Sky pixels are exactly 0 and alpha is exactly 1 elsewhere."""
import numpy as np
import pytest

import oracle
from luxgi_amd import abi
from tests import _scenes, _svgf_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _atlases(u, seed):
    rng = np.random.default_rng(seed)
    irr = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    irr[..., :3] = rng.uniform(0.05, 1.0, irr[..., :3].shape)
    dep = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    dep[..., 0] = rng.uniform(0.05, 0.5, dep.shape[:2])
    dep[..., 1] = dep[..., 0] ** 2 + rng.uniform(0, 0.02, dep.shape[:2])
    return irr, dep


@pytest.mark.parametrize("counts,W,H,seed,storage", [((8, 8, 8), 96, 64, 3, abi.STORAGE_FP32),
                                                     ((5, 4, 6), 77, 45, 4, abi.STORAGE_FP32),
                                                     ((8, 8, 8), 64, 48, 5, abi.STORAGE_FP16)])
def test_sample_probes_matches_oracle(counts, W, H, seed, storage):
    from luxgi_amd.ddgi import DDGIVolume
    u = _scenes.small_uniform(rays=64, counts=counts)
    irr, dep = _atlases(u, seed)
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, seed)
    vol = DDGIVolume(u, atlas_storage=storage)
    tdt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    vol.irradiance[vol.ping_pong].copy_(torch.from_numpy(irr).to(tdt))
    vol.depth[vol.ping_pong].copy_(torch.from_numpy(dep).to(tdt))
    if storage == abi.STORAGE_FP16:  # the oracle reads what the GPU reads
        irr = irr.astype(np.float16).astype(np.float32)
        dep = dep.astype(np.float16).astype(np.float32)
    depth = torch.from_numpy(g["depth"]).cuda()
    normals = torch.from_numpy(g["normals"]).cuda()
    out = vol.sample_probes(depth, normals, cam["eye"], cam["VPI"])
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    p = abi.SampleProbeParams()
    p.viewProjInv = abi.mat4_colmajor(cam["VPI"])
    p.cameraPosition[:] = [*cam["eye"], 1.0]
    ref = oracle.sample_probes(u, p, W, H, g["depth"], g["normals"], irr, dep)
    sky = g["depth"] == 1.0
    assert sky.any() and (~sky).any()
    assert np.all(got[sky] == 0)
    assert np.all(got[~sky][:, 3] == 1.0)
    a, b = got[~sky][:, :3], ref[~sky][:, :3]
    err = np.abs(a - b)
    assert (err <= 1e-5 + 1e-3 * np.abs(b)).mean() >= 0.995
    assert np.all(err <= 1e-4 + 1e-2 * np.abs(b)), float((err / (np.abs(b) + 1e-6)).max())
    assert float((err / (np.abs(b) + 1e-6)).mean()) <= 2e-4


def test_sample_probes_after_update_fp16_output():
    """A full frame (trace + blend) followed by sampling into an RGBA16F target: the half output is
    the fp32 output rounded to half."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=64, counts=(8, 8, 8))
    vol = DDGIVolume(u)
    vol.update(DeviceScene(g, sdf, mip, atlas), _scenes.trace_params(rot_seed=9))
    W, H = 80, 60
    gb, *_rest, cam = _svgf_inputs.make_pair(W, H, 6)
    depth = torch.from_numpy(gb["depth"]).cuda()
    normals = torch.from_numpy(gb["normals"]).cuda()
    o32 = vol.sample_probes(depth, normals, cam["eye"], cam["VPI"]).cpu().numpy()
    o16 = vol.sample_probes(depth, normals, cam["eye"], cam["VPI"], out_storage=abi.STORAGE_FP16).cpu().numpy()
    assert np.isfinite(o32).all() and o32[..., :3].max() > 0
    np.testing.assert_array_equal(o16.astype(np.float32), o32.astype(np.float16).astype(np.float32))


@pytest.mark.parametrize("storage", [abi.STORAGE_FP32, abi.STORAGE_FP16])
def test_sample_probes_relocated_and_classified_matches_oracle(storage):
    """Probe offsets (up to 0.4 of the spacing) and 30% inactive probes in the atlas descriptor:
    the GPU gather reads neighbours at grid + offset and skips inactive ones, as the oracle does
    (tests/test_oracle_probe_sample.py pins that rule). Same tolerances as above; pixels whose 8
    neighbours are all inactive are exactly 0 on both sides."""
    from luxgi_amd.ddgi import DDGIVolume
    u = _scenes.small_uniform(rays=64, counts=(6, 5, 7))
    irr, dep = _atlases(u, 8)
    W, H = 96, 64
    g, prev, noisy, hist, hmom, mask, cam = _svgf_inputs.make_pair(W, H, 8)
    rng = np.random.default_rng(8)
    off = np.zeros((u.num_probes, 4), np.float32)
    off[:, :3] = rng.uniform(-0.4, 0.4, (u.num_probes, 3)) * np.array([u.step[0], u.step[1], u.step[2]])
    states = (rng.uniform(size=u.num_probes) < 0.3).astype(np.uint8)
    vol = DDGIVolume(u, atlas_storage=storage)
    tdt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
    vol.irradiance[vol.ping_pong].copy_(torch.from_numpy(irr).to(tdt))
    vol.depth[vol.ping_pong].copy_(torch.from_numpy(dep).to(tdt))
    vol.offsets.copy_(torch.from_numpy(off))
    vol.states.copy_(torch.from_numpy(states))
    vol.enable_relocation(True)
    if storage == abi.STORAGE_FP16:
        irr = irr.astype(np.float16).astype(np.float32)
        dep = dep.astype(np.float16).astype(np.float32)
    out = vol.sample_probes(torch.from_numpy(g["depth"]).cuda(), torch.from_numpy(g["normals"]).cuda(), cam["eye"],
                            cam["VPI"])
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    p = abi.SampleProbeParams()
    p.viewProjInv = abi.mat4_colmajor(cam["VPI"])
    p.cameraPosition[:] = [*cam["eye"], 1.0]
    ref = oracle.sample_probes(u, p, W, H, g["depth"], g["normals"], irr, dep, offsets=off, states=states)
    base = oracle.sample_probes(u, p, W, H, g["depth"], g["normals"], irr, dep)
    sky = g["depth"] == 1.0
    a, b = got[~sky][:, :3], ref[~sky][:, :3]
    assert np.abs(b - base[~sky][:, :3]).max() > 1e-2          # the relocation changes the answer
    zero = np.all(b == 0, axis=1)
    assert np.all(a[zero] == 0)
    err = np.abs(a - b)
    assert (err <= 1e-5 + 1e-3 * np.abs(b)).mean() >= 0.995
    # With 30% of the probes inactive a pixel can keep one or two small weights, and the
    # normalisation by their sum amplifies the ill-conditioned world position: the oracle's own
    # FP32 instantiation differs from FP64 by up to 8.2e-3 relative here (2.2e-3 without
    # relocation), so the per-channel bound is 5e-2 and the mean stays at 3e-4.
    assert np.all(err <= 1e-4 + 5e-2 * np.abs(b)), float((err / (np.abs(b) + 1e-6)).max())
    assert float((err / (np.abs(b) + 1e-6)).mean()) <= 3e-4
    vol.enable_relocation(False)
    out0 = vol.sample_probes(torch.from_numpy(g["depth"]).cuda(), torch.from_numpy(g["normals"]).cuda(),
                             cam["eye"], cam["VPI"]).cpu().numpy()
    e0 = np.abs(out0[~sky][:, :3] - base[~sky][:, :3])
    assert (e0 <= 1e-5 + 1e-3 * np.abs(base[~sky][:, :3])).mean() >= 0.995

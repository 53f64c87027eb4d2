"""GPU parity: probe relocation + classification (luxgi_ddgi_relocate_classify) against
the oracle on the same ray buffers. States are compared bit for bit (the rule compares
fp32 ray values exactly as the oracle does); offsets to 1e-6 absolute.
The reference has no relocation or classification: the rule is this library's own
(Majercik et al., JCGT 10(2) 2021, restated for an SDF tracer)."""
import numpy as np
import pytest

import oracle
from luxgi_amd import abi
from tests import _scenes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("relocate", [0, 1])
def test_relocate_classify_matches_oracle(relocate):
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=96, counts=(8, 8, 8))
    ds = DeviceScene(g, sdf, mip, atlas)
    vol = DDGIVolume(u, flags_buffer=True)
    vol.trace_rays(ds, _scenes.trace_params(rot_seed=4))
    rng = np.random.default_rng(3)
    off0 = np.zeros((u.num_probes, 4), np.float32)
    off0[:, :3] = rng.uniform(-0.02, 0.02, (u.num_probes, 3))
    vol.offsets.copy_(torch.from_numpy(off0))
    prm = abi.RelocateParams(0.25, 0.1, 0.45, relocate)
    vol.relocate_classify(prm)
    torch.cuda.synchronize()
    dd = vol.dir_dist.cpu().numpy()
    fl = vol.ray_flags.cpu().numpy()
    ref_off, ref_states = oracle.relocate(u, dd, fl, prm, off0)
    states = vol.states.cpu().numpy()
    assert np.array_equal(states, ref_states)
    assert 0 < states.sum() < u.num_probes  # the room has probes inside walls and outside it
    np.testing.assert_allclose(vol.offsets.cpu().numpy(), ref_off, atol=1e-6)
    if relocate:
        moved = np.any(np.abs(ref_off - off0) > 0, axis=1)
        assert moved.any()
        lim = 0.45 * np.array([u.step[0], u.step[1], u.step[2]], np.float32)
        assert np.all(np.abs(ref_off[:, :3]) <= lim + 1e-7)
    else:
        assert np.array_equal(vol.offsets.cpu().numpy(), off0)


def test_relocated_partially_inactive_volume_traced_blended_and_sampled():
    """The relocation loop end to end against the oracle at each stage, on the room scene:
    frame 1 traces from the grid and classifies/relocates; frame 2 traces from the moved probes
    (the offsets reach the ray origins), blends, and the consumers sample the atlas with the
    offsets and states (inactive probes take no weight). Each stage is compared with the oracle
    fed the GPU's previous-stage buffers: trace at the trace tolerances (class >= 99%, distance
    and radiance >= 98%), states bit-exact, offsets 1e-6, blend 2e-5, sampling as the sampling
    tests (>= 99.5% of channels within 1e-3)."""
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene
    from tests import _svgf_inputs
    prims, sdf, mip, g, atlas = _scenes.small_scene()
    u = _scenes.small_uniform(rays=96, counts=(8, 8, 8))
    ds = DeviceScene(g, sdf, mip, atlas)
    vol = DDGIVolume(u, flags_buffer=True)
    p1 = _scenes.trace_params(rot_seed=4)
    vol.trace_rays(ds, p1)
    prm = abi.RelocateParams(0.25, 0.1, 0.45, 1)
    vol.relocate_classify(prm)
    torch.cuda.synchronize()
    off, states = vol.offsets.cpu().numpy(), vol.states.cpu().numpy()
    ref_off, ref_states = oracle.relocate(u, vol.dir_dist.cpu().numpy(), vol.ray_flags.cpu().numpy(), prm,
                                          np.zeros((u.num_probes, 4), np.float32))
    assert np.array_equal(states, ref_states) and 0 < states.sum() < u.num_probes
    np.testing.assert_allclose(off, ref_off, atol=1e-6)
    assert np.any(np.abs(off[:, :3]) > 0)
    # frame 2 from the relocated probes
    vol.enable_relocation(True)
    p2 = _scenes.trace_params(rot_seed=6)
    vol.trace_rays(ds, p2)
    vol.probe_update(fuse_border=True)
    vol.end_frame()
    torch.cuda.synchronize()
    rad, dd = vol.radiance.cpu().numpy(), vol.dir_dist.cpu().numpy()
    orad, odd, ofl, _ = _scenes.oracle_trace(u, g, sdf, mip, atlas, p2, probe_offsets=off)
    cls = np.where(odd[..., 3] >= 59999.0, 0, 1)
    gcls = np.where(dd[..., 3] >= 59999.0, 0, 1)
    assert (cls == gcls).mean() >= 0.99
    close = np.abs(dd[..., 3] - odd[..., 3]) <= 1e-4 * (1 + np.abs(odd[..., 3]))
    assert close.mean() >= 0.98
    # the origins moved: the same rays from the grid give other distances
    grad, gdd, _, _ = _scenes.oracle_trace(u, g, sdf, mip, atlas, p2)
    moved = np.abs(off[:, :3]).max(axis=1) > 1e-3
    assert np.mean(np.abs(gdd[moved][..., 3] - odd[moved][..., 3]) > 1e-3) > 0.05
    z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    oi, od = oracle.probe_update(u, rad, dd, z_i, z_d, 1, 7)
    gi, gd = vol.current_irradiance.cpu().numpy(), vol.current_depth.cpu().numpy()
    np.testing.assert_allclose(gi, oi, rtol=2e-5, atol=1e-6)
    np.testing.assert_allclose(gd, od, rtol=2e-5, atol=1e-6)
    # consumers: per-pixel sampling with offsets and states
    W, H = 96, 64
    gb, *_rest, cam = _svgf_inputs.make_pair(W, H, 12)
    out = vol.sample_probes(torch.from_numpy(gb["depth"]).cuda(), torch.from_numpy(gb["normals"]).cuda(), cam["eye"],
                            cam["VPI"])
    torch.cuda.synchronize()
    got = out.cpu().numpy().astype(np.float64)
    sp = abi.SampleProbeParams()
    sp.viewProjInv = abi.mat4_colmajor(cam["VPI"])
    sp.cameraPosition[:] = [*cam["eye"], 1.0]
    ref = oracle.sample_probes(u, sp, W, H, gb["depth"], gb["normals"], gi, gd, offsets=off, states=states)
    sky = gb["depth"] == 1.0
    a, b = got[~sky][:, :3], ref[~sky][:, :3]
    err = np.abs(a - b)
    assert (err <= 1e-5 + 1e-3 * np.abs(b)).mean() >= 0.995
    assert np.all(a[np.all(b == 0, axis=1)] == 0)

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

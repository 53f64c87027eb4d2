"""GPU check of luxgi_ddgi_probe_update_noalloc (the probe update with its workspace in a static
device array) against luxgi_ddgi_probe_update with a caller workspace, through the C ABI.

Both run the same kernels on the same inputs, so the outputs must be identical bit for bit. Two
calls in a row, the second on a smaller volume, so the second call runs over the first call's
leftovers in the shared array; a third call with the reduced-precision flag covers the other
weight layout; a volume whose workspace exceeds the array is rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

import oracle
from luxgi_amd import abi, scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _inputs(counts, R, ni, nd, seed):
    u = abi.make_uniform(counts, (-1, -1, -1), (0.3, 0.4, 0.5), R, irr_oct=ni, depth_oct=nd, hysteresis=0.97)
    P = u.num_probes
    dirs = oracle.ray_directions(abi.mat4_colmajor(scene.random_rotation(seed)), R).astype(np.float32)
    rng = np.random.default_rng(seed)
    rad = np.zeros((P, R, 4), np.float32)
    dd = np.zeros((P, R, 4), np.float32)
    dd[:, :, :3] = dirs[None]
    for p in range(0, P, 5):  # some probes with their own directions: the per-probe path runs too
        d = dirs + rng.normal(scale=0.05, size=dirs.shape).astype(np.float32)
        dd[p, :, :3] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rad[:, :, :3] = rng.uniform(0, 2, (P, R, 3))
    dd[:, :, 3] = rng.uniform(0.05, 3.0, (P, R))
    dd[:, ::13, 3] = 60000.0
    prev_i = rng.uniform(0, 1, (u.irradianceTextureHeight, u.irradianceTextureWidth, 4)).astype(np.float32)
    prev_d = rng.uniform(0, 1, (u.depthTextureHeight, u.depthTextureWidth, 2)).astype(np.float32)
    return u, rad, dd, prev_i, prev_d


def _both(u, rad, dd, prev_i, prev_d, flags):
    from luxgi_amd import _lib
    from luxgi_amd.ddgi import ptr, stream_handle
    L = _lib.lib()
    t_rad = torch.from_numpy(rad).cuda()
    t_dd = torch.from_numpy(dd).cuda()
    rays = abi.RayBuffers(ptr(t_rad), ptr(t_dd), rad.shape[1], rad.shape[0], abi.STORAGE_FP32, 0)
    p_i = torch.from_numpy(prev_i).cuda()
    p_d = torch.from_numpy(prev_d).cuda()
    prev = abi.ProbeAtlas(ptr(p_i), ptr(p_d), abi.STORAGE_FP32, 0)
    outs = []
    for noalloc in (False, True):
        o_i = torch.full_like(p_i, -7.0)
        o_d = torch.full_like(p_d, -7.0)
        out = abi.ProbeAtlas(ptr(o_i), ptr(o_d), abi.STORAGE_FP32, 0)
        if noalloc:
            rc = L.luxgi_ddgi_probe_update_noalloc(C.byref(u), C.byref(rays), C.byref(prev), C.byref(out), 0, flags,
                                                   stream_handle())
        else:
            ws = torch.empty(L.luxgi_ddgi_probe_update_workspace_size(C.byref(u)), dtype=torch.uint8, device="cuda")
            rc = L.luxgi_ddgi_probe_update(C.byref(u), C.byref(rays), C.byref(prev), C.byref(out), 0, flags,
                                           ptr(ws), ws.numel(), stream_handle())
        assert rc == 0
        torch.cuda.synchronize()
        outs.append((o_i.cpu().numpy(), o_d.cpu().numpy()))
    return outs


def test_noalloc_update_equals_workspace_update_on_consecutive_calls():
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH | abi.UPDATE_FUSE_BORDER
    for counts, R, ni, nd, fl, seed in [((8, 4, 8), 256, 8, 16, flags, 1),
                                        ((3, 2, 2), 96, 6, 14, flags, 2),
                                        ((4, 4, 2), 128, 8, 16, flags | abi.UPDATE_HALF, 3)]:
        print(f"noalloc update: counts={counts} rays={R} oct={ni}/{nd} flags={fl:#x}", flush=True)
        (ri, rd), (ni_, nd_) = _both(*_inputs(counts, R, ni, nd, seed), fl)
        np.testing.assert_array_equal(ni_.view(np.uint32), ri.view(np.uint32))
        np.testing.assert_array_equal(nd_.view(np.uint32), rd.view(np.uint32))
        assert np.all(ri[1:-1, 1:-1] != -7.0) and np.all(rd[1:-1, 1:-1] != -7.0)  # interiors and borders written


def test_noalloc_update_rejects_a_volume_beyond_the_scratch_limit():
    from luxgi_amd import _lib
    L = _lib.lib()
    u = abi.make_uniform((64, 64, 64), (-1, -1, -1), (0.1, 0.1, 0.1), 1024, irr_oct=16, depth_oct=16)
    assert L.luxgi_ddgi_probe_update_workspace_size(C.byref(u)) > (4 << 20)
    dummy = abi.RayBuffers(None, None, 1024, u.num_probes, abi.STORAGE_FP32, 0)
    atl = abi.ProbeAtlas(None, None, abi.STORAGE_FP32, 0)
    rc = L.luxgi_ddgi_probe_update_noalloc(C.byref(u), C.byref(dummy), C.byref(atl), C.byref(atl), 1,
                                           abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH, None)
    assert rc == -4  # LUXGI_ERR_WORKSPACE_TOO_SMALL, before any launch

"""Host side of the batched reflection denoise (CPU): luxgi_svgf_batch_prepare's unit offsets,
header and argument checks (it makes no GPU call), and the descriptor layout."""
import ctypes as C

import pytest

from luxgi_amd import abi


def _items(sizes):
    items = (abi.ReflectionDenoiseItem * len(sizes))()
    fake = 0x1000  # never dereferenced by the host-side preparation
    for k, (w, h) in enumerate(sizes):
        it = items[k]
        it.width, it.height = w, h
        it.filter = abi.ReflectionAtrousParams(1, 0, 10.0, 32.0, 1.0, 1)
        it.cur = abi.GBuffer(fake, fake, fake, fake)
        it.prev = abi.GBuffer(fake, fake, fake, fake)
        it.noisy = fake
        it.tileMap = fake
        for j in range(2):
            it.outputs[j] = it.moments[j] = it.atrous[j] = fake
    return items


def test_prepare_sets_unit_offsets_and_header():
    from luxgi_amd._lib import lib
    sizes = [(77, 45), (8, 8), (1920, 1080), (1, 1), (17, 16)]
    items = _items(sizes)
    hdr = abi.DenoiseBatchHeader()
    assert lib().luxgi_svgf_batch_prepare(C.byref(items), len(sizes), C.byref(hdr)) == 0
    units = [((w + 15) // 16) * ((h + 15) // 16) for w, h in sizes]
    start = 0
    for k, n in enumerate(units):
        assert items[k].blockStart == start
        start += n
    assert hdr.count == len(sizes) and hdr.totalBlocks == sum(units) == 15 + 1 + 8160 + 1 + 2


@pytest.mark.parametrize("field", ["width", "radius", "noisy", "atrous1", "cur.depth", "count"])
def test_prepare_rejects_bad_items(field):
    from luxgi_amd._lib import lib
    items = _items([(64, 32), (16, 16)])
    n = 2
    if field == "width":
        items[1].width = 0
    elif field == "radius":
        items[0].filter.radius = 3
    elif field == "noisy":
        items[1].noisy = None
    elif field == "atrous1":
        items[0].atrous[1] = None
    elif field == "cur.depth":
        items[1].cur.depth = None
    else:
        n = abi.MAX_DENOISE_BATCH + 1
    hdr = abi.DenoiseBatchHeader()
    assert lib().luxgi_svgf_batch_prepare(C.byref(items), n, C.byref(hdr)) < 0


def test_descriptor_layout_matches_the_header():
    assert C.sizeof(abi.ReflectionDenoiseItem) == 336
    assert abi.ReflectionDenoiseItem.cur.offset == 208
    assert abi.ReflectionDenoiseItem.tileMap.offset == 328
    assert C.sizeof(abi.DenoiseBatchHeader) == 16


def _trace_volumes(shapes, storage=abi.STORAGE_FP32):
    vols = (abi.TraceVolume * len(shapes))()
    for k, (counts, rays) in enumerate(shapes):
        u = abi.make_uniform(counts, (-1.0, -1.0, -1.0), [0.5, 0.5, 0.5], rays)
        vols[k].uniform = u
        P = counts[0] * counts[1] * counts[2]
        vols[k].rays = abi.RayBuffers(0x1000, 0x2000, rays, P, storage, 0)  # never dereferenced here
    return vols


def test_trace_batch_prepare_sets_ray_starts():
    from luxgi_amd._lib import lib
    shapes = [((5, 4, 6), 64), ((3, 3, 3), 96), ((6, 2, 4), 128)]
    vols = _trace_volumes(shapes)
    total = C.c_int64(-1)
    assert lib().luxgi_ddgi_trace_batch_prepare(C.byref(vols), 3, abi.STORAGE_FP32, C.byref(total)) == 0
    starts = [0, 120 * 64, 120 * 64 + 27 * 96]
    assert [vols[k].rayStart for k in range(3)] == starts
    assert total.value == starts[-1] + 48 * 128
    assert lib().luxgi_ddgi_trace_batch_workspace_size(total.value, 10) > 0
    assert lib().luxgi_ddgi_trace_batch_workspace_size(0, 10) == 0


@pytest.mark.parametrize("case", ["storage", "width", "height", "null", "count", "counts"])
def test_trace_batch_prepare_rejects_bad_volumes(case):
    from luxgi_amd._lib import lib
    vols = _trace_volumes([((4, 4, 4), 64), ((2, 2, 2), 32)])
    n, storage = 2, abi.STORAGE_FP32
    if case == "storage":
        vols[1].rays.storage = abi.STORAGE_FP16
    elif case == "width":
        vols[0].rays.width = 63
    elif case == "height":
        vols[1].rays.height = 7
    elif case == "null":
        vols[0].rays.radiance = None
    elif case == "count":
        n = abi.MAX_TRACE_BATCH + 1
    else:
        vols[1].uniform.probeCounts[2] = 0
    total = C.c_int64(0)
    assert lib().luxgi_ddgi_trace_batch_prepare(C.byref(vols), n, storage, C.byref(total)) < 0


def test_batched_entry_points_reject_bad_arguments_before_any_launch():
    """Argument checks that return before touching the GPU (null descriptors, grid, order, counts)."""
    from luxgi_amd._lib import lib
    L = lib()
    res = C.c_int32(0)
    fake = C.c_void_p(0x1000)
    assert L.luxgi_svgf_reflection_denoise_batched(None, fake, 4, abi.STORAGE_FP32, 64, 0, C.byref(res), None) < 0
    assert L.luxgi_svgf_reflection_denoise_batched(fake, fake, 4, abi.STORAGE_FP32, 0, 0, C.byref(res), None) < 0
    assert L.luxgi_svgf_reflection_denoise_batched(fake, fake, 4, abi.STORAGE_FP32, 64, 3, C.byref(res), None) < 0
    assert L.luxgi_svgf_reflection_denoise_batched(fake, fake, 17, abi.STORAGE_FP32, 64, 0, C.byref(res), None) < 0
    p = abi.TraceParams()
    assert L.luxgi_ddgi_trace_sdf_batched(None, 1, 100, abi.STORAGE_FP32, None, None, None, None, C.byref(p), None, 0,
                                          None) < 0
    assert L.luxgi_ddgi_trace_sdf_batched(fake, 0, 100, abi.STORAGE_FP32, None, None, None, None, C.byref(p), None, 0,
                                          None) < 0
    assert L.luxgi_ddgi_trace_sdf_batched(fake, 1, 0, abi.STORAGE_FP32, None, None, None, None, C.byref(p), None, 0,
                                          None) < 0
    assert L.luxgi_ddgi_trace_sdf_batched(fake, 1, 100, abi.STORAGE_FP32, None, None, None, None, C.byref(p), None, 0,
                                          None) < 0  # no scene

"""CPU checks of the C ABI: libluxgi.so loads and exports every function that
include/luxgi.h and include/luxgi_scene.h declare; struct sizes match the
reference layouts. No compute calls (no GPU here)."""
import ctypes as C
import os
import re

import pytest

from luxgi_amd import _lib, abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in ("luxgi.h", "luxgi_scene.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(luxgi_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declarations_have_python_signatures():
    names = declared_functions()
    assert len(names) >= 18
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libluxgi.so not built")
def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    for n in declared_functions():
        assert hasattr(L, n), n
    assert b"gfx950" in _lib.lib().luxgi_version()


def test_struct_sizes():
    assert C.sizeof(abi.DDGIUniform) == 96
    assert C.sizeof(abi.GlobalSDFData) == 96
    assert C.sizeof(abi.AtlasObject) == 128
    assert C.sizeof(abi.AtlasTile) == 96
    assert C.sizeof(abi.SurfaceAtlasData) == 32


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libluxgi.so not built")
def test_host_only_entry_points():
    u = abi.make_uniform((32, 16, 32), (-1, -1, -1), (0.1, 0.1, 0.1), 256)
    u.irradianceTextureWidth = 0
    assert _lib.lib().luxgi_ddgi_fill_atlas_sizes(C.byref(u)) == 0
    assert (u.irradianceTextureWidth, u.irradianceTextureHeight, u.depthTextureWidth, u.depthTextureHeight) == \
        (5122, 322, 9218, 578)
    assert _lib.lib().luxgi_svgf_tile_count(1920, 1080) == 240 * 135
    assert _lib.lib().luxgi_ddgi_probe_update_workspace_size(C.byref(u)) >= 256 * (64 + 256) * 4
    # invalid arguments are rejected before any device work
    assert _lib.lib().luxgi_ddgi_probe_update(None, None, None, None, 0, 3, None, 0, None) < 0
    bad = abi.make_uniform((2, 2, 2), (0, 0, 0), (1, 1, 1), 64)
    bad.depthTextureWidth += 1
    rays = abi.RayBuffers(None, None, 64, 8, 0, 0)
    assert _lib.lib().luxgi_ddgi_probe_update(C.byref(bad), C.byref(rays), None, None, 1, 3, None, 0, None) == -1


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libluxgi.so not built")
def test_brick_layout_sizes_and_rejections():
    L = _lib.lib()
    assert L.luxgi_sdf_bricks_bytes(256, 256, 256, abi.STORAGE_FP32) == 64 ** 3 * 128 * 4
    assert L.luxgi_sdf_bricks_bytes(512, 256, 256, abi.STORAGE_FP16) == 128 * 64 * 64 * 128 * 2
    assert L.luxgi_sdf_bricks_bytes(30, 32, 32, abi.STORAGE_FP32) == 0
    assert L.luxgi_sdf_bricks_bytes(32, 32, 32, abi.STORAGE_UNORM8) == 0
    # conversions and packs reject bad shapes before any device work
    assert L.luxgi_sdf_to_bricks(None, 32, 32, 32, abi.STORAGE_FP32, None, None, None) < 0
    assert L.luxgi_atlas_pack(None, None, 2048, None, None) < 0
    dummy = C.c_void_p(16)
    assert L.luxgi_atlas_pack(dummy, dummy, 2046, dummy, None) < 0


def _declarations():
    """name -> (return type, [parameter types]) of every function the headers declare."""
    out = {}
    for h in ("luxgi.h", "luxgi_scene.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#[^\n]*", "", text)
        for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(luxgi_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
            ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
            plist = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
            types = []
            for p in plist:
                p = re.sub(r"\[[^\]]*\]", "*", p)  # float rot[16] -> pointer
                p = re.sub(r"\b[A-Za-z_]\w*$", "", p).strip() if not p.endswith("*") else p
                types.append(p)
            out[name] = (ret.split()[-1] if ret else ret, types)
    return out


def _ctype_of(c_decl):
    """The ctypes type a C parameter/return type must bind to."""
    if "*" in c_decl:
        return "ptr"
    base = c_decl.replace("const", "").strip()
    return {"int": C.c_int, "int32_t": C.c_int32, "uint32_t": C.c_uint32, "size_t": C.c_size_t,
            "int64_t": C.c_int64, "float": C.c_float, "uint64_t": C.c_uint64}[base]


def test_ctypes_argument_lists_match_the_header():
    decls = _declarations()
    assert set(decls) == set(_lib.SIGNATURES), set(decls) ^ set(_lib.SIGNATURES)
    for name, (ret, params) in decls.items():
        res, args = _lib.SIGNATURES[name]
        assert len(args) == len(params), (name, len(args), len(params), params)
        for i, (c_t, py_t) in enumerate(zip(params, args)):
            want = _ctype_of(c_t)
            if want == "ptr":
                assert py_t in (C.c_void_p, C.c_char_p), (name, i, c_t, py_t)
            else:
                assert C.sizeof(py_t) == C.sizeof(want) and (py_t in (C.c_float, C.c_double)) == \
                    (want in (C.c_float, C.c_double)), (name, i, c_t, py_t)
        if ret.endswith("*"):
            assert res in (C.c_char_p, C.c_void_p), name
        else:
            assert C.sizeof(res) == C.sizeof(_ctype_of(ret)), (name, ret, res)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libluxgi.so not built")
def test_bf16_blend_needs_the_reduced_form():
    """LUXGI_UPDATE_BF16 without LUXGI_UPDATE_HALF is rejected (-1) before the workspace check (-4),
    so before any launch."""
    L = _lib.lib()
    u = abi.make_uniform((2, 2, 2), (-1, -1, -1), (0.3, 0.4, 0.5), 64, max_distance=1.5)
    dummy = C.c_void_p(256)
    rays = abi.RayBuffers(dummy, dummy, 64, u.num_probes, abi.STORAGE_FP32, 0)
    atl = abi.ProbeAtlas(dummy, dummy, abi.STORAGE_FP32, 0)
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH
    assert L.luxgi_ddgi_probe_update(C.byref(u), C.byref(rays), C.byref(atl), C.byref(atl), 1,
                                     flags | abi.UPDATE_BF16, None, 0, None) == -1
    assert L.luxgi_ddgi_probe_update(C.byref(u), C.byref(rays), C.byref(atl), C.byref(atl), 1,
                                     flags | abi.UPDATE_BF16 | abi.UPDATE_HALF, None, 0, None) == -4

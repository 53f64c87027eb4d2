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


def test_batched_update_prepare_checks_every_volume_on_the_host():
    """luxgi_ddgi_update_batch_prepare is host-only: it rejects what luxgi_ddgi_probe_update rejects,
    volumes outside the shared-direction MFMA path, mixed storages, forced per-probe weights and
    counts out of range, before any device work."""
    import ctypes as C
    from luxgi_amd import _lib
    from tests import _scenes
    lib = _lib.lib()
    k = 3
    vols = (abi.UpdateVolume * k)()
    fake = 1 << 40  # device addresses are not dereferenced by prepare
    for i in range(k):
        vols[i].uniform = _scenes.small_uniform(rays=64, counts=(3, 2, 2))
        vols[i].rays = abi.RayBuffers(fake, fake, 64, 12, abi.STORAGE_FP32, 0)
        vols[i].out = abi.ProbeAtlas(fake, fake, abi.STORAGE_FP32, 0)
        vols[i].firstFrame = 1
    flags = abi.UPDATE_IRRADIANCE | abi.UPDATE_DEPTH
    ws = lib.luxgi_ddgi_update_batch_workspace_size(C.byref(vols), k)
    host = (C.c_uint8 * lib.luxgi_ddgi_update_batch_descriptor_bytes(k))()
    b = abi.UpdateBatch()

    def prep(f=flags, count=k, wsb=ws):
        return lib.luxgi_ddgi_update_batch_prepare(C.byref(vols), count, f, C.c_void_p(fake), wsb, C.byref(host),
                                                   C.byref(b))
    assert prep() == 0 and b.count == k and b.groups == 3 and b.maxRays == 64
    assert prep(wsb=ws - 1) != 0                                       # workspace too small
    assert prep(count=0) != 0 and prep(count=abi.MAX_UPDATE_BATCH + 1) != 0
    assert lib.luxgi_ddgi_update_batch_descriptor_bytes(0) == 0
    assert prep(f=flags | abi.UPDATE_PER_PROBE) != 0                  # not the MFMA path
    assert prep(f=flags | abi.UPDATE_BF16) != 0                       # BF16 needs HALF
    assert prep(f=0) != 0                                             # nothing to do
    vols[1].rays.storage = abi.STORAGE_FP16
    assert prep() != 0                                                # mixed ray storage
    vols[1].rays.storage = abi.STORAGE_FP32
    vols[2].firstFrame = 0
    assert prep() != 0                                                # no previous atlas
    vols[2].prev = abi.ProbeAtlas(fake, fake, abi.STORAGE_FP32, 0)
    assert prep() == 0
    big = _scenes.small_uniform(rays=64, counts=(3, 2, 2), ni=10, nd=14)  # 100 texels > 64
    vols[0].uniform = big
    assert prep() != 0


def _struct_layout_tu():
    """A translation unit that includes the header and asserts, for every ctypes mirror of a header
    struct, the C size and the C offset of each field as ctypes lays them out."""
    import shutil
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    header = open(os.path.join(ROOT, "include", "luxgi.h")).read()
    lines = ['#include <stddef.h>', '#include "luxgi.h"',
             '#ifdef __cplusplus', '#define LUXGI_SA static_assert', '#else', '#define LUXGI_SA _Static_assert',
             '#endif']
    checked = 0
    for name in dir(abi):
        cls = getattr(abi, name)
        if not (isinstance(cls, type) and issubclass(cls, C.Structure)):
            continue
        cname = "Luxgi" + name
        if not re.search(r"typedef struct %s\b" % cname, header):
            continue
        lines.append(f'LUXGI_SA(sizeof({cname}) == {C.sizeof(cls)}, "sizeof {cname}");')
        for f, _ in cls._fields_:
            lines.append(f'LUXGI_SA(offsetof({cname}, {f}) == {getattr(cls, f).offset}, "{cname}.{f}");')
        checked += 1
    return "\n".join(lines) + "\n", checked


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_header_struct_layouts_match_the_ctypes_mirrors_in_c_and_cpp(lang, tmp_path):
    """Every struct of include/luxgi.h that luxgi_amd/abi.py mirrors has the same size and the same
    field offsets in C (gcc -std=c11) and C++ (g++ -std=c++17), and the header compiles warning-free
    in both."""
    import subprocess
    tu, checked = _struct_layout_tu()
    assert checked >= 30
    src = tmp_path / ("layout.c" if lang == "c" else "layout.cpp")
    src.write_text(tu)
    cmd = (["gcc", "-std=c11"] if lang == "c" else ["g++", "-std=c++17"]) + \
        ["-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def _header_param_counts():
    counts = {}
    for h in ("luxgi.h", "luxgi_scene.h"):
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", h)).read(), flags=re.S)
        for m in re.finditer(r"\b(luxgi_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
            params = m.group(2).strip()
            counts[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return counts


def test_integration_snippets_call_declared_functions_with_their_arity():
    """Every luxgi_* call in INTEGRATION.md's code blocks names a function of the headers and passes
    as many arguments as the header declares."""
    counts = _header_param_counts()
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```(?:cpp|c\+\+|c)\n(.*?)```", doc, flags=re.S)
    assert blocks
    seen = 0
    for b in blocks:
        code = re.sub(r"//[^\n]*", "", b)
        for m in re.finditer(r"\b(luxgi_[a-z0-9_]+)\s*\(", code):
            name = m.group(1)
            assert name in counts, name
            depth, i, args, cur = 1, m.end(), 0, ""
            while depth:
                ch = code[i]
                if ch in "([{":
                    depth += 1
                elif ch in ")]}":
                    depth -= 1
                elif ch == "," and depth == 1:
                    args += 1
                if depth:
                    cur += ch
                i += 1
            n = 0 if not cur.strip() else args + 1
            assert n == counts[name], (name, n, counts[name])
            seen += 1
    assert seen >= 30

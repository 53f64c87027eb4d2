"""Loader for libluxgi.so (the HIP kernels behind include/luxgi.h).

The library is required: every entry point raises if libluxgi.so is missing
or fails to load. There is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libluxgi.so")
_lib = None

P = C.c_void_p
I = C.c_int32
U = C.c_uint32

# name -> (restype, argtypes); mirrors include/luxgi.h and include/luxgi_scene.h
SIGNATURES = {
    "luxgi_version": (C.c_char_p, []),
    "luxgi_ddgi_fill_atlas_sizes": (C.c_int, [P]),
    "luxgi_ddgi_trace_workspace_size": (C.c_size_t, [P, I]),
    "luxgi_ddgi_trace_sdf": (C.c_int, [P, P, P, P, P, P, P, P, P, P, C.c_size_t, P]),
    "luxgi_ddgi_trace_sdf_count": (C.c_int, [P, P, P, P, P, P, P, P, P, P, C.c_size_t, P, P]),
    "luxgi_ddgi_probe_update_workspace_size": (C.c_size_t, [P]),
    "luxgi_ddgi_probe_update": (C.c_int, [P, P, P, P, U, U, P, C.c_size_t, P]),
    "luxgi_ddgi_probe_update_noalloc": (C.c_int, [P, P, P, P, U, U, P]),
    "luxgi_ddgi_update_batch_descriptor_bytes": (C.c_size_t, [I]),
    "luxgi_ddgi_update_batch_workspace_size": (C.c_size_t, [P, I]),
    "luxgi_ddgi_update_batch_prepare": (C.c_int, [P, I, U, P, C.c_size_t, P, P]),
    "luxgi_ddgi_probe_update_batched": (C.c_int, [P, P, P]),
    "luxgi_ddgi_border_update": (C.c_int, [P, P, U, P]),
    "luxgi_ddgi_relocate_classify": (C.c_int, [P, P, P, P, P, P, P]),
    "luxgi_ddgi_sample_probes": (C.c_int, [P, P, P, I, I, P, P, I, P, P]),
    "luxgi_sdf_build_mip": (C.c_int, [P, I, I, I, C.c_float, I, P, P, P]),
    "luxgi_sdf_rasterize": (C.c_int, [P, P, P, P, I, P, P, P, I, I, I, P, P]),
    "luxgi_sdf_bricks_bytes": (C.c_size_t, [I, I, I, I]),
    "luxgi_sdf_to_bricks": (C.c_int, [P, I, I, I, I, P, P, P]),
    "luxgi_sdf_build_guard": (C.c_int, [P, I, I, I, P, P]),
    "luxgi_sdf_cells_bytes": (C.c_size_t, [I, I, I, I]),
    "luxgi_sdf_to_cells": (C.c_int, [P, I, I, I, I, P, P]),
    "luxgi_atlas_pack": (C.c_int, [P, P, I, P, P]),
    "luxgi_atlas_cull_workspace_size": (C.c_size_t, []),
    "luxgi_atlas_cull_chunks": (C.c_int, [P, P, P, P, P, C.c_size_t, P]),
    "luxgi_surface_light": (C.c_int, [P, P, P, I, P, I, P, P, P, P, P, I, P, P]),
    "luxgi_sdf_shadow_mask": (C.c_int, [P, P, I, I, P, P, P, P, P, P]),
    "luxgi_sdf_reflection_workspace_size": (C.c_size_t, [I]),
    "luxgi_sdf_reflection": (C.c_int, [P, P, I, I, P, P, P, P, P, P, P, I, P, P, C.c_size_t, P]),
    "luxgi_svgf_tile_count": (I, [I, I]),
    "luxgi_svgf_reflection_reproject": (C.c_int, [P, I, I, I, P, P, P, P, P, P, P, P, P]),
    "luxgi_svgf_reflection_atrous": (C.c_int, [P, I, I, I, P, P, P, P, P, P]),
    "luxgi_svgf_reflection_denoise": (C.c_int, [P, P, I, I, I, I, P, P, P, P, P, P, P, I, P, P, P, P, P]),
    "luxgi_svgf_shadow_reproject": (C.c_int, [P, I, I, I, P, P, P, P, P, P, P, P, P]),
    "luxgi_svgf_shadow_atrous": (C.c_int, [P, I, I, I, P, P, P, P, P, P]),
    "luxgi_svgf_shadow_denoise": (C.c_int, [P, P, I, I, I, I, P, P, P, P, P, P, P, I, P, P, P, P, P]),
    "luxgi_ddgi_trace_batch_prepare": (C.c_int, [P, I, I, P]),
    "luxgi_ddgi_trace_batch_workspace_size": (C.c_size_t, [C.c_int64, I]),
    "luxgi_ddgi_trace_sdf_batched": (C.c_int, [P, I, C.c_int64, I, P, P, P, P, P, P, C.c_size_t, P]),
    "luxgi_ddgi_trace_scenes_descriptor_bytes": (C.c_size_t, [I]),
    "luxgi_ddgi_trace_scenes_prepare": (C.c_int, [P, I, P, I, I, P, P]),
    "luxgi_ddgi_trace_scenes_workspace_size": (C.c_size_t, [P]),
    "luxgi_ddgi_trace_sdf_scenes": (C.c_int, [P, P, P, P, P, C.c_size_t, P]),
    "luxgi_svgf_batch_prepare": (C.c_int, [P, I, P]),
    "luxgi_svgf_reflection_denoise_batched": (C.c_int, [P, P, I, I, I, I, P, P]),
    "luxgi_svgf_learned_atrous": (C.c_int, [P, P, I, I, P, P, P, P, P]),
    "luxgi_svgf_learned_atrous_backward": (C.c_int, [P, P, I, I, P, P, P, P, P, P, P, P]),
    "luxgi_adam_step": (C.c_int, [P, P, P, P, I, C.c_float, C.c_float, C.c_float, C.c_float, I, P]),
    "luxgi_svgf_reflection_atrous_scaled": (C.c_int, [P, I, I, I, P, P, P, P, P, P, P]),
    "luxgi_learned_pixel_scales": (C.c_int, [P, I, I, P, P, P, P, P, P]),
    "luxgi_learned_pixel_pass_backward_workspace_size": (C.c_size_t, [I, I]),
    "luxgi_learned_pixel_pass_backward": (C.c_int, [P, I, I, P, P, P, P, P, P, P, P, C.c_size_t, P]),
    "luxgi_learned_pixel_mlp_backward_workspace_size": (C.c_size_t, []),
    "luxgi_learned_pixel_mlp_backward": (C.c_int, [P, I, I, P, P, P, P, P, C.c_size_t, P]),
    "luxgi_scene_bake_sdf": (C.c_int, [P, I, I, P, C.c_float, P, P]),
    "luxgi_scene_gbuffer": (C.c_int, [P, I, P, P, I, P, P, I, I, U, P, P, P, P, P, P, P]),
}


def lib():
    """Load libluxgi.so; raises RuntimeError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libluxgi.so not found at {LIB_PATH}: build it with "
                               "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class LuxgiError(RuntimeError):
    pass


def check(rc, what):
    if rc != 0:
        raise LuxgiError(f"{what} returned status {rc}")
    return rc

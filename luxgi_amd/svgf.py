"""Host mirror of LuxGI's SVGF denoise systems over the C ABI (libluxgi.so).

ReflectionDenoiser mirrors Engine/Raytrace/RaytracedReflection.cpp:
  ReflectionPipeline outputs[2] / currentMoments[2] / pingPong, AtrousFiler atrousFilter[2]
  denoise::system (:496-561) = resetArgs + accumulation (:362-417) + atrousFilter (:419-494)
ShadowDenoiser mirrors Engine/Raytrace/RaytracedShadow.cpp:406-610 (scalar visibility,
`power` on the last iteration, copy tiles write 0).
Parameters default to the reference's (RaytracedReflection.h:27-34, RaytracedShadow.cpp:94-115).
"""
import ctypes as C

import torch

from . import abi
from ._lib import check, lib
from .ddgi import ptr, stream_handle


def gbuffer_struct(g):
    """g: dict of cuda tensors color [H][W][4], normals, pbr, depth [H][W]."""
    return abi.GBuffer(ptr(g["color"]), ptr(g["normals"]), ptr(g["pbr"]), ptr(g["depth"]))


class ReflectionDenoiser:
    def __init__(self, width, height, storage=abi.STORAGE_FP32, iterations=4, alpha=0.01, moments_alpha=0.2,
                 radius=1, phi_color=10.0, phi_normal=32.0, sigma_depth=1.0):
        self.W, self.H = width, height
        self.storage = storage
        dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
        z = lambda: torch.zeros((height, width, 4), dtype=dt, device="cuda")  # noqa: E731
        self.outputs = [z(), z()]
        self.moments = [z(), z()]
        self.atrous = [torch.ones((height, width, 4), dtype=dt, device="cuda") for _ in range(2)]
        self.tile_map = torch.zeros(lib().luxgi_svgf_tile_count(width, height), dtype=torch.uint8, device="cuda")
        self.ping_pong = 0
        self.iterations = iterations
        self.rp = abi.ReflectionReprojectParams()
        self.rp.alpha = alpha
        self.rp.momentsAlpha = moments_alpha
        self.rp.approximateWithDDGI = 1
        self.ap = abi.ReflectionAtrousParams(radius, 0, phi_color, phi_normal, sigma_depth, 1)
        self.result = None

    def set_camera(self, view_proj_inv, prev_view_proj, camera_pos, camera_delta):
        self.rp.viewProjInv = abi.mat4_colmajor(view_proj_inv)
        self.rp.prevViewProj = abi.mat4_colmajor(prev_view_proj)
        self.rp.cameraPos[:] = [float(camera_pos[0]), float(camera_pos[1]), float(camera_pos[2]), 1.0]
        self.rp.cameraDelta[:] = [float(v) for v in camera_delta]

    def denoise(self, cur, prev, noisy):
        """denoise::system for one frame; returns the output tensor."""
        res = C.c_int32(-1)
        check(lib().luxgi_svgf_reflection_denoise(
            C.byref(self.rp), C.byref(self.ap), self.iterations, self.W, self.H, self.storage, C.byref(gbuffer_struct(cur)),
            C.byref(gbuffer_struct(prev)), ptr(noisy), ptr(self.outputs[0]), ptr(self.outputs[1]),
            ptr(self.moments[0]), ptr(self.moments[1]), self.ping_pong, ptr(self.atrous[0]), ptr(self.atrous[1]),
            ptr(self.tile_map), C.byref(res), stream_handle()), "luxgi_svgf_reflection_denoise")
        self.ping_pong = 1 - self.ping_pong
        self.result = self.outputs[0] if res.value < 0 else self.atrous[res.value]
        return self.result


class ShadowDenoiser:
    def __init__(self, width, height, storage=abi.STORAGE_FP32, iterations=4, alpha=0.01, moments_alpha=0.2,
                 radius=1, phi_visibility=10.0, phi_normal=32.0, sigma_depth=1.0, power=1.2, near=0.1, far=10000.0):
        self.W, self.H = width, height
        self.storage = storage
        dt = torch.float16 if storage == abi.STORAGE_FP16 else torch.float32
        self.outputs = [torch.zeros((height, width, 2), dtype=dt, device="cuda") for _ in range(2)]
        self.moments = [torch.zeros((height, width, 4), dtype=dt, device="cuda") for _ in range(2)]
        self.atrous = [torch.ones((height, width, 2), dtype=dt, device="cuda") for _ in range(2)]
        self.tile_map = torch.zeros(lib().luxgi_svgf_tile_count(width, height), dtype=torch.uint8, device="cuda")
        self.ping_pong = 0
        self.iterations = iterations
        self.power = power
        self.rp = abi.ShadowReprojectParams()
        self.rp.alpha = alpha
        self.rp.momentsAlpha = moments_alpha
        self.ap = abi.ShadowAtrousParams(radius, 1, phi_visibility, phi_normal, sigma_depth, power, near, far)
        self.result = None

    def set_camera(self, view_proj_inv):
        self.rp.viewProjInv = abi.mat4_colmajor(view_proj_inv)

    def denoise(self, cur, prev, mask):
        """denoise::system (RaytracedShadow.cpp:543-609) for one frame; returns the output tensor."""
        res = C.c_int32(-1)
        check(lib().luxgi_svgf_shadow_denoise(
            C.byref(self.rp), C.byref(self.ap), self.iterations, self.W, self.H, self.storage,
            C.byref(gbuffer_struct(cur)), C.byref(gbuffer_struct(prev)), ptr(mask), ptr(self.outputs[0]),
            ptr(self.outputs[1]), ptr(self.moments[0]), ptr(self.moments[1]), self.ping_pong, ptr(self.atrous[0]),
            ptr(self.atrous[1]), ptr(self.tile_map), C.byref(res), stream_handle()), "luxgi_svgf_shadow_denoise")
        self.ping_pong = 1 - self.ping_pong
        self.result = self.outputs[0] if res.value < 0 else self.atrous[res.value]
        return self.result


def make_gbuffer(prims_dev, n_prims, sdf_data, sdf, res, view_proj_inv, prev_view_proj, width, height, seed,
                 with_mask=True):
    """G-buffer + noisy signal (+ shadow mask) from SDF primary rays (luxgi_scene_gbuffer)."""
    g = {k: torch.zeros((height, width, 4), dtype=torch.float32, device="cuda") for k in ("color", "normals", "pbr")}
    g["depth"] = torch.ones((height, width), dtype=torch.float32, device="cuda")
    noisy = torch.zeros((height, width, 4), dtype=torch.float32, device="cuda")
    mask = None
    if with_mask:
        mask = torch.zeros(((height + 3) // 4, (width + 7) // 8), dtype=torch.int32, device="cuda")
    vpi = abi.mat4_colmajor(view_proj_inv)
    pvp = abi.mat4_colmajor(prev_view_proj)
    check(lib().luxgi_scene_gbuffer(ptr(prims_dev), n_prims, C.byref(sdf_data), ptr(sdf), res, vpi, pvp, width, height,
                                    seed, ptr(g["color"]), ptr(g["normals"]), ptr(g["pbr"]), ptr(g["depth"]),
                                    ptr(noisy), ptr(mask), stream_handle()), "luxgi_scene_gbuffer")
    return g, noisy, mask

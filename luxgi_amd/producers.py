"""Host mirror of the noisy-signal producers traced through the global SDF (libluxgi.so).

Mirrors:
  raytraced_shadow soft trace (Engine/Raytrace/RaytracedShadow.cpp:340-356, SDF/SDFShadow.comp)
      -> sdf_shadow_mask
  raytraced_reflection::soft_ray_trace (Engine/Raytrace/RaytracedReflection.cpp:307-344,
      SDF/SDFReflection.comp) -> sdf_reflection
Blue-noise tables are the engine's RGBA8 textures (Engine/Noise/BlueNoise.h:15-26) as uint8 arrays.
"""
import ctypes as C

import numpy as np
import torch

from . import abi
from ._lib import check, lib
from .ddgi import ptr, stream_handle, to_device


class BlueNoiseTables:
    """sobol_256_4d ([256][4] uint8) and scrambling_ranking_128x128_2d_1spp ([128][128][4] uint8) on the GPU."""

    def __init__(self, sobol, scrambling):
        sobol = np.ascontiguousarray(sobol, dtype=np.uint8).reshape(256, 4)
        scrambling = np.ascontiguousarray(scrambling, dtype=np.uint8).reshape(128, 128, 4)
        self.sobol = to_device(sobol)
        self.scrambling = to_device(scrambling)
        self.struct = abi.BlueNoise(ptr(self.sobol), ptr(self.scrambling))


def sdf_shadow_mask(params, noise, depth, normals, scene, mask=None):
    """luxgi_sdf_shadow_mask: depth [H][W] and normals [H][W][4] FP32 device tensors, scene a
    ddgi.DeviceScene (global SDF). Returns the R32UI mask [ceil(H/4)][ceil(W/8)] (int32 tensor)."""
    H, W = depth.shape
    if mask is None:
        mask = torch.zeros(((H + 3) // 4, (W + 7) // 8), dtype=torch.int32, device="cuda")
    check(lib().luxgi_sdf_shadow_mask(C.byref(params), C.byref(noise.struct), W, H, ptr(depth), ptr(normals),
                                      C.byref(scene.sdf_data), C.byref(scene.volume), ptr(mask), stream_handle()),
          "luxgi_sdf_shadow_mask")
    return mask


def sdf_reflection(params, noise, gbuffer, scene, out, volume=None):
    """luxgi_sdf_reflection: gbuffer a dict of FP32 device tensors ("normals", "pbr", "depth", and
    optionally "color"), scene a ddgi.DeviceScene (global SDF, surface atlas, sky), out an [H][W][4]
    FP32 or FP16 tensor (sky pixels are left untouched), volume a ddgi.DDGIVolume for
    approximateWithDDGI. Allocates only on the first call for a scene."""
    H, W = gbuffer["depth"].shape
    gb = abi.GBuffer(ptr(gbuffer.get("color")), ptr(gbuffer["normals"]), ptr(gbuffer["pbr"]), ptr(gbuffer["depth"]))
    # the workspace is kept on the scene and grown when needed (calls on one scene share it, so
    # calls that overlap on different streams race)
    ws_bytes = max(lib().luxgi_sdf_reflection_workspace_size(scene.n_objects), 256)
    ws = getattr(scene, "_reflection_ws", None)
    if ws is None or ws.numel() < ws_bytes:
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")
        scene._reflection_ws = ws
    storage = abi.STORAGE_FP16 if out.dtype == torch.float16 else abi.STORAGE_FP32
    u = C.byref(volume.u) if volume is not None else None
    probes = C.byref(volume.atlas[volume.ping_pong]) if volume is not None else None
    check(lib().luxgi_sdf_reflection(C.byref(params), C.byref(noise.struct), W, H, C.byref(gb), C.byref(scene.sdf_data),
                                     C.byref(scene.volume), C.byref(scene.atlas), C.byref(scene.skybox), u, probes,
                                     storage, ptr(out), ptr(ws), ws.numel(), stream_handle()), "luxgi_sdf_reflection")
    return out

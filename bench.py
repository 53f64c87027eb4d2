"""Benchmark: LuxGI GI hot path on one MI355X, plain FP32 GPU form at full size.

Workload (one "step" = one frame of the hot path on resident inputs):
  * DDGI update, SDF branch: 32x16x32 probes x 256 rays (4,194,304 rays) traced
    through a 256^3 FP32 global SDF of 2,000 seeded primitives (+ 64^3 coarse
    mip) with surface-atlas radiance at hits (2048^2 RGBA8 atlas, 2% emissive
    texels), then the irradiance (8x8) / depth (16x16) blend with hysteresis and
    the border copy fused into the same pass.
  * SVGF reflection denoise at 1920x1080 on a G-buffer from SDF primary rays:
    reprojection + moments + 17x17 clip + tile classification, 5 a-trous passes.
Reported beside the step (not in value): the reduced-precision form of the same
frame (FP16 SDF bricks + FP32 hit re-evaluation, FP16-MFMA blend with one rounding per
operand, RGBA16F SVGF; reduced_precision), and per-pixel probe sampling of the updated
volume at 1920x1080 (probe_sample_ms; bytes = G-buffer depth + normal + RGBA32F
output per pixel + both atlases once), and the surface-atlas passes (surface_atlas): chunk
culling of the objects and the relight of every atlas tile (emissive copy, one directional
light with an SDF shadow ray per lit texel, DDGI bounce) into an RGBA16F light cache, and
the SDF soft-shadow mask at 1920x1080 (sdf_shadow_mask_ms: one blue-noise disk ray per pixel), its
shadow SVGF frame (shadow_svgf_ms: reprojection + 4 a-trous with power on the last), and
the SDF reflection rays at 1920x1080 (sdf_reflection_ms: one GGX-sampled ray per pixel with the
surface-atlas lookup at its hit, RGBA16F out), and the global SDF build from mesh SDFs (sdf_build:
2,000 instances of three baked 32^3 meshes rasterized into the 256^3 cascade), the learned a-trous
chain at 1920x1080 (learned_atrous: 4 passes, and one training step), the batched forms
(svgf_batched: 16 and 64 frames of uneven sizes per-frame / batched / graph-replayed, schedule
from the autotuner; ddgi_batched: eight uneven probe volumes one after another, on four streams,
in one trace launch sequence and as one graph replay; batched_scenes: 64 volumes from 64 seeded
scenes, SDF 64^3-256^3 in FP16, per-volume calls against one graph replay of the several-scene
trace and the batched blend), and the per-pixel edge-stopping network (learned_pixel: trained on
rendered config-1-style views, evaluated at 1920x1080 against plain SVGF).
value = rays / DDGI-update time (Grays/s). All timing uses HIP events on the
stream the kernels are launched on (torch's current stream), W warm-up steps
then exactly K timed steps between device synchronisations.

roofline: the trace kernel (dominant), HBM-bound; achieved = algorithmic bytes
per launch (from a counting run of the same kernel on the same frame:
mip+fine lookups x 8 voxels x 4 B + normal evaluations x 48 x 4 B + tile
gathers x 4 texels x (4 B depth + 4 B colour) + rays x 32 B of output) / the
mean trace-launch time. traffic = HBM bytes per launch measured with rocprofv3
FETCH_SIZE / WRITE_SIZE (profiles/trace_hbm_traffic.json), null when absent.
cpu_baseline: the FP32 CPU transcription (oracle/liboracle.so, std::thread over
probes) on a 32x16xk slab of the same volume (trace + blend, k sized to ~10 s of
CPU work), on the host cores, scaled to Grays/s; with it (error_vs_fp64) the error
of one first frame of the plain and reduced-precision forms against the FP64
oracle on the two lowest probe layers.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores)")
    ap.add_argument("--svgf-iterations", type=int, default=5)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the reduced-precision form")
    ap.add_argument("--no-surface", action="store_true", help="skip the surface-atlas passes")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched SVGF measurement")
    ap.add_argument("--sdf-layout", choices=["cell8", "brick4", "linear"], default="cell8",
                    help="SDF layout of both forms (cell8: one 8-corner record per lookup)")
    ap.add_argument("--step-only", action="store_true",
                    help="time the step (DDGI update + SVGF) and the counting run only: no side passes, batched "
                         "forms, reduced-precision form or CPU baseline (for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after the run, launch a plain copy kernel (256 MiB read + 256 MiB written, 3 times) "
                         "whose rocprofv3 FETCH_SIZE / WRITE_SIZE scale the trace counters to bytes")
    return ap.parse_args()


def device_copy_gbs():
    """Measured HBM bandwidth of a plain device copy (1 GiB read + 1 GiB written), beside the nominal peak."""
    import torch
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(10):
        dst.copy_(src)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 10
    del src, dst
    torch.cuda.empty_cache()
    return 2 * (1 << 28) * 4 / (ms * 1e-3) / 1e9


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)



def bench_svgf_batched(args, prims_dev, n_prims, sdf_data, sdf, res, q):
    """Per-frame vs batched vs graph-replayed reflection denoise on two batches of uneven frames."""
    import tempfile

    import numpy as np
    import torch

    from luxgi_amd import scene
    from luxgi_amd.autotune import Autotuner, tune_svgf_batch
    from luxgi_amd.batch import ReflectionDenoiseBatch
    from luxgi_amd.svgf import ReflectionDenoiser, make_gbuffer
    workloads = {
        "16 frames 2560x1440..640x360": [(2560, 1440), (1920, 1080), (1280, 720), (640, 360)] * 4,
        "64 frames 480x270..160x90": [(480, 270), (320, 180), (240, 135), (160, 90)] * 16,
    }
    out = {}
    tuner = Autotuner(path=os.path.join(tempfile.mkdtemp(), "autotune.json"))
    for name, sizes in workloads.items():
        frames = []
        for k, (w, h) in enumerate(sizes):
            VP, VPp, eye, delta = scene.camera_pair(w, h, seed=4 + k, eye=(0.0, 0.3, 2.4))
            VPI = np.linalg.inv(VP)
            cur, noisy, _ = make_gbuffer(prims_dev, n_prims, sdf_data, sdf, res, VPI, VPp, w, h, seed=1 + k,
                                         with_mask=False)
            prev, _, _ = make_gbuffer(prims_dev, n_prims, sdf_data, sdf, res, np.linalg.inv(VPp), VPp, w, h,
                                      seed=100 + k, with_mask=False)
            frames.append((cur, prev, noisy, (VPI, VPp, eye, delta)))
        singles = []
        for (w, h), (cur, prev, noisy, cam) in zip(sizes, frames):
            d = ReflectionDenoiser(w, h, iterations=4)
            d.set_camera(*cam)
            singles.append(d)
        batch = ReflectionDenoiseBatch(sizes, iterations=4)
        graphed = ReflectionDenoiseBatch(sizes, iterations=4)
        for b in (batch, graphed):
            for k, (cur, prev, noisy, cam) in enumerate(frames):
                b.set_frame(k, cur, prev, noisy, *cam)
        choice = tune_svgf_batch(batch, tuner)
        graphed.grid_blocks, graphed.order = batch.grid_blocks, batch.order
        graphed.capture()

        def per_frame():
            for d, (cur, prev, noisy, _) in zip(singles, frames):
                d.denoise(cur, prev, noisy)

        def timed(fn, reps):
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            t0 = time.perf_counter()
            ev[0].record()
            for _ in range(reps):
                fn()
            ev[1].record()
            torch.cuda.synchronize()
            return ev[0].elapsed_time(ev[1]) / reps, (time.perf_counter() - t0) * 1e3 / reps

        reps = max(args.steps, 5)
        t_single = timed(per_frame, reps)
        t_batch = timed(batch.denoise, reps)
        t_graph = timed(graphed.denoise, reps)
        # the batched outputs equal the per-frame ones (same inputs, same state sequence)
        out[name] = {"frames": len(sizes), "pixels": int(sum(w * h for w, h in sizes)),
                     "per_frame_calls_ms": t_single[0], "batched_ms": t_batch[0], "graph_replay_ms": t_graph[0],
                     "per_frame_calls_wall_ms": t_single[1], "batched_wall_ms": t_batch[1],
                     "graph_replay_wall_ms": t_graph[1],
                     "launches": {"per_frame": 5 * len(sizes), "batched": 5, "graph": 1},
                     "speedup_batched_vs_per_frame": t_single[0] / t_batch[0],
                     "speedup_graph_vs_per_frame": t_single[0] / t_graph[0],
                     "schedule": {"gridBlocks": choice[0], "order": choice[1],
                                  "candidates_ms": tuner.last_timings or None}}
        log(f"[bench] svgf batched {name}: per-frame {t_single[0]:.3f} ms, batched {t_batch[0]:.3f} ms, "
            f"graph {t_graph[0]:.3f} ms (schedule {choice})", q)
        del singles, batch, graphed, frames
    return out


def bench_batched_scenes(args, q):
    """64 probe volumes from 64 seeded scenes (SDF 64^3-256^3 stored as FP16 cell records, atlases
    of their own; probe grids 4^3 to 32x16x32; 64-512 rays per probe), one frame each per step:
    per-volume calls (a trace and a blend per volume, the reference's dispatch pattern) against
    the several-scene trace + batched blend captured as one graph replay."""
    import numpy as np
    import torch

    from luxgi_amd import abi, scene
    from luxgi_amd.batch import DDGIBatch
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene, bake_sdf, build_mip, default_trace_params
    rng = np.random.default_rng(64)
    grids = [(4, 4, 4), (8, 4, 8), (8, 8, 8), (16, 8, 16), (16, 16, 16), (32, 16, 32)]
    t0 = time.perf_counter()
    scenes, unis, res_list = [], [], []
    for k in range(64):
        res = int(rng.choice([64, 96, 128, 192, 256]))
        n = int(rng.integers(150, 600))
        prims = scene.scene_random(n=n, seed=1000 + k)
        sdf = bake_sdf(prims, res)
        mip = build_mip(sdf, res, max_distance=2.0)
        atlas = scene.build_atlas(prims, resolution=1024 if n < 300 else 2048, tile=16,
                                  color_storage=abi.STORAGE_UNORM8, seed=1000 + k, threshold=1.05 * 2.0 / res)
        scenes.append(DeviceScene(abi.make_sdf_data(resolution=res), sdf, mip, atlas, sdf_storage=abi.STORAGE_FP16,
                                  layout=abi.LAYOUT_CELL8))
        del sdf, mip
        c = grids[int(rng.integers(0, len(grids)))]
        r = int(rng.choice([64, 128, 256, 512]))
        unis.append(abi.make_uniform(c, (-1.0, -1.0, -1.0), [2.0 / (m - 1) for m in c], r, irr_oct=8,
                                     depth_oct=16, hysteresis=0.98))
        res_list.append(res)
    build_s = time.perf_counter() - t0
    singles = [DDGIVolume(u) for u in unis]
    graphed = DDGIBatch(unis, scenes, scene_of=list(range(64)), streams=0, trace="batched", blend="batched")
    frames = [[scene.random_rotation(int(s)) for s in rng.integers(0, 2 ** 31, size=64)]
              for _ in range(args.warmup + args.steps + 1)]

    def per_volume(rots):
        for vol, ds, r in zip(singles, scenes, rots):
            vol.update(ds, default_trace_params(r, frame=vol.frames))

    per_volume(frames[0])
    graphed.update(frames[0])
    graphed.capture()
    for i in range(1, args.warmup + 1):
        per_volume(frames[i])
        graphed.update(frames[i])
    torch.cuda.synchronize()
    out = {"volumes": 64, "scenes": 64, "rays": int(sum(u.num_probes * u.raysPerProbe for u in unis)),
           "sdf_resolutions": res_list, "probe_grids": [list(u.probeCounts) for u in unis],
           "rays_per_probe": [int(u.raysPerProbe) for u in unis], "sdf_storage": "fp16 cell records",
           "scene_build_s": build_s}
    for name, fn in (("per_volume_calls_ms", per_volume), ("graph_replay_ms", graphed.update)):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        t1 = time.perf_counter()
        ev[0].record()
        for i in range(args.warmup + 1, args.warmup + 1 + args.steps):
            fn(frames[i])
        ev[1].record()
        torch.cuda.synchronize()
        out[name] = ev[0].elapsed_time(ev[1]) / args.steps
        out[name.replace("_ms", "_wall_ms")] = (time.perf_counter() - t1) * 1e3 / args.steps
    out["graph_equals_per_volume"] = bool(all(
        torch.equal(a.current_irradiance, b.current_irradiance) and torch.equal(a.current_depth, b.current_depth)
        for a, b in zip(singles, graphed.volumes)))
    out["speedup_graph_vs_per_volume"] = out["per_volume_calls_ms"] / out["graph_replay_ms"]
    out["graph_grays_per_s"] = out["rays"] / (out["graph_replay_ms"] * 1e-3) / 1e9
    log(f"[bench] 64 volumes / 64 scenes ({out['rays'] / 1e6:.1f} M rays): per-volume calls "
        f"{out['per_volume_calls_ms']:.3f} ms, graph replay {out['graph_replay_ms']:.3f} ms "
        f"(equal: {out['graph_equals_per_volume']})", q)
    del singles, graphed, scenes
    torch.cuda.empty_cache()
    return out


def bench_learned_pixel(args, q):
    """The per-pixel 12-32-32-5 edge-stopping network: trained on config-1-style views the method
    renders (reduced here to 24 scenes at 480x270 and 400 steps; the full run is
    `python -m luxgi_amd.learned_pixel`), evaluated at 1920x1080 on two held-out scenes against
    plain SVGF, both feeding back their own filtered history; its 1080p cost beside the reference
    chain's."""
    from luxgi_amd import learned_pixel as lp
    rep, _ = lp.run(train_scenes=24, eval_scenes=2, width=480, height=270, steps=400,
                    log=None if q else (lambda m: log(f"[bench] learned pixel: {m}", q)))
    log(f"[bench] learned pixel: held-out MSE plain {rep['eval_mse']['plain']:.6f}, learned "
        f"{rep['eval_mse']['learned']:.6f} ({rep['learned_over_plain']:.3f}x); 1080p chain "
        f"{rep['timing_ms']['learned_chain_ms']:.3f} ms vs reference {rep['timing_ms']['reference_chain_ms']:.3f} ms", q)
    return rep


def bench_ddgi_batched(args, ds, q):
    """Eight probe volumes of uneven probe and ray counts over the bench scene: per-volume updates on
    one stream and on four streams, all volumes traced in one launch sequence (blends on four
    streams), and that frame as one graph replay (orientations read from device memory)."""
    import numpy as np
    import torch

    from luxgi_amd import abi, scene
    from luxgi_amd.batch import DDGIBatch
    shapes = [((8, 8, 8), 128), ((16, 8, 16), 128), ((8, 4, 8), 256), ((12, 6, 12), 192), ((6, 6, 6), 64),
              ((10, 5, 10), 96), ((4, 4, 4), 256), ((14, 7, 14), 128)]
    unis = [abi.make_uniform(c, (-1.0, -1.0, -1.0), [2.0 / (n - 1) for n in c], r, irr_oct=8, depth_oct=16)
            for c, r in shapes]
    serial = DDGIBatch(unis, ds, streams=0, trace="per_volume")
    streamed = DDGIBatch(unis, ds, trace="per_volume")
    eager = DDGIBatch(unis, ds)
    graphed = DDGIBatch(unis, ds)
    rng = np.random.default_rng(99)
    frames = [[scene.random_rotation(int(s)) for s in rng.integers(0, 2 ** 31, size=len(unis))]
              for _ in range(args.warmup + args.steps + 1)]
    for b in (serial, streamed, eager, graphed):
        b.update(frames[0])
    graphed.capture()
    for i in range(1, args.warmup + 1):
        for b in (serial, streamed, eager, graphed):
            b.update(frames[i])
    torch.cuda.synchronize()
    out = {"volumes": len(unis), "rays": int(sum(u.num_probes * u.raysPerProbe for u in unis)),
           "shapes": [[list(c), r] for c, r in shapes]}
    for name, b in (("per_volume_one_stream_ms", serial), ("per_volume_4_streams_ms", streamed),
                    ("one_trace_launch_ms", eager), ("graph_replay_ms", graphed)):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for i in range(args.warmup + 1, args.warmup + 1 + args.steps):
            b.update(frames[i])
        ev[1].record()
        torch.cuda.synchronize()
        out[name] = ev[0].elapsed_time(ev[1]) / args.steps
    same = all(torch.equal(a.current_irradiance, b.current_irradiance) and
               torch.equal(a.current_irradiance, c.current_irradiance) and
               torch.equal(a.current_irradiance, d.current_irradiance)
               for a, b, c, d in zip(eager.volumes, graphed.volumes, serial.volumes, streamed.volumes))
    out["graph_equals_per_volume"] = bool(same)
    out["speedup_graph_vs_one_stream"] = out["per_volume_one_stream_ms"] / out["graph_replay_ms"]
    log(f"[bench] ddgi batched {len(unis)} volumes: one stream {out['per_volume_one_stream_ms']:.3f} ms, "
        f"4 streams {out['per_volume_4_streams_ms']:.3f} ms, one trace launch {out['one_trace_launch_ms']:.3f} ms, "
        f"graph {out['graph_replay_ms']:.3f} ms", q)
    return out

def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (probe volumes would shard by "
              "probe z-slices); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from luxgi_amd import abi, scene
    from luxgi_amd.ddgi import DDGIVolume, DeviceScene, bake_sdf, build_mip, default_trace_params, to_device
    from luxgi_amd.svgf import ReflectionDenoiser, make_gbuffer

    q = args.quiet
    torch.cuda.set_device(0)
    t_setup = time.perf_counter()
    # ---------------- scene (not timed) ----------------
    res = 256
    prims = scene.scene_random(n=2000, seed=2)
    sdf = bake_sdf(prims, res)
    mip = build_mip(sdf, res, max_distance=2.0)
    g = abi.make_sdf_data(resolution=res)
    log(f"[bench] sdf {res}^3 baked", q)
    atlas = scene.build_atlas(prims, resolution=2048, tile=16, color_storage=abi.STORAGE_UNORM8, seed=3,
                              threshold=1.05 * 2.0 / res, surface_gbuffer=not args.no_surface)
    log("[bench] atlas 2048^2 built", q)
    layout = {"cell8": abi.LAYOUT_CELL8, "brick4": abi.LAYOUT_BRICK4, "linear": abi.LAYOUT_LINEAR}[args.sdf_layout]
    ds = DeviceScene(g, sdf, mip, atlas, layout=layout)
    counts = (32, 16, 32)
    step = [2.0 / (c - 1) for c in counts]
    u = abi.make_uniform(counts, (-1.0, -1.0, -1.0), step, 256, irr_oct=8, depth_oct=16, hysteresis=0.98)
    vol = DDGIVolume(u)
    rays_per_frame = u.num_probes * u.raysPerProbe
    # SVGF inputs: 1080p G-buffer pair from SDF primary rays
    W, H = 1920, 1080
    VP, VPp, eye, delta = scene.camera_pair(W, H, seed=4, eye=(0.0, 0.3, 2.4))
    VPI = np.linalg.inv(VP)
    prims_dev = to_device(np.frombuffer(prims.tobytes(), dtype=np.uint8))
    cur, noisy, _ = make_gbuffer(prims_dev, len(prims), g, sdf, res, VPI, VPp, W, H, seed=1, with_mask=False)
    VPI_prev = np.linalg.inv(VPp)
    prev, _, _ = make_gbuffer(prims_dev, len(prims), g, sdf, res, VPI_prev, VPp, W, H, seed=2, with_mask=False)
    dn = ReflectionDenoiser(W, H, iterations=args.svgf_iterations)
    dn.set_camera(VPI, VPp, eye, delta)
    torch.cuda.synchronize()
    log(f"[bench] setup {time.perf_counter() - t_setup:.1f}s", q)

    rng = np.random.default_rng(12345)
    rotations = [scene.random_rotation(int(s)) for s in rng.integers(0, 2 ** 31, size=args.warmup + args.steps)]

    def frame(i, ev=None):
        p = default_trace_params(rotations[i], frame=vol.frames)
        if ev:
            ev[0].record()
        vol.trace_rays(ds, p)
        if ev:
            ev[1].record()
        vol.probe_update(fuse_border=True)
        if ev:
            ev[2].record()
        vol.end_frame()
        dn.denoise(cur, prev, noisy)
        if ev:
            ev[3].record()

    for i in range(args.warmup):
        frame(i)
    torch.cuda.synchronize()
    log(f"[bench] warmup {args.warmup} steps done", q)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        frame(args.warmup + k, evs[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_trace = np.array([e[0].elapsed_time(e[1]) for e in evs])
    t_upd = np.array([e[1].elapsed_time(e[2]) for e in evs])
    t_svgf = np.array([e[2].elapsed_time(e[3]) for e in evs])
    ms_step = wall * 1e3 / args.steps
    ddgi_ms = float(np.mean(t_trace + t_upd))
    log(f"[bench] timed {args.steps} steps: trace {t_trace.mean():.3f} ms, update {t_upd.mean():.3f} ms, "
        f"svgf {t_svgf.mean():.3f} ms, wall/step {ms_step:.3f} ms", q)

    # the passes reported beside the step (skipped with --step-only, e.g. for profiler runs)
    sample_ms = sample_bytes = learned = surf = shadow_ms = shadow_svgf_ms = refl_ms = sdf_build = None
    if not args.step_only:
        # ---------------- probe sampling at 1080p (sample_probe pass; reported beside the step) ----------------
        pev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        irr_out = torch.empty((H, W, 4), dtype=torch.float32, device="cuda")
        for _ in range(2):
            vol.sample_probes(cur["depth"], cur["normals"], eye, VPI, out=irr_out)
        pev[0].record()
        for _ in range(args.steps):
            vol.sample_probes(cur["depth"], cur["normals"], eye, VPI, out=irr_out)
        pev[1].record()
        torch.cuda.synchronize()
        sample_ms = pev[0].elapsed_time(pev[1]) / args.steps
        # G-buffer depth (4 B) + octahedral normal (8 B) + RGBA32F output (16 B) per pixel, atlases read once
        atlas_bytes = sum(t.numel() * t.element_size() for t in (vol.current_irradiance, vol.current_depth))
        sample_bytes = W * H * (4 + 8 + 16) + atlas_bytes

        # ---------------- learned a-trous chain at 1080p (reported beside the step) ----------------
        # 4 passes of the learned edge-stopping filter on the accumulated reflection signal, against the
        # same 4 passes of the reference a-trous (tile map from the frame's reprojection)
        from luxgi_amd._lib import lib as _L
        from luxgi_amd.ddgi import ptr, stream_handle
        from luxgi_amd.learned import LearnedAtrous
        from luxgi_amd.svgf import gbuffer_struct
        la = LearnedAtrous(W, H, iterations=4)
        la_in = dn.outputs[0].float().contiguous()
        gbs = gbuffer_struct(cur)
        ref_bufs = [torch.empty_like(la_in), torch.empty_like(la_in)]

        def ref_chain():
            src = la_in
            for i in range(4):
                ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
                _L().luxgi_svgf_reflection_atrous(ctypes.byref(ap), W, H, abi.STORAGE_FP32, ctypes.byref(gbs),
                                                  ptr(dn.tile_map), ptr(src), ptr(ref_bufs[i & 1]), None,
                                                  stream_handle())
                src = ref_bufs[i & 1]

        lev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        for _ in range(2):
            la.forward(cur, la_in)
            ref_chain()
        lev[0].record()
        for _ in range(args.steps):
            la.forward(cur, la_in)
        lev[1].record()
        for _ in range(args.steps):
            ref_chain()
        lev[2].record()
        torch.cuda.synchronize()
        # one training step (4 passes forward, 4 backward with the parameter-gradient reduction, Adam)
        from luxgi_amd.learned import filtered_mask
        la_mask = filtered_mask(cur)
        la_target = la_in.clone()
        for _ in range(2):
            la.train_step(cur, la_in, la_target, la_mask, lr=1e-4)
        lev[4].record()
        for _ in range(args.steps):
            la.train_step(cur, la_in, la_target, la_mask, lr=1e-4)
        lev[5].record()
        torch.cuda.synchronize()
        learned = {"learned_4_passes_ms": lev[0].elapsed_time(lev[1]) / args.steps,
                   "reference_4_passes_ms": lev[1].elapsed_time(lev[2]) / args.steps,
                   "train_step_ms": lev[4].elapsed_time(lev[5]) / args.steps,
                   "mlp": "6-10-1 ReLU, 81 parameters, one per tap", "image": [W, H]}
        log(f"[bench] learned a-trous 1080p x4 {learned['learned_4_passes_ms']:.3f} ms "
            f"(reference {learned['reference_4_passes_ms']:.3f} ms), train step {learned['train_step_ms']:.3f} ms", q)
        del la, la_in, ref_bufs, la_target

        # ---------------- surface atlas passes (reported beside the step) ----------------
        # chunk culling of the 2,000 objects, then the surface-cache relight of all 12,000 tiles
        # (16x16 texels) with one directional light (SDF shadow ray per lit texel), the emissive
        # copy and the DDGI bounce from the updated volume, into an RGBA16F light cache
        surf = None
        if not args.no_surface:
            from luxgi_amd.surface import SurfaceCache, cull_chunks
            cap = int(atlas.culled.size) + 64
            for _ in range(2):
                cull_chunks(atlas.data, ds.objects, atlas.n_objects, cap)
            cev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            cev[0].record()
            for _ in range(args.steps):
                cull_chunks(atlas.data, ds.objects, atlas.n_objects, cap)
            cev[1].record()
            sc = SurfaceCache(atlas, out_storage=abi.STORAGE_FP16)
            sun = (abi.Light * 1)()
            sun[0].color[:] = [1.0, 0.95, 0.9, 1.0]
            dsun = np.array([0.3, -0.8, -0.52]) / np.linalg.norm([0.3, -0.8, -0.52])
            sun[0].direction[:] = [float(dsun[0]), float(dsun[1]), float(dsun[2]), 0.0]
            sun[0].intensity, sun[0].type = 2.0, abi.LIGHT_DIRECTIONAL
            sc.set_lights(sun)
            passes = abi.SURFACE_EMISSIVE | abi.SURFACE_DIRECT | abi.SURFACE_INDIRECT
            for _ in range(2):
                sc.light(ds, passes=passes, camera=eye, shadow_bias=2.0 * 2.0 / res, volume=vol)
            sev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            sev[0].record()
            for _ in range(args.steps):
                sc.light(ds, passes=passes, camera=eye, shadow_bias=2.0 * 2.0 / res, volume=vol)
            sev[1].record()
            torch.cuda.synchronize()
            light_ms = sev[0].elapsed_time(sev[1]) / args.steps
            texels = sc.n_draws * 16 * 16
            lit = int(sum(1 for k in range(sc.n_draws) if sc.draws_host[k].lightMask))
            surf = {"cull_ms": cev[0].elapsed_time(cev[1]) / args.steps, "surface_light_ms": light_ms,
                    "tiles": sc.n_draws, "tiles_with_light": lit, "texels": texels,
                    "gtexels_per_s": texels / (light_ms * 1e-3) / 1e9,
                    "passes": "emissive + 1 directional light with SDF shadow rays + DDGI bounce", "out": "RGBA16F"}
            log(f"[bench] surface atlas: cull {surf['cull_ms']:.3f} ms, relight {light_ms:.3f} ms", q)
            del sc

        # ---------------- SDF soft-shadow mask at 1080p (the shadow SVGF's input; reported beside the step) ----
        # one ray per pixel toward a blue-noise point of a sun disk; synthetic blue-noise tables of the
        # engine's shapes (a permutation per Sobol dimension, random scrambling-ranking bytes)
        from luxgi_amd.producers import BlueNoiseTables, sdf_shadow_mask
        nrng = np.random.default_rng(21)
        noise = BlueNoiseTables(np.stack([nrng.permutation(256) for _ in range(4)], axis=1).astype(np.uint8),
                                nrng.integers(0, 256, (128, 128, 4), dtype=np.uint8))
        sp = abi.SdfShadowParams()
        sp.light.color[:] = [1.0, 1.0, 1.0, 1.0]
        sp.light.direction[:] = [0.3, -0.8, -0.52, 0.02]
        sp.light.intensity, sp.light.type = 1.0, abi.LIGHT_DIRECTIONAL
        sp.viewProjInv = abi.mat4_colmajor(VPI)
        sp.shadowBias = 2.0 * 2.0 / res
        smask = torch.zeros(((H + 3) // 4, (W + 7) // 8), dtype=torch.int32, device="cuda")
        for f in range(2):
            sp.numFrames = f
            sdf_shadow_mask(sp, noise, cur["depth"], cur["normals"], ds, mask=smask)
        mev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        mev[0].record()
        for f in range(args.steps):
            sp.numFrames = f
            sdf_shadow_mask(sp, noise, cur["depth"], cur["normals"], ds, mask=smask)
        mev[1].record()
        torch.cuda.synchronize()
        shadow_ms = mev[0].elapsed_time(mev[1]) / args.steps
        log(f"[bench] sdf shadow mask 1080p {shadow_ms:.3f} ms", q)
        # the shadow SVGF frame at 1080p (RaytracedShadow.cpp:543-609): reprojection of the mask, 4
        # a-trous passes with the history copy after iteration 1 and `power` on the last one
        from luxgi_amd.svgf import ShadowDenoiser
        sdn = ShadowDenoiser(W, H)
        sdn.set_camera(VPI)
        for _ in range(2):
            sdn.denoise(cur, prev, smask)
        sev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        sev[0].record()
        for _ in range(args.steps):
            sdn.denoise(cur, prev, smask)
        sev[1].record()
        torch.cuda.synchronize()
        shadow_svgf_ms = sev[0].elapsed_time(sev[1]) / args.steps
        log(f"[bench] shadow svgf 1080p {shadow_svgf_ms:.3f} ms", q)
        del sdn
        # SDF reflection rays at 1080p (the reflection SVGF's noisy input): one GGX-sampled ray per pixel
        # (the G-buffer's roughness lies in (0.05, 0.45)), coloured by the surface-atlas lookup or the sky
        from luxgi_amd.producers import sdf_reflection
        rp = abi.SdfReflectionParams()
        rp.trim, rp.roughDDGIIntensity = 1.0, 1.0
        rp.cameraPosition[:] = [float(eye[0]), float(eye[1]), float(eye[2]), 1.0]
        rp.viewProjInv = abi.mat4_colmajor(VPI)
        rp.skyColor[:] = [0.6, 0.7, 0.9, 1.0]
        refl_out = torch.zeros((H, W, 4), dtype=torch.float16, device="cuda")
        for f in range(2):
            rp.numFrames = f
            sdf_reflection(rp, noise, cur, ds, refl_out)
        rev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        rev[0].record()
        for f in range(args.steps):
            rp.numFrames = f
            sdf_reflection(rp, noise, cur, ds, refl_out)
        rev[1].record()
        torch.cuda.synchronize()
        refl_ms = rev[0].elapsed_time(rev[1]) / args.steps
        log(f"[bench] sdf reflection 1080p {refl_ms:.3f} ms", q)

        # ---------------- global SDF build from mesh SDFs (reported beside the step) ----------------
        # the 2,000 primitives as instances of three baked 32^3 unit meshes (+2 mips), rasterized into
        # all 512 chunks of the 256^3 cascade (luxgi_sdf_rasterize; FP32 volume, as the trace reads it)
        from luxgi_amd import meshsdf
        msh, mobj, mbounds, mwhich = meshsdf.primitive_instances(prims)
        cas0 = meshsdf.Cascade(0, (0.0, 0.0, 0.0), 1.0, res)
        mlists = meshsdf.chunk_lists(cas0, mbounds)
        builder = meshsdf.GlobalSDFBuilder(msh, mobj, len(prims), which=mwhich)
        built = torch.empty((res, res, res), dtype=torch.float32, device="cuda")
        for _ in range(2):
            builder.rasterize(cas0, mlists, built)
        bev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        bev[0].record()
        for _ in range(args.steps):
            builder.rasterize(cas0, mlists, built)
        bev[1].record()
        torch.cuda.synchronize()
        build_ms = bev[0].elapsed_time(bev[1]) / args.steps
        # against the analytic SDF on the voxels within one voxel of a surface: the instances scale unit
        # meshes non-uniformly (distances scaled by one factor, as the engine does) and a chunk is only
        # rasterized with the objects whose bounds + 4 voxels reach it (other chunks clear to 1), so
        # report the distribution in voxels and the cleared fraction rather than a maximum
        vox = 1.0 / res  # one voxel in the stored (distance / 2) units
        near = torch.abs(sdf) < vox
        d_near = (torch.abs(built - sdf)[near] / vox).float()
        sdf_build = {"rasterize_ms": build_ms, "objects": len(prims), "chunk_object_pairs": int(mlists[1][-1]),
                     "voxels": res ** 3,
                     "near_surface_abs_diff_voxels": {"median": float(d_near.median()),
                                                      "p99": float(torch.quantile(d_near[:1 << 24], 0.99))},
                     "near_surface_cleared_fraction": float((built[near] >= 1.0).float().mean())}
        log(f"[bench] global SDF rasterize {build_ms:.3f} ms", q)
        del builder, built

    # ---------------- batched SVGF over frames of uneven sizes (reported beside the step) ----------------
    # the same reflection denoise (reprojection + 4 a-trous) for many frames at once: per-frame calls
    # (1 + 4 launches each) against the batched persistent launches (5 in all) and their graph replay;
    # the persistent grid and unit order chosen by the autotuner
    if args.step_only:
        args.no_batch = args.no_mixed = args.no_cpu = True
    svgf_batched = None if args.no_batch else bench_svgf_batched(args, prims_dev, len(prims), g, sdf, res, q)
    ddgi_batched = None if args.no_batch else bench_ddgi_batched(args, ds, q)
    batched_scenes = None if args.no_batch else bench_batched_scenes(args, q)
    learned_pixel = None if args.no_batch else bench_learned_pixel(args, q)
    # ---------------- reduced-precision form (reported beside the step, not in value) ----------------
    # FP16 SDF bricks with per-brick scale and FP32 re-evaluation at hits, FP16-MFMA blend (one
    # rounding per operand, FP32 accumulation), SVGF with RGBA16F intermediates. Ray buffers stay
    # FP32: RGBA16F distances lose the depth moments of near hits to the cancellation in t - 0.01.
    mixed = None
    if not args.no_mixed:
        ds16 = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_FP16, layout=layout,
                           brick_scale=layout == abi.LAYOUT_BRICK4,
                           refine=True)
        vol16 = DDGIVolume(u)
        dn16 = ReflectionDenoiser(W, H, storage=abi.STORAGE_FP16, iterations=args.svgf_iterations)
        dn16.set_camera(VPI, VPp, eye, delta)
        noisy16 = noisy.half()

        def frame16(i, ev=None):
            p = default_trace_params(rotations[i], frame=vol16.frames)
            if ev:
                ev[0].record()
            vol16.trace_rays(ds16, p)
            if ev:
                ev[1].record()
            vol16.probe_update(fuse_border=True, half=True)
            if ev:
                ev[2].record()
            vol16.end_frame()
            dn16.denoise(cur, prev, noisy16)
            if ev:
                ev[3].record()

        for i in range(args.warmup):
            frame16(i)
        ev16 = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
        torch.cuda.synchronize()
        for k in range(args.steps):
            frame16(args.warmup + k, ev16[k])
        torch.cuda.synchronize()
        m_trace = float(np.mean([e[0].elapsed_time(e[1]) for e in ev16]))
        m_upd = float(np.mean([e[1].elapsed_time(e[2]) for e in ev16]))
        m_svgf = float(np.mean([e[2].elapsed_time(e[3]) for e in ev16]))
        mixed = {"trace_ms": m_trace, "probe_update_ms": m_upd, "svgf_ms": m_svgf, "ddgi_update_ms": m_trace + m_upd,
                 "grays_per_s": rays_per_frame / ((m_trace + m_upd) * 1e-3) / 1e9,
                 "speedup_ddgi_vs_fp32": ddgi_ms / (m_trace + m_upd),
                 "speedup_step_vs_fp32": (ddgi_ms + float(np.mean(t_svgf))) / (m_trace + m_upd + m_svgf)}
        log(f"[bench] reduced precision: trace {m_trace:.3f} ms, update {m_upd:.3f} ms, svgf {m_svgf:.3f} ms", q)
        del vol16, dn16  # ds16 stays for the error report beside the CPU baseline
        # the same with BF16 SDF cells (FP32 refine at hits) and the BF16-operand MFMA blend
        dsbf = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_BF16, layout=layout, refine=True)
        volbf = DDGIVolume(u)
        tb = []
        for k in range(args.warmup + args.steps):
            e3 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e3[0].record()
            volbf.trace_rays(dsbf, default_trace_params(rotations[k], frame=volbf.frames))
            e3[1].record()
            volbf.probe_update(fuse_border=True, half=True, bf16=True)
            e3[2].record()
            volbf.end_frame()
            if k >= args.warmup:
                tb.append(e3)
        torch.cuda.synchronize()
        b_trace = float(np.mean([e[0].elapsed_time(e[1]) for e in tb]))
        b_upd = float(np.mean([e[1].elapsed_time(e[2]) for e in tb]))
        mixed["bf16"] = {"trace_ms": b_trace, "probe_update_ms": b_upd, "ddgi_update_ms": b_trace + b_upd,
                         "speedup_ddgi_vs_fp32": ddgi_ms / (b_trace + b_upd)}
        log(f"[bench] reduced precision bf16: trace {b_trace:.3f} ms, update {b_upd:.3f} ms", q)
        del volbf
        # FP16 cells without the FP32 refine at hits (the hits keep the FP16 distances)
        dsnr = DeviceScene(g, sdf, mip, atlas, sdf_storage=abi.STORAGE_FP16, layout=layout)
        volnr = DDGIVolume(u)
        tn = []
        for k in range(args.warmup + args.steps):
            e3 = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
            e3[0].record()
            volnr.trace_rays(dsnr, default_trace_params(rotations[k], frame=volnr.frames))
            e3[1].record()
            volnr.probe_update(fuse_border=True, half=True)
            e3[2].record()
            volnr.end_frame()
            if k >= args.warmup:
                tn.append(e3)
        torch.cuda.synchronize()
        n_trace = float(np.mean([e[0].elapsed_time(e[1]) for e in tn]))
        n_upd = float(np.mean([e[1].elapsed_time(e[2]) for e in tn]))
        mixed["fp16_no_refine"] = {"trace_ms": n_trace, "probe_update_ms": n_upd, "ddgi_update_ms": n_trace + n_upd,
                                   "speedup_ddgi_vs_fp32": ddgi_ms / (n_trace + n_upd)}
        log(f"[bench] reduced precision fp16 without refine: trace {n_trace:.3f} ms, update {n_upd:.3f} ms", q)
        del volnr

    # ---------------- counting run (algorithmic bytes of the trace) ----------------
    counters = torch.zeros(8, dtype=torch.int64, device="cuda")
    vol.trace_rays(ds, default_trace_params(rotations[-1], frame=vol.frames), counters=counters)
    torch.cuda.synchronize()
    c = counters.cpu().numpy().astype(np.float64)
    n_lookup, n_normal, n_tile, n_rays = c[0] + c[1], c[2], c[4], c[5]
    trace_bytes = n_lookup * 8 * 4 + n_normal * 48 * 4 + n_tile * 4 * (4 + 4) + n_rays * 32
    trace_s = float(np.mean(t_trace)) * 1e-3
    achieved = trace_bytes / trace_s / 1e9
    traffic = None
    copy_gbs = None if args.step_only else device_copy_gbs()
    tpath = os.path.join(ROOT, "profiles", "trace_hbm_traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    # ---------------- CPU baseline (bounded sample, FP32 transcription) ----------------
    # and the error of the plain and reduced-precision forms against the FP64 oracle: one first
    # frame (zero history) of fresh volumes, compared on the two lowest probe layers
    cpu = None
    if not args.no_cpu:
        first = {}
        forms = [("fp32", ds, False, False)] + ([] if mixed is None else [("reduced", ds16, True, False),
                                                                         ("reduced_bf16", dsbf, True, True),
                                                                         ("reduced_fp16_no_refine", dsnr, True, False)])
        for name, scene_, half, bf16 in forms:
            v = DDGIVolume(u)
            v.trace_rays(scene_, default_trace_params(rotations[-1]))
            v.probe_update(fuse_border=True, half=half, bf16=bf16)
            torch.cuda.synchronize()
            first[name] = (v.irradiance[1].cpu().numpy(), v.depth[1].cpu().numpy())
            del v
        hc = {k: t.cpu().numpy() for k, t in cur.items()}
        hp = {k: t.cpu().numpy() for k, t in prev.items()}
        noisy_h = noisy.cpu().numpy()
        cpu = cpu_baseline(u, g, sdf, mip, res, atlas, rotations[-1], args.cpu_threads, q, first,
                           svgf=(dn.rp, args.svgf_iterations, W, H, hc, hp, noisy_h))
        cpu["svgf_error_vs_fp64"] = svgf_error_report(dn.rp, args.svgf_iterations, W, H, cur, prev, noisy, hc, hp,
                                                      noisy_h, (VPI, VPp, eye, delta), q)

    out = {
        "metric": "DDGI update Grays/s (32x16x32 probes x 256 rays) + 1080p SVGF ms",
        "value": rays_per_frame / (ddgi_ms * 1e-3) / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded primitives, atlas and G-buffer)",
        "config": {"workload": "plain FP32 GPU form: DDGI 32x16x32 probes x 256 rays, 256^3 SDF of 2000 "
                               "primitives, 2048^2 RGBA8 atlas, oct 8/16 + fused border; SVGF 1920x1080 "
                               f"x{args.svgf_iterations} a-trous",
                   "probes": list(counts), "rays_per_probe": 256, "sdf_resolution": res,
                   "svgf": [W, H, args.svgf_iterations], "parallelism": "single GPU"},
        "ddgi_update_ms": ddgi_ms,
        "trace_ms": float(np.mean(t_trace)),
        "probe_update_ms": float(np.mean(t_upd)),
        "svgf_ms": float(np.mean(t_svgf)),
        "probe_sample_ms": sample_ms,
        "learned_atrous": learned,
        "reduced_precision": mixed,
        "probe_sample_gbs": None if sample_ms is None else sample_bytes / (sample_ms * 1e-3) / 1e9,
        "surface_atlas": surf,
        "sdf_shadow_mask_ms": shadow_ms,
        "shadow_svgf_ms": shadow_svgf_ms,
        "sdf_reflection_ms": refl_ms,
        "sdf_build": sdf_build,
        "svgf_batched": svgf_batched,
        "ddgi_batched": ddgi_batched,
        "batched_scenes": batched_scenes,
        "learned_pixel": learned_pixel,
        "trace_grays_per_s": rays_per_frame / trace_s / 1e9,
        "trace_counts": {"lookups": n_lookup, "normals": n_normal, "objects": c[3], "tile_gathers": n_tile,
                         "rays": n_rays, "steps": c[6], "hits": c[7]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "measured_copy_gbs": copy_gbs,
                     "kernel": "luxgi_ddgi_trace_sdf launch sequence (march + hit sort + shade)",
                     "algorithmic_bytes_per_launch": trace_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(u_full, g, sdf_t, mip_t, res, atlas, rotation, threads, quiet, first=None, svgf=None):
    """svgf: (rp, iterations, W, H, cur, prev, noisy) host arrays of the bench's SVGF frame, or None.
    The host share of one GPU is 16 cores, so the pools use min(16, hardware threads) unless
    --cpu-threads says otherwise; every timing is the median of three runs."""
    import numpy as np

    import oracle
    from luxgi_amd import abi
    from luxgi_amd.ddgi import default_trace_params

    hw = oracle.threads()
    n = threads if threads > 0 else min(16, hw)
    sdf = sdf_t.cpu().numpy()
    mip = mip_t.cpu().numpy()
    p = default_trace_params(rotation)

    def run(nz):
        # a 32x16xnz slab of the same volume (same spacing, centred in z)
        counts = (u_full.probeCounts[0], u_full.probeCounts[1], nz)
        z0 = u_full.startPosition[2] + u_full.step[2] * ((u_full.probeCounts[2] - nz) // 2)
        u = abi.make_uniform(counts, (u_full.startPosition[0], u_full.startPosition[1], z0),
                             (u_full.step[0], u_full.step[1], u_full.step[2]), u_full.raysPerProbe, irr_oct=8,
                             depth_oct=16, hysteresis=0.98)
        t0 = time.perf_counter()
        rad, dd, fl, _ = oracle.trace(u, g, sdf, mip, res, (res + 3) // 4, atlas.data, atlas.chunks, atlas.culled,
                                      atlas.objects, atlas.tiles, atlas.depth, atlas.color_f32, p, precision="f32",
                                      nthreads=n)
        t1 = time.perf_counter()
        z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
        z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
        oracle.probe_update(u, rad, dd, z_i, z_d, 0, 7, precision="f32", nthreads=n)
        t2 = time.perf_counter()
        return u, t1 - t0, t2 - t1

    # size the sample to ~5 s of CPU work: probe 2 z-layers, then scale (at most the whole volume);
    # three runs of that sample, the median reported
    u, tt, tu = run(2)
    nz = int(min(u_full.probeCounts[2], max(2, 2 * 5.0 / max(tt + tu, 1e-3))))
    runs = [run(nz) for _ in range(3)]
    totals = [a[1] + a[2] for a in runs]
    u, tt, tu = runs[int(np.argsort(totals)[1])]
    rays = u.num_probes * u.raysPerProbe
    log(f"[bench] cpu baseline: {u.num_probes} probes, trace {tt:.2f}s update {tu:.2f}s on {n} threads "
        f"(median of 3: {', '.join(f'{t:.2f}' for t in totals)} s)", quiet)
    result = {"value": rays / (tt + tu) / 1e9, "unit": "Grays/s", "cores": n, "kind": "port",
              "sample": f"FP32 C++ transcription, {u.num_probes} probes (32x16x{nz} slab of the same volume) x "
                        f"{u.raysPerProbe} rays: trace {tt:.2f} s + blend {tu:.2f} s on {n} threads, median of "
                        f"3 runs (host reports {hw} hardware threads; one GPU's host share is 16)",
              "trace_s": tt, "update_s": tu, "runs_s": totals, "hardware_threads": hw, "cpu_model": cpu_model()}
    if svgf is not None:
        # SVGF on the CPU: the FP32 transcription of the reflection reprojection and a-trous passes
        # over image rows (std::thread), the bench's 1080p frame, median of three
        rp, iters, W, H, hc, hp, noisy = svgf
        z = np.zeros((H, W, 4), np.float32)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            c, m, tiles = oracle.reflection_reproject(rp, W, H, hc, hp, noisy, z, z, precision="f32", nthreads=n)
            x = c
            for i in range(iters):
                ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
                x = oracle.reflection_atrous(ap, W, H, hc["normals"], hc["pbr"], tiles, x, precision="f32",
                                             nthreads=n)
            ts.append(time.perf_counter() - t0)
        result["svgf_ms"] = float(np.median(ts)) * 1e3
        result["svgf_sample"] = (f"FP32 C++ transcription of the reflection denoise (reprojection + {iters} a-trous) "
                                 f"at {W}x{H} on {n} threads over image rows, median of 3 runs")
        log(f"[bench] cpu baseline svgf {W}x{H}: {result['svgf_ms']:.1f} ms", quiet)
    if first:
        result["error_vs_fp64"] = precision_report(u_full, g, sdf, mip, res, atlas, p, n, first)
        for name, e in result["error_vs_fp64"].items():
            log(f"[bench] {name} vs FP64: irradiance normwise {e['irradiance_normwise_rel']:.2e}, p99 "
                f"{e['irradiance_p99_rel']:.2e}, max rel {e['irradiance_max_rel']:.2e}; depth normwise "
                f"{e['depth_normwise_rel']:.2e}", quiet)
    return result


def svgf_error_report(rp, iters, W, H, cur, prev, noisy, hc, hp, noisy_h, cam, quiet):
    """One reflection denoise of the bench frame (fresh history) in FP32 and in FP16 storage
    against the FP64 oracle's chain: normwise and per-pixel relative errors of the colour, and the
    fraction of pixels within 2e-3 + 2e-3 |ref| (FP32) or 1e-2 + 1e-2 |ref| (FP16)."""
    import numpy as np
    import torch

    import oracle
    from luxgi_amd import abi
    from luxgi_amd.svgf import ReflectionDenoiser
    z = np.zeros((H, W, 4), np.float32)
    c, m, tiles = oracle.reflection_reproject(rp, W, H, hc, hp, noisy_h, z, z)
    x = c.astype(np.float32)
    for i in range(iters):
        ap = abi.ReflectionAtrousParams(1, 1 << i, 10.0, 32.0, 1.0, 1)
        x = oracle.reflection_atrous(ap, W, H, hc["normals"], hc["pbr"], tiles, x).astype(np.float32)
    out = {}
    for name, st, nz, tol in (("fp32", abi.STORAGE_FP32, noisy, 2e-3), ("fp16", abi.STORAGE_FP16, noisy.half(), 1e-2)):
        d = ReflectionDenoiser(W, H, storage=st, iterations=iters)
        d.set_camera(*cam)
        got = d.denoise(cur, prev, nz).float().cpu().numpy()
        torch.cuda.synchronize()
        e = oracle.relative_error_report(got[..., :3], x[..., :3])
        e["agree_fraction"] = float(np.all(np.abs(got - x) <= tol + tol * np.abs(x), axis=-1).mean())
        e["agree_tolerance"] = tol
        out[name] = e
        log(f"[bench] svgf {name} vs FP64: normwise {e['normwise_rel']:.2e}, p99 {e['p99_rel']:.2e}, "
            f"agree {e['agree_fraction']:.4f}", quiet)
        del d
    return out


def precision_report(u_full, g, sdf, mip, res, atlas, p, n, first):
    """Error of first-frame GPU atlases against the FP64 oracle on the two lowest probe layers (the
    same probes and rays: a 32x16x2 volume with the full volume's start and spacing). Those layers
    fill the first 2 (side + 2) + 1 rows of both atlases (the atlas border row 0, then a tile row
    of side + 2 per layer). The bound statistic is the normwise relative error ||got - ref|| /
    ||ref|| over the compared texels; per-texel relative errors (max, p99, p99.9) are taken
    against max(|ref|, 1e-2)."""
    import numpy as np

    import oracle
    from luxgi_amd import abi
    counts = (u_full.probeCounts[0], u_full.probeCounts[1], 2)
    u = abi.make_uniform(counts, tuple(u_full.startPosition[:3]), tuple(u_full.step[:3]), u_full.raysPerProbe,
                         irr_oct=8, depth_oct=16, hysteresis=0.98)
    rad, dd, fl, _ = oracle.trace(u, g, sdf, mip, res, (res + 3) // 4, atlas.data, atlas.chunks, atlas.culled,
                                  atlas.objects, atlas.tiles, atlas.depth, atlas.color_f32, p, nthreads=n)
    z_i = np.zeros((u.irradianceTextureHeight, u.irradianceTextureWidth, 4), np.float32)
    z_d = np.zeros((u.depthTextureHeight, u.depthTextureWidth, 2), np.float32)
    ref_i, ref_d = oracle.probe_update(u, rad, dd, z_i, z_d, 1, 7, nthreads=n)
    hi, hd = 2 * (8 + 2) + 1, 2 * (16 + 2) + 1  # border row 0 and the tiles of the two layers
    out = {}
    for name, (gi, gd) in first.items():
        ri = ref_i[:hi, :, :3]
        di = np.abs(gi[:hi, :, :3].astype(np.float64) - ri)
        rel = di / np.maximum(np.abs(ri), 1e-2)
        dd_ = np.abs(gd[:hd].astype(np.float64) - ref_d[:hd])
        ri_rep = oracle.relative_error_report(gi[:hi, :, :3], ri)
        rd_rep = oracle.relative_error_report(gd[:hd], ref_d[:hd])
        out[name] = {"irradiance_normwise_rel": ri_rep["normwise_rel"], "irradiance_p99_rel": ri_rep["p99_rel"],
                     "irradiance_max_rel": float(rel.max()), "irradiance_p999_rel": float(np.quantile(rel, 0.999)),
                     "irradiance_mean_abs": float(di.mean()), "depth_normwise_rel": rd_rep["normwise_rel"],
                     "depth_p99_rel": rd_rep["p99_rel"], "depth_max_abs": float(dd_.max()),
                     "depth_mean_abs": float(dd_.mean()), "texels": int(ri.shape[0] * ri.shape[1]),
                     "rays_compared": int(u.num_probes * u.raysPerProbe)}
    return out


if __name__ == "__main__":
    main()

"""Schedule autotuner: times every valid schedule of a launch on the device and keeps the fastest.

A schedule only changes how work is mapped to the GPU (persistent grid size, unit order, shading
order), never the result, so candidates are compared on time alone; callers that need proof pass
`check`, which compares each candidate's outputs with the first one's. Choices are cached per
device, kernel and problem signature in a JSON file ($LUXGI_AUTOTUNE_CACHE, default
~/.cache/luxgi_amd/autotune.json), so a later process reuses them without timing.
"""
import json
import os

import torch


def _device_key():
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}|{p.multi_processor_count}cu"


def time_ms(run, warmup=2, reps=5):
    """Mean milliseconds of run() over `reps` launches on the current stream (HIP events)."""
    for _ in range(warmup):
        run()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        run()
    ev[1].record()
    torch.cuda.synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


class Autotuner:
    def __init__(self, path=None):
        self.path = path or os.environ.get("LUXGI_AUTOTUNE_CACHE") or os.path.join(
            os.path.expanduser("~"), ".cache", "luxgi_amd", "autotune.json")
        self.cache = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                self.cache = json.load(f)
        self.last_timings = {}

    def save(self):
        os.makedirs(os.path.dirname(self.path), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.cache, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)

    def best(self, kernel, signature, candidates, run, check=None, warmup=2, reps=5, retune=False):
        """The fastest of `candidates` (JSON-able values) for run(candidate). check(candidate) -> bool,
        if given, is called after each candidate's timing and must hold."""
        key = f"{_device_key()}|{kernel}|{signature}"
        if not retune and key in self.cache and self.cache[key]["choice"] in candidates:
            return self.cache[key]["choice"]
        timings = {}
        for c in candidates:
            timings[json.dumps(c)] = time_ms(lambda: run(c), warmup, reps)
            if check is not None and not check(c):
                raise RuntimeError(f"{kernel}: schedule {c} changed the result")
        choice = min(candidates, key=lambda c: timings[json.dumps(c)])
        self.cache[key] = {"choice": choice, "ms": timings}
        self.last_timings = timings
        self.save()
        return choice


def svgf_batch_candidates(batch):
    """Persistent grids of 1, 2, 4 and 8 blocks per compute unit and of one block per unit of work,
    each with the strided and the contiguous walk, and the loop-free one-unit-per-block launch."""
    cu = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    total = int(batch.header.totalBlocks) or sum(((w + 15) // 16) * ((h + 15) // 16) for w, h in batch.sizes)
    grids = sorted({cu, 2 * cu, 4 * cu, 8 * cu, max(total, 1)})
    return [[g, o] for g in grids for o in (0, 1)] + [[max(total, 1), 2]]


def tune_svgf_batch(batch, tuner=None, check=True):
    """Chooses (gridBlocks, order) for a ReflectionDenoiseBatch whose frames are set, leaves the batch
    on the choice and its temporal state as it was. Returns the choice."""
    tuner = tuner or Autotuner()
    batch.upload()
    saved = batch.state()
    cands = svgf_batch_candidates(batch)
    sig = f"{batch.storage}|{batch.iterations}|" + ",".join(f"{w}x{h}" for w, h in batch.sizes)
    ref = []

    def run(c):
        batch.restore(saved)
        batch.grid_blocks, batch.order = int(c[0]), int(c[1])
        batch.launch()

    def same(c):
        run(c)
        torch.cuda.synchronize()
        out = [t.clone() for t in batch.results()]
        if not ref:
            ref.extend(out)
            return True
        return all(torch.equal(a, b) for a, b in zip(out, ref))

    choice = tuner.best("svgf_reflection_denoise_batched", sig, cands, run, check=same if check else None)
    batch.grid_blocks, batch.order = int(choice[0]), int(choice[1])
    batch.restore(saved)
    torch.cuda.synchronize()
    return choice


def tune_trace_shade_order(volume, scene, params, tuner=None):
    """Chooses the trace's shading order (0: hits sorted by surface-atlas chunk, 1: ray order) for a
    DDGIVolume and scene; both give identical ray data. Sets params.shadeOrder and returns it."""
    tuner = tuner or Autotuner()
    u = volume.u
    sig = f"{u.probeCounts[0]}x{u.probeCounts[1]}x{u.probeCounts[2]}x{u.raysPerProbe}|{scene.n_objects}"

    def run(c):
        params.shadeOrder = int(c)
        volume.trace_rays(scene, params)

    choice = tuner.best("ddgi_trace_sdf.shadeOrder", sig, [0, 1], run)
    params.shadeOrder = int(choice)
    return choice

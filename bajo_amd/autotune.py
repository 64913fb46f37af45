"""Launch-schedule autotuner for wavefront renders.

bajo has no counterpart: its renders use one fixed chunk of paths in flight
(GPU_RT_DEFAULT_CAPACITY, rt/gpu/resources.mojo:24).  Here the number of paths
in flight per chunk (`RenderSettings.path_capacity`) is the schedule knob.
Whole-frame chunks give every launch the most parallel work. Smaller chunks
keep each bounce's working set (the environments and BVH nodes that the
chunk's paths touch) closer to the 256 MB Infinity Cache and shrink the path
arena.  Images do not depend on the chunking: every path draws its own random
streams and owns its sample slot, so the tuner only measures time.

`tune_render` times each candidate with HIP events on the context's stream
(one untimed render first, which also sizes the path arena) and returns the
fastest settings. Results are cached per scene object (a weak reference, so
the entry goes with the scene) and, within it, per (native handle, build
count, image size, spp, depth, algorithm): a rebuilt scene is timed again.
"""

from __future__ import annotations

import weakref
from dataclasses import replace

import torch

_CACHE: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()


def default_candidates(total_paths: int, spp: int) -> list[int]:
    """Whole pixels per chunk: the whole frame, then halvings down to 1 M paths (at least 2 chunks apart)."""
    out = [0]
    cap = 1 << 20
    while cap < total_paths:
        out.append((cap // spp) * spp)
        cap <<= 1
    return out


def _time(fn, stream, reps: int) -> float:
    fn()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(reps):
        fn()
    ev[1].record(stream)
    ev[1].synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def tune_render(settings, camera, scene, algorithm: int, candidates: list[int] | None = None, reps: int = 2,
                pixels: torch.Tensor | None = None) -> tuple:
    """Times `enqueue_render_gpu` over `candidates` path capacities (0 = the whole frame) and returns
    (best settings, {capacity: milliseconds}).  The render output goes to `pixels` (allocated if None)."""
    from .rt import enqueue_render_gpu

    envs = getattr(scene, "env_count", 1)
    total = settings.image_width * settings.image_height * settings.samples_per_pixel * envs
    key = (scene.handle.value, getattr(scene, "builds", 0), scene.triangle_count, envs, settings.image_width,
           settings.image_height, settings.samples_per_pixel, settings.max_depth, algorithm)
    per_scene = _CACHE.setdefault(scene, {})
    if key in per_scene and candidates is None:
        best, table = per_scene[key]
        return replace(settings, path_capacity=best), dict(table)
    cands = candidates if candidates is not None else default_candidates(total, settings.samples_per_pixel)
    stream = scene.ctx.stream
    out = pixels
    table = {}
    for cap in cands:
        st = replace(settings, path_capacity=int(cap))
        if out is None:
            out = enqueue_render_gpu(st, camera, scene, algorithm).pixels
        table[int(cap)] = _time(lambda: enqueue_render_gpu(st, camera, scene, algorithm, pixels=out), stream, reps)
    best = min(table, key=table.get)
    if candidates is None:
        per_scene[key] = (best, dict(table))
    return replace(settings, path_capacity=best), table

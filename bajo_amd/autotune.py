"""Launch-schedule autotuner for wavefront renders.

bajo has no counterpart: its renders use one fixed chunk of paths in flight
(GPU_RT_DEFAULT_CAPACITY, rt/gpu/resources.mojo:24) and fixed launch shapes.
Here six launch knobs of `RenderSettings` form the schedule:

- ``extend_waves``: the occupancy the bounce >= 1 extension kernel is compiled
  for (4, 6 or 8 waves per SIMD), with its LDS traversal stack sized to match
  (32, 24 or 16 entries per lane; deeper entries spill to device memory).
  Fewer waves give each lane more VGPRs and a deeper on-chip stack; more
  waves hide more memory latency.  The best point depends on how incoherent
  the scene's rays are.
- ``shadow_group``: lanes per shadow ray on dense bounces (1, 8, 16 or 32).
  A group of G lanes takes G/8 nodes off its stack per step and tests all the
  entered leaves' records at once, so a ray's serial chain shortens while
  fewer rays are in flight; sparse bounces always get a wave per ray.
- ``shadow_walk``: with one lane per ray, the any-hit walk: one node per
  step or two (the stack top loaded with the current node), leaf records one
  ahead or four at a time.  Shadow rays leave a surface, so their walks are
  long serial chains; the forms trade registers (occupancy) for memory-level
  parallelism within a ray.
- ``tail_group``: lanes per path in the tail kernel (1, 16, 32 or 64).  A
  lane group walks a path's extension ray (exactly as one lane would decide
  it) and its shadow ray a node group at a time (one node per eight lanes);
  one lane per path keeps more paths in flight.
- ``tail_paths``: the live-path count at which the tail kernel (one launch
  that finishes a chunk's remaining bounces with paired path / shadow lanes)
  takes over from the per-bounce launches.  Earlier saves launches on long
  thin tails; later keeps the wide per-bounce kernels busy.
- ``path_capacity``: paths in flight per chunk.  Whole-frame chunks give every
  launch the most parallel work; smaller chunks keep each bounce's working set
  (the environments and BVH nodes its paths touch) closer to the 256 MB
  Infinity Cache and shrink the path arena.

Images do not depend on the schedule: every path draws its own random streams
and owns its sample slot, and the tail keeps the wavefront's order, so the
tuner only measures time (the tests check the images stay bit-identical).

`tune_render` is a coordinate search: starting from the defaults it times
every candidate of one knob with the others at their best so far, keeps the
fastest, and moves to the next knob.  Each candidate is timed with HIP events
on the context's stream (one untimed render first, which also sizes the path
arena).  Results are cached per scene object (a weak reference, so the entry
goes with the scene) and, within it, per device and per (native handle, build
count, image size, spp, depth, algorithm): a rebuilt scene is timed again.
"""

from __future__ import annotations

import weakref
from dataclasses import replace

import torch

_CACHE: "weakref.WeakKeyDictionary" = weakref.WeakKeyDictionary()

KNOBS = ("extend_waves", "shadow_group", "shadow_walk", "tail_group", "tail_paths", "path_capacity")


def default_candidates(total_paths: int, spp: int) -> list[int]:
    """Whole pixels per chunk: the whole frame, then halvings down to 1 M paths (at least 2 chunks apart)."""
    out = [0]
    cap = 1 << 20
    while cap < total_paths:
        out.append((cap // spp) * spp)
        cap <<= 1
    return out


def default_space(total_paths: int, spp: int) -> dict[str, list[int]]:
    """Candidate values per knob (0 = the library's default for that knob)."""
    return {
        "extend_waves": [0, 4, 8],
        "shadow_group": [0, 1, 8, 32],
        "shadow_walk": [0, 1, 2, 4],
        "tail_group": [0, 1, 16, 64],
        "tail_paths": [0, 1 << 15, 1 << 17, 1 << 19, 1 << 21],
        "path_capacity": default_candidates(total_paths, spp),
    }


def _time(fn, stream, reps: int) -> float:
    fn()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(reps):
        fn()
    ev[1].record(stream)
    ev[1].synchronize()
    return ev[0].elapsed_time(ev[1]) / reps


def _device_key(ctx) -> tuple:
    return (int(ctx.device), torch.cuda.get_device_name(ctx.torch_device))


def tune_render(settings, camera, scene, algorithm: int, candidates: list[int] | None = None, reps: int = 2,
                pixels: torch.Tensor | None = None, space: dict[str, list[int]] | None = None) -> tuple:
    """Tunes the render schedule of `enqueue_render_gpu(settings, camera, scene, algorithm)`.

    Returns (best settings, {(knob, value): milliseconds}).  `space` maps knobs to candidate values
    (default: `default_space`); `candidates` alone restricts the search to those path capacities.  The
    render output goes to `pixels` (allocated if None).  Only full default searches are cached."""
    from .rt import enqueue_render_gpu

    envs = getattr(scene, "env_count", 1)
    total = settings.image_width * settings.image_height * settings.samples_per_pixel * envs
    cacheable = candidates is None and space is None
    if candidates is not None:
        space = {"path_capacity": list(candidates)}
    elif space is None:
        space = default_space(total, settings.samples_per_pixel)
    key = (_device_key(scene.ctx), scene.handle.value, getattr(scene, "builds", 0), scene.triangle_count, envs,
           settings.image_width, settings.image_height, settings.samples_per_pixel, settings.max_depth, algorithm)
    per_scene = _CACHE.setdefault(scene, {})
    if cacheable and key in per_scene:
        best, table = per_scene[key]
        return replace(settings, **best), dict(table)
    stream = scene.ctx.stream
    out = pixels
    best = {k: getattr(settings, k) for k in KNOBS if k in space}
    table: dict[tuple[str, int], float] = {}
    for knob in KNOBS:
        if knob not in space:
            continue
        times = {}
        for value in space[knob]:
            st = replace(settings, **{**best, knob: int(value)})
            if out is None:
                out = enqueue_render_gpu(st, camera, scene, algorithm).pixels
            times[int(value)] = _time(lambda: enqueue_render_gpu(st, camera, scene, algorithm, pixels=out), stream,
                                      reps)
            table[(knob, int(value))] = times[int(value)]
        best[knob] = min(times, key=times.get)
    if cacheable:
        per_scene[key] = (dict(best), dict(table))
    return replace(settings, **best), table

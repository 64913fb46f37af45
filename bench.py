#!/usr/bin/env python3
"""Benchmark of the MI355X BVH / ray-tracing hot path.

Workload (one GPU): a displaced geodesic icosphere of 4,194,304 FP32
triangles (frequency 458, radial value-noise amplitude 0.05).  One step =
LBVH build (bounds, Morton, onesweep sort, Karras, refit) followed by one
batch of 1920 x 1080 x 8 = 16,588,800 incoherent closest-hit rays (origins
uniform in the scene box, directions uniform on the sphere).  Inputs are
resident in HBM before timing starts.

Prints one JSON line with whole-step MRays/s as `value`, the build Mtris/s
and sort GB/s beside it, the roofline of the dominant kernel (traversal,
measured live with HIP events) and the CPU oracle timed on the host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MRays/s incoherent @4M tris; build Mtris/s; equal-time MSE guided vs plain"
N_TRIS = 4_194_304
WIDTH, HEIGHT, SPP = 1920, 1080, 8
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 64  # both child boxes + children, one node record
TRI_BYTES = 48  # v0, e1, e2, prim
RAY_IO_BYTES = 64  # 32 B packed ray in + 32 B hit out
BUILD_BYTES_PER_TRI = 452  # minimal binary-LBVH traffic per triangle
SORT_BYTES_PER_KEY_PASS = 16  # 8 B read + 8 B write of a 64-bit key


def incoherent_rays(n: int, lo: np.ndarray, hi: np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rays = np.empty((8, n), np.float32)
    for k in range(3):
        rays[k] = lo[k] + (hi[k] - lo[k]) * rng.random(n, dtype=np.float32)
    d = rng.standard_normal((3, n), dtype=np.float32)
    d /= np.linalg.norm(d, axis=0, keepdims=True)
    rays[3] = 1e-4
    rays[4:7] = d
    rays[7] = np.finfo(np.float32).max
    return rays.reshape(-1)


def cpu_baseline(vertices: np.ndarray, rays: np.ndarray, sample: int, threads: int) -> dict:
    import oracle

    t0 = time.perf_counter()
    lbvh = oracle.build_lbvh(vertices)
    build_s = time.perf_counter() - t0
    sub = rays.reshape(8, -1)[:, :sample].reshape(-1)
    t0 = time.perf_counter()
    oracle.trace_bvh2(vertices, lbvh, sub, mode=0, threads=threads)
    trace_s = time.perf_counter() - t0
    return {
        "value": sample / trace_s / 1e6,
        "unit": "MRays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{sample} of the step's incoherent rays traced by the C oracle BVH2 tracer on {threads} threads; "
        f"oracle LBVH build of all {len(vertices) // 9} triangles on 1 thread took {build_s:.2f} s "
        f"({len(vertices) // 9 / build_s / 1e6:.2f} Mtris/s)",
        "build_mtris_s_1core": len(vertices) // 9 / build_s / 1e6,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument(
        "--copy-calibration-mib",
        type=int,
        default=0,
        help="after the run, copy a buffer of this many MiB once (scales rocprofv3 FETCH_SIZE/WRITE_SIZE)",
    )
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: this build has no multi-GPU path (a replicated BVH with "
            "ray ranges split per GPU is described in DESIGN.md). Run with --gpus 1.",
            file=sys.stderr,
        )
        return 2

    import torch

    from bajo_amd import scenes
    from bajo_amd.bvh import TRACE, build_triangle_bvh
    from bajo_amd.context import default_context

    ctx = default_context(0)
    dev = ctx.torch_device
    print(f"[bench] device {ctx.name()}", file=sys.stderr, flush=True)

    vertices = scenes.icosphere(N_TRIS)
    lo, hi = scenes.bounds(vertices)
    R = WIDTH * HEIGHT * SPP
    rays = incoherent_rays(R, lo, hi, seed=2026)
    d_vertices = torch.from_numpy(vertices).to(dev)
    d_rays = torch.from_numpy(rays).to(dev)
    d_hits = torch.empty(8 * R, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    print("[bench] inputs resident", file=sys.stderr, flush=True)

    bvh = build_triangle_bvh(d_vertices, ctx=ctx)
    stream = torch.cuda.current_stream(dev)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        bvh.rebuild(d_vertices)
        if ev is not None:
            ev[1].record(stream)
        bvh.launch_rays(d_rays, d_hits, R, TRACE.CLOSEST_HIT)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    build_ms = np.array([e[0].elapsed_time(e[1]) for e in events])
    trace_ms = np.array([e[1].elapsed_time(e[2]) for e in events])
    ms_per_step = wall * 1e3 / args.steps
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    # per-stage build breakdown (one extra measured build, outside the timed region)
    stages = bvh.rebuild(d_vertices, timings=True)
    # traversal counters from the instrumented kernel over the same rays
    stats = torch.empty(5 * R, dtype=torch.int32, device=dev)
    bvh.launch_rays(d_rays, d_hits, R, TRACE.CLOSEST_HIT, stats=stats)
    st = stats.view(-1, 5).to(torch.float64)
    visits = float(st[:, 0].mean())
    leaf_tests = float(st[:, 3].mean())
    max_stack = int(stats.view(-1, 5)[:, 4].max())
    hit_frac = float((d_hits.view(torch.int32).view(-1, 8)[:, 2] != -1).float().mean())

    bytes_per_ray = RAY_IO_BYTES + visits * NODE_BYTES + leaf_tests * TRI_BYTES
    trace_avg_s = float(trace_ms.mean()) / 1e3
    achieved = bytes_per_ray * R / trace_avg_s / 1e9
    build_s = float(build_ms.mean()) / 1e3
    sort_s = stages["sort_ns"] / 1e9

    line = {
        "metric": METRIC,
        "value": R / (ms_per_step / 1e3) / 1e6,
        "unit": "MRays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "icosphere-4M: LBVH build + 1920x1080x8 incoherent closest-hit rays per step",
            "triangles": N_TRIS,
            "rays_per_step": R,
            "rays": "origins U(scene box), directions U(S^2), t in [1e-4, FLT_MAX]",
            "bvh": "binary LBVH (Karras), 64 B nodes, 48 B sorted triangles",
        },
        "trace_mrays_s": R / trace_avg_s / 1e6,
        "trace_ms": float(trace_ms.mean()),
        "build_ms": float(build_ms.mean()),
        "build_mtris_s": N_TRIS / build_s / 1e6,
        "build_stages_ms": {k: v / 1e6 for k, v in stages.items()},
        "sort_gbs": N_TRIS * 4 * SORT_BYTES_PER_KEY_PASS / sort_s / 1e9,
        "build_gbs_algorithmic": N_TRIS * BUILD_BYTES_PER_TRI / build_s / 1e9,
        "traversal": {"node_visits": visits, "leaf_tests": leaf_tests, "max_stack": max_stack, "hit_fraction": hit_frac},
        "roofline": {
            "kernel": "trace_kernel<closest-hit>",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "bytes_per_ray": bytes_per_ray,
        },
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        line["cpu_baseline"] = cpu_baseline(vertices, rays, min(args.cpu_sample, R), threads)
    print(json.dumps(line), flush=True)
    bvh.free()
    if args.copy_calibration_mib:
        nbytes = args.copy_calibration_mib << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] calibration copy of {nbytes} bytes done", file=sys.stderr, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

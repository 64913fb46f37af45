"""Parity at the benchmark's full sizes.

- The 4M-triangle icosphere frame exactly as the benchmark renders it
  (1920x1080x8 spp, depth 5, seed 2026, 8-wide LBVH with leaves of 4): NEE and
  PATH images against the oracle's render of the same scene and streams
  (RMSE <= 1e-3, > 99 % bit-identical pixels, equal ray counts).
- The 16.7M-triangle soup grid (64 x 64 copies): the H-PLOC binary layout word
  for word against the oracle's H-PLOC, a second build bit-identical to the
  first (the cross-XCD hand-offs of the clustering and the refit), and
  closest hits of 200,000 incoherent rays through the device's 8-wide
  hierarchy bit-identical to the oracle's trace of its own collapse of the
  same binary layout.
- A render target's HIP-graph replay against the oracle's render.
"""

import os

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def image_parity(got, want):
    got = np.asarray(got, np.float32).reshape(-1)
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == np.asarray(want, np.float32).view(np.uint32)))
    return rmse, exact


@pytest.fixture(scope="module")
def icosphere():
    from bajo_amd import scenes

    return scenes.icosphere_scene(4_194_304)


@pytest.mark.parametrize("alg", ["NEE", "PATH"])
def test_icosphere_full_frame_matches_oracle(gpu_ctx, icosphere, alg):
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import RENDER, GpuScene, RenderSettings, enqueue_render_gpu

    announce(f"icosphere 4M full frame {alg}")
    algorithm = getattr(RENDER, alg)
    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.41421356).flatten()
    g = GpuScene(icosphere, gpu_ctx, node_width=8, leaf_size=4)
    res = enqueue_render_gpu(RenderSettings(1920, 1080, 8, 2026, 5), cam, g, algorithm, with_stats=True)
    got = res.numpy()
    want, rays = oracle.render(icosphere.vertices, icosphere.surfaces, icosphere.lambertian, icosphere.emissive, 1920,
                               1080, 8, 5, 2026, algorithm, cam, THREADS, node_width=8, leaf_size=4)
    rmse, exact = image_parity(got, want)
    print(f"{alg}: rmse {rmse:.3e}, bit-identical pixels {exact:.5f}")
    assert rmse <= 1e-3 and exact > 0.99
    assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
    assert res.stats["status"] == 0 and res.stats["traversal_status"] == 0
    g.free()


def test_soup_grid_full_size_hploc_and_hits(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import TRACE, build_triangle_bvh

    v = scenes.soup_grid(64, 64, 4096)
    n = v.size // 9
    announce(f"soup grid 64x64x4096 n={n}: oracle H-PLOC")
    want = oracle.build_hploc(v)
    announce("soup grid full size: device H-PLOC + 8-wide build")
    vd = to_dev(v)
    bvh = build_triangle_bvh(vd, node_width=8, leaf_size=4, method=1)
    b = bvh.binary()
    assert b["root"] == want["root"]
    for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
        np.testing.assert_array_equal(b[key], want[key], err_msg=key)
    np.testing.assert_array_equal(b["node_bounds"].view(np.uint32), want["node_bounds"].view(np.uint32))
    wide = bvh.wide()
    bvh.rebuild(vd)
    b2 = bvh.binary()
    for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
        np.testing.assert_array_equal(b2[key], b[key], err_msg=f"second build {key}")
    np.testing.assert_array_equal(b2["node_bounds"].view(np.uint32), b["node_bounds"].view(np.uint32))
    R = 200_000
    rng = np.random.default_rng(44)
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    o = lo + (hi - lo) * rng.random((R, 3))
    d = rng.normal(size=(R, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((8, R), np.float32)
    rays[0:3], rays[3], rays[4:7], rays[7] = o.T, 0.0, d.T, np.finfo(np.float32).max
    rays = rays.reshape(-1)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    announce(f"soup grid full size: {R} closest-hit rays")
    bvh.launch_rays(to_dev(rays), hits, R, TRACE.CLOSEST_HIT)
    got = to_host(hits, np.float32)
    mine, _, over = oracle.trace_wide(v, b["leaf_ids"], wide, rays)
    assert over == 0
    np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 8), mine.view(np.uint32))
    theirs, _, over = oracle.trace_wide(v, want["leaf_ids"], oracle.collapse_wide(want, 4), rays)
    assert over == 0
    np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 8), theirs.view(np.uint32))
    hit = float(np.mean(oracle.hit_prims(got) != 0xFFFFFFFF))
    print(f"hit fraction {hit:.3f}")
    assert 0.05 < hit < 0.999
    bvh.free()


def test_render_target_replay_matches_oracle(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import RENDER, GpuRtRenderTarget, GpuScene, RenderSettings

    announce("render target replay vs oracle")
    sc = scenes.material_box_scene()
    g = GpuScene(sc, gpu_ctx, node_width=8, leaf_size=4)
    st = RenderSettings(64, 48, 8, 31, 6)
    ca = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    cb = Camera((0.4, 1.2, -2.8), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    t = GpuRtRenderTarget(st, ca, RENDER.MIS)
    t.capture(g)
    t.update_camera(cb)
    t.pixels.zero_()
    got = t.replay()
    torch.cuda.synchronize()
    want, _ = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 64, 48, 8, 6, 31, RENDER.MIS, cb,
                            THREADS, node_width=8, leaf_size=4, metal=sc.metal, dielectric=sc.dielectric)
    rmse, exact = image_parity(got.cpu().numpy(), want)
    print(f"replay: rmse {rmse:.3e}, bit-identical pixels {exact:.5f}")
    assert rmse <= 1e-3 and exact > 0.99
    t.free()
    g.free()

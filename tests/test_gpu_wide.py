"""Device 8-wide BVH vs the oracle's collapse and traversal (oracle/orc_wide.c).

bajo's wide collapse numbers nodes in the order its waves reserve them, so
its tests compare only structure and traversal results
(test/bvh/gpu/test_hploc_wide.mojo).  Here the device hierarchy must equal
the oracle's byte for byte once both are put in the same depth-first order
(quantized boxes, exponents, leaf ranges), and the device traversal must
return the oracle's 8-wide hits and counters bit for bit; against the binary
traversal the hits may differ only on near-ties at shared edges and faces.
"""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host
from test_gpu_trace import random_rays
from wide_check import compare_hits, validate_wide

pytestmark = pytest.mark.gpu


def soup(n, seed):
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3), dtype=np.float32)
    e = (rng.random((n, 2, 3), dtype=np.float32) - 0.5) * 0.05
    return np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)


def build_both(v, leaf):
    from bajo_amd.bvh import build_triangle_bvh

    t = {}
    bvh = build_triangle_bvh(to_dev(v), timings=t, node_width=8, leaf_size=leaf)
    lb = oracle.build_lbvh(v)
    return bvh, lb, oracle.collapse_wide(lb, leaf), t


@pytest.mark.parametrize("n,leaf", [(1, 4), (2, 4), (3, 1), (100, 4), (4096, 1), (4096, 4), (4096, 8), (50_000, 4)])
def test_wide_structure_equals_oracle(gpu_ctx, n, leaf):
    announce(f"wide build n={n} leaf={leaf}")
    v = soup(n, seed=n + leaf)
    bvh, lb, want, t = build_both(v, leaf)
    got = bvh.wide()
    assert bvh.info().node_width == 8 and bvh.info().wide_node_count == len(got)
    assert len(got) == len(want)
    validate_wide(got, v, lb["leaf_ids"], leaf)
    a, b = oracle.canonical_wide(got), oracle.canonical_wide(want)
    assert a.tobytes() == b.tobytes()
    assert t["collapse_ns"] > 0 and t["total_ns"] >= t["collapse_ns"]
    bvh.free()


@pytest.mark.parametrize("n,leaf,seed", [(1, 4, 0), (2, 4, 1), (500, 4, 2), (65536, 4, 3), (65536, 1, 4), (65536, 8, 5)])
def test_wide_trace_equals_oracle(gpu_ctx, n, leaf, seed):
    from bajo_amd.bvh import TRACE

    announce(f"wide trace n={n} leaf={leaf}")
    v = soup(n, seed)
    bvh, lb, w, _ = build_both(v, leaf)
    R = 100_000
    rays = random_rays(R, 0.0, 1.0, seed + 10, t_min=1e-4)
    rays.reshape(8, -1)[7, ::3] = 0.05  # finite t_max on every third ray
    d_rays = to_dev(rays)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    bvh.launch_rays(d_rays, hits, R, TRACE.CLOSEST_HIT, stats=stats)
    want, want_stats, over = oracle.trace_wide(v, lb["leaf_ids"], w, rays, mode=0, with_stats=True)
    assert over == 0
    got = to_host(hits, np.float32).reshape(-1, 8)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    st = to_host(stats, np.uint32).reshape(-1, 5)
    np.testing.assert_array_equal(st[:, [0, 1, 3, 4]], want_stats[:, [0, 1, 3, 4]])
    # against the binary traversal: near-ties only
    compare_hits(got, oracle.trace_bvh2(v, lb, rays, mode=0), min_identical=0.998)
    bvh.launch_rays(d_rays, hits, R, TRACE.ANY_HIT)
    want_any, _, _ = oracle.trace_wide(v, lb["leaf_ids"], w, rays, mode=1)
    np.testing.assert_array_equal(to_host(hits, np.uint32).reshape(-1, 8), want_any.view(np.uint32))
    bvh.free()


def test_wide_camera_icosphere_4m(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera, build_triangle_bvh

    n = 4_194_304
    announce(f"wide camera icosphere n={n}")
    v = scenes.icosphere(n)
    t = {}
    bvh = build_triangle_bvh(to_dev(v), timings=t, node_width=8, leaf_size=4)
    print("build", {k: val / 1e6 for k, val in t.items()}, "wide nodes", bvh.info().wide_node_count)
    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.41421356).flatten()
    w, h = 1920, 1080
    R = w * h
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    bvh.launch_camera(to_dev(cam), hits, R, w, h, stats=stats)
    lb = oracle.build_lbvh(v)
    ow = oracle.collapse_wide(lb, 4)
    assert len(ow) == bvh.info().wide_node_count
    sample = np.arange(0, R, 13)
    rays = oracle.camera_rays(cam, R, w, h).reshape(8, -1)[:, sample].reshape(-1)
    want, want_stats, over = oracle.trace_wide(v, lb["leaf_ids"], ow, rays, with_stats=True)
    assert over == 0
    got = to_host(hits, np.float32).reshape(-1, 8)[sample]
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    st = to_host(stats, np.uint32).reshape(-1, 5)[sample]
    np.testing.assert_array_equal(st[:, [0, 1, 3, 4]], want_stats[:, [0, 1, 3, 4]])
    info = compare_hits(got, oracle.trace_bvh2(v, lb, rays, mode=0, threads=8), min_identical=0.999)
    print(info, "visits/ray", st[:, 0].mean(), "tests/ray", st[:, 3].mean(), "max stack", st[:, 4].max())
    bvh.free()


def test_wide_rebuild_switches_layout(gpu_ctx):
    """A handle rebuilt in place keeps its width; binary and wide handles over
    the same vertices trace the same scene."""
    from bajo_amd.bvh import TRACE, build_triangle_bvh

    announce("wide rebuild")
    v = soup(3000, 9)
    dv = to_dev(v)
    wide = build_triangle_bvh(dv, node_width=8)
    binary = build_triangle_bvh(dv)
    v2 = soup(5000, 10)
    wide.rebuild(to_dev(v2))
    assert wide.info().node_width == 8 and wide.prim_count == 5000
    R = 20_000
    rays = random_rays(R, 0.0, 1.0, 3, t_min=1e-4)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    wide.launch_rays(to_dev(rays), hits, R, TRACE.CLOSEST_HIT)
    lb = oracle.build_lbvh(v2)
    want, _, _ = oracle.trace_wide(v2, lb["leaf_ids"], oracle.collapse_wide(lb, 4), rays)
    np.testing.assert_array_equal(to_host(hits, np.uint32).reshape(-1, 8), want.view(np.uint32))
    binary.launch_rays(to_dev(rays[: 8 * R]), hits, R, TRACE.CLOSEST_HIT)
    wide.free()
    binary.free()


def test_wide_rejects_bad_options(gpu_ctx):
    from bajo_amd._lib import BajoError
    from bajo_amd.bvh import build_triangle_bvh

    dv = to_dev(soup(10, 1))
    for kw in ({"node_width": 3}, {"node_width": 16}, {"node_width": 8, "leaf_size": 9}, {"node_width": 8, "leaf_size": -1}):
        with pytest.raises(BajoError):
            build_triangle_bvh(dv, **kw)
    binary = build_triangle_bvh(dv)
    with pytest.raises(BajoError):
        binary.wide()
    binary.free()


ALGS = {"PATH": 0, "NORMALS": 1, "AO": 2, "NEE": 3, "MIS": 4}


@pytest.mark.parametrize("name", ["PATH", "NEE", "MIS", "AO"])
def test_wide_render_box_matches_oracle(gpu_ctx, name):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce(f"wide render box {name}")
    sc = scenes.box_scene()
    g = GpuScene(sc, gpu_ctx, node_width=8)
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    w, h, spp, depth, seed = 40, 30, 8, 5, 2026
    res = enqueue_render_gpu(RenderSettings(w, h, spp, seed, depth), cam, g, ALGS[name], with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(
        sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed, ALGS[name], cam, 8,
        node_width=8, leaf_size=4,
    )
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    print(f"{name}: rmse {rmse:.3e}, bit-identical pixels {exact:.4f}")
    assert rmse <= 1e-3 and exact > 0.99
    assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
    assert res.stats["traversal_status"] == 0
    g.free()


def test_wide_render_icosphere_matches_oracle(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    n = 1 << 16
    announce(f"wide render icosphere n={n} NEE")
    sc = scenes.icosphere_scene(n, frequency=58)
    g = GpuScene(sc, gpu_ctx, node_width=8)
    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.41421356).flatten()
    w, h, spp, depth, seed = 48, 27, 4, 5, 2026
    res = enqueue_render_gpu(RenderSettings(w, h, spp, seed, depth), cam, g, ALGS["NEE"], with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(
        sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed, ALGS["NEE"], cam, 8,
        node_width=8, leaf_size=4,
    )
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    print(f"icosphere wide NEE: rmse {rmse:.3e}, identical {exact:.4f}")
    assert rmse <= 1e-3 and exact > 0.99
    assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
    g.free()


@pytest.mark.parametrize("n,leaf", [(1, 4), (2, 4), (3, 1), (100, 4), (4096, 4), (50_000, 4), (50_000, 1)])
def test_wide4_structure_and_hits_equal_oracle(gpu_ctx, n, leaf):
    """Width 4 (bajo's render_gpu default W): the device collapse equals the oracle's 4-wide collapse in
    canonical order, no node holds more than 4 children, and closest / any hits equal the oracle's."""
    from bajo_amd.bvh import TRACE, build_triangle_bvh

    announce(f"4-wide build n={n} leaf={leaf}")
    v = soup(n, seed=7 * n + leaf)
    bvh = build_triangle_bvh(to_dev(v), node_width=4, leaf_size=leaf)
    lb = oracle.build_lbvh(v)
    want = oracle.collapse_wide(lb, leaf, width=4)
    got = bvh.wide()
    assert bvh.info().node_width == 4 and len(got) == len(want)
    assert got["count"].max() <= 4
    validate_wide(got, v, lb["leaf_ids"], leaf)
    assert oracle.canonical_wide(got).tobytes() == oracle.canonical_wide(want).tobytes()
    R = 20_000
    rays = random_rays(R, -0.1, 1.1, n + 3)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    for mode in (TRACE.CLOSEST_HIT, TRACE.ANY_HIT):
        bvh.launch_rays(to_dev(rays), hits, R, mode)
        w, _, over = oracle.trace_wide(v, lb["leaf_ids"], got, rays, mode=mode)
        assert over == 0
        np.testing.assert_array_equal(to_host(hits, np.uint32).reshape(-1, 8), w.view(np.uint32))
    bvh.free()


@pytest.mark.parametrize("name", ["NEE", "MIS"])
def test_wide4_render_matches_oracle(gpu_ctx, name):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce(f"4-wide render {name}")
    sc = scenes.material_box_scene()
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    g = GpuScene(sc, gpu_ctx, node_width=4)
    res = enqueue_render_gpu(RenderSettings(40, 30, 8, 2027, 6), cam, g, ALGS[name], with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 40, 30, 8, 6, 2027, ALGS[name],
                               cam, 8, node_width=4, leaf_size=4, metal=sc.metal, dielectric=sc.dielectric)
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    assert rmse <= 1e-3 and exact > 0.99
    assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
    g.free()

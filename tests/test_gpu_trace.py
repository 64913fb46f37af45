"""GPU BVH2 traversal vs the CPU oracle: closest-hit prim ids equal except on
exact-t ties, relative t error <= 1e-6 (in practice bit-exact), any-hit
occlusion equal, traversal counters equal (same near-first order)."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu
F32_MAX = np.finfo(np.float32).max


def random_rays(n, lo, hi, seed, t_min=0.0):
    rng = np.random.default_rng(seed)
    o = lo + (hi - lo) * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((8, n), np.float32)
    rays[0:3] = o.T
    rays[3] = t_min
    rays[4:7] = d.T
    rays[7] = F32_MAX
    return rays.reshape(-1)


def compare_hits(got, want, rtol=1e-6):
    g, w = got.reshape(-1, 8), want.reshape(-1, 8)
    gp, wp = g[:, 2].view(np.uint32), w[:, 2].view(np.uint32)
    gt, wt = g[:, 7], w[:, 7]
    same_t = np.abs(gt - wt) <= rtol * np.maximum(np.abs(wt), 1e-30)
    # prim may differ only where t ties exactly
    bad = (gp != wp) & ~(gt == wt)
    assert not bad.any(), f"{bad.sum()} prim mismatches, first rows {np.nonzero(bad)[0][:8]}"
    assert same_t.all(), f"{(~same_t).sum()} t mismatches"
    hit = (gp == wp) & (wp != 0xFFFFFFFF)
    np.testing.assert_array_equal(g[hit][:, [0, 1, 4, 5, 6]], w[hit][:, [0, 1, 4, 5, 6]])
    return hit.mean()


def test_camera_soup_65536(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera, build_triangle_bvh

    announce("trace camera soup 65536 1024x1024")
    v = scenes.soup(65536, initstate=2027, initseq=54)
    bvh = build_triangle_bvh(to_dev(v))
    cam = Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()
    w = h = 1024
    R = w * h
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    bvh.launch_camera(to_dev(cam), hits, R, w, h, stats=stats)
    lb = oracle.build_lbvh(v)
    rays = oracle.camera_rays(cam, R, w, h)
    want, want_stats = oracle.trace_bvh2(v, lb, rays, stats=True)
    frac = compare_hits(to_host(hits, np.float32), want)
    assert frac > 0.3
    st = to_host(stats, np.uint32).reshape(-1, 5)
    ws = want_stats.reshape(-1, 5)
    np.testing.assert_array_equal(st[:, [0, 1, 3, 4]], ws[:, [0, 1, 3, 4]])


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (500, 2), (65536, 3)])
def test_incoherent_rays_closest_and_any(gpu_ctx, n, seed):
    from bajo_amd.bvh import TRACE, build_triangle_bvh

    announce(f"trace incoherent n={n}")
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3), dtype=np.float32)
    e = (rng.random((n, 2, 3), dtype=np.float32) - 0.5) * 0.05
    v = np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)
    bvh = build_triangle_bvh(to_dev(v))
    R = 200_000
    rays = random_rays(R, 0.0, 1.0, seed + 10, t_min=1e-4)
    # a finite t_max on every 3rd ray exercises the any-hit range
    rays.reshape(8, -1)[7, ::3] = 0.05
    d_rays = to_dev(rays)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    bvh.launch_rays(d_rays, hits, R, TRACE.CLOSEST_HIT)
    lb = oracle.build_lbvh(v)
    compare_hits(to_host(hits, np.float32), oracle.trace_bvh2(v, lb, rays, mode=0))
    bvh.launch_rays(d_rays, hits, R, TRACE.ANY_HIT)
    got = to_host(hits, np.float32)
    want = oracle.trace_bvh2(v, lb, rays, mode=1)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_icosphere_4m_incoherent(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import TRACE, build_triangle_bvh

    n = 4_194_304
    announce(f"trace icosphere n={n}")
    v = scenes.icosphere(n)
    bvh = build_triangle_bvh(to_dev(v))
    R = 1 << 20
    rays = random_rays(R, -1.2, 1.2, 77, t_min=1e-4)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    bvh.launch_rays(to_dev(rays), hits, R, TRACE.CLOSEST_HIT)
    lb = oracle.build_lbvh(v)
    sample = np.arange(0, R, 7)
    sub = rays.reshape(8, -1)[:, sample].reshape(-1)
    want = oracle.trace_bvh2(v, lb, sub, mode=0, threads=8)
    got = to_host(hits, np.float32).reshape(-1, 8)[sample].reshape(-1)
    frac = compare_hits(got, want)
    assert frac > 0.3  # about 44% of box-origin rays reach the surface


@pytest.mark.parametrize("width", [2, 8])
def test_grid_fixture_checksum_equals_bajo_published(gpu_ctx, width):
    """bajo's 65,536-triangle grid fixture and its 262,144 rays (bajo/benchmark/bvh_fixtures.mojo:11-58):
    the GPU hits give bajo's published checksum (bench/results/bvh_cpu/bvh_cpu.md) and equal the oracle's."""
    from bajo_amd import scenes
    from bajo_amd.bvh import build_triangle_bvh

    announce(f"grid fixture width {width}")
    v, rays = scenes.grid_fixture()
    bvh = build_triangle_bvh(torch.from_numpy(v).cuda(), node_width=width)
    n = rays.size // 8
    hits = torch.empty(8 * n, dtype=torch.float32, device="cuda")
    bvh.launch_rays(torch.from_numpy(rays).cuda(), hits, n)
    got = to_host(hits, np.float32)
    assert scenes.grid_checksum(got) == scenes.GRID_CHECKSUM
    want = oracle.trace_bvh2(v, oracle.build_lbvh(v), rays)
    compare_hits(got, want)
    bvh.free()

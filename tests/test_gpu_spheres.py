"""Sphere BVHs (bajo's build_sphere_bvh / GpuSphereBvh, bvh/gpu/sphere_bvh.mojo:
274-478) against the oracle: the box-build layout over Sphere.bounds bit-exact
(LBVH, H-PLOC, 8-wide in canonical order), closest / any hits of rays and of a
camera bit-exact, sphere BLASes in a TLAS (alone and mixed with triangle
BLASes) bit-exact, and bajo's GPU sphere-TLAS cases
(test/bvh/gpu/test_tlas.mojo:328-419)."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host
from test_gpu_lbvh import assert_same_lbvh
from test_gpu_tlas import launch, random_rays_world
from test_oracle_spheres import rays_of
from test_oracle_tlas import blas_box, random_instances, translation

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def spheres(n, seed, lo=0.0, hi=10.0, rmin=0.02, rmax=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, (n, 3))
    r = rng.uniform(rmin, rmax, (n, 1))
    return np.concatenate([c, r], 1).astype(np.float32).reshape(-1)


def sphere_box(sp):
    b = oracle.sphere_bounds(sp).reshape(-1, 6)
    return np.concatenate([b[:, :3].min(0), b[:, 3:].max(0)]).astype(np.float32)


@pytest.mark.parametrize("n,width,method", [(1, 2, 0), (2, 2, 0), (777, 2, 0), (20000, 2, 0), (20000, 8, 0),
                                            (5000, 2, 1), (5000, 8, 1)])
def test_sphere_bvh_layout_and_hits(gpu_ctx, n, width, method):
    from bajo_amd.bvh import TRACE, build_sphere_bvh

    announce(f"sphere bvh n={n} width={width} method={method}")
    sp = spheres(n, 10 + n)
    bvh = build_sphere_bvh(to_dev(sp), node_width=width, leaf_size=4, method=method)
    boxes = oracle.sphere_bounds(sp)
    want = oracle.build_lbvh_bounds(boxes) if method == 0 else oracle.build_hploc_bounds(boxes)
    got = bvh.binary()
    if method == 0:
        assert_same_lbvh(got, want)
    else:
        for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
            np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    if width == 8:
        assert oracle.canonical_wide(bvh.wide()).tobytes() == \
            oracle.canonical_wide(oracle.collapse_wide(want, 4)).tobytes()
    R = 30_000
    rng = np.random.default_rng(n)
    o = rng.uniform(0, 10, (R, 3))
    d = rng.normal(size=(R, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_of(o, d, 1e-4)
    rays.reshape(8, -1)[7, ::3] = 0.5
    d_rays = to_dev(rays)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    for mode in (TRACE.CLOSEST_HIT, TRACE.ANY_HIT):
        st = stats if mode == TRACE.CLOSEST_HIT else None
        bvh.launch_rays(d_rays, hits, R, mode, stats=st)
        g = to_host(hits, np.float32).reshape(-1, 8)
        if width == 8:
            w, ws, over = oracle.trace_spheres_wide(sp, want["leaf_ids"], bvh.wide(), rays, mode=int(mode),
                                                    with_stats=True)
            assert over == 0
        else:
            w, ws = oracle.trace_spheres(sp, want, rays, mode=int(mode), stats=True)
        np.testing.assert_array_equal(g.view(np.uint32), w.view(np.uint32), err_msg=f"mode {mode}")
        if st is not None:
            np.testing.assert_array_equal(to_host(stats, np.uint32).reshape(-1, 5)[:, [0, 1, 3, 4]], ws[:, [0, 1, 3, 4]])
    if n > 2:
        assert (g[:, 7] == 0.0).mean() > 0.05
    bvh.free()


def test_sphere_camera(gpu_ctx):
    from bajo_amd.bvh import Camera, build_sphere_bvh

    announce("sphere camera")
    sp = spheres(4000, 77, 0.0, 1.0, 0.005, 0.03)
    bvh = build_sphere_bvh(to_dev(sp), node_width=8, leaf_size=4)
    cam = Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()
    w, h = 160, 120
    hits = torch.empty(8 * w * h, dtype=torch.float32, device="cuda:0")
    bvh.launch_camera(to_dev(cam), hits, w * h, w, h)
    got = to_host(hits, np.float32).reshape(-1, 8)
    want, _, _ = oracle.trace_spheres_wide(sp, bvh.binary()["leaf_ids"], bvh.wide(), oracle.camera_rays(cam, w * h, w, h))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert (got[:, 2].view(np.uint32) != EMPTY).mean() > 0.2
    bvh.free()


@pytest.mark.parametrize("mixed", [False, True])
def test_sphere_tlas_matches_oracle(gpu_ctx, mixed):
    from bajo_amd import scenes
    from bajo_amd.bvh import Instance, build_sphere_bvh, build_triangle_bvh, build_triangle_tlas

    announce(f"sphere tlas mixed={mixed}")
    rng = np.random.default_rng(12)
    sps = [spheres(n, 900 + k, -1.0, 1.0, 0.05, 0.3) for k, n in enumerate((50, 1, 400))]
    gpu, ora, boxes = [], [], []
    for k, sp in enumerate(sps):
        b = build_sphere_bvh(to_dev(sp), node_width=8 if k == 2 else 2, leaf_size=4)
        d = dict(b.binary())
        d["spheres"] = sp
        if k == 2:
            d["wide"] = b.wide()
        gpu.append(b)
        ora.append(d)
        boxes.append(sphere_box(sp))
    if mixed:
        v = 2.0 * (scenes.soup(300, initstate=950, edge_min=0.2, edge_max=0.8) - 0.5)
        tb = build_triangle_bvh(to_dev(v), node_width=8, leaf_size=4)
        d = dict(tb.binary())
        d["vertices"] = v
        d["wide"] = tb.wide()
        gpu.append(tb)
        ora.append(d)
        boxes.append(blas_box(v))
    xforms, idx = random_instances(rng, 80, len(gpu))
    tlas = build_triangle_tlas(gpu, [Instance(x, i) for x, i in zip(xforms, idx)])
    ib = oracle.tlas_instance_bounds(np.stack(xforms), np.stack(boxes), idx)
    np.testing.assert_array_equal(tlas.instance_bounds().view(np.uint32), ib.view(np.uint32))
    rays = random_rays_world(30_000, 21)
    top = oracle.build_lbvh_bounds(ib)
    for mode in (0, 1):
        got = launch(tlas, rays, mode)
        want = oracle.tlas_trace(ora, np.stack(xforms).reshape(-1), np.asarray(idx, np.uint32), top, rays, mode)
        np.testing.assert_array_equal(got.view(np.uint32), want.reshape(-1, 8).view(np.uint32), err_msg=f"mode {mode}")
    assert tlas.status() == 0


def test_bajo_sphere_tlas_cases(gpu_ctx):
    from bajo_amd.bvh import Instance, build_sphere_bvh, build_triangle_tlas

    announce("bajo sphere tlas cases")
    b = build_sphere_bvh(to_dev(np.array([0, 0, 2, 1], np.float32)), node_width=2)
    # test/bvh/gpu/test_tlas.mojo:328-376: one instance moved to x = 10, a ray from (10, 0, 0) along +z
    t = build_triangle_tlas([b], [Instance(translation(10.0), 0)])
    h = launch(t, rays_of([10, 0, 0], [0, 0, 1]))[0]
    assert h[7] == pytest.approx(1.0) and h[2].view(np.uint32) == 0 and h[3].view(np.uint32) == 0
    np.testing.assert_allclose(h[4:7], [0, 0, -1], atol=1e-6)
    # :379-419: non-uniform scale (2, 1, 1), a ray from (1, 0, 0) along +z
    m = np.array([[2, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)
    t = build_triangle_tlas([b], [Instance(m, 0)])
    h = launch(t, rays_of([1, 0, 0], [0, 0, 1]))[0]
    assert h[7] == pytest.approx(1.1339746, abs=1e-5)
    np.testing.assert_allclose(h[4:7], [0.2773501, 0.0, -0.9607689], atol=1e-5)


def test_sphere_bvh_rejects_bad_use(gpu_ctx):
    from bajo_amd._lib import BajoError
    from bajo_amd.bvh import build_binary_bvh, build_sphere_bvh

    announce("sphere bvh argument errors")
    b = build_sphere_bvh(to_dev(spheres(10, 1)))
    with pytest.raises(BajoError, match="rebuilt by bajo_build_sphere_bvh"):
        from bajo_amd.bvh import GpuBinaryBoundsBvh
        GpuBinaryBoundsBvh.rebuild(b, to_dev(oracle.sphere_bounds(spheres(10, 1))))
    boxes = build_binary_bvh(to_dev(oracle.sphere_bounds(spheres(10, 1))))
    with pytest.raises(BajoError, match="not a sphere BVH"):
        from bajo_amd.bvh import GpuSphereBvh
        GpuSphereBvh.rebuild(boxes, to_dev(spheres(10, 1)))
    b.free()
    boxes.free()

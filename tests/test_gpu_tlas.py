"""Instances under a top-level hierarchy (bajo's build_triangle_tlas /
GpuTriangleTlas, bvh/gpu/tlas.mojo:54-233, 482-645) against the oracle's
restatement: instance world boxes and the top-level layout bit-exact, closest
and any hits of random rays and of a camera bit-exact, over binary and 8-wide
BLASes (mixed in one TLAS) and both top-level builders; bajo's own GPU cases
(test/bvh/gpu/test_tlas.mojo:97-141, test_blas_tlas.mojo:128-273); and the
argument errors."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host
from test_oracle_tlas import TRI, blas_box, random_instances, rays_of, translation

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def gpu_blas(v, width, method=0):
    from bajo_amd.bvh import build_triangle_bvh

    bvh = build_triangle_bvh(to_dev(v), node_width=width, leaf_size=4 if width == 8 else 0, method=method)
    d = dict(bvh.binary())
    d["vertices"] = np.asarray(v, np.float32)
    if width == 8:
        d["wide"] = bvh.wide()
    return bvh, d


def launch(tlas, rays, mode=0):
    R = rays.size // 8
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    tlas.launch_rays(to_dev(rays), hits, R, mode)
    return to_host(hits, np.float32).reshape(-1, 8)


def scene(widths, n_inst, seed, method=0):
    from bajo_amd import scenes
    from bajo_amd.bvh import Instance, build_triangle_tlas

    rng = np.random.default_rng(seed)
    sizes = (60, 300, 1, 2000)[: len(widths)]
    verts = [2.0 * (scenes.soup(n, initstate=400 + k, edge_min=0.2, edge_max=0.8) - 0.5) for k, n in enumerate(sizes)]
    built = [gpu_blas(v, w) for v, w in zip(verts, widths)]
    xforms, idx = random_instances(rng, n_inst, len(verts))
    tlas = build_triangle_tlas([b for b, _ in built], [Instance(x, i) for x, i in zip(xforms, idx)], method=method)
    boxes = oracle.tlas_instance_bounds(np.stack(xforms), np.stack([blas_box(v) for v in verts]), idx)
    return tlas, [d for _, d in built], np.stack(xforms).reshape(-1), np.asarray(idx, np.uint32), boxes


def random_rays_world(R, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-7, 7, (R, 3))
    d = rng.uniform(-3, 3, (R, 3)) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_of(o, d, 0.0, 1e30)
    rays.reshape(8, -1)[7, ::4] = 6.0  # finite t_max on every fourth ray
    rays.reshape(8, -1)[3, 1::4] = 2.0  # and a t_min on another quarter
    return rays


@pytest.mark.parametrize(
    "widths,method",
    [((2, 2, 2), 0), ((8, 8, 8), 0), ((2, 8, 2, 8), 0), ((8, 2, 8), 1)],
    ids=["binary-lbvh", "wide-lbvh", "mixed-lbvh", "mixed-hploc"],
)
def test_tlas_matches_oracle(gpu_ctx, widths, method):
    announce(f"tlas widths={widths} method={method}")
    tlas, blases, xf, idx, boxes = scene(widths, 64, 23 + method, method)
    np.testing.assert_array_equal(tlas.instance_bounds().view(np.uint32), boxes.view(np.uint32))
    top = tlas.top().binary()
    want_top = oracle.build_lbvh_bounds(boxes) if method == 0 else oracle.build_hploc_bounds(boxes)
    assert top["root"] == want_top["root"]
    for key in ("node_meta", "leaf_ids"):
        np.testing.assert_array_equal(top[key], want_top[key], err_msg=key)
    np.testing.assert_array_equal(top["node_bounds"].view(np.uint32), want_top["node_bounds"].view(np.uint32))
    rays = random_rays_world(50_000, 31 + method)
    for mode in (0, 1):
        got = launch(tlas, rays, mode)
        want = oracle.tlas_trace(blases, xf, idx, want_top, rays, mode).reshape(-1, 8)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"mode {mode}")
    assert (got[:, 7] == 0.0).mean() > 0.2  # the any-hit query saw occluders
    assert tlas.status() == 0


def test_tlas_camera_matches_oracle(gpu_ctx):
    from bajo_amd.bvh import Camera

    announce("tlas camera 256x192")
    tlas, blases, xf, idx, boxes = scene((8, 2, 8), 200, 41)
    cam = Camera((0.0, 2.0, -14.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.6).flatten()
    w, h = 256, 192
    hits = torch.empty(8 * w * h, dtype=torch.float32, device="cuda:0")
    tlas.launch_camera(to_dev(cam), hits, w * h, w, h)
    got = to_host(hits, np.float32).reshape(-1, 8)
    want = oracle.tlas_trace(blases, xf, idx, oracle.build_lbvh_bounds(boxes), oracle.camera_rays(cam, w * h, w, h))
    np.testing.assert_array_equal(got.view(np.uint32), want.reshape(-1, 8).view(np.uint32))
    assert (got[:, 2].view(np.uint32) != EMPTY).mean() > 0.2
    assert tlas.status() == 0


@pytest.mark.parametrize("width", [2, 8])
def test_tlas_bajo_cases(gpu_ctx, width):
    from bajo_amd.bvh import Instance, build_triangle_tlas

    announce(f"tlas bajo cases width={width}")
    b0, _ = gpu_blas(TRI, width)
    b1, _ = gpu_blas(TRI, width)
    # test/bvh/gpu/test_tlas.mojo:97-141: one instance moved to x = 10, a ray from (10, 0, 0) along +z
    t = build_triangle_tlas([b0], [Instance(translation(10.0), 0)])
    h = launch(t, rays_of([10, 0, 0], [0, 0, 1]))[0]
    assert h[7] == pytest.approx(2.0) and h[2].view(np.uint32) == 0 and h[3].view(np.uint32) == 0
    np.testing.assert_allclose(h[4:7], [0, 0, 1], atol=1e-7)
    # test_blas_tlas.mojo:128-168 / 201-237: the instance's BLAS index selects the BLAS; the closest of
    # two instances (different BLASes) wins
    t = build_triangle_tlas([b0, b1], [Instance(translation(tz=6.0), 0), Instance(translation(), 1)])
    h = launch(t, rays_of([[0, 0, 0], [0, 0, 9]], [[0, 0, 1], [0, 0, -1]]))
    assert h[0, 3].view(np.uint32) == 1 and h[0, 7] == pytest.approx(2.0)
    assert h[1, 3].view(np.uint32) == 0 and h[1, 7] == pytest.approx(1.0)
    a = launch(t, rays_of([0, 0, 0], [0, 0, 1], tmax=1.5), mode=1)[0]
    assert a[7] == pytest.approx(1.5) and a[3].view(np.uint32) == EMPTY  # short shadow ray: unoccluded
    a = launch(t, rays_of([0, 0, 0], [0, 0, 1], tmax=3.0), mode=1)[0]
    assert a[7] == 0.0 and a[3].view(np.uint32) == 1
    assert t.status() == 0


def test_tlas_rejects_bad_arguments(gpu_ctx):
    from bajo_amd._lib import BajoError
    from bajo_amd.bvh import Instance, build_binary_bvh, build_triangle_tlas

    announce("tlas argument errors")
    b, _ = gpu_blas(TRI, 2)
    with pytest.raises(BajoError, match="references BLAS"):
        build_triangle_tlas([b], [Instance(translation(), 3)])
    singular = np.array([[1, 2, 3, 0], [2, 4, 6, 0], [0, 0, 1, 0]], np.float32)
    with pytest.raises(BajoError, match="not invertible"):
        build_triangle_tlas([b], [Instance(singular, 0)])
    boxes = build_binary_bvh(to_dev(np.array([0, 0, 0, 1, 1, 1], np.float32)))
    with pytest.raises(BajoError, match="not a triangle or sphere BVH"):
        build_triangle_tlas([boxes], [Instance(translation(), 0)])

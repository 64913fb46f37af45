"""Segmented BLAS sets (bajo's build_triangle_blas_set, bvh/gpu/triangle_bvh.mojo:
366-491): every segment of one segmented build equals the standalone oracle LBVH
of its mesh bit for bit (binary layout; 8-wide collapse in canonical order),
traces bit-identically to the oracle, and enters a TLAS like a standalone BVH
(test/bvh/gpu/test_blas_tlas.mojo:128-273, 320-426 cover bajo's BLAS sets)."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host
from test_gpu_lbvh import assert_same_lbvh
from test_gpu_tlas import launch, random_rays_world
from test_gpu_trace import random_rays
from test_gpu_wide import soup
from test_oracle_tlas import blas_box, random_instances

pytestmark = pytest.mark.gpu

SIZES = (1, 2, 3, 100, 5000, 1, 777, 65536, 4)


def meshes(sizes, seed):
    return [soup(n, seed + k) for k, n in enumerate(sizes)]


@pytest.mark.parametrize("width", [2, 8])
def test_segments_equal_standalone_builds(gpu_ctx, width):
    from bajo_amd.bvh import build_triangle_blas_set

    announce(f"blas set layout width={width}")
    ms = meshes(SIZES, 100)
    t = {}
    bs = build_triangle_blas_set([to_dev(m) for m in ms], node_width=width, leaf_size=4, timings=t)
    assert t["total_ns"] > 0
    for k, m in enumerate(ms):
        view = bs.blas(k)
        lb = oracle.build_lbvh(m)
        assert_same_lbvh(view.binary(), lb)
        if width == 8:
            got = oracle.canonical_wide(view.wide())
            assert got.tobytes() == oracle.canonical_wide(oracle.collapse_wide(lb, 4)).tobytes(), f"segment {k}"
    bs.free()


@pytest.mark.parametrize("width", [2, 8])
def test_segments_trace_like_the_oracle(gpu_ctx, width):
    from bajo_amd.bvh import TRACE, build_triangle_blas_set

    announce(f"blas set trace width={width}")
    ms = meshes(SIZES, 200)
    bs = build_triangle_blas_set([to_dev(m) for m in ms], node_width=width, leaf_size=4)
    R = 20_000
    rays = random_rays(R, 0.0, 1.0, 7, t_min=1e-4)
    d_rays = to_dev(rays)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    for k, m in enumerate(ms):
        view = bs.blas(k)
        lb = oracle.build_lbvh(m)
        for mode in (TRACE.CLOSEST_HIT, TRACE.ANY_HIT):
            view.launch_rays(d_rays, hits, R, mode, stats=stats if mode == TRACE.CLOSEST_HIT else None)
            got = to_host(hits, np.float32).reshape(-1, 8)
            if width == 8:
                want, want_stats, over = oracle.trace_wide(m, lb["leaf_ids"], view.wide(), rays, mode=int(mode),
                                                           with_stats=True)
                assert over == 0
            else:
                want, want_stats = oracle.trace_bvh2(m, lb, rays, mode=int(mode), stats=True)
                want = want.reshape(-1, 8)
                want_stats = want_stats.reshape(-1, 5)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"segment {k} mode {mode}")
            if mode == TRACE.CLOSEST_HIT:
                st = to_host(stats, np.uint32).reshape(-1, 5)
                np.testing.assert_array_equal(st[:, [0, 1, 3, 4]], want_stats[:, [0, 1, 3, 4]])
    bs.free()


@pytest.mark.parametrize("width", [2, 8])
def test_batched_camera_matches_per_mesh_oracle(gpu_ctx, width):
    from bajo_amd.bvh import Camera, build_triangle_blas_set

    announce(f"blas set batched camera width={width}")
    ms = meshes(SIZES, 300)
    bs = build_triangle_blas_set([to_dev(m) for m in ms], node_width=width, leaf_size=4)
    w, h = 48, 32
    cams = np.concatenate([Camera((0.5 + 0.05 * k, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()
                           for k in range(len(ms))])
    R = len(ms) * w * h
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    bs.launch_camera(to_dev(cams), hits, w, h, stats=stats)
    got = to_host(hits, np.float32).reshape(len(ms), w * h, 8)
    st = to_host(stats, np.uint32).reshape(len(ms), w * h, 5)
    for k, m in enumerate(ms):
        lb = oracle.build_lbvh(m)
        rays = oracle.camera_rays(cams[20 * k: 20 * k + 20], w * h, w, h)
        if width == 8:
            want, want_st, _ = oracle.trace_wide(m, lb["leaf_ids"], bs.blas(k).wide(), rays, with_stats=True)
        else:
            want, want_st = oracle.trace_bvh2(m, lb, rays, stats=True)
            want, want_st = want.reshape(-1, 8), want_st.reshape(-1, 5)
        np.testing.assert_array_equal(got[k].view(np.uint32), want.view(np.uint32), err_msg=f"view {k}")
        np.testing.assert_array_equal(st[k][:, [0, 1, 3, 4]], want_st[:, [0, 1, 3, 4]], err_msg=f"view {k}")
    bs.free()


def test_tlas_over_blas_set(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Instance, build_triangle_blas_set, build_triangle_tlas

    announce("tlas over a blas set")
    rng = np.random.default_rng(3)
    verts = [2.0 * (scenes.soup(n, initstate=500 + k, edge_min=0.2, edge_max=0.8) - 0.5)
             for k, n in enumerate((60, 1, 300, 2000))]
    bs = build_triangle_blas_set([to_dev(v) for v in verts], node_width=8, leaf_size=4)
    xforms, idx = random_instances(rng, 96, len(verts))
    tlas = build_triangle_tlas(bs.blases(), [Instance(x, i) for x, i in zip(xforms, idx)])
    boxes = oracle.tlas_instance_bounds(np.stack(xforms), np.stack([blas_box(v) for v in verts]), idx)
    np.testing.assert_array_equal(tlas.instance_bounds().view(np.uint32), boxes.view(np.uint32))
    blases = []
    for k, v in enumerate(verts):
        d = dict(bs.blas(k).binary())
        d["vertices"] = v
        d["wide"] = bs.blas(k).wide()
        blases.append(d)
    rays = random_rays_world(30_000, 5)
    top = oracle.build_lbvh_bounds(boxes)
    for mode in (0, 1):
        got = launch(tlas, rays, mode)
        want = oracle.tlas_trace(blases, np.stack(xforms).reshape(-1), np.asarray(idx, np.uint32), top, rays, mode)
        np.testing.assert_array_equal(got.view(np.uint32), want.reshape(-1, 8).view(np.uint32), err_msg=f"mode {mode}")
    assert tlas.status() == 0
    tlas.free()
    bs.free()


def test_blas_set_rejects_bad_arguments(gpu_ctx):
    import ctypes

    from bajo_amd import _lib
    from bajo_amd._lib import BajoError
    from bajo_amd.bvh import build_options, build_triangle_blas_set
    from bajo_amd.context import default_context

    announce("blas set argument errors")
    ctx = default_context(0)
    v = to_dev(soup(10, 1))
    h = ctypes.c_void_p()

    def call(offsets, method=0, width=2):
        off = np.asarray(offsets, np.uint64)
        opts = build_options(width, 0, method)
        return _lib.lib().bajo_build_triangle_blas_set(ctx.handle, v.data_ptr(), off.ctypes.data_as(ctypes.c_void_p),
                                                       off.size - 1, ctypes.byref(opts), None, ctypes.byref(h))

    assert call([0, 4, 4, 10]) != 0 and b"empty" in _lib.lib().bajo_last_error()
    assert call([1, 10]) != 0
    assert call([0, 10], method=5) != 0 and b"H-PLOC" in _lib.lib().bajo_last_error()
    bs = build_triangle_blas_set([v])
    view = bs.blas(0)
    assert _lib.lib().bajo_bvh_free(view.handle) != 0  # a view is released with its set
    with pytest.raises(BajoError, match="cannot be rebuilt"):
        view.rebuild(v)
    with pytest.raises(BajoError):
        bs.blas(1)
    bs.free()


def _preorder(b):
    """Preorder walk of an exported binary layout: (refs with leaf positions, inner boxes, inner fences)."""
    meta = b["node_meta"].reshape(-1, 4)
    nb = b["node_bounds"].reshape(-1, 12)
    refs, boxes, fences = [], [], []
    stack = [int(b["root"])]
    while stack:
        r = stack.pop()
        refs.append(r if r & 0x80000000 else -1)
        if r & 0x80000000:
            continue
        boxes.append(nb[r].view(np.uint32))
        fences.append(int(meta[r, 3]))
        stack.append(int(meta[r, 2]))  # meta rows: parent, left, right, fence
        stack.append(int(meta[r, 1]))
    return np.array(refs), np.array(boxes), np.array(fences)


@pytest.mark.parametrize("width", [2, 8])
def test_hploc_segments_equal_standalone_hploc(gpu_ctx, width):
    """bajo's default set build is H-PLOC (triangle_bvh.mojo:469-491, segmented_build.mojo:117-130): every
    segment equals the standalone H-PLOC build of its mesh (same tree in preorder: leaf order, node boxes,
    fences; same 8-wide collapse; identical hits) and the oracle's H-PLOC topology."""
    from bajo_amd.bvh import build_triangle_blas_set, build_triangle_bvh

    announce(f"hploc blas set width={width}")
    ms = meshes(SIZES, 300)
    bs = build_triangle_blas_set([to_dev(m) for m in ms], node_width=width, leaf_size=4, method=1)
    for k, m in enumerate(ms):
        view = bs.blas(k)
        one = build_triangle_bvh(to_dev(m), node_width=width, leaf_size=4, method=1)
        got, want = view.binary(), one.binary()
        np.testing.assert_array_equal(got["leaf_ids"], want["leaf_ids"], err_msg=f"segment {k}")
        if m.size // 9 > 1:
            ga, gb, gf = _preorder(got)
            wa, wb, wf = _preorder(want)
            np.testing.assert_array_equal(ga, wa, err_msg=f"segment {k}")
            np.testing.assert_array_equal(gb, wb, err_msg=f"segment {k}")
            np.testing.assert_array_equal(gf, wf, err_msg=f"segment {k}")
            assert got["root"] == 0
            ref = oracle.build_hploc(m)
            np.testing.assert_array_equal(oracle.preorder_binary(got["root"], got["node_meta"], got["leaf_ids"]),
                                          oracle.preorder_topology(ref["topology"]), err_msg=f"segment {k}")
        if width == 8:
            assert oracle.canonical_wide(view.wide()).tobytes() == oracle.canonical_wide(one.wide()).tobytes()
        R = 4096
        rays = to_dev(random_rays(R, -0.1, 1.1, 40 + k, t_min=1e-4))
        ha = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
        hb = torch.empty_like(ha)
        view.launch_rays(rays, ha, R)
        one.launch_rays(rays, hb, R)
        np.testing.assert_array_equal(to_host(ha, np.uint32), to_host(hb, np.uint32), err_msg=f"segment {k}")
        one.free()
    bs.free()

"""The builders over caller boxes (bajo's GpuBinaryBoundsBvh + build_binary_bvh,
bvh/gpu/builder/binary_layout.mojo:303-448, binary_builder.mojo:21-45)
against the oracle's box-input restatement: the LBVH and the H-PLOC binary
layouts bit-exact (keys, leaf ids, node meta and boxes, leaf parents and
counts), the 8-wide collapse equal in canonical order, a triangle build equal
to the box build of its triangles' boxes, and ray launches refused on a boxes
hierarchy."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev
from test_gpu_lbvh import assert_same_lbvh

pytestmark = pytest.mark.gpu


def random_boxes(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)) * 10.0
    e = rng.random((n, 3)) ** 3 * 0.5
    return np.concatenate([c - e, c + e], axis=1).astype(np.float32).reshape(-1)


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (100, 3), (5000, 4), (200_000, 5)])
def test_lbvh_over_boxes_bit_exact(gpu_ctx, n, seed):
    from bajo_amd.bvh import build_binary_bvh

    announce(f"bounds lbvh n={n}")
    b = random_boxes(n, seed)
    got = build_binary_bvh(to_dev(b)).binary()
    assert_same_lbvh(got, oracle.build_lbvh_bounds(b))


@pytest.mark.parametrize("n,seed", [(2, 6), (3, 7), (1000, 8), (65536, 9)])
def test_hploc_over_boxes_bit_exact(gpu_ctx, n, seed):
    from bajo_amd.bvh import GpuBvhBuildMethod, build_binary_bvh

    announce(f"bounds hploc n={n}")
    b = random_boxes(n, seed)
    got = build_binary_bvh(to_dev(b), method=GpuBvhBuildMethod.HPLOC).binary()
    want = oracle.build_hploc_bounds(b)
    assert got["root"] == want["root"]
    for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    np.testing.assert_array_equal(got["node_bounds"].view(np.uint32), want["node_bounds"].view(np.uint32))


def test_wide_over_boxes_equals_oracle_collapse(gpu_ctx):
    from bajo_amd.bvh import build_binary_bvh

    announce("bounds wide")
    b = random_boxes(30_000, 10)
    bvh = build_binary_bvh(to_dev(b), node_width=8, leaf_size=4)
    want = oracle.collapse_wide(oracle.build_lbvh_bounds(b), 4)
    assert oracle.canonical_wide(bvh.wide()).tobytes() == oracle.canonical_wide(want).tobytes()


def test_triangle_build_equals_box_build_of_its_boxes(gpu_ctx):
    from bajo_amd.bvh import build_binary_bvh, build_triangle_bvh

    announce("bounds vs triangles")
    rng = np.random.default_rng(11)
    v0 = rng.random((20_000, 3), dtype=np.float32)
    e = (rng.random((20_000, 2, 3), dtype=np.float32) - 0.5) * 0.05
    v = np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)
    tri = build_triangle_bvh(to_dev(v)).binary()
    box = build_binary_bvh(to_dev(oracle.triangle_bounds(v))).binary()
    assert_same_lbvh(box, tri)


def test_ray_launch_refused_on_boxes(gpu_ctx):
    from bajo_amd._lib import BajoError
    from bajo_amd.bvh import build_binary_bvh

    bvh = build_binary_bvh(to_dev(random_boxes(10, 12)))
    rays = torch.zeros(8 * 4, dtype=torch.float32, device="cuda:0")
    hits = torch.empty(8 * 4, dtype=torch.float32, device="cuda:0")
    with pytest.raises(BajoError):
        bvh.launch_rays(rays, hits, 4)

"""The oracle's box-input build equals its triangle build on the triangles'
boxes (the triangle build is pinned in test_oracle.py), and its node boxes are
the exact unions of the input boxes below them.  CPU only."""

import numpy as np

import oracle
from test_oracle import soup


def test_box_build_equals_triangle_build():
    v = soup(3000, seed=21)
    a = oracle.build_lbvh(v)
    b = oracle.build_lbvh_bounds(oracle.triangle_bounds(v))
    for key in ("sorted_keys", "leaf_ids", "node_meta", "leaf_parent", "leaf_counts", "node_bounds", "scene_bounds"):
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    assert a["root"] == b["root"]


def test_box_build_unions():
    rng = np.random.default_rng(3)
    c = rng.random((500, 3)) * 4
    e = rng.random((500, 3)) * 0.3
    boxes = np.concatenate([c - e, c + e], axis=1).astype(np.float32)
    b = oracle.build_lbvh_bounds(boxes.reshape(-1))
    meta = b["node_meta"].reshape(-1, 4)
    nb = b["node_bounds"].reshape(-1, 12)
    sorted_boxes = boxes[b["leaf_ids"]]
    for x in range(meta.shape[0]):
        last = meta[x, 3]
        first = last + 1 - b["leaf_counts"][x]
        lo = np.minimum(nb[x, :3], nb[x, 6:9])
        hi = np.maximum(nb[x, 3:6], nb[x, 9:12])
        np.testing.assert_array_equal(lo, sorted_boxes[first:last + 1, :3].min(0))
        np.testing.assert_array_equal(hi, sorted_boxes[first:last + 1, 3:].max(0))
    h = oracle.build_hploc_bounds(boxes.reshape(-1))
    assert sorted(h["leaf_ids"].tolist()) == list(range(500))

"""Status words of the build (CPU): the refit climb's corruption checks restated
by the oracle (the library's refit raises BAJO_STATUS_REFIT_CORRUPT = 5 on the same
conditions): a valid hierarchy climbs to the same boxes as the closed-form refit and
reports 0; a node that does not list its child, a parent outside the node array and
a parent cycle each report 5 instead of being silently accepted."""

import numpy as np

import oracle
from bajo_amd import scenes


def _build(n=3000, seed=41):
    v = scenes.soup(n, initstate=seed, initseq=54)
    bvh = oracle.build_lbvh(v)
    boxes = oracle.triangle_bounds(v).reshape(-1, 6)[bvh["leaf_ids"]]
    return bvh, boxes


def test_valid_hierarchy_climbs_to_the_closed_form_boxes():
    bvh, boxes = _build()
    status, nb = oracle.refit_climb(boxes, bvh["node_meta"], bvh["leaf_parent"])
    assert status == 0
    np.testing.assert_array_equal(nb.view(np.uint32), bvh["node_bounds"].view(np.uint32))


def test_corrupted_hierarchies_raise_the_refit_status():
    bvh, boxes = _build()
    meta = bvh["node_meta"].reshape(-1, 4)
    # a node whose left child no longer lists the leaf / node that climbs into it
    bad = meta.copy()
    k = 17
    bad[k, 1] = bad[k, 2]
    assert oracle.refit_climb(boxes, bad.reshape(-1), bvh["leaf_parent"])[0] == 5
    # a leaf whose parent index lies outside the node array
    lp = bvh["leaf_parent"].copy()
    lp[5] = meta.shape[0] + 3
    assert oracle.refit_climb(boxes, bvh["node_meta"], lp)[0] == 5
    # a parent cycle: an inner node's parent points back at one of its children
    cyc = meta.copy()
    child = int(cyc[k, 1])
    if child & oracle.LEAF_FLAG == 0:
        cyc[k, 0] = child
        assert oracle.refit_climb(boxes, cyc.reshape(-1), bvh["leaf_parent"])[0] == 5


def test_single_leaf_has_nothing_to_climb():
    status, nb = oracle.refit_climb(np.zeros(6, np.float32), np.zeros(0, np.uint32),
                                    np.full(1, oracle.SENTINEL, np.uint32))
    assert status == 0 and nb.size == 0

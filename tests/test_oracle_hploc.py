"""Pins the oracle's H-PLOC topology (oracle/orc_hploc.c) before it checks the
device builder.  CPU only.

- bajo's exact-topology fixtures (test/bvh/gpu/test_hploc_topology.mojo:
  28-100): hierarchical merging with threshold 2, strict-tie ordering,
  duplicate codes with a permuted id list;
- bajo's quality gate (:103-150): on its 64-triangle scene H-PLOC is no worse
  than the LBVH over the same sorted codes, and the root boxes agree;
- the structural invariants of HplocTopology.validate (bvh/cpu/builder/
  hploc.mojo:245-322) on random soups, and the preorder serialization used
  to compare topologies across node numberings.
"""

import numpy as np
import pytest

import oracle
from test_oracle import soup


def box(cx):
    return [cx - 0.1, -0.1, -0.1, cx + 0.1, 0.1, 0.1]


def validate(t, n):
    """bajo's HplocTopology.validate, restated over the oracle's arrays."""
    m = 2 * n - 1
    assert t["left"].size == m
    root = t["root"]
    assert t["parent"][root] == oracle.SENTINEL and t["leaf_count"][root] == n
    seen_nodes = np.zeros(m, bool)
    seen_leaves = np.zeros(n, bool)
    stack = [root]
    b = t["bounds"].reshape(-1, 6)
    while stack:
        x = stack.pop()
        assert not seen_nodes[x]
        seen_nodes[x] = True
        if t["left"][x] == oracle.SENTINEL:
            assert t["right"][x] == oracle.SENTINEL and t["leaf_count"][x] == 1
            lid = t["leaf_id"][x]
            assert lid < n and not seen_leaves[lid]
            seen_leaves[lid] = True
            continue
        a, c = int(t["left"][x]), int(t["right"][x])
        assert a != c and t["parent"][a] == x and t["parent"][c] == x
        assert t["leaf_id"][x] == oracle.SENTINEL
        assert t["leaf_count"][x] == t["leaf_count"][a] + t["leaf_count"][c]
        np.testing.assert_array_equal(b[x, :3], np.minimum(b[a, :3], b[c, :3]))
        np.testing.assert_array_equal(b[x, 3:], np.maximum(b[a, 3:], b[c, 3:]))
        # bottom-up numbering: children were created before their parent
        assert a < x and c < x
        stack += [a, c]
    assert seen_nodes.all() and seen_leaves.all()


def test_hierarchical_merging_exact_topology():
    centers = [0.0, 1.0, 10.0, 11.0, 100.0, 101.0, 110.0, 111.0]
    bounds = np.array([box(c) for c in centers], np.float32).reshape(-1)
    codes = np.array([0, 1, 2, 3, 256, 257, 258, 259], np.uint32)
    t = oracle.hploc_topology(bounds, codes, np.arange(8, dtype=np.uint32), search_radius=8, merging_threshold=2)
    validate(t, 8)
    assert t["root"] == 14 and t["left"].size == 15
    st = t["stats"]
    assert (st["guide_nodes"], st["merge_calls"], st["hierarchical_rounds"], st["final_rounds"]) == (7, 3, 2, 2)
    assert list(t["left"][8:]) == [0, 2, 4, 6, 8, 10, 12]
    assert list(t["right"][8:]) == [1, 3, 5, 7, 9, 11, 13]


def test_strict_ties_prefer_right_then_near():
    bounds = np.array([box(0.0)] * 4, np.float32).reshape(-1)
    t = oracle.hploc_topology(bounds, np.full(4, 7, np.uint32), np.arange(4, dtype=np.uint32), merging_threshold=16)
    validate(t, 4)
    assert (t["left"][4], t["right"][4]) == (2, 3)
    assert (t["left"][5], t["right"][5]) == (1, 4)
    assert (t["left"][6], t["right"][6]) == (0, 5)
    assert t["root"] == 6


def test_duplicate_codes_and_permutation_are_deterministic():
    bounds = np.array([box(-6.0), box(-2.0), box(2.0), box(6.0)], np.float32).reshape(-1)
    codes = np.full(4, 11, np.uint32)
    ids = np.array([3, 1, 2, 0], np.uint32)
    a = oracle.hploc_topology(bounds, codes, ids, merging_threshold=2)
    b = oracle.hploc_topology(bounds, codes, ids, merging_threshold=2)
    validate(a, 4)
    for k in ("parent", "left", "right", "leaf_id"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["parent"][a["root"]] == oracle.SENTINEL


def quality_triangle_bounds():
    """The 64-triangle scene of bajo's quality gate, as triangle boxes."""
    out = []
    for i in range(64):
        x = np.float32((i % 8) * 4 - 14)
        y = np.float32(((i // 8) % 8) * 4 - 14)
        z = np.float32((i * 7) % 5)
        s = np.float32(2.5) if i % 9 == 0 else np.float32(0.7)
        p = np.array([[x - s, y - s, z], [x + s, y - s, z], [x, y + s, z + np.float32(0.1)]], np.float32)
        out.append(np.concatenate([p.min(0), p.max(0)]))
    return np.array(out, np.float32)


def binary_quality(lb):
    """bajo's measure_binary_bvh_quality (bvh/gpu/quality.mojo:63-108)."""
    nb = lb["node_bounds"].reshape(-1, 12).astype(np.float64)
    lo, hi = np.minimum(nb[:, :3], nb[:, 6:9]), np.maximum(nb[:, 3:6], nb[:, 9:12])
    d = hi - lo
    internal = (2 * (d[:, 0] * d[:, 1] + d[:, 0] * d[:, 2] + d[:, 1] * d[:, 2])).sum()
    return internal, lo, hi


def test_quality_gate_against_lbvh():
    tb = quality_triangle_bounds()
    # the LBVH over the same sorted (code, id) pairs, built by the oracle from the boxes
    sb = oracle.scene_bounds(tb.reshape(-1))
    codes = oracle.morton_codes(tb.reshape(-1), sb[6:12])
    keys = (codes.astype(np.uint64) << np.uint64(32)) | np.arange(64, dtype=np.uint64)
    keys, _ = oracle.radix_sort_pairs_u64(keys, None)
    sc = (keys >> np.uint64(32)).astype(np.uint32)
    ids = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    t = oracle.hploc_topology(tb.reshape(-1), sc, ids)
    validate(t, 64)
    # LBVH quality: internal-node areas (from the Karras topology + refit) plus leaf areas
    root, meta, _, counts = oracle.karras_topology(sc)
    nb = np.empty(12 * 63, np.float32)
    oracle.lib().orc_refit(root, oracle._p(tb.reshape(-1)), oracle._p(ids), oracle._p(meta), oracle._p(nb))
    internal, lo, hi = binary_quality({"node_bounds": nb})
    d = tb[:, 3:].astype(np.float64) - tb[:, :3]
    leaf_area = (2 * (d[:, 0] * d[:, 1] + d[:, 0] * d[:, 2] + d[:, 1] * d[:, 2])).sum()
    rd = hi[root] - lo[root]
    root_area = 2 * (rd[0] * rd[1] + rd[0] * rd[2] + rd[1] * rd[2])
    lbvh_quality = (internal + leaf_area) / root_area
    hb = t["bounds"].reshape(-1, 6)[t["root"]].astype(np.float64)
    hd = hb[3:] - hb[:3]
    assert abs(2 * (hd[0] * hd[1] + hd[0] * hd[2] + hd[1] * hd[2]) - root_area) <= 1e-4
    q = oracle.hploc_quality(t)
    print("H-PLOC quality", q, "LBVH quality", lbvh_quality)
    assert q <= lbvh_quality + 1e-5


@pytest.mark.parametrize("n,seed,thr", [(1, 0, 16), (2, 1, 16), (3, 2, 2), (17, 3, 16), (33, 4, 16), (500, 5, 16),
                                        (500, 6, 4), (4000, 7, 16), (65, 8, 32), (4000, 9, 32)])
def test_invariants_on_soups(n, seed, thr):
    v = soup(n, seed=seed)
    lb = oracle.build_lbvh(v)
    codes = (lb["sorted_keys"] >> np.uint64(32)).astype(np.uint32)
    t = oracle.hploc_topology(oracle.triangle_bounds(v), codes, lb["leaf_ids"], merging_threshold=thr)
    validate(t, n)
    pre = oracle.preorder_topology(t)
    assert pre.size == 2 * n - 1
    assert np.sum(pre != oracle.SENTINEL) == n


def test_preorder_of_lbvh_binary_layout():
    """The binary-layout serialization agrees with the topology form on the
    same tree (the LBVH, written both ways)."""
    v = soup(300, seed=9)
    lb = oracle.build_lbvh(v)
    n = 300
    meta = lb["node_meta"].reshape(-1, 4)
    # topology form: leaves 0..n-1 by sorted position, inner node i -> n + i
    left = np.full(2 * n - 1, oracle.SENTINEL, np.uint32)
    right = left.copy()
    leaf_id = np.concatenate([lb["leaf_ids"], np.full(n - 1, oracle.SENTINEL, np.uint32)])

    def conv(ref):
        return ref & 0x7FFFFFFF if ref & oracle.LEAF_FLAG else n + ref

    for i in range(n - 1):
        left[n + i], right[n + i] = conv(int(meta[i, 1])), conv(int(meta[i, 2]))
    t = {"root": n + lb["root"], "left": left, "right": right, "leaf_id": leaf_id}
    a = oracle.preorder_topology(t)
    b = oracle.preorder_binary(lb["root"], lb["node_meta"], lb["leaf_ids"])
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("n,seed", [(2, 11), (3, 12), (40, 13), (777, 14), (5000, 15)])
def test_lowest_leaf_numbering(n, seed):
    """The device builder's counter-free inner-node numbering, restated by the
    oracle: a bijection onto 0..n-2, equal (after relabeling) to the
    creation-order layout, and equal to a replay of the merges in creation
    order where each cluster carries the set of its sorted positions."""
    v = soup(n, seed=seed)
    lb = oracle.build_lbvh(v)
    codes = (lb["sorted_keys"] >> np.uint64(32)).astype(np.uint32)
    t = oracle.hploc_topology(oracle.triangle_bounds(v), codes, lb["leaf_ids"])
    num = oracle.hploc_lowest_leaf_numbering(t, lb["leaf_ids"])
    inner = num[n:]
    assert sorted(inner.tolist()) == list(range(n - 1))
    # replay: positions held by each cluster, merged in creation order
    pos_of_leaf = np.empty(n, np.int64)
    pos_of_leaf[lb["leaf_ids"]] = np.arange(n)
    held = {x: {int(pos_of_leaf[x])} for x in range(n)}
    for x in range(n, 2 * n - 1):
        a, b = held.pop(int(t["left"][x])), held.pop(int(t["right"][x]))
        assert num[x] == max(min(a), min(b)) - 1
        held[x] = a | b
    assert list(held) == [t["root"]] and min(held[t["root"]]) == 0
    # the same tree under both numberings
    by_creation = oracle.hploc_binary(t)
    by_position = oracle.hploc_binary(t, num)
    np.testing.assert_array_equal(by_creation["leaf_ids"], by_position["leaf_ids"])
    np.testing.assert_array_equal(
        oracle.preorder_binary(by_creation["root"], by_creation["node_meta"], by_creation["leaf_ids"]),
        oracle.preorder_binary(by_position["root"], by_position["node_meta"], by_position["leaf_ids"]),
    )
    relabel = np.empty(max(n - 1, 1), np.int64)
    relabel[np.arange(n - 1)] = inner[np.arange(n - 1)]  # creation index i -> position index
    m1 = by_creation["node_meta"].reshape(-1, 4)
    m2 = by_position["node_meta"].reshape(-1, 4)
    np.testing.assert_array_equal(m2[relabel], np.where(
        (m1 & oracle.LEAF_FLAG) | (m1 == oracle.SENTINEL) | (np.arange(4) == 3), m1,
        relabel[np.minimum(m1, n - 2)]).astype(np.uint32))
    np.testing.assert_array_equal(by_position["node_bounds"].reshape(-1, 12)[relabel],
                                  by_creation["node_bounds"].reshape(-1, 12))
    np.testing.assert_array_equal(by_position["leaf_parent"], relabel[by_creation["leaf_parent"]])


def test_threshold_32_is_the_limit():
    """Lists of two halves at threshold 32 fill 64 slots (one wave); 33 is refused."""
    v = soup(200, seed=11)
    lb = oracle.build_lbvh(v)
    codes = (lb["sorted_keys"] >> np.uint64(32)).astype(np.uint32)
    t = oracle.hploc_topology(oracle.triangle_bounds(v), codes, lb["leaf_ids"], merging_threshold=32)
    validate(t, 200)
    with pytest.raises(RuntimeError):
        oracle.hploc_topology(oracle.triangle_bounds(v), codes, lb["leaf_ids"], merging_threshold=33)

"""Structural checks of an 8-wide hierarchy (device build or oracle collapse).

Follows the invariants bajo checks on its wide collapse (every leaf reached
exactly once, test/bvh/gpu/test_hploc_wide.mojo) plus the box guarantees of
the quantized layout: every dequantized child box contains the exact bounds
of the triangles beneath it (in exact arithmetic) and is at most two grid
steps wider than them on each side.
"""

from __future__ import annotations

import numpy as np

WIDE_LEAF = 0x80000000
START_MASK = 0x0FFFFFFF


def scale(exp: np.ndarray) -> np.ndarray:
    """2^(e - 127) for biased exponents e (float64)."""
    return np.ldexp(1.0, exp.astype(np.int64) - 127)


def validate_wide(nodes: np.ndarray, vertices: np.ndarray, leaf_ids: np.ndarray, leaf_size: int) -> dict:
    n = leaf_ids.size
    tri = np.asarray(vertices, np.float32).reshape(-1, 3, 3)[leaf_ids].astype(np.float64)  # sorted order
    tri_lo, tri_hi = tri.min(axis=1), tri.max(axis=1)
    count = len(nodes)
    assert count >= 1
    covered = np.zeros(n, np.int64)
    refs = np.zeros(count, np.int64)
    refs[0] = 1
    box_lo = np.full((count, 3), np.inf)
    box_hi = np.full((count, 3), -np.inf)
    first = np.full(count, -1, np.int64)
    tris_under = np.zeros(count, np.int64)
    max_slack = 0.0
    leaf_children = inner_children = 0
    # children always carry larger indices than their parent, so a reverse
    # sweep sees every inner child before its parent
    for k in range(count - 1, -1, -1):
        nd = nodes[k]
        c = int(nd["count"])
        assert 1 <= c <= 8, (k, c)
        if n > 1:
            assert c >= 2, (k, c)
        o = nd["origin"].astype(np.float64)
        s = scale(nd["exp"])
        lo_q = nd["lo"].astype(np.float64)  # [3, 8]
        hi_q = nd["hi"].astype(np.float64)
        spans = []
        for i in range(c):
            ref = int(nd["child"][i])
            if ref & WIDE_LEAF:
                start = ref & START_MASK
                size = ((ref >> 28) & 7) + 1
                assert size <= max(leaf_size, 1) and start + size <= n, (k, i, start, size)
                covered[start : start + size] += 1
                clo, chi = tri_lo[start : start + size].min(axis=0), tri_hi[start : start + size].max(axis=0)
                spans.append((start, size))
                leaf_children += 1
            else:
                assert k < ref < count, (k, ref)
                refs[ref] += 1
                assert tris_under[ref] > leaf_size, (ref, tris_under[ref])
                clo, chi = box_lo[ref], box_hi[ref]
                spans.append((first[ref], tris_under[ref]))
                inner_children += 1
            dlo = o + lo_q[:, i] * s
            dhi = o + hi_q[:, i] * s
            assert np.all(dlo <= clo) and np.all(dhi >= chi), (k, i, dlo, clo, dhi, chi)
            slack = max(np.max((clo - dlo) / s), np.max((dhi - chi) / s))
            max_slack = max(max_slack, float(slack))
            box_lo[k] = np.minimum(box_lo[k], clo)
            box_hi[k] = np.maximum(box_hi[k], chi)
        # unused slots hold an empty box
        for i in range(c, 8):
            assert np.all(nd["lo"][:, i] == 255) and np.all(nd["hi"][:, i] == 0) and nd["child"][i] == 0
        spans.sort()
        first[k] = spans[0][0]
        tris_under[k] = sum(sz for _, sz in spans)
        # the children of a node partition one contiguous sorted range
        pos = spans[0][0]
        for st, sz in spans:
            assert st == pos, (k, spans)
            pos += sz
        # the node origin is the lower corner of its children's boxes
        assert np.all(o == box_lo[k]) or n == 1, (k, o, box_lo[k])
    assert np.all(covered == 1), "every sorted triangle is referenced by exactly one leaf child"
    assert np.all(refs == 1), "every node except the root has exactly one parent"
    assert first[0] == 0 and tris_under[0] == n
    # floor/ceil plus one step of padding: at most two steps beyond the exact box
    assert max_slack <= 2.0 + 1e-6, max_slack
    return {"nodes": count, "leaf_children": leaf_children, "inner_children": inner_children, "max_slack": max_slack}


def compare_hits(got: np.ndarray, want: np.ndarray, min_identical: float = 0.999, rel_t: float = 1e-4) -> dict:
    """Closest hits of two traversals that may disagree only on near-ties: the
    same prim and t on all but a small fraction of rays, and on those the hit
    distances agree to `rel_t` (rays grazing a shared edge or box face)."""
    got = got.reshape(-1, 8)
    want = want.reshape(-1, 8)
    gp, wp = got[:, 2].view(np.uint32), want[:, 2].view(np.uint32)
    same = (gp == wp) & (got[:, 7].view(np.uint32) == want[:, 7].view(np.uint32))
    frac = float(same.mean())
    bad = ~same
    both = bad & (gp != 0xFFFFFFFF) & (wp != 0xFFFFFFFF)
    if both.any():
        t1, t2 = got[both, 7].astype(np.float64), want[both, 7].astype(np.float64)
        assert np.all(np.abs(t1 - t2) <= rel_t * np.maximum(np.abs(t2), 1e-3)), (t1[:8], t2[:8])
    assert frac >= min_identical, frac
    return {"identical": frac, "differing": int(bad.sum()), "hit_vs_miss": int((bad & ~both).sum())}

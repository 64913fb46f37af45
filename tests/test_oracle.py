"""Pins the CPU oracle (oracle/) against bajo's known answers and against
independent restatements (closed forms, brute force) before it is trusted as
the checker for the GPU path.  CPU only."""

import json
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def soup(n, seed=7):
    """Small random triangle soups (numpy PCG64, independent of the product generator)."""
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3), dtype=np.float32)
    e = (rng.random((n, 2, 3), dtype=np.float32) - 0.5) * 0.1
    v = np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32)
    return v.reshape(-1)


# ---- Morton (test/core/test_morton.mojo) -------------------------------------

def test_expand_bits_known_answers():
    for x, want in GOLD["expand_bits_3d"]["cases"]:
        assert oracle.lib().orc_expand_bits(x) == want


def test_expand_bits_is_bit_interleave():
    # closed form: bit k of x lands on bit 3k
    for x in [0, 1, 2, 3, 5, 511, 512, 1023, 0x2AA, 0x155]:
        want = sum(((x >> k) & 1) << (3 * k) for k in range(10))
        assert oracle.lib().orc_expand_bits(x) == want


def test_morton3_known_answers():
    for (x, y, z), want in GOLD["morton3"]["cases"]:
        assert oracle.lib().orc_morton(x, y, z) == want


def test_morton3_clamps_and_truncates():
    m = oracle.lib().orc_morton
    assert m(1.5, 2.0, 100.0) == (1 << 30) - 1  # clamp to 1023 per axis
    assert m(0.99999994, 0.0, 0.0) == 0x09249249  # 1023 spread on x
    assert m(1.0 / 1024 - 1e-9, 0.0, 0.0) == 0  # truncation, not rounding
    assert m(0.5, 0.0, 0.0) == oracle.lib().orc_expand_bits(512)


def test_common_prefix_known_answers():
    for args, want in GOLD["morton_common_prefix"]["cases"]:
        assert oracle.lib().orc_common_prefix(*args) == want


def test_morton_codes_normalization_matches_numpy():
    v = soup(3000, seed=1)
    b = oracle.triangle_bounds(v)
    sb = oracle.scene_bounds(b)
    codes = oracle.morton_codes(b, sb[6:12])
    # numpy float32 restatement of lbvh.mojo:46-58 + core/morton.mojo:61-89
    bb = b.reshape(-1, 6)
    c = (bb[:, :3] + bb[:, 3:]) * np.float32(0.5)
    cmin, cmax = sb[6:9], sb[9:12]
    ext = (cmax - cmin).astype(np.float32)
    inv = np.where(np.abs(ext) > np.float32(1e-20), np.float32(1) / ext, np.float32(0)).astype(np.float32)
    q = ((c - cmin) * inv).astype(np.float32)
    u = np.clip(np.floor(q * np.float32(1024)), 0, 1023).astype(np.uint32)

    def spread(x):
        out = np.zeros_like(x)
        for k in range(10):
            out |= ((x >> k) & 1) << (3 * k)
        return out

    want = (spread(u[:, 2]) << 2) | (spread(u[:, 1]) << 1) | spread(u[:, 0])
    np.testing.assert_array_equal(codes, want)


def test_scene_bounds_closed_form():
    v = soup(500, seed=2)
    b = oracle.triangle_bounds(v)
    t = v.reshape(-1, 3, 3)
    np.testing.assert_array_equal(b.reshape(-1, 6)[:, :3], t.min(axis=1))
    np.testing.assert_array_equal(b.reshape(-1, 6)[:, 3:], t.max(axis=1))
    sb = oracle.scene_bounds(b)
    bb = b.reshape(-1, 6)
    c = (bb[:, :3] + bb[:, 3:]) * np.float32(0.5)
    np.testing.assert_array_equal(sb[:3], bb[:, :3].min(0))
    np.testing.assert_array_equal(sb[3:6], bb[:, 3:].max(0))
    np.testing.assert_array_equal(sb[6:9], c.min(0))
    np.testing.assert_array_equal(sb[9:12], c.max(0))


# ---- intersection (test/core/test_intersect.mojo) ----------------------------

def test_ray_aabb_known_answers():
    for case in GOLD["ray_aabb"]["cases"]:
        t = np.zeros(1, np.float32)
        hit = oracle.lib().orc_intersect_aabb(
            oracle._p(oracle.f32(case["o"])), oracle._p(oracle.f32(case["rcp"])), oracle._p(oracle.f32(case["box"])),
            case["t_max"], oracle._p(t),
        )
        assert bool(hit) == case["hit"]
        if case["hit"]:
            assert abs(t[0] - case["t"]) <= case["atol"]


def _tri(fn, o, d, v0, e1, e2, t_max, t_min=1e-4, nout=3):
    out = np.zeros(nout, np.float32)
    m = fn(*(oracle._p(oracle.f32(x)) for x in (o, d, v0, e1, e2)), t_max, t_min, oracle._p(out))
    return bool(m), out


def test_ray_triangle_known_answers():
    c = GOLD["ray_triangle"]["packed_triangle"]
    v = np.array(c["v"], np.float32)
    hit, tuv = _tri(oracle.lib().orc_intersect_tri_eager, c["o"], c["d"], v[:3], v[3:6] - v[:3], v[6:] - v[:3], c["t_max"])
    assert hit and abs(tuv[0] - c["t"]) < 1e-6


def test_scaled_matches_eager_both_windings():
    c = GOLD["ray_triangle"]["scaled_vs_eager"]
    for e1, e2 in ((c["edge_x"], c["edge_y"]), (c["edge_y"], c["edge_x"])):
        he, tuv = _tri(oracle.lib().orc_intersect_tri_eager, c["o"], c["d"], c["v0"], e1, e2, c["t_max"])
        hs, dtuv = _tri(oracle.lib().orc_intersect_tri_scaled, c["o"], c["d"], c["v0"], e1, e2, c["t_max"], nout=4)
        assert he == hs
        inv = np.float32(1) / dtuv[0]
        np.testing.assert_allclose([dtuv[1] * inv, dtuv[2] * inv, dtuv[3] * inv], tuv, atol=c["atol"])
    he, _ = _tri(oracle.lib().orc_intersect_tri_eager, c["miss_origin"], c["d"], c["v0"], c["edge_x"], c["edge_y"], c["t_max"])
    hs, _ = _tri(oracle.lib().orc_intersect_tri_scaled, c["miss_origin"], c["d"], c["v0"], c["edge_x"], c["edge_y"], c["t_max"], nout=4)
    assert not he and not hs


def test_rcp_direction_epsilon():
    for d, want in GOLD["rcp_direction"]["cases"]:
        out = np.zeros(3, np.float32)
        oracle.lib().orc_rcp_direction(oracle._p(oracle.f32(d)), oracle._p(out))
        np.testing.assert_array_equal(out, np.array(want, np.float32))


# ---- sort (test/sort/*) ------------------------------------------------------

def composite_keys(size):
    i = np.arange(size, dtype=np.uint64)
    seg = (i * 7 % 19) & 0xFFFFFFFF
    mort = ((i * 2654435761) ^ (i << 11) ^ (i >> 3)) & 0xFFFFFFFF
    return (seg << np.uint64(32)) | mort


def test_sort_bitonic_fixture():
    g = GOLD["sort_bitonic_fixture"]
    k, v = oracle.radix_sort_pairs_u64(np.array(g["keys"], np.uint64), np.arange(8, dtype=np.uint32), 0, 32)
    assert k.tolist() == g["expected_keys"]
    assert v.tolist() == g["expected_values"]


def test_sort_composite_keys_stable():
    size = GOLD["sort_composite_keys"]["size"]
    keys = composite_keys(size)
    k, v = oracle.radix_sort_pairs_u64(keys, np.arange(size, dtype=np.uint32))
    np.testing.assert_array_equal(k, keys[v])
    assert np.all(k[1:] >= k[:-1])
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(v, order)


def test_sort_bit_range_is_stable_partial_sort():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 2**63, 5000, dtype=np.uint64)
    k, v = oracle.radix_sort_pairs_u64(keys, np.arange(5000, dtype=np.uint32), 16, 40)
    digit = (keys >> np.uint64(16)) & np.uint64((1 << 24) - 1)
    np.testing.assert_array_equal(v, np.argsort(digit, kind="stable"))


# ---- LBVH topology and refit -------------------------------------------------

def brute_split(codes, first, last):
    """Mathematical definition of the Karras split: the last position whose
    (code, index) key shares more leading bits with `first` than `last` does."""

    def cp(i, j):
        x = ((int(codes[i]) ^ int(codes[j])) << 32) | (i ^ j)
        return 64 if x == 0 else 64 - x.bit_length()

    node = cp(first, last)
    return max(s for s in range(first, last) if cp(first, s) > node) if any(
        cp(first, s) > node for s in range(first, last)
    ) else first


def brute_tree(codes, first, last, out):
    if first == last:
        return
    s = brute_split(codes, first, last)
    out.append((first, last, s))
    brute_tree(codes, first, s, out)
    brute_tree(codes, s + 1, last, out)


@pytest.mark.parametrize("n,seed,dup", [(2, 0, 0), (3, 1, 0), (7, 2, 0), (12, 3, 0), (12, 4, 1), (40, 5, 2)])
def test_karras_matches_brute_force(n, seed, dup):
    rng = np.random.default_rng(seed)
    codes = np.sort(rng.integers(0, 1 << 30, n).astype(np.uint32) >> (dup * 8)).astype(np.uint32)
    root, meta, parent, counts = oracle.karras_topology(codes)
    want = []
    brute_tree(codes, 0, n - 1, want)
    got = oracle.canonical_karras(root, meta, counts)
    np.testing.assert_array_equal(got, np.array(want, np.uint32))


def check_topology_invariants(lb, n):
    """bvh/gpu/validate.mojo:62-124: one root, parent/child agreement, every leaf once."""
    meta = lb["node_meta"].reshape(-1, 4)
    m = n - 1
    assert lb["root"] == 0
    assert meta[0, 0] == oracle.SENTINEL
    seen = np.zeros(n, np.int32)
    for i in range(m):
        for c in meta[i, 1:3]:
            if c & oracle.LEAF_FLAG:
                leaf = c & 0x7FFFFFFF
                seen[leaf] += 1
                assert lb["leaf_parent"][leaf] == i
            else:
                assert meta[c, 0] == i
    assert np.all(seen == 1)
    assert np.sum(meta[1:, 0] == oracle.SENTINEL) == 0


def test_lbvh_invariants_and_refit_closed_form():
    n = 3000
    v = soup(n, seed=11)
    lb = oracle.build_lbvh(v)
    check_topology_invariants(lb, n)
    # sorted order == stable sort by code
    b = oracle.triangle_bounds(v)
    codes = oracle.morton_codes(b, oracle.scene_bounds(b)[6:])
    np.testing.assert_array_equal(lb["leaf_ids"], np.argsort(codes, kind="stable"))
    np.testing.assert_array_equal(lb["sorted_keys"] >> np.uint64(32), codes[lb["leaf_ids"]])
    # refit: each child box is the exact min/max over its sorted-leaf range
    canon = oracle.canonical_karras(lb["root"], lb["node_meta"], lb["leaf_counts"])
    bb = b.reshape(-1, 6)[lb["leaf_ids"]]
    meta = lb["node_meta"].reshape(-1, 4)
    nb = lb["node_bounds"].reshape(-1, 12)
    # walk nodes by index: a node's range is [first, fence]; first from leaf_counts
    for i in range(n - 1):
        last = int(meta[i, 3])
        first = last - int(lb["leaf_counts"][i]) + 1
        for half, c in enumerate(meta[i, 1:3]):
            if c & oracle.LEAF_FLAG:
                lo = hi = int(c & 0x7FFFFFFF)
            else:
                hi = int(meta[c, 3])
                lo = hi - int(lb["leaf_counts"][c]) + 1
            assert first <= lo <= hi <= last
            box = nb[i, 6 * half: 6 * half + 6]
            np.testing.assert_array_equal(box[:3], bb[lo: hi + 1, :3].min(0))
            np.testing.assert_array_equal(box[3:], bb[lo: hi + 1, 3:].max(0))
    assert canon.shape == (n - 1, 3)


@pytest.mark.parametrize("n", [2, 5, 100, 2000])
def test_karras_equals_topdown_cpu_lbvh(n):
    v = soup(n, seed=n)
    lb = oracle.build_lbvh(v)
    b = oracle.triangle_bounds(v)
    codes = (lb["sorted_keys"] >> np.uint64(32)).astype(np.uint32)
    td = oracle.build_lbvh_topdown(codes, lb["leaf_ids"], b, leaf_size=1)
    np.testing.assert_array_equal(
        oracle.canonical_karras(lb["root"], lb["node_meta"], lb["leaf_counts"]), oracle.canonical_topdown(td)
    )
    # root box of the top-down build equals the union of the Karras root's children
    nb = lb["node_bounds"].reshape(-1, 12)[0]
    np.testing.assert_array_equal(td["boxes"][:3], np.minimum(nb[:3], nb[6:9]))
    np.testing.assert_array_equal(td["boxes"][3:6], np.maximum(nb[3:6], nb[9:12]))


def test_duplicate_centroids_give_valid_tree():
    # test/bvh/fixtures.mojo:72-78: twelve identical triangles
    t = np.array([-0.5, -0.5, 2.0, 0.5, -0.5, 2.0, 0.0, 0.5, 2.0], np.float32)
    v = np.tile(t, 12)
    lb = oracle.build_lbvh(v)
    check_topology_invariants(lb, 12)
    np.testing.assert_array_equal(lb["leaf_ids"], np.arange(12))


def test_leaf_size_four_topdown_covers_every_item():
    n = 1000
    v = soup(n, seed=4)
    lb = oracle.build_lbvh(v)
    b = oracle.triangle_bounds(v)
    codes = (lb["sorted_keys"] >> np.uint64(32)).astype(np.uint32)
    td = oracle.build_lbvh_topdown(codes, lb["leaf_ids"], b, leaf_size=4)
    leaves = td["count"] > 0
    assert td["count"][leaves].max() <= 4
    assert td["count"][leaves].sum() == n


# ---- traversal -----------------------------------------------------------------

def random_rays(n, lo, hi, seed, t_min=1e-4):
    rng = np.random.default_rng(seed)
    o = lo + (hi - lo) * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((8, n), np.float32)
    rays[0:3] = o.T
    rays[3] = t_min
    rays[4:7] = d.T
    rays[7] = np.finfo(np.float32).max
    return rays.reshape(-1)


def edge_tie(hits_a, hits_b, rows):
    """A hit within 1e-5 of a triangle edge may be accepted by one MT variant only."""
    ha, hb = hits_a.reshape(-1, 8)[rows], hits_b.reshape(-1, 8)[rows]
    def near(h):
        u, v = h[:, 0], h[:, 1]
        return (np.abs(u) < 1e-5) | (np.abs(v) < 1e-5) | (np.abs(u + v - 1) < 1e-5)
    return near(ha) | near(hb)


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (64, 2), (1500, 3)])
def test_bvh_trace_matches_brute_force(n, seed):
    v = soup(n, seed=seed)
    lb = oracle.build_lbvh(v)
    rays = random_rays(2000, -0.1, 1.1, seed + 100)
    got = oracle.trace_bvh2(v, lb, rays)
    want = oracle.trace_brute(v, rays)
    gp, wp = oracle.hit_prims(got), oracle.hit_prims(want)
    gt, wt = oracle.hit_t(got), oracle.hit_t(want)
    bad = np.nonzero((gp != wp) & ~np.isclose(gt, wt, rtol=1e-6, atol=0))[0]
    assert len(bad) == 0 or np.all(edge_tie(got, want, bad)), bad[:10]
    both = (gp == wp) & (gp != oracle.SENTINEL)
    np.testing.assert_array_equal(gt[both], wt[both])
    if n >= 64:
        assert both.sum() > 0


def test_any_hit_agrees_with_closest_hit():
    v = soup(800, seed=9)
    lb = oracle.build_lbvh(v)
    rays = random_rays(3000, 0.0, 1.0, 5)
    closest = oracle.trace_bvh2(v, lb, rays, mode=0)
    anyh = oracle.trace_bvh2(v, lb, rays, mode=1).reshape(-1, 8)
    occluded = (anyh[:, 2].view(np.uint32) == oracle.SENTINEL) & (anyh[:, 7] == 0)
    hit = oracle.hit_prims(closest) != oracle.SENTINEL
    np.testing.assert_array_equal(occluded, hit)


def test_camera_rays_follow_make_ray_raster():
    cam = oracle.camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 0.41421356)
    np.testing.assert_allclose(cam[3:6], [0, 0, 1], atol=1e-7)
    np.testing.assert_allclose(cam[6:9], [-1, 0, 0], atol=1e-7)  # right = forward x up
    np.testing.assert_allclose(cam[9:12], [0, 1, 0], atol=1e-7)
    w, h = 8, 4
    rays = oracle.camera_rays(cam, w * h, w, h).reshape(8, -1)
    # pixel (px, py): screen ((2(px+.5)-w)/h, 1-2(py+.5)/h), direction normalized in float64 for the check
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    sx = (2 * (px.ravel() + 0.5) - w) / h
    sy = 1 - 2 * (py.ravel() + 0.5) / h
    d = np.stack([-sx * 0.41421356, sy * 0.41421356, np.ones_like(sx)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    np.testing.assert_allclose(rays[4:7].T, d, atol=2e-7)
    assert np.all(rays[3] == 0) and np.all(rays[7] == np.finfo(np.float32).max)


def test_grid_fixture_checksum_equals_bajo_published():
    """bajo's 65,536-triangle grid fixture (bajo/benchmark/bvh_fixtures.mojo:11-58) traced closest-hit
    through the oracle's LBVH: the published checksum of bench/results/bvh_cpu/bvh_cpu.md (hits 196,608,
    checksum 6443188224.000)."""
    from bajo_amd import scenes

    v, rays = scenes.grid_fixture()
    hits = oracle.trace_bvh2(v, oracle.build_lbvh(v), rays)
    assert int((oracle.hit_prims(hits) != oracle.SENTINEL).sum()) == 196608
    assert scenes.grid_checksum(hits) == scenes.GRID_CHECKSUM

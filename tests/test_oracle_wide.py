"""Pins the oracle's 8-wide collapse and traversal (oracle/orc_wide.c) before
it is used as the checker for the device build.  CPU only.

- grid exponent and quantization against their closed forms;
- the collapsed hierarchy against the structural invariants bajo checks on
  its wide collapse (every triangle reached once, test/bvh/gpu/
  test_hploc_wide.mojo) and the quantized-box guarantees (containment in
  exact arithmetic, at most two grid steps of slack);
- closest and any hit over the 8-wide nodes against the binary oracle
  traversal and brute force (the layouts differ, so rays grazing a shared
  edge or box face may resolve to an equally close neighbour);
- the oracle renderer over the 8-wide nodes against the binary one.
"""

import numpy as np
import pytest

import oracle
from test_oracle import random_rays, soup
from wide_check import compare_hits, validate_wide


def test_grid_exponent_closed_forms():
    ge = oracle.grid_exponent
    assert ge(0.0, 254.0) == 127  # 254 steps of 1.0 exactly
    assert ge(0.0, 254.5) == 128  # needs steps of 2.0
    assert ge(0.0, 127.0) == 126  # 254 steps of 0.5
    assert ge(1.0, 1.0) == 1  # zero extent: smallest normal grid
    assert ge(-3.0, 5.0) == 127 - 4  # 8 / 254 -> 2^-4 * 254 = 15.875 >= 8, 2^-5 * 254 = 7.9375 < 8
    rng = np.random.default_rng(4)
    for _ in range(2000):
        o = np.float32(rng.normal() * 10.0 ** rng.integers(-3, 4))
        ext = np.float32(abs(rng.normal()) * 10.0 ** rng.integers(-6, 3))
        hi = np.float32(o + ext)
        e = ge(float(o), float(hi))
        s = np.ldexp(1.0, e - 127)
        # o + 254 s >= hi in FP32 (the test the traversal's grid relies on) ...
        assert np.float32(o) + np.float32(254.0) * np.float32(s) >= hi
        # ... and no finer grid spans the extent in exact arithmetic
        if e > 1:
            assert 254.0 * (s / 2) < float(hi) - float(o)


def test_quantize_is_conservative_and_tight():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        o = np.float32(rng.normal() * 100)
        ext = np.float32(abs(rng.normal()) * 10.0 ** rng.integers(-5, 2))
        top = np.float32(o + ext)
        e = oracle.grid_exponent(float(o), float(top))
        s = np.ldexp(1.0, e - 127)
        a, b = sorted(np.float32(o + rng.random(2) * (top - o)))
        a, b = max(a, o), min(b, top)
        ql, qh = oracle.quantize(float(o), e, float(a), float(b))
        assert 0 <= ql <= qh <= 255
        # exact arithmetic: the grid box contains [a, b] ...
        assert float(o) + ql * s <= float(a) and float(o) + qh * s >= float(b)
        # ... with at most two steps of slack on either side
        assert float(a) - (float(o) + ql * s) <= 2 * s + 1e-12 * abs(float(o))
        assert (float(o) + qh * s) - float(b) <= 2 * s + 1e-12 * abs(float(o))


def test_quantize_padding_steps():
    # an interior box on an exact grid gets exactly one step of padding
    ql, qh = oracle.quantize(0.0, 127, 10.0, 20.0)
    assert (ql, qh) == (9, 21)
    # clamped at the grid ends
    assert oracle.quantize(0.0, 127, 0.0, 254.0) == (0, 255)


@pytest.mark.parametrize("n,leaf", [(2, 4), (3, 1), (9, 4), (64, 4), (777, 1), (777, 4), (777, 8), (5000, 4)])
def test_collapse_invariants(n, leaf):
    v = soup(n, seed=n + leaf)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, leaf)
    info = validate_wide(w, v, lb["leaf_ids"], leaf)
    # every node below the root was opened from a subtree of more than `leaf`
    # triangles, so there are fewer nodes than n / leaf
    assert info["nodes"] <= max(1, n // max(leaf, 1))
    if n >= 64:
        # opening fills nodes: the average fan-out is well above binary (with
        # single-triangle leaves, subtrees of 2-7 triangles end as small nodes)
        fan_out = (info["leaf_children"] + info["inner_children"]) / info["nodes"]
        assert fan_out > (3.5 if leaf == 1 else 5.0), fan_out


def test_collapse_single_triangle():
    v = soup(1, seed=3)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, 4)
    assert len(w) == 1 and w[0]["count"] == 1 and w[0]["child"][0] == oracle.WIDE_LEAF
    validate_wide(w, v, lb["leaf_ids"], 4)
    rays = random_rays(500, -0.1, 1.1, 9)
    got, _, over = oracle.trace_wide(v, lb["leaf_ids"], w, rays)
    want = oracle.trace_bvh2(v, lb, rays)
    assert over == 0
    np.testing.assert_array_equal(got.view(np.uint32), want.reshape(-1, 8).view(np.uint32))


def test_collapse_duplicate_centroids():
    # 300 identical triangles: the Karras tree splits on the index bits
    one = soup(1, seed=11)
    v = np.tile(one, 300)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, 4)
    validate_wide(w, v, lb["leaf_ids"], 4)


def test_canonical_form_ignores_numbering():
    v = soup(3000, seed=12)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, 4)
    # renumber the nodes below the root with a random permutation
    rng = np.random.default_rng(0)
    perm = np.concatenate([[0], 1 + rng.permutation(len(w) - 1)])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(w))
    shuffled = w[inv].copy()
    for rec in shuffled:
        for j in range(rec["count"]):
            c = int(rec["child"][j])
            if not c & oracle.WIDE_LEAF:
                rec["child"][j] = perm[c]
    a, b = oracle.canonical_wide(w), oracle.canonical_wide(shuffled)
    assert a.tobytes() == b.tobytes()
    # and it does see a changed box byte
    shuffled[len(w) // 2]["hi"][0, 0] ^= 1
    assert oracle.canonical_wide(shuffled).tobytes() != a.tobytes()


@pytest.mark.parametrize("n,leaf,seed", [(64, 4, 1), (2000, 1, 2), (2000, 4, 3), (2000, 8, 4)])
def test_wide_trace_matches_binary_and_brute_force(n, leaf, seed):
    v = soup(n, seed=seed)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, leaf)
    rays = random_rays(4000, -0.1, 1.1, seed + 50)
    got, stats, over = oracle.trace_wide(v, lb["leaf_ids"], w, rays, with_stats=True)
    assert over == 0
    binary = oracle.trace_bvh2(v, lb, rays)
    compare_hits(got, binary, min_identical=0.998)
    brute = oracle.trace_brute(v, rays)
    gp, bp = got[:, 2].view(np.uint32), oracle.hit_prims(brute)
    assert np.mean(gp == bp) > 0.99
    # the counters are consistent: every ray visits the root; leaf tests are
    # bounded by the triangles of the visited leaves
    assert np.all(stats[:, 0] >= 1)
    assert np.all(stats[:, 3] <= stats[:, 0] * 8 * leaf)


def test_wide_any_hit_agrees_with_closest():
    v = soup(1500, seed=21)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, 4)
    rays = random_rays(3000, 0.0, 1.0, 6)
    closest, _, _ = oracle.trace_wide(v, lb["leaf_ids"], w, rays, mode=0)
    anyh, _, _ = oracle.trace_wide(v, lb["leaf_ids"], w, rays, mode=1)
    occluded = (anyh[:, 2].view(np.uint32) == oracle.SENTINEL) & (anyh[:, 7] == 0)
    hit = closest[:, 2].view(np.uint32) != oracle.SENTINEL
    np.testing.assert_array_equal(occluded, hit)


def test_wide_camera_trace_on_icosphere():
    from bajo_amd import scenes

    v = scenes.icosphere(20480, frequency=58)
    lb = oracle.build_lbvh(v)
    w = oracle.collapse_wide(lb, 4)
    validate_wide(w, v, lb["leaf_ids"], 4)
    cam = oracle.camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0, 1, 0), 0.41421356)
    rays = oracle.camera_rays(cam, 96 * 54, 96, 54)
    got, stats, over = oracle.trace_wide(v, lb["leaf_ids"], w, rays, with_stats=True)
    assert over == 0
    info = compare_hits(got, oracle.trace_bvh2(v, lb, rays), min_identical=0.999)
    print(info, "visits/ray", stats[:, 0].mean(), "tests/ray", stats[:, 3].mean())


def test_oracle_render_wide_matches_binary():
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera

    sc = scenes.box_scene()
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    args = (sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 32, 24, 4, 5, 2026, 3, cam, 8)
    a, ra = oracle.render(*args)
    b, rb = oracle.render(*args, node_width=8, leaf_size=4)
    same = np.mean(a.view(np.uint32) == b.view(np.uint32))
    assert same > 0.97, same
    assert np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) < 0.05


@pytest.mark.parametrize("n,leaf,seed", [(2, 4, 1), (64, 4, 2), (2000, 1, 3), (2000, 4, 4), (5000, 8, 5)])
def test_collapse_width4(n, leaf, seed):
    """Width 4: every node holds at most 4 children and opens its largest children first (nodes are fuller
    than binary), and traversal of the 4-wide hierarchy returns the binary hierarchy's hits."""
    v = soup(n, seed=seed)
    lb = oracle.build_lbvh(v)
    w4 = oracle.collapse_wide(lb, leaf, width=4)
    info = validate_wide(w4, v, lb["leaf_ids"], leaf)
    assert w4["count"].max() <= 4
    if n >= 64:
        fan_out = (info["leaf_children"] + info["inner_children"]) / info["nodes"]
        assert fan_out > (2.5 if leaf == 1 else 3.0), fan_out
    rays = random_rays(2000, -0.1, 1.1, seed + 11)
    got, _, over = oracle.trace_wide(v, lb["leaf_ids"], w4, rays)
    assert over == 0
    want = oracle.trace_bvh2(v, lb, rays)
    compare_hits(got.reshape(-1), want)

"""Device H-PLOC builder vs the oracle's restatement of bajo's CPU H-PLOC
topology (oracle/orc_hploc.c), the protocol of test/bvh/gpu/
test_hploc_binary.mojo: the device topology must equal the reference
topology exactly (child order included; compared through a preorder
serialization of leaf ids), node boxes must be the unions of their leaves,
and quality must match.  Beyond bajo's protocol the whole binary layout is
compared word for word: inner nodes are numbered by the counter-free
lowest-leaf rule, which the oracle restates, so the device output is
deterministic and equal to the oracle's.  Traversal over the device hierarchy is checked
bit-for-bit against the oracle traversing the same (exported) hierarchy, and
renders against the oracle renderer on its own H-PLOC hierarchy.
"""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host
from test_gpu_trace import random_rays
from wide_check import compare_hits, validate_wide

pytestmark = pytest.mark.gpu
HPLOC = 1


def soup(n, seed):
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3), dtype=np.float32)
    e = (rng.random((n, 2, 3), dtype=np.float32) - 0.5) * 0.05
    return np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)


def bajo_triangles(count):
    """The triangle generator of test/bvh/gpu/test_hploc_binary.mojo:236-251."""
    out = []
    for i in range(count):
        x = np.float32((i % 19) * 4 - 36)
        y = np.float32(((i // 19) % 14) * 4 - 26)
        z = np.float32((i * 7) % 11)
        s = np.float32(2.2) if i % 17 == 0 else np.float32(0.7)
        out += [x - s, y - s, z, x + s, y - s, z, x, y + s, z + np.float32(0.1)]
    return np.array(out, np.float32)


def check_binary(b, v):
    """Layout invariants: depth-first leaf ranges, fences, counts, parent
    links, child boxes equal to the unions of their leaves' boxes."""
    n = b["leaf_ids"].size
    assert sorted(b["leaf_ids"].tolist()) == list(range(n))
    if n == 1:
        return
    meta = b["node_meta"].reshape(-1, 4)
    nb = b["node_bounds"].reshape(-1, 12)
    tb = oracle.triangle_bounds(v).reshape(-1, 6)[b["leaf_ids"]]
    root = b["root"]
    assert meta[root, 0] == oracle.SENTINEL
    seen = np.zeros(n - 1, bool)
    stack = [(root, 0)]
    while stack:
        x, first = stack.pop()
        assert not seen[x]
        seen[x] = True
        cnt = b["leaf_counts"][x]
        assert meta[x, 3] == first + cnt - 1  # fence = last depth-first position
        pos = first
        for side in (1, 2):
            c = int(meta[x, side])
            if c & oracle.LEAF_FLAG:
                p = c & 0x7FFFFFFF
                assert p == pos and b["leaf_parent"][p] == x
                lo, hi = tb[p, :3], tb[p, 3:]
                pos += 1
            else:
                assert meta[c, 0] == x
                k = b["leaf_counts"][c]
                lo, hi = tb[pos : pos + k, :3].min(0), tb[pos : pos + k, 3:].max(0)
                stack.append((c, pos))
                pos += k
            box = nb[x, 6 * (side - 1) : 6 * side]
            np.testing.assert_array_equal(box[:3], lo)
            np.testing.assert_array_equal(box[3:], hi)
        assert pos == first + cnt
    assert seen.all()


def device_hploc(v, node_width=2, merging_threshold=0):
    from bajo_amd.bvh import build_triangle_bvh

    t = {}
    bvh = build_triangle_bvh(to_dev(v), timings=t, node_width=node_width, method=HPLOC,
                             merging_threshold=merging_threshold)
    b = bvh.binary()
    return bvh, b, t


@pytest.mark.parametrize(
    "name,n", [("one", 1), ("two", 2), ("three", 3), ("bajo-257", 257), ("soup-1000", 1000),
               ("soup-65536", 65536), ("icosphere-1M", 1 << 20)]
)
def test_hploc_topology_equals_reference(gpu_ctx, name, n):
    announce(f"hploc topology {name} n={n}")
    if name.startswith("bajo"):
        v = bajo_triangles(n)
    elif name.startswith("icosphere"):
        from bajo_amd import scenes

        v = scenes.icosphere(n, frequency=230)
    else:
        v = soup(n, n)
    bvh, got, t = device_hploc(v)
    want = oracle.build_hploc(v)
    if n <= 100_000:
        check_binary(got, v)
    a = oracle.preorder_binary(got["root"], got["node_meta"], got["leaf_ids"])
    w = oracle.preorder_topology(want["topology"])
    assert a.size == 2 * n - 1
    np.testing.assert_array_equal(a, w)
    if n >= 2:
        assert got["root"] == want["root"]
        for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
            np.testing.assert_array_equal(got[key], want[key], err_msg=key)
        np.testing.assert_array_equal(got["node_bounds"].view(np.uint32), want["node_bounds"].view(np.uint32))
        # keys follow their leaves to depth-first order
        lb_keys = oracle.build_lbvh(v)["sorted_keys"] if n <= 100_000 else None
        if lb_keys is not None:
            code_of = np.empty(n, np.uint64)
            code_of[(lb_keys & np.uint64(0xFFFFFFFF)).astype(np.int64)] = lb_keys >> np.uint64(32)
            ids = want["leaf_ids"].astype(np.uint64)
            np.testing.assert_array_equal(got["sorted_keys"], (code_of[ids.astype(np.int64)] << np.uint64(32)) | ids)
    print(name, {k: val / 1e6 for k, val in t.items()})
    bvh.free()


@pytest.mark.parametrize("name,n", [("bajo-257", 257), ("soup-65536", 65536), ("icosphere-1M", 1 << 20)])
def test_hploc_threshold32_equals_reference(gpu_ctx, name, n):
    """Merging threshold 32 (bajo's GPU build on wave64): two lists of up to 32 clusters fill the
    wave; the layout equals the oracle's at the same threshold word for word."""
    announce(f"hploc threshold 32 {name} n={n}")
    if name.startswith("bajo"):
        v = bajo_triangles(n)
    elif name.startswith("icosphere"):
        from bajo_amd import scenes

        v = scenes.icosphere(n, frequency=230)
    else:
        v = soup(n, n)
    bvh, got, t = device_hploc(v, merging_threshold=32)
    want = oracle.build_hploc(v, merging_threshold=32)
    if n <= 100_000:
        check_binary(got, v)
    assert got["root"] == want["root"]
    for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    np.testing.assert_array_equal(got["node_bounds"].view(np.uint32), want["node_bounds"].view(np.uint32))
    # a different tree from threshold 16's
    sixteen = oracle.build_hploc(v)
    assert not np.array_equal(sixteen["node_meta"], want["node_meta"])
    print(name, "threshold 32", {k: val / 1e6 for k, val in t.items()})
    bvh.free()


def test_hploc_threshold_is_validated(gpu_ctx):
    from bajo_amd.bvh import build_triangle_bvh

    with pytest.raises(RuntimeError, match="merging_threshold"):
        build_triangle_bvh(to_dev(soup(100, 1)), method=HPLOC, merging_threshold=24)


def test_hploc_quality_not_worse_than_lbvh(gpu_ctx):
    from bajo_amd.bvh import build_triangle_bvh

    announce("hploc quality")
    v = soup(50_000, 7)
    lb = build_triangle_bvh(to_dev(v)).binary()
    hp = device_hploc(v)[1]

    def sah(b):
        nb = b["node_bounds"].reshape(-1, 12).astype(np.float64)
        lo, hi = np.minimum(nb[:, :3], nb[:, 6:9]), np.maximum(nb[:, 3:6], nb[:, 9:12])
        d = hi - lo
        area = 2 * (d[:, 0] * d[:, 1] + d[:, 0] * d[:, 2] + d[:, 1] * d[:, 2])
        return area.sum() / area[b["root"]]

    q_lbvh, q_hploc = sah(lb), sah(hp)
    print("inner-node SAH ratio: LBVH", q_lbvh, "H-PLOC", q_hploc)
    assert q_hploc <= q_lbvh


@pytest.mark.parametrize("n,seed", [(2, 1), (500, 2), (65536, 3)])
def test_hploc_trace_equals_oracle_on_same_tree(gpu_ctx, n, seed):
    from bajo_amd.bvh import TRACE

    announce(f"hploc trace n={n}")
    v = soup(n, seed)
    bvh, b, _ = device_hploc(v)
    R = 100_000
    rays = random_rays(R, 0.0, 1.0, seed + 20, t_min=1e-4)
    d_rays = to_dev(rays)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    stats = torch.empty(5 * R, dtype=torch.int32, device="cuda:0")
    bvh.launch_rays(d_rays, hits, R, TRACE.CLOSEST_HIT, stats=stats)
    want, want_stats = oracle.trace_bvh2(v, b, rays, stats=True)
    got = to_host(hits, np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    st = to_host(stats, np.uint32).reshape(-1, 5)
    np.testing.assert_array_equal(st[:, [0, 1, 3, 4]], want_stats.reshape(-1, 5)[:, [0, 1, 3, 4]])
    # and the same scene through brute force: equal but for edge ties
    brute = oracle.trace_brute(v, rays)
    assert np.mean(oracle.hit_prims(got) == oracle.hit_prims(brute)) > 0.999
    bvh.launch_rays(d_rays, hits, R, TRACE.ANY_HIT)
    np.testing.assert_array_equal(
        to_host(hits, np.uint32), oracle.trace_bvh2(v, b, rays, mode=1).view(np.uint32)
    )
    bvh.free()


def test_hploc_wide_equals_oracle_collapse(gpu_ctx):
    announce("hploc + wide")
    n = 30_000
    v = soup(n, 11)
    bvh, b, _ = device_hploc(v, node_width=8)
    got = bvh.wide()
    b["scene_bounds"] = oracle.build_lbvh(v)["scene_bounds"]
    want = oracle.collapse_wide(b, 4)
    validate_wide(got, v, b["leaf_ids"], 4)
    assert oracle.canonical_wide(got).tobytes() == oracle.canonical_wide(want).tobytes()
    R = 50_000
    rays = random_rays(R, 0.0, 1.0, 5, t_min=1e-4)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    bvh.launch_rays(to_dev(rays), hits, R)
    wh, _, over = oracle.trace_wide(v, b["leaf_ids"], want, rays)
    assert over == 0
    np.testing.assert_array_equal(to_host(hits, np.uint32).reshape(-1, 8), wh.view(np.uint32))
    compare_hits(to_host(hits, np.float32), oracle.trace_brute(v, rays), min_identical=0.998)
    bvh.free()


@pytest.mark.parametrize("width", [2, 8])
def test_hploc_render_matches_oracle(gpu_ctx, width):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    n = 1 << 16
    announce(f"hploc render icosphere n={n} width={width}")
    sc = scenes.icosphere_scene(n, frequency=58)
    g = GpuScene(sc, gpu_ctx, node_width=width, method=HPLOC)
    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.41421356).flatten()
    w, h, spp, depth, seed = 48, 27, 4, 5, 2026
    res = enqueue_render_gpu(RenderSettings(w, h, spp, seed, depth), cam, g, 3, with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(
        sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed, 3, cam, 8,
        node_width=width, leaf_size=4, method=HPLOC,
    )
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    print(f"hploc width {width} NEE: rmse {rmse:.3e}, identical {exact:.4f}")
    assert rmse <= 1e-3 and exact > 0.99
    assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
    g.free()

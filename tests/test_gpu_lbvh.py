"""GPU LBVH build vs the CPU oracle: Morton keys, sorted order, Karras
topology, leaf parents/counts and refit boxes bit-exact (bvh/gpu/builder/lbvh.mojo
semantics), at small sizes, at 65,536 triangles and at 4,194,304 triangles."""

import numpy as np
import pytest

import oracle
from conftest import announce
from gpu_util import to_dev

pytestmark = pytest.mark.gpu


def soup_np(n, seed):
    rng = np.random.default_rng(seed)
    v0 = rng.random((n, 3), dtype=np.float32)
    e = (rng.random((n, 2, 3), dtype=np.float32) - 0.5) * 0.05
    return np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)


def assert_same_lbvh(got: dict, want: dict):
    assert got["root"] == want["root"]
    np.testing.assert_array_equal(got["sorted_keys"], want["sorted_keys"])
    np.testing.assert_array_equal(got["leaf_ids"], want["leaf_ids"])
    np.testing.assert_array_equal(got["node_meta"], want["node_meta"])
    np.testing.assert_array_equal(got["leaf_parent"], want["leaf_parent"])
    np.testing.assert_array_equal(got["leaf_counts"], want["leaf_counts"])
    # node boxes bit-exact (compare the raw bits)
    np.testing.assert_array_equal(got["node_bounds"].view(np.uint32), want["node_bounds"].view(np.uint32))
    # scene boxes: equal as floats (the order of a parallel min may pick -0 or +0)
    np.testing.assert_array_equal(got["scene_bounds"], want["scene_bounds"])


def build_both(v):
    from bajo_amd.bvh import build_triangle_bvh

    bvh = build_triangle_bvh(to_dev(v))
    return bvh, bvh.binary(), oracle.build_lbvh(v)


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (7, 3), (100, 4), (4097, 5), (20000, 6)])
def test_lbvh_bit_exact_small(gpu_ctx, n, seed):
    announce(f"lbvh small n={n}")
    v = soup_np(n, seed)
    _, got, want = build_both(v)
    assert_same_lbvh(got, want)


def test_lbvh_duplicate_centroids(gpu_ctx):
    # test/bvh/fixtures.mojo:72-78 -- twelve identical triangles
    announce("lbvh duplicate centroids")
    t = np.array([-0.5, -0.5, 2.0, 0.5, -0.5, 2.0, 0.0, 0.5, 2.0], np.float32)
    _, got, want = build_both(np.tile(t, 12))
    assert_same_lbvh(got, want)


def test_lbvh_strip_and_small_scene(gpu_ctx):
    # test/bvh/fixtures.mojo:29-69 scenes
    announce("lbvh fixture scenes")
    def tri(cx, cy, z):
        return [cx - 1, cy - 1, z, cx + 1, cy - 1, z, cx, cy + 1, z]
    strip = np.array(sum((tri(i * 4 - 16 * 2, 0, 2.0) for i in range(16)), []), np.float32)
    small = np.array(sum((tri(x, y, z) for z in (2.0, 4.0) for y in (-1.0, 1.0) for x in (-1.0, 1.0)), []), np.float32)
    for v in (strip, small):
        _, got, want = build_both(v)
        assert_same_lbvh(got, want)


def test_lbvh_soup_65536_bit_exact(gpu_ctx):
    from bajo_amd import scenes

    announce("lbvh soup 65536")
    v = scenes.soup(65536, initstate=2027, initseq=54)
    bvh, got, want = build_both(v)
    assert_same_lbvh(got, want)


def test_lbvh_rebuild_reuses_handle(gpu_ctx):
    announce("lbvh rebuild")
    from bajo_amd.bvh import build_triangle_bvh

    a, b = soup_np(5000, 1), soup_np(3000, 2)
    bvh = build_triangle_bvh(to_dev(a))
    t = bvh.rebuild(to_dev(b), timings=True)
    assert t["total_ns"] > 0
    assert_same_lbvh(bvh.binary(), oracle.build_lbvh(b))


def test_lbvh_icosphere_4m_bit_exact(gpu_ctx):
    from bajo_amd import scenes

    n = 4_194_304
    announce(f"lbvh icosphere n={n}")
    v = scenes.icosphere(n)
    bvh, got, want = build_both(v)
    assert_same_lbvh(got, want)
    # size-independent properties: sorted keys, prim ids form a permutation
    k = got["sorted_keys"]
    assert np.all(k[1:] > k[:-1])
    assert np.array_equal(np.sort(got["leaf_ids"]), np.arange(n, dtype=np.uint32))

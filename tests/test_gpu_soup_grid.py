"""The instanced-soup workload (a grid of 4,096-triangle soups flattened to one
mesh; 64 x 64 copies = 16.7M triangles at full size) at 8 x 8 copies: H-PLOC
build collapsed to 8-wide quantized nodes, incoherent closest-hit rays
(origins uniform in the scene box, directions uniform on the sphere) and 4
ambient-occlusion rays per hit traced any-hit (bajo's AO ray,
rt/gpu/path_shading.mojo:650-665, t in [0.001, 4]).

Checks: the H-PLOC binary layout equals the oracle's word for word; the
8-wide hierarchy equals the oracle's collapse of it (canonical order);
closest hits and AO rays are bit-exact against the oracle traversing the
exported hierarchy; AO occlusion is bit-exact; closest hits agree with brute
force up to exact-t ties (test/bvh/gpu/test_bounds_bvh.mojo:340-343).
"""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce
from gpu_util import to_dev, to_host
from wide_check import validate_wide

pytestmark = pytest.mark.gpu
F32_MAX = np.finfo(np.float32).max
HPLOC = 1


def box_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = lo + (hi - lo) * rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.empty((8, n), np.float32)
    rays[0:3] = o.T
    rays[3] = 0.0
    rays[4:7] = d.T
    rays[7] = F32_MAX
    return rays.reshape(-1)


@pytest.fixture(scope="module")
def grid():
    from bajo_amd import scenes

    v = scenes.soup_grid(8, 8, 4096)
    want = oracle.build_hploc(v)
    return v, want


def test_soup_grid_hploc_wide_closest_and_ao(gpu_ctx, grid):
    from bajo_amd.bvh import TRACE, build_triangle_bvh, make_ao_rays

    v, want = grid
    n = v.size // 9
    announce(f"soup grid 8x8x4096 n={n}: H-PLOC + 8-wide build")
    bvh = build_triangle_bvh(to_dev(v), node_width=8, leaf_size=4, method=HPLOC)
    b = bvh.binary()
    assert b["root"] == want["root"]
    for key in ("node_meta", "leaf_ids", "leaf_counts", "leaf_parent"):
        np.testing.assert_array_equal(b[key], want[key], err_msg=key)
    np.testing.assert_array_equal(b["node_bounds"].view(np.uint32), want["node_bounds"].view(np.uint32))
    wide = bvh.wide()
    validate_wide(wide, v, b["leaf_ids"], 4)
    assert oracle.canonical_wide(wide).tobytes() == oracle.canonical_wide(oracle.collapse_wide(want, 4)).tobytes()

    R = 100_000
    lo, hi = v.reshape(-1, 3).min(0), v.reshape(-1, 3).max(0)
    rays = box_rays(R, lo, hi, 33)
    d_rays = to_dev(rays)
    hits = torch.empty(8 * R, dtype=torch.float32, device="cuda:0")
    announce(f"soup grid closest hit R={R}")
    bvh.launch_rays(d_rays, hits, R, TRACE.CLOSEST_HIT)
    got = to_host(hits, np.float32)
    wh, _, over = oracle.trace_wide(v, b["leaf_ids"], wide, rays)
    assert over == 0
    np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 8), wh.view(np.uint32))
    sub = 500  # brute force over all triangles for the first rays
    brute = oracle.trace_brute(v, rays.reshape(8, R)[:, :sub].reshape(-1))
    gp, bp = oracle.hit_prims(got), oracle.hit_prims(brute)
    gt, bt = got.reshape(-1, 8)[:sub, 7], brute.reshape(-1, 8)[:, 7]
    assert not ((gp[:sub] != bp) & (gt != bt)).any()
    hit_frac = float(np.mean(gp != 0xFFFFFFFF))
    assert 0.05 < hit_frac < 0.999

    S, seed = 4, 2029
    announce(f"soup grid AO rays samples={S}")
    ao = make_ao_rays(d_rays, hits, R, S, seed)
    ao_h = to_host(ao, np.float32)
    want_ao = oracle.ao_rays(rays, got, S, seed)
    np.testing.assert_array_equal(ao_h.view(np.uint32), want_ao.view(np.uint32))
    occ = torch.empty(8 * R * S, dtype=torch.float32, device="cuda:0")
    bvh.launch_rays(ao, occ, R * S, TRACE.ANY_HIT)
    occ_h = to_host(occ, np.uint32).reshape(-1, 8)
    w_occ, _, over = oracle.trace_wide(v, b["leaf_ids"], wide, want_ao, mode=1)
    assert over == 0
    np.testing.assert_array_equal(occ_h, w_occ.view(np.uint32))
    # rays of missed primary rays (t_max = -1) report a miss
    missed = np.repeat(gp == 0xFFFFFFFF, S)
    assert (occ_h[missed, 2] == 0xFFFFFFFF).all()
    occluded = (occ_h[:, 2] == 0xFFFFFFFF) & (occ_h[:, 7] == 0)
    frac = float(occluded[~missed].mean())
    print(f"hits {hit_frac:.3f}, AO occluded {frac:.3f}")
    assert 0.0 < frac < 1.0
    bvh.free()

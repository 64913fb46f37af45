"""Oracle sphere queries (bajo's intersect_ray_sphere, core/intersect.mojo:309-350,
and the sphere leaf of bvh/gpu/sphere_bvh.mojo:423-478) pinned by closed forms
and by brute force over every sphere."""

import numpy as np
import pytest

import oracle

EMPTY = 0xFFFFFFFF


def rays_of(o, d, tmin=0.0, tmax=1e30):
    o = np.atleast_2d(np.asarray(o, np.float32))
    d = np.atleast_2d(np.asarray(d, np.float32))
    out = np.empty((8, o.shape[0]), np.float32)
    out[0:3] = o.T
    out[3] = tmin
    out[4:7] = d.T
    out[7] = tmax
    return out.reshape(-1)


def one(sphere, o, d, tmin=0.0, tmax=1e30):
    return oracle.trace_brute_spheres(np.asarray(sphere, np.float32), rays_of(o, d, tmin, tmax))[0]


def test_sphere_closed_forms():
    unit = [0, 0, 5, 1]
    h = one(unit, [0, 0, 0], [0, 0, 1])
    assert h[7] == 4.0 and h[2].view(np.uint32) == 0
    np.testing.assert_array_equal(h[4:7], [0, 0, -1])
    assert h[0] == 0.0 and h[1] == 0.0 and h[3].view(np.uint32) == EMPTY
    # from inside: the far root
    h = one(unit, [0, 0, 5], [0, 0, 1])
    assert h[7] == 1.0
    np.testing.assert_array_equal(h[4:7], [0, 0, 1])
    # t_min past the near root: the far root; t_max before the near root: a miss
    assert one(unit, [0, 0, 0], [0, 0, 1], tmin=4.5)[7] == 6.0
    assert one(unit, [0, 0, 0], [0, 0, 1], tmax=3.9)[2].view(np.uint32) == EMPTY
    # behind the origin, and a miss to the side
    assert one(unit, [0, 0, 10], [0, 0, 1])[2].view(np.uint32) == EMPTY
    assert one(unit, [2, 0, 0], [0, 0, 1])[2].view(np.uint32) == EMPTY
    # tangent ray (det = 0) hits at the tangent point
    h = one(unit, [1, 0, 0], [0, 0, 1])
    assert h[2].view(np.uint32) == 0 and h[7] == 5.0
    # unnormalized direction: t scales by 1 / |d|
    assert one(unit, [0, 0, 0], [0, 0, 2])[7] == 2.0


def test_sphere_hit_matches_float64_quadratic():
    rng = np.random.default_rng(3)
    for _ in range(500):
        c = rng.uniform(-2, 2, 3)
        r = rng.uniform(0.2, 1.5)
        o = rng.uniform(-6, 6, 3)
        d = c + rng.uniform(-1, 1, 3) * r - o
        d /= np.linalg.norm(d)
        h = one(np.r_[c, r], o, d)
        oc = o - c
        b = oc @ d
        disc = b * b - (oc @ oc - r * r)
        if disc < 0:
            continue
        ts = [t for t in (-b - np.sqrt(disc), -b + np.sqrt(disc)) if t > 0]
        if not ts or h[2].view(np.uint32) == EMPTY:
            continue
        assert h[7] == pytest.approx(ts[0], rel=2e-5, abs=2e-5)
        p = o + ts[0] * d
        n = (p - c) / np.linalg.norm(p - c)
        np.testing.assert_allclose(h[4:7], n, atol=2e-3)


@pytest.mark.parametrize("wide", [False, True])
def test_sphere_bvh_matches_brute_force(wide):
    rng = np.random.default_rng(8)
    n = 3000
    c = rng.uniform(0, 10, (n, 3))
    r = rng.uniform(0.02, 0.3, (n, 1))
    sp = np.concatenate([c, r], 1).astype(np.float32).reshape(-1)
    lb = oracle.build_lbvh_bounds(oracle.sphere_bounds(sp))
    R = 4000
    o = rng.uniform(0, 10, (R, 3))
    d = rng.normal(size=(R, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_of(o, d, 1e-4)
    ref = oracle.trace_brute_spheres(sp, rays)
    if wide:
        got, _, over = oracle.trace_spheres_wide(sp, lb["leaf_ids"], oracle.collapse_wide(lb, 4), rays)
        assert over == 0
    else:
        got = oracle.trace_spheres(sp, lb, rays)
    assert (ref[:, 2].view(np.uint32) != EMPTY).mean() > 0.3
    same = np.all(got.view(np.uint32) == ref.view(np.uint32), axis=1)
    assert same.mean() > 0.999, same.mean()
    np.testing.assert_array_equal(got[~same, 7], ref[~same, 7])  # differing rows are exact-t ties
    occ = oracle.trace_spheres_wide(sp, lb["leaf_ids"], oracle.collapse_wide(lb, 4), rays, mode=1)[0] if wide \
        else oracle.trace_spheres(sp, lb, rays, mode=1)
    np.testing.assert_array_equal(occ[:, 7] == 0.0, ref[:, 2].view(np.uint32) != EMPTY)

"""Oracle instances (orc_tlas.c) pinned against bajo's own TLAS cases
(test/bvh/cpu/test_tlas.mojo, test/core/test_aabb.mojo:42-61), linear
algebra, and world-space brute force over the instanced triangles."""

import numpy as np
import pytest

import oracle
from bajo_amd import scenes

EMPTY = 0xFFFFFFFF


def translation(tx=0.0, ty=0.0, tz=0.0):
    return np.array([[1, 0, 0, tx], [0, 1, 0, ty], [0, 0, 1, tz]], np.float32)


def rays_of(o, d, tmin=0.0, tmax=np.inf):
    o = np.atleast_2d(np.asarray(o, np.float32))
    d = np.atleast_2d(np.asarray(d, np.float32))
    R = o.shape[0]
    out = np.empty((8, R), np.float32)
    out[0:3] = o.T
    out[3] = tmin
    out[4:7] = d.T
    out[7] = tmax
    return out.reshape(-1)


def blas_of(v, wide=False):
    b = dict(oracle.build_lbvh(v))
    b["vertices"] = np.asarray(v, np.float32)
    if wide:
        b["wide"] = oracle.collapse_wide(b, 4)
    return b


def blas_box(v):
    t = np.asarray(v, np.float32).reshape(-1, 3)
    return np.concatenate([t.min(0), t.max(0)]).astype(np.float32)


def tlas_of(blases, xforms, inst_blas, method="lbvh"):
    boxes = oracle.tlas_instance_bounds(np.stack(xforms), np.stack([blas_box(b["vertices"]) for b in blases]),
                                        inst_blas)
    top = oracle.build_lbvh_bounds(boxes) if method == "lbvh" else oracle.build_hploc_bounds(boxes)
    return dict(blases=blases, transforms=np.stack(xforms).reshape(-1), inst_blas=np.asarray(inst_blas, np.uint32),
                top=top)


def trace(t, rays, mode=0):
    return oracle.tlas_trace(t["blases"], t["transforms"], t["inst_blas"], t["top"], rays, mode).reshape(-1, 8)


# bajo's test triangle (test/bvh/cpu/test_tlas.mojo:36-41): z = 2, facing +z
TRI = np.array([-1, -1, 2, 1, -1, 2, 0, 1, 2], np.float32)


def as_u32(x):
    return np.asarray(x, np.float32).view(np.uint32)


def test_instance_inverse_round_trip():
    # test/bvh/cpu/test_tlas.mojo:19-33: scale (2, 4, 5), point (1.5, -2, 3)
    m = np.diag([2.0, 4.0, 5.0]).astype(np.float32)
    m = np.concatenate([m, np.zeros((3, 1), np.float32)], 1)
    inv, ok = oracle.affine_inverse(m)
    assert ok
    p = np.array([1.5, -2.0, 3.0], np.float64)
    world = m[:, :3].astype(np.float64) @ p + m[:, 3]
    back = inv.reshape(3, 4)[:, :3].astype(np.float64) @ world + inv.reshape(3, 4)[:, 3]
    np.testing.assert_allclose(back, p, atol=1e-6)


def test_affine_inverse_matches_linear_algebra():
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = rng.uniform(-1, 1, (3, 3)) + np.eye(3) * 2.0
        t = rng.uniform(-10, 10, 3)
        m = np.concatenate([a, t[:, None]], 1).astype(np.float32)
        inv, ok = oracle.affine_inverse(m)
        assert ok
        full = np.eye(4)
        full[:3] = m.astype(np.float64)
        want = np.linalg.inv(full)[:3]
        np.testing.assert_allclose(inv.reshape(3, 4), want, rtol=1e-5, atol=1e-5)


def test_affine_inverse_rejects_singular():
    m = np.array([[1, 2, 3, 0], [2, 4, 6, 0], [0, 0, 1, 0]], np.float32)  # rank 2
    _, ok = oracle.affine_inverse(m)
    assert not ok
    _, ok = oracle.affine_inverse(np.diag([1e-3, 1e-3, 1e-3, 0])[:3].astype(np.float32))  # det 1e-9 <= 1e-8
    assert not ok


def test_transform_box_rotation_golden():
    # test/core/test_aabb.mojo:42-61: [-1, 1]^3 turned 45 degrees about z
    c = np.float32(np.cos(np.pi / 4))
    m = np.array([[c, -c, 0, 0], [c, c, 0, 0], [0, 0, 1, 0]], np.float32)
    b = oracle.transform_box(m, [-1, -1, -1, 1, 1, 1])
    r2 = np.sqrt(2.0)
    np.testing.assert_allclose(b, [-r2, -r2, -1, r2, r2, 1], rtol=1e-6)


def test_transform_box_is_the_box_of_the_corners():
    rng = np.random.default_rng(9)
    for _ in range(200):
        m = rng.uniform(-2, 2, (3, 4)).astype(np.float32)
        lo = rng.uniform(-3, 3, 3)
        box = np.concatenate([lo, lo + rng.uniform(0, 2, 3)]).astype(np.float32)
        got = oracle.transform_box(m, box).astype(np.float64)
        corners = np.array([[box[3 * ((k >> j) & 1) + j] for j in range(3)] for k in range(8)], np.float64)
        w = corners @ m[:, :3].astype(np.float64).T + m[:, 3]
        np.testing.assert_allclose(got[:3], w.min(0), atol=1e-5)
        np.testing.assert_allclose(got[3:], w.max(0), atol=1e-5)


@pytest.mark.parametrize("wide", [False, True])
def test_single_instance_cases(wide):
    # test/bvh/cpu/test_tlas.mojo:82-101
    b = blas_of(TRI, wide)
    ident = tlas_of([b], [translation()], [0])
    h = trace(ident, rays_of([0, 0, 0], [0, 0, 1]))[0]
    assert as_u32(h[2]) == 0 and as_u32(h[3]) == 0
    assert h[7] == pytest.approx(2.0)
    assert h[6] == pytest.approx(1.0)
    moved = tlas_of([b], [translation(5.0)], [0])
    h = trace(moved, rays_of([[5, 0, 0], [0, 0, 0]], [[0, 0, 1], [0, 0, 1]]))
    assert as_u32(h[0, 2]) == 0 and as_u32(h[0, 3]) == 0 and h[0, 7] == pytest.approx(2.0)
    assert as_u32(h[1, 2]) == EMPTY


@pytest.mark.parametrize("wide", [False, True])
def test_two_and_three_instance_cases(wide):
    # test/bvh/cpu/test_tlas.mojo:104-155
    b0, b1 = blas_of(TRI, wide), blas_of(TRI, wide)
    near_far = tlas_of([b0, b1], [translation(), translation(tz=6.0)], [0, 1])
    h = trace(near_far, rays_of([0, 0, 0], [0, 0, 1]))[0]
    assert as_u32(h[3]) == 0 and h[7] == pytest.approx(2.0)
    left_right = tlas_of([b0, b1], [translation(-5.0), translation(5.0)], [0, 1])
    h = trace(left_right, rays_of([5, 0, 0], [0, 0, 1]))[0]
    assert as_u32(h[3]) == 1 and as_u32(h[2]) == 0 and h[7] == pytest.approx(2.0)
    three = tlas_of([b0], [translation(-5.0), translation(5.0), translation(10.0)], [0, 0, 0])
    h = trace(three, rays_of([5, 0, 0], [0, 0, 1]))[0]
    assert as_u32(h[3]) == 1 and h[7] == pytest.approx(2.0)
    a = trace(three, rays_of([5, 0, 0], [0, 0, 1]), mode=1)[0]
    assert a[7] == 0.0 and as_u32(a[2]) == EMPTY and as_u32(a[3]) == 1  # bajo's shadow hit carrying the instance


def test_shadow_cases():
    # test/bvh/cpu/test_tlas.mojo:175-185
    t = tlas_of([blas_of(TRI)], [translation(5.0)], [0])
    a = trace(t, rays_of([[5, 0, 0], [0, 0, 0]], [[0, 0, 1], [0, 0, 1]], tmax=1e30), mode=1)
    assert a[0, 7] == 0.0
    assert a[1, 7] == np.float32(1e30) and as_u32(a[1, 3]) == EMPTY


def test_nonuniform_scale_normal():
    # a tilted triangle under a non-uniform scale: the world normal is the
    # normal of the transformed triangle (inverse transpose of the local one)
    tri = np.array([0, 0, 1, 1, 0, 2, 0, 1, 1], np.float32)
    m = np.array([[2, 0, 0, 0.3], [0, 1, 0, -0.2], [0, 0, 0.5, 0.1]], np.float32)
    t = tlas_of([blas_of(tri)], [m], [0])
    world = (tri.reshape(3, 3).astype(np.float64) @ m[:, :3].T.astype(np.float64) + m[:, 3])
    n = np.cross(world[1] - world[0], world[2] - world[0])
    n /= np.linalg.norm(n)
    c = world.mean(0)
    o = c - 3.0 * n
    h = trace(t, rays_of(o, n))[0]
    assert as_u32(h[2]) == 0
    assert h[7] == pytest.approx(3.0, rel=1e-5)
    np.testing.assert_allclose(np.abs(h[4:7]), np.abs(n), atol=2e-6)


def world_soup(blas_verts, xforms, inst_blas):
    parts, owner = [], []
    for i, b in enumerate(inst_blas):
        v = np.asarray(blas_verts[b], np.float64).reshape(-1, 3)
        m = np.asarray(xforms[i], np.float64).reshape(3, 4)
        parts.append((v @ m[:, :3].T + m[:, 3]).astype(np.float32).reshape(-1))
        n = v.shape[0] // 3
        owner.append(np.stack([np.full(n, i), np.arange(n)], 1))
    return np.concatenate(parts), np.concatenate(owner)


def random_instances(rng, n_inst, n_blas):
    xforms, idx = [], []
    for i in range(n_inst):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        s = np.diag(rng.uniform(0.5, 1.5, 3))
        a = q @ s
        t = rng.uniform(-4, 4, 3)
        xforms.append(np.concatenate([a, t[:, None]], 1).astype(np.float32))
        idx.append(int(rng.integers(0, n_blas)))
    return xforms, idx


@pytest.mark.parametrize("wide,method", [(False, "lbvh"), (True, "lbvh"), (False, "hploc")])
def test_random_instances_match_world_brute_force(wide, method):
    rng = np.random.default_rng(17 + wide)
    verts = [2.0 * (scenes.soup(n, initstate=300 + k, edge_min=0.2, edge_max=0.8) - 0.5)
             for k, n in enumerate((60, 300, 1))]
    blases = [blas_of(v, wide) for v in verts]
    xforms, idx = random_instances(rng, 48, len(verts))
    t = tlas_of(blases, xforms, idx, method)
    R = 3000
    o = rng.uniform(-7, 7, (R, 3))
    target = rng.uniform(-3, 3, (R, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = rays_of(o, d, 0.0, 1e30)
    got = trace(t, rays)
    soup, owner = world_soup(verts, xforms, idx)
    ref = oracle.trace_brute(soup, rays).reshape(-1, 8)
    gp = got[:, 2].view(np.uint32)
    rp = ref[:, 2].view(np.uint32)
    hit = rp != EMPTY
    assert hit.sum() > R // 3
    same = np.zeros(R, bool)
    same[~hit] = gp[~hit] == EMPTY
    ok = hit & (gp != EMPTY)
    same[ok] = (owner[rp[ok], 0] == got[ok, 3].view(np.uint32)) & (owner[rp[ok], 1] == gp[ok])
    assert same.mean() > 0.99, same.mean()
    np.testing.assert_allclose(got[same & hit, 7], ref[same & hit, 7], rtol=1e-4, atol=1e-5)
    # world normals agree with the brute-force geometric normal up to orientation
    dots = np.abs(np.sum(got[same & hit, 4:7] * ref[same & hit, 4:7], 1))
    assert np.all(dots > 1 - 1e-4)
    # any hit: occlusion agrees with the closest hit wherever the closest hit is unambiguous
    occ = trace(t, rays, mode=1)
    assert np.all((occ[:, 7] == 0.0) == (gp != EMPTY))
    assert np.all(occ[occ[:, 7] == 0.0, 3].view(np.uint32) < 48)

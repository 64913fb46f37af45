"""Pins the oracle's AO rays (bajo's _make_ao_ray, rt/gpu/path_shading.mojo:
650-665) by what the definition fixes: origin = o + t d of the hit, t range
[0.001, 4.0], unit directions on the hemisphere of the normal oriented against
the incoming ray (rt/geometry.mojo:20-30), uniform on that hemisphere (mean
cosine 1/2, mean of each tangent component 0), one stream per (ray, sample),
and t_max = -1 for rays that hit nothing.  CPU only."""

import numpy as np

import oracle


def synthetic(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    t = rng.uniform(0.1, 5.0, n).astype(np.float32)
    rays = np.empty((8, n), np.float32)
    rays[0:3], rays[3], rays[4:7], rays[7] = o.T, 0.0, d.T, np.finfo(np.float32).max
    hits = np.zeros((n, 8), np.float32)
    hits[:, 4:7], hits[:, 7] = nrm, t
    prim = np.arange(n, dtype=np.uint32)
    prim[::7] = 0xFFFFFFFF  # misses
    hits[:, 2] = prim.view(np.float32)
    hits[:, 3] = np.full(n, 0xFFFFFFFF, np.uint32).view(np.float32)
    return rays.reshape(-1), hits.reshape(-1), o, d, nrm, t, prim


def test_ao_rays_definition():
    n, S = 20_000, 4
    rays, hits, o, d, nrm, t, prim = synthetic(n, 3)
    ao = oracle.ao_rays(rays, hits, S, 2029).reshape(8, n * S)
    r = np.repeat(np.arange(n), S)
    hit = prim[r] != 0xFFFFFFFF
    # misses: the ray can hit nothing
    assert (ao[7, ~hit] == -1.0).all() and (ao[3, ~hit] == np.float32(0.001)).all()
    # origin = o + t d, evaluated in FP32 without contraction
    po = o[r] + d[r] * t[r][:, None]
    np.testing.assert_array_equal(ao[0:3, hit].T, po[hit])
    assert (ao[3, hit] == np.float32(0.001)).all() and (ao[7, hit] == 4.0).all()
    dirs = ao[4:7, hit].T.astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=2e-6)
    nn = np.where((np.sum(d * nrm, axis=1) < 0)[:, None], nrm, -nrm)[r][hit].astype(np.float64)
    cos = np.sum(dirs * nn, axis=1)
    assert (cos >= -1e-7).all()
    # uniform on the hemisphere: E[cos] = 1/2, E[cos^2] = 1/3
    m = cos.size
    assert abs(cos.mean() - 0.5) < 4 * np.sqrt(1 / 12 / m)
    assert abs((cos ** 2).mean() - 1 / 3) < 4 * np.sqrt(4 / 45 / m)
    # samples of one ray are different streams
    a = ao[4:7].T.reshape(n, S, 3)[prim != 0xFFFFFFFF]
    assert (np.abs(a[:, 0] - a[:, 1]).sum(axis=1) > 0).mean() > 0.999


def test_ao_rays_seed_and_ray_count_independence():
    n, S = 500, 2
    rays, hits, *_ = synthetic(n, 4)
    a = oracle.ao_rays(rays, hits, S, 7).reshape(8, n * S)
    b = oracle.ao_rays(rays, hits, S, 8).reshape(8, n * S)
    assert not np.array_equal(a[4:7], b[4:7])
    # ray j's stream depends on (seed, j) only: a prefix launch gives the prefix rays
    half = n // 2
    rh = rays.reshape(8, n)[:, :half].reshape(-1)
    c = oracle.ao_rays(rh, hits[: 8 * half], S, 7).reshape(8, half * S)
    np.testing.assert_array_equal(c, a[:, : half * S])

"""The GGX conductor (BAJO_MAT_GGX; bajo has none, so its own mathematics pins it),
checked on the oracle's FP32 restatement that the device follows bit for bit:
- a known answer: at normal incidence and exit, f = F0 / (4 pi alpha^2);
- reciprocity f(wo, wi) = f(wi, wo);
- the sampler's weight equals f cos / pdf at the sampled direction;
- the visible-normal pdf integrates (quadrature over the hemisphere) to the
  fraction of samples that stay above the surface (pins D and G1 jointly);
- sampled directions follow that pdf (moments of the samples against quadrature);
- white furnace: with F0 = 1 the albedo (quadrature of f cos, and the mean sample
  weight) stays <= 1 and approaches 1 as the surface gets smooth;
- the FP32 renderer with GGX surfaces agrees with the independent FP64 estimator
  (PATH, NEE, MIS)."""

import numpy as np
import pytest

import oracle

N = np.array([0.0, 0.0, 1.0], np.float32)


def ray_toward(theta):
    # a ray arriving at the surface from direction wo = (sin, 0, cos)
    return np.array([-np.sin(theta), 0.0, -np.cos(theta)], np.float32)


def hemisphere_grid(n_t=180, n_p=360):
    t = (np.arange(n_t) + 0.5) * (np.pi / 2) / n_t
    p = (np.arange(n_p) + 0.5) * 2 * np.pi / n_p
    T, P = np.meshgrid(t, p, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / n_t) * (2 * np.pi / n_p)).reshape(-1)
    return dirs.astype(np.float32), dw


def quad(theta, alpha, f0=(1.0, 1.0, 1.0), n_t=120, n_p=240):
    dirs, dw = hemisphere_grid(n_t, n_p)
    rd = ray_toward(theta)
    val = np.empty(len(dirs))
    pdf = np.empty(len(dirs))
    for i, d in enumerate(dirs):
        v, p = oracle.eval_ggx(rd, N, f0, alpha, d)
        val[i] = v[0] * d[2]
        pdf[i] = p
    return dirs, dw, val, pdf


def samples(theta, alpha, n, seed, f0=(1.0, 1.0, 1.0)):
    rng = np.random.default_rng(seed)
    u = rng.random((n, 2)).astype(np.float32)
    rd = ray_toward(theta)
    out = [oracle.sample_ggx(rd, N, f0, alpha, a, b) for a, b in u]
    ok = np.array([o[3] for o in out])
    dirs = np.array([o[0] for o in out], np.float64)
    w = np.array([o[1][0] for o in out], np.float64)
    pdf = np.array([o[2] for o in out], np.float64)
    return ok, dirs, w, pdf


@pytest.mark.parametrize("alpha", [0.05, 0.2, 0.5])
def test_normal_incidence_known_answer(alpha):
    f0 = np.array([0.9, 0.6, 0.3], np.float32)
    v, p = oracle.eval_ggx(ray_toward(0.0), N, f0, alpha, N)
    np.testing.assert_allclose(v, f0 / (4 * np.pi * alpha * alpha), rtol=2e-6)
    np.testing.assert_allclose(p, 1 / (4 * np.pi * alpha * alpha), rtol=2e-6)


def test_reciprocity():
    rng = np.random.default_rng(3)
    f0 = np.array([0.8, 0.5, 0.2], np.float32)
    for _ in range(300):
        a = rng.normal(size=3); a[2] = abs(a[2]) + 0.05; a /= np.linalg.norm(a)
        b = rng.normal(size=3); b[2] = abs(b[2]) + 0.05; b /= np.linalg.norm(b)
        alpha = rng.uniform(0.05, 0.6)
        fab, _ = oracle.eval_ggx(-a.astype(np.float32), N, f0, alpha, b.astype(np.float32))
        fba, _ = oracle.eval_ggx(-b.astype(np.float32), N, f0, alpha, a.astype(np.float32))
        np.testing.assert_allclose(fab, fba, rtol=3e-5)


def test_sample_weight_is_f_cos_over_pdf():
    f0 = np.array([0.9, 0.7, 0.4], np.float32)
    for theta, alpha in [(0.3, 0.1), (1.0, 0.3), (1.4, 0.5)]:
        rd = ray_toward(theta)
        rng = np.random.default_rng(int(theta * 100))
        for u1, u2 in rng.random((200, 2)):
            d, w, p, ok = oracle.sample_ggx(rd, N, f0, alpha, u1, u2)
            if not ok:
                continue
            v, pe = oracle.eval_ggx(rd, N, f0, alpha, d)
            np.testing.assert_allclose(pe, p, rtol=2e-4)
            np.testing.assert_allclose(w, v * d[2] / pe, rtol=3e-4)


@pytest.mark.parametrize("theta,alpha", [(0.0, 0.3), (0.7, 0.2), (1.2, 0.4), (1.45, 0.5)])
def test_pdf_integral_equals_sample_survival_and_moments(theta, alpha):
    dirs, dw, _, pdf = quad(theta, alpha)
    integral = float(np.sum(pdf * dw))
    assert integral <= 1.0 + 2e-3
    ok, sd, _, _ = samples(theta, alpha, 6000, seed=11)
    frac = ok.mean()
    sigma = np.sqrt(frac * (1 - frac) / ok.size) + 1e-3
    assert abs(integral - frac) < 4 * sigma + 3e-3, (integral, frac)
    # moments of the accepted samples against the pdf
    for g in (lambda x: x[:, 2], lambda x: x[:, 0], lambda x: x[:, 0] ** 2):
        want = float(np.sum(g(dirs.astype(np.float64)) * pdf * dw) / integral)
        got = g(sd[ok])
        assert abs(got.mean() - want) < 4 * got.std() / np.sqrt(got.size) + 3e-3


@pytest.mark.parametrize("theta", [0.1, 0.8, 1.3])
def test_white_furnace(theta):
    """F0 = 1: single-scattering GGX keeps at most all the energy, a near-mirror keeps
    almost all of it, and rougher surfaces lose more (the albedo falls with alpha)."""
    albedos = []
    for alpha in (0.05, 0.25, 0.6, 1.0):
        _, dw, val, _ = quad(theta, alpha)
        albedo = float(np.sum(val * dw))
        ok, _, w, _ = samples(theta, alpha, 4000, seed=5)
        mc = float(np.where(ok, w, 0.0).mean())
        assert albedo <= 1.0 + 5e-3 and mc <= 1.0 + 1e-6
        assert abs(albedo - mc) < 0.02 + 4 * np.where(ok, w, 0.0).std() / np.sqrt(ok.size)
        if alpha <= 0.05 and theta < 1.0:
            assert mc > 0.97
        albedos.append(albedo)
    assert all(a >= b - 5e-3 for a, b in zip(albedos, albedos[1:])), albedos
    assert albedos[-1] > 0.3


def test_fp32_renderer_with_ggx_agrees_with_fp64_estimator():
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera

    sc = scenes.ggx_box_scene()
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    args = (sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 12, 8, 128, 6)
    b = np.array([oracle.render_fp64(*args, 90 + s, 3, cam, 8, ggx=sc.ggx).reshape(-1, 3).mean(0) for s in range(6)])
    for alg in (0, 3, 4):
        a = np.array([oracle.render(*args, s, alg, cam, 8, ggx=sc.ggx)[0].reshape(-1, 3).mean(0) for s in range(6)])
        z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
        assert np.all(np.abs(z) < 4.5), (alg, z, a.mean(0), b.mean(0))

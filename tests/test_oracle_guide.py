"""The path-guiding oracle pinned on the CPU: BF16 rounding known answers, the
encoding against numpy's sin / cos, the analytic gradient against central
finite differences, Adam's closed form under a constant gradient, vMF
densities integrating to one, the vMF sampler against its closed-form CDF
(Kolmogorov-Smirnov), the mixture sampler against its own density, and the
one-sample 50/50 MIS estimator against BSDF sampling."""

import math

import numpy as np
import pytest

import oracle

RNG = np.random.default_rng(2030)


def unit_vectors(n, rng):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def random_params(rng, scale=1.0):
    p = np.zeros(oracle.GUIDE_PARAMS)
    off = oracle.GUIDE_OFFSETS
    p[off["W1"]:off["b1"]] = rng.uniform(-1, 1, off["b1"] - off["W1"]) * math.sqrt(6 / 216) * scale
    p[off["b1"]:off["W2"]] = rng.uniform(-0.1, 0.1, 64)
    for w, b in (("W2", "b2"), ("W3", "b3")):
        nxt = {"b2": "W3", "b3": "W4"}[b]
        p[off[w]:off[b]] = rng.uniform(-1, 1, off[b] - off[w]) * math.sqrt(6 / 64) * scale
        p[off[b]:off[nxt]] = rng.uniform(-0.1, 0.1, 64)
    p[off["W4"]:off["b4"]] = rng.uniform(-1, 1, off["b4"] - off["W4"]) * math.sqrt(6 / 64)
    p[off["b4"]:] = rng.uniform(-0.5, 0.5, 40)
    return p


def random_samples(n, rng):
    s = np.zeros((n, 13), np.float32)
    s[:, 0:3] = rng.uniform(0, 1, (n, 3))
    s[:, 3:6] = unit_vectors(n, rng)
    s[:, 6:9] = unit_vectors(n, rng)
    s[:, 9:12] = unit_vectors(n, rng)
    s[:, 12] = rng.uniform(0.1, 2.0, n)
    return s


def test_bf16_rounding_known_answers():
    one = 1.0
    cases = {
        one: one,
        one + 2.0**-8: one,                       # tie -> even (mantissa 0)
        one + 3 * 2.0**-8: one + 2 * 2.0**-7,     # tie -> even (rounds up)
        one + 2.0**-8 + 2.0**-20: one + 2.0**-7,  # above the tie
        -2.5: -2.5,
        -(one + 3 * 2.0**-9): -(one + 2.0**-7),   # rounds away from zero in magnitude
    }
    for x, want in cases.items():
        assert float(oracle.bf16_round(np.float32(x))[()]) == want, x
    assert float(oracle.lib().orc_bf16_round(np.float32(1.0 + 2.0**-8))) == 1.0


def test_encoding_matches_numpy():
    x = np.concatenate([RNG.uniform(0, 1, 3), RNG.uniform(-1, 1, 6)]).astype(np.float32)
    feat = oracle.guide_encode(x)
    want = np.empty(216)
    for d in range(9):
        for f in range(12):
            a = math.pi * 2.0**f * float(x[d])
            want[24 * d + 2 * f] = math.sin(a)
            want[24 * d + 2 * f + 1] = math.cos(a)
    np.testing.assert_allclose(feat, want, atol=2e-5)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(5)
    p = random_params(rng)
    s = random_samples(24, rng)
    loss, g = oracle.guide_loss_grad(p, s, emulate_bf16=False)
    assert np.isfinite(loss) and np.all(np.isfinite(g))
    off = oracle.GUIDE_OFFSETS
    picks = []
    for a, b in (("W1", "b1"), ("b1", "W2"), ("W2", "b2"), ("b2", "W3"), ("W3", "b3"), ("b3", "W4"), ("W4", "b4")):
        picks += list(rng.integers(off[a], off[b], 12))
    picks += list(range(off["b4"], off["b4"] + 40))  # every output bias: every output's gradient
    h = 1e-6
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (oracle.guide_loss_grad(pp, s, with_grad=False)[0] - oracle.guide_loss_grad(pm, s, with_grad=False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 2e-5 * abs(fd), (k, fd, g[k])


def test_gradient_of_clamped_concentration_is_zero():
    rng = np.random.default_rng(6)
    p = random_params(rng)
    off = oracle.GUIDE_OFFSETS
    p[off["b4"] + 3] = 20.0  # lobe 0's log concentration far above the clamp
    s = random_samples(8, rng)
    _, g = oracle.guide_loss_grad(p, s)
    assert g[off["b4"] + 3] == 0.0
    assert np.all(g[off["W4"] + 3 * 64: off["W4"] + 4 * 64] == 0.0)


def test_adam_constant_gradient_closed_form():
    rng = np.random.default_rng(7)
    p0 = rng.normal(size=oracle.GUIDE_PARAMS)
    g = rng.normal(size=oracle.GUIDE_PARAMS)
    g[:10] = 0.0
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t in (1, 2, 3):
        oracle.guide_adam(p, m, v, g, t, lr=1e-2)
    # with a constant gradient m_hat = g and v_hat = g^2 at every step
    want = p0 - 3 * 1e-2 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p, want, rtol=0, atol=1e-12)
    assert np.all(p[:10] == p0[:10])


@pytest.mark.parametrize("kappa", [0.02, 0.5, 5.0, 100.0, 1000.0])
def test_vmf_density_integrates_to_one(kappa):
    from scipy.integrate import quad

    total, _ = quad(lambda t: 2 * math.pi * oracle.vmf_pdf64(kappa, t), -1, 1, points=[1 - 10 / kappa] if kappa > 10 else None,
                    limit=200)
    assert abs(total - 1.0) < 1e-9


@pytest.mark.parametrize("kappa", [0.05, 1.0, 20.0, 400.0])
def test_vmf_sampler_ks(kappa):
    from scipy.stats import kstest

    u = np.random.default_rng(int(kappa * 10)).uniform(0, 1, 20000)
    t = np.array([oracle.vmf_sample_cos64(kappa, x) for x in u])
    cdf = lambda x: (np.exp(kappa * (np.asarray(x) - 1)) - math.exp(-2 * kappa)) / (-math.expm1(-2 * kappa))
    assert kstest(t, cdf).pvalue > 1e-3


def mixture_outputs(rng, kappa_log_range=(-1.0, 3.0)):
    o = np.zeros(40, np.float32)
    for j in range(8):
        o[5 * j: 5 * j + 3] = rng.normal(size=3)
        o[5 * j + 3] = rng.uniform(*kappa_log_range)
        o[5 * j + 4] = rng.normal()
    return o


def sphere_grid(n_theta=600, n_phi=1200):
    ct = np.linspace(-1, 1, n_theta + 1)
    ct = 0.5 * (ct[1:] + ct[:-1])
    ph = (np.arange(n_phi) + 0.5) * 2 * math.pi / n_phi
    st = np.sqrt(1 - ct**2)
    d = np.stack([np.outer(st, np.cos(ph)), np.outer(st, np.sin(ph)), np.outer(ct, np.ones_like(ph))], -1)
    return d.reshape(-1, 3), (2.0 / n_theta) * (2 * math.pi / n_phi)


def test_mixture_density_integrates_to_one():
    rng = np.random.default_rng(8)
    dirs, dw = sphere_grid(300, 600)
    for _ in range(2):
        o = mixture_outputs(rng)
        vals = np.array([oracle.guide_mixture_pdf(o, d) for d in dirs])
        assert abs(vals.sum() * dw - 1.0) < 2e-3


def test_mixture_sampler_consistent_with_density():
    rng = np.random.default_rng(9)
    o = mixture_outputs(rng, (0.0, 2.0))
    n = 40000
    u = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    inv = np.empty(n)
    for i in range(n):
        d, pdf = oracle.guide_mixture_sample(o, u[i])
        assert abs(np.linalg.norm(d) - 1) < 1e-6
        assert pdf == oracle.guide_mixture_pdf(o, d)
        inv[i] = 1.0 / pdf
    # E[1 / q(w)] over w ~ q is the sphere's area
    mean, err = inv.mean(), inv.std() / math.sqrt(n)
    assert abs(mean - 4 * math.pi) < 4 * err


def test_one_sample_mis_weights_and_unbiasedness():
    """Half BSDF, half guide: the balance weights of the two strategies sum to one, and the combined
    estimator of the GGX directional albedo agrees with BSDF sampling alone."""
    rng = np.random.default_rng(10)
    n_vec = np.array([0.0, 0.0, 1.0], np.float32)
    ray_d = np.array([0.6, 0.0, -0.8], np.float32)
    f0, alpha = np.array([0.9, 0.6, 0.3], np.float32), 0.3
    o = mixture_outputs(rng, (0.0, 2.0))
    m = 60000
    mis, plain = np.zeros((m, 3)), np.zeros((m, 3))
    for i in range(m):
        _, weight, _, ok = oracle.sample_ggx(ray_d, n_vec, f0, alpha, rng.uniform(), rng.uniform())
        if ok:
            plain[i] = weight
        u = rng.uniform(0, 1, 5).astype(np.float32)
        if u[0] < 0.5:
            d, _, _, ok = oracle.sample_ggx(ray_d, n_vec, f0, alpha, u[1], u[2])
            if not ok:
                continue
            d = np.asarray(d, np.float32)
        else:
            d, _ = oracle.guide_mixture_sample(o, u[2:5])
        value, pb = oracle.eval_ggx(ray_d, n_vec, f0, alpha, d)
        pg = oracle.guide_mixture_pdf(o, d)
        wb, wg = 0.5 * pb / (0.5 * pb + 0.5 * pg), 0.5 * pg / (0.5 * pb + 0.5 * pg)
        assert abs(wb + wg - 1.0) < 1e-6
        cos = max(float(np.dot(d, n_vec)), 0.0)
        if cos > 0 and pb > 0:
            mis[i] = np.asarray(value) * cos / (0.5 * pb + 0.5 * pg)
    for c in range(3):
        a, b = mis[:, c], plain[:, c]
        err = math.sqrt(a.var() / m + b.var() / m)
        assert abs(a.mean() - b.mean()) < 4 * err, (c, a.mean(), b.mean(), err)

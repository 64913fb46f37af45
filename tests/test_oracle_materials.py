"""Pins the oracle's metal and dielectric (bajo's rt/shading.mojo:30-260, as
shaded on the GPU by rt/gpu/path_shading.mojo:669-783) before it checks the
device renderer.  CPU only.

- bajo's known answers (test/rt/test_cpu.mojo:102-109, 243-291): reflect,
  Schlick reflectance at normal incidence, a mirror metal returns its albedo
  as a delta sample with pdf 1 and is absorbed from the back, a dielectric
  sample has weight 1, is delta and has pdf in (0, 1];
- closed forms: pow against numpy; the Phong lobe's sampled cosine has CDF
  c^(e+1) and a uniform azimuth, its pdf is (e+1)/(2 pi) cos^e and the sample
  weight albedo (e+2)/(e+1) cos_n; refraction obeys Snell's law and Schlick's
  formula; total internal reflection reflects with pdf 1;
- the FP32 renderer on a scene with every material kind agrees with the
  independent FP64 estimator (PATH), for PATH, NEE and MIS.
"""

import numpy as np

import oracle


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_bajo_known_answers():
    np.testing.assert_array_equal(oracle.reflect([1.0, -1.0, 0.0], [0.0, 1.0, 0.0]), [1.0, 1.0, 0.0])
    assert abs(oracle.reflectance(1.0, 1.5) - 0.04) < 1e-5
    albedo = [0.8, 0.7, 0.6]
    s = oracle.sample_metal([0.0, 0.0, -1.0], [0.0, 0.0, 1.0], albedo, 0.0, 0.3, 0.6)
    assert s["ok"] and s["delta"] and s["pdf"] == 1.0
    np.testing.assert_array_equal(s["weight"], np.float32(albedo))
    assert s["dir"][2] > 0
    back = oracle.sample_metal([0.0, 0.0, -1.0], [0.0, 0.0, -1.0], albedo, 0.0, 0.3, 0.6)
    assert not back["ok"]
    for ru in (0.01, 0.5, 0.99):
        d = oracle.sample_dielectric([0.0, 0.0, -1.0], [0.0, 0.0, 1.0], 1.5, True, ru)
        assert d["ok"] and d["delta"] and 0.0 < d["pdf"] <= 1.0
        np.testing.assert_array_equal(d["weight"], [1.0, 1.0, 1.0])
        assert np.linalg.norm(d["dir"]) > 0


def test_pow_unit_against_numpy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(4000), 10.0 ** rng.uniform(-30, 0, 2000), [1.0, 0.5, 1e-38]]).astype(np.float32)
    y = np.concatenate([rng.uniform(0, 3, 3000), rng.uniform(0, 2000, 3003)]).astype(np.float32)
    got = np.array([oracle.pow_unit(float(a), float(b)) for a, b in zip(x, y)])
    want = np.power(x.astype(np.float64), y.astype(np.float64))
    big = want >= 2.0 ** -120
    # FP32 evaluation of t = y log2 x carries an absolute error ~ ulp(t); e^(ulp(t) ln 2) bounds the rest
    t = np.abs(y.astype(np.float64) * np.log2(x.astype(np.float64)))
    tol = (4e-7 + 1.2e-7 * t) * want
    assert np.all(np.abs(got[big] - want[big]) <= tol[big]), np.max((np.abs(got - want) / tol)[big])
    assert np.all(got[~big] <= 2.0 ** -119)
    assert oracle.pow_unit(0.0, 0.0) == 1.0 and oracle.pow_unit(0.0, 2.0) == 0.0 and oracle.pow_unit(0.3, 0.0) == 1.0


def test_metal_lobe_distribution_and_weight():
    rng = np.random.default_rng(2)
    n = unit([0.1, 0.2, 1.0]).astype(np.float32)
    ray_d = unit([0.3, -0.1, -1.0]).astype(np.float32)
    albedo = np.float32([0.8, 0.65, 0.3])
    for fuzz in (0.25, 0.6):
        e = max(2.0 / np.float32(fuzz) ** 2 - 2.0, 0.0)
        refl = unit(ray_d - 2 * np.dot(ray_d, n) * n)
        t = unit(np.cross([0.0, 1.0, 0.0], refl))
        b = np.cross(refl, t)
        cosr, phi = [], []
        for _ in range(3000):
            s = oracle.sample_metal(ray_d, n, albedo, fuzz, rng.random(), rng.random())
            d = s["dir"].astype(np.float64)
            c = float(np.dot(refl, d))
            cosr.append(c)
            phi.append(np.arctan2(np.dot(d, b), np.dot(d, t)))
            if s["ok"]:
                cn = float(np.dot(n, d))
                np.testing.assert_allclose(s["pdf"], (e + 1) / (2 * np.pi) * c ** e, rtol=2e-4)
                np.testing.assert_allclose(s["weight"], albedo * (e + 2) / (e + 1) * cn, rtol=2e-4, atol=1e-7)
                ev = oracle.eval_metal(ray_d, n, albedo, fuzz, d)
                np.testing.assert_allclose(ev["pdf"], s["pdf"], rtol=1e-6)
            else:
                assert np.dot(n, d) <= 0 or c <= 0
        # Kolmogorov-Smirnov against the CDF c^(e+1) and a uniform azimuth
        c = np.sort(np.array(cosr))
        ks = np.max(np.abs(np.arange(1, c.size + 1) / c.size - c ** (e + 1)))
        assert ks < 1.63 / np.sqrt(c.size), ks
        p = np.sort((np.array(phi) + np.pi) / (2 * np.pi))
        ks = np.max(np.abs(np.arange(1, p.size + 1) / p.size - p))
        assert ks < 1.63 / np.sqrt(p.size), ks


def test_dielectric_snell_schlick_and_total_internal_reflection():
    n = np.float32([0.0, 0.0, 1.0])
    for theta in (0.1, 0.5, 0.9, 1.3):
        d = np.float32([np.sin(theta), 0.0, -np.cos(theta)])
        for front in (True, False):
            ri = 1 / 1.5 if front else 1.5
            cos_i = np.cos(theta)
            r0 = ((1 - ri) / (1 + ri)) ** 2
            R = r0 + (1 - r0) * (1 - cos_i) ** 5
            tir = ri * np.sin(theta) > 1
            refr = oracle.sample_dielectric(d, n, 1.5, front, 0.999999 if not tir else 0.5)  # refraction
            refl = oracle.sample_dielectric(d, n, 1.5, front, 0.0)  # reflection
            if tir:
                np.testing.assert_allclose(refl["dir"], [np.sin(theta), 0, np.cos(theta)], atol=2e-6)
                assert refr["pdf"] == 1.0 and refl["pdf"] == 1.0
                np.testing.assert_allclose(refr["dir"], [np.sin(theta), 0, np.cos(theta)], atol=2e-6)
                continue
            # ru = 0: reflection (R > 0), pdf R; ru ~ 1: refraction, pdf 1 - R
            np.testing.assert_allclose(refl["pdf"], R, rtol=1e-5)
            np.testing.assert_allclose(refl["dir"], [np.sin(theta), 0, np.cos(theta)], atol=2e-6)
            np.testing.assert_allclose(refr["pdf"], 1 - R, rtol=1e-5)
            out = refr["dir"].astype(np.float64)
            np.testing.assert_allclose(np.linalg.norm(out[:2]), ri * np.sin(theta), atol=3e-6)
            assert out[2] < 0


def test_fp32_renderer_with_materials_agrees_with_fp64_path():
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera

    sc = scenes.material_box_scene()
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    args = (sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 12, 8, 128, 6)
    b = np.array([oracle.render_fp64(*args, 90 + s, 0, cam, 8, metal=sc.metal, dielectric=sc.dielectric)
                  .reshape(-1, 3).mean(0) for s in range(6)])
    for alg in (0, 3, 4):
        a = np.array([oracle.render(*args, s, alg, cam, 8, metal=sc.metal, dielectric=sc.dielectric)[0]
                      .reshape(-1, 3).mean(0) for s in range(6)])
        z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
        assert np.all(np.abs(z) < 4.5), (alg, z, a.mean(0), b.mean(0))

"""Oracle renderer over mixed triangle + sphere scenes (rt/gpu/mixed_path.mojo,
rt/gpu/scene_trace.mojo:113-193, sphere lights of rt/types.mojo:490-530 and
rt/lighting.mojo) pinned by closed forms: a Lambertian floor under a sphere
light has radiance albedo * L * (R / D)^2 * cos(theta) (a uniform spherical
emitter fully above the horizon), a directly seen emissive sphere shows its
radiance exactly, the light table's sphere weights are 4 pi r^2 times the mean
radiance; and the FP32 estimator agrees statistically with the independent
FP64 one on a scene with every material kind and a negative-radius sphere."""

import numpy as np
import pytest

import oracle

NEE, PATH, MIS = 3, 0, 4


def floor_light_scene(albedo=0.5, radiance=4.0, center=(0.0, 1.0, 0.0), radius=0.25):
    from bajo_amd.rt import MAT, SceneData, surface_id
    from bajo_amd.scenes import quad

    floor = quad((-50, 0, -50), (-50, 0, 50), (50, 0, 50), (50, 0, -50))
    return SceneData(
        floor, np.full(2, surface_id(MAT.LAMBERTIAN, 0), np.uint32),
        np.array([albedo] * 3, np.float32), np.array([radiance] * 3, np.float32),
        spheres=np.array([*center, radius], np.float32),
        sphere_surfaces=np.array([surface_id(MAT.EMISSIVE, 0)], np.uint32),
    )


def centre_rays(cam, w, h):
    """Pixel-centre rays of bajo's pinhole camera (bvh/camera.mojo:137-154) in float64."""
    c = np.asarray(cam, np.float64)
    px, py = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    sx = (2.0 * px - w) / h
    sy = 1.0 - 2.0 * py / h
    d = c[3:6] + (sx * c[12])[..., None] * c[6:9] + (sy * c[12])[..., None] * c[9:12]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return c[0:3], d.reshape(-1, 3)


def sphere_t(o, d, c, r):
    oc = o - c
    b = d @ oc
    disc = b * b - (oc @ oc - r * r)
    return np.where(disc > 0, -b - np.sqrt(np.maximum(disc, 0)), np.inf), disc


def test_sphere_light_table_weights():
    from bajo_amd import scenes

    sc = scenes.mixed_box_scene()
    fields, kinds, total = oracle.light_table(sc.vertices, sc.surfaces, sc.emissive, sc.spheres, sc.sphere_surfaces)
    # two quad triangles (0.36 * 12 in total) then the emissive ball of radius 0.12, radiance 6 * (1, 0.8, 0.5)
    assert len(kinds) == 3 and list(kinds & 0xF) == [1, 1, 0]
    f = fields.reshape(-1, 14)
    np.testing.assert_array_equal(f[2, 0:3], np.float32([-0.55, 1.2, 0.6]))
    assert f[2, 9] == np.float32(0.12)
    sphere_w = 4.0 * np.pi * 0.12 ** 2 * 6.0 * (1.0 + 0.8 + 0.5) / 3.0
    assert total == pytest.approx(0.36 * 12.0 + sphere_w, rel=1e-5)
    # alias-table probabilities: the sphere's share of the total weight
    prob = f[:, 13].astype(np.float64)
    alias = kinds >> 4
    share = np.zeros(3)
    for col in range(3):
        share[col] += prob[col] / 3
        share[alias[col]] += (1 - prob[col]) / 3
    assert share[2] == pytest.approx(sphere_w / (0.36 * 12.0 + sphere_w), rel=1e-5)


def test_nee_from_a_sphere_light_matches_closed_form():
    from bajo_amd.bvh import Camera

    a, L, c, R = 0.5, 4.0, np.array([0.0, 1.0, 0.0]), 0.25
    sc = floor_light_scene(a, L, tuple(c), R)
    cam = Camera((1.6, 2.2, 0.3), (1.2, 0.0, 0.1), (0.0, 1.0, 0.0), 0.3).flatten()
    w, h = 16, 12
    px, _ = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, 512, 1, 77, NEE, cam, 8,
                          spheres=sc.spheres, sphere_surfaces=sc.sphere_surfaces)
    o, d = centre_rays(cam, w, h)
    t_floor = -o[1] / d[:, 1]
    assert np.all(t_floor > 0)
    ts, disc = sphere_t(o, d, c, R)
    assert np.all(disc < -0.05), "the sphere light must stay out of view"
    p = o + t_floor[:, None] * d
    to_c = c - p
    D = np.linalg.norm(to_c, axis=1)
    want = a * L * (R / D) ** 2 * (to_c[:, 1] / D)
    got = px.reshape(-1, 3).astype(np.float64)
    np.testing.assert_allclose(got.mean(0), [want.mean()] * 3, rtol=0.03)
    # per pixel: Monte Carlo error of 512 area samples of the sphere stays within 25 %
    np.testing.assert_allclose(got[:, 0], want, rtol=0.25)


def test_directly_seen_emissive_sphere_shows_its_radiance():
    from bajo_amd.bvh import Camera

    L, c, R = 4.0, np.array([0.0, 1.0, 0.0]), 0.25
    sc = floor_light_scene(0.5, L, tuple(c), R)
    cam = Camera((0.0, 1.0, -2.0), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 0.2).flatten()
    w, h = 24, 24
    for alg in (PATH, NEE, MIS):
        px, _ = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, 4, 1, 5, alg, cam, 8,
                              spheres=sc.spheres, sphere_surfaces=sc.sphere_surfaces)
        o, d = centre_rays(cam, w, h)
        ts, disc = sphere_t(o, d, c, R)
        # pixels well inside the silhouette (every jittered sample hits the sphere)
        inside = disc > 0.02
        assert inside.sum() > 40
        np.testing.assert_array_equal(px.reshape(-1, 3)[inside], np.float32(L))


def test_negative_radius_sphere_normals_point_inward():
    """NORMALS faces the normal toward the ray, so a sphere seen from outside looks the same with either radius
    sign; seen from inside through its own surface the visible normal is the same too - the sign shows in the
    dielectric's front/back choice, which the FP64 comparison below covers."""
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import MAT, SceneData, surface_id

    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.4).flatten()
    out = []
    for r in (0.8, -0.8):
        sc = SceneData(np.zeros(0, np.float32), np.zeros(0, np.uint32), np.array([0.5] * 3, np.float32),
                       np.zeros(0, np.float32), spheres=np.array([0, 0, 0, r], np.float32),
                       sphere_surfaces=np.array([surface_id(MAT.LAMBERTIAN, 0)], np.uint32))
        px, _ = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 16, 16, 1, 1, 3, 1, cam, 8,
                              spheres=sc.spheres, sphere_surfaces=sc.sphere_surfaces)
        out.append(px)
    np.testing.assert_array_equal(out[0], out[1])
    n = 2.0 * out[0].reshape(-1, 3).astype(np.float64) - 1.0
    hit = np.abs(np.linalg.norm(n, axis=1) - 1.0) < 1e-3
    assert hit.sum() > 60 and np.all(n[hit, 2] < 0)  # facing the camera at -z


@pytest.mark.parametrize("alg", [PATH, NEE, MIS])
def test_mixed_fp32_oracle_agrees_with_fp64_estimator(alg):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera

    sc = scenes.mixed_box_scene()
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    kw = dict(metal=sc.metal, dielectric=sc.dielectric, spheres=sc.spheres, sphere_surfaces=sc.sphere_surfaces)
    a = np.array([oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 12, 8, 128, 6, s, alg, cam, 8,
                                **kw)[0].reshape(-1, 3).mean(0) for s in range(6)])
    b = np.array([oracle.render_fp64(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 12, 8, 128, 6, 50 + s, 0,
                                     cam, 8, **kw).reshape(-1, 3).mean(0) for s in range(6)])
    z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
    assert np.all(np.abs(z) < 4.5), (a.mean(0), b.mean(0), z)


def test_nee_and_mis_agree_with_path_on_a_sphere_light():
    """NEE, MIS and plain path sampling of the floor under a sphere light estimate the same image (sky included
    at depth 2): pins the sphere light's solid-angle pdf in the MIS weights and emission on BSDF-sampled hits."""
    from bajo_amd.bvh import Camera

    sc = floor_light_scene()
    cam = Camera((1.6, 2.2, 0.3), (1.2, 0.0, 0.1), (0.0, 1.0, 0.0), 0.3).flatten()
    kw = dict(spheres=sc.spheres, sphere_surfaces=sc.sphere_surfaces)
    means = {}
    for alg in (PATH, NEE, MIS):
        means[alg] = np.array([oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 8, 6, 256, 2,
                                             100 * alg + s, alg, cam, 8, **kw)[0].reshape(-1, 3).mean(0)
                               for s in range(6)])
    for alg in (NEE, MIS):
        a, b = means[alg], means[PATH]
        z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
        assert np.all(np.abs(z) < 4.5), (alg, a.mean(0), b.mean(0), z)

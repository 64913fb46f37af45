"""GPU wavefront renderer vs the CPU oracle with identical per-path streams
(bajo's protocol, test/rt/test_gpu.mojo:346-730): image RMSE <= 1e-3 (in
practice bit-identical pixels), chunked == whole-frame bit-exact
(test/rt/test_gpu.mojo:446-479), ray counts equal, light table equal, and
radiance statistically equal to the independent FP64 estimator."""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu

ALGS = {"PATH": 0, "NORMALS": 1, "AO": 2, "NEE": 3, "MIS": 4}


def box_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()


@pytest.fixture(scope="module")
def box(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuScene

    sc = scenes.box_scene()
    return sc, GpuScene(sc, gpu_ctx)


def render_both(sc, gscene, alg, w, h, spp, depth, seed, cam, capacity=0):
    from bajo_amd.rt import RenderSettings, enqueue_render_gpu

    st = RenderSettings(w, h, spp, seed, depth, capacity)
    res = enqueue_render_gpu(st, cam, gscene, alg, with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed, alg, cam, 8)
    return got, want, res.stats, rays


@pytest.mark.parametrize("name", ["PATH", "NEE", "MIS", "NORMALS", "AO"])
def test_box_image_matches_oracle(box, name):
    sc, g = box
    announce(f"render box {name}")
    got, want, stats, rays = render_both(sc, g, ALGS[name], 40, 30, 8, 5, 2026, box_camera())
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    print(f"{name}: rmse {rmse:.3e}, bit-identical pixels {exact:.4f}")
    assert rmse <= 1e-3
    assert exact > 0.99
    assert stats["extension_rays"] == rays[0]
    assert stats["shadow_rays"] == rays[1]


def test_chunked_equals_full_frame(box):
    sc, g = box
    from bajo_amd.rt import RenderSettings, enqueue_render_gpu

    announce("render chunked vs full")
    cam = box_camera()
    full = enqueue_render_gpu(RenderSettings(24, 16, 4, 99, 5, 0), cam, g, ALGS["NEE"]).numpy()
    for cap in (4, 4 * 7, 4 * 64, 1000):
        part = enqueue_render_gpu(RenderSettings(24, 16, 4, 99, 5, cap), cam, g, ALGS["NEE"]).numpy()
        np.testing.assert_array_equal(part.view(np.uint32), full.view(np.uint32))


def test_light_table_equals_oracle(box):
    sc, g = box
    fields, kinds, total = g.light_table()
    of, ok, ot = oracle.light_table(sc.vertices, sc.surfaces, sc.emissive)
    np.testing.assert_array_equal(fields, of)
    np.testing.assert_array_equal(kinds, ok)
    assert total == ot


@pytest.mark.parametrize("name", ["PATH", "NEE"])
def test_radiance_matches_fp64_estimator(box, name):
    """Image means over independent seeds agree with the FP64 estimator (|z| < 4.5)."""
    sc, g = box
    from bajo_amd.rt import RenderSettings, enqueue_render_gpu

    announce(f"render radiance vs fp64 {name}")
    cam = box_camera()
    alg = ALGS[name]
    w, h = 16, 12
    a = np.array([enqueue_render_gpu(RenderSettings(w, h, 512, 1000 + s, 5), cam, g, alg).numpy().reshape(-1, 3).mean(0) for s in range(8)])
    b = np.array([oracle.render_fp64(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, 256, 5, 77 + s, alg, cam, 8).reshape(-1, 3).mean(0) for s in range(8)])
    z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 8 + b.var(0, ddof=1) / 8)
    print(name, a.mean(0), b.mean(0), z)
    assert np.all(np.abs(z) < 4.5), z


def test_icosphere_scene_small_image_matches_oracle(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene

    n = 1 << 16
    announce(f"render icosphere n={n} NEE")
    sc = scenes.icosphere_scene(n, frequency=58)
    g = GpuScene(sc, gpu_ctx)
    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.41421356).flatten()
    got, want, stats, rays = render_both(sc, g, ALGS["NEE"], 48, 27, 4, 5, 2026, cam)
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    print(f"icosphere NEE: rmse {rmse:.3e}, identical {exact:.4f}, rays {stats}")
    assert rmse <= 1e-3 and exact > 0.99
    assert stats["extension_rays"] == rays[0] and stats["shadow_rays"] == rays[1]

"""GPU renderer with the GGX conductor (BAJO_MAT_GGX) beside Lambertian surfaces
against the oracle renderer with identical per-path streams: image RMSE <= 1e-3 and
> 99 % bit-identical pixels, equal ray counts, at both node widths, through the tail
kernel (depth 8); radiance statistically equal to the independent FP64 estimator."""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu
ALGS = {"PATH": 0, "NEE": 3, "MIS": 4}


def camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()


@pytest.fixture(scope="module")
def scene():
    from bajo_amd import scenes

    return scenes.ggx_box_scene()


@pytest.mark.parametrize("width", [2, 8])
@pytest.mark.parametrize("name", ["PATH", "NEE", "MIS"])
def test_ggx_box_matches_oracle(gpu_ctx, scene, width, name):
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce(f"ggx {name} width {width}")
    g = GpuScene(scene, gpu_ctx, node_width=width)
    w, h, spp, depth, seed = 40, 30, 8, 8, 2026
    res = enqueue_render_gpu(RenderSettings(w, h, spp, seed, depth), camera(), g, ALGS[name], with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(scene.vertices, scene.surfaces, scene.lambertian, scene.emissive, w, h, spp, depth,
                               seed, ALGS[name], camera(), 8, node_width=width, leaf_size=4, ggx=scene.ggx)
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    print(f"{name} width {width}: rmse {rmse:.3e}, bit-identical pixels {exact:.4f}")
    assert rmse <= 1e-3 and exact > 0.99
    assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
    assert res.stats["status"] == 0 and res.stats["traversal_status"] == 0
    g.free()


def test_ggx_box_radiance_matches_fp64(gpu_ctx, scene):
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce("ggx radiance vs FP64")
    g = GpuScene(scene, gpu_ctx, node_width=8)
    w, h = 12, 8
    b = np.array([oracle.render_fp64(scene.vertices, scene.surfaces, scene.lambertian, scene.emissive, w, h, 256, 6,
                                     300 + s, 3, camera(), 8, ggx=scene.ggx).reshape(-1, 3).mean(0) for s in range(6)])
    for name in ("PATH", "NEE", "MIS"):
        a = np.array([enqueue_render_gpu(RenderSettings(w, h, 256, 700 + s, 6), camera(), g, ALGS[name])
                      .numpy().reshape(-1, 3).mean(0) for s in range(6)])
        z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
        print(name, a.mean(0), b.mean(0), z)
        assert np.all(np.abs(z) < 4.5), (name, z)
    g.free()

"""GPU renderer over mixed triangle + sphere scenes (rt/gpu/mixed_path.mojo:
spheres and triangles traced through one scene trace, rt/gpu/scene_trace.mojo:113-193;
sphere lights from SceneData, rt/types.mojo:490-530) against the oracle renderer
with identical per-path streams: image RMSE <= 1e-3 and > 99 % bit-identical
pixels, equal ray counts, at both node widths, for a box of triangles with
spheres of every material kind (one of negative radius, one emissive) and for a
spheres-only field; the light table equal to the oracle's; radiance
statistically equal to the independent FP64 estimator; invalid spheres rejected."""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu
ALGS = {"PATH": 0, "NEE": 3, "MIS": 4, "AO": 2, "NORMALS": 1}


def box_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()


def field_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 2.0, -7.0), (0.0, 0.3, 0.0), (0.0, 1.0, 0.0), 0.6).flatten()


@pytest.fixture(scope="module")
def box():
    from bajo_amd import scenes

    return scenes.mixed_box_scene()


@pytest.fixture(scope="module")
def field():
    from bajo_amd import scenes

    return scenes.sphere_field_scene()


def render_pair(ctx, sc, cam, name, width, w=40, h=30, spp=8, depth=8, seed=2026):
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    g = GpuScene(sc, ctx, node_width=width)
    res = enqueue_render_gpu(RenderSettings(w, h, spp, seed, depth), cam, g, ALGS[name], with_stats=True)
    got = res.numpy().reshape(-1)
    want, rays = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed,
                               ALGS[name], cam, 8, node_width=width, leaf_size=4, metal=sc.metal,
                               dielectric=sc.dielectric, spheres=sc.spheres, sphere_surfaces=sc.sphere_surfaces)
    g.free()
    return got, want, rays, res.stats


def check(got, want, rays, stats, label):
    rmse = float(np.sqrt(np.mean((got.astype(np.float64) - want) ** 2)))
    exact = float(np.mean(got.view(np.uint32) == want.view(np.uint32)))
    print(f"{label}: rmse {rmse:.3e}, bit-identical pixels {exact:.4f}")
    assert rmse <= 1e-3 and exact > 0.99
    assert stats["extension_rays"] == rays[0] and stats["shadow_rays"] == rays[1]
    assert stats["status"] == 0 and stats["traversal_status"] == 0


@pytest.mark.parametrize("width", [2, 8])
@pytest.mark.parametrize("name", ["PATH", "NEE", "MIS", "AO", "NORMALS"])
def test_mixed_box_matches_oracle(gpu_ctx, box, width, name):
    announce(f"mixed box {name} width {width}")
    check(*render_pair(gpu_ctx, box, box_camera(), name, width), f"mixed box {name} width {width}")


@pytest.mark.parametrize("width", [2, 8])
@pytest.mark.parametrize("name", ["PATH", "NEE", "MIS"])
def test_sphere_field_matches_oracle(gpu_ctx, field, width, name):
    announce(f"sphere field {name} width {width}")
    check(*render_pair(gpu_ctx, field, field_camera(), name, width, w=48, h=32, spp=4, depth=6),
          f"sphere field {name} width {width}")


def test_mixed_light_table_matches_oracle(gpu_ctx, box):
    from bajo_amd.rt import GpuScene

    announce("mixed light table")
    g = GpuScene(box, gpu_ctx)
    f, k, total = g.light_table()
    wf, wk, wtotal = oracle.light_table(box.vertices, box.surfaces, box.emissive, box.spheres, box.sphere_surfaces)
    np.testing.assert_array_equal(k, wk)
    np.testing.assert_array_equal(f, wf)
    assert np.float32(total) == np.float32(wtotal)
    g.free()


def test_mixed_box_radiance_matches_fp64(gpu_ctx, box):
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce("mixed radiance vs FP64")
    g = GpuScene(box, gpu_ctx, node_width=8)
    w, h = 12, 8
    kw = dict(metal=box.metal, dielectric=box.dielectric, spheres=box.spheres, sphere_surfaces=box.sphere_surfaces)
    b = np.array([oracle.render_fp64(box.vertices, box.surfaces, box.lambertian, box.emissive, w, h, 256, 6, 300 + s,
                                     0, box_camera(), 8, **kw).reshape(-1, 3).mean(0) for s in range(6)])
    for name in ("PATH", "NEE", "MIS"):
        a = np.array([enqueue_render_gpu(RenderSettings(w, h, 256, 700 + s, 6), box_camera(), g, ALGS[name])
                      .numpy().reshape(-1, 3).mean(0) for s in range(6)])
        z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
        print(name, a.mean(0), b.mean(0), z)
        assert np.all(np.abs(z) < 4.5), (name, z)
    g.free()


def test_mixed_scene_render_target_replay_equals_render(gpu_ctx, box):
    import torch

    from bajo_amd.rt import GpuRtRenderTarget, GpuScene, RenderSettings, enqueue_render_gpu

    announce("mixed render target")
    g = GpuScene(box, gpu_ctx, node_width=8)
    st = RenderSettings(32, 24, 4, 91, 8)
    want = enqueue_render_gpu(st, box_camera(), g, ALGS["MIS"]).numpy()
    t = GpuRtRenderTarget(st, box_camera(), ALGS["MIS"], gpu_ctx)
    t.capture(g)
    for _ in range(2):
        got = t.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), want)
    t.free()
    g.free()


def test_invalid_spheres_are_rejected(gpu_ctx, box):
    from dataclasses import replace

    from bajo_amd.rt import MAT, GpuScene, surface_id

    announce("mixed scene validation")
    sp = box.spheres.copy()
    sp[3] = 0.0
    with pytest.raises(RuntimeError, match="radius 0"):
        GpuScene(replace(box, spheres=sp), gpu_ctx)
    ss = box.sphere_surfaces.copy()
    ss[1] = surface_id(MAT.METAL, 9)
    with pytest.raises(RuntimeError, match="metal index out of range"):
        GpuScene(replace(box, sphere_surfaces=ss), gpu_ctx)
    with pytest.raises(ValueError, match="one surface id each"):
        GpuScene(replace(box, sphere_surfaces=ss[:-1]), gpu_ctx)

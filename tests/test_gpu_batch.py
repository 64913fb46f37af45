"""Batched environments: many triangle scenes rendered in one wavefront over one
segmented BVH build (bajo has no counterpart; its closest code is the segmented
BLAS build, bvh/gpu/builder/segmented_build.mojo:89-162, and render_gpu per scene,
rt/gpu/render.mojo:322-420).  Environment e of a batch render is bit-identical
to a render of its scene alone with seed + e (the per-scene launch loop), at
both node widths; it matches the oracle renderer (image RMSE <= 1e-3, > 99 %
bit-identical pixels); ray counts add up; a render target replays a batch
(HIP graph) bit-identically and refuses a batch of another size."""

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu
ALGS = {"PATH": 0, "NEE": 3, "MIS": 4}
W, H, SPP, DEPTH, SEED = 24, 16, 4, 3, 4242


def camera():
    from bajo_amd.bvh import Camera

    return Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()


@pytest.fixture(scope="module")
def batch():
    from bajo_amd import scenes

    return scenes.batched_environments(count=9, seed=77, lo=40, hi=900)


def settings():
    from bajo_amd.rt import RenderSettings

    return RenderSettings(W, H, SPP, SEED, DEPTH)


@pytest.mark.parametrize("width", [2, 8])
@pytest.mark.parametrize("name", ["PATH", "NEE", "MIS"])
def test_batch_equals_per_scene_loop(gpu_ctx, batch, width, name):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuScene, GpuSceneBatch, RenderSettings, enqueue_render_gpu

    announce(f"batch {name} width {width}")
    sc, off = batch
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=width)
    res = enqueue_render_gpu(settings(), camera(), g, ALGS[name], with_stats=True)
    got = res.numpy()
    assert got.shape == (9, H, W, 3)
    ext = 0
    for e in range(9):
        one = GpuScene(scenes.environment(sc, off, e), gpu_ctx, node_width=width)
        r = enqueue_render_gpu(RenderSettings(W, H, SPP, SEED + e, DEPTH), camera(), one, ALGS[name],
                               with_stats=True)
        np.testing.assert_array_equal(got[e], r.numpy(), err_msg=f"environment {e}")
        ext += r.stats["extension_rays"]
        one.free()
    assert res.stats["extension_rays"] == ext
    assert res.stats["status"] == 0 and res.stats["traversal_status"] == 0
    g.free()


@pytest.mark.parametrize("name", ["NEE", "MIS"])
def test_batch_environments_match_oracle(gpu_ctx, batch, name):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuSceneBatch, enqueue_render_gpu

    announce(f"batch vs oracle {name}")
    sc, off = batch
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=8)
    got = enqueue_render_gpu(settings(), camera(), g, ALGS[name]).numpy()
    for e in (0, 4, 8):
        env = scenes.environment(sc, off, e)
        want, _ = oracle.render(env.vertices, env.surfaces, env.lambertian, env.emissive, W, H, SPP, DEPTH, SEED + e,
                                ALGS[name], camera(), 8, node_width=8, leaf_size=4)
        img = got[e].reshape(-1)
        rmse = float(np.sqrt(np.mean((img.astype(np.float64) - want) ** 2)))
        exact = float(np.mean(img.view(np.uint32) == want.view(np.uint32)))
        print(f"environment {e}: rmse {rmse:.3e}, bit-identical {exact:.4f}")
        assert rmse <= 1e-3 and exact > 0.99
    g.free()


def test_batch_light_table_is_per_environment(gpu_ctx, batch):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuSceneBatch

    announce("batch light table")
    sc, off = batch
    g = GpuSceneBatch(sc, off, gpu_ctx)
    f, k, _ = g.light_table()
    assert k.size == 2 * 9
    for e in range(9):
        env = scenes.environment(sc, off, e)
        wf, wk, _ = oracle.light_table(env.vertices, env.surfaces, env.emissive)
        np.testing.assert_array_equal(k[2 * e:2 * e + 2], wk)
        np.testing.assert_array_equal(f[28 * e:28 * e + 28], wf)
    g.free()


def test_batch_render_target_replay(gpu_ctx, batch):
    import torch

    from bajo_amd.rt import GpuRtRenderTarget, GpuSceneBatch, enqueue_render_gpu

    announce("batch render target")
    sc, off = batch
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=8)
    want = enqueue_render_gpu(settings(), camera(), g, ALGS["NEE"]).numpy()
    t = GpuRtRenderTarget(settings(), camera(), ALGS["NEE"], gpu_ctx, environments=9)
    t.capture(g)
    for _ in range(2):
        got = t.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy(), want)
    t.free()
    single = GpuRtRenderTarget(settings(), camera(), ALGS["NEE"], gpu_ctx)
    with pytest.raises(RuntimeError, match="environments"):
        single.enqueue(g)
    single.free()
    g.free()


def test_autotuned_schedule_renders_the_same_image(gpu_ctx, batch):
    from dataclasses import replace

    from bajo_amd.autotune import tune_render
    from bajo_amd.rt import GpuSceneBatch, enqueue_render_gpu

    announce("batch autotune")
    sc, off = batch
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=8)
    want = enqueue_render_gpu(settings(), camera(), g, ALGS["MIS"]).numpy()
    tuned, table = tune_render(settings(), camera(), g, ALGS["MIS"], candidates=[0, 1536, 6144], reps=1)
    assert set(table) == {("path_capacity", c) for c in (0, 1536, 6144)} and all(v > 0 for v in table.values())
    assert tuned.path_capacity == min((c for _, c in table), key=lambda c: table[("path_capacity", c)])
    # every knob of the schedule, each value on its own: images stay bit-identical
    space = {"extend_waves": [0, 4, 6, 8], "shadow_group": [0, 1, 8, 16, 32], "shadow_walk": [0, 1, 2, 3, 4],
             "tail_group": [0, 1, 16, 32, 64], "tail_paths": [0, 1, 64, 1 << 20], "path_capacity": [0, 1536]}
    for knob, values in space.items():
        for value in values:
            got = enqueue_render_gpu(replace(settings(), **{knob: value}), camera(), g, ALGS["MIS"]).numpy()
            np.testing.assert_array_equal(got, want, err_msg=f"{knob} = {value}")
    # the full search: a cached result, an image equal to the default schedule's
    best, full = tune_render(settings(), camera(), g, ALGS["MIS"], reps=1)
    assert {k for k, _ in full} == {"extend_waves", "shadow_group", "shadow_walk", "tail_group", "tail_paths",
                                    "path_capacity"}
    again, cached = tune_render(settings(), camera(), g, ALGS["MIS"], reps=1)
    assert again == best and cached == full
    np.testing.assert_array_equal(enqueue_render_gpu(best, camera(), g, ALGS["MIS"]).numpy(), want)
    with pytest.raises(RuntimeError, match="extend_waves"):
        enqueue_render_gpu(replace(settings(), extend_waves=5), camera(), g, ALGS["MIS"])
    with pytest.raises(RuntimeError, match="shadow_walk"):
        enqueue_render_gpu(replace(settings(), shadow_walk=5), camera(), g, ALGS["MIS"])
    with pytest.raises(RuntimeError, match="shadow_group"):
        enqueue_render_gpu(replace(settings(), shadow_group=4), camera(), g, ALGS["MIS"])
    with pytest.raises(RuntimeError, match="tail_group"):
        enqueue_render_gpu(replace(settings(), tail_group=8), camera(), g, ALGS["MIS"])
    g.free()


@pytest.mark.parametrize("width", [2, 8])
def test_batch_with_a_camera_per_environment(gpu_ctx, batch, width):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene, GpuSceneBatch, RenderSettings, enqueue_render_gpu

    announce(f"batch per-environment cameras width {width}")
    sc, off = batch
    cams = np.stack([Camera((0.5 + 0.3 * np.cos(e), 0.6, 0.5 - 1.8 * np.cos(0.2 * e)), (0.5, 0.45, 0.5),
                            (0.0, 1.0, 0.0), 0.5).flatten() for e in range(9)])
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=width)
    shared_before = enqueue_render_gpu(settings(), camera(), g, ALGS["NEE"]).numpy()
    got = enqueue_render_gpu(settings(), cams, g, ALGS["NEE"]).numpy()
    for e in range(9):
        one = GpuScene(scenes.environment(sc, off, e), gpu_ctx, node_width=width)
        want = enqueue_render_gpu(RenderSettings(W, H, SPP, SEED + e, DEPTH), cams[e], one, ALGS["NEE"]).numpy()
        np.testing.assert_array_equal(got[e], want, err_msg=f"environment {e}")
        one.free()
    # a shared-camera render afterwards still uses the shared camera: the per-environment
    # cameras did not overwrite it
    np.testing.assert_array_equal(enqueue_render_gpu(settings(), camera(), g, ALGS["NEE"]).numpy(), shared_before)
    with pytest.raises(ValueError, match="cameras for 9 environments"):
        enqueue_render_gpu(settings(), cams[:4], g, ALGS["NEE"])
    g.free()

"""Builds are enqueue-only, so a scene's rebuild and its render are captured in one HIP graph
(bajo has no counterpart; its closest code is the segmented build, bvh/gpu/builder/
segmented_build.mojo:89-162, and the enqueue-only render, rt/gpu/render.mojo:61-235).

For a batch (the segmented LBVH and H-PLOC builds of every environment plus the batched render)
and for single scenes (LBVH and H-PLOC at widths 2, 4 and 8), the graph is replayed several times.
Between replays the non-emissive triangles move (a device copy on the stream) and eager work runs
on the same stream.  Every replay must equal, bit for bit, a fresh scene built eagerly from the same
triangles and rendered with the same settings; the context status word must stay clear."""

from dataclasses import replace

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu
NEE = 3
W, H, SPP, DEPTH, SEED = 24, 16, 4, 3, 4242


def batch_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()


def box_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()


def settings():
    from bajo_amd.rt import RenderSettings

    return RenderSettings(W, H, SPP, SEED, DEPTH)


def moved(scene, k: int) -> np.ndarray:
    """The scene's triangles with every non-emissive one scaled about the origin by 1 + 0.03 k."""
    from bajo_amd.rt import MAT

    v = np.ascontiguousarray(scene.vertices, np.float32).reshape(-1, 9).copy()
    kind = np.asarray(scene.surfaces, np.uint32) >> 28
    keep = kind != MAT.EMISSIVE
    v[keep] *= np.float32(1.0 + 0.03 * k)
    return v.reshape(-1)


def replay_and_compare(ctx, scene_data, make_scene, camera, environments):
    import torch

    from bajo_amd.rt import GpuRtRenderTarget, enqueue_render_gpu

    g = make_scene(scene_data)
    t = GpuRtRenderTarget(settings(), camera, NEE, ctx, environments=environments)
    t.capture(g, rebuild=True)
    images = []
    for k in (1, 2, 0):
        vk = moved(scene_data, k)
        g.set_vertices(torch.from_numpy(vk).to(ctx.torch_device))
        # eager work on the same stream between replays
        other = enqueue_render_gpu(settings(), camera, make_scene(scene_data), NEE).numpy()
        assert np.isfinite(other).all()
        got = t.replay()
        ctx.synchronize()
        img = got.cpu().numpy().copy()
        ref = make_scene(replace(scene_data, vertices=vk))
        want = enqueue_render_gpu(settings(), camera, ref, NEE).numpy()
        np.testing.assert_array_equal(img, want, err_msg=f"replay with scale step {k}")
        images.append(img)
        ref.free()
    assert not np.array_equal(images[0], images[1]), "moving the triangles must change the image"
    assert ctx.status() == 0
    t.free()
    g.free()


@pytest.mark.parametrize("method", [0, 1], ids=["lbvh", "hploc"])
@pytest.mark.parametrize("width", [2, 8])
def test_batch_rebuild_and_render_in_one_graph(gpu_ctx, method, width):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuSceneBatch

    announce(f"graph: batch rebuild + render, method {method}, width {width}")
    sc, off = scenes.batched_environments(count=9, seed=77, lo=40, hi=900)
    replay_and_compare(gpu_ctx, sc, lambda d: GpuSceneBatch(d, off, gpu_ctx, node_width=width, method=method),
                       batch_camera(), 9)


@pytest.mark.parametrize("method,width", [(0, 4), (0, 8), (1, 2), (1, 4), (1, 8)])
def test_scene_rebuild_and_render_in_one_graph(gpu_ctx, method, width):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuScene

    announce(f"graph: scene rebuild + render, method {method}, width {width}")
    sc = scenes.box_scene()
    replay_and_compare(gpu_ctx, sc, lambda d: GpuScene(d, gpu_ctx, node_width=width, method=method), box_camera(), 1)


def test_large_batch_graph_matches_eager(gpu_ctx):
    """512 environments of 256-16,384 triangles: deeper segments than the small batch."""
    import torch

    from bajo_amd import scenes
    from bajo_amd.rt import GpuRtRenderTarget, GpuSceneBatch, enqueue_render_gpu

    announce("graph: 512-environment batch rebuild + render")
    sc, off = scenes.batched_environments(count=512, seed=91, lo=256, hi=16384)
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=8)
    t = GpuRtRenderTarget(settings(), batch_camera(), NEE, gpu_ctx, environments=512)
    t.capture(g, rebuild=True)
    vk = moved(sc, 1)
    g.set_vertices(torch.from_numpy(vk).to(gpu_ctx.torch_device))
    got = t.replay()
    gpu_ctx.synchronize()
    img = got.cpu().numpy().copy()
    ref = GpuSceneBatch(replace(sc, vertices=vk), off, gpu_ctx, node_width=8)
    want = enqueue_render_gpu(settings(), batch_camera(), ref, NEE).numpy()
    np.testing.assert_array_equal(img, want)
    assert gpu_ctx.status() == 0
    ref.free()
    t.free()
    g.free()

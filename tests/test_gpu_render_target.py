"""Persistent render target (bajo's GpuRtRenderTarget / update_gpu_camera /
enqueue-only submission, rt/gpu/resources.mojo:38-155, rt/gpu/render.mojo:61-235):
a target's renders equal bajo_render's bit for bit, camera updates take effect,
a captured HIP graph replays the same image several times with eager work and
synchronizations between replays (and picks up a camera update), two targets
render one scene concurrently, and chunked targets equal chunked renders."""

import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu


def cams():
    from bajo_amd.bvh import Camera

    a = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    b = Camera((0.4, 1.2, -2.8), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), 0.55).flatten()
    return a, b


def eager(settings, cam, gscene, alg):
    from bajo_amd.rt import enqueue_render_gpu

    out = enqueue_render_gpu(settings, cam, gscene, alg).pixels
    torch.cuda.synchronize()
    return out.cpu().numpy().copy()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().copy()


@pytest.mark.parametrize("width,alg", [(2, "NEE"), (8, "MIS"), (8, "PATH")])
def test_target_matches_bajo_render_and_replays(gpu_ctx, width, alg):
    from bajo_amd import scenes
    from bajo_amd.rt import RENDER, GpuRtRenderTarget, GpuScene, RenderSettings

    announce(f"render target width={width} alg={alg}")
    algorithm = getattr(RENDER, alg)
    gscene = GpuScene(scenes.material_box_scene(), node_width=width)
    settings = RenderSettings(48, 32, 4, 11, 6)
    ca, cb = cams()
    want_a = eager(settings, ca, gscene, algorithm)
    want_b = eager(settings, cb, gscene, algorithm)
    assert not np.array_equal(want_a, want_b)
    t = GpuRtRenderTarget(settings, ca, algorithm)
    np.testing.assert_array_equal(host(t.enqueue(gscene)), want_a)
    t.update_camera(cb)
    np.testing.assert_array_equal(host(t.enqueue(gscene)), want_b)
    t.capture(gscene)
    for k in range(3):
        t.pixels.zero_()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(host(t.replay()), want_b, err_msg=f"replay {k}")
        eager(settings, ca, gscene, algorithm)  # eager work between replays
    t.update_camera(ca)
    np.testing.assert_array_equal(host(t.replay()), want_a, err_msg="replay after a camera update")
    np.testing.assert_array_equal(host(t.replay()), want_a, err_msg="second replay after a camera update")
    t.free()
    gscene.free()


def test_two_targets_and_chunked_target(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.rt import RENDER, GpuRtRenderTarget, GpuScene, RenderSettings

    announce("render targets: concurrent and chunked")
    gscene = GpuScene(scenes.box_scene(), node_width=8)
    ca, cb = cams()
    settings = RenderSettings(40, 30, 3, 5, 5)
    want_a = eager(settings, ca, gscene, RENDER.NEE)
    want_b = eager(settings, cb, gscene, RENDER.NEE)
    t1 = GpuRtRenderTarget(settings, ca, RENDER.NEE)
    t2 = GpuRtRenderTarget(settings, cb, RENDER.NEE)
    t1.enqueue(gscene)
    t2.enqueue(gscene)
    np.testing.assert_array_equal(host(t1.pixels), want_a)
    np.testing.assert_array_equal(host(t2.pixels), want_b)
    chunked = RenderSettings(40, 30, 3, 5, 5, path_capacity=1200)
    want_c = eager(chunked, ca, gscene, RENDER.NEE)
    t3 = GpuRtRenderTarget(chunked, ca, RENDER.NEE)
    np.testing.assert_array_equal(host(t3.enqueue(gscene)), want_c)
    t3.capture(gscene)
    np.testing.assert_array_equal(host(t3.replay()), want_c)
    np.testing.assert_array_equal(want_c, want_a)  # chunking does not change the image
    for t in (t1, t2, t3):
        t.free()
    gscene.free()


def test_target_rejects_bad_use(gpu_ctx):
    from bajo_amd._lib import BajoError
    from bajo_amd.rt import RENDER, GpuRtRenderTarget, RenderSettings

    announce("render target argument errors")
    ca, _ = cams()
    with pytest.raises(BajoError, match="positive"):
        GpuRtRenderTarget(RenderSettings(8, 8, 0, 1, 2), ca, RENDER.NEE)
    t = GpuRtRenderTarget(RenderSettings(8, 8, 1, 1, 2), ca, RENDER.NEE)
    with pytest.raises(BajoError, match="nothing captured"):
        t.replay()
    t.free()


def test_enqueue_of_another_scene_never_leaves_a_stale_graph(gpu_ctx):
    """A captured graph holds the target's arena buffers.  Enqueueing a scene whose
    launches need a larger arena (another node width or geometry can change the
    tail kernel's resident grid) must drop the graph before the arena is reallocated:
    afterwards a replay either reports that nothing is captured or still renders the
    captured scene exactly."""
    from bajo_amd import scenes
    from bajo_amd._lib import BajoError
    from bajo_amd.rt import RENDER, GpuRtRenderTarget, GpuScene, RenderSettings

    announce("render target: capture, then enqueue scenes of other widths and geometry")
    settings = RenderSettings(32, 24, 2, 9, 5)
    ca, _ = cams()
    binary = GpuScene(scenes.box_scene(), node_width=2)
    want = eager(settings, ca, binary, RENDER.NEE)
    others = [GpuScene(scenes.box_scene(), node_width=8), GpuScene(scenes.mixed_box_scene(), node_width=8),
              GpuScene(scenes.mixed_box_scene(), node_width=2)]
    t = GpuRtRenderTarget(settings, ca, RENDER.NEE)
    t.capture(binary)
    np.testing.assert_array_equal(host(t.replay()), want)
    for k, other in enumerate(others):
        np.testing.assert_array_equal(host(t.enqueue(other)), eager(settings, ca, other, RENDER.NEE),
                                      err_msg=f"enqueue of scene {k}")
        try:
            got = host(t.replay())
        except BajoError as e:
            assert "nothing captured" in str(e)
            t.capture(binary)  # capture again for the next scene
            got = host(t.replay())
        np.testing.assert_array_equal(got, want, err_msg=f"replay after scene {k}")
    t.free()
    for g in (binary, *others):
        g.free()

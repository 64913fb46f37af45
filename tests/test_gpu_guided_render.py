"""Guided rendering (NEE with one-sample 50/50 MIS of the BSDF and the learned
vMF guide at every Lambertian / GGX vertex): without a guide it is plain NEE
bit for bit; with a fixed guide it is deterministic; its radiance equals the
independent FP64 estimator's (any guide keeps it unbiased); training frames
collect samples and take Adam steps; and a trained guide lowers the error
against a converged reference below plain NEE's at equal samples per pixel."""

import numpy as np
import pytest
import torch

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room():
    from bajo_amd import scenes

    return scenes.room_scene(cells_per_tile=2, sphere_frequency=6)


def settings(w, h, spp, seed, depth=5):
    from bajo_amd.rt import RenderSettings

    return RenderSettings(w, h, spp, seed, depth)


def train_guide(gpu_ctx, gscene, cam, frames, w=48, h=27, spp=4, **kw):
    from bajo_amd.guide import PathGuide, render_guided

    guide = PathGuide(gpu_ctx, **kw)
    stats = None
    for f in range(frames):
        _, stats = render_guided(gscene, guide, settings(w, h, spp, 9000 + f), cam, train=True, with_stats=True)
    return guide, stats


@pytest.mark.parametrize("width", [2, 8])
def test_without_guide_is_plain_nee(gpu_ctx, room, width):
    from bajo_amd import scenes
    from bajo_amd.guide import render_guided
    from bajo_amd.rt import RENDER, GpuScene, enqueue_render_gpu

    announce(f"guided render without guide, width {width}")
    g = GpuScene(room, gpu_ctx, node_width=width)
    cam = scenes.room_camera()
    st = settings(40, 24, 4, 77)
    a, _ = render_guided(g, None, st, cam)
    b = enqueue_render_gpu(st, cam, g, RENDER.NEE).pixels
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    g.free()


def test_fixed_guide_is_deterministic_and_training_reports(gpu_ctx, room):
    from bajo_amd import scenes
    from bajo_amd.guide import render_guided
    from bajo_amd.rt import GpuScene

    announce("guided render determinism + training stats")
    g = GpuScene(room, gpu_ctx, node_width=8)
    cam = scenes.room_camera()
    guide, stats = train_guide(gpu_ctx, g, cam, 3, steps_per_frame=3)
    assert stats.training_samples > 100 and stats.adam_steps == 3 and np.isfinite(stats.loss)
    assert stats.guided_vertices > 48 * 27 * 4
    st = settings(40, 24, 8, 5)
    a, s1 = render_guided(g, guide, st, cam, with_stats=True)
    a = a.clone()
    b, _ = render_guided(g, guide, st, cam)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert s1.adam_steps == 0 and s1.training_samples == 0 and s1.infer_ns > 0
    assert torch.isfinite(a).all() and float(a.mean()) > 0
    guide.free()
    g.free()


@pytest.mark.parametrize("scene_name", ["room", "ggx_box"])
def test_guided_radiance_matches_fp64(gpu_ctx, room, scene_name):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.guide import render_guided
    from bajo_amd.rt import GpuScene

    announce(f"guided radiance vs FP64 ({scene_name})")
    if scene_name == "room":
        sc, cam, w, h, depth = room, scenes.room_camera(), 12, 8, 5
    else:
        sc = scenes.ggx_box_scene()
        cam, w, h, depth = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten(), 12, 8, 6
    g = GpuScene(sc, gpu_ctx, node_width=8)
    guide, _ = train_guide(gpu_ctx, g, cam, 6)
    reps, spp = 6, 512
    got = np.array([render_guided(g, guide, settings(w, h, spp, 700 + s, depth), cam)[0].cpu().numpy()
                    .reshape(-1, 3).mean(0) for s in range(reps)])
    want = np.array([oracle.render_fp64(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth,
                                        300 + s, 3, cam, 8, ggx=sc.ggx).reshape(-1, 3).mean(0) for s in range(reps)])
    for c in range(3):
        err = np.sqrt(got[:, c].var(ddof=1) / reps + want[:, c].var(ddof=1) / reps)
        print(f"{scene_name} channel {c}: guided {got[:, c].mean():.5f} fp64 {want[:, c].mean():.5f} +- {err:.5f}")
        assert abs(got[:, c].mean() - want[:, c].mean()) < 4.5 * err + 1e-6
    guide.free()
    g.free()


def test_trained_guide_beats_plain_nee_at_equal_spp(gpu_ctx, room):
    from bajo_amd import scenes
    from bajo_amd.guide import render_guided
    from bajo_amd.rt import RENDER, GpuScene, enqueue_render_gpu

    announce("guided vs plain MSE on the room")
    g = GpuScene(room, gpu_ctx, node_width=8)
    cam = scenes.room_camera()
    w, h = 48, 27
    ref = np.mean([enqueue_render_gpu(settings(w, h, 1024, 50 + s), cam, g, RENDER.NEE).numpy()
                   for s in range(4)], axis=0)
    guide, _ = train_guide(gpu_ctx, g, cam, 100, w=96, h=54, steps_per_frame=8)
    spp = 16
    mse_g = np.mean([np.mean((render_guided(g, guide, settings(w, h, spp, 500 + s), cam)[0].cpu().numpy() - ref) ** 2)
                     for s in range(4)])
    mse_p = np.mean([np.mean((enqueue_render_gpu(settings(w, h, spp, 600 + s), cam, g, RENDER.NEE).numpy() - ref) ** 2)
                     for s in range(4)])
    print(f"room {w}x{h} at {spp} spp: MSE guided {mse_g:.4e}, plain NEE {mse_p:.4e}, ratio {mse_g / mse_p:.3f}")
    assert mse_g < 0.8 * mse_p
    # the last training frame's samples: finite, positive weights, unit directions, inputs in range
    smp = guide.training_samples()
    assert len(smp) > 1000 and np.all(np.isfinite(smp)) and np.all(smp[:, 12] > 0)
    np.testing.assert_allclose(np.linalg.norm(smp[:, 9:12], axis=1), 1.0, atol=1e-5)
    assert smp[:, :3].min() >= 0 and smp[:, :3].max() <= 1
    # and they point through the doorway far more often than cosine sampling would
    lo, ext = np.array([0, 0, 0.0]), np.array([10, 3, 6.0])
    p = lo + smp[:, :3] * ext
    main = p[:, 0] < 7.9
    to_door = np.array([8.0, 1.0, 3.0]) - p
    to_door /= np.linalg.norm(to_door, axis=1, keepdims=True)
    near = (to_door * smp[:, 9:12]).sum(1) > np.cos(np.radians(30))
    w = smp[:, 12]
    assert w[main & near].sum() > 0.3 * w[main].sum()
    guide.free()
    g.free()


def test_training_is_deterministic(gpu_ctx, room):
    """Two guides trained on the same frames end bit-identical: the frame's samples are stored in a fixed
    order and every reduction of the training step runs in a fixed order."""
    from bajo_amd import scenes
    from bajo_amd.guide import render_guided
    from bajo_amd.rt import GpuScene

    announce("guided training determinism")
    g = GpuScene(room, gpu_ctx, node_width=8)
    cam = scenes.room_camera()
    a, _ = train_guide(gpu_ctx, g, cam, 6, w=64, h=36, steps_per_frame=4)
    b, _ = train_guide(gpu_ctx, g, cam, 6, w=64, h=36, steps_per_frame=4)
    np.testing.assert_array_equal(a.params(), b.params())
    np.testing.assert_array_equal(a.training_samples(), b.training_samples())
    ia, _ = render_guided(g, a, settings(32, 18, 2, 3), cam)
    ib, _ = render_guided(g, b, settings(32, 18, 2, 3), cam)
    np.testing.assert_array_equal(ia.cpu().numpy(), ib.cpu().numpy())
    a.free()
    b.free()
    g.free()

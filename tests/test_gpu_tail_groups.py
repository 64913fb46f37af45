"""The lane-group tail gives the per-lane tail's images bit for bit.

The tail kernel carries a chunk's last paths through its remaining bounces.  With `tail_group = 16`
(or 32, 64) each path gets that many lanes: the extension ray through GroupClosest (which defers to Walk8 when two
candidates lie within rounding of each other, since Walk8's winner can then depend on test order),
shading on the group's leader, the shadow ray through GroupWalk.  Images must equal the one-lane
tail's, and the ray counts the oracle's (bajo's wavefront protocol, test/rt/test_gpu.mojo:346-730).
A scene whose every triangle is duplicated with a different albedo makes every hit an exact tie,
so the deferred path decides every pixel there."""

from dataclasses import replace

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu
PATH, NEE, MIS = 0, 3, 4


def box_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()


def duplicated(sc):
    """Every triangle twice: the copy is the same triangle with its own Lambertian albedo (emitters kept)."""
    from bajo_amd.rt import MAT, surface_id

    v = np.asarray(sc.vertices, np.float32).reshape(-1, 9)
    s = np.asarray(sc.surfaces, np.uint32)
    lamb = np.asarray(sc.lambertian, np.float32).reshape(-1, 3)
    kind = s >> 28
    extra = np.where(kind == MAT.LAMBERTIAN)[0]
    copy_s = s.copy()
    new_albedo = (lamb[(s[extra] & 0x0FFFFFFF)] * np.float32(0.5) + np.float32(0.25)).astype(np.float32)
    copy_s[extra] = np.array([surface_id(MAT.LAMBERTIAN, lamb.shape[0] + i) for i in range(extra.size)], np.uint32)
    return replace(sc, vertices=np.concatenate([v, v]).reshape(-1), surfaces=np.concatenate([s, copy_s]),
                   lambertian=np.concatenate([lamb, new_albedo]).reshape(-1))


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("alg", [PATH, NEE, MIS], ids=["PATH", "NEE", "MIS"])
@pytest.mark.parametrize("dup", [False, True], ids=["box", "duplicated-box"])
def test_tail_groups_equal_per_lane_tail(gpu_ctx, width, alg, dup):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce(f"tail groups width {width} alg {alg} duplicated {dup}")
    sc = scenes.box_scene()
    if dup:
        sc = duplicated(sc)
    g = GpuScene(sc, gpu_ctx, node_width=width, leaf_size=4)
    w, h, spp, depth, seed = 40, 30, 4, 5, 91
    # the tail takes over from bounce 1 (every path), so the group tail traces most extension rays
    base = RenderSettings(w, h, spp, seed, depth, tail_paths=1 << 30)
    ref = enqueue_render_gpu(replace(base, tail_group=1), box_camera(), g, alg, with_stats=True)
    want, rays = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed, alg,
                               box_camera(), 8, node_width=width, leaf_size=4)
    img = ref.numpy().reshape(-1)
    rmse = float(np.sqrt(np.mean((img.astype(np.float64) - want) ** 2)))
    assert rmse <= 1e-3 and float(np.mean(img.view(np.uint32) == want.view(np.uint32))) > 0.99
    for tg in (16, 32, 64, 0):
        res = enqueue_render_gpu(replace(base, tail_group=tg), box_camera(), g, alg, with_stats=True)
        np.testing.assert_array_equal(res.numpy().view(np.uint32), ref.numpy().view(np.uint32),
                                      err_msg=f"tail_group {tg}")
        assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
        assert res.stats["rays_per_bounce"] == ref.stats["rays_per_bounce"]
        assert res.stats["status"] == 0 and res.stats["traversal_status"] == 0
    g.free()


def test_tail_groups_icosphere_and_batch(gpu_ctx):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuScene, GpuSceneBatch, RenderSettings, enqueue_render_gpu

    announce("tail groups icosphere and batch")
    sc = scenes.icosphere_scene(1 << 16, frequency=58)
    g = GpuScene(sc, gpu_ctx, node_width=8)
    cam = Camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 0.41421356).flatten()
    for alg in (PATH, NEE):
        base = RenderSettings(96, 54, 4, 2026, 5)
        imgs = [enqueue_render_gpu(replace(base, tail_group=tg, tail_paths=tp), cam, g, alg).numpy()
                for tg, tp in ((1, 0), (16, 0), (16, 1 << 30), (32, 1 << 30), (64, 1 << 30), (0, 0))]
        for im in imgs[1:]:
            np.testing.assert_array_equal(im, imgs[0])
    g.free()
    bsc, off = scenes.batched_environments(count=9, seed=77, lo=40, hi=900)
    gb = GpuSceneBatch(bsc, off, gpu_ctx, node_width=8)
    bcam = Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()
    base = RenderSettings(24, 16, 4, 4242, 3, tail_paths=1 << 30)
    a = enqueue_render_gpu(replace(base, tail_group=1), bcam, gb, NEE).numpy()
    for tg in (16, 32, 64):
        b = enqueue_render_gpu(replace(base, tail_group=tg), bcam, gb, NEE).numpy()
        np.testing.assert_array_equal(a, b, err_msg=f"tail_group {tg}")
    gb.free()

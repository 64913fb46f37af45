"""Every any-hit walk of the renderer's shadow rays gives the oracle's image.

The shadow-ray kernel has four per-lane walks (one or two nodes per step, leaf records one ahead
or four at a time), lane-group walks (8, 16 or 32 lanes per ray, one node per 8 lanes per step)
and a one-ray-per-wave walk for sparse bounces (up to eight nodes and all their leaf records per
step).  Occlusion does not depend on the order a walk visits nodes and
records in, so every form must give the default schedule's image bit for bit, with the oracle's
ray counts, at widths 4 and 8, for a single scene and for a batch of environments; the default
image meets the renderer's oracle bar (bajo's wavefront protocol, test/rt/test_gpu.mojo:346-730)."""

from dataclasses import replace

import numpy as np
import pytest

import oracle
from conftest import announce

pytestmark = pytest.mark.gpu
NEE, AO, MIS = 3, 2, 4
NEVER, ALWAYS = 0xFFFFFFFF, 1 << 30


def box_camera():
    from bajo_amd.bvh import Camera

    return Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("alg", [NEE, MIS, AO], ids=["NEE", "MIS", "AO"])
def test_shadow_walks_match_oracle(gpu_ctx, width, alg):
    from bajo_amd import scenes
    from bajo_amd.rt import GpuScene, RenderSettings, enqueue_render_gpu

    announce(f"shadow walks width {width} alg {alg}")
    sc = scenes.box_scene()
    g = GpuScene(sc, gpu_ctx, node_width=width, leaf_size=4)
    w, h, spp, depth, seed = 40, 30, 4, 5 if alg != AO else 1, 77
    want, rays = oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, w, h, spp, depth, seed, alg,
                               box_camera(), 8, node_width=width, leaf_size=4)
    base = RenderSettings(w, h, spp, seed, depth)
    ref = enqueue_render_gpu(base, box_camera(), g, alg).numpy().reshape(-1)
    # the default schedule against the oracle (the renderer's bar: RMSE <= 1e-3, > 99 % identical pixels)
    rmse = float(np.sqrt(np.mean((ref.astype(np.float64) - want) ** 2)))
    assert rmse <= 1e-3 and float(np.mean(ref.view(np.uint32) == want.view(np.uint32))) > 0.99
    forms = [(walk, 1, sparse) for walk in (1, 2, 3, 4) for sparse in (NEVER, 0, ALWAYS)]
    forms += [(0, group, sparse) for group in (8, 16, 32) for sparse in (NEVER, 0)]
    for walk, group, sparse in forms:
        if True:
            st = replace(base, shadow_walk=walk, sparse_shadow_rays=sparse, shadow_group=group)
            res = enqueue_render_gpu(st, box_camera(), g, alg, with_stats=True)
            got = res.numpy().reshape(-1)
            np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32),
                                          err_msg=f"walk {walk}, group {group}, sparse {sparse:#x}")
            assert res.stats["extension_rays"] == rays[0] and res.stats["shadow_rays"] == rays[1]
            assert res.stats["status"] == 0 and res.stats["traversal_status"] == 0
            assert res.stats["shadow_node_visits"] > 0
    g.free()


@pytest.mark.parametrize("width", [4, 8])
def test_batch_sparse_shadows_equal_per_lane(gpu_ctx, width):
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera
    from bajo_amd.rt import GpuSceneBatch, RenderSettings, enqueue_render_gpu

    announce(f"batch sparse shadows width {width}")
    sc, off = scenes.batched_environments(count=9, seed=77, lo=40, hi=900)
    g = GpuSceneBatch(sc, off, gpu_ctx, node_width=width)
    cam = Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()
    base = RenderSettings(24, 16, 4, 4242, 3)
    imgs = [enqueue_render_gpu(replace(base, sparse_shadow_rays=s, shadow_group=gr), cam, g, NEE).numpy()
            for s, gr in ((NEVER, 1), (ALWAYS, 1), (NEVER, 8), (NEVER, 32))]
    for img in imgs[1:]:
        np.testing.assert_array_equal(img, imgs[0])
    g.free()

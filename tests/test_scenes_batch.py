"""Host-side pieces of the batched-environment render: the environment generator's layout and the
autotuner's candidate schedules (no GPU)."""

import numpy as np

from bajo_amd import scenes
from bajo_amd.autotune import KNOBS, default_candidates, default_space
from bajo_amd.rt import MAT, surface_id


def test_batched_environments_layout():
    sc, off = scenes.batched_environments(count=5, seed=9, lo=20, hi=200)
    soups, soup_off = scenes.batched_soups(5, 9, 20, 200)
    assert off.dtype == np.uint64 and off.size == 6 and off[0] == 0
    np.testing.assert_array_equal(np.diff(off.astype(np.int64)), np.diff(soup_off.astype(np.int64)) + 2)
    tri = sc.vertices.reshape(-1, 9)
    light = scenes.quad((0.2, 1.2, 0.2), (0.8, 1.2, 0.2), (0.8, 1.2, 0.8), (0.2, 1.2, 0.8)).reshape(2, 9)
    for k in range(5):
        a, b = int(off[k]), int(off[k + 1])
        np.testing.assert_array_equal(tri[a:b - 2], soups.reshape(-1, 9)[int(soup_off[k]):int(soup_off[k + 1])])
        np.testing.assert_array_equal(tri[b - 2:b], light)
        assert np.all(sc.surfaces[a:b - 2] == surface_id(MAT.LAMBERTIAN, k))
        assert np.all(sc.surfaces[b - 2:b] == surface_id(MAT.EMISSIVE, 0))
        env = scenes.environment(sc, off, k)
        assert env.vertices.size == 9 * (b - a)
    # the light faces down: (p1 - p0) x (p2 - p0) points to -y
    p = light[0].reshape(3, 3).astype(np.float64)
    assert np.cross(p[1] - p[0], p[2] - p[0])[1] < 0
    assert sc.lambertian.size == 15 and np.all((sc.lambertian >= 0.2) & (sc.lambertian <= 0.8))


def test_autotune_candidates():
    assert default_candidates(1000, 4) == [0]
    c = default_candidates(67_108_864, 4)
    assert c[0] == 0 and c[1:] == [1 << k for k in range(20, 26)]
    assert all(x % 3 == 0 for x in default_candidates(10_000_000, 3)[1:])


def test_autotune_space_covers_every_knob():
    sp = default_space(67_108_864, 4)
    assert tuple(sp) == KNOBS
    assert sp["extend_waves"] == [0, 4, 8] and sp["tail_paths"][0] == 0 and sp["shadow_walk"][0] == 0
    assert sp["shadow_group"][0] == 0 and sp["tail_group"] == [0, 1, 16, 64]
    assert sp["path_capacity"] == default_candidates(67_108_864, 4)

"""The guided-rendering room generator: sizes, the exact Lambertian / GGX
split, GGX parameter ranges, a closed pair of rooms (no ray from inside
escapes) and a downward-facing light visible only through the doorway."""

import numpy as np

import oracle
from bajo_amd import scenes
from bajo_amd.rt import MAT


def test_room_default_size_and_material_split():
    s = scenes.room_scene()
    assert s.triangle_count == 1_201_794
    kind = s.surfaces >> 28
    assert (kind == MAT.EMISSIVE).sum() == 2
    assert (kind == MAT.LAMBERTIAN).sum() == (kind == MAT.GGX).sum() == (s.triangle_count - 2) // 2
    g = s.ggx.reshape(-1, 4)
    assert g[:, 3].min() >= 0.05 and g[:, 3].max() <= 0.5
    assert g[:, :3].min() >= 0.5 and g[:, :3].max() <= 0.95
    lo, hi = scenes.bounds(s.vertices)
    np.testing.assert_array_equal(lo, [0, 0, 0])
    np.testing.assert_array_equal(hi, [10, 3, 6])


def test_room_is_closed_and_lit_through_the_doorway():
    s = scenes.room_scene(cells_per_tile=2, sphere_frequency=6)
    v = s.vertices
    bvh = oracle.build_lbvh(v)
    rng = np.random.default_rng(3)
    n = 4000
    o = np.column_stack([rng.uniform(0.2, 7.8, n), rng.uniform(0.2, 2.8, n), rng.uniform(0.2, 5.8, n)])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((8, n), np.float32)  # field-major planes: origin, t_min, direction, t_max
    rays[0:3], rays[3], rays[4:7], rays[7] = o.T, 1e-4, d.T, 3.0e38
    hits = oracle.trace_bvh2(v, bvh, rays.reshape(-1))
    prims = oracle.hit_prims(hits)
    assert np.all(prims != oracle.SENTINEL), "a ray from inside the main room escaped"
    # the light faces down into the light room
    tri = v.reshape(-1, 3, 3)[s.surfaces >> 28 == MAT.EMISSIVE]
    nrm = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert np.all(nrm[:, 1] < 0)
    # rays from the main room toward the light reach it only through the doorway
    light_c = tri.reshape(-1, 3).mean(axis=0)
    to_l = light_c - o
    dist = np.linalg.norm(to_l, axis=1)
    rays[4:7] = (to_l / dist[:, None]).T
    hits = oracle.trace_bvh2(v, bvh, rays.reshape(-1))
    lit = (s.surfaces[oracle.hit_prims(hits)] >> 28) == MAT.EMISSIVE
    p_door = o + (8.0 - o[:, :1]) / (light_c[0] - o[:, :1]) * to_l  # where the segment crosses x = 8
    through = (p_door[:, 2] > 2.5) & (p_door[:, 2] < 3.5) & (p_door[:, 1] < 2.0)
    assert lit.any()
    assert not np.any(lit & ~through)

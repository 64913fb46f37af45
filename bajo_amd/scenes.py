"""Seeded synthetic scenes (host numpy arrays of f32[9N] triangles).

The generators run in the library's host code (bajo_amd/csrc/scenes.cpp):
PCG32 streams, a random triangle soup and a displaced geodesic icosphere.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

# Soup configuration: vertices uniform in [0,1]^3, edge lengths U[0.002, 0.02].
SOUP_EDGE_MIN = 0.002
SOUP_EDGE_MAX = 0.02


def soup(n: int, initstate: int = 2027, initseq: int = 54, edge_min=SOUP_EDGE_MIN, edge_max=SOUP_EDGE_MAX) -> np.ndarray:
    out = np.empty(9 * n, dtype=np.float32)
    _lib.check(_lib.lib().bajo_scene_soup(initstate, initseq, n, edge_min, edge_max, out.ctypes.data_as(ctypes.c_void_p)))
    return out


def soup_grid(rows: int = 64, cols: int = 64, per_copy: int = 4096, initstate: int = 2029) -> np.ndarray:
    """A rows x cols grid of triangle soups flattened to one mesh (the 16.7M-triangle H-PLOC / 8-wide
    workload at the default 64 x 64 x 4096): copy c = row * cols + col is soup(per_copy) from PCG32 stream
    (initstate, initseq = c), translated by (col, row, 0) into its unit cell."""
    out = np.empty((rows * cols, per_copy, 3, 3), dtype=np.float32)
    for row in range(rows):
        for col in range(cols):
            c = row * cols + col
            tri = soup(per_copy, initstate=initstate, initseq=c).reshape(per_copy, 3, 3)
            tri[:, :, 0] += np.float32(col)
            tri[:, :, 1] += np.float32(row)
            out[c] = tri
    return out.reshape(-1)


def icosphere(n: int, frequency: int = 458, noise_seed: int = 2028, amplitude: float = 0.05, noise_frequency: float = 4.0):
    if n > 20 * frequency * frequency:
        raise ValueError(f"frequency {frequency} yields only {20 * frequency * frequency} triangles")
    out = np.empty(9 * n, dtype=np.float32)
    _lib.check(
        _lib.lib().bajo_scene_icosphere(
            frequency, n, noise_seed, amplitude, noise_frequency, out.ctypes.data_as(ctypes.c_void_p)
        )
    )
    return out


def batched_soups(count: int = 4096, seed: int = 2031, lo: int = 256, hi: int = 65536) -> tuple[np.ndarray, np.ndarray]:
    """`count` independent soups (the batched-environment workload): triangle counts
    round(exp(U[ln lo, ln hi])) from numpy's PCG64 stream `seed`, soup k drawn by the soup generator from
    PCG32 stream (seed, k).  Returns the meshes back to back (f32[9 N]) and their offsets u64[count + 1]."""
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    out = np.empty(9 * int(offsets[-1]), dtype=np.float32)
    for k, n in enumerate(counts):
        out[9 * int(offsets[k]): 9 * int(offsets[k + 1])] = soup(int(n), initstate=seed, initseq=k)
    return out, offsets


# bajo's regular-grid BVH fixture (bajo/benchmark/bvh_fixtures.mojo:11-58): 256 x 256 triangles on the
# plane z = 2 and 4 rays per triangle, one in four a miss; its published closest-hit checksum
# (sum over hits of t + u + v + normal + prim) is 6443188224.0 (bench/results/bvh_cpu/bvh_cpu.md).
GRID_SIDE = 256
GRID_CHECKSUM = 6443188224.0


def grid_fixture() -> tuple[np.ndarray, np.ndarray]:
    """(vertices f32[9 * 65536], rays f32[8 * 262144] field-major: origin, t_min, direction, t_max)."""
    i = np.arange(GRID_SIDE * GRID_SIDE)
    cx = ((i % GRID_SIDE).astype(np.float32) - np.float32(GRID_SIDE) * np.float32(0.5)) * np.float32(3.0)
    cy = ((i // GRID_SIDE).astype(np.float32) - np.float32(GRID_SIDE) * np.float32(0.5)) * np.float32(3.0)
    z = np.full_like(cx, 2.0)
    v = np.stack([cx - 0.75, cy - 0.75, z, cx + 0.75, cy - 0.75, z, cx, cy + 0.75, z], axis=1).astype(np.float32)
    r = np.arange(4 * GRID_SIDE * GRID_SIDE)
    prim = r % (GRID_SIDE * GRID_SIDE)
    miss = r % 4 == 0
    rays = np.zeros((8, r.size), np.float32)
    rays[0] = np.where(miss, np.float32(10000.0) + r.astype(np.float32), cx[prim])
    rays[1] = np.where(miss, np.float32(10000.0), cy[prim])
    rays[2] = 0.0
    rays[3] = 0.0
    rays[6] = 1.0
    rays[7] = np.finfo(np.float32).max
    return v.reshape(-1), rays.reshape(-1)


def grid_checksum(hits: np.ndarray) -> float:
    """bajo's checksum of closest hits f32[8 R] (u, v, prim bits, inst, normal.xyz, t per hit;
    bench/bvh/bench_cpu_bvh_grid.mojo:35-52): sum over hits of t + u + v + normal.x + normal.y +
    normal.z + prim, in float64."""
    h = np.ascontiguousarray(hits, np.float32).reshape(-1, 8)
    prim = h[:, 2].view(np.uint32)
    hit = prim != 0xFFFFFFFF
    hh = h[hit].astype(np.float64)
    return float((hh[:, 7] + hh[:, 0] + hh[:, 1] + hh[:, 4] + hh[:, 5] + hh[:, 6]).sum()
                 + prim[hit].astype(np.float64).sum())


def bounds(vertices: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    v = vertices.reshape(-1, 3)
    return v.min(axis=0), v.max(axis=0)


# ---- render scenes ---------------------------------------------------------------

def quad(p0, p1, p2, p3) -> np.ndarray:
    """Two triangles (p0, p1, p2), (p0, p2, p3); normal = (p1-p0) x (p2-p0)."""
    return np.array([*p0, *p1, *p2, *p0, *p2, *p3], np.float32)


def _scene(tri_blocks, kinds, lambertian, emissive):
    from .rt import MAT, SceneData, surface_id

    vertices = np.concatenate([np.asarray(b, np.float32).reshape(-1) for b in tri_blocks])
    surfaces = np.concatenate(kinds).astype(np.uint32)
    assert surfaces.size * 9 == vertices.size
    return SceneData(vertices, surfaces, np.asarray(lambertian, np.float32).reshape(-1), np.asarray(emissive, np.float32).reshape(-1))


def icosphere_scene(n_sphere: int = 4_194_304, frequency: int = 458, albedo_seed: int = 2029, emission: float = 15.0):
    """Displaced icosphere (Lambertian, albedo per triangle U[0.2, 0.8]^3) lit
    by one downward-facing quad light of the given emission above it."""
    from .rt import MAT, surface_id

    sphere = icosphere(n_sphere, frequency=frequency)
    albedo = np.random.default_rng(albedo_seed).uniform(0.2, 0.8, (n_sphere, 3)).astype(np.float32)
    light = quad((-0.5, 1.6, -0.5), (0.5, 1.6, -0.5), (0.5, 1.6, 0.5), (-0.5, 1.6, 0.5))
    kinds = [
        (np.uint32(MAT.LAMBERTIAN) << 28) | np.arange(n_sphere, dtype=np.uint32),
        np.full(2, surface_id(MAT.EMISSIVE, 0), np.uint32),
    ]
    return _scene([sphere, light], kinds, albedo, [emission, emission, emission])


def box_scene(n_soup: int = 120, seed: int = 5, emission: float = 12.0):
    """Small closed-ish box (floor, back wall, two side walls) with a soup of
    small triangles inside and a quad light under the ceiling; for parity and
    radiance tests."""
    from .rt import MAT, surface_id

    floor = quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    back = quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1))
    left = quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))
    right = quad((1, 0, 1), (1, 2, 1), (1, 2, -1), (1, 0, -1))
    light = quad((-0.3, 1.9, -0.3), (0.3, 1.9, -0.3), (0.3, 1.9, 0.3), (-0.3, 1.9, 0.3))
    rng = np.random.default_rng(seed)
    v0 = rng.uniform([-0.8, 0.05, -0.6], [0.8, 1.4, 0.8], (n_soup, 3))
    e = rng.uniform(-0.25, 0.25, (n_soup, 2, 3))
    soup_t = np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32)
    walls = np.concatenate([floor, back, left, right])
    n_walls = walls.size // 9
    lamb = np.concatenate([
        np.array([[0.7, 0.7, 0.7], [0.7, 0.7, 0.7], [0.6, 0.2, 0.2], [0.2, 0.6, 0.2]], np.float32),
        rng.uniform(0.2, 0.8, (n_soup, 3)).astype(np.float32),
    ])
    wall_mat = np.repeat(np.arange(4, dtype=np.uint32), 2)
    kinds = [
        (np.uint32(MAT.LAMBERTIAN) << 28) | wall_mat,
        (np.uint32(MAT.LAMBERTIAN) << 28) | (4 + np.arange(n_soup, dtype=np.uint32)),
        np.full(2, surface_id(MAT.EMISSIVE, 0), np.uint32),
    ]
    assert n_walls == 8
    return _scene([walls, soup_t, light], kinds, lamb, [emission, emission, emission])


def material_box_scene(seed: int = 7, n_soup: int = 60, emission: float = 12.0):
    """The box scene's walls and light with all four of bajo's material kinds:
    a glass ball (dielectric, index 1.5), a rough metal ball (fuzz 0.25), a
    mirror panel (fuzz 0) on the back wall, and a soup of small triangles
    cycling Lambertian / rough metal / glass; for material parity and
    radiance tests."""
    from .rt import MAT, SceneData, surface_id

    floor = quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    back = quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1))
    left = quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))
    right = quad((1, 0, 1), (1, 2, 1), (1, 2, -1), (1, 0, -1))
    mirror = quad((-0.5, 0.4, 0.99), (-0.5, 1.4, 0.99), (0.5, 1.4, 0.99), (0.5, 0.4, 0.99))
    light = quad((-0.3, 1.9, -0.3), (0.3, 1.9, -0.3), (0.3, 1.9, 0.3), (-0.3, 1.9, 0.3))
    ball = icosphere(320, frequency=4, amplitude=0.0).reshape(-1, 3)
    glass = (ball * 0.3 + np.array([0.4, 0.3, 0.1], np.float32)).astype(np.float32).reshape(-1)
    metal_ball = (ball * 0.25 + np.array([-0.45, 0.25, 0.3], np.float32)).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(seed)
    v0 = rng.uniform([-0.8, 0.6, -0.6], [0.8, 1.5, 0.6], (n_soup, 3))
    e = rng.uniform(-0.2, 0.2, (n_soup, 2, 3))
    soup_t = np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)
    lamb = np.array([[0.7, 0.7, 0.7], [0.7, 0.7, 0.7], [0.6, 0.2, 0.2], [0.2, 0.6, 0.2], [0.5, 0.4, 0.3]], np.float32)
    metal = np.array([[0.9, 0.9, 0.9, 0.0], [0.8, 0.65, 0.3, 0.25], [0.7, 0.75, 0.8, 0.35]], np.float32)
    dielectric = np.array([1.5], np.float32)
    wall_mat = np.repeat(np.arange(4, dtype=np.uint32), 2)
    cycle = np.array([surface_id(MAT.LAMBERTIAN, 4), surface_id(MAT.METAL, 2), surface_id(MAT.DIELECTRIC, 0)],
                     np.uint32)
    kinds = [
        (np.uint32(MAT.LAMBERTIAN) << 28) | wall_mat,
        np.full(2, surface_id(MAT.METAL, 0), np.uint32),
        np.full(glass.size // 9, surface_id(MAT.DIELECTRIC, 0), np.uint32),
        np.full(metal_ball.size // 9, surface_id(MAT.METAL, 1), np.uint32),
        cycle[np.arange(n_soup) % 3],
        np.full(2, surface_id(MAT.EMISSIVE, 0), np.uint32),
    ]
    walls = np.concatenate([floor, back, left, right])
    vertices = np.concatenate([walls, mirror, glass, metal_ball, soup_t, light]).astype(np.float32)
    surfaces = np.concatenate(kinds).astype(np.uint32)
    assert surfaces.size * 9 == vertices.size
    return SceneData(vertices, surfaces, lamb.reshape(-1), np.array([emission] * 3, np.float32), metal.reshape(-1),
                     dielectric)


def ggx_box_scene(seed: int = 17, n_soup: int = 60, emission: float = 12.0):
    """The box scene with GGX conductors beside Lambertian surfaces: a glossy
    GGX floor (roughness 0.2), a GGX ball (roughness 0.08), a Lambertian
    ball, and a soup cycling Lambertian / GGX (roughness 0.05-0.5); for GGX
    parity and radiance tests."""
    from .rt import MAT, SceneData, surface_id

    floor = quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    back = quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1))
    left = quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))
    right = quad((1, 0, 1), (1, 2, 1), (1, 2, -1), (1, 0, -1))
    light = quad((-0.3, 1.9, -0.3), (0.3, 1.9, -0.3), (0.3, 1.9, 0.3), (-0.3, 1.9, 0.3))
    ball = icosphere(320, frequency=4, amplitude=0.0).reshape(-1, 3)
    gball = (ball * 0.3 + np.array([0.4, 0.3, 0.1], np.float32)).astype(np.float32).reshape(-1)
    lball = (ball * 0.25 + np.array([-0.45, 0.25, 0.3], np.float32)).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(seed)
    v0 = rng.uniform([-0.8, 0.6, -0.6], [0.8, 1.5, 0.6], (n_soup, 3))
    e = rng.uniform(-0.2, 0.2, (n_soup, 2, 3))
    soup_t = np.concatenate([v0, v0 + e[:, 0], v0 + e[:, 1]], axis=1).astype(np.float32).reshape(-1)
    lamb = np.array([[0.7, 0.7, 0.7], [0.6, 0.2, 0.2], [0.2, 0.6, 0.2], [0.5, 0.4, 0.3]], np.float32)
    ggx = np.array([[0.8, 0.8, 0.8, 0.2], [0.95, 0.7, 0.35, 0.08], [0.6, 0.7, 0.9, 0.05], [0.9, 0.9, 0.9, 0.5]],
                   np.float32)
    soup_kinds = np.array([surface_id(MAT.LAMBERTIAN, 3), surface_id(MAT.GGX, 2), surface_id(MAT.GGX, 3)], np.uint32)
    kinds = [
        np.full(2, surface_id(MAT.GGX, 0), np.uint32),
        np.full(2, surface_id(MAT.LAMBERTIAN, 0), np.uint32),
        np.full(2, surface_id(MAT.LAMBERTIAN, 1), np.uint32),
        np.full(2, surface_id(MAT.LAMBERTIAN, 2), np.uint32),
        np.full(gball.size // 9, surface_id(MAT.GGX, 1), np.uint32),
        np.full(lball.size // 9, surface_id(MAT.LAMBERTIAN, 0), np.uint32),
        soup_kinds[np.arange(n_soup) % 3],
        np.full(2, surface_id(MAT.EMISSIVE, 0), np.uint32),
    ]
    vertices = np.concatenate([floor, back, left, right, gball, lball, soup_t, light]).astype(np.float32)
    surfaces = np.concatenate(kinds).astype(np.uint32)
    assert surfaces.size * 9 == vertices.size
    return SceneData(vertices, surfaces, lamb.reshape(-1), np.array([emission] * 3, np.float32), ggx=ggx.reshape(-1))


def mixed_box_scene(seed: int = 11, n_small: int = 24, emission: float = 12.0, sphere_emission: float = 6.0):
    """The box scene's walls and quad light with spheres of every material kind (bajo's mixed
    triangle + sphere scenes, rt/gpu/mixed_path.mojo): a Lambertian ball, a rough metal ball, a glass
    ball with a hollow inner sphere (negative radius: the normal points inward), an emissive ball
    (a sphere light), and `n_small` small spheres cycling Lambertian / metal / glass; for mixed-scene
    parity and radiance tests."""
    from .rt import MAT, SceneData, surface_id

    floor = quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    back = quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1))
    left = quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))
    right = quad((1, 0, 1), (1, 2, 1), (1, 2, -1), (1, 0, -1))
    light = quad((-0.3, 1.9, -0.3), (0.3, 1.9, -0.3), (0.3, 1.9, 0.3), (-0.3, 1.9, 0.3))
    walls = np.concatenate([floor, back, left, right])
    vertices = np.concatenate([walls, light]).astype(np.float32)
    surfaces = np.concatenate([
        (np.uint32(MAT.LAMBERTIAN) << 28) | np.repeat(np.arange(4, dtype=np.uint32), 2),
        np.full(2, surface_id(MAT.EMISSIVE, 0), np.uint32),
    ]).astype(np.uint32)
    lamb = np.array([[0.7, 0.7, 0.7], [0.7, 0.7, 0.7], [0.6, 0.2, 0.2], [0.2, 0.6, 0.2], [0.5, 0.4, 0.3]], np.float32)
    metal = np.array([[0.8, 0.65, 0.3, 0.25], [0.7, 0.75, 0.8, 0.0]], np.float32)
    dielectric = np.array([1.5], np.float32)
    emit = np.array([[emission] * 3, [sphere_emission, 0.8 * sphere_emission, 0.5 * sphere_emission]], np.float32)
    big = [
        ((-0.5, 0.3, 0.2, 0.3), surface_id(MAT.LAMBERTIAN, 4)),
        ((0.45, 0.25, 0.35, 0.25), surface_id(MAT.METAL, 0)),
        ((0.1, 0.35, -0.35, 0.3), surface_id(MAT.DIELECTRIC, 0)),
        ((0.1, 0.35, -0.35, -0.25), surface_id(MAT.DIELECTRIC, 0)),
        ((-0.55, 1.2, 0.6, 0.12), surface_id(MAT.EMISSIVE, 1)),
    ]
    rng = np.random.default_rng(seed)
    c = rng.uniform([-0.8, 0.6, -0.6], [0.8, 1.5, 0.6], (n_small, 3))
    r = rng.uniform(0.04, 0.1, n_small)
    cycle = np.array([surface_id(MAT.LAMBERTIAN, 4), surface_id(MAT.METAL, 1), surface_id(MAT.DIELECTRIC, 0)],
                     np.uint32)
    spheres = np.concatenate([np.array([b[0] for b in big], np.float32).reshape(-1),
                              np.concatenate([c, r[:, None]], axis=1).astype(np.float32).reshape(-1)])
    sphere_surfaces = np.concatenate([np.array([b[1] for b in big], np.uint32), cycle[np.arange(n_small) % 3]])
    return SceneData(vertices, surfaces, lamb.reshape(-1), emit.reshape(-1), metal.reshape(-1), dielectric,
                     spheres.astype(np.float32), sphere_surfaces.astype(np.uint32))


def sphere_field_scene(n: int = 2000, seed: int = 13, emission: float = 4.0):
    """A spheres-only scene (no triangles): a large Lambertian ground sphere, `n` small spheres of every
    material kind scattered above it and a few emissive spheres as the lights."""
    from .rt import MAT, SceneData, surface_id

    rng = np.random.default_rng(seed)
    xz = rng.uniform(-4.0, 4.0, (n, 2))
    r = rng.uniform(0.05, 0.2, n)
    c = np.stack([xz[:, 0], r, xz[:, 1]], axis=1)
    kinds = rng.integers(0, 3, n)
    lamb_n = int((kinds == 0).sum())
    lamb = np.concatenate([[[0.5, 0.5, 0.5]], rng.uniform(0.1, 0.9, (lamb_n, 3))]).astype(np.float32)
    metal = np.array([[0.7, 0.6, 0.5, 0.0], [0.8, 0.8, 0.9, 0.3]], np.float32)
    surf = np.empty(n, np.uint32)
    li = 1
    for i, k in enumerate(kinds):
        if k == 0:
            surf[i] = surface_id(MAT.LAMBERTIAN, li)
            li += 1
        elif k == 1:
            surf[i] = surface_id(MAT.METAL, i % 2)
        else:
            surf[i] = surface_id(MAT.DIELECTRIC, 0)
    lights = np.array([[-2.0, 3.0, -1.0, 0.6], [2.5, 2.5, 1.5, 0.4], [0.0, 4.0, 3.0, 0.8]], np.float32)
    spheres = np.concatenate([
        np.array([0.0, -1000.0, 0.0, 1000.0], np.float32),
        np.concatenate([c, r[:, None]], axis=1).astype(np.float32).reshape(-1),
        lights.reshape(-1),
    ])
    sphere_surfaces = np.concatenate([[surface_id(MAT.LAMBERTIAN, 0)], surf,
                                      np.full(3, surface_id(MAT.EMISSIVE, 0), np.uint32)]).astype(np.uint32)
    return SceneData(np.zeros(0, np.float32), np.zeros(0, np.uint32), lamb.reshape(-1),
                     np.array([emission] * 3, np.float32), metal.reshape(-1), np.array([1.5], np.float32),
                     spheres.astype(np.float32), sphere_surfaces)


def batched_environments(count: int = 4096, seed: int = 2031, lo: int = 256, hi: int = 65536, emission: float = 6.0,
                         soups=None):
    """The batched-environment render scenes: environment k is soup k of `batched_soups(count, seed, lo, hi)`
    (Lambertian, albedo U[0.2, 0.8]^3 per environment from numpy's PCG64 stream seed + 1) under one
    downward-facing quad light over the unit cube (y = 1.2, x and z in [0.2, 0.8], the given emission).
    Returns (SceneData of every environment back to back, offsets u64[count + 1]).  `soups` may pass the
    result of `batched_soups(count, seed, lo, hi)` computed before."""
    from .rt import MAT, SceneData, surface_id

    soups, soup_off = soups if soups is not None else batched_soups(count, seed, lo, hi)
    light = quad((0.2, 1.2, 0.2), (0.8, 1.2, 0.2), (0.8, 1.2, 0.8), (0.2, 1.2, 0.8))
    n = np.diff(soup_off.astype(np.int64))
    offsets = np.concatenate([[0], np.cumsum(n + 2)]).astype(np.uint64)
    tri = np.empty((int(offsets[-1]), 9), np.float32)
    surf = np.empty(int(offsets[-1]), np.uint32)
    src = soups.reshape(-1, 9)
    for k in range(count):
        a, b = int(offsets[k]), int(offsets[k + 1])
        tri[a:b - 2] = src[int(soup_off[k]):int(soup_off[k + 1])]
        tri[b - 2:b] = light.reshape(2, 9)
        surf[a:b - 2] = surface_id(MAT.LAMBERTIAN, k)
        surf[b - 2:b] = surface_id(MAT.EMISSIVE, 0)
    albedo = np.random.default_rng(seed + 1).uniform(0.2, 0.8, (count, 3)).astype(np.float32)
    return SceneData(tri.reshape(-1), surf, albedo.reshape(-1), np.array([emission] * 3, np.float32)), offsets


def environment(batch, offsets, k: int):
    """Environment k of a batch as a scene of its own (materials shared)."""
    from dataclasses import replace

    a, b = int(offsets[k]), int(offsets[k + 1])
    return replace(batch, vertices=batch.vertices[9 * a:9 * b].copy(), surfaces=batch.surfaces[a:b].copy())


# ---- the guided-rendering room -------------------------------------------------------

# Room layout (metres): the main room is [0, 8] x [0, 3] x [0, 6] (x, y, z); a small light room
# [8, 10] x [0, 3] x [2, 4] sits behind the main room's x = 8 wall, which has a doorway
# z in [2.5, 3.5], y in [0, 2]; the only emitter is a small downward-facing quad under the light
# room's ceiling, so the main room is lit through the doorway.
ROOM_TILE = 0.5
ROOM_LIGHT = ((9.2, 2.95, 2.8), (9.6, 2.95, 2.8), (9.6, 2.95, 3.2), (9.2, 2.95, 3.2))
ROOM_CAMERA = ((0.8, 1.7, 0.5), (8.0, 1.1, 3.3), (0.0, 1.0, 0.0), 0.7)
ROOM_SPHERES = 8


def _room_walls():
    """Wall pieces as (origin, u axis, v axis, tiles along u, tiles along v, skipped tiles); every piece
    faces the inside of its room (u x v is the inward normal)."""
    T = ROOM_TILE
    pieces = [
        # main room
        ((0, 0, 0), (0, 0, T), (T, 0, 0), 12, 16, ()),          # floor y = 0
        ((0, 3, 0), (T, 0, 0), (0, 0, T), 16, 12, ()),          # ceiling y = 3
        ((0, 0, 0), (T, 0, 0), (0, T, 0), 16, 6, ()),           # wall z = 0
        ((0, 0, 6), (0, T, 0), (T, 0, 0), 6, 16, ()),           # wall z = 6
        ((0, 0, 0), (0, T, 0), (0, 0, T), 6, 12, ()),           # wall x = 0
        # wall x = 8 (both rooms see it): doorway tiles z in [2.5, 3.5] (iv 5, 6), y in [0, 2] (iu 0..3)
        ((8, 0, 0), (0, 0, T), (0, T, 0), 12, 6, tuple((iu, iv) for iu in (5, 6) for iv in range(4))),
        # light room
        ((8, 0, 2), (T, 0, 0), (0, 0, T), 4, 4, ()),            # floor y = 0
        ((8, 3, 2), (0, 0, T), (T, 0, 0), 4, 4, ()),            # ceiling y = 3
        ((8, 0, 2), (0, T, 0), (T, 0, 0), 6, 4, ()),            # wall z = 2
        ((8, 0, 4), (T, 0, 0), (0, T, 0), 4, 6, ()),            # wall z = 4
        ((10, 0, 2), (0, 0, T), (0, T, 0), 4, 6, ()),           # wall x = 10
    ]
    return pieces


def room_scene(seed: int = 2030, cells_per_tile: int = 16, sphere_frequency: int = 70, emission: float = 100.0):
    """The guided-rendering room: tessellated walls of the two rooms above (tiles of 0.5 m, each
    `cells_per_tile`^2 cells of two triangles), `ROOM_SPHERES` displaced icospheres of frequency
    `sphere_frequency` on the main room's floor, and one emissive quad in the light room.  Wall tiles
    are checkered per wall piece (even tiles Lambertian, odd tiles GGX) and spheres alternate, so
    exactly half of the non-emissive triangles are Lambertian and half GGX.  Lambertian albedo
    U[0.3, 0.8]^3, GGX F0 U[0.5, 0.95]^3 and roughness U[0.05, 0.5], sphere radius U[0.35, 0.6] and
    positions (rejection-sampled apart) come from numpy's PCG64 stream `seed`.  At the defaults:
    417,792 wall + 784,000 sphere + 2 light = 1,201,794 triangles."""
    from .rt import MAT, SceneData, surface_id

    rng = np.random.default_rng(seed)
    c = int(cells_per_tile)
    tris, surf, lamb, ggx = [], [], [], []
    g = np.arange(c, dtype=np.float64)
    cu, cv = np.meshgrid(g, g, indexing="ij")
    cu, cv = cu.reshape(-1), cv.reshape(-1)
    for origin, u, v, nu, nv, skip in _room_walls():
        o, u, v = (np.asarray(x, np.float64) for x in (origin, u, v))
        for iu in range(nu):
            for iv in range(nv):
                if (iu, iv) in skip:
                    continue
                p00 = o + np.outer(iu + cu / c, u) + np.outer(iv + cv / c, v)
                p10 = p00 + u / c
                p11 = p10 + v / c
                p01 = p00 + v / c
                t = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)])
                tris.append(t.reshape(-1, 9))
                if (iu + iv) % 2 == 0:
                    sid = surface_id(MAT.LAMBERTIAN, len(lamb))
                    lamb.append(rng.uniform(0.3, 0.8, 3))
                else:
                    sid = surface_id(MAT.GGX, len(ggx))
                    ggx.append(np.concatenate([rng.uniform(0.5, 0.95, 3), rng.uniform(0.05, 0.5, 1)]))
                surf.append(np.full(2 * c * c, sid, np.uint32))
    unit = icosphere(20 * sphere_frequency * sphere_frequency, frequency=sphere_frequency).reshape(-1, 3)
    placed = []
    while len(placed) < ROOM_SPHERES:
        r = rng.uniform(0.35, 0.6)
        x, z = rng.uniform(1.0, 7.0), rng.uniform(1.0, 5.0)
        if all((x - px) ** 2 + (z - pz) ** 2 > (r + pr + 0.1) ** 2 for px, pz, pr in placed):
            placed.append((x, z, r))
    for k, (x, z, r) in enumerate(placed):
        ball = unit.astype(np.float64) * (r / 1.05) + np.array([x, r, z])
        tris.append(ball.reshape(-1, 9))
        if k % 2 == 0:
            sid = surface_id(MAT.LAMBERTIAN, len(lamb))
            lamb.append(rng.uniform(0.3, 0.8, 3))
        else:
            sid = surface_id(MAT.GGX, len(ggx))
            ggx.append(np.concatenate([rng.uniform(0.5, 0.95, 3), rng.uniform(0.05, 0.5, 1)]))
        surf.append(np.full(ball.shape[0] // 3, sid, np.uint32))
    light = quad(*ROOM_LIGHT).reshape(-1, 9)
    if np.cross(light[0, 3:6] - light[0, 0:3], light[0, 6:9] - light[0, 0:3])[1] > 0:
        light = quad(*ROOM_LIGHT[::-1]).reshape(-1, 9)  # face down, into the light room
    tris.append(light.astype(np.float64))
    surf.append(np.full(2, surface_id(MAT.EMISSIVE, 0), np.uint32))
    vertices = np.concatenate(tris).astype(np.float32).reshape(-1)
    return SceneData(vertices, np.concatenate(surf).astype(np.uint32), np.asarray(lamb, np.float32).reshape(-1),
                     np.array([emission] * 3, np.float32), ggx=np.asarray(ggx, np.float32).reshape(-1))


def room_camera():
    """The room's camera: in the main room's corner, looking at the doorway."""
    from .bvh import Camera

    o, t, up, fov = ROOM_CAMERA
    return Camera(o, t, up, fov).flatten()

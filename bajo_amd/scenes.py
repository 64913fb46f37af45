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


def bounds(vertices: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    v = vertices.reshape(-1, 3)
    return v.min(axis=0), v.max(axis=0)

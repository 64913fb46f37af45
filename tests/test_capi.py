"""The C-ABI library loads and exports every entry point of include/bajo_hip.h.
CPU only: no compute call is made, only symbol lookup and host helpers."""

import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "bajo_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bajo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from bajo_amd._lib import SIGNATURES, lib

    handle = lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(handle, name), f"{name} declared in bajo_hip.h but not exported"
    assert set(names) == set(SIGNATURES), "ctypes signatures and header disagree"


def test_status_strings_and_errors_without_gpu():
    from bajo_amd._lib import lib

    assert lib().bajo_status_string(0) == b"ok"
    assert lib().bajo_status_string(1) == b"traversal stack overflow"
    # argument validation happens before any device call
    assert lib().bajo_build_triangle_bvh(None, None, 0, 0, None, None) == -1
    assert b"NULL" in lib().bajo_last_error()


def test_camera_make_matches_oracle():
    import oracle
    from bajo_amd.bvh import Camera

    got = Camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356).flatten()
    want = oracle.camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.41421356)
    np.testing.assert_array_equal(got, want)


def test_scene_generators_are_deterministic_and_bounded():
    from bajo_amd import scenes

    a = scenes.soup(4096, initstate=5, initseq=54)
    b = scenes.soup(4096, initstate=5, initseq=54)
    np.testing.assert_array_equal(a, b)
    t = a.reshape(-1, 3, 3).astype(np.float64)
    assert t[:, 0].min() >= 0 and t[:, 0].max() < 1
    for e in (1, 2):
        length = np.linalg.norm(t[:, e] - t[:, 0], axis=1)
        assert length.min() > 0.002 - 1e-6 and length.max() < 0.02 + 1e-6


def test_icosphere_is_crack_free():
    from bajo_amd import scenes

    f = 6
    v = scenes.icosphere(20 * f * f, frequency=f)
    tris = v.reshape(-1, 3, 3)
    # every undirected edge (bit-identical endpoints) is shared by exactly two triangles
    keys = {}
    for tri in tris:
        pts = [tuple(p.view(np.uint32)) for p in tri]
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = tuple(sorted((pts[a], pts[b])))
            keys[e] = keys.get(e, 0) + 1
    assert set(keys.values()) == {2}
    r = np.linalg.norm(tris.reshape(-1, 3), axis=1)
    assert r.min() > 0.94 and r.max() < 1.06

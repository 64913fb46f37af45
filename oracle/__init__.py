"""CPU oracle -- test infrastructure only.

ctypes wrapper of ``oracle/liboracle.so``, the plain-C restatement of bajo's
BVH algorithms (orc_lbvh.c, orc_trace.c).  Imported by the tests, the smoke
check and the CPU-baseline leg of bench.py, as the checker; the product
package ``bajo_amd`` never imports it.
"""

from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_float, c_int, c_int64, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
SENTINEL = 0xFFFFFFFF
LEAF_FLAG = 0x80000000

_SIG = {
    "orc_expand_bits": (c_uint32, [c_uint32]),
    "orc_morton": (c_uint32, [c_float, c_float, c_float]),
    "orc_common_prefix": (c_int, [c_uint32, c_uint32, c_uint32, c_uint32]),
    "orc_triangle_bounds": (None, [c_void_p, c_uint64, c_void_p]),
    "orc_scene_bounds": (None, [c_void_p, c_uint64, c_void_p]),
    "orc_morton_codes": (None, [c_void_p, c_uint64, c_void_p, c_void_p]),
    "orc_radix_sort_pairs_u64": (c_int, [c_void_p, c_void_p, c_uint64, c_int, c_int]),
    "orc_radix_sort_morton_pairs": (c_int, [c_void_p, c_void_p, c_uint64]),
    "orc_lbvh_find_split": (c_int64, [c_void_p, c_int64, c_int64, c_int64]),
    "orc_karras_topology": (c_uint32, [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p]),
    "orc_refit": (None, [c_uint32, c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_refit_climb": (c_uint32, [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p]),
    "orc_build_lbvh": (
        c_uint32,
        [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "orc_build_lbvh_bounds": (
        c_uint32,
        [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "orc_build_lbvh_topdown": (c_uint32, [c_void_p, c_void_p, c_void_p, c_uint64, c_int, c_void_p, c_void_p, c_void_p]),
    "orc_canonical_karras": (c_uint64, [c_uint32, c_void_p, c_void_p, c_void_p]),
    "orc_canonical_topdown": (c_uint64, [c_void_p, c_void_p, c_void_p]),
    "orc_camera_make": (None, [c_void_p, c_void_p, c_void_p, c_float, c_float, c_void_p]),
    "orc_camera_rays": (None, [c_void_p, c_uint64, c_uint32, c_uint32, c_void_p]),
    "orc_trace_bvh2": (
        None,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_uint64, c_int, c_void_p, c_void_p],
    ),
    "orc_trace_bvh2_mt": (
        None,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_uint64, c_int, c_void_p, c_int],
    ),
    "orc_trace_brute": (None, [c_void_p, c_uint64, c_void_p, c_uint64, c_void_p]),
    "orc_intersect_aabb": (c_int, [c_void_p, c_void_p, c_void_p, c_float, c_void_p]),
    "orc_intersect_tri_eager": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_void_p]),
    "orc_intersect_tri_scaled": (
        c_int,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_void_p],
    ),
    "orc_rcp_direction": (None, [c_void_p, c_void_p]),
    "orc_philox4x32_10": (None, [c_void_p, c_uint32, c_uint32, c_void_p]),
    "orc_rng_stage_floats": (None, [c_uint64, c_uint32, c_uint32, c_uint32, c_void_p]),
    "orc_sincos_turns": (None, [c_float, c_void_p, c_void_p]),
    "orc_alias_table": (None, [c_void_p, c_uint32, c_float, c_void_p, c_void_p]),
    "orc_light_table": (None, [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_grid_exponent": (c_uint32, [c_float, c_float]),
    "orc_quantize": (None, [c_float, c_uint32, c_float, c_float, c_void_p, c_void_p]),
    "orc_collapse_wide": (
        c_uint64, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_void_p, c_uint32, c_void_p, c_uint64]
    ),
    "orc_trace_wide": (
        c_uint64,
        [c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_void_p, c_uint64, c_int, c_void_p, c_void_p],
    ),
    "orc_hploc_topology": (
        c_uint32,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p],
    ),
    "orc_preorder_topology": (c_uint64, [c_uint32, c_void_p, c_void_p, c_void_p, c_uint64, c_void_p]),
    "orc_preorder_binary": (c_uint64, [c_uint32, c_void_p, c_void_p, c_uint64, c_void_p]),
    "orc_hploc_binary": (
        c_uint32,
        [c_uint32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p],
    ),
    "orc_hploc_lowest_leaf_numbering": (None, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p]),
    "orc_ao_rays": (None, [c_void_p, c_void_p, c_uint64, c_uint32, c_uint64, c_void_p]),
    "orc_render": (
        c_int, [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_int],
    ),
    "orc_render_fp64": (
        c_int, [c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_int],
    ),
    "orc_trace_bvh2_spheres": (
        None,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_uint64, c_int, c_void_p, c_void_p],
    ),
    "orc_trace_brute_spheres": (None, [c_void_p, c_uint64, c_void_p, c_uint64, c_void_p]),
    "orc_trace_wide_spheres": (
        c_uint64,
        [c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_void_p, c_uint64, c_int, c_void_p, c_void_p],
    ),
    "orc_affine_inverse": (c_int, [c_void_p, c_void_p]),
    "orc_transform_box": (None, [c_void_p, c_void_p, c_void_p]),
    "orc_tlas_trace": (
        None,
        [c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_void_p, c_uint32, c_void_p,
         c_uint64, c_int, c_void_p],
    ),
    "orc_render_mixed": (
        c_int, [c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p, c_void_p, c_void_p, c_int],
    ),
    "orc_eval_ggx": (None, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    "orc_collapse_wide_n": (
        c_uint64,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_void_p, c_uint32, c_uint32, c_void_p, c_uint64],
    ),
    "orc_render_set_timing": (None, [c_int]),
    "orc_render_best_seconds": (ctypes.c_double, []),
    "orc_bf16_round": (c_float, [c_float]),
    "orc_guide_encode": (None, [c_void_p, c_void_p]),
    "orc_guide_forward": (None, [c_void_p, c_void_p, c_uint64, c_int, c_void_p]),
    "orc_guide_loss_grad": (ctypes.c_double, [c_void_p, c_void_p, c_uint64, c_int, c_void_p]),
    "orc_guide_adam": (
        None,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
         ctypes.c_double],
    ),
    "orc_guide_mixture": (c_float, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "orc_vmf_pdf64": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
    "orc_vmf_sample_cos64": (ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
    "orc_sample_ggx": (None, [c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_void_p]),
    "orc_render_fp64_mixed": (
        c_int, [c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                c_void_p, c_void_p, c_void_p, c_int],
    ),
    "orc_light_table_mixed": (
        None, [c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
               c_void_p],
    ),
    "orc_pow_unit": (c_float, [c_float, c_float]),
    "orc_reflectance": (c_float, [c_float, c_float]),
    "orc_reflect3": (None, [c_void_p, c_void_p, c_void_p]),
    "orc_sample_metal": (None, [c_void_p, c_void_p, c_void_p, c_float, c_float, c_float, c_void_p]),
    "orc_eval_metal": (None, [c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    "orc_sample_dielectric": (None, [c_void_p, c_void_p, c_float, c_int, c_float, c_void_p]),
}


class RenderSettings(ctypes.Structure):
    _fields_ = [
        ("width", c_uint32),
        ("height", c_uint32),
        ("spp", c_uint32),
        ("max_depth", c_uint32),
        ("seed", c_uint64),
        ("algorithm", c_int),
        ("node_width", c_int),
        ("leaf_size", c_int),
        ("method", c_int),
        ("ggx", c_void_p),
    ]

_lib = None


def build() -> Path:
    """Compiles liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load(path: Path) -> ctypes.CDLL:
    h = ctypes.CDLL(str(path))
    for name, (res, args) in _SIG.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    return h


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _load(LIB_PATH)
    return _lib


NATIVE_PATH = HERE / "liboracle_native.so"


def native_lib() -> ctypes.CDLL:
    """The same sources built -O3 for the host's own instruction set (the CPU baseline's build): compiled here,
    on the machine that runs it."""
    subprocess.run(["make", "-s", "-C", str(HERE), "native"], check=True)
    return _load(NATIVE_PATH)


class use_lib:
    """Context manager: routes the module's functions through another build of the oracle."""

    def __init__(self, handle: ctypes.CDLL):
        self.handle = handle

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = self.handle
        return self.handle

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved


def _p(a: np.ndarray) -> c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(c_void_p)


def f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


# ---- LBVH -----------------------------------------------------------------

def triangle_bounds(vertices: np.ndarray) -> np.ndarray:
    v = f32(vertices)
    n = v.size // 9
    out = np.empty(6 * n, np.float32)
    lib().orc_triangle_bounds(_p(v), n, _p(out))
    return out


def scene_bounds(bounds6: np.ndarray) -> np.ndarray:
    b = f32(bounds6)
    out = np.empty(12, np.float32)
    lib().orc_scene_bounds(_p(b), b.size // 6, _p(out))
    return out


def morton_codes(bounds6: np.ndarray, centroid_box6: np.ndarray) -> np.ndarray:
    b = f32(bounds6)
    cb = f32(centroid_box6)
    out = np.empty(b.size // 6, np.uint32)
    lib().orc_morton_codes(_p(b), b.size // 6, _p(cb), _p(out))
    return out


def radix_sort_pairs_u64(keys: np.ndarray, values: np.ndarray | None, begin_bit=0, end_bit=64):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    v = None if values is None else np.ascontiguousarray(values, dtype=np.uint32).copy()
    rc = lib().orc_radix_sort_pairs_u64(_p(k), _p(v) if v is not None else None, k.size, begin_bit, end_bit)
    assert rc == 0
    return k, v


def build_lbvh(vertices: np.ndarray) -> dict:
    """Single-segment LBVH in bajo's GPU binary layout (GpuBinaryBoundsBvh)."""
    v = f32(vertices)
    return _build_lbvh(v, v.size // 9, lib().orc_build_lbvh)


def build_lbvh_bounds(leaf_bounds: np.ndarray) -> dict:
    """The same build over leaf boxes f32[6N] (lo.xyz, hi.xyz)."""
    b = f32(leaf_bounds)
    return _build_lbvh(b, b.size // 6, lib().orc_build_lbvh_bounds)


def _build_lbvh(v, n, fn) -> dict:
    m = max(n - 1, 0)
    out = {
        "sorted_keys": np.empty(n, np.uint64),
        "leaf_ids": np.empty(n, np.uint32),
        "node_meta": np.empty(max(4 * m, 1), np.uint32),
        "leaf_parent": np.empty(n, np.uint32),
        "leaf_counts": np.empty(max(m, 1), np.uint32),
        "node_bounds": np.empty(max(12 * m, 1), np.float32),
        "scene_bounds": np.empty(12, np.float32),
    }
    root = fn(
        _p(v), n, _p(out["sorted_keys"]), _p(out["leaf_ids"]), _p(out["node_meta"]), _p(out["leaf_parent"]),
        _p(out["leaf_counts"]), _p(out["node_bounds"]), _p(out["scene_bounds"]),
    )
    out["node_meta"] = out["node_meta"][: 4 * m]
    out["leaf_counts"] = out["leaf_counts"][:m]
    out["node_bounds"] = out["node_bounds"][: 12 * m]
    out["root"] = int(root)
    return out


def refit_climb(leaf_bounds_sorted: np.ndarray, node_meta: np.ndarray, leaf_parent: np.ndarray):
    """The GPU refit's bottom-up climb with its corruption checks: (status, node_bounds f32[12(n-1)])."""
    b = f32(leaf_bounds_sorted)
    n = b.size // 6
    meta = np.ascontiguousarray(node_meta, dtype=np.uint32)
    par = np.ascontiguousarray(leaf_parent, dtype=np.uint32)
    out = np.zeros(12 * max(n - 1, 1), np.float32)
    status = lib().orc_refit_climb(_p(b), n, _p(meta), _p(par), _p(out))
    return int(status), out[: 12 * max(n - 1, 0)]


def karras_topology(sorted_codes: np.ndarray):
    c = np.ascontiguousarray(sorted_codes, dtype=np.uint32)
    n = c.size
    m = max(n - 1, 1)
    meta = np.empty(4 * m, np.uint32)
    parent = np.empty(max(n, 1), np.uint32)
    counts = np.empty(m, np.uint32)
    root = lib().orc_karras_topology(_p(c), n, _p(meta), _p(parent), _p(counts))
    return int(root), meta[: 4 * (n - 1)], parent[:n], counts[: n - 1]


def build_lbvh_topdown(sorted_codes, sorted_ids, bounds6, leaf_size=1) -> dict:
    c = np.ascontiguousarray(sorted_codes, dtype=np.uint32)
    ids = np.ascontiguousarray(sorted_ids, dtype=np.uint32)
    b = f32(bounds6)
    n = c.size
    cap = max(2 * n - 1, 1)
    boxes = np.empty(6 * cap, np.float32)
    first = np.empty(cap, np.uint32)
    count = np.empty(cap, np.uint32)
    used = lib().orc_build_lbvh_topdown(_p(c), _p(ids), _p(b), n, leaf_size, _p(boxes), _p(first), _p(count))
    return {"boxes": boxes[: 6 * used], "first": first[:used], "count": count[:used], "node_count": int(used)}


def canonical_karras(root: int, node_meta: np.ndarray, leaf_counts: np.ndarray) -> np.ndarray:
    meta = np.ascontiguousarray(node_meta, dtype=np.uint32)
    cnt = np.ascontiguousarray(leaf_counts, dtype=np.uint32)
    m = meta.size // 4
    out = np.empty(max(3 * m, 3), np.uint32)
    k = lib().orc_canonical_karras(root, _p(meta) if meta.size else None, _p(cnt) if cnt.size else None, _p(out))
    return out[: 3 * k].reshape(-1, 3)


def canonical_topdown(tree: dict) -> np.ndarray:
    n_nodes = tree["node_count"]
    out = np.empty(max(3 * n_nodes, 3), np.uint32)
    k = lib().orc_canonical_topdown(_p(tree["first"]), _p(tree["count"]), _p(out))
    return out[: 3 * k].reshape(-1, 3)


# ---- rays -------------------------------------------------------------------

def camera(origin, target, up=(0.0, 1.0, 0.0), fov_scale=0.75, focus_dist=1.0) -> np.ndarray:
    out = np.empty(20, np.float32)
    lib().orc_camera_make(_p(f32(origin)), _p(f32(target)), _p(f32(up)), fov_scale, focus_dist, _p(out))
    return out


def camera_rays(cams: np.ndarray, ray_count: int, width: int, height: int) -> np.ndarray:
    c = f32(cams)
    out = np.empty(8 * ray_count, np.float32)
    lib().orc_camera_rays(_p(c), ray_count, width, height, _p(out))
    return out


def trace_bvh2(vertices, lbvh: dict, rays: np.ndarray, mode: int = 0, stats: bool = False, threads: int = 1):
    v = f32(vertices)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    meta = lbvh["node_meta"] if lbvh["node_meta"].size else np.zeros(4, np.uint32)
    nb = lbvh["node_bounds"] if lbvh["node_bounds"].size else np.zeros(12, np.float32)
    if threads > 1 and not stats:
        lib().orc_trace_bvh2_mt(_p(v), _p(lbvh["leaf_ids"]), _p(meta), _p(nb), lbvh["root"], _p(r), R, mode, _p(hits), threads)
        return hits
    st = np.empty(5 * R, np.uint32) if stats else None
    lib().orc_trace_bvh2(
        _p(v), _p(lbvh["leaf_ids"]), _p(meta), _p(nb), lbvh["root"], _p(r), R, mode, _p(hits),
        _p(st) if st is not None else None,
    )
    return (hits, st) if stats else hits


def trace_brute(vertices, rays: np.ndarray) -> np.ndarray:
    v = f32(vertices)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    lib().orc_trace_brute(_p(v), v.size // 9, _p(r), R, _p(hits))
    return hits


def hit_prims(hits: np.ndarray) -> np.ndarray:
    return hits.reshape(-1, 8)[:, 2].view(np.uint32)


def hit_t(hits: np.ndarray) -> np.ndarray:
    return hits.reshape(-1, 8)[:, 7]


# ---- renderer -----------------------------------------------------------------

def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    out = np.empty(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), int(key[0]), int(key[1]), _p(out))
    return out


def rng_stage_floats(seed: int, path_id: int, stage: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    lib().orc_rng_stage_floats(seed, path_id, stage, count, _p(out))
    return out


def sincos_turns(u: float):
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    lib().orc_sincos_turns(u, _p(s), _p(c))
    return float(s[0]), float(c[0])


def alias_table(weights) -> tuple[np.ndarray, np.ndarray]:
    w = f32(weights)
    n = w.size
    total = np.float32(0)
    for x in w:
        total = np.float32(total + x)
    prob = np.empty(max(n, 1), np.float32)
    alias = np.empty(max(n, 1), np.uint32)
    lib().orc_alias_table(_p(w), n, float(total), _p(prob), _p(alias))
    return prob[:n], alias[:n]


def light_table(vertices, surfaces, emissive, spheres=None, sphere_surfaces=None):
    v = f32(vertices)
    n = v.size // 9
    surf = np.ascontiguousarray(surfaces, dtype=np.uint32)
    e = f32(emissive) if len(emissive) else np.zeros(3, np.float32)
    sp, ss, ns = _spheres(spheres, sphere_surfaces)
    fields = np.empty(14 * max(n + ns, 1), np.float32)
    kinds = np.empty(max(n + ns, 1), np.uint32)
    count = np.zeros(1, np.uint32)
    total = np.zeros(1, np.float32)
    lib().orc_light_table_mixed(_p(v), n, _p(surf), _p(sp), ns, _p(ss), _p(e), _p(fields), _p(kinds), _p(count),
                                _p(total))
    k = int(count[0])
    return fields[: 14 * k], kinds[:k], float(total[0])


def _spheres(spheres, sphere_surfaces):
    if spheres is None or len(spheres) == 0:
        return np.zeros(4, np.float32), np.zeros(1, np.uint32), 0
    sp = f32(spheres)
    return sp, np.ascontiguousarray(sphere_surfaces, np.uint32), sp.size // 4


def _materials(lambertian, emissive, metal, dielectric):
    def arr(a, k):
        return f32(a) if a is not None and len(a) else np.zeros(k, np.float32)

    return arr(lambertian, 3), arr(emissive, 3), arr(metal, 4), arr(dielectric, 1)


def render(vertices, surfaces, lambertian, emissive, width, height, spp, max_depth, seed, algorithm, cam20, threads=1,
           node_width=2, leaf_size=0, method=0, metal=None, dielectric=None, spheres=None, sphere_surfaces=None,
           ggx=None):
    """The FP32 wavefront renderer restated (Lambertian, metal f32[4 per material], dielectric f32[1],
    emissive, GGX f32[4 per material]; optional spheres f32[4S] with signed radii and their surface ids):
    (pixels, (extension rays, shadow rays))."""
    v = f32(vertices) if len(vertices) else np.zeros(9, np.float32)
    n = len(vertices) // 9
    surf = np.ascontiguousarray(surfaces, dtype=np.uint32) if n else np.zeros(1, np.uint32)
    lamb, emit, met, die = _materials(lambertian, emissive, metal, dielectric)
    sp, ss, ns = _spheres(spheres, sphere_surfaces)
    g = f32(ggx) if ggx is not None and len(ggx) else np.zeros(4, np.float32)
    st = RenderSettings(width, height, spp, max_depth, seed, algorithm, node_width, leaf_size, method, _p(g))
    pixels = np.empty(3 * width * height, np.float32)
    rays = np.zeros(2, np.uint64)
    rc = lib().orc_render_mixed(
        _p(v), n, _p(surf), _p(sp), ns, _p(ss), _p(lamb), _p(met), _p(die), _p(emit), ctypes.byref(st),
        _p(f32(cam20)), _p(pixels), _p(rays), threads,
    )
    if rc != 0:
        raise RuntimeError(f"orc_render failed ({rc})")
    return pixels, rays


# 96-byte 8-wide node record of orc_wide.c (and of the device library).
WIDE_NODE_DTYPE = np.dtype(
    [
        ("origin", "<f4", 3),
        ("exp", "u1", 3),
        ("count", "u1"),
        ("lo", "u1", (3, 8)),
        ("hi", "u1", (3, 8)),
        ("child", "<u4", 8),
    ]
)
WIDE_LEAF = 0x80000000
WIDE_START_MASK = 0x0FFFFFFF


def grid_exponent(o: float, hi: float) -> int:
    return int(lib().orc_grid_exponent(ctypes.c_float(o), ctypes.c_float(hi)))


def quantize(o: float, e: int, lo: float, hi: float):
    a, b = ctypes.c_uint8(), ctypes.c_uint8()
    lib().orc_quantize(ctypes.c_float(o), e, ctypes.c_float(lo), ctypes.c_float(hi), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def collapse_wide(bvh: dict, leaf_size: int = 4, width: int = 8) -> np.ndarray:
    """W-wide nodes (W = 8 or 4) collapsed from an oracle LBVH (build_lbvh's dict), breadth-first numbering."""
    n = bvh["leaf_ids"].size
    cap = max(n, 1)
    out = np.zeros(cap, WIDE_NODE_DTYPE)
    meta = np.ascontiguousarray(bvh["node_meta"], np.uint32)
    nb = f32(bvh["node_bounds"])
    counts = np.ascontiguousarray(bvh["leaf_counts"], np.uint32)
    k = lib().orc_collapse_wide_n(
        _p(meta) if meta.size else None, _p(nb) if nb.size else None, _p(counts) if counts.size else None, n,
        int(bvh["root"]) if n > 1 else 0, _p(f32(bvh["scene_bounds"])), leaf_size, width,
        out.ctypes.data_as(ctypes.c_void_p), cap,
    )
    if k == 0:
        raise RuntimeError("orc_collapse_wide failed")
    return out[:k].copy()


def trace_wide(vertices, leaf_ids, wide: np.ndarray, rays: np.ndarray, mode: int = 0, with_stats: bool = False):
    """Traverses 8-wide nodes: (hits f32[R, 8], stats u32[R, 5] or None, stack overflows)."""
    v = f32(vertices)
    ids = np.ascontiguousarray(leaf_ids, np.uint32)
    w = np.ascontiguousarray(wide, WIDE_NODE_DTYPE)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    stats = np.empty(5 * R, np.uint32) if with_stats else None
    over = lib().orc_trace_wide(
        _p(v), v.size // 9, _p(ids), w.ctypes.data_as(ctypes.c_void_p), w.size, _p(r), R, mode, _p(hits),
        _p(stats) if stats is not None else None,
    )
    return hits.reshape(R, 8), (stats.reshape(R, 5) if stats is not None else None), int(over)


def canonical_wide(nodes: np.ndarray) -> np.ndarray:
    """Depth-first renumbering (children in slot order) of 8-wide nodes, so two
    builds that differ only in node numbering compare equal."""
    order, stack = [], [0]
    while stack:
        k = stack.pop()
        order.append(k)
        c = nodes[k]["child"][: nodes[k]["count"]]
        inner = [int(x) for x in c if not (int(x) & WIDE_LEAF)]
        stack.extend(reversed(inner))
    pos = {k: i for i, k in enumerate(order)}
    out = nodes[order].copy()
    for i in range(out.size):
        cnt = int(out[i]["count"])
        for j in range(cnt):
            c = int(out[i]["child"][j])
            if not (c & WIDE_LEAF):
                out[i]["child"][j] = pos[c]
    return out


def render_fp64(vertices, surfaces, lambertian, emissive, width, height, spp, max_depth, seed, algorithm, cam20, threads=1,
                metal=None, dielectric=None, spheres=None, sphere_surfaces=None, ggx=None):
    v = f32(vertices) if len(vertices) else np.zeros(9, np.float32)
    n = len(vertices) // 9
    surf = np.ascontiguousarray(surfaces, dtype=np.uint32) if n else np.zeros(1, np.uint32)
    lamb, emit, met, die = _materials(lambertian, emissive, metal, dielectric)
    sp, ss, ns = _spheres(spheres, sphere_surfaces)
    g = f32(ggx) if ggx is not None and len(ggx) else np.zeros(4, np.float32)
    st = RenderSettings(width, height, spp, max_depth, seed, algorithm, 2, 0, 0, _p(g))
    pixels = np.empty(3 * width * height, np.float64)
    lib().orc_render_fp64_mixed(_p(v), n, _p(surf), _p(sp), ns, _p(ss), _p(lamb), _p(met), _p(die), _p(emit),
                                ctypes.byref(st), _p(f32(cam20)), _p(pixels), threads)
    return pixels


# ---- GGX conductor (the library's BAJO_MAT_GGX) -----------------------------------
def eval_ggx(ray_d, n, f0, alpha, direction):
    """FP32 GGX f and pdf toward `direction` (value.xyz, pdf)."""
    out = np.empty(4, np.float32)
    lib().orc_eval_ggx(_p(f32(ray_d)), _p(f32(n)), _p(f32(f0)), float(alpha), _p(f32(direction)), _p(out))
    return out[:3].copy(), float(out[3])


def sample_ggx(ray_d, n, f0, alpha, u1, u2):
    """FP32 visible-normal GGX sample: (direction, weight, pdf, ok)."""
    out = np.empty(8, np.float32)
    lib().orc_sample_ggx(_p(f32(ray_d)), _p(f32(n)), _p(f32(f0)), float(alpha), float(u1), float(u2), _p(out))
    return out[:3].copy(), out[3:6].copy(), float(out[6]), bool(out[7])


# ---- materials (rt/shading.mojo) ----------------------------------------------

def pow_unit(x: float, y: float) -> float:
    return float(lib().orc_pow_unit(x, y))


def reflectance(cosine: float, ref_idx: float) -> float:
    return float(lib().orc_reflectance(cosine, ref_idx))


def reflect(v, n) -> np.ndarray:
    out = np.empty(3, np.float32)
    lib().orc_reflect3(_p(f32(v)), _p(f32(n)), _p(out))
    return out


def sample_metal(ray_d, n, albedo, fuzz, ru, rv) -> dict:
    out = np.empty(9, np.float32)
    lib().orc_sample_metal(_p(f32(ray_d)), _p(f32(n)), _p(f32(albedo)), fuzz, ru, rv, _p(out))
    return {"dir": out[:3], "weight": out[3:6], "pdf": float(out[6]), "delta": bool(out[7]), "ok": bool(out[8])}


def eval_metal(ray_d, n, albedo, fuzz, out_dir) -> dict:
    out = np.empty(5, np.float32)
    lib().orc_eval_metal(_p(f32(ray_d)), _p(f32(n)), _p(f32(albedo)), fuzz, _p(f32(out_dir)), _p(out))
    return {"value": out[:3], "pdf": float(out[3]), "delta": bool(out[4])}


def sample_dielectric(ray_d, n, ior, front, ru) -> dict:
    out = np.empty(9, np.float32)
    lib().orc_sample_dielectric(_p(f32(ray_d)), _p(f32(n)), ior, int(front), ru, _p(out))
    return {"dir": out[:3], "weight": out[3:6], "pdf": float(out[6]), "delta": bool(out[7]), "ok": bool(out[8])}


# ---- H-PLOC ------------------------------------------------------------------

HPLOC_SEARCH_RADIUS = 8
HPLOC_MERGING_THRESHOLD = 16


def hploc_topology(leaf_bounds6, sorted_codes, sorted_ids, search_radius=HPLOC_SEARCH_RADIUS,
                   merging_threshold=HPLOC_MERGING_THRESHOLD) -> dict:
    """bajo's CPU H-PLOC topology: nodes 0..n-1 are leaves (by leaf id), inner
    nodes n..2n-2 in creation order."""
    b = f32(leaf_bounds6)
    codes = np.ascontiguousarray(sorted_codes, np.uint32)
    ids = np.ascontiguousarray(sorted_ids, np.uint32)
    n = codes.size
    m = 2 * n - 1
    out = {k: np.empty(m, np.uint32) for k in ("parent", "left", "right", "leaf_id", "leaf_count")}
    out["bounds"] = np.empty(6 * m, np.float32)
    stats = np.zeros(6, np.uint64)
    root = lib().orc_hploc_topology(
        _p(b), _p(codes), _p(ids), n, search_radius, merging_threshold, _p(out["parent"]), _p(out["left"]),
        _p(out["right"]), _p(out["leaf_id"]), _p(out["leaf_count"]), _p(out["bounds"]), _p(stats),
    )
    if root == SENTINEL:
        raise RuntimeError("orc_hploc_topology failed")
    out["root"] = int(root)
    out["stats"] = dict(zip(
        ("guide_nodes", "merge_calls", "merge_rounds", "hierarchical_rounds", "final_rounds", "max_cluster_count"),
        (int(x) for x in stats),
    ))
    return out


def hploc_quality(t: dict) -> float:
    """Sum of node surface areas over the root's (bajo's HplocTopology.quality)."""
    b = t["bounds"].reshape(-1, 6).astype(np.float64)
    d = b[:, 3:] - b[:, :3]
    area = 2 * (d[:, 0] * d[:, 1] + d[:, 0] * d[:, 2] + d[:, 1] * d[:, 2])
    root = area[t["root"]]
    return float(area.sum() / root) if root > 0 else 0.0


def preorder_topology(t: dict) -> np.ndarray:
    m = t["left"].size
    out = np.empty(m, np.uint32)
    k = lib().orc_preorder_topology(t["root"], _p(t["left"]), _p(t["right"]), _p(t["leaf_id"]), m, _p(out))
    return out[:k]


def preorder_binary(root: int, node_meta: np.ndarray, leaf_ids: np.ndarray) -> np.ndarray:
    meta = np.ascontiguousarray(node_meta, np.uint32)
    ids = np.ascontiguousarray(leaf_ids, np.uint32)
    n = ids.size
    out = np.empty(2 * n, np.uint32)
    k = lib().orc_preorder_binary(root, _p(meta) if meta.size else None, _p(ids), n, _p(out))
    return out[:k]


def hploc_lowest_leaf_numbering(t: dict, sorted_ids: np.ndarray) -> np.ndarray:
    """The device builder's inner-node indices for topology `t`: a cluster is
    named by its lowest sorted leaf position, a merge retires the larger name
    and the new node takes index (retired - 1)."""
    n = (t["left"].size + 1) // 2
    out = np.empty(2 * n - 1, np.uint32)
    ids = np.ascontiguousarray(sorted_ids, np.uint32)
    lib().orc_hploc_lowest_leaf_numbering(_p(t["left"]), _p(t["right"]), _p(ids), n, _p(out))
    return out


def hploc_binary(t: dict, numbering: np.ndarray | None = None) -> dict:
    """The topology in the binary layout with depth-first leaf positions (the
    device H-PLOC builder's output form).  Inner nodes are numbered by
    `numbering` (from hploc_lowest_leaf_numbering: then the layout equals the
    device's word for word) or, when None, in creation order."""
    n = (t["left"].size + 1) // 2
    m = max(n - 1, 1)
    num = None if numbering is None else np.ascontiguousarray(numbering, np.uint32)
    out = {
        "node_meta": np.zeros(4 * m, np.uint32),
        "node_bounds": np.zeros(12 * m, np.float32),
        "leaf_ids": np.zeros(n, np.uint32),
        "leaf_counts": np.zeros(m, np.uint32),
        "leaf_parent": np.zeros(n, np.uint32),
    }
    root = lib().orc_hploc_binary(
        t["root"], _p(t["left"]), _p(t["right"]), _p(t["leaf_id"]), _p(t["leaf_count"]), _p(t["bounds"]), n,
        _p(num) if num is not None else None, _p(out["node_meta"]), _p(out["node_bounds"]), _p(out["leaf_ids"]), _p(out["leaf_counts"]),
        _p(out["leaf_parent"]),
    )
    out["root"] = int(root)
    if n == 1:
        for k in ("node_meta", "node_bounds", "leaf_counts"):
            out[k] = out[k][:0]
    return out


def build_hploc(vertices: np.ndarray, merging_threshold: int = HPLOC_MERGING_THRESHOLD) -> dict:
    """Oracle H-PLOC over a triangle soup: the LBVH's sorted keys, bajo's CPU
    topology at the given merging threshold (16, or 32 as bajo's GPU build on
    wave64), and its depth-first binary layout (usable by trace_bvh2 and
    collapse_wide)."""
    return build_hploc_bounds(triangle_bounds(f32(vertices)), merging_threshold)


def build_hploc_bounds(leaf_bounds: np.ndarray, merging_threshold: int = HPLOC_MERGING_THRESHOLD) -> dict:
    """Oracle H-PLOC over leaf boxes f32[6N]."""
    bounds = f32(leaf_bounds)
    lb = build_lbvh_bounds(bounds)
    codes = (lb["sorted_keys"] >> np.uint64(32)).astype(np.uint32)
    t = hploc_topology(bounds, codes, lb["leaf_ids"], merging_threshold=merging_threshold)
    b = hploc_binary(t, hploc_lowest_leaf_numbering(t, lb["leaf_ids"]))
    b["scene_bounds"] = lb["scene_bounds"]
    b["topology"] = t
    return b


def ao_rays(rays: np.ndarray, hits: np.ndarray, samples: int, seed: int) -> np.ndarray:
    """AO rays of a closest-hit launch (the renderer's AO ray per sample, rt/gpu/path_shading.mojo:650-665);
    field-major f32[8 * R * samples]."""
    r = f32(rays)
    h = f32(hits)
    n = r.size // 8
    out = np.empty(8 * n * samples, np.float32)
    lib().orc_ao_rays(_p(r), _p(h), n, samples, seed, _p(out))
    return out


# ---- instances ------------------------------------------------------------------

class OrcBvh(ctypes.Structure):
    """orc_bvh: one BLAS for orc_tlas_trace (binary layout, or 8-wide when `wide` is set; sphere leaves when
    `spheres` is set)."""

    _fields_ = [
        ("vertices", c_void_p),
        ("leaf_ids", c_void_p),
        ("node_meta", c_void_p),
        ("node_bounds", c_void_p),
        ("root", c_uint32),
        ("wide", c_void_p),
        ("wide_count", c_uint64),
        ("prim_count", c_uint64),
        ("spheres", c_void_p),
    ]


def affine_inverse(m3x4) -> tuple[np.ndarray, bool]:
    """Affine3.inverse of a row-major 3x4: (f32[12], invertible)."""
    a = f32(m3x4).reshape(-1)
    o = np.empty(12, np.float32)
    ok = lib().orc_affine_inverse(_p(a), _p(o))
    return o, bool(ok)


def transform_box(m3x4, box6) -> np.ndarray:
    """AABB.apply_transform of a box (lo.xyz, hi.xyz) by a row-major 3x4."""
    m = f32(m3x4).reshape(-1)
    b = f32(box6).reshape(-1)
    o = np.empty(6, np.float32)
    lib().orc_transform_box(_p(m), _p(b), _p(o))
    return o


def tlas_instance_bounds(transforms: np.ndarray, blas_boxes: np.ndarray, inst_blas: np.ndarray) -> np.ndarray:
    """World boxes f32[6I] of instances: each BLAS root box through its instance transform."""
    xf = f32(transforms).reshape(-1, 12)
    boxes = f32(blas_boxes).reshape(-1, 6)
    return np.concatenate([transform_box(xf[i], boxes[int(b)]) for i, b in enumerate(inst_blas)]).astype(np.float32)


def tlas_trace(blases: list, transforms: np.ndarray, inst_blas: np.ndarray, top: dict, rays: np.ndarray,
               mode: int = 0) -> np.ndarray:
    """Rays through instances of BLASes.  blases: dicts with 'vertices' (or 'spheres' f32[4n]) and the
    binary layout of build_lbvh / build_hploc ('leaf_ids', 'node_meta', 'node_bounds', 'root'), plus
    'wide' (8-wide nodes over the same leaf_ids) to traverse that layout instead; transforms f32[12I] BLAS -> world;
    top: binary layout over the instance boxes whose leaf_ids are instance indices."""
    keep = []
    descs = (OrcBvh * len(blases))()
    for k, b in enumerate(blases):
        sph = b.get("spheres") is not None
        v = f32(b["spheres"] if sph else b["vertices"])
        ids = np.ascontiguousarray(b["leaf_ids"], np.uint32)
        meta = np.ascontiguousarray(b["node_meta"] if b["node_meta"].size else np.zeros(4), np.uint32)
        nb = f32(b["node_bounds"] if b["node_bounds"].size else np.zeros(12))
        keep += [v, ids, meta, nb]
        d = descs[k]
        d.leaf_ids, d.node_meta, d.node_bounds = (x.ctypes.data for x in (ids, meta, nb))
        if sph:
            d.spheres = v.ctypes.data
            d.prim_count = v.size // 4
        else:
            d.vertices = v.ctypes.data
            d.prim_count = v.size // 9
        d.root = int(b["root"])
        if b.get("wide") is not None:
            w = np.ascontiguousarray(b["wide"], WIDE_NODE_DTYPE)
            keep.append(w)
            d.wide = w.ctypes.data
            d.wide_count = w.size
    xf = f32(transforms).reshape(-1, 12)
    inv = np.empty_like(xf)
    for i in range(xf.shape[0]):
        inv[i], ok = affine_inverse(xf[i])
        if not ok:
            raise ValueError(f"transform of instance {i} is not invertible")
    ib = np.ascontiguousarray(inst_blas, np.uint32)
    tmeta = np.ascontiguousarray(top["node_meta"] if top["node_meta"].size else np.zeros(4), np.uint32)
    tnb = f32(top["node_bounds"] if top["node_bounds"].size else np.zeros(12))
    tids = np.ascontiguousarray(top["leaf_ids"], np.uint32)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    inv = np.ascontiguousarray(inv.reshape(-1))
    lib().orc_tlas_trace(ctypes.cast(descs, c_void_p), len(blases), _p(inv), _p(ib), ib.size, _p(tids), _p(tmeta),
                         _p(tnb), int(top["root"]), _p(r), R, int(mode), _p(hits))
    del keep
    return hits


# ---- spheres ----------------------------------------------------------------------

def sphere_bounds(spheres) -> np.ndarray:
    """Sphere.bounds (bvh/types.mojo:122-124) of f32[4N] spheres: f32[6N] boxes."""
    sp = f32(spheres).reshape(-1, 4)
    c, r = sp[:, :3], sp[:, 3:4]
    return np.ascontiguousarray(np.concatenate([c - r, c + r], axis=1).reshape(-1), np.float32)


def trace_spheres(spheres, bvh: dict, rays: np.ndarray, mode: int = 0, stats: bool = False):
    """Near-first / unordered queries over a sphere BVH (binary layout of the box build over sphere_bounds)."""
    sp = f32(spheres)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    meta = bvh["node_meta"] if bvh["node_meta"].size else np.zeros(4, np.uint32)
    nb = bvh["node_bounds"] if bvh["node_bounds"].size else np.zeros(12, np.float32)
    st = np.empty(5 * R, np.uint32) if stats else None
    lib().orc_trace_bvh2_spheres(_p(sp), _p(np.ascontiguousarray(bvh["leaf_ids"], np.uint32)),
                                 _p(np.ascontiguousarray(meta, np.uint32)), _p(f32(nb)), int(bvh["root"]), _p(r), R,
                                 mode, _p(hits), _p(st) if st is not None else None)
    return (hits.reshape(R, 8), st.reshape(R, 5)) if stats else hits.reshape(R, 8)


def trace_spheres_wide(spheres, leaf_ids, wide: np.ndarray, rays: np.ndarray, mode: int = 0, with_stats: bool = False):
    sp = f32(spheres)
    ids = np.ascontiguousarray(leaf_ids, np.uint32)
    w = np.ascontiguousarray(wide, WIDE_NODE_DTYPE)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    stats = np.empty(5 * R, np.uint32) if with_stats else None
    over = lib().orc_trace_wide_spheres(_p(sp), sp.size // 4, _p(ids), w.ctypes.data_as(ctypes.c_void_p), w.size,
                                        _p(r), R, mode, _p(hits), _p(stats) if stats is not None else None)
    return hits.reshape(R, 8), (stats.reshape(R, 5) if stats is not None else None), int(over)


def trace_brute_spheres(spheres, rays: np.ndarray) -> np.ndarray:
    sp = f32(spheres)
    r = f32(rays)
    R = r.size // 8
    hits = np.empty(8 * R, np.float32)
    lib().orc_trace_brute_spheres(_p(sp), sp.size // 4, _p(r), R, _p(hits))
    return hits.reshape(R, 8)


# ---- learned path guiding ------------------------------------------------------------
GUIDE_PARAMS = 24808
GUIDE_OUTPUTS = 40
GUIDE_FEATURES = 216
# offsets of W1, b1, W2, b2, W3, b3, W4, b4 in the parameter vector
GUIDE_OFFSETS = {"W1": 0, "b1": 13824, "W2": 13888, "b2": 17984, "W3": 18048, "b3": 22144, "W4": 22208, "b4": 24768}


def bf16_round(x) -> np.ndarray:
    """FP32 -> BF16 (round to nearest even) -> FP32, elementwise."""
    a = f32(x)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def guide_encode(x9) -> np.ndarray:
    """The 216 features of one sample's 9 inputs (FP32 series, device order)."""
    x = f32(x9)
    out = np.empty(GUIDE_FEATURES, np.float32)
    lib().orc_guide_encode(_p(x), _p(out))
    return out


def guide_forward(params, inputs, emulate_bf16=True) -> np.ndarray:
    """Network outputs f64[n, 40] for inputs f32[n, 9] (FP64 evaluation; emulate_bf16 rounds operands as the
    device does)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = f32(inputs).reshape(-1, 9)
    out = np.empty((x.shape[0], GUIDE_OUTPUTS), np.float64)
    lib().orc_guide_forward(_p(p), _p(x), x.shape[0], int(emulate_bf16), _p(out))
    return out


def guide_loss_grad(params, samples, emulate_bf16=False, with_grad=True):
    """(normalized loss, gradient f64[24808] or None) of samples f32[n, 13] (inputs, direction, weight)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    s = f32(samples).reshape(-1, 13)
    g = np.empty(GUIDE_PARAMS, np.float64) if with_grad else None
    loss = lib().orc_guide_loss_grad(_p(p), _p(s), s.shape[0], int(emulate_bf16), _p(g) if with_grad else None)
    return loss, g


def guide_adam(params, m, v, grad, step, lr=1e-2, beta1=0.9, beta2=0.999, eps=1e-8):
    """One FP64 Adam step in place (step is 1-based)."""
    for a in (params, m, v):
        assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    g = np.ascontiguousarray(grad, dtype=np.float64)
    lib().orc_guide_adam(_p(params), _p(m), _p(v), _p(g), int(step), lr, beta1, beta2, eps)


def guide_mixture_sample(outputs40, u3):
    """(direction f32[3], density) drawn from the mixture of one output row (FP32, device operations)."""
    o = f32(outputs40)
    u = f32(u3)
    d = np.empty(3, np.float32)
    pdf = lib().orc_guide_mixture(_p(o), _p(u), _p(d), None)
    return d, pdf


def guide_mixture_pdf(outputs40, direction) -> float:
    o = f32(outputs40)
    d = f32(direction)
    return float(lib().orc_guide_mixture(_p(o), None, None, _p(d)))


def vmf_pdf64(kappa: float, cos_t: float) -> float:
    return float(lib().orc_vmf_pdf64(kappa, cos_t))


def vmf_sample_cos64(kappa: float, u: float) -> float:
    return float(lib().orc_vmf_sample_cos64(kappa, u))


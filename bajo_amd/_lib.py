"""ctypes binding of ``libbajo_hip.so`` (the C ABI declared in include/bajo_hip.h).

The library is built in-tree by ``python -m bajo_amd.build``.  There is no
fallback: if the shared object is missing every entry point raises.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int32, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbajo_hip.so"

BAJO_OK = 0
BAJO_TRACE_CLOSEST_HIT = 0
BAJO_TRACE_ANY_HIT = 1
BAJO_BUILD_LBVH = 0
BAJO_BUILD_HPLOC = 1
BAJO_WIDTH_BINARY = 2
BAJO_WIDTH_WIDE8 = 8
BAJO_STATUS_SORT_LOOKBACK = 4
BAJO_STATUS_REFIT_CORRUPT = 5
BAJO_STATUS_HPLOC_FAILED = 6
BAJO_STATUS_COLLAPSE_FAILED = 7
BAJO_CAPTURE_REBUILD = 1


class BajoError(RuntimeError):
    """A non-zero bajo status (argument error, HIP error or device status)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"bajo status {status}: {message}")
        self.status = status


class BuildTimings(ctypes.Structure):
    _fields_ = [
        ("bounds_ns", c_uint64),
        ("morton_ns", c_uint64),
        ("sort_ns", c_uint64),
        ("topology_ns", c_uint64),
        ("refit_ns", c_uint64),
        ("collapse_ns", c_uint64),
        ("total_ns", c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BvhInfo(ctypes.Structure):
    _fields_ = [
        ("prim_count", c_uint64),
        ("node_count", c_uint64),
        ("root", c_uint32),
        ("node_width", c_uint32),
        ("nodes", c_void_p),
        ("triangles", c_void_p),
        ("sorted_keys", c_void_p),
        ("leaf_parent", c_void_p),
        ("leaf_counts", c_void_p),
        ("scene_bounds", c_void_p),
        ("wide_node_count", c_uint64),
        ("wide_nodes", c_void_p),
    ]


class BuildOptions(ctypes.Structure):
    _fields_ = [
        ("method", c_int32),
        ("node_width", c_int32),
        ("leaf_size", c_int32),
        ("merging_threshold", c_int32),
    ]


class Materials(ctypes.Structure):
    """bajo_materials: the flattened material arrays of SceneData."""

    _fields_ = [
        ("lambertian", c_void_p),
        ("n_lambertian", c_uint32),
        ("metal", c_void_p),
        ("n_metal", c_uint32),
        ("dielectric", c_void_p),
        ("n_dielectric", c_uint32),
        ("emissive", c_void_p),
        ("n_emissive", c_uint32),
        ("ggx", c_void_p),
        ("n_ggx", c_uint32),
    ]


class RenderSettings(ctypes.Structure):
    _fields_ = [
        ("image_width", c_uint32),
        ("image_height", c_uint32),
        ("samples_per_pixel", c_uint32),
        ("max_depth", c_uint32),
        ("rng_seed", c_uint64),
        ("algorithm", c_int32),
        ("path_capacity", c_uint32),
        ("extend_waves", c_uint32),
        ("tail_paths", c_uint32),
        ("shadow_walk", c_uint32),
        ("sparse_shadow_rays", c_uint32),
        ("shadow_group", c_uint32),
        ("tail_group", c_uint32),
    ]


class RenderStats(ctypes.Structure):
    _fields_ = [
        ("extension_rays", c_uint64),
        ("shadow_rays", c_uint64),
        ("incoherent_rays", c_uint64),
        ("rays_per_bounce", c_uint64 * 16),
        ("status", c_uint32),
        ("traversal_status", c_uint32),
        ("extension_node_visits", c_uint64),
        ("extension_leaf_tests", c_uint64),
        ("shadow_node_visits", c_uint64),
        ("shadow_leaf_tests", c_uint64),
        ("total_ms", c_float),
        ("extend_ms", c_float),
        ("shadow_ms", c_float),
        ("extend_launches", c_uint32),
        ("shadow_launches", c_uint32),
        ("extend_bounce_ms", c_float * 16),
        ("shadow_bounce_ms", c_float * 16),
        ("max_extension_visits", c_uint32 * 16),
        ("max_shadow_visits", c_uint32 * 16),
    ]


class GuideOptions(ctypes.Structure):
    """bajo_guide_options."""

    _fields_ = [
        ("learning_rate", c_float),
        ("beta1", c_float),
        ("beta2", c_float),
        ("epsilon", c_float),
        ("batch_size", c_uint32),
        ("steps_per_frame", c_uint32),
        ("seed", c_uint64),
    ]


class GuideFrameStats(ctypes.Structure):
    """bajo_guide_frame_stats."""

    _fields_ = [
        ("guided_vertices", c_uint64),
        ("training_samples", c_uint64),
        ("adam_steps", c_uint32),
        ("loss", c_float),
        ("infer_ns", c_uint64),
        ("train_ns", c_uint64),
        ("render_ns", c_uint64),
    ]


# name -> (restype, argtypes); the exported surface of include/bajo_hip.h
SIGNATURES = {
    "bajo_ctx_create": (c_int32, [c_int32, c_void_p, POINTER(c_void_p)]),
    "bajo_ctx_destroy": (c_int32, [c_void_p]),
    "bajo_ctx_synchronize": (c_int32, [c_void_p]),
    "bajo_ctx_status": (c_int32, [c_void_p, POINTER(c_uint32)]),
    "bajo_ctx_set_spin_limit": (c_int32, [c_void_p, c_uint32]),
    "bajo_last_error": (c_char_p, []),
    "bajo_status_string": (c_char_p, [c_int32]),
    "bajo_device_name": (c_int32, [c_int32, c_char_p, c_int32]),
    "bajo_sort_workspace_bytes": (c_uint64, [c_uint64, c_int32, c_int32]),
    "bajo_device_radix_sort_pairs_u64_u32": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p, c_uint64]),
    "bajo_device_radix_sort_pairs_u32_u32": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p, c_uint64]),
    "bajo_device_radix_sort_keys_u64": (c_int32, [c_void_p, c_void_p, c_uint64, c_void_p, c_uint64]),
    "bajo_device_radix_sort_keys_u32": (c_int32, [c_void_p, c_void_p, c_uint64, c_void_p, c_uint64]),
    "bajo_onesweep_radix_sort_keys_u64": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, c_int32, c_int32, c_void_p, c_uint64],
    ),
    "bajo_onesweep_radix_sort_pairs_u64_u32": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_int32, c_int32, c_void_p, c_uint64],
    ),
    "bajo_onesweep_radix_sort_pairs_u64_u32_noalloc": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_uint64, c_int32, c_int32],
    ),
    "bajo_guide_create": (c_int32, [c_void_p, POINTER(GuideOptions), POINTER(c_void_p)]),
    "bajo_guide_free": (c_int32, [c_void_p]),
    "bajo_guide_get_params": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "bajo_guide_set_params": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "bajo_guide_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p]),
    "bajo_guide_gradient": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_void_p, POINTER(c_float)]),
    "bajo_guide_train": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64]),
    "bajo_guide_training_samples": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, POINTER(c_uint64)]),
    "bajo_guide_mixture_sample": (c_int32, [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_void_p]),
    "bajo_render_guided": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, POINTER(RenderSettings), POINTER(c_float), c_void_p, c_int32,
         POINTER(GuideFrameStats)],
    ),
    "bajo_build_triangle_bvh": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, c_int32, POINTER(BuildTimings), POINTER(c_void_p)],
    ),
    "bajo_build_triangle_bvh_ex": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, POINTER(BuildOptions), POINTER(BuildTimings), POINTER(c_void_p)],
    ),
    "bajo_build_binary_bounds": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, POINTER(BuildOptions), POINTER(BuildTimings), POINTER(c_void_p)],
    ),
    "bajo_build_sphere_bvh": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, POINTER(BuildOptions), POINTER(BuildTimings), POINTER(c_void_p)],
    ),
    "bajo_bvh_free": (c_int32, [c_void_p]),
    "bajo_bvh_get_info": (c_int32, [c_void_p, POINTER(BvhInfo)]),
    "bajo_bvh_copy_binary": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "bajo_bvh_copy_wide": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, POINTER(c_uint64)]),
    "bajo_trace_rays": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_uint64, c_void_p]),
    "bajo_ao_rays": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint64, c_void_p]),
    "bajo_trace_camera": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32, c_void_p],
    ),
    "bajo_build_triangle_blas_set": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_uint32, POINTER(BuildOptions), POINTER(BuildTimings), POINTER(c_void_p)],
    ),
    "bajo_blas_set_count": (c_int32, [c_void_p, POINTER(c_uint32)]),
    "bajo_blas_set_get": (c_int32, [c_void_p, c_uint32, POINTER(c_void_p)]),
    "bajo_blas_set_free": (c_int32, [c_void_p]),
    "bajo_blas_set_trace_camera": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p]),
    "bajo_tlas_create": (
        c_int32, [c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, c_uint32, POINTER(BuildOptions), POINTER(c_void_p)],
    ),
    "bajo_tlas_trace_rays": (c_int32, [c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_uint64]),
    "bajo_tlas_trace_camera": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_uint64, c_uint32, c_uint32]),
    "bajo_tlas_get_top": (c_int32, [c_void_p, POINTER(c_void_p)]),
    "bajo_tlas_copy_instance_bounds": (c_int32, [c_void_p, c_void_p]),
    "bajo_tlas_status": (c_int32, [c_void_p, c_void_p, POINTER(c_uint32)]),
    "bajo_tlas_free": (c_int32, [c_void_p]),
    "bajo_camera_make": (
        c_int32,
        [POINTER(c_float), POINTER(c_float), POINTER(c_float), c_float, c_float, POINTER(c_float)],
    ),
    "bajo_scene_create": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_uint32, c_void_p, c_uint32, POINTER(c_void_p)],
    ),
    "bajo_scene_create_materials": (
        c_int32, [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, POINTER(BuildOptions), POINTER(c_void_p)],
    ),
    "bajo_scene_create_batch": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p, POINTER(BuildOptions), POINTER(c_void_p)],
    ),
    "bajo_scene_create_mixed": (
        c_int32,
        [c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, POINTER(BuildOptions),
         POINTER(c_void_p)],
    ),
    "bajo_scene_create_ex": (
        c_int32,
        [
            c_void_p, c_void_p, c_uint64, c_void_p, c_void_p, c_uint32, c_void_p, c_uint32,
            POINTER(BuildOptions), POINTER(c_void_p),
        ],
    ),
    "bajo_scene_rebuild": (c_int32, [c_void_p, c_void_p, POINTER(BuildTimings)]),
    "bajo_scene_free": (c_int32, [c_void_p]),
    "bajo_scene_get_bvh": (c_int32, [c_void_p, POINTER(c_void_p)]),
    "bajo_scene_light_table": (c_int32, [c_void_p, c_void_p, c_void_p, POINTER(c_uint32), POINTER(c_float)]),
    "bajo_render": (
        c_int32,
        [c_void_p, c_void_p, POINTER(RenderSettings), POINTER(c_float), c_void_p, POINTER(RenderStats)],
    ),
    "bajo_render_target_create": (
        c_int32, [c_void_p, POINTER(RenderSettings), POINTER(c_float), c_void_p, POINTER(c_void_p)],
    ),
    "bajo_render_cameras": (
        c_int32, [c_void_p, c_void_p, POINTER(RenderSettings), c_void_p, c_uint32, c_void_p, POINTER(RenderStats)],
    ),
    "bajo_render_target_create_envs": (
        c_int32, [c_void_p, POINTER(RenderSettings), c_uint32, POINTER(c_float), c_void_p, POINTER(c_void_p)],
    ),
    "bajo_render_target_set_camera": (c_int32, [c_void_p, c_void_p, POINTER(c_float)]),
    "bajo_render_enqueue": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "bajo_render_target_capture": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "bajo_render_target_capture_ex": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint32]),
    "bajo_scene_vertices": (c_int32, [c_void_p, POINTER(c_void_p), POINTER(c_uint64)]),
    "bajo_scene_set_vertices": (c_int32, [c_void_p, c_void_p, c_void_p, c_uint64]),
    "bajo_render_target_replay": (c_int32, [c_void_p, c_void_p]),
    "bajo_render_target_free": (c_int32, [c_void_p]),
    "bajo_rng_stage_floats": (c_int32, [c_uint64, c_uint32, c_uint32, c_uint32, c_void_p]),
    "bajo_scene_soup": (c_int32, [c_uint64, c_uint64, c_uint64, c_float, c_float, c_void_p]),
    "bajo_scene_icosphere": (c_int32, [c_uint32, c_uint64, c_uint64, c_float, c_float, c_void_p]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m bajo_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH), mode=ctypes.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int) -> None:
    if status != BAJO_OK:
        msg = lib().bajo_last_error().decode(errors="replace")
        if not msg:
            msg = lib().bajo_status_string(status).decode()
        raise BajoError(status, msg)

"""Triangle BVH build and traversal with bajo's ``bajo/bvh/gpu`` interface.

Mirrors ``build_triangle_bvh`` / ``build_triangle_bvh_measured``
(bvh/gpu/triangle_bvh.mojo:617-646) and ``GpuTriangleBvh.launch_rays`` /
``launch_camera`` / ``launch_camera_instrumented`` (:520-614), plus bajo's
``Camera`` record (bvh/camera.mojo:19-231).  Buffers follow bajo's ABIs:
vertices f32[9N], rays f32[8R] field-major, hits f32[8R] AoS, camera
f32[20 * views], traversal stats u32[5R].
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .context import DeviceContext, default_context, ptr, require_device

EMPTY_LANE = 0xFFFFFFFF
LBVH_SENTINEL = 0xFFFFFFFF
RAY_STRIDE = 8
HIT_STRIDE = 8
CAMERA_STRIDE = 20
STATS_STRIDE = 5
F32_MAX = float(np.finfo(np.float32).max)


class TRACE:
    """bvh/constants.mojo:9-13."""

    CLOSEST_HIT = _lib.BAJO_TRACE_CLOSEST_HIT
    ANY_HIT = _lib.BAJO_TRACE_ANY_HIT


class GpuBvhBuildMethod:
    """bvh/gpu/builder GpuBvhBuildMethod."""

    LBVH = _lib.BAJO_BUILD_LBVH
    HPLOC = _lib.BAJO_BUILD_HPLOC


# One 8-wide node (96 bytes): child i spans origin + [lo, hi] * 2^(exp - 127)
# per axis; child bit 31 = leaf of ((child >> 28) & 7) + 1 sorted triangles
# starting at child & 0x0FFFFFFF, else a wide node index.
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
assert WIDE_NODE_DTYPE.itemsize == 96


def build_options(node_width: int = 2, leaf_size: int = 0, method: int = GpuBvhBuildMethod.LBVH,
                  merging_threshold: int = 0):
    """H-PLOC's merging threshold: 0 / 16 = bajo's CPU definition, 32 = its GPU build on wave64."""
    return _lib.BuildOptions(int(method), int(node_width), int(leaf_size), int(merging_threshold))


@dataclass
class Camera:
    """bajo's pinhole camera (bvh/camera.mojo:86-106), flattened to 20 floats."""

    origin: tuple
    target: tuple
    world_up: tuple = (0.0, 1.0, 0.0)
    fov_scale: float = 0.75
    focus_dist: float = 1.0

    def flatten(self) -> np.ndarray:
        f3 = ctypes.c_float * 3
        out = (ctypes.c_float * CAMERA_STRIDE)()
        _lib.check(
            _lib.lib().bajo_camera_make(
                f3(*self.origin), f3(*self.target), f3(*self.world_up), self.fov_scale, self.focus_dist, out
            )
        )
        return np.frombuffer(out, dtype=np.float32).copy()


class GpuTriangleBvh:
    """Ready-to-trace triangle BVH (bvh/gpu/triangle_bvh.mojo:494-614)."""

    merging_threshold = 0  # H-PLOC merging threshold of its builds (0 = 16)

    def __init__(self, handle: ctypes.c_void_p, ctx: DeviceContext, prim_count: int, node_width: int = 2,
                 leaf_size: int = 0, method: int = GpuBvhBuildMethod.LBVH):
        self._handle = handle
        self.ctx = ctx
        self.prim_count = prim_count
        self.node_width = node_width
        self.leaf_size = leaf_size
        self.method = method

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._handle:
            raise RuntimeError("BVH already freed")
        return self._handle

    def info(self) -> _lib.BvhInfo:
        out = _lib.BvhInfo()
        _lib.check(_lib.lib().bajo_bvh_get_info(self.handle, ctypes.byref(out)))
        return out

    def rebuild(self, vertices: torch.Tensor, timings: bool = False):
        """Rebuilds in place over new vertices, reusing device memory when it fits."""
        n = _check_vertices(vertices, self.ctx)
        t = _lib.BuildTimings() if timings else None
        handle = ctypes.c_void_p(self.handle.value)
        opts = build_options(self.node_width, self.leaf_size, self.method, self.merging_threshold)
        _lib.check(
            _lib.lib().bajo_build_triangle_bvh_ex(
                self.ctx.handle, ptr(vertices), n, ctypes.byref(opts), ctypes.byref(t) if t else None,
                ctypes.byref(handle),
            )
        )
        self._handle = handle
        self.prim_count = n
        return t.as_dict() if t else None

    def launch_rays(self, rays, hits, ray_count=None, mode=TRACE.CLOSEST_HIT, stats=None) -> None:
        """launch_rays[mode] (bvh/gpu/triangle_bvh.mojo:587-614)."""
        r = rays.numel() // RAY_STRIDE if ray_count is None else int(ray_count)
        for name, buf, stride in (("rays", rays, RAY_STRIDE), ("hits", hits, HIT_STRIDE)):
            require_device(buf, self.ctx, name)
            if buf.numel() < r * stride:
                raise ValueError(f"{name} buffer is too short for {r} rays")
        if rays.numel() != r * RAY_STRIDE:
            raise ValueError("rays must hold exactly 8 field-major planes of ray_count entries")
        if stats is not None and stats.numel() < r * STATS_STRIDE:
            raise ValueError("stats buffer is too short")
        _lib.check(
            _lib.lib().bajo_trace_rays(self.ctx.handle, self.handle, int(mode), ptr(rays), ptr(hits), r, ptr(stats))
        )

    def launch_camera(self, camera_params, hits, ray_count, width, height, stats=None) -> None:
        """launch_camera (bvh/gpu/triangle_bvh.mojo:520-548)."""
        require_device(camera_params, self.ctx, "camera_params")
        require_device(hits, self.ctx, "hits")
        views = -(-int(ray_count) // (int(width) * int(height)))
        if camera_params.numel() < views * CAMERA_STRIDE:
            raise ValueError("camera parameter buffer is too short")
        if hits.numel() < int(ray_count) * HIT_STRIDE:
            raise ValueError("hit output buffer is too short")
        if stats is not None and stats.numel() < int(ray_count) * STATS_STRIDE:
            raise ValueError("traversal stats output buffer is too short")
        _lib.check(
            _lib.lib().bajo_trace_camera(
                self.ctx.handle, self.handle, ptr(camera_params), ptr(hits), int(ray_count), int(width), int(height),
                ptr(stats),
            )
        )

    def launch_camera_instrumented(self, camera_params, hits, stats, ray_count, width, height) -> None:
        """launch_camera_instrumented (bvh/gpu/triangle_bvh.mojo:550-585)."""
        self.launch_camera(camera_params, hits, ray_count, width, height, stats=stats)

    def binary(self) -> dict:
        """Host copy of the binary LBVH in bajo's GpuBinaryBoundsBvh layout."""
        n = self.prim_count
        m = max(n - 1, 0)
        out = {
            "sorted_keys": np.empty(n, np.uint64),
            "leaf_ids": np.empty(n, np.uint32),
            "node_meta": np.empty(4 * m, np.uint32),
            "node_bounds": np.empty(12 * m, np.float32),
            "leaf_parent": np.empty(n, np.uint32),
            "leaf_counts": np.empty(m, np.uint32),
            "scene_bounds": np.empty(12, np.float32),
        }
        p = {k: v.ctypes.data_as(ctypes.c_void_p) for k, v in out.items()}
        _lib.check(
            _lib.lib().bajo_bvh_copy_binary(
                self.ctx.handle, self.handle, p["sorted_keys"], p["leaf_ids"], p["node_meta"], p["node_bounds"],
                p["leaf_parent"], p["leaf_counts"], p["scene_bounds"],
            )
        )
        out["root"] = self.info().root
        return out

    def wide(self) -> np.ndarray:
        """Host copy of the 8-wide nodes (WIDE_NODE_DTYPE records, root = record 0)."""
        count = ctypes.c_uint64()
        _lib.check(_lib.lib().bajo_bvh_copy_wide(self.ctx.handle, self.handle, None, 0, ctypes.byref(count)))
        out = np.empty(count.value, WIDE_NODE_DTYPE)
        _lib.check(
            _lib.lib().bajo_bvh_copy_wide(
                self.ctx.handle, self.handle, out.ctypes.data_as(ctypes.c_void_p), count.value, ctypes.byref(count)
            )
        )
        return out

    def free(self) -> None:
        if self._handle:
            if getattr(self, "_owner", None) is None:
                _lib.lib().bajo_bvh_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuSphereBvh(GpuTriangleBvh):
    """GpuSphereBvh (bvh/gpu/sphere_bvh.mojo:274-318): the box builder over Sphere.bounds with each sorted
    leaf's sphere packed for traversal.  Launches trace bajo's ray-sphere test (hit u = v = 0,
    normal = normalize(p - center)); `binary()` / `wide()` export the hierarchy like GpuBinaryBoundsBvh."""

    def rebuild(self, spheres: torch.Tensor, timings: bool = False):
        n = _check_spheres(spheres, self.ctx)
        t = _lib.BuildTimings() if timings else None
        handle = ctypes.c_void_p(self.handle.value)
        opts = build_options(self.node_width, self.leaf_size, self.method, self.merging_threshold)
        _lib.check(_lib.lib().bajo_build_sphere_bvh(self.ctx.handle, ptr(spheres), n, ctypes.byref(opts),
                                                    ctypes.byref(t) if t else None, ctypes.byref(handle)))
        self._handle = handle
        self.prim_count = n
        return t.as_dict() if t else None


def _check_spheres(spheres: torch.Tensor, ctx: DeviceContext) -> int:
    require_device(spheres, ctx, "spheres")
    if spheres.dtype != torch.float32:
        raise ValueError("spheres must be float32")
    if spheres.numel() % 4 != 0 or spheres.numel() == 0:
        raise ValueError("spheres hold four floats (center.xyz, radius) per sphere")
    return spheres.numel() // 4


def build_sphere_bvh(spheres: torch.Tensor, ctx: DeviceContext | None = None, node_width: int = 2, leaf_size: int = 0,
                     method: int = GpuBvhBuildMethod.LBVH, timings: dict | None = None) -> GpuSphereBvh:
    """build_sphere_bvh[WORLD, W, L] (bvh/gpu/sphere_bvh.mojo:320-371) over spheres f32[4N] on the device."""
    ctx = ctx or default_context(spheres.device.index or 0)
    n = _check_spheres(spheres, ctx)
    t = _lib.BuildTimings() if timings is not None else None
    handle = ctypes.c_void_p()
    opts = build_options(node_width, leaf_size, method)
    _lib.check(_lib.lib().bajo_build_sphere_bvh(ctx.handle, ptr(spheres), n, ctypes.byref(opts),
                                                ctypes.byref(t) if t else None, ctypes.byref(handle)))
    if t is not None:
        timings.update(t.as_dict())
    return GpuSphereBvh(handle, ctx, n, node_width, leaf_size, method)


class GpuTriangleBlasSet:
    """GpuBlasSet of triangles (bvh/gpu/blas_storage.mojo:21-120) made by build_triangle_blas_set
    (bvh/gpu/triangle_bvh.mojo:469-491): several meshes built in one segmented pass.  `blas(k)` is mesh k
    as a ready-to-trace BVH that shares the set's storage (bajo's BlasDesc); it stays valid until the set
    is freed."""

    merging_threshold = 0  # H-PLOC merging threshold of its builds (0 = 16)

    def __init__(self, handle: ctypes.c_void_p, ctx: DeviceContext, counts: list, node_width: int, leaf_size: int,
                 method: int = 0):
        self._handle = handle
        self.ctx = ctx
        self.counts = list(counts)
        self.node_width = node_width
        self.leaf_size = leaf_size
        self.method = method

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._handle:
            raise RuntimeError("BLAS set already freed")
        return self._handle

    @property
    def blas_count(self) -> int:
        return len(self.counts)

    def blas(self, k: int) -> GpuTriangleBvh:
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bajo_blas_set_get(self.handle, int(k), ctypes.byref(h)))
        view = GpuTriangleBvh(h, self.ctx, self.counts[k], self.node_width, self.leaf_size, self.method)
        view.merging_threshold = self.merging_threshold
        view._owner = self
        return view

    def blases(self) -> list:
        return [self.blas(k) for k in range(self.blas_count)]

    def rebuild(self, vertices: torch.Tensor, offsets: np.ndarray, timings: bool = False):
        """Rebuilds in place (same mesh count) over meshes stored back to back; views stay valid."""
        if len(offsets) - 1 != self.blas_count:
            raise ValueError("a rebuild keeps the mesh count")
        t = {} if timings else None
        _build_set(vertices, offsets, self.ctx, self.node_width, self.leaf_size, t, ctypes.c_void_p(self.handle.value),
                   self.method, self.merging_threshold)
        self.counts = np.diff(np.asarray(offsets, np.int64)).tolist()
        return t

    def launch_camera(self, cameras, hits, width, height, stats=None) -> None:
        """One camera launch over every mesh: view e (cameras f32[20 S]) traces mesh e; hits f32[8 S w h]."""
        require_device(cameras, self.ctx, "cameras")
        require_device(hits, self.ctx, "hits")
        R = self.blas_count * int(width) * int(height)
        if cameras.numel() < self.blas_count * CAMERA_STRIDE:
            raise ValueError("one 20-float camera per mesh is needed")
        if hits.numel() < R * HIT_STRIDE or (stats is not None and stats.numel() < R * STATS_STRIDE):
            raise ValueError("hit / stats output buffer is too short")
        _lib.check(_lib.lib().bajo_blas_set_trace_camera(self.ctx.handle, self.handle, ptr(cameras), ptr(hits),
                                                         int(width), int(height), ptr(stats)))

    def free(self) -> None:
        if self._handle:
            _lib.lib().bajo_blas_set_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _build_set(flat: torch.Tensor, offsets: np.ndarray, ctx: DeviceContext, node_width: int, leaf_size: int,
               timings: dict | None, handle: ctypes.c_void_p, method: int = 0,
               merging_threshold: int = 0) -> ctypes.c_void_p:
    _check_vertices(flat, ctx)
    off = np.ascontiguousarray(offsets, np.uint64)
    if off.size < 2 or int(off[-1]) * 9 != flat.numel():
        raise ValueError("segment offsets must end at the triangle count of the vertex buffer")
    t = _lib.BuildTimings() if timings is not None else None
    opts = build_options(node_width, leaf_size, method, merging_threshold)
    _lib.check(_lib.lib().bajo_build_triangle_blas_set(
        ctx.handle, ptr(flat), off.ctypes.data_as(ctypes.c_void_p), off.size - 1, ctypes.byref(opts),
        ctypes.byref(t) if t else None, ctypes.byref(handle)))
    if t is not None:
        timings.update(t.as_dict())
    return handle


def build_triangle_blas_set(vertex_sets: list, ctx: DeviceContext | None = None, node_width: int = 2,
                            leaf_size: int = 0, timings: dict | None = None,
                            method: int = GpuBvhBuildMethod.LBVH) -> GpuTriangleBlasSet:
    """build_triangle_blas_set[W, L, METHOD] (bvh/gpu/triangle_bvh.mojo:469-491): one segmented build (LBVH or
    H-PLOC) over the meshes `vertex_sets` (float32 tensors of 9 floats per triangle on one device); node_width
    2 or 8."""
    if not vertex_sets:
        raise ValueError("a BLAS set needs at least one mesh")
    ctx = ctx or default_context(vertex_sets[0].device.index or 0)
    counts = [_check_vertices(v, ctx) for v in vertex_sets]
    flat = torch.cat([v.reshape(-1) for v in vertex_sets]).contiguous()
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    return build_triangle_blas_set_flat(flat, offsets, ctx, node_width, leaf_size, timings, method)


def build_triangle_blas_set_flat(vertices: torch.Tensor, offsets: np.ndarray, ctx: DeviceContext | None = None,
                                 node_width: int = 2, leaf_size: int = 0,
                                 timings: dict | None = None,
                                 method: int = GpuBvhBuildMethod.LBVH,
                                 merging_threshold: int = 0) -> GpuTriangleBlasSet:
    """The same build over meshes stored back to back in `vertices`, mesh k = triangles
    [offsets[k], offsets[k + 1])."""
    ctx = ctx or default_context(vertices.device.index or 0)
    handle = _build_set(vertices, offsets, ctx, node_width, leaf_size, timings, ctypes.c_void_p(), method,
                        merging_threshold)
    counts = np.diff(np.asarray(offsets, np.int64)).tolist()
    out = GpuTriangleBlasSet(handle, ctx, counts, node_width, leaf_size, method)
    out.merging_threshold = merging_threshold
    return out


@dataclass
class Instance:
    """bajo's Instance (bvh/types.mojo:153-190): a BLAS placed in world space by an affine `transform`
    (3x4 or 4x4, row-major, BLAS -> world)."""

    transform: np.ndarray
    blas_idx: int

    def matrix3x4(self) -> np.ndarray:
        m = np.asarray(self.transform, np.float32)
        return np.ascontiguousarray(m.reshape(-1, 4)[:3], np.float32)


class GpuTriangleTlas:
    """build_triangle_tlas / GpuTriangleTlas (bvh/gpu/tlas.mojo:482-645): instances of triangle BVHs under a
    top-level hierarchy.  The BLAS objects are kept alive by this object."""

    def __init__(self, handle: ctypes.c_void_p, ctx: DeviceContext, blases: list, instance_count: int):
        self._handle = handle
        self.ctx = ctx
        self.blases = list(blases)
        self.instance_count = instance_count

    @property
    def handle(self) -> ctypes.c_void_p:
        if not self._handle:
            raise RuntimeError("TLAS already freed")
        return self._handle

    def launch_rays(self, rays, hits, ray_count=None, mode=TRACE.CLOSEST_HIT) -> None:
        r = rays.numel() // RAY_STRIDE if ray_count is None else int(ray_count)
        require_device(rays, self.ctx, "rays")
        require_device(hits, self.ctx, "hits")
        if rays.numel() != r * RAY_STRIDE or hits.numel() < r * HIT_STRIDE:
            raise ValueError("rays must hold 8 planes of ray_count entries and hits 8 floats per ray")
        _lib.check(_lib.lib().bajo_tlas_trace_rays(self.ctx.handle, self.handle, int(mode), ptr(rays), ptr(hits), r))

    def launch_camera(self, camera_params, hits, ray_count, width, height) -> None:
        """GpuTriangleTlas.launch_camera (bvh/gpu/tlas.mojo:499-548)."""
        require_device(camera_params, self.ctx, "camera_params")
        require_device(hits, self.ctx, "hits")
        if hits.numel() < int(ray_count) * HIT_STRIDE:
            raise ValueError("hit output buffer is too short")
        _lib.check(_lib.lib().bajo_tlas_trace_camera(self.ctx.handle, self.handle, ptr(camera_params), ptr(hits),
                                                     int(ray_count), int(width), int(height)))

    def top(self) -> "GpuBinaryBoundsBvh":
        """The top-level boxes hierarchy (owned by the TLAS): its leaf_ids are instance indices."""
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bajo_tlas_get_top(self.handle, ctypes.byref(h)))
        view = GpuBinaryBoundsBvh(h, self.ctx, self.instance_count)
        view._owner = self
        return view

    def instance_bounds(self) -> np.ndarray:
        out = np.empty(6 * self.instance_count, np.float32)
        _lib.check(_lib.lib().bajo_tlas_copy_instance_bounds(self.handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def status(self) -> int:
        st = ctypes.c_uint32()
        _lib.check(_lib.lib().bajo_tlas_status(self.ctx.handle, self.handle, ctypes.byref(st)))
        return st.value

    def free(self) -> None:
        if self._handle:
            _lib.lib().bajo_tlas_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def build_triangle_tlas(blases: list, instances: list, ctx: DeviceContext | None = None,
                        method: int = GpuBvhBuildMethod.LBVH) -> GpuTriangleTlas:
    """build_triangle_tlas (bvh/gpu/tlas.mojo:618-645) over triangle BVHs and Instance records."""
    if not blases or not instances:
        raise ValueError("a TLAS needs BLASes and instances")
    ctx = ctx or blases[0].ctx
    handles = (ctypes.c_void_p * len(blases))(*[b.handle.value for b in blases])
    xf = np.ascontiguousarray(np.stack([inst.matrix3x4() for inst in instances]).reshape(-1), np.float32)
    idx = np.ascontiguousarray([int(inst.blas_idx) for inst in instances], np.uint32)
    opts = build_options(2, 0, method)
    handle = ctypes.c_void_p()
    _lib.check(_lib.lib().bajo_tlas_create(ctx.handle, handles, len(blases), xf.ctypes.data_as(ctypes.c_void_p),
                                           idx.ctypes.data_as(ctypes.c_void_p), len(instances), ctypes.byref(opts),
                                           ctypes.byref(handle)))
    return GpuTriangleTlas(handle, ctx, blases, len(instances))


def _check_vertices(vertices: torch.Tensor, ctx: DeviceContext) -> int:
    require_device(vertices, ctx, "vertices")
    if vertices.dtype != torch.float32:
        raise ValueError("vertices must be float32")
    if vertices.numel() % 9 != 0:
        raise ValueError("triangle vertex buffer must contain complete triangle records")
    n = vertices.numel() // 9
    if n == 0:
        raise ValueError("standalone triangle BVH requires nonempty input")
    return n


def build_triangle_bvh(
    vertices: torch.Tensor, ctx: DeviceContext | None = None, timings: dict | None = None, node_width: int = 2,
    leaf_size: int = 0, method: int = GpuBvhBuildMethod.LBVH, merging_threshold: int = 0,
):
    """build_triangle_bvh[WORLD, W, L, method] (bvh/gpu/triangle_bvh.mojo:617-630).

    method LBVH or HPLOC builds the binary hierarchy; node_width 2 keeps it,
    4 or 8 collapses it into quantized wide nodes whose leaf children hold up
    to `leaf_size` triangles (0 = 4).  H-PLOC merges with `merging_threshold`
    16 (0) or 32."""
    ctx = ctx or default_context(vertices.device.index or 0)
    n = _check_vertices(vertices, ctx)
    t = _lib.BuildTimings() if timings is not None else None
    handle = ctypes.c_void_p()
    opts = build_options(node_width, leaf_size, method, merging_threshold)
    _lib.check(
        _lib.lib().bajo_build_triangle_bvh_ex(
            ctx.handle, ptr(vertices), n, ctypes.byref(opts), ctypes.byref(t) if t else None, ctypes.byref(handle)
        )
    )
    if t is not None:
        timings.update(t.as_dict())
    out = GpuTriangleBvh(handle, ctx, n, node_width, leaf_size, method)
    out.merging_threshold = merging_threshold
    return out


class GpuBinaryBoundsBvh(GpuTriangleBvh):
    """A binary (or collapsed 8-wide) hierarchy over caller boxes: bajo's GpuBinaryBoundsBvh
    (bvh/gpu/builder/binary_layout.mojo:303-448).  `binary()` exports bajo's layout; its leaf_ids index the
    input boxes, through which bajo's payloads are looked up."""

    def leaf_payloads(self, payloads: np.ndarray) -> np.ndarray:
        """The payload of each sorted leaf."""
        return np.asarray(payloads)[self.binary()["leaf_ids"]]

    def rebuild(self, leaf_bounds: torch.Tensor, timings: bool = False):
        n = _check_bounds(leaf_bounds, self.ctx)
        t = _lib.BuildTimings() if timings else None
        handle = ctypes.c_void_p(self.handle.value)
        opts = build_options(self.node_width, self.leaf_size, self.method, self.merging_threshold)
        _lib.check(_lib.lib().bajo_build_binary_bounds(self.ctx.handle, ptr(leaf_bounds), n, ctypes.byref(opts),
                                                       ctypes.byref(t) if t else None, ctypes.byref(handle)))
        self._handle = handle
        self.prim_count = n
        return t.as_dict() if t else None


def _check_bounds(leaf_bounds: torch.Tensor, ctx: DeviceContext) -> int:
    require_device(leaf_bounds, ctx, "leaf_bounds")
    if leaf_bounds.dtype != torch.float32:
        raise ValueError("leaf bounds must be float32")
    if leaf_bounds.numel() % 6 != 0 or leaf_bounds.numel() == 0:
        raise ValueError("leaf bounds hold six floats (lo.xyz, hi.xyz) per leaf")
    return leaf_bounds.numel() // 6


def build_binary_bvh(leaf_bounds: torch.Tensor, ctx: DeviceContext | None = None, method: int = GpuBvhBuildMethod.LBVH,
                     timings: dict | None = None, node_width: int = 2, leaf_size: int = 0) -> GpuBinaryBoundsBvh:
    """build_binary_bvh[method] over GpuBinaryBoundsBvh(ctx, leaf_bounds, ...)
    (bvh/gpu/builder/binary_builder.mojo:21-45): the LBVH or H-PLOC builder over boxes f32[6N]."""
    ctx = ctx or default_context(leaf_bounds.device.index or 0)
    n = _check_bounds(leaf_bounds, ctx)
    t = _lib.BuildTimings() if timings is not None else None
    handle = ctypes.c_void_p()
    opts = build_options(node_width, leaf_size, method)
    _lib.check(_lib.lib().bajo_build_binary_bounds(ctx.handle, ptr(leaf_bounds), n, ctypes.byref(opts),
                                                   ctypes.byref(t) if t else None, ctypes.byref(handle)))
    if t is not None:
        timings.update(t.as_dict())
    return GpuBinaryBoundsBvh(handle, ctx, n, node_width, leaf_size, method)


def build_triangle_bvh_measured(vertices: torch.Tensor, ctx: DeviceContext | None = None, node_width: int = 2,
                                leaf_size: int = 0, method: int = GpuBvhBuildMethod.LBVH):
    """build_triangle_bvh_measured (bvh/gpu/triangle_bvh.mojo:633-646): returns (bvh, timings)."""
    t: dict = {}
    bvh = build_triangle_bvh(vertices, ctx=ctx, timings=t, node_width=node_width, leaf_size=leaf_size, method=method)
    return bvh, t


def make_ao_rays(rays: torch.Tensor, hits: torch.Tensor, ray_count: int, samples: int, seed: int,
                 out: torch.Tensor | None = None, ctx: DeviceContext | None = None) -> torch.Tensor:
    """AO rays of a closest-hit launch (bajo's _make_ao_ray, rt/gpu/path_shading.mojo:650-665): `samples`
    hemisphere rays per primary ray, field-major, for launch_rays(mode=TRACE.ANY_HIT).  Rays of primary
    rays without a hit get t_max = -1."""
    ctx = ctx or default_context(rays.device.index or 0)
    r, s = int(ray_count), int(samples)
    require_device(rays, ctx, "rays")
    require_device(hits, ctx, "hits")
    if rays.numel() != r * RAY_STRIDE or hits.numel() < r * HIT_STRIDE:
        raise ValueError("rays must hold 8 planes of ray_count entries and hits 8 floats per ray")
    if out is None:
        out = torch.empty(r * s * RAY_STRIDE, dtype=torch.float32, device=rays.device)
    require_device(out, ctx, "out")
    if out.numel() != r * s * RAY_STRIDE:
        raise ValueError("AO ray buffer must hold 8 planes of ray_count * samples entries")
    _lib.check(_lib.lib().bajo_ao_rays(ctx.handle, ptr(rays), ptr(hits), r, s, int(seed), ptr(out)))
    return out

"""Wavefront path tracing with bajo's ``bajo/rt`` interface.

Mirrors ``render_gpu[ALGORITHM](settings, camera, world) -> RenderResult``
(rt/gpu/render.mojo:322-420), ``RenderSettings`` (rt/types.mojo:324-357),
the ``RENDER`` / ``MAT`` enums (rt/types.mojo:24-57) and ``SceneData`` for
scenes of triangles and spheres with Lambertian, metal, dielectric and
emissive surfaces (rt/types.mojo:382-526).
The renderer runs entirely in libbajo_hip.so (bajo_amd/csrc/render.hip).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .context import DeviceContext, default_context, ptr, require_device


class RENDER:
    PATH = 0
    NORMALS = 1
    AO = 2
    NEE = 3
    MIS = 4


class MAT:
    LAMBERTIAN = 0
    METAL = 1
    DIELECTRIC = 2
    EMISSIVE = 3
    GGX = 4  # this library's GGX conductor (F0.rgb, roughness); bajo's MAT stops at EMISSIVE


SURFACE_KIND_SHIFT = 28


def surface_id(kind: int, index: int) -> int:
    """SurfaceId packing: kind in the top 4 bits, material index below."""
    return (int(kind) << SURFACE_KIND_SHIFT) | int(index)


@dataclass
class RenderSettings:
    image_width: int
    image_height: int
    samples_per_pixel: int
    rng_seed: int
    max_depth: int = 8
    path_capacity: int = 0  # paths in flight per chunk; 0 = the whole frame
    extend_waves: int = 0  # occupancy of the bounce >= 1 extension kernel: 4, 6 or 8 waves per SIMD; 0 = 6
    tail_paths: int = 0  # live paths at which the tail kernel takes over; 0 = its resident threads
    shadow_walk: int = 0  # shadow any-hit walk: 1 one node, 2 two nodes, 3 / 4 with 4-record leaf loads; 0 = default
    sparse_shadow_rays: int = 0  # shadow rays per bounce traced one per wave at or below; 0 = 8192, 2**32-1 never
    shadow_group: int = 0  # lanes per shadow ray above that: 1, 8, 16 or 32; 0 = default
    tail_group: int = 0  # lanes per path in the tail kernel: 1, 16, 32 or 64; 0 = default

    def as_c(self, algorithm: int) -> _lib.RenderSettings:
        return _lib.RenderSettings(
            self.image_width,
            self.image_height,
            self.samples_per_pixel,
            self.max_depth,
            self.rng_seed,
            algorithm,
            self.path_capacity,
            self.extend_waves,
            self.tail_paths,
            self.shadow_walk,
            self.sparse_shadow_rays,
            self.shadow_group,
            self.tail_group,
        )


@dataclass
class SceneData:
    """Host description of a scene: triangles and, optionally, spheres (SceneData, rt/types.mojo:382-452)."""

    vertices: np.ndarray  # f32[9N]
    surfaces: np.ndarray  # u32[N], surface_id(kind, index)
    lambertian: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # f32[3L] albedo
    emissive: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # f32[3E] radiance
    metal: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # f32[4M] albedo, fuzz
    dielectric: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # f32[D] refraction index
    spheres: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # f32[4S] center, signed radius
    sphere_surfaces: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))  # u32[S]
    ggx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))  # f32[4G] F0.rgb, roughness

    @property
    def triangle_count(self) -> int:
        return self.vertices.size // 9

    @property
    def sphere_count(self) -> int:
        return np.asarray(self.spheres).size // 4


class GpuScene:
    """Device-resident scene: vertices, BVH, materials and the light table."""

    def __init__(self, scene: SceneData, ctx: DeviceContext | None = None, node_width: int = 2, leaf_size: int = 0,
                 method: int = 0):
        from .bvh import build_options

        self.ctx = ctx or default_context(0)
        v = np.ascontiguousarray(scene.vertices, dtype=np.float32)
        s = np.ascontiguousarray(scene.surfaces, dtype=np.uint32)
        lamb = np.ascontiguousarray(scene.lambertian, dtype=np.float32).reshape(-1)
        emit = np.ascontiguousarray(scene.emissive, dtype=np.float32).reshape(-1)
        met = np.ascontiguousarray(getattr(scene, "metal", np.zeros(0)), dtype=np.float32).reshape(-1)
        die = np.ascontiguousarray(getattr(scene, "dielectric", np.zeros(0)), dtype=np.float32).reshape(-1)
        ggx = np.ascontiguousarray(getattr(scene, "ggx", np.zeros(0)), dtype=np.float32).reshape(-1)
        sp = np.ascontiguousarray(getattr(scene, "spheres", np.zeros(0)), dtype=np.float32).reshape(-1)
        ss = np.ascontiguousarray(getattr(scene, "sphere_surfaces", np.zeros(0)), dtype=np.uint32).reshape(-1)
        if s.size != v.size // 9 or v.size % 9:
            raise ValueError("one surface id per triangle is required")
        if sp.size % 4 or ss.size != sp.size // 4:
            raise ValueError("spheres are f32[4] (center, signed radius) with one surface id each")
        if lamb.size % 3 or emit.size % 3 or met.size % 4 or ggx.size % 4:
            raise ValueError("materials are f32[3] per Lambertian / emissive and f32[4] per metal / GGX")
        handle = ctypes.c_void_p()
        opts = build_options(node_width, leaf_size, method)
        mats = _lib.Materials(
            lamb.ctypes.data_as(ctypes.c_void_p), lamb.size // 3, met.ctypes.data_as(ctypes.c_void_p), met.size // 4,
            die.ctypes.data_as(ctypes.c_void_p), die.size, emit.ctypes.data_as(ctypes.c_void_p), emit.size // 3,
            ggx.ctypes.data_as(ctypes.c_void_p), ggx.size // 4,
        )
        _lib.check(
            _lib.lib().bajo_scene_create_mixed(
                self.ctx.handle,
                v.ctypes.data_as(ctypes.c_void_p),
                v.size // 9,
                s.ctypes.data_as(ctypes.c_void_p),
                sp.ctypes.data_as(ctypes.c_void_p),
                sp.size // 4,
                ss.ctypes.data_as(ctypes.c_void_p),
                ctypes.byref(mats),
                ctypes.byref(opts),
                ctypes.byref(handle),
            )
        )
        self._keep = (lamb, emit, met, die, ggx)
        self._handle = handle
        self.triangle_count = v.size // 9
        self.sphere_count = sp.size // 4
        self.node_width = node_width

    @property
    def handle(self):
        if not self._handle:
            raise RuntimeError("scene already freed")
        return self._handle

    builds = 1  # BVH builds so far (the autotuner's cache keys on it)

    def rebuild(self, timings: bool = False):
        t = _lib.BuildTimings() if timings else None
        _lib.check(_lib.lib().bajo_scene_rebuild(self.ctx.handle, self.handle, ctypes.byref(t) if t else None))
        self.builds += 1
        return t.as_dict() if t else None

    def set_vertices(self, vertices: torch.Tensor) -> None:
        """Enqueues a copy of device triangles f32[n, 9] into the scene (the next rebuild's input)."""
        require_device(vertices, self.ctx, "vertices")
        if vertices.dtype != torch.float32 or not vertices.is_contiguous() or vertices.numel() != 9 * self.triangle_count:
            raise ValueError("vertices must be contiguous f32 with 9 floats per scene triangle")
        _lib.check(_lib.lib().bajo_scene_set_vertices(self.ctx.handle, self.handle, ptr(vertices), self.triangle_count))

    def light_table(self):
        count = ctypes.c_uint32()
        total = ctypes.c_float()
        _lib.check(_lib.lib().bajo_scene_light_table(self.handle, None, None, ctypes.byref(count), ctypes.byref(total)))
        fields = np.empty(14 * max(count.value, 1), np.float32)
        kinds = np.empty(max(count.value, 1), np.uint32)
        _lib.check(
            _lib.lib().bajo_scene_light_table(
                self.handle, fields.ctypes.data_as(ctypes.c_void_p), kinds.ctypes.data_as(ctypes.c_void_p),
                ctypes.byref(count), ctypes.byref(total),
            )
        )
        return fields[: 14 * count.value], kinds[: count.value], total.value

    def free(self):
        if getattr(self, "_handle", None):
            _lib.lib().bajo_scene_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GpuSceneBatch(GpuScene):
    """A batch of triangle environments rendered in one wavefront (the batched many-environment render;
    bajo renders one scene per ``render_gpu`` call, rt/gpu/render.mojo:322-420).  Environment e is the
    triangles [offsets[e], offsets[e+1]) of ``scene`` with its own lights; all share ``scene``'s
    materials; their BVHs are one segmented build, LBVH or H-PLOC (``method``,
    bvh/gpu/builder/segmented_build.mojo:89-162).
    A render writes f32[E, H, W, 3]; environment e's image equals a render of its scene alone with
    seed ``rng_seed + e``."""

    def __init__(self, scene: SceneData, offsets, ctx: DeviceContext | None = None, node_width: int = 2,
                 leaf_size: int = 0, method: int = 0):
        from .bvh import build_options

        self.ctx = ctx or default_context(0)
        v = np.ascontiguousarray(scene.vertices, dtype=np.float32).reshape(-1)
        s = np.ascontiguousarray(scene.surfaces, dtype=np.uint32).reshape(-1)
        off = np.ascontiguousarray(offsets, dtype=np.uint64).reshape(-1)
        lamb = np.ascontiguousarray(scene.lambertian, dtype=np.float32).reshape(-1)
        emit = np.ascontiguousarray(scene.emissive, dtype=np.float32).reshape(-1)
        met = np.ascontiguousarray(scene.metal, dtype=np.float32).reshape(-1)
        die = np.ascontiguousarray(scene.dielectric, dtype=np.float32).reshape(-1)
        ggx = np.ascontiguousarray(getattr(scene, "ggx", np.zeros(0)), dtype=np.float32).reshape(-1)
        if v.size % 9 or s.size != v.size // 9:
            raise ValueError("one surface id per triangle is required")
        if off.size < 2 or int(off[-1]) != s.size:
            raise ValueError("offsets must be u64[E + 1] ending at the triangle count")
        if np.asarray(scene.spheres).size:
            raise ValueError("environment batches hold triangles only")
        if lamb.size % 3 or emit.size % 3 or met.size % 4:
            raise ValueError("materials are f32[3] per Lambertian / emissive and f32[4] per metal")
        handle = ctypes.c_void_p()
        opts = build_options(node_width, leaf_size, method)
        mats = _lib.Materials(
            lamb.ctypes.data_as(ctypes.c_void_p), lamb.size // 3, met.ctypes.data_as(ctypes.c_void_p), met.size // 4,
            die.ctypes.data_as(ctypes.c_void_p), die.size, emit.ctypes.data_as(ctypes.c_void_p), emit.size // 3,
            ggx.ctypes.data_as(ctypes.c_void_p), ggx.size // 4,
        )
        _lib.check(_lib.lib().bajo_scene_create_batch(
            self.ctx.handle, v.ctypes.data_as(ctypes.c_void_p), off.ctypes.data_as(ctypes.c_void_p), off.size - 1,
            s.ctypes.data_as(ctypes.c_void_p), ctypes.byref(mats), ctypes.byref(opts), ctypes.byref(handle)))
        self._keep = (lamb, emit, met, die, ggx)
        self._handle = handle
        self.triangle_count = s.size
        self.sphere_count = 0
        self.env_count = off.size - 1
        self.offsets = off.copy()
        self.node_width = node_width


def _environments(scene) -> int:
    return getattr(scene, "env_count", 1)


@dataclass
class RenderResult:
    pixels: torch.Tensor  # f32[H, W, 3] on the device
    stats: dict | None = None

    def numpy(self) -> np.ndarray:
        torch.cuda.synchronize(self.pixels.device)
        return self.pixels.cpu().numpy()


def enqueue_render_gpu(
    settings: RenderSettings, camera: np.ndarray, scene: GpuScene, algorithm: int = RENDER.PATH,
    pixels: torch.Tensor | None = None, with_stats: bool = False,
) -> RenderResult:
    """Asynchronous render (rt/gpu/render.mojo:61-235): enqueues and returns."""
    ctx = scene.ctx
    w, h = settings.image_width, settings.image_height
    envs = _environments(scene)
    if pixels is None:
        shape = (envs, h, w, 3) if isinstance(scene, GpuSceneBatch) else (h, w, 3)
        pixels = torch.empty(shape, dtype=torch.float32, device=ctx.torch_device)
    require_device(pixels, ctx, "pixels")
    if pixels.numel() < 3 * w * h * envs:
        raise ValueError("pixel buffer too small")
    cams = np.ascontiguousarray(np.asarray(camera, np.float32).reshape(-1))
    st = settings.as_c(algorithm)
    stats = _lib.RenderStats() if with_stats else None
    if cams.size > 20:
        # one camera per environment of a batch: f32[E, 20]
        if cams.size != 20 * envs:
            raise ValueError(f"{cams.size // 20} cameras for {envs} environments")
        _lib.check(_lib.lib().bajo_render_cameras(
            ctx.handle, scene.handle, ctypes.byref(st), cams.ctypes.data_as(ctypes.c_void_p), envs, ptr(pixels),
            ctypes.byref(stats) if stats else None))
    else:
        cam = (ctypes.c_float * 20)(*[float(x) for x in cams[:20]])
        _lib.check(_lib.lib().bajo_render(ctx.handle, scene.handle, ctypes.byref(st), cam, ptr(pixels),
                                          ctypes.byref(stats) if stats else None))
    out = None
    if stats is not None:
        out = {name: getattr(stats, name) for name, _ in _lib.RenderStats._fields_}
        out["rays_per_bounce"] = list(stats.rays_per_bounce)
        out["extend_bounce_ms"] = list(stats.extend_bounce_ms)
        out["shadow_bounce_ms"] = list(stats.shadow_bounce_ms)
        out["max_extension_visits"] = list(stats.max_extension_visits)
        out["max_shadow_visits"] = list(stats.max_shadow_visits)
    return RenderResult(pixels, out)


def _cam20(camera) -> ctypes.Array:
    return (ctypes.c_float * 20)(*[float(x) for x in np.asarray(camera, np.float32).reshape(-1)[:20]])


class GpuRtRenderTarget:
    """GpuRtRenderTarget (rt/gpu/resources.mojo:38-100): path queues, the frame's camera and the pixel output
    kept on the device across frames, for enqueue-only submission (`enqueue`) and HIP-graph replay
    (`capture` once, then `replay` per frame; `update_camera` between replays takes effect)."""

    def __init__(self, settings: RenderSettings, camera, algorithm: int = RENDER.PATH, ctx: DeviceContext | None = None,
                 pixels: torch.Tensor | None = None, environments: int = 1):
        self.ctx = ctx or default_context(0)
        w, h = settings.image_width, settings.image_height
        shape = (h, w, 3) if environments == 1 else (environments, h, w, 3)
        self.pixels = pixels if pixels is not None else torch.empty(shape, dtype=torch.float32,
                                                                   device=self.ctx.torch_device)
        require_device(self.pixels, self.ctx, "pixels")
        if self.pixels.numel() < 3 * w * h * environments:
            raise ValueError("pixel buffer too small")
        self.settings = settings
        self.algorithm = algorithm
        self.environments = environments
        st = settings.as_c(algorithm)
        self._handle = ctypes.c_void_p()
        _lib.check(_lib.lib().bajo_render_target_create_envs(self.ctx.handle, ctypes.byref(st), environments,
                                                             _cam20(camera), ptr(self.pixels),
                                                             ctypes.byref(self._handle)))

    @property
    def handle(self):
        if not self._handle:
            raise RuntimeError("render target already freed")
        return self._handle

    def update_camera(self, camera) -> None:
        """update_gpu_camera (rt/gpu/resources.mojo:103-115), ordered on the stream instead of synchronizing."""
        _lib.check(_lib.lib().bajo_render_target_set_camera(self.ctx.handle, self.handle, _cam20(camera)))

    def enqueue(self, scene: "GpuScene") -> torch.Tensor:
        """Enqueues one render of `scene` into `pixels` and returns without synchronizing."""
        _lib.check(_lib.lib().bajo_render_enqueue(self.ctx.handle, self.handle, scene.handle))
        return self.pixels

    def capture(self, scene: "GpuScene", rebuild: bool = False) -> None:
        """Records the render of `scene` (with `rebuild`, its BVH rebuild first) into one HIP graph."""
        flags = _lib.BAJO_CAPTURE_REBUILD if rebuild else 0
        _lib.check(_lib.lib().bajo_render_target_capture_ex(self.ctx.handle, self.handle, scene.handle, flags))
        self._scene = scene  # the graph holds the scene's device pointers

    def replay(self) -> torch.Tensor:
        _lib.check(_lib.lib().bajo_render_target_replay(self.ctx.handle, self.handle))
        return self.pixels

    def free(self) -> None:
        if self._handle:
            _lib.lib().bajo_render_target_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def update_gpu_camera(target: GpuRtRenderTarget, camera) -> None:
    """update_gpu_camera (rt/gpu/resources.mojo:103-115)."""
    target.update_camera(camera)


def render_gpu(settings: RenderSettings, camera: np.ndarray, world, algorithm: int = RENDER.PATH, ctx=None) -> RenderResult:
    """render_gpu[ALGORITHM](settings, camera, world) (rt/gpu/render.mojo:322-420)."""
    scene = world if isinstance(world, GpuScene) else GpuScene(world, ctx)
    return enqueue_render_gpu(settings, camera, scene, algorithm, with_stats=True)


def rng_stage_floats(seed: int, path_id: int, stage: int, count: int) -> np.ndarray:
    """Host view of one path stage's random stream (the same numbers the kernels draw)."""
    out = np.empty(count, np.float32)
    _lib.check(_lib.lib().bajo_rng_stage_floats(seed, path_id, stage, count, out.ctypes.data_as(ctypes.c_void_p)))
    return out


# rt/wavefront_contract.mojo:107-134
def wavefront_rng_subsequence(path_id: int, stage: int) -> int:
    return (int(path_id) << 32) | int(stage)


def wavefront_rng_stage(bounce: int) -> int:
    return (0 << 30) | (bounce + 1)


def wavefront_rng_roulette_stage(bounce: int) -> int:
    return (1 << 30) | (bounce + 1)


def wavefront_rng_light_stage(bounce: int) -> int:
    return (2 << 30) | (bounce + 1)

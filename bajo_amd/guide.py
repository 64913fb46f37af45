"""Learned path guiding: the network, its training and the guided renderer.

bajo has no counterpart: its renderer samples the BSDF only
(rt/gpu/path_shading.mojo:463-535, rt/shading.mojo:92-118).  The network maps
a path vertex (position normalized to the scene box, shading normal,
outgoing direction; 12-frequency sin/cos encoding) through three hidden
layers of 64 ReLU units to an 8-lobe von Mises-Fisher mixture over
directions.  It runs on the BF16 matrix cores with FP32 accumulation and is
trained online by the library's own forward, backward and Adam kernels on the
radiance the path tracer measures.  The guided renderer draws each
Lambertian / GGX continuation from the guide or the BSDF with probability 1/2
and weights by the combined density (one-sample MIS), on top of NEE; without
a guide it is plain NEE.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from .context import DeviceContext, default_context, ptr, require_device

GUIDE_PARAMS = 24808
GUIDE_OUTPUTS = 40
GUIDE_SAMPLE_FLOATS = 13  # inputs 9, direction 3, weight 1
RENDER_GUIDED = 5


class PathGuide:
    """A guiding network on one device (bajo_guide_*)."""

    def __init__(self, ctx: DeviceContext | None = None, learning_rate: float = 1e-2, batch_size: int = 65536,
                 steps_per_frame: int = 4, seed: int = 2030, beta1: float = 0.9, beta2: float = 0.999,
                 epsilon: float = 1e-8):
        self.ctx = ctx or default_context(0)
        opt = _lib.GuideOptions(learning_rate, beta1, beta2, epsilon, batch_size, steps_per_frame, seed)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().bajo_guide_create(self.ctx.handle, ctypes.byref(opt), ctypes.byref(h)))
        self._handle = h
        self.batch_size = batch_size

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def params(self) -> np.ndarray:
        out = np.empty(GUIDE_PARAMS, np.float32)
        _lib.check(_lib.lib().bajo_guide_get_params(self.ctx.handle, self._handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        if p.size != GUIDE_PARAMS:
            raise ValueError(f"expected {GUIDE_PARAMS} parameters")
        _lib.check(_lib.lib().bajo_guide_set_params(self.ctx.handle, self._handle, p.ctypes.data_as(ctypes.c_void_p)))

    def forward(self, inputs: torch.Tensor) -> torch.Tensor:
        """Outputs f32[n, 40] for inputs f32[n, 9] on the device."""
        require_device(inputs, self.ctx, "inputs")
        x = inputs.contiguous().view(-1)
        n = x.numel() // 9
        out = torch.empty((n, GUIDE_OUTPUTS), dtype=torch.float32, device=self.ctx.torch_device)
        _lib.check(_lib.lib().bajo_guide_forward(self.ctx.handle, self._handle, ptr(x), n, ptr(out)))
        return out

    def gradient(self, samples: torch.Tensor) -> tuple[torch.Tensor, float]:
        """(gradient f32[24808], normalized loss) of samples f32[n, 13], without an update."""
        require_device(samples, self.ctx, "samples")
        s = samples.contiguous().view(-1)
        n = s.numel() // GUIDE_SAMPLE_FLOATS
        g = torch.empty(GUIDE_PARAMS, dtype=torch.float32, device=self.ctx.torch_device)
        loss = ctypes.c_float()
        _lib.check(_lib.lib().bajo_guide_gradient(self.ctx.handle, self._handle, ptr(s), n, ptr(g), ctypes.byref(loss)))
        return g, float(loss.value)

    def train(self, samples: torch.Tensor) -> None:
        """Adam steps over samples f32[n, 13] (one per batch_size samples)."""
        require_device(samples, self.ctx, "samples")
        s = samples.contiguous().view(-1)
        n = s.numel() // GUIDE_SAMPLE_FLOATS
        _lib.check(_lib.lib().bajo_guide_train(self.ctx.handle, self._handle, ptr(s), n))

    def training_samples(self, cap: int = 1 << 22) -> np.ndarray:
        """The samples the last training frame collected, f32[n, 13] (inputs 9, direction 3, weight 1)."""
        buf = np.empty((cap, GUIDE_SAMPLE_FLOATS), np.float32)
        count = ctypes.c_uint64()
        _lib.check(_lib.lib().bajo_guide_training_samples(self.ctx.handle, self._handle,
                                                          buf.ctypes.data_as(ctypes.c_void_p), cap,
                                                          ctypes.byref(count)))
        return buf[: min(count.value, cap)].copy()

    def free(self) -> None:
        if getattr(self, "_handle", None):
            _lib.lib().bajo_guide_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def mixture_sample(outputs: torch.Tensor, uniforms: torch.Tensor, ctx: DeviceContext | None = None):
    """Directions f32[n, 3] and densities f32[n] drawn from the mixtures of outputs f32[n, 40] with
    uniforms f32[n, 3] (lobe, then the lobe's two)."""
    ctx = ctx or default_context(outputs.device.index or 0)
    o = outputs.contiguous().view(-1)
    u = uniforms.contiguous().view(-1)
    n = o.numel() // GUIDE_OUTPUTS
    dirs = torch.empty((n, 3), dtype=torch.float32, device=ctx.torch_device)
    pdfs = torch.empty(n, dtype=torch.float32, device=ctx.torch_device)
    _lib.check(_lib.lib().bajo_guide_mixture_sample(ctx.handle, ptr(o), n, ptr(u), ptr(dirs), ptr(pdfs)))
    return dirs, pdfs


def render_guided(scene, guide: PathGuide | None, settings, camera, pixels: torch.Tensor | None = None,
                  train: bool = False, with_stats: bool = False):
    """One frame of the guided renderer (NEE with guided continuation); returns (pixels f32[h, w, 3], stats)."""
    ctx = scene.ctx
    st = settings.as_c(RENDER_GUIDED)
    if pixels is None:
        pixels = torch.empty((settings.image_height, settings.image_width, 3), dtype=torch.float32,
                             device=ctx.torch_device)
    require_device(pixels, ctx, "pixels")
    cam = (ctypes.c_float * 20)(*[float(x) for x in np.asarray(camera, np.float32).reshape(-1)[:20]])
    stats = _lib.GuideFrameStats()
    _lib.check(_lib.lib().bajo_render_guided(
        ctx.handle, scene.handle, guide.handle if guide is not None else None, ctypes.byref(st),
        cam, ptr(pixels), int(train),
        ctypes.byref(stats) if with_stats else None))
    return pixels, (stats if with_stats else None)

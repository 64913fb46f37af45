"""Device context: the MI355X counterpart of bajo's MAX ``DeviceContext``.

A :class:`DeviceContext` wraps one ``bajo_ctx`` bound to a device and a HIP
stream (torch's current stream by default, so library work orders with
torch work on that stream).  Buffers are torch tensors on that device; the
library only sees their device addresses.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib


class DeviceContext:
    def __init__(self, device: int = 0, stream: "torch.cuda.Stream | None" = None):
        if not torch.cuda.is_available():
            raise RuntimeError("bajo_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = int(device)
        self.torch_device = torch.device("cuda", self.device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.torch_device)
        handle = ctypes.c_void_p()
        _lib.check(_lib.lib().bajo_ctx_create(self.device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(handle)))
        self._handle = handle

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._handle

    def synchronize(self) -> None:
        """Waits for the stream; raises BajoError if a kernel raised a device status since the last check."""
        _lib.check(_lib.lib().bajo_ctx_synchronize(self._handle))

    def status(self) -> int:
        """Synchronizes, returns the device status word (0 = none raised) and clears it."""
        word = ctypes.c_uint32()
        _lib.check(_lib.lib().bajo_ctx_status(self._handle, ctypes.byref(word)))
        return word.value

    def set_spin_limit(self, spin_limit: int) -> None:
        """Spin cap of the sort's look-back wait (0 = default); a tile past it raises BAJO_STATUS_SORT_LOOKBACK."""
        _lib.check(_lib.lib().bajo_ctx_set_spin_limit(self._handle, int(spin_limit)))

    def name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _lib.check(_lib.lib().bajo_device_name(self.device, buf, 256))
        return buf.value.decode()

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _lib.lib().bajo_ctx_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.synchronize()
        self.close()


_default: dict[int, DeviceContext] = {}


def default_context(device: int = 0) -> DeviceContext:
    ctx = _default.get(device)
    if ctx is None:
        ctx = DeviceContext(device)
        _default[device] = ctx
    return ctx


def ptr(t: "torch.Tensor | None") -> ctypes.c_void_p:
    """Device address of a contiguous tensor (NULL for None)."""
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_contiguous():
        raise ValueError("bajo_amd buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def require_device(t: torch.Tensor, ctx: DeviceContext, name: str) -> None:
    if t.device != ctx.torch_device:
        raise ValueError(f"{name} must live on {ctx.torch_device}, got {t.device}")

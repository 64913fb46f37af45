"""GPU radix sorts with bajo's ``bajo/sort/gpu`` interface.

Every function sorts device tensors in place, ascending and stable, like the
Mojo functions it mirrors (sort/gpu/radix_sort.mojo:397-609,
sort/gpu/onesweep.mojo:362-572, sort/gpu/__init__.mojo:7-18).  Keys may be
64-bit (torch.int64 / torch.uint64) or 32-bit (torch.int32 / torch.uint32)
tensors; they are sorted as unsigned integers.  Values are 32-bit.
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib
from .context import DeviceContext, default_context, ptr, require_device


def _ctx(ctx: DeviceContext | None, t: torch.Tensor) -> DeviceContext:
    if ctx is None:
        ctx = default_context(t.device.index or 0)
    require_device(t, ctx, "keys")
    return ctx


def _size(keys: torch.Tensor, size: int | None) -> int:
    n = keys.numel() if size is None else int(size)
    if n < 0 or n > keys.numel():
        raise ValueError(f"size {n} outside the key buffer of {keys.numel()} elements")
    return n


def _ws(workspace: torch.Tensor | None):
    if workspace is None:
        return ctypes.c_void_p(0), 0
    return ptr(workspace), workspace.numel() * workspace.element_size()


def sort_workspace_bytes(n: int, key_bytes: int = 8, has_values: bool = True) -> int:
    return int(_lib.lib().bajo_sort_workspace_bytes(n, key_bytes, int(has_values)))


def device_radix_sort_pairs(keys, values, size=None, workspace=None, ctx=None) -> None:
    """device_radix_sort_pairs[K, u32] (sort/gpu/radix_sort.mojo:510-609)."""
    ctx = _ctx(ctx, keys)
    require_device(values, ctx, "values")
    if values.element_size() != 4:
        raise ValueError("values must be 32-bit")
    n = _size(keys, size)
    if values.numel() < n:
        raise ValueError("values buffer shorter than size")
    wp, wb = _ws(workspace)
    fn = (
        _lib.lib().bajo_device_radix_sort_pairs_u64_u32
        if keys.element_size() == 8
        else _lib.lib().bajo_device_radix_sort_pairs_u32_u32
    )
    _lib.check(fn(ctx.handle, ptr(keys), ptr(values), n, wp, wb))


def device_radix_sort_keys(keys, size=None, workspace=None, ctx=None) -> None:
    """device_radix_sort_keys[K] (sort/gpu/radix_sort.mojo:397-471)."""
    ctx = _ctx(ctx, keys)
    n = _size(keys, size)
    wp, wb = _ws(workspace)
    fn = _lib.lib().bajo_device_radix_sort_keys_u64 if keys.element_size() == 8 else _lib.lib().bajo_device_radix_sort_keys_u32
    _lib.check(fn(ctx.handle, ptr(keys), n, wp, wb))


def onesweep_radix_sort_keys(keys, size=None, begin_bit=0, end_bit=64, workspace=None, ctx=None) -> None:
    """onesweep_radix_sort_keys[u64] (sort/gpu/onesweep.mojo:362-451) over bits [begin_bit, end_bit)."""
    ctx = _ctx(ctx, keys)
    if keys.element_size() != 8:
        raise ValueError("onesweep keys must be 64-bit")
    n = _size(keys, size)
    wp, wb = _ws(workspace)
    _lib.check(_lib.lib().bajo_onesweep_radix_sort_keys_u64(ctx.handle, ptr(keys), n, begin_bit, end_bit, wp, wb))


def onesweep_radix_sort_pairs(keys, values, size=None, begin_bit=0, end_bit=64, workspace=None, ctx=None) -> None:
    """onesweep_radix_sort_pairs[u64, u32] (sort/gpu/onesweep.mojo:454-572)."""
    ctx = _ctx(ctx, keys)
    require_device(values, ctx, "values")
    if keys.element_size() != 8 or values.element_size() != 4:
        raise ValueError("onesweep pairs are 64-bit keys with 32-bit values")
    n = _size(keys, size)
    wp, wb = _ws(workspace)
    _lib.check(
        _lib.lib().bajo_onesweep_radix_sort_pairs_u64_u32(
            ctx.handle, ptr(keys), ptr(values), n, begin_bit, end_bit, wp, wb
        )
    )


def gpu_sort_pairs(keys, values, size=None, ctx=None) -> None:
    """gpu_sort_pairs (sort/gpu/__init__.mojo:7-18): one stable radix engine for every size."""
    device_radix_sort_pairs(keys, values, size=size, ctx=ctx)

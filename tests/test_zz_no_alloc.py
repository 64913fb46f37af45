"""The no-allocation sort entry point against the allocating one, through the
C ABI: two calls in a row (the second smaller, so the fixed scratch keeps the
first call's contents), and the size limit."""

import ctypes

import numpy as np
import pytest

from conftest import announce
from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _pairs(n, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    keys[: n // 3] &= np.uint64(0xFFFF0000FF00FF00)  # duplicate digits in some passes
    return keys, np.arange(n, dtype=np.uint32)


def test_noalloc_sort_matches_allocating_sort(gpu_ctx):
    from bajo_amd import _lib

    lib = _lib.lib()
    for n, seed, bits in ((1 << 20, 11, (0, 64)), (70001, 12, (16, 48))):
        announce(f"noalloc sort n={n} bits={bits}")
        keys, vals = _pairs(n, seed)
        ka, va = to_dev(keys), to_dev(vals)
        kb, vb = to_dev(keys), to_dev(vals)
        _lib.check(lib.bajo_onesweep_radix_sort_pairs_u64_u32_noalloc(gpu_ctx.handle, _ptr(ka), _ptr(va), n, *bits))
        _lib.check(lib.bajo_onesweep_radix_sort_pairs_u64_u32(gpu_ctx.handle, _ptr(kb), _ptr(vb), n, *bits, None, 0))
        _lib.check(lib.bajo_ctx_synchronize(gpu_ctx.handle))
        np.testing.assert_array_equal(to_host(ka, np.uint64), to_host(kb, np.uint64))
        np.testing.assert_array_equal(to_host(va, np.uint32), to_host(vb, np.uint32))
        lo, hi = bits
        digits = (keys >> np.uint64(lo)) & np.uint64((1 << (hi - lo)) - 1 if hi - lo < 64 else 2**64 - 1)
        np.testing.assert_array_equal(to_host(va, np.uint32), np.argsort(digits, kind="stable"))


def test_noalloc_sort_rejects_oversize(gpu_ctx):
    from bajo_amd import _lib

    lib = _lib.lib()
    n = (1 << 20) + 1
    rc = lib.bajo_onesweep_radix_sort_pairs_u64_u32_noalloc(gpu_ctx.handle, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                                             n, 0, 64)
    assert rc == -1
    assert b"limit" in lib.bajo_last_error()

"""GPU onesweep radix sort vs the stable-sort contract of bajo/sort/gpu
(checked with numpy's stable argsort and bajo's sort test fixtures)."""

import json
from pathlib import Path

import numpy as np
import pytest

from conftest import announce
from gpu_util import to_dev, to_host

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def composite_keys(size):
    i = np.arange(size, dtype=np.uint64)
    seg = (i * 7 % 19) & 0xFFFFFFFF
    mort = ((i * 2654435761) ^ (i << 11) ^ (i >> 3)) & 0xFFFFFFFF
    return (seg << np.uint64(32)) | mort


def test_bitonic_fixture_pairs(gpu_ctx):
    from bajo_amd import sort

    announce("sort bitonic fixture")
    g = GOLD["sort_bitonic_fixture"]
    k = to_dev(np.array(g["keys"], np.uint32))
    v = to_dev(np.arange(8, dtype=np.uint32))
    sort.device_radix_sort_pairs(k, v)
    assert to_host(k, np.uint32).tolist() == g["expected_keys"]
    assert to_host(v, np.uint32).tolist() == g["expected_values"]


def test_composite_key_pairs_end_to_end(gpu_ctx):
    from bajo_amd import sort

    announce("sort composite keys 10003")
    size = GOLD["sort_composite_keys"]["size"]
    keys = composite_keys(size)
    k, v = to_dev(keys), to_dev(np.arange(size, dtype=np.uint32))
    sort.device_radix_sort_pairs(k, v)
    hk, hv = to_host(k, np.uint64), to_host(v, np.uint32)
    np.testing.assert_array_equal(hk, keys[hv])
    np.testing.assert_array_equal(hv, np.argsort(keys, kind="stable"))


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 4095, 4096, 4097, 12289, 100003, 1 << 20])
def test_pairs_u64_random(gpu_ctx, n):
    from bajo_amd import sort

    announce(f"sort pairs u64 n={n}")
    rng = np.random.default_rng(n + 1)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    keys[: n // 3] &= np.uint64(0xFF00FF00FF00FF00)  # many duplicates in some digits
    k, v = to_dev(keys), to_dev(np.arange(n, dtype=np.uint32))
    sort.device_radix_sort_pairs(k, v)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(to_host(v, np.uint32), order)
    np.testing.assert_array_equal(to_host(k, np.uint64), keys[order])


@pytest.mark.parametrize("n", [1000, 300001])
def test_keys_u32_and_u64(gpu_ctx, n):
    from bajo_amd import sort

    announce(f"sort keys n={n}")
    rng = np.random.default_rng(7)
    k32 = rng.integers(0, 2**32 - 1, n, dtype=np.uint32, endpoint=True)
    d = to_dev(k32)
    sort.device_radix_sort_keys(d)
    np.testing.assert_array_equal(to_host(d, np.uint32), np.sort(k32))
    k64 = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    d = to_dev(k64)
    sort.device_radix_sort_keys(d)
    np.testing.assert_array_equal(to_host(d, np.uint64), np.sort(k64))


@pytest.mark.parametrize("begin,end", [(32, 64), (0, 24), (8, 16), (16, 48)])
def test_onesweep_bit_ranges_are_stable(gpu_ctx, begin, end):
    from bajo_amd import sort

    announce(f"onesweep bits [{begin},{end})")
    n = 200003
    rng = np.random.default_rng(begin * 100 + end)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    digit = (keys >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1)
    order = np.argsort(digit, kind="stable")
    k, v = to_dev(keys), to_dev(np.arange(n, dtype=np.uint32))
    sort.onesweep_radix_sort_pairs(k, v, begin_bit=begin, end_bit=end)
    np.testing.assert_array_equal(to_host(v, np.uint32), order)
    k2 = to_dev(keys)
    sort.onesweep_radix_sort_keys(k2, begin_bit=begin, end_bit=end)
    np.testing.assert_array_equal(to_host(k2, np.uint64), keys[order])


def test_sort_with_caller_workspace(gpu_ctx):
    import torch

    from bajo_amd import sort

    announce("sort caller workspace")
    n = 50000
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    ws_bytes = sort.sort_workspace_bytes(n, 8, True)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda:0")
    k, v = to_dev(keys), to_dev(np.arange(n, dtype=np.uint32))
    sort.device_radix_sort_pairs(k, v, workspace=ws)
    np.testing.assert_array_equal(to_host(v, np.uint32), np.argsort(keys, kind="stable"))
    small = torch.empty(16, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(Exception, match="workspace"):
        sort.device_radix_sort_pairs(k, v, workspace=small)


def test_lookback_give_up_is_reported_never_silent(gpu_ctx):
    """The look-back wait is capped; a tile past the cap raises BAJO_STATUS_SORT_LOOKBACK.
    With the cap forced to 1 spin from the host (every kernel still exits normally), a
    multi-tile sort either comes out right or reports the status: never silently wrong.
    With the default cap the same sort is right and reports nothing."""
    from bajo_amd import sort
    from bajo_amd._lib import BAJO_STATUS_SORT_LOOKBACK
    from bajo_amd.context import DeviceContext

    announce("sort look-back spin cap forced to 1")
    ctx = DeviceContext(0)
    n = 1 << 21
    rng = np.random.default_rng(77)
    keys = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    want = np.sort(keys)
    ctx.set_spin_limit(1)
    silent_wrong = 0
    raised = 0
    for _ in range(3):
        k = to_dev(keys)
        sort.device_radix_sort_keys(k, ctx=ctx)
        status = ctx.status()
        right = np.array_equal(to_host(k, np.uint64), want)
        assert status in (0, BAJO_STATUS_SORT_LOOKBACK)
        raised += status != 0
        silent_wrong += (not right) and status == 0
    assert silent_wrong == 0
    ctx.set_spin_limit(0)
    k = to_dev(keys)
    sort.device_radix_sort_keys(k, ctx=ctx)
    assert ctx.status() == 0
    np.testing.assert_array_equal(to_host(k, np.uint64), want)
    print(f"forced give-ups reported in {raised} of 3 sorts")
    ctx.close()

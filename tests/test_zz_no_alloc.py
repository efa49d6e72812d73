"""GPU check of the allocation-free sort (dsmc_sort_by_key_noalloc) against the
allocating one (dsmc_sort_by_key) through the C ABI: two calls in a row on
one device, the second with a smaller input (the scratch keeps the first
call's contents), with an odd and an even number of radix passes; bit-exact.
Sizes beyond the scratch's maximum are rejected before any device work."""
import ctypes as C

import numpy as np
import pytest

from tests.gpu_util import DeviceArray

pytestmark = pytest.mark.gpu

MAX_PAIRS = 1 << 20


def _sort(lib, fn, keys, vals, bits):
    with DeviceArray(lib, keys) as dk, DeviceArray(lib, vals) as dv:
        st = fn(dk.ptr, dv.ptr, len(keys), bits, None)
        assert st == 0, lib.dsmc_last_error(None)
        # the no-alloc form is asynchronous on the default stream; the copy
        # back (hipMemcpy) waits for it
        return dk.get(), dv.get()


def _case(rng, n, bits):
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals


def test_noalloc_sort_equals_sort_by_key_twice(lib):
    rng = np.random.default_rng(2024)
    # (n, bits): the largest input first, then a smaller one, so the second
    # call runs on scratch that still holds the first call's data
    for n, bits in ((MAX_PAIRS, 22), (100003, 16), (2049, 32), (1, 5)):
        print(f"noalloc sort n={n} bits={bits}", flush=True)
        keys, vals = _case(rng, n, bits)
        k0, v0 = _sort(lib, lib.dsmc_sort_by_key, keys, vals, bits)
        k1, v1 = _sort(lib, lib.dsmc_sort_by_key_noalloc, keys, vals, bits)
        np.testing.assert_array_equal(k1, k0)
        np.testing.assert_array_equal(v1, v0)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(k1, keys[order])
        np.testing.assert_array_equal(v1, vals[order])


def test_noalloc_sort_rejects_oversize(lib):
    print("noalloc sort: size check", flush=True)
    st = lib.dsmc_sort_by_key_noalloc(C.c_void_p(16), C.c_void_p(16), MAX_PAIRS + 1, 20, None)
    assert st == 1   # DSMC_EINVAL, before any device access
    assert b"at most" in lib.dsmc_last_error(None)
    st = lib.dsmc_sort_by_key_noalloc(None, None, 0, 20, None)
    assert st == 0

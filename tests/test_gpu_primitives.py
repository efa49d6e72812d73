"""GPU parity of the device primitives behind the C ABI against independent
CPU routines: Philox known-answer vectors (Random123), numpy's stable argsort
(thrust::sort_by_key, collisions.cu:282-284), numpy cumsum
(thrust::exclusive_scan, main.cu:440-442) and boolean masking
(thrust::remove_if, main.cu:493-498).  Bit-exact."""
import ctypes as C

import numpy as np
import pytest

from tests.gpu_util import DeviceArray

pytestmark = pytest.mark.gpu

# Random123 philox4x32_10 known-answer vectors: (counter, key, output)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def _philox_dev(lib, ctr, key, n=1):
    out = DeviceArray(lib, np.zeros(4 * n, np.uint32))
    c = (C.c_uint * 4)(*ctr)
    k = (C.c_uint * 2)(*key)
    st = lib.dsmc_philox(out.ptr, n, c, k, None)
    assert st == 0, lib.dsmc_last_error(None)
    v = out.get()
    out.free()
    return v


@pytest.mark.parametrize("ctr,key,expect", PHILOX_KAT)
def test_philox_known_answers(lib, ctr, key, expect):
    print(f"philox KAT ctr={ctr}", flush=True)
    assert list(_philox_dev(lib, ctr, key)) == expect


def test_philox_stream_matches_oracle(lib, orc):
    ctr = [123, 7, 9, 0x01000005]
    key = [1, 3]
    dev = _philox_dev(lib, ctr, key, n=4096).reshape(-1, 4)
    for i in (0, 1, 17, 1000, 4095):
        assert list(dev[i]) == orc.philox([ctr[0] + i, ctr[1], ctr[2], ctr[3]], key)


def _sort_dev(lib, keys, vals, bits):
    with DeviceArray(lib, keys) as dk, DeviceArray(lib, vals) as dv:
        st = lib.dsmc_sort_by_key(dk.ptr, dv.ptr, len(keys), bits, None)
        assert st == 0, lib.dsmc_last_error(None)
        return dk.get(), dv.get()


@pytest.mark.parametrize("n", [1, 7, 63, 64, 2047, 2048, 2049, 100003, 1 << 20])
@pytest.mark.parametrize("bits", [5, 8, 16, 22, 32])
def test_sort_by_key_random(lib, n, bits):
    print(f"sort n={n} bits={bits}", flush=True)
    rng = np.random.default_rng(n * 131 + bits)
    hi = (1 << bits) if bits < 32 else (1 << 32)
    keys = rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    k, v = _sort_dev(lib, keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("pattern", ["equal", "sorted", "reversed", "few", "max"])
def test_sort_by_key_patterns(lib, pattern):
    n = 300001
    rng = np.random.default_rng(5)
    if pattern == "equal":
        keys = np.full(n, 12345, np.uint32)
    elif pattern == "sorted":
        keys = np.sort(rng.integers(0, 1 << 20, n).astype(np.uint32))
    elif pattern == "reversed":
        keys = np.sort(rng.integers(0, 1 << 20, n).astype(np.uint32))[::-1].copy()
    elif pattern == "few":
        keys = rng.integers(0, 3, n).astype(np.uint32)
    else:
        keys = np.full(n, 0xFFFFFFFF, np.uint32)
        keys[::7] = 0
    vals = np.arange(n, dtype=np.uint32)
    k, v = _sort_dev(lib, keys, vals, 32 if pattern == "max" else 21)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("n", [1, 5, 2048, 2049, 999999, 2097153])
def test_exclusive_scan(lib, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 50, n).astype(np.int32)
    with DeviceArray(lib, x) as dx, DeviceArray(lib, np.zeros(n, np.int32)) as dy:
        st = lib.dsmc_exclusive_scan(dx.ptr, dy.ptr, n, None)
        assert st == 0
        y = dy.get()
    expect = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32)
    np.testing.assert_array_equal(y, expect)


@pytest.mark.parametrize("n,frac", [(1, 0.5), (2048, 0.0), (2048, 1.0), (100001, 0.3),
                                    (1 << 20, 0.01)])
def test_remove_if(lib, n, frac):
    rng = np.random.default_rng(n)
    vals = rng.integers(0, 1 << 30, n).astype(np.int32)
    stencil = (rng.random(n) < frac).astype(np.int32)
    with DeviceArray(lib, vals) as dv, DeviceArray(lib, stencil) as ds:
        n_out = C.c_int()
        st = lib.dsmc_remove_if(dv.ptr, ds.ptr, n, C.byref(n_out), None)
        assert st == 0
        got = dv.get()[: n_out.value]
    np.testing.assert_array_equal(got, vals[stencil == 0])

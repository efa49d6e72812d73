"""The atom label inside the FP32 state records (csrc/label_bits.hpp),
compiled for the host from the same header: the position remainder rounded
to nearest-even at 12 explicit mantissa bits, checked against an exact
rational rounding of the same float (normal and subnormal floats, ties,
overflow to infinity); the label's three parts written into and read back
from the remainders for labels 0 .. 2^31 - 1; infinity, NaN and signed zero
kept through a round trip."""
import ctypes as C
import math
import subprocess
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = r"""
#include "label_bits.hpp"
extern "C" {
void round_batch(const float *x, int n, float *y)
{
    for (int i = 0; i < n; ++i) y[i] = dsmc::rem_round(x[i]);
}
// remainders r (n x 3) tagged with labels, then read back: values and labels
void tag_batch(const float *r, const int *label, int n, float *stored, float *value, int *back)
{
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            stored[3 * i + k] = dsmc::rem_tag(r[3 * i + k], dsmc::label_part(label[i], k));
            value[3 * i + k] = dsmc::rem_value(stored[3 * i + k]);
        }
        back[i] = dsmc::rec_label(stored[3 * i], stored[3 * i + 1], stored[3 * i + 2]);
    }
}
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("label")
    (d / "t.cpp").write_text(SRC)
    so = d / "libt.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
                    f"-I{ROOT / 'cudadsmc_amd' / 'csrc'}", str(d / "t.cpp"), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    L.round_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.tag_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def rem_round(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib.round_batch(x.ctypes.data, len(x), y.ctypes.data)
    return y


def exact_round(x):
    """x (a finite float32) rounded to nearest-even on the grid of 13
    significant bits of its binade (2^(E - 12) for x in [2^E, 2^(E+1)); the
    subnormals share the grid of the smallest binade, 2^-138), in exact
    rational arithmetic; values that round to 2^128 or beyond are infinite."""
    f = Fraction(float(x))
    if f == 0:
        return np.float32(x)
    E = math.frexp(float(x))[1] - 1   # floor(log2 |x|), exact for a nonzero float
    q = Fraction(2) ** (max(E, -126) - 12)
    n = f / q
    fl = n.numerator // n.denominator
    rem = n - fl
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and fl % 2 == 1):
        fl += 1
    r = fl * q
    if abs(r) >= Fraction(2) ** 128:
        return np.float32(np.inf) if r > 0 else np.float32(-np.inf)
    return np.float32(float(r))


def test_round_matches_exact_rounding(lib):
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 0x7F800000, 4000, dtype=np.uint32)   # finite, every binade
    bits[: 400] = rng.integers(0, 0x00800000, 400, dtype=np.uint32)   # subnormals
    x = bits.view(np.float32)
    x = np.concatenate([x, -x])
    # ties: exactly half a step of the 13-bit grid above an even and an odd value
    base = np.array([1.0, 1.0 + 2.0 ** -12, 3e-12, -7.5e-9], dtype=np.float32)
    ties = (base.view(np.uint32) & ~np.uint32(0x7FF)) | np.uint32(0x400)
    x = np.concatenate([x, ties.view(np.float32),
                        np.array([np.float32(3.4028235e38), np.float32(-3.4028235e38)])])
    y = rem_round(lib, x)
    expect = np.array([exact_round(v) for v in x], dtype=np.float32)
    np.testing.assert_array_equal(y.view(np.uint32), expect.view(np.uint32))
    assert np.all((y.view(np.uint32) & 0x7FF) == 0)
    np.testing.assert_array_equal(rem_round(lib, y).view(np.uint32), y.view(np.uint32))   # idempotent
    # a step of the grid: relative error at most 2^-13 for normal floats
    nrm = np.abs(x) >= np.float32(2.0 ** -126)
    fin = np.isfinite(y)
    rel = np.abs(y[nrm & fin].astype(np.float64) - x[nrm & fin]) / np.abs(x[nrm & fin])
    assert rel.max() <= 2.0 ** -13


def test_labels_round_trip(lib):
    rng = np.random.default_rng(6)
    n = 5000
    labels = rng.integers(0, 2 ** 31, n, dtype=np.int64)
    labels[:6] = [0, 1, 2047, 2048, 2 ** 22 - 1, 2 ** 31 - 1]
    labels = labels.astype(np.int32)
    r = (rng.standard_normal((n, 3)) * 10.0 ** rng.uniform(-20, -3, (n, 3))).astype(np.float32)
    r[0] = [0.0, -0.0, np.inf]
    r[1] = [-np.inf, np.nan, 1e-45]
    stored = np.empty_like(r)
    value = np.empty_like(r)
    back = np.empty(n, dtype=np.int32)
    lib.tag_batch(r.ctypes.data, labels.ctypes.data, n, stored.ctypes.data, value.ctypes.data,
                  back.ctypes.data)
    np.testing.assert_array_equal(back, labels)
    expect = rem_round(lib, r.ravel()).reshape(r.shape)
    np.testing.assert_array_equal(value.view(np.uint32), expect.view(np.uint32))
    assert value[0, 0] == 0 and not np.signbit(value[0, 0])
    assert value[0, 1] == 0 and np.signbit(value[0, 1])
    assert value[0, 2] == np.inf and value[1, 0] == -np.inf and np.isnan(value[1, 1])
    assert value[1, 2] == 0.0   # below the grid of the smallest binade

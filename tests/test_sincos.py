"""The FP32 sine and cosine of the spin substeps (csrc/sincos_f32.hpp),
compiled for the host from the same header and checked against FP64 sin and
cos of the same float arguments: within 2 ulp of the result, or 2^-24
absolute near a zero, over small, typical and large angles, at quadrant
boundaries and beyond the reduction range (the library fallback)."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = r"""
#include "sincos_f32.hpp"
extern "C" void sincos_batch(const float *x, int n, float *s, float *c)
{
    for (int i = 0; i < n; ++i) dsmc::sincos_f32(x[i], s + i, c + i);
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("sincos")
    (d / "t.cpp").write_text(SRC)
    so = d / "libt.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                    f"-I{ROOT / 'cudadsmc_amd' / 'csrc'}", str(d / "t.cpp"), "-o", str(so)], check=True)
    L = C.CDLL(str(so))
    L.sincos_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    return L


def run(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib.sincos_batch(x.ctypes.data, len(x), s.ctypes.data, c.ctypes.data)
    return s, c


def check(got, ref):
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    err = np.abs(got.astype(np.float64) - ref)
    bound = np.maximum(2.0 * ulp, 2.0 ** -24)
    return float(np.max(err / bound))


def test_sincos_f32_against_fp64(lib):
    rng = np.random.default_rng(5)
    pieces = [
        rng.uniform(-0.1, 0.1, 200000),            # the substeps' usual angles
        rng.uniform(-4.0, 4.0, 200000),
        rng.uniform(-8192.0, 8192.0, 200000),      # the whole reduced range
        rng.uniform(-1e6, 1e6, 20000),             # the fallback
        np.array([0.0, -0.0, 1e-30, -1e-30, 1e-8]),
    ]
    # around multiples of pi/4 and pi/2 (quadrant and polynomial boundaries)
    k = np.arange(-5000, 5001)
    for base in (np.pi / 4, np.pi / 2):
        m = (k * base).astype(np.float32)
        pieces += [m, np.nextafter(m, np.float32(np.inf)), np.nextafter(m, np.float32(-np.inf))]
    x = np.concatenate([np.asarray(p, dtype=np.float32) for p in pieces])
    s, c = run(lib, x)
    xd = x.astype(np.float64)
    worst_s = check(s, np.sin(xd))
    worst_c = check(c, np.cos(xd))
    print(f"worst error / bound: sin {worst_s:.3f}, cos {worst_c:.3f}")
    assert worst_s <= 1.0 and worst_c <= 1.0
    # sin^2 + cos^2 = 1 to FP32 rounding
    assert float(np.max(np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1.0))) < 5e-7


def test_sincos_f32_special_values(lib):
    s, c = run(lib, np.array([np.nan, np.inf, -np.inf], dtype=np.float32))
    assert np.all(np.isnan(s)) and np.all(np.isnan(c))
    s, c = run(lib, np.array([0.0], dtype=np.float32))
    assert s[0] == 0.0 and c[0] == 1.0

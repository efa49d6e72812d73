"""Build recipes of the product: the HIP library and the C++ host driver.

`build_all()` compiles, in place:
  * cudadsmc_amd/libdsmc.so      -- HIP kernels + C ABI (hipcc, gfx950)
  * cudadsmc_amd/libdsmc_h5.so   -- HDF5 output in the reference's layout (when libhdf5
                                    with headers is installed)
  * cudadsmc_amd/bin/cudadsmc    -- C++ host driver (CLI of main.cu:36-69)
Every HIP source is compiled with -ffp-contract=off (FP64 results reproducible
by the CPU restatement) and without packed FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libdsmc.so"
DRIVER = PKG / "bin" / "cudadsmc"
H5LIB = PKG / "libdsmc_h5.so"

HIP_SOURCES = ["primitives.hip", "plain.hip", "fused.hip", "context.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if res.returncode != 0:
        err = "\n".join(line for line in res.stderr.splitlines()
                        if "packed-fp32-ops" not in line)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(map(str, cmd))}\n{err}")
    return res


def build_library(jobs: int = 4) -> Path:
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    objs = []

    def compile_one(src):
        obj = objdir / (Path(src).stem + ".o")
        _run([hipcc, *HIPCC_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> Path:
    hipcc = _hipcc()
    DRIVER.parent.mkdir(exist_ok=True)
    src = CSRC / "driver" / "dsmc_main.cpp"
    h5 = ["-DDSMC_WITH_H5", "-ldsmc_h5"] if H5LIB.exists() and _hdf5() is not None else []
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(src), "-o", str(DRIVER),
          f"-L{PKG}", *h5, "-ldsmc", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."])
    del hipcc
    return DRIVER


HDF5_PREFIXES = [os.environ.get("HDF5_DIR"), "/opt/conda", "/usr", "/usr/lib/x86_64-linux-gnu/hdf5/serial"]


def _hdf5() -> Path | None:
    """Prefix holding include/hdf5.h and lib/libhdf5.so, or None."""
    for pre in HDF5_PREFIXES:
        if pre and (Path(pre) / "include" / "hdf5.h").exists() and \
                (Path(pre) / "lib" / "libhdf5.so").exists():
            return Path(pre)
    return None


def build_h5() -> Path | None:
    """libdsmc_h5.so: HDF5 output in the reference's file layout (host C++).
    Skipped (None) when no libhdf5 with headers is installed."""
    pre = _hdf5()
    if pre is None:
        return None
    src = CSRC / "io" / "dsmc_h5.cpp"
    tmp = H5LIB.with_suffix(".so.tmp")
    # libstdc++ is linked statically: the run path that finds libhdf5 must not
    # also supply an older libstdc++ than the one the HIP runtime needs
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", f"-I{ROOT / 'include'}",
          f"-I{pre / 'include'}", str(src), "-o", str(tmp), "-static-libstdc++", "-static-libgcc",
          f"-L{PKG}", "-ldsmc", f"-L{pre / 'lib'}", "-lhdf5", "-Wl,--enable-new-dtags",
          f"-Wl,-rpath,$ORIGIN:{pre / 'lib'}"])
    os.replace(tmp, H5LIB)
    return H5LIB


def build_all(verbose: bool = True) -> None:
    steps = [("libdsmc.so", build_library), ("libdsmc_h5.so", build_h5), ("driver", build_driver)]
    for name, fn in steps:
        out = fn()
        if verbose:
            print(f"[build] {name}: {out}", flush=True)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)

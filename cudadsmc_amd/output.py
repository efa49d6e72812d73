"""HDF5 output in the reference's file layout, and the checks of its
post-processing scripts over such a file.

The writer is libdsmc_h5.so (include/dsmc_h5.h), the counterpart of
CUDADSMC/src/hdf5Helpers.cu:13-163: group "/atomData", rank-3 datasets
(rows, cols, t) with t unlimited and one chunk per frame, the thirteen datasets
of main.cu:202-383 written once per print (main.cu:526-576).

`energy_check` and `temperature` restate CUDADSMC/energyChecker.py:118-124,
170-185, `collision_rate` CUDADSMC/collisionChecker.py:56-64 and
`collision_rate_cumulative` CUDADSMC/collisionComparison.py:70-75 (Python 3,
reading the file through the same library instead of h5py).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Dict

import numpy as np

from . import _abi
from .configs import GS, KB, MRB, MUB

H5_LIB_PATH = Path(__file__).resolve().parent / "libdsmc_h5.so"

DOUBLE, INT, BOOL = 0, 1, 2
_NP = {DOUBLE: np.float64, INT: np.int32, BOOL: np.uint8}

# main.cu:202-383: dataset -> (columns, type); rows are the atom capacity, or
# cells + 1 for the per-cell counts, or 1 for time and atom number
DATASETS = {
    "positions": (3, DOUBLE), "velocities": (3, DOUBLE), "evapPos": (3, DOUBLE),
    "evapVel": (3, DOUBLE), "psiUp": (2, DOUBLE), "psiDn": (2, DOUBLE),
    "collisionCount": (1, INT), "atomCount": (1, INT), "simuatedTime": (1, DOUBLE),
    "atomNumber": (1, INT), "cellID": (1, INT), "atomID": (1, INT), "atomIsSpinUp": (1, BOOL),
}

_h5 = None


def load_library():
    """libdsmc_h5.so, or an error naming what is missing (it is built only
    where libhdf5 with headers is installed)."""
    global _h5
    if _h5 is not None:
        return _h5
    _abi.load_library()   # libdsmc.so first: libdsmc_h5.so links it
    if not H5_LIB_PATH.exists():
        raise RuntimeError(f"{H5_LIB_PATH} is missing: build() makes it where libhdf5 is installed")
    lib = C.CDLL(str(H5_LIB_PATH))
    sig = {
        "dsmc_h5_create": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]),
        "dsmc_h5_dataset": (C.c_int, [C.c_void_p, C.c_char_p, C.c_longlong, C.c_longlong, C.c_int]),
        "dsmc_h5_append": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p]),
        "dsmc_h5_write_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
        "dsmc_h5_close": (None, [C.c_void_p]),
        "dsmc_h5_dims": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_longlong)]),
        "dsmc_h5_read": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, C.c_size_t]),
        "dsmc_h5_last_error": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _h5 = lib
    return lib


def available() -> bool:
    try:
        load_library()
        return True
    except (RuntimeError, OSError):
        return False


def _check(lib, st):
    if st != 0:
        raise RuntimeError(f"dsmc_h5: {lib.dsmc_h5_last_error().decode()}")


class H5Writer:
    """createHDF5File + one handle per dataset (hdf5Helpers.cu:13-112)."""

    def __init__(self, filename: str, group: str = "/atomData"):
        self.lib = load_library()
        self.group = group
        h = C.c_void_p()
        _check(self.lib, self.lib.dsmc_h5_create(str(filename).encode(), group.encode(), C.byref(h)))
        self.h = h

    def dataset(self, name: str, rows: int, cols: int, dtype: int = DOUBLE):
        _check(self.lib, self.lib.dsmc_h5_dataset(self.h, name.encode(), rows, cols, dtype))

    def append(self, name: str, data: np.ndarray, dtype: int = DOUBLE):
        """writeHDF5File (hdf5Helpers.cu:115-163): one more frame."""
        a = np.ascontiguousarray(data, dtype=_NP[dtype])
        _check(self.lib, self.lib.dsmc_h5_append(self.h, name.encode(), a.ctypes.data_as(C.c_void_p)))

    def write_frame(self, sim, time: float):
        """The thirteen datasets of main.cu:526-576 from a Simulation."""
        _check(self.lib, self.lib.dsmc_h5_write_frame(self.h, sim.h, float(time)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.dsmc_h5_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def read(filename: str, name: str, dtype: int = DOUBLE) -> np.ndarray:
    """Whole dataset, shape (rows, cols, t)."""
    lib = load_library()
    dims = (C.c_longlong * 3)()
    _check(lib, lib.dsmc_h5_dims(str(filename).encode(), name.encode(), dims))
    out = np.zeros((dims[0], dims[1], dims[2]), dtype=_NP[dtype])
    _check(lib, lib.dsmc_h5_read(str(filename).encode(), name.encode(), dtype,
                                 out.ctypes.data_as(C.c_void_p), out.nbytes))
    return out


def read_frames(filename: str, group: str = "/atomData") -> Dict[str, np.ndarray]:
    """Every dataset of the reference's layout present in the file."""
    return {k: read(filename, f"{group}/{k}", t) for k, (_, t) in DATASETS.items()}


def ip_field_magnitude(pos: np.ndarray, B0: float, dBdx: float, d2Bdx2: float) -> np.ndarray:
    """|B| of the Ioffe-Pritchard trap at positions (..., 3)
    (energyChecker.py:88-98 and moveAtoms.cu:93-121)."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    Bx = dBdx * x - 0.5 * d2Bdx2 * x * z
    By = -dBdx * y - 0.5 * d2Bdx2 * y * z
    Bz = B0 + 0.5 * d2Bdx2 * (z * z - 0.5 * (x * x + y * y))
    return np.sqrt(Bx * Bx + By * By + Bz * Bz)


def energies(frames: Dict[str, np.ndarray], B0: float = 0.01, dBdx: float = 20.0,
             d2Bdx2: float = 40000.0, mass: float = MRB, gs: float = GS):
    """Mean kinetic, potential and total energy per atom of every frame, in
    micro-kelvin (energyChecker.py:118-124): E_k = m v^2 / 2 and E_p =
    gs muB (|B| isSpinUp - B0) / 2 over the alive rows."""
    n_alive = frames["atomNumber"][0, 0, :]
    T = n_alive.size
    Ek, Ep = np.zeros(T), np.zeros(T)
    for i in range(T):
        n = int(n_alive[i])
        v = frames["velocities"][:n, :, i]
        p = frames["positions"][:n, :, i]
        up = frames["atomIsSpinUp"][:n, 0, i].astype(np.float64)
        Ek[i] = np.mean(0.5 * mass * np.sum(v * v, axis=1)) / KB * 1e6
        B = ip_field_magnitude(p, B0, dBdx, d2Bdx2)
        Ep[i] = np.mean(0.5 * gs * MUB * (B * up - B0)) / KB * 1e6
    return Ek, Ep, Ek + Ep


def energy_check(frames: Dict[str, np.ndarray], tolerance_percent: float = 1e-3, **trap):
    """energyChecker.py:170-185: the largest relative change of the mean total
    energy over the frames, in percent, and whether it is below the tolerance
    (the reference's pass mark is 1e-3 %)."""
    _, _, Et = energies(frames, **trap)
    dE = float(np.max(np.abs(Et - Et[0]) / Et[0] * 100.0))
    return dE, dE < tolerance_percent


def temperature(frames: Dict[str, np.ndarray], mass: float = MRB) -> np.ndarray:
    """T = m <v^2> / (3 kB) of the alive atoms per frame (evaporation.cu:120)."""
    n_alive = frames["atomNumber"][0, 0, :]
    out = np.zeros(n_alive.size)
    for i in range(n_alive.size):
        n = int(n_alive[i])
        v = frames["velocities"][:n, :, i]
        out[i] = mass * np.mean(np.sum(v * v, axis=1)) / (3.0 * KB)
    return out


def collision_rate(frames: Dict[str, np.ndarray]) -> np.ndarray:
    """collisionChecker.py:56-58: per frame, the collisions counted in the
    cells (collisionCount holds the counts of the frame's last collision
    step, collisions.cu:335-436) over the frame interval and the atom number:
    events per atom per second when frames are one collision step apart."""
    time = frames["simuatedTime"][0, 0, :]
    total = frames["collisionCount"][:, 0, :].sum(axis=0).astype(np.float64)
    return total / (time[1] - time[0]) / frames["atomNumber"][0, 0, :]


def collision_rate_cumulative(frames: Dict[str, np.ndarray]) -> np.ndarray:
    """collisionComparison.py:70-72: the rate from counts that accumulate over
    the run, differentiated in time (np.gradient) and divided by the atom
    number."""
    time = frames["simuatedTime"][0, 0, :]
    total = frames["collisionCount"][:, 0, :].sum(axis=0).astype(np.float64)
    return np.gradient(total, time[1] - time[0]) / frames["atomNumber"][0, 0, :]


def rate_summary(rate: np.ndarray, skip: int = 10):
    """Mean and standard deviation of a rate over frames skip .. -2
    (collisionChecker.py:60-61 and collisionComparison.py:74-75 use 10)."""
    r = np.asarray(rate)[skip:-1]
    return float(np.mean(r)), float(np.std(r))

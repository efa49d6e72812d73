"""ctypes view of include/dsmc.h.

The structures and prototypes below mirror the header field for field; the
library is loaded from the package directory and nothing else: if it is
missing the import-time loader raises, there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdsmc.so"

# status codes
DSMC_OK, DSMC_EINVAL, DSMC_EHIP, DSMC_ENOMEM, DSMC_ESTATE, DSMC_ENOTSUP = range(6)
STATUS_NAMES = {0: "DSMC_OK", 1: "DSMC_EINVAL", 2: "DSMC_EHIP", 3: "DSMC_ENOMEM",
                4: "DSMC_ESTATE", 5: "DSMC_ENOTSUP"}

TRAP_IP, TRAP_QUADRUPOLE, TRAP_BOX = 0, 1, 2
FP64_PLAIN, FP32_KAHAN = 0, 1
PAIR_UNIFORM, PAIR_REFERENCE = 0, 1
SORT_INCREMENTAL, SORT_FULL = 0, 1
EVAP_SPIN, EVAP_ENERGY = 1, 2

(FIELD_POS, FIELD_VEL, FIELD_ACC, FIELD_ATOM_ID, FIELD_CELL_ID, FIELD_PERM, FIELD_SPIN,
 FIELD_EVAP_POS, FIELD_EVAP_VEL, FIELD_EVAP_TAG, FIELD_CELL_START_END, FIELD_CELL_COUNT,
 FIELD_CELL_PREFIX, FIELD_SIGVRMAX, FIELD_COLLISION_COUNT, FIELD_TRAP_START, FIELD_POS_COMP,
 FIELD_VEL_COMP, FIELD_OBS, FIELD_SPIN_UP, FIELD_LOCAL_POPS) = range(21)


class Trap(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("n_atoms", C.c_int), ("nc", C.c_int * 3), ("philox_stream", C.c_int),
        ("half_width", C.c_double * 3),
        ("B0", C.c_double), ("dBdx", C.c_double), ("d2Bdx2", C.c_double),
        ("dBdz", C.c_double), ("mu", C.c_double), ("T", C.c_double),
        ("T_axis", C.c_double * 3), ("n_real", C.c_double),
        ("eta_start", C.c_double), ("eta_end", C.c_double), ("T_cut", C.c_double),
        ("eta_steps", C.c_longlong),
    ]


class Config(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("pair_mode", C.c_int), ("evaporation", C.c_int),
        ("spin", C.c_int), ("spin_substeps", C.c_int), ("spin_interp", C.c_int),
        ("lpc", C.c_int), ("use_graph", C.c_int),
        ("dt", C.c_double), ("mass", C.c_double), ("sigma", C.c_double), ("gs", C.c_double),
        ("seed", C.c_ulonglong), ("block_size", C.c_int), ("obs_capacity", C.c_int),
        ("sort_mode", C.c_int), ("mcwf", C.c_int), ("reserved", C.c_int * 4),
    ]


class Observables(C.Structure):
    _fields_ = [
        ("ek", C.c_double), ("ep", C.c_double), ("v2", C.c_double * 3), ("e_evap", C.c_double),
        ("n_alive", C.c_longlong), ("n_evap", C.c_longlong), ("collisions", C.c_longlong),
        ("step", C.c_longlong),
    ]


class Info(C.Structure):
    _fields_ = [
        ("precision", C.c_int), ("n_traps", C.c_int), ("n_capacity", C.c_int),
        ("n_alive", C.c_int), ("n_keys", C.c_int), ("key_bits", C.c_int),
        ("radix_passes", C.c_int), ("block_size", C.c_int),
        ("step", C.c_longlong), ("collision_steps", C.c_longlong), ("obs_capacity", C.c_longlong),
    ]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_longlong), ("total_ms", C.c_double),
                ("bytes_per_launch", C.c_double)]


P = C.c_void_p
I = C.c_int
PROTOTYPES = {
    "dsmc_abi_version": (I, []),
    "dsmc_last_error": (C.c_char_p, [P]),
    "dsmc_create": (I, [C.POINTER(Config), C.POINTER(Trap), I, C.POINTER(P)]),
    "dsmc_destroy": (I, [P]),
    "dsmc_init_distribution": (I, [P]),
    "dsmc_set_state": (I, [P, P, P, I]),
    "dsmc_set_spin": (I, [P, P, I, I]),
    "dsmc_index_atoms": (I, [P]),
    "dsmc_sort_arrays": (I, [P]),
    "dsmc_build_cell_table": (I, [P]),
    "dsmc_collide": (I, [P]),
    "dsmc_evolve_spin": (I, [P]),
    "dsmc_local_populations": (I, [P]),
    "dsmc_flip_atoms": (I, [P, C.POINTER(C.c_longlong)]),
    "dsmc_exponential_decay": (I, [P]),
    "dsmc_move_atoms": (I, [P]),
    "dsmc_calculate_temp": (I, [P, C.POINTER(C.c_double)]),
    "dsmc_evaporation_tag": (I, [P]),
    "dsmc_compact": (I, [P, C.POINTER(I)]),
    "dsmc_step": (I, [P, I]),
    "dsmc_sync": (I, [P]),
    "dsmc_get_info": (I, [P, C.POINTER(Info)]),
    "dsmc_observe": (I, [P, C.POINTER(Observables), I]),
    "dsmc_copy_out": (I, [P, I, P, C.c_size_t]),
    "dsmc_field_bytes": (I, [P, I, C.POINTER(C.c_size_t)]),
    "dsmc_time_steps": (I, [P, I, C.POINTER(C.c_double)]),
    "dsmc_profile_enable": (I, [P, I]),
    "dsmc_profile_read": (I, [P, C.POINTER(KernelStat), I, C.POINTER(I)]),
    "dsmc_autotune": (I, [P, C.POINTER(I)]),
    "dsmc_step_group": (I, [C.POINTER(P), I, I, I]),
    "dsmc_time_steps_group": (I, [C.POINTER(P), I, I, I, C.POINTER(C.c_double)]),
    "dsmc_sort_by_key": (I, [P, P, I, I, P]),
    "dsmc_sort_by_key_noalloc": (I, [P, P, I, I, P]),
    "dsmc_exclusive_scan": (I, [P, P, I, P]),
    "dsmc_remove_if": (I, [P, P, I, C.POINTER(I), P]),
    "dsmc_philox": (I, [P, I, C.POINTER(C.c_uint), C.POINTER(C.c_uint), P]),
    "dsmc_stream_copy": (I, [P, P, C.c_size_t, I, C.POINTER(C.c_double)]),
    "dsmc_gather_copy": (I, [P, P, P, I, I, C.POINTER(C.c_double)]),
    "dsmc_device_alloc": (I, [C.POINTER(P), C.c_size_t]),
    "dsmc_device_free": (I, [P]),
    "dsmc_memcpy_htod": (I, [P, P, C.c_size_t]),
    "dsmc_memcpy_dtoh": (I, [P, P, C.c_size_t]),
}

_lib = None


def load_library(path: Path | None = None):
    """Load libdsmc.so (built in-tree by __graft_entry__.build()).  Raises when
    it is missing: the product path has no CPU fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; "
                           f"g.build()'` (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(str(p))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class DsmcError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def check(lib, status: int, ctx=None):
    if status != DSMC_OK:
        msg = lib.dsmc_last_error(ctx)
        raise DsmcError(status, msg.decode() if msg else "")

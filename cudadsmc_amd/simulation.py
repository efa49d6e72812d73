"""Host-side mirror of the reference driver (CUDADSMC/src/main.cu) over the C ABI.

`Simulation` owns one libdsmc context.  Its methods are the reference's host
calls under the same names (indexAtoms -> index_atoms, sortArrays ->
sort_arrays, ...), in the order main.cu:410-524 issues them; `step()` runs that
whole loop body on the device without host round trips.  Every call goes to
libdsmc.so; there is no Python or CPU implementation behind it.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional

import numpy as np

from . import _abi
from .configs import KB, Run


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Simulation:
    def __init__(self, run: Run, lib=None):
        self.run = run
        self.lib = lib or _abi.load_library()
        self._traps = (_abi.Trap * len(run.traps))(*[t.to_c() for t in run.traps])
        self._cfg = run.to_c()
        h = C.c_void_p()
        st = self.lib.dsmc_create(C.byref(self._cfg), self._traps, len(run.traps), C.byref(h))
        _abi.check(self.lib, st, None)
        self.h = h

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.dsmc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        _abi.check(self.lib, st, self.h)

    # -- initial conditions (setUp.cu) ---------------------------------------
    def init_distribution(self):
        """h_initRNG + h_generateInitialDist + initSigvrmax + indexAtoms."""
        self._check(self.lib.dsmc_init_distribution(self.h))

    def set_state(self, pos: np.ndarray, vel: np.ndarray):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        assert pos.shape == vel.shape and pos.shape[1] == 3
        self._check(self.lib.dsmc_set_state(self.h, _ptr(pos), _ptr(vel), pos.shape[0]))

    def set_spin(self, spin: np.ndarray):
        spin = np.ascontiguousarray(spin, dtype=np.float64)
        self._check(self.lib.dsmc_set_spin(self.h, _ptr(spin), spin.shape[0], spin.shape[1]))

    # -- the step, one method per reference call -----------------------------
    def index_atoms(self):
        self._check(self.lib.dsmc_index_atoms(self.h))

    def sort_arrays(self):
        self._check(self.lib.dsmc_sort_arrays(self.h))

    def build_cell_table(self):
        self._check(self.lib.dsmc_build_cell_table(self.h))

    def collide(self):
        self._check(self.lib.dsmc_collide(self.h))

    def evolve_spin(self):
        self._check(self.lib.dsmc_evolve_spin(self.h))

    def local_populations(self):
        """getLocalPopulations (wavefunction.cu:75-97), cfg.mcwf."""
        self._check(self.lib.dsmc_local_populations(self.h))

    def flip_atoms(self) -> int:
        """flipAtoms (wavefunction.cu:99-154); returns the atoms flipped."""
        n = C.c_longlong()
        self._check(self.lib.dsmc_flip_atoms(self.h, C.byref(n)))
        return n.value

    def exponential_decay(self):
        """exponentialDecay over lpc * dt (wavefunction.cu:175-219)."""
        self._check(self.lib.dsmc_exponential_decay(self.h))

    def move_atoms(self):
        self._check(self.lib.dsmc_move_atoms(self.h))

    def calculate_temp(self) -> float:
        T = C.c_double()
        self._check(self.lib.dsmc_calculate_temp(self.h, C.byref(T)))
        return T.value

    def evaporation_tag(self):
        self._check(self.lib.dsmc_evaporation_tag(self.h))

    def compact(self) -> int:
        n = C.c_int()
        self._check(self.lib.dsmc_compact(self.h, C.byref(n)))
        return n.value

    def step(self, n_steps: int = 1):
        self._check(self.lib.dsmc_step(self.h, int(n_steps)))

    def sync(self):
        self._check(self.lib.dsmc_sync(self.h))

    # -- output surface --------------------------------------------------------
    def info(self) -> _abi.Info:
        inf = _abi.Info()
        self._check(self.lib.dsmc_get_info(self.h, C.byref(inf)))
        return inf

    def observe(self) -> List[_abi.Observables]:
        n = len(self.run.traps)
        out = (_abi.Observables * n)()
        self._check(self.lib.dsmc_observe(self.h, out, n))
        return list(out)

    _FIELD_DTYPES = {
        _abi.FIELD_POS: (np.float64, 3), _abi.FIELD_VEL: (np.float64, 3),
        _abi.FIELD_ACC: (np.float64, 3), _abi.FIELD_ATOM_ID: (np.int32, 1),
        _abi.FIELD_CELL_ID: (np.int32, 1), _abi.FIELD_PERM: (np.int32, 1),
        _abi.FIELD_EVAP_POS: (np.float64, 3), _abi.FIELD_EVAP_VEL: (np.float64, 3),
        _abi.FIELD_EVAP_TAG: (np.int32, 1), _abi.FIELD_CELL_START_END: (np.int32, 2),
        _abi.FIELD_CELL_COUNT: (np.int32, 1), _abi.FIELD_CELL_PREFIX: (np.int32, 1),
        _abi.FIELD_SIGVRMAX: (np.float64, 1), _abi.FIELD_COLLISION_COUNT: (np.int32, 1),
        _abi.FIELD_TRAP_START: (np.int32, 1), _abi.FIELD_POS_COMP: (np.float64, 3),
        _abi.FIELD_VEL_COMP: (np.float64, 3), _abi.FIELD_SPIN_UP: (np.int32, 1),
        _abi.FIELD_LOCAL_POPS: (np.float64, 2),
    }

    def copy_out(self, field: int) -> np.ndarray:
        nbytes = C.c_size_t()
        self._check(self.lib.dsmc_field_bytes(self.h, field, C.byref(nbytes)))
        if field == _abi.FIELD_OBS:
            out = np.zeros(nbytes.value // OBS_DTYPE.itemsize, dtype=OBS_DTYPE)
            self._check(self.lib.dsmc_copy_out(self.h, field, _ptr(out), nbytes.value))
            return out
        if field == _abi.FIELD_SPIN:
            cols = 3 if self.run.precision == "fp32" else 4
            dtype = np.float64
        else:
            dtype, cols = self._FIELD_DTYPES[field]
        out = np.empty(nbytes.value // np.dtype(dtype).itemsize, dtype=dtype)
        self._check(self.lib.dsmc_copy_out(self.h, field, _ptr(out), nbytes.value))
        return out.reshape(-1, cols) if cols > 1 else out

    def obs_log(self) -> np.ndarray:
        """Observables of every logged step, ordered by step, per trap."""
        inf = self.info()
        cap = int(inf.obs_capacity)
        a = self.copy_out(_abi.FIELD_OBS).reshape(cap, inf.n_traps)
        lo = max(0, int(inf.step) - cap)
        steps = np.arange(lo, int(inf.step))
        out = a[steps % cap]
        if len(steps) and not (out["step"][:, 0] == steps).all():
            raise RuntimeError("observables log out of sync with the step counter")
        return out

    # -- measurement -----------------------------------------------------------
    def time_steps(self, n_steps: int) -> float:
        ms = C.c_double()
        self._check(self.lib.dsmc_time_steps(self.h, int(n_steps), C.byref(ms)))
        return ms.value

    def profile(self, enable: bool = True):
        self._check(self.lib.dsmc_profile_enable(self.h, int(enable)))

    def profile_read(self) -> Dict[str, dict]:
        n = C.c_int()
        self._check(self.lib.dsmc_profile_read(self.h, None, 0, C.byref(n)))
        stats = (_abi.KernelStat * max(1, n.value))()
        self._check(self.lib.dsmc_profile_read(self.h, stats, n.value, C.byref(n)))
        return {s.name.decode(): {"launches": s.launches, "total_ms": s.total_ms,
                                  "bytes_per_launch": s.bytes_per_launch}
                for s in list(stats)[: n.value]}

    def autotune(self) -> int:
        bs = C.c_int()
        self._check(self.lib.dsmc_autotune(self.h, C.byref(bs)))
        return bs.value

    # -- derived observables ---------------------------------------------------
    def temperature(self, obs: Optional[_abi.Observables] = None, trap: int = 0) -> float:
        o = obs if obs is not None else self.observe()[trap]
        v2 = o.v2[0] + o.v2[1] + o.v2[2]
        return self.run.mass * v2 / (3.0 * KB * max(o.n_alive, 1))


OBS_DTYPE = np.dtype([("ek", "f8"), ("ep", "f8"), ("v2", "f8", 3), ("e_evap", "f8"),
                      ("n_alive", "i8"), ("n_evap", "i8"), ("collisions", "i8"), ("step", "i8")])
assert OBS_DTYPE.itemsize == C.sizeof(_abi.Observables)


class SimulationGroup:
    """Several contexts stepped together, one per group of independent traps
    (`configs.split_traps`): each context has its own stream and graph, and
    `step` submits their steps in turns (dsmc_step_group) so that the groups
    overlap on the GPU.  A trap's results are those of the same trap in one
    context holding all of them (its Philox stream is kept)."""

    def __init__(self, runs: List[Run], lib=None):
        if not runs:
            raise ValueError("an empty group")
        self.lib = lib or _abi.load_library()
        self.sims = [Simulation(r, lib=self.lib) for r in runs]

    def _handles(self):
        hs = [s.h for s in self.sims]
        return (C.c_void_p * len(hs))(*hs)

    def close(self):
        for s in self.sims:
            s.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init_distribution(self):
        for s in self.sims:
            s.init_distribution()

    def autotune(self) -> List[int]:
        return [s.autotune() for s in self.sims]

    def step(self, n_steps: int = 1, chunk: int = 50):
        _abi.check(self.lib, self.lib.dsmc_step_group(self._handles(), len(self.sims), int(n_steps),
                                                      int(chunk)), self.sims[0].h)

    def time_steps(self, n_steps: int, chunk: int = 50) -> float:
        ms = C.c_double()
        _abi.check(self.lib, self.lib.dsmc_time_steps_group(self._handles(), len(self.sims), int(n_steps),
                                                            int(chunk), C.byref(ms)), self.sims[0].h)
        return ms.value

    def sync(self):
        for s in self.sims:
            s.sync()

    def obs_log(self) -> np.ndarray:
        """The groups' observable logs side by side: (steps, traps) in the
        order of the traps before the split."""
        return np.concatenate([s.obs_log() for s in self.sims], axis=1)


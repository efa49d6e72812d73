"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker (or the CPU baseline),
never by the product package.

Arrays follow the reference layout: positions/velocities (N, 3) float64 AoS
indexed by physical atom id, an int32 atomID list, spinors (N, 2) complex
stored as (N, 2, 2) float64 (re, im).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

TRAP_KINDS = {"ip": 0, "quadrupole": 1, "box": 2}


class Params(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("B0", C.c_double), ("dBdx", C.c_double), ("d2Bdx2", C.c_double),
        ("dBdz", C.c_double), ("mu", C.c_double), ("mass", C.c_double), ("gs", C.c_double),
        ("nc", C.c_int * 3), ("half_width", C.c_double * 3), ("alpha", C.c_double),
        ("sigma", C.c_double), ("dt", C.c_double), ("lpc", C.c_int), ("pair_mode", C.c_int),
        ("periodic", C.c_int), ("seed", C.c_uint64), ("stream", C.c_uint32),
    ]


class D3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class RunStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("atom_steps", C.c_longlong),
                ("collisions", C.c_longlong), ("threads", C.c_int)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            import subprocess
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        _lib = C.CDLL(str(LIB))
        P = C.c_void_p
        sig = {
            "orc_philox": (None, [P, P, P]),
            "orc_u01": (C.c_double, [C.c_uint32, C.c_uint32]),
            "orc_uint_below": (C.c_uint32, [C.c_uint32, C.c_uint32]),
            "orc_index_atoms": (None, [P, P, P, P, C.c_int, C.c_int]),
            "orc_sort_arrays": (None, [P, P, C.c_int, C.c_int]),
            "orc_cell_table": (None, [P, C.c_int, C.c_int, P, P, P]),
            "orc_collide": (C.c_longlong, [P, P, P, P, P, P, C.c_int, C.c_uint32]),
            "orc_move_atoms": (None, [P, P, P, P, P, P, C.c_int]),
            "orc_evolve_wavefunction": (None, [P, P, P, P, P, C.c_int, C.c_double]),
            "orc_normalise": (None, [P, P, P, C.c_int]),
            "orc_rotate_bloch": (None, [P, P, P, P, C.c_int, C.c_double]),
            "orc_evaporation_tag": (None, [P, P, P, P, P, P, P, P, P, C.c_int, C.c_double]),
            "orc_remove_if": (C.c_int, [P, P, C.c_int]),
            "orc_calculate_temp": (C.c_double, [P, P, P, C.c_int]),
            "orc_energy": (None, [P, P, P, P, C.c_int, P]),
            "orc_energy_spin": (None, [P, P, P, P, P, C.c_int, P]),
            "orc_local_populations": (None, [P, P, P, P, P, C.c_int, P]),
            "orc_flip_atoms": (C.c_longlong, [P, P, P, P, P, P, P, P, C.c_int, C.c_uint32]),
            "orc_decoherence_tau": (C.c_double, [P, D3]),
            "orc_exponential_decay": (None, [P, P, P, P, P, C.c_int, C.c_double]),
            "orc_init_distribution": (None, [P, C.c_double, P, P, P, P, P, P, P, P, C.c_int]),
            "orc_sigvrmax_init": (C.c_double, [P, C.c_double]),
            "orc_run": (None, [P, P, P, P, P, P, P, C.c_int, C.c_int, C.c_uint32, P]),
            "orc_num_threads": (C.c_int, []),
            "orc_field": (D3, [P, D3]),
            "orc_potential": (C.c_double, [P, D3]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params_from(run, trap_index: int = 0) -> Params:
    """orc_params of one trap of a cudadsmc_amd.configs.Run."""
    t = run.traps[trap_index]
    p = Params()
    p.kind = TRAP_KINDS[t.kind]
    p.B0, p.dBdx, p.d2Bdx2, p.dBdz = t.B0, t.dBdx, t.d2Bdx2, t.dBdz
    p.mu, p.mass, p.gs = t.mu, run.mass, run.gs
    for k in range(3):
        p.nc[k] = int(t.nc[k])
        p.half_width[k] = float(t.half_width[k])
    p.alpha = t.n_real / float(t.n_atoms)
    p.sigma = run.sigma
    p.dt = run.dt
    p.lpc = int(run.lpc)
    p.pair_mode = 0 if run.pair_mode == "uniform" else 1
    p.periodic = 1 if t.kind == "box" else 0
    p.seed = int(run.seed)
    p.stream = int(trap_index + t.philox_stream)
    return p


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox(C.cast(c, C.c_void_p), C.cast(k, C.c_void_p), C.cast(o, C.c_void_p))
    return list(o)


def u01(hi, lo):
    return load().orc_u01(hi, lo)


def uint_below(r, n):
    return load().orc_uint_below(r, n)


def field(p: Params, r):
    """B(r) of the trap (T)."""
    b = load().orc_field(C.byref(p), D3(*map(float, r)))
    return np.array([b.x, b.y, b.z])


def potential(p: Params, r):
    """U(r) of one atom (J)."""
    return load().orc_potential(C.byref(p), D3(*map(float, r)))


def ncells(p: Params) -> int:
    return p.nc[0] * p.nc[1] * p.nc[2]


def index_atoms(p: Params, pos, atom_id, d1_quirk=False):
    lib = load()
    pos = np.ascontiguousarray(pos, np.float64)
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    cell = np.zeros(len(atom_id), np.int32)
    lib.orc_index_atoms(C.byref(p), _p(pos), _p(cell), _p(atom_id), len(atom_id), int(d1_quirk))
    return cell


def sort_arrays(cell, atom_id, nkeys):
    lib = load()
    cell = np.ascontiguousarray(cell, np.int32).copy()
    atom_id = np.ascontiguousarray(atom_id, np.int32).copy()
    lib.orc_sort_arrays(_p(cell), _p(atom_id), len(cell), int(nkeys))
    return cell, atom_id


def cell_table(cell_sorted, ncells_):
    lib = load()
    cell_sorted = np.ascontiguousarray(cell_sorted, np.int32)
    se = np.zeros((ncells_ + 1, 2), np.int32)
    count = np.zeros(ncells_ + 1, np.int32)
    prefix = np.zeros(ncells_ + 1, np.int32)
    lib.orc_cell_table(_p(cell_sorted), len(cell_sorted), int(ncells_), _p(se), _p(count),
                       _p(prefix))
    return se, count, prefix


def collide(p: Params, vel, sigvrmax, prefix, atom_id, step):
    """In place on vel and sigvrmax; returns (per-cell counts, total)."""
    lib = load()
    assert vel.flags.c_contiguous and vel.dtype == np.float64
    assert sigvrmax.flags.c_contiguous and sigvrmax.dtype == np.float64
    nc = ncells(p)
    counts = np.zeros(nc + 1, np.int32)
    prefix = np.ascontiguousarray(prefix, np.int32)
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    total = lib.orc_collide(C.byref(p), _p(vel), _p(sigvrmax), _p(prefix), _p(counts),
                            _p(atom_id), nc, int(step))
    return counts, total


def move_atoms(p: Params, pos, vel, acc, atom_id, spin_up=None):
    lib = load()
    for a in (pos, vel, acc):
        assert a.flags.c_contiguous and a.dtype == np.float64
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    lib.orc_move_atoms(C.byref(p), _p(pos), _p(vel), _p(acc), _p(spin_up), _p(atom_id), len(atom_id))


def evolve_wavefunction(p: Params, pos, psi_up, psi_dn, atom_id, dt_spin):
    lib = load()
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    lib.orc_evolve_wavefunction(C.byref(p), _p(pos), _p(psi_up), _p(psi_dn), _p(atom_id),
                                len(atom_id), float(dt_spin))


def normalise(psi_up, psi_dn, atom_id):
    lib = load()
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    lib.orc_normalise(_p(psi_up), _p(psi_dn), _p(atom_id), len(atom_id))


def rotate_bloch(p: Params, pos, s, atom_id, dt_spin):
    lib = load()
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    lib.orc_rotate_bloch(C.byref(p), _p(pos), _p(s), _p(atom_id), len(atom_id), float(dt_spin))


def evaporation_tag(p: Params, pos, vel, atom_id, psi_up=None, psi_dn=None, e_cut=0.0,
                    evap_pos=None, evap_vel=None):
    lib = load()
    n = len(atom_id)
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    tag = np.zeros(n, np.int32)
    if evap_pos is None:
        evap_pos = np.zeros_like(pos)
    if evap_vel is None:
        evap_vel = np.zeros_like(vel)
    lib.orc_evaporation_tag(C.byref(p), _p(pos), _p(vel), _p(evap_pos), _p(evap_vel), _p(psi_up),
                            _p(psi_dn), _p(atom_id), _p(tag), n, float(e_cut))
    return tag, evap_pos, evap_vel


def remove_if(atom_id, tag):
    lib = load()
    a = np.ascontiguousarray(atom_id, np.int32).copy()
    tag = np.ascontiguousarray(tag, np.int32)
    n = lib.orc_remove_if(_p(a), _p(tag), len(a))
    return a[:n]


def calculate_temp(p: Params, vel, atom_id):
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    return load().orc_calculate_temp(C.byref(p), _p(vel), _p(atom_id), len(atom_id))


def energy(p: Params, pos, vel, atom_id, spin_up=None):
    """(E_k, E_p, sum v^2); with spin states, U_down = -U_up - 2 mu B0."""
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    out = np.zeros(3)
    if spin_up is None:
        load().orc_energy(C.byref(p), _p(pos), _p(vel), _p(atom_id), len(atom_id), _p(out))
    else:
        up = np.ascontiguousarray(spin_up, np.uint8)
        load().orc_energy_spin(C.byref(p), _p(pos), _p(vel), _p(up), _p(atom_id), len(atom_id),
                               _p(out))
    return out


# Majorana spin-flip model (wavefunction.cu:75-237)
def local_populations(p: Params, pos, psi_up, psi_dn, atom_id, pops=None):
    """getLocalPopulations: (P_up, P_down) per label, shape (n_labels, 2)."""
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    if pops is None:
        pops = np.zeros((len(pos), 2))
    load().orc_local_populations(C.byref(p), _p(pos), _p(psi_up), _p(psi_dn), _p(atom_id),
                                 len(atom_id), _p(pops))
    return pops


def flip_atoms(p: Params, pos, vel, psi_up, psi_dn, pops, spin_up, atom_id, step):
    """flipAtoms, in place on vel and spin_up; returns the number flipped."""
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    assert vel.flags.c_contiguous and spin_up.dtype == np.uint8
    return load().orc_flip_atoms(C.byref(p), _p(pos), _p(vel), _p(psi_up), _p(psi_dn), _p(pops),
                                 _p(spin_up), _p(atom_id), len(atom_id), int(step))


def decoherence_tau(p: Params, r):
    return load().orc_decoherence_tau(C.byref(p), D3(*map(float, r)))


def exponential_decay(p: Params, pos, psi_up, psi_dn, atom_id, dt):
    atom_id = np.ascontiguousarray(atom_id, np.int32)
    load().orc_exponential_decay(C.byref(p), _p(pos), _p(psi_up), _p(psi_dn), _p(atom_id),
                                 len(atom_id), float(dt))


def init_distribution(p: Params, n, T, T_axis=None, spin=False):
    lib = load()
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    acc = np.zeros((n, 3))
    atom_id = np.zeros(n, np.int32)
    up = np.zeros(n, np.uint8)
    psi_up = np.zeros((n, 2)) if spin else None
    psi_dn = np.zeros((n, 2)) if spin else None
    tax = None if T_axis is None else np.ascontiguousarray(T_axis, np.float64)
    lib.orc_init_distribution(C.byref(p), float(T), _p(tax), _p(pos), _p(vel), _p(acc),
                              _p(psi_up), _p(psi_dn), _p(atom_id), _p(up), int(n))
    out = dict(pos=pos, vel=vel, acc=acc, atom_id=atom_id, spin_up=up)
    if spin:
        out.update(psi_up=psi_up, psi_dn=psi_dn)
    return out


def sigvrmax_init(p: Params, T):
    return load().orc_sigvrmax_init(C.byref(p), float(T))


def run(p: Params, state, n_collision_steps, step0=0, sigvrmax=None):
    """orc_run over a state dict from init_distribution (modified in place)."""
    lib = load()
    n = len(state["atom_id"])
    cell = np.zeros(n, np.int32)
    if sigvrmax is None:
        raise ValueError("sigvrmax array required")
    st = RunStats()
    lib.orc_run(C.byref(p), _p(state["pos"]), _p(state["vel"]), _p(state["acc"]),
                _p(state["atom_id"]), _p(cell), _p(sigvrmax), n, int(n_collision_steps),
                int(step0), C.byref(st))
    return dict(seconds=st.seconds, atom_steps=st.atom_steps, collisions=st.collisions,
                threads=st.threads)


def num_threads():
    return load().orc_num_threads()


class Reference:
    """The reference's loop state (main.cu:88-140) on the CPU, single trap.

    Holds AoS state indexed by physical atom id plus the atomID list and
    steps it exactly like main.cu:410-524 (without the dead median and the
    output; the Majorana spin-flip model when run.mcwf), using the oracle
    functions above."""

    def __init__(self, run, pos=None, vel=None, state=None):
        self.run = run
        self.p = params_from(run)
        t = run.traps[0]
        self.nc = ncells(self.p)
        if state is None:
            if pos is None:
                state = init_distribution(self.p, t.n_atoms, t.T,
                                          None if not any(t.T_axis) else t.T_axis,
                                          spin=run.spin)
            else:
                n = len(pos)
                state = dict(pos=np.ascontiguousarray(pos, np.float64).copy(),
                             vel=np.ascontiguousarray(vel, np.float64).copy(),
                             atom_id=np.arange(n, dtype=np.int32), spin_up=np.ones(n, np.uint8))
                state["acc"] = np.zeros((n, 3))
                self._accel_all(state)
        self.s = state
        Tm = t.T if not any(t.T_axis) else sum(t.T_axis) / 3.0
        self.sigvrmax = np.full(self.nc + 1, sigvrmax_init(self.p, Tm))
        self.step_index = 0
        self.collisions = 0
        self.cell = None
        self.evap_pos = np.zeros_like(self.s["pos"])
        self.evap_vel = np.zeros_like(self.s["vel"])
        self.coll_steps = 0
        self.flips = 0
        self.pops = np.zeros((len(self.s["pos"]), 2))

    def _accel_all(self, state):
        # a(x) for every atom: one Verlet step with dt = 0 computes exactly
        # updateAcc at the current position and leaves x and v unchanged.
        q = Params.from_buffer_copy(self.p)
        q.dt = 0.0
        q.periodic = 0
        move_atoms(q, state["pos"], state["vel"], state["acc"], state["atom_id"],
                   state.get("spin_up"))

    def collision_step(self):
        self.cell = index_atoms(self.p, self.s["pos"], self.s["atom_id"])
        self.cell, self.s["atom_id"] = sort_arrays(self.cell, self.s["atom_id"], self.nc + 1)
        se, count, prefix = cell_table(self.cell, self.nc)
        counts, total = collide(self.p, self.s["vel"], self.sigvrmax, prefix, self.s["atom_id"],
                                self.step_index)
        self.collisions += total
        self.coll_steps += 1
        if getattr(self.run, "mcwf", False) and self.run.spin:
            local_populations(self.p, self.s["pos"], self.s["psi_up"], self.s["psi_dn"],
                              self.s["atom_id"], self.pops)
        return dict(start_end=se, count=count, prefix=prefix, coll_count=counts, total=total)

    def motion_step(self):
        run = self.run
        if run.spin:
            # spin substeps at the pre-move position, or at substep midpoints
            # x + v_half (m + 1/2) dt_s along the coming drift
            S = run.spin_substeps
            dts = run.dt / S
            vh = self.s["vel"] + (0.5 * self.s["acc"]) * run.dt
            for m in range(S):
                pos_m = self.s["pos"] + vh * ((m + 0.5) * dts) if run.spin_interp else self.s["pos"]
                pos_m = np.ascontiguousarray(pos_m)
                evolve_wavefunction(self.p, pos_m, self.s["psi_up"], self.s["psi_dn"],
                                    self.s["atom_id"], dts)
        move_atoms(self.p, self.s["pos"], self.s["vel"], self.s["acc"], self.s["atom_id"],
                   self.s.get("spin_up"))
        if run.spin_loss or run.energy_cut:
            tag, _, _ = evaporation_tag(self.p, self.s["pos"], self.s["vel"], self.s["atom_id"],
                                        self.s.get("psi_up") if run.spin_loss else None,
                                        self.s.get("psi_dn") if run.spin_loss else None, 0.0,
                                        self.evap_pos, self.evap_vel)
            self.s["atom_id"] = remove_if(self.s["atom_id"], tag)
        if getattr(run, "mcwf", False) and run.spin and (self.step_index + 1) % run.lpc == 0:
            self.flips += flip_atoms(self.p, self.s["pos"], self.s["vel"], self.s["psi_up"],
                                     self.s["psi_dn"], self.pops, self.s["spin_up"],
                                     self.s["atom_id"], self.coll_steps)
            exponential_decay(self.p, self.s["pos"], self.s["psi_up"], self.s["psi_dn"],
                              self.s["atom_id"], run.lpc * run.dt)
        if run.spin and (self.step_index + 1) % run.lpc == 0:
            normalise(self.s["psi_up"], self.s["psi_dn"], self.s["atom_id"])
        self.step_index += 1

    def step(self, n=1):
        for _ in range(n):
            if self.step_index % self.run.lpc == 0:
                self.collision_step()
            self.motion_step()

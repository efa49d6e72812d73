"""Physical constants, trap/run descriptions and the named workloads.

Constants are those of include/deviceSystemParameters.cuh:14-19 and
include/initialSystemParameters.cuh:14-33 (including the truncated pi,
3.14159265, which enters the cross-section and sigma*v_r,max).

Named workloads
---------------
reference(Nc, N)  the reference's own run (main.cu:45-69, 165-168): Ioffe-Pritchard
                  trap, moment 0.5*gs*muB, 20 uK, grid +-1 mm, dt = 1e-7 s,
                  10 motion substeps per collision step, spinor evolution,
                  spin-projection loss and the Majorana spin-flip model of
                  main.cu:457-521, alpha = 3e6 / N, its biased pair choice.
quadrupole_1e4    N = 1e4 in a spherical quadrupole, B' = 2.5 T/m, moment muB/2,
                  20 uK, 35^3 cells over +-5 thermal radii, dt = 1e-6 s, FP64,
                  no spin, no evaporation (the CPU-runnable case).
quadrupole_1e6    N = 1e6, same trap, 100^3 cells (precision form).
box_thermalise    N = 2^22 in a periodic 1 mm^3 box at n = 1e19 m^-3,
                  Tx = 2 Ty = 2 Tz = 40 uK, 128^3 cells.
spin_evaporation  N = 4e6 at 50 uK with spin vectors along B, 20 spin substeps of
                  5e-8 s, spin-flip loss plus an energy cut ramped 10 -> 4 kB T0
                  over 2e5 steps.
batched_traps     256 independent quadrupole traps, N log-uniform in [1e3, 1e5],
                  B' ~ U[1, 5] T/m, T ~ U[5, 50] uK, 32^3..64^3 cells chosen per
                  trap, and per trap an energy-cut schedule: eta ramps linearly
                  from eta_start ~ U[7, 10] down by U[1, 2] over U[1e4, 5e4]
                  steps (E_cut = eta kB T).
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

from . import _abi

# include/deviceSystemParameters.cuh:14-19
GS = 0.5
MUB = 9.27400915e-24
MRB = 1.443160648e-25
PI_REF = 3.14159265
A_SCATTER = 5.3e-9
KB = 1.3806503e-23
HBAR = 1.05457148e-34
SIGMA = 8. * PI_REF * A_SCATTER * A_SCATTER  # collisions.cu:312, initialSystemParameters.cuh:25
N_REAL_REF = 3e6                             # h_N, initialSystemParameters.cuh:26

TRAP_KINDS = {"ip": _abi.TRAP_IP, "quadrupole": _abi.TRAP_QUADRUPOLE, "box": _abi.TRAP_BOX}


@dataclass
class Trap:
    kind: str = "quadrupole"
    n_atoms: int = 10000
    nc: Sequence[int] = (35, 35, 35)
    half_width: Sequence[float] = (1e-3, 1e-3, 1e-3)
    B0: float = 0.0
    dBdx: float = 0.0
    d2Bdx2: float = 0.0
    dBdz: float = 0.0
    mu: float = 0.5 * MUB
    T: float = 20e-6
    T_axis: Sequence[float] = (0.0, 0.0, 0.0)
    n_real: float = N_REAL_REF
    eta_start: float = 0.0
    eta_end: float = 0.0
    T_cut: float = 0.0
    eta_steps: int = 0
    philox_stream: int = 0

    def thermal_radius(self) -> float:
        return KB * self.T / (self.mu * self.dBdz)

    def to_c(self) -> _abi.Trap:
        t = _abi.Trap()
        t.kind = TRAP_KINDS[self.kind]
        t.n_atoms = int(self.n_atoms)
        for k in range(3):
            t.nc[k] = int(self.nc[k])
            t.half_width[k] = float(self.half_width[k])
            t.T_axis[k] = float(self.T_axis[k])
        t.B0, t.dBdx, t.d2Bdx2, t.dBdz = self.B0, self.dBdx, self.d2Bdx2, self.dBdz
        t.mu, t.T, t.n_real = self.mu, self.T, self.n_real
        t.eta_start, t.eta_end, t.T_cut = self.eta_start, self.eta_end, self.T_cut
        t.eta_steps = int(self.eta_steps)
        t.philox_stream = int(self.philox_stream)
        return t


@dataclass
class Run:
    traps: List[Trap] = field(default_factory=lambda: [Trap()])
    precision: str = "fp64"            # "fp64" (plain) or "fp32" (Kahan)
    pair_mode: str = "uniform"         # or "reference" (collisions.cu:400-413)
    spin: bool = False
    spin_substeps: int = 1
    spin_interp: bool = False
    spin_loss: bool = False
    energy_cut: bool = False
    lpc: int = 1
    dt: float = 1e-6
    mass: float = MRB
    sigma: float = SIGMA
    gs: float = GS
    seed: int = 1
    use_graph: bool = True
    block_size: int = 0
    obs_capacity: int = 0
    sort_mode: str = "incremental"     # or "full": radix sort of every key each step
    mcwf: bool = False                 # FP64 form: Majorana spin flips (wavefunction.cu:75-237)
    steps: int = 1000
    name: str = "custom"

    @property
    def n_atoms(self) -> int:
        return sum(int(t.n_atoms) for t in self.traps)

    def with_(self, **kw) -> "Run":
        r = copy.deepcopy(self)
        for k, v in kw.items():
            setattr(r, k, v)
        return r

    def to_c(self) -> _abi.Config:
        c = _abi.Config()
        c.precision = _abi.FP64_PLAIN if self.precision == "fp64" else _abi.FP32_KAHAN
        c.pair_mode = _abi.PAIR_UNIFORM if self.pair_mode == "uniform" else _abi.PAIR_REFERENCE
        c.evaporation = (_abi.EVAP_SPIN if self.spin_loss else 0) | \
                        (_abi.EVAP_ENERGY if self.energy_cut else 0)
        c.spin = int(self.spin)
        c.spin_substeps = int(self.spin_substeps)
        c.spin_interp = int(self.spin_interp)
        c.lpc = int(self.lpc)
        c.use_graph = int(self.use_graph)
        c.dt, c.mass, c.sigma, c.gs = self.dt, self.mass, self.sigma, self.gs
        c.seed = int(self.seed)
        c.block_size = int(self.block_size)
        c.obs_capacity = int(self.obs_capacity)
        c.sort_mode = _abi.SORT_INCREMENTAL if self.sort_mode == "incremental" else _abi.SORT_FULL
        c.mcwf = int(self.mcwf)
        return c


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------
def quadrupole_trap(n_atoms: int, nc: int, T: float = 20e-6, dBdz: float = 2.5,
                    mu: float = 0.5 * MUB, widths: float = 5.0, n_real: float = N_REAL_REF) -> Trap:
    rth = KB * T / (mu * dBdz)
    hw = widths * rth
    return Trap(kind="quadrupole", n_atoms=int(n_atoms), nc=(nc, nc, nc), half_width=(hw, hw, hw),
                dBdz=dBdz, mu=mu, T=T, n_real=n_real)


def reference(Nc: int = 50, N: int = 10000) -> Run:
    """main.cu:45-69, 165-168 with the parameters of deviceSystemParameters.cuh."""
    trap = Trap(kind="ip", n_atoms=N, nc=(Nc, Nc, Nc), half_width=(1e-3, 1e-3, 1e-3),
                B0=0.01, dBdx=20., d2Bdx2=40000., mu=0.5 * GS * MUB, T=20e-6, n_real=N_REAL_REF)
    return Run(traps=[trap], precision="fp64", pair_mode="reference", spin=True, spin_substeps=1,
               spin_interp=False, spin_loss=True, mcwf=True, lpc=10, dt=1e-7, seed=1234,
               steps=int(math.ceil(5e-3 / 1e-7)), name=f"reference-{Nc}-{N}")


def quadrupole_1e4() -> Run:
    return Run(traps=[quadrupole_trap(10000, 35)], precision="fp64", dt=1e-6, steps=1000, seed=1,
               name="quadrupole-1e4-35c")


def quadrupole_1e6(precision: str = "fp32") -> Run:
    return Run(traps=[quadrupole_trap(1000000, 100)], precision=precision, dt=1e-6, steps=10000,
               seed=1, name="quadrupole-1e6-100c")


def box_thermalise(precision: str = "fp32", n_atoms: int = 1 << 22, nc: int = 128) -> Run:
    n_density = 1e19
    L = 1e-3
    Tx, Ty = 40e-6, 20e-6
    trap = Trap(kind="box", n_atoms=n_atoms, nc=(nc, nc, nc), half_width=(L / 2, L / 2, L / 2),
                T=(Tx + 2 * Ty) / 3, T_axis=(Tx, Ty, Ty), n_real=n_density * L ** 3, mu=0.0)
    return Run(traps=[trap], precision=precision, dt=1e-6, steps=10000, seed=1,
               name=f"box-thermalise-{n_atoms}")


def spin_evaporation(precision: str = "fp32", n_atoms: int = 4000000, nc: int = 100,
                     steps: int = 200000) -> Run:
    T0 = 50e-6
    trap = quadrupole_trap(n_atoms, nc, T=T0, n_real=float(n_atoms))
    trap.eta_start, trap.eta_end, trap.T_cut, trap.eta_steps = 10.0, 4.0, T0, 200000
    return Run(traps=[trap], precision=precision, spin=True, spin_substeps=20, spin_interp=True,
               spin_loss=True, energy_cut=True, dt=1e-6, steps=steps, seed=1,
               name=f"spin-evaporation-{n_atoms}")


def cells_for_trap(n_atoms: int, target_peak: float = 8.0, lo: int = 32, hi: int = 64) -> int:
    """Cells per dimension of a quadrupole trap gridded over +-5 thermal radii.

    The peak density is N / (4 pi r_th^3), so the central cell of a (10 r_th /
    Nc)^3 grid holds about 1000 N / (4 pi Nc^3) test particles.  Nc is chosen
    so that it holds `target_peak`, clamped to [lo, hi]."""
    nc = (1000.0 * n_atoms / (4.0 * math.pi * target_peak)) ** (1.0 / 3.0)
    return int(min(hi, max(lo, round(nc))))


def split_traps(run: Run, groups: int) -> List[Run]:
    """`run` as `groups` runs of consecutive traps with about equal atom
    counts, for stepping on one GPU as a group of contexts.  Every trap keeps
    its Philox stream (its index in `run` plus its philox_stream), so its
    results do not change."""
    import numpy as np

    T = len(run.traps)
    if not 1 <= groups <= T:
        raise ValueError(f"{groups} groups for {T} traps")
    n = np.array([t.n_atoms for t in run.traps], dtype=np.float64)
    cum = np.cumsum(n)
    cuts = [0]
    for g in range(1, groups):
        c = int(np.searchsorted(cum, cum[-1] * g / groups)) + 1
        cuts.append(min(max(c, cuts[-1] + 1), T - (groups - g)))
    cuts.append(T)
    out = []
    for a, b in zip(cuts, cuts[1:]):
        traps = [copy.deepcopy(t) for t in run.traps[a:b]]
        for t in traps:
            t.philox_stream = int(t.philox_stream) + a
        out.append(run.with_(traps=traps, name=f"{run.name}[{a}:{b}]"))
    return out


def batched_traps(n_traps: int = 256, seed: int = 1, steps: int = 5000) -> Run:
    import numpy as np

    rng = np.random.default_rng(seed)
    rng_cut = np.random.default_rng([seed, 1])
    traps = []
    for _ in range(n_traps):
        n = int(math.floor(math.exp(rng.uniform(math.log(1e3), math.log(1e5)))))
        dBdz = float(rng.uniform(1.0, 5.0))
        T = float(rng.uniform(5e-6, 50e-6))
        eta0 = float(rng.uniform(7.0, 10.0))
        # the ramp from a second stream, so that N, B' and T stay those of
        # the seed's first stream
        eta1 = eta0 - float(rng_cut.uniform(1.0, 2.0))
        ramp = int(rng_cut.uniform(1e4, 5e4))
        tr = quadrupole_trap(n, cells_for_trap(n), T=T, dBdz=dBdz, n_real=float(n))
        tr.eta_start, tr.eta_end, tr.T_cut, tr.eta_steps = eta0, eta1, T, ramp
        traps.append(tr)
    return Run(traps=traps, precision="fp32", energy_cut=True, dt=1e-6, steps=steps, seed=seed,
               name=f"batched-{n_traps}-traps")


WORKLOADS = {
    "reference": reference,
    "quadrupole_1e4": quadrupole_1e4,
    "quadrupole_1e6": quadrupole_1e6,
    "box_thermalise": box_thermalise,
    "spin_evaporation": spin_evaporation,
    "batched_traps": batched_traps,
}


def analytic_trap_rate(n_real: float, T: float, dBdz: float, mu: float = 0.5 * MUB,
                       sigma: float = SIGMA, mass: float = MRB) -> float:
    """Collision events per atom per second in a spherical quadrupole,
    R = N sigma (mu B')^3 / (16 pi^{3/2} sqrt(m) (kT)^{5/2}).  With
    mu = gs muB / 2 it equals collisionChecker.py:34."""
    return n_real * sigma * (mu * dBdz) ** 3 / (16.0 * PI_REF ** 1.5 * math.sqrt(mass) * (KB * T) ** 2.5)


def analytic_homogeneous_rate(n: float, T: float, sigma: float = SIGMA, mass: float = MRB) -> float:
    """Events per atom per second in a homogeneous gas, n sigma <v_r> / 2 with
    <v_r> = sqrt(16 kT / (pi m)) (homogeneousGasCollisions.py:33)."""
    return n * sigma * math.sqrt(16.0 * KB * T / (PI_REF * mass)) / 2.0

"""GPU checks of the spin-and-evaporation workload (FP32 Kahan form with FP64
spin renormalisation).

* Full size: N = 4e6 atoms in the B' = 2.5 T/m quadrupole at 50 uK, spins
  along the local field, 20 spin substeps of 5e-8 s per 1 us step at the
  substep midpoints, spin-projection loss (evaporation.cu:90-94), an energy
  cut ramped from 10 to 4 kB T0 over 2000 steps (the workload's 2e5 steps
  shortened so that the ramp completes), collisions on; 2500 steps.  Energy
  bookkeeping (alive + removed) |dE/E| <= 1e-4, |s| = 1 within 1e-6, and every
  survivor on the kept side of both predicates.
* Exact predicates: one step from a given state with spins tilted from the
  field by angles uniform in [0, pi] and a cut at 5 kB T, no collisions: the
  removed set equals the oracle's (its FP64 spinor update, wavefunction.cu:
  39-73, Verlet move, the projection of evaporation.cu:90-92 and the energy
  E = m v^2 / 2 + mu B' rho > 5 kB T) for every atom farther than 1e-4 from
  the projection threshold and 1e-6 (relative) from the energy threshold.
"""
import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs

pytestmark = pytest.mark.gpu


def quad_dir(trap, pos):
    """Unit field direction of the spherical quadrupole, B ~ (x, y, -2 z)."""
    b = np.stack([pos[:, 0], pos[:, 1], -2.0 * pos[:, 2]], axis=1)
    return b / np.linalg.norm(b, axis=1)[:, None]


def quad_energy(run, trap, pos, vel):
    rho = np.sqrt(pos[:, 0] ** 2 + pos[:, 1] ** 2 + 4.0 * pos[:, 2] ** 2)
    return 0.5 * run.mass * np.sum(vel * vel, axis=1) + trap.mu * trap.dBdz * rho


def bloch_of(up, dn):
    return np.stack([2 * (up[:, 0] * dn[:, 0] + up[:, 1] * dn[:, 1]),
                     2 * (up[:, 0] * dn[:, 1] - up[:, 1] * dn[:, 0]),
                     (up ** 2).sum(1) - (dn ** 2).sum(1)], axis=1)


def spinor_of(s):
    """(psi_up, psi_dn) with Bloch vector s: (cos th/2, e^{i phi} sin th/2)."""
    th = np.arccos(np.clip(s[:, 2], -1.0, 1.0))
    ph = np.arctan2(s[:, 1], s[:, 0])
    up = np.stack([np.cos(0.5 * th), np.zeros(len(s))], axis=1)
    dn = np.stack([np.sin(0.5 * th) * np.cos(ph), np.sin(0.5 * th) * np.sin(ph)], axis=1)
    return np.ascontiguousarray(up), np.ascontiguousarray(dn)


def test_spin_evaporation_full_size():
    run = configs.spin_evaporation("fp32")
    tr = run.traps[0]
    assert tr.n_atoms == 4000000 and run.spin_substeps == 20 and run.spin_loss and run.energy_cut
    tr.eta_steps = 2000
    steps = 2500
    print(f"spin + evaporation N={tr.n_atoms}, {steps} steps, ramp {tr.eta_start}->{tr.eta_end} kT "
          f"over {tr.eta_steps}", flush=True)
    with Simulation(run.with_(obs_capacity=steps + 16)) as sim:
        sim.init_distribution()
        for k in range(5):
            sim.step(steps // 5)
            print(f"  step {(k + 1) * steps // 5}", flush=True)
        lg = sim.obs_log()[:, 0]
        pos = sim.copy_out(_abi.FIELD_POS)
        vel = sim.copy_out(_abi.FIELD_VEL)
        spin = sim.copy_out(_abi.FIELD_SPIN)
        n_alive = sim.info().n_alive
    E = lg["ek"] + lg["ep"] + np.cumsum(lg["e_evap"])
    drift = float(np.max(np.abs(E - E[0]) / abs(E[0])))
    removed = int(lg["n_evap"].sum())
    print(f"|dE/E| {drift:.2e}; removed {removed}; alive {n_alive}; collisions "
          f"{int(lg['collisions'].sum())}", flush=True)
    assert drift <= 1e-4
    assert int(lg["collisions"].sum()) > 0
    assert int(lg["n_evap"][100:2000].sum()) > 0          # the ramp keeps removing atoms
    assert n_alive == len(pos) == int(lg["n_alive"][-1])
    assert np.all(np.diff(lg["n_alive"]) <= 0)
    norm_err = float(np.max(np.abs(np.linalg.norm(spin, axis=1) - 1.0)))
    print(f"max ||s| - 1| {norm_err:.2e}", flush=True)
    assert norm_err <= 1e-6
    # survivors of the last step: below the final cut and spin along B
    cut = tr.eta_end * configs.KB * tr.T_cut
    assert float(quad_energy(run, tr, pos, vel).max()) <= cut * (1 + 1e-6)
    assert float(np.min(np.sum(quad_dir(tr, pos) * spin, axis=1))) >= -1e-5


def test_loss_predicates_one_step_exact(orc):
    n = 40000
    run = configs.spin_evaporation("fp32", n_atoms=n, nc=40).with_(
        sigma=0.0, spin_interp=False, obs_capacity=64)
    tr = run.traps[0]
    tr.eta_start = tr.eta_end = 5.0
    tr.eta_steps = 0
    run64 = run.with_(precision="fp64", energy_cut=False)
    p = orc.params_from(run64)
    st = orc.init_distribution(p, n, tr.T)
    pos, vel = st["pos"], st["vel"]
    # spins tilted from B by angles uniform in [0, pi] about a random axis
    rng = np.random.default_rng(17)
    b = quad_dir(tr, pos)
    e = np.cross(b, rng.normal(size=(n, 3)))
    e /= np.linalg.norm(e, axis=1)[:, None]
    th = rng.uniform(0.0, np.pi, n)
    s0 = np.cos(th)[:, None] * b + np.sin(th)[:, None] * e
    print("loss predicates, one step", flush=True)
    with Simulation(run) as sim:
        sim.set_state(pos, vel)
        sim.set_spin(s0)
        sim.step(1)
        alive = np.zeros(n, bool)
        alive[sim.copy_out(_abi.FIELD_ATOM_ID)] = True
    # the oracle: spinor substeps at the pre-move position, the move, the
    # spin-projection tag; the energy cut on its post-move state
    ref = orc.Reference(run64, pos=pos, vel=vel)
    up, dn = spinor_of(s0)
    ref.s["psi_up"], ref.s["psi_dn"] = up, dn
    ref.step(1)
    spin_kept = np.zeros(n, bool)
    spin_kept[ref.s["atom_id"]] = True
    s1 = bloch_of(ref.s["psi_up"], ref.s["psi_dn"])
    proj = np.sum(quad_dir(tr, ref.s["pos"]) * s1, axis=1)
    E1 = quad_energy(run, tr, ref.s["pos"], ref.s["vel"])
    cut = 5.0 * configs.KB * tr.T
    expect_alive = spin_kept & (E1 <= cut)
    far = (np.abs(proj) > 1e-4) & (np.abs(E1 - cut) > 1e-6 * cut)
    # the oracle's own projection test agrees with the formula away from 0
    assert np.array_equal(spin_kept[far], proj[far] >= 0)
    lost_spin = int((~spin_kept & far).sum())
    lost_energy = int((spin_kept & (E1 > cut) & far).sum())
    mismatch = int((alive[far] != expect_alive[far]).sum())
    print(f"lost by spin {lost_spin}, by energy {lost_energy}, near a threshold "
          f"{int((~far).sum())}, mismatches {mismatch}", flush=True)
    assert lost_spin > 1000 and lost_energy > 1000
    assert mismatch == 0

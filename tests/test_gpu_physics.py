"""GPU checks of the physics against kinetic theory and the reference's own
checker formulas, for both numerical forms:

* homogeneous-gas collision rate n sigma <v_r> / 2 per atom
  (homogeneousGasCollisions.py:33);
* quadrupole-trap collision rate N sigma (mu B')^3 / (16 pi^1.5 sqrt(m) (kT)^2.5)
  per atom (collisionChecker.py:34 with the moment explicit);
* relaxation of a temperature anisotropy Tx = 2 Ty = 2 Tz in a periodic box:
  d(Tx - T)/dt = -(4/5) n sigma <v_r> (Tx - T) in the linear regime
  (hard spheres, isotropic scattering), checked on the late-time decay;
* energy conservation |dE/E| <= 1e-4 with collisions (energyChecker.py:170,
  185 uses 1e-5 over its shorter runs).

Rates are measured after the per-cell sigma v_r,max has converged (its start
value, sigma <v_r>, under-counts pairs faster than it until trials have seen
them, exactly as in the reference).
"""
import math

import numpy as np
import pytest

from cudadsmc_amd import Simulation, configs

pytestmark = pytest.mark.gpu


def box_run(precision, n_atoms, nc, T=20e-6, T_axis=(0.0, 0.0, 0.0), n_density=1e19, seed=21):
    L = 1e-3
    trap = configs.Trap(kind="box", n_atoms=n_atoms, nc=(nc, nc, nc),
                        half_width=(L / 2, L / 2, L / 2), T=T, T_axis=T_axis,
                        n_real=n_density * L ** 3, mu=0.0)
    return configs.Run(traps=[trap], precision=precision, dt=1e-6, seed=seed,
                       obs_capacity=4096, name="box")


def measured_rate(sim, warm, steps, dt):
    sim.step(warm)
    sim.step(steps)
    lg = sim.obs_log()[-steps:, 0]
    coll = int(lg["collisions"].sum())
    n = float(lg["n_alive"].mean())
    return coll / (n * steps * dt), coll


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_homogeneous_collision_rate(precision):
    run = box_run(precision, 1 << 20, 64)
    expect = configs.analytic_homogeneous_rate(1e19, 20e-6)
    print(f"homogeneous rate {precision}", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        rate, coll = measured_rate(sim, 300, 300, run.dt)
    err = 1.0 / math.sqrt(coll)
    print(f"rate {rate:.2f} /s vs {expect:.2f} /s ({coll} events)", flush=True)
    assert abs(rate / expect - 1.0) < 4 * err + 0.01


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_trap_collision_rate(precision):
    n, n_real, T, dBdz = 200000, 1e8, 20e-6, 2.5
    trap = configs.quadrupole_trap(n, 100, T=T, dBdz=dBdz, n_real=n_real)
    run = configs.Run(traps=[trap], precision=precision, dt=1e-6, seed=22, obs_capacity=4096)
    expect = configs.analytic_trap_rate(n_real, T, dBdz, mu=trap.mu)
    print(f"trap rate {precision}", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        rate, coll = measured_rate(sim, 200, 200, run.dt)
    err = 1.0 / math.sqrt(coll)
    print(f"rate {rate:.2f} /s vs {expect:.2f} /s ({coll} events)", flush=True)
    assert abs(rate / expect - 1.0) < 4 * err + 0.03


def test_box_thermalisation_time():
    """Homogeneous-gas configuration: N = 2^22 atoms, 1 mm^3 periodic box at
    n = 1e19 m^-3, 128^3 cells, Tx = 40 uK, Ty = Tz = 20 uK.  The decay time
    of Tx - Ty over [1, 3.5] tau must be tau = 5 / (4 n sigma <v_r>) within 10 %."""
    run = configs.box_thermalise("fp32").with_(obs_capacity=8192)
    t = run.traps[0]
    n_density = t.n_real / (2 * t.half_width[0]) ** 3
    vr = math.sqrt(16 * configs.KB * t.T / (configs.PI_REF * run.mass))
    tau = 5.0 / (4.0 * n_density * run.sigma * vr)
    steps = int(4.0 * tau / run.dt)
    print(f"thermalisation: tau = {tau * 1e3:.3f} ms, {steps} steps", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(steps)
        lg = sim.obs_log()[-steps:, 0]
    v2 = lg["v2"]
    Tx = run.mass * v2[:, 0] / (configs.KB * lg["n_alive"])
    Ty = run.mass * 0.5 * (v2[:, 1] + v2[:, 2]) / (configs.KB * lg["n_alive"])
    time = (np.arange(steps) + 1) * run.dt
    d = Tx - Ty
    win = (time > 1.0 * tau) & (time < 3.5 * tau)
    slope = np.polyfit(time[win], np.log(d[win]), 1)[0]
    tau_fit = -1.0 / slope
    print(f"tau fit {tau_fit * 1e3:.3f} ms vs {tau * 1e3:.3f} ms; "
          f"T_mean drift {abs((Tx + 2 * Ty)[-1] / (Tx + 2 * Ty)[0] - 1):.2e}", flush=True)
    assert abs(tau_fit / tau - 1.0) < 0.10
    # energy is conserved: the mean temperature stays put
    Tm = (Tx + 2 * Ty) / 3
    assert abs(Tm[-1] / Tm[0] - 1.0) < 1e-3


def test_fp64_plain_energy_conservation():
    """FP64 drop-in form, N = 1e5, 1e4 steps with collisions: |dE/E| <= 1e-4."""
    run = configs.Run(traps=[configs.quadrupole_trap(100000, 60)], precision="fp64", dt=1e-6,
                      seed=23, obs_capacity=10050)
    print("fp64 energy conservation, 1e4 steps", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(10000)
        lg = sim.obs_log()[-10000:, 0]
    E = lg["ek"] + lg["ep"]
    drift = float(np.max(np.abs(E - E[0])) / E[0])
    print(f"max |dE/E| = {drift:.3e}, collisions {int(lg['collisions'].sum())}", flush=True)
    assert drift <= 1e-4
    assert int(lg["collisions"].sum()) > 0

"""GPU checks of the batched workload at its full size: 256 independent
quadrupole traps, N log-uniform in [1e3, 1e5] (5 220 993 atoms), B' ~ U[1, 5]
T/m, T ~ U[5, 50] uK, 32^3..64^3 cells per trap and a per-trap energy-cut
schedule (eta ramping down over 1e4..5e4 steps), FP32 Kahan form, 1e4
steps replayed from the HIP graph.

* Energy: |dE/E| <= 1e-4 for every trap, counting the energy of the removed
  atoms (the form of energyChecker.py:170 with evaporation).
* Collisions: for the 20 traps with the most events, the collisions counted
  in five 2000-step windows equal the NTC expectation of the state at the
  window's middle, c1 sigma sum over cells and pairs of |v_i - v_j|
  (collisions.cu:312-322: Ncol trials, each accepted with probability
  alpha lpc dt sigma v_r / (Vc lambda)), within 4 sigma + 3 % per trap and
  over the 20 together; and, with the cut off, the rate of the same traps
  equals the analytic trap rate N sigma (mu B')^3 / (16 pi^1.5 sqrt(m)
  (kT)^2.5) (collisionChecker.py:34), within the same bounds.  (With the cut
  on, the truncated cloud is not the Boltzmann one the analytic rate
  assumes: 3 to 40 % off at 10 to 6 kT.)  The workload is nearly
  collisionless (alpha = 1: a few events per atom per second), so the
  busiest trap sees about 2000 events in 1e4 steps.
* Independence: five sampled traps equal, bitwise, the same trap run alone
  for the same 1e4 steps.
"""
import copy
import math

import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs

pytestmark = pytest.mark.gpu

STEPS = 10000


def pair_speed_sum(pos, vel, trap):
    """Sum over the trap's grid cells and unordered pairs of |v_i - v_j|
    (cells as indexAtoms forms them; atoms outside the grid collide with
    nobody)."""
    hw = np.asarray(trap.half_width, float)
    nc = np.asarray(trap.nc, int)
    idx = np.floor((pos + hw) / (2 * hw / nc)).astype(np.int64)
    ok = np.all((idx >= 0) & (idx < nc), axis=1)
    key = (idx[ok, 2] * nc[1] + idx[ok, 1]) * nc[0] + idx[ok, 0]
    v = vel[ok]
    order = np.argsort(key, kind="stable")
    key, v = key[order], v[order]
    starts = np.r_[0, np.flatnonzero(np.diff(key)) + 1]
    sizes = np.diff(np.r_[starts, len(key)])
    total = 0.0
    for k in np.unique(sizes[sizes >= 2]):
        s = starts[sizes == k]
        g = v[s[:, None] + np.arange(k)[None, :]]                  # (cells, k, 3)
        d = np.linalg.norm(g[:, :, None, :] - g[:, None, :, :], axis=3)
        total += d.sum() / 2.0
    return total


def expected_collisions_per_step(pos, vel, trap, run):
    """NTC expectation of one collision step from this state."""
    vc = np.prod(2 * np.asarray(trap.half_width) / np.asarray(trap.nc))
    alpha = trap.n_real / trap.n_atoms
    c1s = alpha * run.lpc * run.dt / vc * run.sigma
    return c1s * pair_speed_sum(pos, vel, trap)


def test_batched_workload_full_size():
    run = configs.batched_traps()
    assert len(run.traps) == 256 and run.n_atoms == 5220993
    assert all(t.eta_start > t.eta_end and t.eta_steps > 0 for t in run.traps)
    snaps = (1000, 3000, 5000, 7000, 9000)
    half = 1000
    states = {}
    print(f"batched workload: {run.n_atoms} atoms, {STEPS} steps", flush=True)
    with Simulation(run.with_(obs_capacity=STEPS + 16)) as sim:
        sim.init_distribution()
        done = 0
        for s in snaps:
            sim.step(s - done)
            done = s
            ts = sim.copy_out(_abi.FIELD_TRAP_START)
            states[s] = (ts, sim.copy_out(_abi.FIELD_POS), sim.copy_out(_abi.FIELD_VEL))
            print(f"  step {s}", flush=True)
        sim.step(STEPS - done)
        lg = sim.obs_log()
        final = [sim.copy_out(f) for f in (_abi.FIELD_TRAP_START, _abi.FIELD_POS, _abi.FIELD_VEL,
                                           _abi.FIELD_ATOM_ID)]
    assert lg.shape == (STEPS, 256)

    # energy of every trap, alive plus removed
    E = lg["ek"] + lg["ep"] + np.cumsum(lg["e_evap"], axis=0)
    drift = np.max(np.abs(E - E[0]) / np.abs(E[0]), axis=0)
    removed = lg["n_evap"][1:].sum(axis=0)     # after the first step's cut of the tail
    print(f"max |dE/E| {drift.max():.2e} (median {np.median(drift):.2e}); traps losing atoms "
          f"after step 0: {int((removed > 0).sum())}", flush=True)
    assert drift.max() <= 1e-4
    assert int((removed > 0).sum()) >= 200     # the cuts ramp down through the tail

    # collisions in each window against the NTC expectation of its state
    coll = lg["collisions"]
    busiest = np.argsort(coll.sum(axis=0))[-20:]
    tot_obs = tot_exp = 0.0
    for t in busiest:
        tr = run.traps[t]
        obs = exp = 0.0
        for s, (ts, pos, vel) in states.items():
            sl = slice(ts[t], ts[t + 1])
            exp += expected_collisions_per_step(pos[sl], vel[sl], tr, run) * 2 * half
            obs += float(coll[s - half:s + half, t].sum())
        print(f"  trap {t}: N0 {tr.n_atoms}, collisions {obs:.0f} vs NTC expectation {exp:.1f}",
              flush=True)
        assert obs > 100
        assert abs(obs / exp - 1.0) < 4.0 / math.sqrt(obs) + 0.03
        tot_obs += obs
        tot_exp += exp
    print(f"  20 traps: {tot_obs:.0f} vs {tot_exp:.1f}", flush=True)
    assert abs(tot_obs / tot_exp - 1.0) < 4.0 / math.sqrt(tot_obs) + 0.03

    # five sampled traps against the same trap run alone
    ts, pos, vel, ids = final
    begin = np.cumsum([0] + [t.n_atoms for t in run.traps])
    rng = np.random.default_rng(3)
    for t in sorted(rng.choice(256, 5, replace=False)):
        tr = copy.deepcopy(run.traps[t])
        tr.philox_stream = int(t)
        solo = run.with_(traps=[tr], obs_capacity=STEPS + 16)
        print(f"  trap {t} alone ({tr.n_atoms} atoms)", flush=True)
        with Simulation(solo) as s1:
            s1.init_distribution()
            s1.step(STEPS)
            l1 = s1.obs_log()
            sl = slice(ts[t], ts[t + 1])
            np.testing.assert_array_equal(ids[sl] - begin[t], s1.copy_out(_abi.FIELD_ATOM_ID))
            np.testing.assert_array_equal(pos[sl], s1.copy_out(_abi.FIELD_POS))
            np.testing.assert_array_equal(vel[sl], s1.copy_out(_abi.FIELD_VEL))
        for f in ("collisions", "n_evap", "n_alive"):
            np.testing.assert_array_equal(lg[f][:, t], l1[f][:, 0])
        # the FP64 sums run over other block boundaries
        np.testing.assert_allclose(lg["ek"][:, t], l1["ek"][:, 0], rtol=1e-12)


def test_batched_collision_rate_analytic():
    """The same 256 traps without the cut: the rate of the 20 busiest traps
    over steps 300..10300 against the analytic trap rate."""
    warm, steps = 300, 10000
    run = configs.batched_traps().with_(energy_cut=False, obs_capacity=warm + steps + 16)
    print("batched traps, no cut: collision rates", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(warm + steps)
        lg = sim.obs_log()[-steps:]
    coll = lg["collisions"].sum(axis=0)
    assert int(lg["n_evap"].sum()) == 0
    tot_obs = tot_exp = 0.0
    for t in np.argsort(coll)[-20:]:
        tr = run.traps[t]
        n = float(lg["n_alive"][:, t].mean())
        expect = configs.analytic_trap_rate(tr.n_real, tr.T, tr.dBdz, mu=tr.mu, sigma=run.sigma,
                                            mass=run.mass)
        rate = coll[t] / (n * steps * run.dt)
        print(f"  trap {t}: rate {rate:.3f} /s vs {expect:.3f} /s ({coll[t]} events)", flush=True)
        assert coll[t] > 100
        assert abs(rate / expect - 1.0) < 4.0 / math.sqrt(coll[t]) + 0.03
        tot_obs += coll[t]
        tot_exp += expect * n * steps * run.dt
    print(f"  20 traps: {tot_obs:.0f} vs {tot_exp:.1f}", flush=True)
    assert abs(tot_obs / tot_exp - 1.0) < 4.0 / math.sqrt(tot_obs) + 0.03

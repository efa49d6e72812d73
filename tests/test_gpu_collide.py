"""The FP32 collide (k_collide_fused, collisions.cu:291-445) pinned on one
collision step from a given state.

A periodic box of 1e5 atoms in 40^3 cells is copied into 64 traps of one
context; trap t draws from Philox stream t, so one dsmc_collide gives 64
independent collision steps of the same state.  The density (alpha = 1e6 real
atoms per test particle) keeps every cell at one trial per step
(ceil(X) = 1, X < 0.25 even in the fullest cell, so the acceptance
probability stays below 1), so a cell's two changed atoms are the pair of
its accepted collision.

* Per pair: total momentum and |v_a - v_b| are conserved to FP64 rounding
  of value + remainder (collisions.cu:426-436).
* Per cell, against the oracle's FP64 collide (orc_collide) on the same
  sorted state and the same Philox stream: the same trials (pair and
  uniform), so the same cells accept except where the acceptance
  probability and the uniform lie closer than the FP32 rounding; no such
  cell is expected, one in 10^4 is allowed.
* Per cell, against the NTC acceptance probability of the state,
  p = alpha lpc dt sigma / Vc * sum over pairs |v_a - v_b| (one trial:
  collisions.cu:318-322), summed over the 64 streams: chi-square over 20
  bins of equal expectation below 50 (p ~ 2e-4 for 20 degrees of freedom),
  and the total within 4 sigma.
"""
import copy

import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs

pytestmark = pytest.mark.gpu

S, N, NC, ALPHA = 64, 100000, 40, 1e6


def box_trap():
    L = 1e-3
    return configs.Trap(kind="box", n_atoms=N, nc=(NC, NC, NC), half_width=(L / 2,) * 3, T=20e-6,
                        mu=0.0, n_real=ALPHA * N)


def test_fp32_collide_pairs_oracle_and_acceptance(orc):
    trap = box_trap()
    one = configs.Run(traps=[trap], precision="fp64", dt=1e-6, seed=31)
    p0 = orc.params_from(one)
    st = orc.init_distribution(p0, N, trap.T)
    pos, vel = st["pos"], st["vel"]
    run = configs.Run(traps=[copy.deepcopy(trap) for _ in range(S)], precision="fp32", dt=1e-6,
                      seed=31, obs_capacity=16)
    print(f"FP32 collide: {S} streams x {N} atoms, {NC}^3 cells", flush=True)
    with Simulation(run) as sim:
        sim.set_state(np.tile(pos, (S, 1)), np.tile(vel, (S, 1)))
        keys = sim.copy_out(_abi.FIELD_CELL_ID).astype(np.int64)
        v0 = sim.copy_out(_abi.FIELD_VEL)
        assert np.array_equal(sim.copy_out(_abi.FIELD_ATOM_ID), np.arange(S * N))
        sim.sort_arrays()
        sim.build_cell_table()
        sim.collide()
        v1 = sim.copy_out(_abi.FIELD_VEL)
    ncell = NC ** 3
    local = (keys - np.repeat(np.arange(S) * (ncell + 1), N)).reshape(S, N)
    assert np.all(local == local[0])          # the same cells in every copy
    cell = local[0]
    v0 = v0.reshape(S, N, 3)
    v1 = v1.reshape(S, N, 3)
    changed = np.any(v1 != v0, axis=2)        # (S, N)

    # one trial per cell: X = Mc c1 sigma v_r,max0 < 0.25 in the fullest cell
    occ = np.bincount(cell, minlength=ncell + 1)[:ncell]
    vc = (1e-3 / NC) ** 3
    c1 = ALPHA * run.lpc * run.dt / vc
    smax0 = float(np.float32(run.sigma * np.sqrt(16 * configs.KB * trap.T / (configs.PI_REF * run.mass))))
    kmax = int(occ.max())
    assert kmax * (kmax - 1) / 2 * c1 * smax0 < 0.25

    # accepted collisions per stream and cell: 0 or 2 changed atoms
    acc = np.zeros((S, ncell), np.int64)
    for t in range(S):
        acc[t] = np.bincount(cell[changed[t]], minlength=ncell + 1)[:ncell]
    assert set(np.unique(acc)) <= {0, 2}
    acc //= 2
    total = int(acc.sum())
    print(f"accepted collisions: {total}", flush=True)
    assert total > 10000

    # per pair: momentum and relative speed
    order = np.argsort(cell, kind="stable")
    for t in range(S):
        idx = order[changed[t][order]]        # changed atoms, grouped by cell
        a, b = idx[0::2], idx[1::2]
        assert np.array_equal(cell[a], cell[b])
        pa, pb = v0[t, a] + v0[t, b], v1[t, a] + v1[t, b]
        ra = np.linalg.norm(v0[t, a] - v0[t, b], axis=1)
        rb = np.linalg.norm(v1[t, a] - v1[t, b], axis=1)
        scale = np.abs(v0[t]).max()
        assert np.abs(pa - pb).max() < 1e-13 * scale
        assert np.abs(ra - rb).max() < 1e-13 * scale

    # the oracle's FP64 collide on the same sorted state, streams 0..3
    mism = 0
    multi = 0
    for t in range(4):
        p = orc.params_from(run, t)
        ids = np.arange(N, dtype=np.int32)
        c = orc.index_atoms(p, pos, ids)
        np.testing.assert_array_equal(c, cell)
        c, ids = orc.sort_arrays(c, ids, ncell + 1)
        _, count, prefix = orc.cell_table(c, ncell)
        v = vel.copy()
        sv = np.full(ncell + 1, orc.sigvrmax_init(p, trap.T))
        counts, tot = orc.collide(p, v, sv, prefix, ids, 0)
        mism += int(np.count_nonzero(counts[:ncell] != acc[t]))
        multi += int(np.count_nonzero(count[:ncell] >= 2))
        if t == 0:
            print(f"stream 0: oracle {tot} collisions, FP32 {int(acc[0].sum())}", flush=True)
    print(f"cells with another outcome than the oracle: {mism} of {multi}", flush=True)
    assert mism <= max(1, multi // 10000)

    # against the NTC acceptance probability of each cell
    vs = vel[order]
    cs = cell[order]
    starts = np.r_[0, np.flatnonzero(np.diff(cs)) + 1]
    sizes = np.diff(np.r_[starts, len(cs)])
    pc = np.zeros(ncell)
    c1s = c1 * run.sigma
    for k in np.unique(sizes[sizes >= 2]):
        sel = starts[sizes == k]
        g = vs[sel[:, None] + np.arange(k)[None, :]]
        d = np.linalg.norm(g[:, :, None, :] - g[:, None, :, :], axis=3).sum(axis=(1, 2)) / 2.0
        pc[cs[sel]] = c1s * d
    obs = acc.sum(axis=0).astype(float)
    exp = S * pc
    var = S * pc * (1 - pc)
    m = pc > 0
    srt = np.argsort(pc[m])
    cum = np.cumsum(exp[m][srt])
    bins = np.minimum((cum / cum[-1] * 20).astype(int), 19)
    O = np.bincount(bins, weights=obs[m][srt], minlength=20)
    E = np.bincount(bins, weights=exp[m][srt], minlength=20)
    V = np.bincount(bins, weights=var[m][srt], minlength=20)
    chi2 = float(np.sum((O - E) ** 2 / V))
    print(f"chi2 over 20 bins {chi2:.1f}; total {obs.sum():.0f} vs {exp.sum():.1f}", flush=True)
    assert chi2 < 50.0
    assert abs(obs.sum() - exp.sum()) < 4 * np.sqrt(var.sum())
    assert obs[~m].sum() == 0

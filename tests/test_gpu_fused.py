"""GPU checks of the FP32 Kahan form (single and batched traps).

Exact checks: FP64 cell ids of the stored positions, the sort permutation
(numpy stable argsort), the occupied-cell table, HIP-graph replay against
eager launches (bitwise), and a trap inside a batch against the same trap run
alone (bitwise: each trap has its own Philox stream and the work of different
traps never mixes).  Tolerance checks against the FP64 path and the oracle:
collision-free trajectories, energy conservation |dE/E| <= 1e-4 over 1e4
steps, and the energy-cut survivor set.
"""
import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs

pytestmark = pytest.mark.gpu


def quad(n, nc, precision="fp32", **kw):
    run = configs.Run(traps=[configs.quadrupole_trap(n, nc)], precision=precision, dt=1e-6,
                      seed=3, name="fused-test")
    return run.with_(**kw) if kw else run


def by_label(sim, field, n=None):
    ids = sim.copy_out(_abi.FIELD_ATOM_ID)
    a = sim.copy_out(field)
    out = np.full(((n or sim.run.n_atoms),) + a.shape[1:], np.nan)
    out[ids] = a
    return out


def test_fused_index_sort_table_exact(orc):
    run = quad(30000, 40)
    print("fused index/sort/table", flush=True)
    ref = orc.Reference(run.with_(precision="fp64"))
    with Simulation(run) as sim:
        sim.set_state(ref.s["pos"], ref.s["vel"])
        pos = sim.copy_out(_abi.FIELD_POS)  # float + remainder, as FP64
        keys = sim.copy_out(_abi.FIELD_CELL_ID)
        # the labels travel in the position remainders, which keep 13
        # significant bits: positions to 2^-37 relative, velocities exact
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_ATOM_ID), np.arange(run.n_atoms))
        assert np.all(np.abs(pos - ref.s["pos"]) <= 2.0 ** -36 * np.abs(ref.s["pos"]))
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_VEL),
                                      ref.s["vel"].astype(np.float32).astype(np.float64)
                                      + (ref.s["vel"] - ref.s["vel"].astype(np.float32))
                                      .astype(np.float32))
        expect = orc.index_atoms(ref.p, pos, np.arange(run.n_atoms, dtype=np.int32))
        np.testing.assert_array_equal(keys, expect)
        sim.sort_arrays()
        order = np.argsort(expect, kind="stable")
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_PERM), order)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_ID), expect[order])
        sim.build_cell_table()
        se = sim.copy_out(_abi.FIELD_CELL_START_END)
        se_o, count, _ = orc.cell_table(expect[order], orc.ncells(ref.p))
        occ = count > 0
        np.testing.assert_array_equal(se[: len(count)][occ], se_o[occ])
        ts = sim.copy_out(_abi.FIELD_TRAP_START)
        assert list(ts) == [0, run.n_atoms]
        sim.collide()
        sim.move_atoms()


def test_move_keys_equal_fp64_index(orc):
    """The cell keys the move writes (FP32 quotient, exact FP64 formula within
    1e-4 of a face) equal the oracle's FP64 index of the stored positions:
    every trap of a batch after some steps, and the golden fixture of atoms
    on faces, one ulp either side, at the grid edges, outside, NaN and inf
    (moved with dt = 1e-30 s, which leaves every position as it is)."""
    from pathlib import Path

    run = _batch_run()
    key_base = np.cumsum([0] + [int(np.prod(t.nc)) + 1 for t in run.traps])
    print("move keys vs FP64 index: batch", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(25)
        keys = sim.copy_out(_abi.FIELD_CELL_ID).astype(np.int64)
        pos = sim.copy_out(_abi.FIELD_POS)
        ts = sim.copy_out(_abi.FIELD_TRAP_START)
    for t in range(len(run.traps)):
        sl = slice(ts[t], ts[t + 1])
        p = orc.params_from(run, t)
        expect = orc.index_atoms(p, np.ascontiguousarray(pos[sl]), np.arange(ts[t + 1] - ts[t], dtype=np.int32))
        np.testing.assert_array_equal(keys[sl] - key_base[t], expect)

    fx = np.load(Path(__file__).resolve().parent / "golden" / "index_fixture.npz")
    base = configs.quadrupole_1e4()
    trap = base.traps[0]
    trap.n_atoms = len(fx["pos"])
    run = base.with_(traps=[trap], precision="fp32", dt=1e-30, obs_capacity=16)
    print("move keys vs FP64 index: face fixture", flush=True)
    with Simulation(run) as sim:
        sim.set_state(fx["pos"], np.zeros_like(fx["vel"]))
        stored = sim.copy_out(_abi.FIELD_POS)
        sim.step(1)
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        keys = sim.copy_out(_abi.FIELD_CELL_ID)
        after = sim.copy_out(_abi.FIELD_POS)
    fin = np.all(np.isfinite(stored[ids]), axis=1)   # inf positions turn NaN when moved
    np.testing.assert_array_equal(after[fin], stored[ids][fin])
    p = orc.params_from(run)
    expect = orc.index_atoms(p, np.ascontiguousarray(stored), np.arange(len(stored), dtype=np.int32))
    np.testing.assert_array_equal(keys, expect[ids])
    assert int(np.sum(expect == orc.ncells(p))) > 0   # atoms outside the grid


# tolerances (position over the thermal radius for the quadrupole, over the
# cloud's extent otherwise; velocity over the largest speed) after 2000
# collision-free steps from the same state
MOTION_TOL = {"quadrupole": (1e-6, 1e-5), "ip": (1e-7, 1e-6), "box": (1e-6, 1e-12)}


@pytest.mark.parametrize("kind", ["quadrupole", "ip", "box"])
def test_fused_motion_matches_fp64(orc, kind):
    """Collision-free motion (sigma = 0): FP32 Kahan vs the FP64 path, in
    each trap kind: the quadrupole, the Ioffe-Pritchard field of the
    reference (moveAtoms.cu:93-157) and the periodic box."""
    if kind == "quadrupole":
        run = quad(50000, 40, sigma=0.0)
    elif kind == "ip":
        run = configs.reference(Nc=40, N=50000).with_(
            precision="fp32", sigma=0.0, spin=False, spin_loss=False, mcwf=False, lpc=1, dt=1e-6)
    else:
        run = configs.box_thermalise("fp32", n_atoms=50000, nc=32).with_(sigma=0.0)
    ref = orc.Reference(run.with_(precision="fp64"))
    print(f"fused motion vs fp64, {kind}, 2000 steps", flush=True)
    with Simulation(run) as s32, Simulation(run.with_(precision="fp64")) as s64:
        s32.set_state(ref.s["pos"], ref.s["vel"])
        s64.set_state(ref.s["pos"], ref.s["vel"])
        s32.step(2000)
        s64.step(2000)
        p32, p64 = by_label(s32, _abi.FIELD_POS), by_label(s64, _abi.FIELD_POS)
        v32, v64 = by_label(s32, _abi.FIELD_VEL), by_label(s64, _abi.FIELD_VEL)
    # the quadrupole over its thermal radius; the others over the cloud
    scale = run.traps[0].thermal_radius() if kind == "quadrupole" else np.abs(p64).max()
    dx = np.abs(p32 - p64).max() / scale
    dv = np.abs(v32 - v64).max() / np.abs(v64).max()
    print(f"{kind}: max |dx| / scale = {dx:.3e}, max |dv| / v = {dv:.3e}", flush=True)
    tx, tv = MOTION_TOL[kind]
    assert dx < tx and dv < tv


@pytest.mark.parametrize("n,nc,steps", [(1000000, 100, 10000)])
def test_fused_energy_conservation(n, nc, steps):
    """|dE/E| <= 1e-4 over 1e4 steps with collisions on (N = 1e6, 100^3 cells)."""
    run = configs.quadrupole_1e6("fp32")
    print(f"energy conservation N={n} {nc}^3 {steps} steps", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        e0 = sim.observe()[0]
        E0 = e0.ek + e0.ep
        drift = 0.0
        coll = 0
        for chunk in range(steps // 1000):
            sim.step(1000)
            log = sim.obs_log()[-1000:, 0]
            E = log["ek"] + log["ep"]
            drift = max(drift, float(np.max(np.abs(E - E0)) / E0))
            coll += int(log["collisions"].sum())
        print(f"max |dE/E| = {drift:.3e}, collisions = {coll}", flush=True)
    assert drift <= 1e-4
    assert coll > 0


@pytest.mark.parametrize("pair_mode", ["uniform", "reference"])
def test_wave_collide_equals_serial(pair_mode, monkeypatch):
    """Cells above the wave-trial threshold are run by a whole wave
    speculatively; the result must be bitwise that of one lane per cell.
    Dense cells (n_real = 3e9 on 1e5 atoms, 12^3 cells) give hundreds of
    trials per cell."""
    trap = configs.quadrupole_trap(100000, 12, n_real=3e9)
    run = configs.Run(traps=[trap], precision="fp32", dt=1e-6, seed=5, pair_mode=pair_mode,
                      name="wave-collide")
    out = {}
    for thr in ("1000000000", "0", "16"):
        monkeypatch.setenv("DSMC_COLLIDE_WAVE_TRIALS", thr)
        print(f"wave collide threshold {thr}, pair mode {pair_mode}", flush=True)
        with Simulation(run) as sim:
            sim.init_distribution()
            sim.step(20)
            out[thr] = (sim.copy_out(_abi.FIELD_POS), sim.copy_out(_abi.FIELD_VEL),
                        sim.obs_log()["collisions"].copy())
    ser = out["1000000000"]
    assert int(ser[2].sum()) > 20 * 1000
    for thr in ("0", "16"):
        for a, b in zip(ser, out[thr]):
            np.testing.assert_array_equal(a, b)


def test_incremental_sort_is_the_stable_sort():
    """After a sorted move only the atoms that changed cell are re-sorted and
    merged; the permutation must be numpy's stable argsort of the new keys."""
    run = quad(200000, 40)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(7)
        for rep in range(3):
            keys = sim.copy_out(_abi.FIELD_CELL_ID).astype(np.int64)
            sim.sort_arrays()
            perm = sim.copy_out(_abi.FIELD_PERM)
            srt = sim.copy_out(_abi.FIELD_CELL_ID)
            order = np.argsort(keys, kind="stable")[: len(perm)]
            moved = int(np.count_nonzero(np.diff(keys) < 0))
            print(f"re-sort {rep}: n={len(keys)}, descents={moved}", flush=True)
            assert moved > 0
            np.testing.assert_array_equal(perm, order)
            np.testing.assert_array_equal(srt, keys[order])
            sim.build_cell_table()
            sim.collide()
            sim.move_atoms()


@pytest.mark.parametrize("n", [(1 << 24) - 1, 1 << 24])
def test_sort_at_the_incremental_limit(n):
    """The incremental re-sort at its largest size (2^24 - 1 slots: 8192
    tiles of 2048, the locate's LDS table of tile samples full) and one slot
    beyond it (the step's sort then falls back to the radix sort): after some
    steps the sort equals numpy's stable argsort of the keys, and the log's
    energy stays within 1e-5 of the initial energy."""
    run = quad(n, 128)
    print(f"sort at the incremental limit: n = {n}", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        e0 = sim.observe()[0]
        sim.step(6)
        keys = sim.copy_out(_abi.FIELD_CELL_ID).astype(np.int64)
        sim.sort_arrays()
        perm = sim.copy_out(_abi.FIELD_PERM)
        srt = sim.copy_out(_abi.FIELD_CELL_ID)
        log = sim.obs_log()
    assert len(keys) == n and int(np.count_nonzero(np.diff(keys) < 0)) > 0
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(perm, order)
    np.testing.assert_array_equal(srt, keys[order])
    E0 = e0.ek + e0.ep
    assert np.max(np.abs(log["ek"][:, 0] + log["ep"][:, 0] - E0)) / E0 < 1e-5


def test_incremental_sort_equals_full_sort(monkeypatch):
    """Whole batched runs (evaporation on, graph replay, a block-size autotune
    half way) with the incremental and the full sort are bitwise identical,
    with the incremental re-sort's first phase in the finalize launch or in
    the re-sort itself."""
    run = _batch_run()
    out = {}
    for mode in ("incremental", "incremental-prep-in-resort", "full"):
        print(f"batched run, sort mode {mode}", flush=True)
        monkeypatch.setenv("DSMC_RESORT_PREP_IN_FINALIZE",
                           "0" if mode == "incremental-prep-in-resort" else "1")
        with Simulation(run.with_(sort_mode=mode.split("-")[0], block_size=64)) as sim:
            sim.init_distribution()
            sim.step(20)
            # changes the move block size between steps: the tuner takes
            # shapes of at least 512 atoms per block only
            assert sim.autotune() in (256, 512, 1024)
            sim.step(21)
            out[mode] = [sim.copy_out(f) for f in (_abi.FIELD_POS, _abi.FIELD_VEL,
                                                   _abi.FIELD_ATOM_ID)]
            out[mode].append(sim.obs_log()["collisions"].copy())
            out[mode].append(sim.obs_log()["n_evap"].copy())
    assert int(out["full"][4].sum()) > 0
    for other in ("incremental-prep-in-resort", "full"):
        for a, b in zip(out["incremental"], out[other]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("case", ["batched", "dense-64", "dense-256"])
def test_fused_step_equals_split(case, monkeypatch):
    """The fused step (DSMC_STEP_FUSED=1: collisions inside the move kernel on
    an LDS copy of each wave's velocities and its neighbours', oversized cells
    across span boundaries in an edge pass) against the collide and move
    kernels launched one after the other: bitwise equal states, labels,
    collision counts, removals and the removed atoms' records.  "batched": three traps with evaporation,
    graph replay and a block-size autotune half way; "dense": cells of tens to
    hundreds of atoms (many cross span boundaries, many run wave-parallel),
    with 64- and 256-thread blocks."""
    if case == "batched":
        run, steps = _batch_run(), (20, 21)
    else:
        trap = configs.quadrupole_trap(100000, 12, n_real=3e9)
        bs = 64 if case == "dense-64" else 256
        run = configs.Run(traps=[trap], precision="fp32", dt=1e-6, seed=5, block_size=bs,
                          name="dense-cells")
        steps = (10, 11)
    out = {}
    for split in ("1", "0"):
        monkeypatch.setenv("DSMC_STEP_FUSED", "0" if split == "1" else "1")
        print(f"fused vs split step: {case}, split={split}", flush=True)
        with Simulation(run) as sim:
            sim.init_distribution()
            sim.step(steps[0])
            if case == "batched":
                sim.autotune()
            sim.step(steps[1])
            lg = sim.obs_log()
            out[split] = [sim.copy_out(f) for f in (_abi.FIELD_POS, _abi.FIELD_VEL, _abi.FIELD_ATOM_ID)]
            out[split] += [lg["collisions"].copy(), lg["n_evap"].copy(), lg["ek"].copy(),
                           sim.copy_out(_abi.FIELD_EVAP_POS), sim.copy_out(_abi.FIELD_EVAP_VEL)]
    assert int(out["1"][3].sum()) > 100
    for a, b in zip(out["1"][:5] + out["1"][6:], out["0"][:5] + out["0"][6:]):
        np.testing.assert_array_equal(a, b)   # the removed atoms' records at their labels too
    # the FP64 energy sums are per move block: the autotuner may choose
    # another block size for the fused kernel, which changes only the order
    # of the additions
    np.testing.assert_allclose(out["1"][5], out["0"][5], rtol=1e-14)


@pytest.mark.parametrize("n", [400, 900, 100000])
def test_incremental_sort_with_crowded_gap(n):
    """Atoms aimed at one cell: on the second step nearly every atom changes
    cell into the same cell, so one gap between two stays receives ~all the
    movers, and on the third step they leave it again.  n = 400 and 900: the
    bucket is ranked in the merge's LDS table with hundreds of ties; 1e5: far
    beyond the bucket's room (room and overflow list gathered and sorted in
    scratch).  Bitwise equal to the full radix sort."""
    L, nc, dt = 1e-3, 16, 1e-6
    trap = configs.Trap(kind="box", n_atoms=n, nc=(nc, nc, nc), half_width=(L / 2,) * 3,
                        T=20e-6, mu=0.0, n_real=1e10)
    run = configs.Run(traps=[trap], precision="fp32", dt=dt, seed=9, sigma=0.0,
                      name="crowded-gap")
    rng = np.random.default_rng(4)
    pos = rng.uniform(-0.4 * L, 0.4 * L, (n, 3))
    target = np.array([0.5, 0.5, 0.5]) * (L / nc)   # centre of a cell
    vel = (target - pos) / (2 * dt)
    out = {}
    for mode in ("incremental", "full"):
        print(f"crowded gap, sort mode {mode}", flush=True)
        with Simulation(run.with_(sort_mode=mode)) as sim:
            sim.set_state(pos, vel)
            sim.step(2)
            keys = sim.copy_out(_abi.FIELD_CELL_ID)
            crowd = int(np.max(np.bincount(keys.astype(np.int64))))
            sim.step(2)
            out[mode] = [sim.copy_out(f) for f in (_abi.FIELD_POS, _abi.FIELD_VEL,
                                                   _abi.FIELD_ATOM_ID)]
        print(f"largest cell after two steps: {crowd}", flush=True)
        assert crowd > n // 2
    for a, b in zip(out["incremental"], out["full"]):
        np.testing.assert_array_equal(a, b)


def test_trap_ranges_after_evaporation():
    """The trap ranges and alive count that the step keeps (from the atoms
    each trap lost) describe the sorted order exactly: checked against the
    sorted keys themselves and against the removal counts in the log."""
    run = _batch_run()
    key_base = np.cumsum([0] + [int(np.prod(t.nc)) + 1 for t in run.traps])
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(31)
        sim.sort_arrays()
        keys = sim.copy_out(_abi.FIELD_CELL_ID).astype(np.int64)
        ts = sim.copy_out(_abi.FIELD_TRAP_START)
        log = sim.obs_log()
    removed = int(log["n_evap"].sum())
    print(f"alive {len(keys)}, removed {removed}", flush=True)
    assert removed > 0
    assert len(keys) == run.n_atoms - removed
    assert np.all(keys < key_base[-1])          # no removed atom among them
    assert np.all(np.diff(keys) >= 0)
    np.testing.assert_array_equal(ts, np.searchsorted(keys, key_base))


def test_graph_replay_equals_eager():
    run = quad(40000, 30)
    with Simulation(run) as a, Simulation(run.with_(use_graph=False)) as b:
        a.init_distribution()
        b.init_distribution()
        a.step(51)
        b.step(51)
        for f in (_abi.FIELD_POS, _abi.FIELD_VEL, _abi.FIELD_ATOM_ID):
            np.testing.assert_array_equal(a.copy_out(f), b.copy_out(f))
        la, lb = a.obs_log(), b.obs_log()
        np.testing.assert_array_equal(la["collisions"], lb["collisions"])
        assert int(la["collisions"].sum()) > 0


def _batch_run():
    traps = []
    for n, dBdz, T in [(3000, 1.5, 30e-6), (20000, 4.0, 8e-6), (7000, 2.5, 15e-6)]:
        t = configs.quadrupole_trap(n, configs.cells_for_trap(n), T=T, dBdz=dBdz, n_real=3e6)
        t.eta_start = t.eta_end = 6.0
        t.T_cut = T
        traps.append(t)
    return configs.Run(traps=traps, precision="fp32", energy_cut=True, dt=1e-6, seed=11,
                       name="batch-test")


def test_batched_trap_equals_standalone():
    run = _batch_run()
    print("batched vs standalone", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(60)
        ts = sim.copy_out(_abi.FIELD_TRAP_START)
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        pos = sim.copy_out(_abi.FIELD_POS)
        vel = sim.copy_out(_abi.FIELD_VEL)
        log = sim.obs_log()
    begin = np.cumsum([0] + [t.n_atoms for t in run.traps])
    total_coll = 0
    for t in range(len(run.traps)):
        tr = run.traps[t]
        tr.philox_stream = t
        solo = configs.Run(traps=[tr], precision="fp32", energy_cut=True, dt=1e-6, seed=11)
        with Simulation(solo) as s1:
            s1.init_distribution()
            s1.step(60)
            ids1 = s1.copy_out(_abi.FIELD_ATOM_ID)
            np.testing.assert_array_equal(ids[ts[t]:ts[t + 1]] - begin[t], ids1)
            np.testing.assert_array_equal(pos[ts[t]:ts[t + 1]], s1.copy_out(_abi.FIELD_POS))
            np.testing.assert_array_equal(vel[ts[t]:ts[t + 1]], s1.copy_out(_abi.FIELD_VEL))
            l1 = s1.obs_log()
            np.testing.assert_array_equal(log["collisions"][:, t], l1["collisions"][:, 0])
            np.testing.assert_array_equal(log["n_evap"][:, t], l1["n_evap"][:, 0])
            np.testing.assert_allclose(log["ek"][:, t], l1["ek"][:, 0], rtol=1e-12)
            total_coll += int(l1["collisions"].sum())
    assert total_coll > 0
    assert int(log["n_evap"].sum()) > 0


def test_host_reads_between_graph_replays():
    """A host read between two runs of graph-replayed steps leaves the run
    bitwise as it is without the read (host reads change no device state that
    later steps use)."""
    run = _batch_run()
    out = {}
    for mode in ("plain", "read"):
        print(f"host reads between replays: {mode}", flush=True)
        with Simulation(run) as sim:
            sim.init_distribution()
            if mode == "read":
                sim.step(30)
                _ = sim.copy_out(_abi.FIELD_POS)
                sim.step(30)
            else:
                sim.step(60)
            out[mode] = [sim.copy_out(f) for f in (_abi.FIELD_ATOM_ID, _abi.FIELD_POS,
                                                   _abi.FIELD_VEL, _abi.FIELD_CELL_ID)]
            out[mode].append(sim.obs_log()["n_evap"].copy())
    assert int(out["plain"][4].sum()) > 0
    for a, b in zip(out["plain"], out["read"]):
        np.testing.assert_array_equal(a, b)


def test_reads_after_a_removing_step():
    """Host reads right after a step whose move removed atoms in several traps:
    the removed atoms keep their slots (dead key) until the next sort, and
    every read skips them.  dsmc_observe per trap equals the last log record
    (the log sums the move's FP32 per-atom squares and potentials, dsmc_observe
    FP64 ones of value + remainder: within 1e-6), its alive counts equal the
    trap ranges of copy_out, and the survivors' velocities from copy_out give
    its sums; no survivor lies above the cut.  dsmc_observe and
    dsmc_calculate_temp give the same bits on repeated calls."""
    traps = []
    for n, dBdz, T in [(3000, 1.5, 30e-6), (20000, 4.0, 8e-6), (7000, 2.5, 15e-6), (12000, 3.0, 20e-6)]:
        t = configs.quadrupole_trap(n, configs.cells_for_trap(n), T=T, dBdz=dBdz, n_real=3e6)
        # a cut ramped down through the thermal tail: atoms go in every step
        t.eta_start, t.eta_end, t.T_cut, t.eta_steps = 5.0, 2.0, T, 40
        traps.append(t)
    run = configs.Run(traps=traps, precision="fp32", energy_cut=True, dt=1e-6, seed=13,
                      name="reads-after-removal")
    for steps in (25, 26):   # last step replayed from the graph, then eager
        print(f"reads after a removing step, {steps} steps", flush=True)
        with Simulation(run) as sim:
            sim.init_distribution()
            sim.step(steps)
            last = sim.obs_log()[-1]
            obs = sim.observe()
            # fixed-order sums: the same bits on every call
            again = sim.observe()
            for a, b in zip(obs, again):
                assert (a.ek, a.ep, tuple(a.v2), a.n_alive) == (b.ek, b.ep, tuple(b.v2), b.n_alive)
            T1, T2 = sim.calculate_temp(), sim.calculate_temp()
            assert T1 == T2 and T1 > 0
            ts = sim.copy_out(_abi.FIELD_TRAP_START)
            vel = sim.copy_out(_abi.FIELD_VEL)
            pos = sim.copy_out(_abi.FIELD_POS)
            ids = sim.copy_out(_abi.FIELD_ATOM_ID)
            info = sim.info()
        assert int((last["n_evap"] > 0).sum()) >= 2, last["n_evap"]
        assert info.n_alive == ts[-1] == len(ids) == len(np.unique(ids))
        for t, tr in enumerate(run.traps):
            o = obs[t]
            assert o.n_alive == last["n_alive"][t] == ts[t + 1] - ts[t]
            np.testing.assert_allclose([o.v2[0], o.v2[1], o.v2[2], o.ep], list(last["v2"][t]) + [last["ep"][t]],
                                       rtol=1e-6)
            v = vel[ts[t]:ts[t + 1]]
            np.testing.assert_allclose(list(o.v2), (v ** 2).sum(axis=0), rtol=1e-12)
            x, y, z = pos[ts[t]:ts[t + 1]].T
            E = 0.5 * run.mass * (v ** 2).sum(axis=1) + tr.mu * tr.dBdz * np.sqrt(x * x + y * y + 4 * z * z)
            f = min((steps - 1) / tr.eta_steps, 1.0)
            cut = (tr.eta_start + (tr.eta_end - tr.eta_start) * f) * configs.KB * tr.T_cut
            assert E.max() <= cut * (1 + 1e-6)


def test_empty_and_emptied_traps():
    """A batch with traps of no atoms (first, middle and last) and a trap whose
    cut (0.05 kB T) removes every atom in its first step: their slot ranges
    stay empty, their log records and dsmc_observe sums read zero, and the
    populated traps are bitwise the same as when run alone."""
    spec = [(0, 2.0, 20e-6, 0.0), (3000, 1.5, 30e-6, 6.0), (0, 2.0, 20e-6, 0.0),
            (2000, 2.5, 15e-6, 0.05), (20000, 4.0, 8e-6, 6.0), (0, 2.0, 20e-6, 0.0)]
    traps = []
    for n, dBdz, T, eta in spec:
        t = configs.quadrupole_trap(n, configs.cells_for_trap(n), T=T, dBdz=dBdz, n_real=3e6)
        t.eta_start = t.eta_end = eta
        t.T_cut = T
        traps.append(t)
    run = configs.Run(traps=traps, precision="fp32", energy_cut=True, dt=1e-6, seed=17,
                      name="empty-traps")
    steps = 30
    print("empty and emptied traps: batch", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(steps)
        ts = sim.copy_out(_abi.FIELD_TRAP_START)
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        pos = sim.copy_out(_abi.FIELD_POS)
        vel = sim.copy_out(_abi.FIELD_VEL)
        log = sim.obs_log()
        obs = sim.observe()
        n_alive = sim.info().n_alive
    assert ts[0] == ts[1] == 0 and ts[2] == ts[3] == ts[4]
    assert ts[5] == ts[6] == n_alive == len(ids) == len(np.unique(ids))
    for t in (0, 2, 5):
        for f in ("n_alive", "n_evap", "collisions", "ek", "ep"):
            assert not log[f][:, t].any(), (t, f)
        assert (obs[t].n_alive, obs[t].ek, obs[t].ep) == (0, 0.0, 0.0)
    assert int(log["n_evap"][:, 3].sum()) == spec[3][0]
    assert log["n_alive"][0, 3] == 0 and obs[3].n_alive == 0 and obs[3].ek == 0.0
    begin = np.cumsum([0] + [t.n_atoms for t in traps])
    total_coll = 0
    for t in (1, 4):
        tr = traps[t]
        tr.philox_stream = t
        print(f"empty and emptied traps: trap {t} alone", flush=True)
        solo = configs.Run(traps=[tr], precision="fp32", energy_cut=True, dt=1e-6, seed=17)
        with Simulation(solo) as s1:
            s1.init_distribution()
            s1.step(steps)
            np.testing.assert_array_equal(ids[ts[t]:ts[t + 1]] - begin[t],
                                          s1.copy_out(_abi.FIELD_ATOM_ID))
            np.testing.assert_array_equal(pos[ts[t]:ts[t + 1]], s1.copy_out(_abi.FIELD_POS))
            np.testing.assert_array_equal(vel[ts[t]:ts[t + 1]], s1.copy_out(_abi.FIELD_VEL))
            l1 = s1.obs_log()
        np.testing.assert_array_equal(log["collisions"][:, t], l1["collisions"][:, 0])
        np.testing.assert_array_equal(log["n_evap"][:, t], l1["n_evap"][:, 0])
        np.testing.assert_array_equal(log["n_alive"][:, t], l1["n_alive"][:, 0])
        np.testing.assert_allclose(log["ek"][:, t], l1["ek"][:, 0], rtol=1e-12)
        total_coll += int(l1["collisions"].sum())
    assert total_coll > 0


@pytest.mark.parametrize("n_traps", [600, 8192])
def test_many_traps(n_traps):
    """Batches of many small traps (64 atoms, 8^3 cells each): 600 traps,
    beyond the 511 up to which the collide runs one-wave blocks, and 8192,
    the most a context takes (every kernel's trap table fills 32 KB of LDS).
    The trap ranges hold 64 atoms each, and the first, middle and last traps
    are bitwise the same as when run alone."""
    traps = [configs.quadrupole_trap(64, 8, T=10e-6 + 20e-6 * (i % 7) / 6, dBdz=2.5, n_real=3e8)
             for i in range(n_traps)]
    run = configs.Run(traps=traps, precision="fp32", dt=1e-6, seed=21, obs_capacity=16,
                      name="many-traps")
    steps = 12
    print(f"many traps: {n_traps}", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(steps)
        ts = sim.copy_out(_abi.FIELD_TRAP_START)
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        pos = sim.copy_out(_abi.FIELD_POS)
        vel = sim.copy_out(_abi.FIELD_VEL)
        log = sim.obs_log()
        n_alive = sim.info().n_alive
    assert len(ts) == n_traps + 1 and ts[0] == 0 and ts[-1] == n_alive == 64 * n_traps
    assert np.all(np.diff(ts) == 64)
    for t in (0, n_traps // 2, n_traps - 1):
        tr = traps[t]
        tr.philox_stream = t
        print(f"many traps: {n_traps}, trap {t} alone", flush=True)
        solo = configs.Run(traps=[tr], precision="fp32", dt=1e-6, seed=21, obs_capacity=16)
        with Simulation(solo) as s1:
            s1.init_distribution()
            s1.step(steps)
            np.testing.assert_array_equal(ids[ts[t]:ts[t + 1]] - 64 * t, s1.copy_out(_abi.FIELD_ATOM_ID))
            np.testing.assert_array_equal(pos[ts[t]:ts[t + 1]], s1.copy_out(_abi.FIELD_POS))
            np.testing.assert_array_equal(vel[ts[t]:ts[t + 1]], s1.copy_out(_abi.FIELD_VEL))
            l1 = s1.obs_log()
        np.testing.assert_array_equal(log["collisions"][:, t], l1["collisions"][:, 0])
        np.testing.assert_allclose(log["ek"][:, t], l1["ek"][:, 0], rtol=1e-12)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_every_atom_removed(precision):
    """A cut of 0.05 kB T removes every atom of a single trap in its first
    step; the steps after it run on no atoms: the log reads zero alive atoms
    from the first record on, the removed energy equals the initial energy
    (the step conserves it), every atom has its evaporation record, reads
    return empty arrays, and dsmc_observe and dsmc_calculate_temp read zero.
    Both numerical forms."""
    run = quad(20000, 30, precision=precision, energy_cut=True)
    tr = run.traps[0]
    tr.eta_start = tr.eta_end = 0.05
    tr.T_cut = tr.T
    print(f"every atom removed, {precision}", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        e0 = sim.observe()[0]
        sim.step(21)
        log = sim.obs_log()
        obs = sim.observe()[0]
        assert sim.info().n_alive == 0 and sim.calculate_temp() == 0.0
        assert sim.copy_out(_abi.FIELD_POS).shape == (0, 3)
        assert sim.copy_out(_abi.FIELD_ATOM_ID).size == 0
        # a field the form does not keep is still refused, not read as empty
        lacking = _abi.FIELD_ACC if precision == "fp32" else _abi.FIELD_POS_COMP
        with pytest.raises(_abi.DsmcError) as err:
            sim.copy_out(lacking)
        assert err.value.status == _abi.DSMC_ENOTSUP
        ev = sim.copy_out(_abi.FIELD_EVAP_POS).reshape(-1, 3)
    assert not log["n_alive"][:, 0].any()
    assert int(log["n_evap"][0, 0]) == run.n_atoms and not log["n_evap"][1:, 0].any()
    assert not log["collisions"][1:, 0].any() and not log["ek"][1:, 0].any()
    np.testing.assert_allclose(log["e_evap"][0, 0], e0.ek + e0.ep, rtol=1e-4)
    assert ev.shape == (run.n_atoms, 3) and np.isfinite(ev).all()
    assert (obs.n_alive, obs.ek, obs.ep) == (0, 0.0, 0.0)


def test_energy_cut_survivors(orc):
    """One step with a fixed cut: survivors = atoms whose FP64 energy after the
    move is below eta kB T (oracle move of the same state), away from the
    threshold by more than the FP32 trajectory error; each removed atom's
    evaporation record, stored at its label (which the state carries in its
    position remainders), holds its position and velocity after that move."""
    run = quad(50000, 40, sigma=0.0, energy_cut=True)
    T = run.traps[0].T
    run.traps[0].eta_start = run.traps[0].eta_end = 5.0
    run.traps[0].T_cut = T
    ref = orc.Reference(run.with_(precision="fp64", energy_cut=False))
    with Simulation(run) as sim:
        sim.set_state(ref.s["pos"], ref.s["vel"])
        sim.step(1)
        sim.step(1)  # the second sort removes the atoms tagged by the first move
        alive = set(sim.copy_out(_abi.FIELD_ATOM_ID).tolist())
        nev = int(sim.obs_log()["n_evap"].sum())
        ev_pos = sim.copy_out(_abi.FIELD_EVAP_POS).reshape(-1, 3)
        ev_vel = sim.copy_out(_abi.FIELD_EVAP_VEL).reshape(-1, 3)
    ref.step(1)
    p = ref.p
    E = np.array([0.0])
    ek = 0.5 * run.mass * np.sum(ref.s["vel"] ** 2, axis=1)
    x, y, z = ref.s["pos"].T
    ep = run.traps[0].mu * run.traps[0].dBdz * np.sqrt(x * x + y * y + 4.0 * z * z)
    E = ek + ep
    cut = 5.0 * configs.KB * T
    far = np.abs(E - cut) > 1e-6 * cut
    expect_alive = set(np.where((E <= cut) & far)[0].tolist())
    expect_dead = set(np.where((E > cut) & far)[0].tolist())
    assert expect_alive <= alive
    assert not (expect_dead & alive)
    assert nev >= len(expect_dead)
    dead = np.array(sorted(expect_dead))
    assert len(dead) > 100
    np.testing.assert_allclose(ev_pos[dead], ref.s["pos"][dead], rtol=1e-6, atol=1e-10)
    np.testing.assert_allclose(ev_vel[dead], ref.s["vel"][dead], rtol=1e-6, atol=1e-7)
    del p


def _bloch(u, d):
    return np.stack([2 * (u[:, 0] * d[:, 0] + u[:, 1] * d[:, 1]),
                     2 * (u[:, 0] * d[:, 1] - u[:, 1] * d[:, 0]),
                     (u ** 2).sum(1) - (d ** 2).sum(1)], axis=1)


def test_spin_precession_and_loss_match_oracle(orc):
    """Spin-evaporation configuration without collisions (N = 2e4, 50 uK,
    20 spin substeps per 1 us step, rotation at substep midpoints): the FP32
    Bloch-vector rotation against the oracle's FP64 spinor update
    (wavefunction.cu:39-73) and the spin-flip loss set (evaporation.cu:90-94)
    after 300 steps."""
    run = configs.spin_evaporation("fp32", n_atoms=20000, nc=40).with_(
        sigma=0.0, energy_cut=False, obs_capacity=512)
    ref = orc.Reference(run.with_(precision="fp64"))
    s0 = _bloch(ref.s["psi_up"], ref.s["psi_dn"])
    print("spin precession vs oracle, 300 steps", flush=True)
    with Simulation(run) as sim:
        sim.set_state(ref.s["pos"], ref.s["vel"])
        sim.set_spin(s0)
        sim.step(300)
        sim.step(1)                       # remove the atoms tagged by the last move
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        spin = sim.copy_out(_abi.FIELD_SPIN)
    ref.step(300)
    alive_ref = set(ref.s["atom_id"].tolist())
    lost_ref = 20000 - len(alive_ref)
    # the GPU ran one more step: compare the atoms both still hold
    both = np.array(sorted(set(ids.tolist()) & alive_ref))
    diff = len(alive_ref ^ set(ids.tolist()))
    print(f"lost: oracle {lost_ref}, mismatched labels {diff}", flush=True)
    assert lost_ref > 0
    assert diff <= max(3, lost_ref // 10 + 2)
    # advance the oracle by the same extra step for the spin comparison
    ref.step(1)
    s_ref = _bloch(ref.s["psi_up"], ref.s["psi_dn"])
    by = np.full((20000, 3), np.nan)
    by[ids] = spin
    common = np.array(sorted(set(ids.tolist()) & set(ref.s["atom_id"].tolist())))
    err = np.abs(by[common] - s_ref[common]).max(axis=1)
    print(f"max |ds| = {err.max():.2e}, 99.9% quantile {np.quantile(err, 0.999):.2e}", flush=True)
    assert np.quantile(err, 0.999) < 1e-3
    del both

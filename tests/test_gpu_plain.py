"""GPU parity of the FP64 drop-in form against the CPU oracle.

Both sides start from the same positions and velocities (sampled by the
oracle), use the same Philox counters and the same FP64 operand order with no
fused multiply-adds, so cell ids, the sort permutation, the cell table, the
compaction survivors and (on the quadrupole and box traps) the whole FP64
trajectory including collisions are compared bit for bit.  Spin rotations use
sin/cos, whose last-ulp results differ between the CPU libm and the device
library, so spinors are compared to 1e-13.
"""
import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs

pytestmark = pytest.mark.gpu


def quad_run(n=20000, nc=35, **kw):
    run = configs.Run(traps=[configs.quadrupole_trap(n, nc)], precision="fp64", dt=1e-6, seed=7,
                      name="test-quad")
    return run.with_(**kw) if kw else run


def start(orc, run):
    """Oracle-sampled initial state loaded into a fresh GPU context."""
    ref = orc.Reference(run)
    sim = Simulation(run)
    sim.set_state(ref.s["pos"], ref.s["vel"])
    return ref, sim


def by_label(sim, field):
    """GPU per-slot field reordered to atom-label order (the reference's
    physical-index layout); rows of removed atoms are NaN."""
    ids = sim.copy_out(_abi.FIELD_ATOM_ID)
    a = sim.copy_out(field)
    out = np.full((sim.run.n_atoms,) + a.shape[1:], np.nan)
    out[ids] = a
    return out, ids


def test_index_sort_table_bit_exact(orc):
    run = quad_run()
    print("index/sort/table quadrupole N=20000 35^3", flush=True)
    ref, sim = start(orc, run)
    with sim:
        sim.index_atoms()
        cell_gpu = sim.copy_out(_abi.FIELD_CELL_ID)
        cell_orc = orc.index_atoms(ref.p, ref.s["pos"], ref.s["atom_id"])
        np.testing.assert_array_equal(cell_gpu, cell_orc)

        sim.sort_arrays()
        k_orc, id_orc = orc.sort_arrays(cell_orc, ref.s["atom_id"], orc.ncells(ref.p) + 1)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_ID), k_orc)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_ATOM_ID), id_orc)

        sim.build_cell_table()
        se, count, prefix = orc.cell_table(k_orc, orc.ncells(ref.p))
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_START_END), se)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_COUNT), count)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_PREFIX), prefix)


@pytest.mark.parametrize("pair_mode", ["uniform", "reference"])
def test_collide_bit_exact(orc, pair_mode):
    # dense cells so that every cell runs many trials and accepts often
    run = quad_run(n=50000, nc=12, pair_mode=pair_mode)
    run.traps[0].n_real = 5e9
    print(f"collide bit-exact {pair_mode}", flush=True)
    ref, sim = start(orc, run)
    with sim:
        sim.index_atoms(); sim.sort_arrays(); sim.build_cell_table(); sim.collide()
        info = ref.collision_step()
        assert info["total"] > 1000
        vel_gpu, ids = by_label(sim, _abi.FIELD_VEL)
        np.testing.assert_array_equal(vel_gpu, ref.s["vel"])
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_COLLISION_COUNT)[:-1],
                                      info["coll_count"][:-1])
        sv = sim.copy_out(_abi.FIELD_SIGVRMAX)
        np.testing.assert_array_equal(sv[: len(ref.sigvrmax)], ref.sigvrmax)


@pytest.mark.parametrize("kind", ["quadrupole", "ip", "box"])
def test_trajectory_bit_exact(orc, kind):
    if kind == "quadrupole":
        run = quad_run(n=20000, nc=30)
    elif kind == "ip":
        run = configs.reference(20, 20000).with_(spin=False, spin_loss=False, mcwf=False, lpc=1, dt=1e-6,
                                                 pair_mode="uniform")
    else:
        run = configs.box_thermalise("fp64", n_atoms=30000, nc=24)
    print(f"trajectory bit-exact {kind}", flush=True)
    ref, sim = start(orc, run)
    with sim:
        steps = 40
        sim.step(steps)
        ref.step(steps)
        pos, ids = by_label(sim, _abi.FIELD_POS)
        vel, _ = by_label(sim, _abi.FIELD_VEL)
        acc, _ = by_label(sim, _abi.FIELD_ACC)
        np.testing.assert_array_equal(np.sort(ids), np.arange(run.n_atoms))
        np.testing.assert_array_equal(pos, ref.s["pos"])
        np.testing.assert_array_equal(vel, ref.s["vel"])
        np.testing.assert_array_equal(acc, ref.s["acc"])
        log = sim.obs_log()
        assert int(log["collisions"].sum()) == ref.collisions
        e = orc.energy(ref.p, ref.s["pos"], ref.s["vel"], ref.s["atom_id"])
        np.testing.assert_allclose(log["ek"][-1, 0], e[0], rtol=1e-12)
        np.testing.assert_allclose(log["ep"][-1, 0], e[1], rtol=1e-12)


def test_reference_run_spin_and_evaporation(orc):
    """The reference's own loop (IP trap, spinors, spin-projection loss, lpc 10,
    biased pair choice) against the oracle for 30 motion steps."""
    run = configs.reference(20, 20000)
    print("reference loop with spin + evaporation", flush=True)
    ref = orc.Reference(run)
    # tilt a tenth of the spinors so that some atoms are lost
    psi_u, psi_d = ref.s["psi_up"], ref.s["psi_dn"]
    flip = np.arange(run.n_atoms) % 10 == 3
    psi_u[flip], psi_d[flip] = psi_d[flip].copy(), -psi_u[flip].copy()
    sim = Simulation(run)
    with sim:
        sim.set_state(ref.s["pos"], ref.s["vel"])
        sim.set_spin(np.concatenate([psi_u, psi_d], axis=1))
        sim.step(30)
        ref.step(30)
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        assert set(ids.tolist()) == set(ref.s["atom_id"].tolist())
        assert len(ids) < run.n_atoms
        pos, _ = by_label(sim, _abi.FIELD_POS)
        alive = ref.s["atom_id"]
        np.testing.assert_array_equal(pos[alive], ref.s["pos"][alive])
        ev = sim.copy_out(_abi.FIELD_EVAP_POS)
        lost = np.setdiff1d(np.arange(run.n_atoms), alive)
        np.testing.assert_array_equal(ev[lost], ref.evap_pos[lost])
        sp, _ = by_label(sim, _abi.FIELD_SPIN)
        np.testing.assert_allclose(sp[alive, :2], psi_u[alive], rtol=0, atol=1e-13)
        np.testing.assert_allclose(sp[alive, 2:], psi_d[alive], rtol=0, atol=1e-13)


def test_init_distribution_matches_oracle(orc):
    run = quad_run(n=20000)
    print("init distribution", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        pos, ids = by_label(sim, _abi.FIELD_POS)
        vel, _ = by_label(sim, _abi.FIELD_VEL)
    p = orc.params_from(run)
    st = orc.init_distribution(p, run.n_atoms, run.traps[0].T)
    np.testing.assert_allclose(pos, st["pos"], rtol=1e-13, atol=1e-20)
    np.testing.assert_allclose(vel, st["vel"], rtol=1e-13, atol=1e-18)


def _bloch(up, dn):
    return np.stack([2 * (up[:, 0] * dn[:, 0] + up[:, 1] * dn[:, 1]),
                     2 * (up[:, 0] * dn[:, 1] - up[:, 1] * dn[:, 0]),
                     (up ** 2).sum(1) - (dn ** 2).sum(1)], axis=1)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("kind", ["ip", "box", "quadrupole"])
def test_init_samplers_match_oracle(orc, kind, precision):
    """The device samplers of setUp.cu:129-252 against the oracle's, from the
    same Philox draws: the Ioffe-Pritchard rejection sampler
    (setUp.cu:178-205) with spins aligned to the local field
    (setUp.cu:236-252), the periodic box (uniform positions, per-axis
    Maxwell velocities) and the quadrupole; in both numerical forms (the
    FP32 form stores the FP64 sample as value + remainder, and its spin as
    the Bloch vector of the aligned spinor).  To 1e-13 (the device log, exp
    and sin differ from glibc in the last ulp); the FP32 form's positions to
    2^-36, the precision of value + remainder when the remainder also holds
    the atom's label (13 significant bits)."""
    n = 20000
    if kind == "ip":
        run = configs.reference(Nc=40, N=n).with_(mcwf=False, spin_loss=False)
        if precision == "fp32":
            run = run.with_(lpc=1)
    elif kind == "box":
        L = 1e-3
        trap = configs.Trap(kind="box", n_atoms=n, nc=(32, 32, 32), half_width=(L / 2,) * 3,
                            T=(40e-6 + 2 * 20e-6) / 3, T_axis=(40e-6, 20e-6, 20e-6), n_real=1e10,
                            mu=0.0)
        run = configs.Run(traps=[trap], dt=1e-6, seed=11)
    else:
        run = quad_run(n=n).with_(spin=True)
    run = run.with_(precision=precision, obs_capacity=16)
    t = run.traps[0]
    print(f"init sampler {kind} {precision}", flush=True)
    with Simulation(run) as sim:
        sim.init_distribution()
        pos, ids = by_label(sim, _abi.FIELD_POS)
        vel, _ = by_label(sim, _abi.FIELD_VEL)
        spin = by_label(sim, _abi.FIELD_SPIN)[0] if run.spin else None
    assert np.array_equal(np.sort(ids), np.arange(n))
    p = orc.params_from(run.with_(precision="fp64"))
    st = orc.init_distribution(p, n, t.T, None if not any(t.T_axis) else t.T_axis, spin=run.spin)
    np.testing.assert_allclose(pos, st["pos"], rtol=2.0 ** -36 if precision == "fp32" else 1e-13,
                               atol=1e-20)
    np.testing.assert_allclose(vel, st["vel"], rtol=1e-13, atol=1e-18)
    if run.spin:
        if precision == "fp64":
            np.testing.assert_allclose(spin[:, :2], st["psi_up"], rtol=0, atol=1e-13)
            np.testing.assert_allclose(spin[:, 2:], st["psi_dn"], rtol=0, atol=1e-13)
        else:
            np.testing.assert_allclose(spin, _bloch(st["psi_up"], st["psi_dn"]), rtol=0, atol=1e-6)


def test_golden_index_fixture(orc):
    """tests/golden/index_fixture.npz: 4096 positions of the 35^3 quadrupole
    grid with atoms on cell faces, one ulp either side, on the grid edges,
    outside the grid, NaN and inf; expected cell ids, permutation, cell table
    and compaction survivors were written by the oracle."""
    import ctypes as C
    from pathlib import Path
    from tests.gpu_util import DeviceArray

    fx = np.load(Path(__file__).resolve().parent / "golden" / "index_fixture.npz")
    base = configs.quadrupole_1e4()
    trap = base.traps[0]
    trap.n_atoms = len(fx["pos"])
    run = base.with_(traps=[trap])
    print("golden index fixture", flush=True)
    with Simulation(run) as sim:
        sim.set_state(fx["pos"], fx["vel"])
        sim.index_atoms()
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_ID), fx["cell"])
        sim.sort_arrays()
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_ID), fx["sorted_cell"])
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_ATOM_ID), fx["sorted_id"])
        sim.build_cell_table()
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_START_END), fx["start_end"])
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_COUNT), fx["count"])
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_PREFIX), fx["prefix"])
    lib = _abi.load_library()
    with DeviceArray(lib, fx["sorted_id"]) as vals, DeviceArray(lib, fx["stencil"]) as sten:
        n_out = C.c_int()
        assert lib.dsmc_remove_if(vals.ptr, sten.ptr, len(fx["sorted_id"]), C.byref(n_out), None) == 0
        got = vals.get()[: n_out.value]
    np.testing.assert_array_equal(got, fx["survivors"])


def _tilted(orc, p, pos, q):
    """Spinors with spin-down population q along the local field."""
    n = len(pos)
    u = np.zeros((n, 2)); d = np.zeros((n, 2))
    for i in range(n):
        B = orc.field(p, pos[i])
        bx, by, bz = B / np.linalg.norm(B)
        H = np.array([[bz, bx - 1j * by], [bx + 1j * by, -bz]])
        w, V = np.linalg.eigh(H)
        psi = np.sqrt(1 - q) * V[:, 1] + np.sqrt(q) * V[:, 0]
        u[i] = psi[0].real, psi[0].imag
        d[i] = psi[1].real, psi[1].imag
    return u, d


def test_spin_flip_calls_match_oracle(orc):
    """dsmc_local_populations, dsmc_flip_atoms and dsmc_exponential_decay one
    call at a time against the oracle: populations of aligned spins, then a
    30 % spin-down admixture flips atoms (bit-exact velocities and spin
    states), then decoherence over lpc * dt (spinors within 1e-13)."""
    run = configs.reference(20, 20000).with_(mcwf=True)
    n = run.n_atoms
    p = orc.params_from(run)
    st = orc.init_distribution(p, n, run.traps[0].T, spin=True)
    ids = np.arange(n, dtype=np.int32)
    print("spin-flip calls vs oracle, n=20000", flush=True)
    with Simulation(run) as sim:
        sim.set_state(st["pos"], st["vel"])
        sim.set_spin(np.concatenate([st["psi_up"], st["psi_dn"]], axis=1))
        sim.local_populations()
        pops = orc.local_populations(p, st["pos"], st["psi_up"], st["psi_dn"], ids)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_LOCAL_POPS), pops)
        u, d = _tilted(orc, p, st["pos"], 0.3)
        sim.set_spin(np.concatenate([u, d], axis=1))
        vel = st["vel"].copy()
        up = np.ones(n, np.uint8)
        k_ref = orc.flip_atoms(p, st["pos"], vel, u, d, pops, up, ids, 0)
        k = sim.flip_atoms()
        assert k == k_ref and 0.25 * n < k < 0.35 * n
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_SPIN_UP), up)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_VEL), vel)
        sim.exponential_decay()
        orc.exponential_decay(p, st["pos"], u, d, ids, run.lpc * run.dt)
        sp = sim.copy_out(_abi.FIELD_SPIN)
        np.testing.assert_allclose(sp[:, :2], u, rtol=0, atol=1e-13)
        np.testing.assert_allclose(sp[:, 2:], d, rtol=0, atol=1e-13)


def test_reference_run_with_spin_flips(orc):
    """The reference loop with the spin-flip model (main.cu:457-521): 40
    motion steps (4 collision steps) from spinors with a 2 % spin-down
    admixture, against the oracle (positions, velocities and spin states
    bit-exact)."""
    run = configs.reference(20, 20000).with_(mcwf=True)
    print("reference loop with spin flips", flush=True)
    ref = orc.Reference(run)
    u, d = _tilted(orc, ref.p, ref.s["pos"], 0.02)
    ref.s["psi_up"][:], ref.s["psi_dn"][:] = u, d
    with Simulation(run) as sim:
        sim.set_state(ref.s["pos"], ref.s["vel"])
        sim.set_spin(np.concatenate([u, d], axis=1))
        sim.step(40)
        ref.step(40)
        ids = sim.copy_out(_abi.FIELD_ATOM_ID)
        assert set(ids.tolist()) == set(ref.s["atom_id"].tolist())
        alive = ref.s["atom_id"]
        pos, _ = by_label(sim, _abi.FIELD_POS)
        vel, _ = by_label(sim, _abi.FIELD_VEL)
        np.testing.assert_array_equal(pos[alive], ref.s["pos"][alive])
        np.testing.assert_array_equal(vel[alive], ref.s["vel"][alive])
        up, _ = by_label(sim, _abi.FIELD_SPIN_UP)
        np.testing.assert_array_equal(up[alive], ref.s["spin_up"][alive])
        # flips are rare here (adiabatic following); the call-by-call test
        # above exercises them in bulk
        print(f"flips {ref.flips}, alive {len(alive)}", flush=True)

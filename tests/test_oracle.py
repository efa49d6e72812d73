"""CPU pins of the oracle (oracle/dsmc_oracle.c): every function is checked
against something other than a retyped formula -- published known answers,
library routines, closed forms of the dynamics, conservation laws, and the
reference's own published numbers (homogeneousGasCollisions.py:27-33,
collisionChecker.py:34, energyChecker.py:185).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from cudadsmc_amd import configs

GOLDEN = Path(__file__).resolve().parent / "golden"


# ---------------------------------------------------------------------------
# Philox4x32-10 and the draws
# ---------------------------------------------------------------------------
def test_philox_known_answers(orc):
    kat = json.loads((GOLDEN / "philox_kat.json").read_text())
    for case in kat["vectors"]:
        assert orc.philox(case["ctr"], case["key"]) == case["out"]


def test_u01_grid(orc):
    # (k + 1/2) 2^-52 from the top 52 of 64 bits: both ends strictly inside (0,1)
    assert orc.u01(0, 0) == 2.0 ** -53
    assert orc.u01(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert orc.u01(0x80000000, 0) == 0.5 + 2.0 ** -53
    assert orc.u01(0x7FFFFFFF, 0xFFFFF000) == 0.5 - 2.0 ** -53


def test_uint_below(orc):
    for n in (1, 2, 3, 7, 1000, 2 ** 31):
        for r in (0, 1, 0x7FFFFFFF, 0xFFFFFFFF):
            v = orc.uint_below(r, n)
            assert 0 <= v < n
            assert v == (r * n) >> 32


# ---------------------------------------------------------------------------
# Trap field and motion
# ---------------------------------------------------------------------------
def _accel(orc, p, r):
    q = orc.Params.from_buffer_copy(p)
    q.dt = 0.0
    pos = np.array([r], dtype=np.float64)
    vel = np.zeros((1, 3))
    acc = np.zeros((1, 3))
    orc.move_atoms(q, pos, vel, acc, np.array([0], np.int32))
    return acc[0]


def _potential(run, r):
    t = run.traps[0]
    x, y, z = r
    if t.kind == "quadrupole":
        return t.mu * t.dBdz * math.sqrt(x * x + y * y + 4 * z * z)
    Bx = t.dBdx * x - 0.5 * t.d2Bdx2 * x * z
    By = -t.dBdx * y - 0.5 * t.d2Bdx2 * y * z
    Bz = t.B0 + 0.5 * t.d2Bdx2 * (z * z - 0.5 * (x * x + y * y))
    return t.mu * (math.sqrt(Bx * Bx + By * By + Bz * Bz) - t.B0)


@pytest.mark.parametrize("kind", ["ip", "quadrupole"])
def test_force_is_minus_gradient_of_potential(orc, kind):
    """a = -grad U / m by central differences of an independently written U
    (IP field of moveAtoms.cu:135-142 and the quadrupole |B| = B' rho)."""
    run = configs.reference() if kind == "ip" else configs.quadrupole_1e4()
    p = orc.params_from(run)
    rng = np.random.default_rng(1)
    for _ in range(20):
        r = rng.normal(0, 5e-5, 3)
        a = _accel(orc, p, r)
        g = np.zeros(3)
        h = 1e-10
        for d in range(3):
            e = np.zeros(3); e[d] = h
            g[d] = (_potential(run, r + e) - _potential(run, r - e)) / (2 * h)
        np.testing.assert_allclose(a, -g / run.mass, rtol=2e-5, atol=1e-9 * np.abs(g).max() / run.mass)


def test_verlet_exact_for_constant_force(orc):
    """On the +x axis of the quadrupole the force is the constant -mu B'/m x^:
    velocity Verlet reproduces x0 + v0 t - g t^2 / 2 to rounding."""
    run = configs.quadrupole_1e4()
    p = orc.params_from(run)
    t = run.traps[0]
    g = t.mu * t.dBdz / run.mass
    n = 64
    x0 = np.linspace(1e-4, 2e-4, n)
    v0 = np.linspace(-0.01, 0.05, n)
    pos = np.zeros((n, 3)); pos[:, 0] = x0
    vel = np.zeros((n, 3)); vel[:, 0] = v0
    acc = np.zeros((n, 3)); acc[:, 0] = -g
    ids = np.arange(n, dtype=np.int32)
    steps = 200
    for _ in range(steps):
        orc.move_atoms(p, pos, vel, acc, ids)
    T = steps * run.dt
    np.testing.assert_allclose(pos[:, 0], x0 + v0 * T - 0.5 * g * T * T, rtol=1e-12)
    np.testing.assert_allclose(vel[:, 0], v0 - g * T, rtol=1e-11, atol=1e-15)
    assert np.all(pos[:, 1:] == 0.0)


def test_periodic_free_flight(orc):
    run = configs.box_thermalise("fp64", n_atoms=1000, nc=8).with_(sigma=0.0)
    p = orc.params_from(run)
    hw = run.traps[0].half_width[0]
    rng = np.random.default_rng(3)
    pos = rng.uniform(-hw, hw, (1000, 3))
    vel = rng.normal(0, 0.05, (1000, 3))
    acc = np.zeros((1000, 3))
    x0, v0 = pos.copy(), vel.copy()
    ids = np.arange(1000, dtype=np.int32)
    for _ in range(500):
        orc.move_atoms(p, pos, vel, acc, ids)
    expect = (x0 + v0 * 500 * run.dt + hw) % (2 * hw) - hw
    np.testing.assert_allclose(pos, expect, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(vel, v0)
    assert np.all(pos >= -hw) and np.all(pos < hw)


def _energy_series(orc, run, steps, every=10):
    ref = orc.Reference(run)
    E = []
    for k in range(steps):
        if k % every == 0:
            e = orc.energy(ref.p, ref.s["pos"], ref.s["vel"], ref.s["atom_id"])
            E.append(e[0] + e[1])
        ref.motion_step()
    return np.array(E)


def test_energy_conservation_ip(orc):
    """energyChecker.py:170,185: max |dE/E| below 1e-5 (its 1e-3 %) for the
    reference's own trap and step without collisions."""
    run = configs.reference(20, 3000).with_(sigma=0.0, spin=False, spin_loss=False)
    E = _energy_series(orc, run, 2000)
    assert np.max(np.abs(E - E[0]) / E[0]) < 1e-5


def test_energy_error_second_order_in_dt(orc):
    """energyComparison.py:60: log-log slope of max |dE| against dt is 2 for
    velocity Verlet (1 for symplectic Euler)."""
    base = configs.reference(20, 2000).with_(sigma=0.0, spin=False, spin_loss=False)
    errs = []
    dts = [4e-6, 2e-6, 1e-6]
    total = 4e-3
    for dt in dts:
        E = _energy_series(orc, base.with_(dt=dt), int(round(total / dt)), every=1)
        errs.append(np.max(np.abs(E - E[0])))
    slope = np.polyfit(np.log(dts), np.log(errs), 1)[0]
    assert 1.7 < slope < 2.3, slope


# ---------------------------------------------------------------------------
# Index, sort, cell table, compaction
# ---------------------------------------------------------------------------
def test_index_hand_placed(orc):
    """Cell centres, lower faces, one ulp below a face, outside and NaN
    (collisions.cu:187-219): ids built from the placing indices."""
    run = configs.quadrupole_1e4()
    p = orc.params_from(run)
    nc = p.nc[0]
    hw = p.half_width[0]
    L = (2.0 * hw) / nc
    rng = np.random.default_rng(5)
    idx = rng.integers(0, nc, (500, 3))
    centre = -hw + (idx + 0.5) * L
    expect = idx[:, 2] * nc * nc + idx[:, 1] * nc + idx[:, 0]
    np.testing.assert_array_equal(orc.index_atoms(p, centre, np.arange(500, dtype=np.int32)),
                                  expect)
    # exactly on the lower face of cell 0 along x, and beyond the grid
    pts = np.array([[-hw, 0, 0], [hw, 0, 0], [np.nextafter(-hw, -1), 0, 0], [0, 0, 2 * hw],
                    [np.nan, 0, 0], [0, np.inf, 0]])
    cells = orc.index_atoms(p, pts, np.arange(len(pts), dtype=np.int32))
    mid = nc // 2
    assert cells[0] == mid * nc * nc + mid * nc + 0
    assert all(c == nc ** 3 for c in cells[1:])


def test_sort_matches_stable_argsort(orc):
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1001, 50000).astype(np.int32)
    ids = rng.permutation(50000).astype(np.int32)
    k, v = orc.sort_arrays(keys, ids, 1001)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, ids[order])


def test_cell_table_matches_bincount(orc):
    rng = np.random.default_rng(8)
    nc = 9 ** 3
    keys = np.sort(rng.integers(0, nc + 1, 20000)).astype(np.int32)
    se, count, prefix = orc.cell_table(keys, nc)
    np.testing.assert_array_equal(count, np.bincount(keys, minlength=nc + 1))
    np.testing.assert_array_equal(prefix, np.concatenate([[0], np.cumsum(count)[:-1]]))
    occ = count > 0
    assert np.all(se[~occ] == -1)
    np.testing.assert_array_equal(se[occ, 1] - se[occ, 0] + 1, count[occ])
    np.testing.assert_array_equal(keys[se[occ, 0]], np.nonzero(occ)[0])


def test_remove_if_matches_mask(orc):
    rng = np.random.default_rng(9)
    ids = rng.permutation(10000).astype(np.int32)
    tag = (rng.random(10000) < 0.37).astype(np.int32)
    np.testing.assert_array_equal(orc.remove_if(ids, tag), ids[tag == 0])


def test_d1_quirk_equals_intended_only_at_identity(orc):
    """findAtomIndex writes cellID[atomID[slot]] (collisions.cu:181) while the
    sort pairs cellID[slot] with atomID[slot]: identical to the intended
    key[slot] = cell(atom in slot) only while atomID is the identity."""
    run = configs.quadrupole_1e4()
    p = orc.params_from(run)
    st = orc.init_distribution(p, 5000, run.traps[0].T)
    ident = np.arange(5000, dtype=np.int32)
    a = orc.index_atoms(p, st["pos"], ident, d1_quirk=True)
    b = orc.index_atoms(p, st["pos"], ident, d1_quirk=False)
    np.testing.assert_array_equal(a, b)
    perm = np.random.default_rng(2).permutation(5000).astype(np.int32)
    a = orc.index_atoms(p, st["pos"], perm, d1_quirk=True)
    b = orc.index_atoms(p, st["pos"], perm, d1_quirk=False)
    assert not np.array_equal(a, b)
    np.testing.assert_array_equal(np.sort(a), np.sort(b))


# ---------------------------------------------------------------------------
# Collisions
# ---------------------------------------------------------------------------
def _dense_collision_state(orc, n=20000, nc=6, n_real=1e10, seed=4, pair_mode="uniform"):
    run = configs.Run(traps=[configs.quadrupole_trap(n, nc, n_real=n_real)], precision="fp64",
                      dt=1e-6, seed=seed, pair_mode=pair_mode)
    ref = orc.Reference(run)
    return run, ref


@pytest.mark.parametrize("pair_mode", ["uniform", "reference"])
def test_collide_conserves_momentum_and_energy(orc, pair_mode):
    run, ref = _dense_collision_state(orc, pair_mode=pair_mode)
    v0 = ref.s["vel"].copy()
    info = ref.collision_step()
    v1 = ref.s["vel"]
    assert info["total"] > 500
    changed = np.any(v0 != v1, axis=1)
    assert changed.sum() >= info["total"]
    np.testing.assert_allclose(v1.sum(axis=0), v0.sum(axis=0), rtol=0,
                               atol=1e-13 * np.abs(v0).sum())
    e0, e1 = np.sum(v0 ** 2), np.sum(v1 ** 2)
    assert abs(e1 - e0) <= 1e-13 * e0
    # sigma v_r,max only grows
    init = orc.sigvrmax_init(ref.p, run.traps[0].T)
    assert np.all(ref.sigvrmax >= init)


def test_scatter_directions_isotropic(orc):
    """Marsaglia (1972) directions replace the three normalised Gaussians of
    collisions.cu:537-546: mean 0 and second moments delta_ij / 3."""
    hw = 1e-3
    ncell = 30
    n = 2 * ncell ** 3
    run = configs.Run(traps=[configs.Trap(kind="box", n_atoms=n, nc=(ncell,) * 3,
                                          half_width=(hw,) * 3, T=20e-6, mu=0.0, n_real=n * 2e9)],
                      precision="fp64", dt=1e-6, seed=9)
    p = orc.params_from(run)
    L = 2 * hw / ncell
    idx = np.repeat(np.arange(ncell ** 3), 2)
    pos = np.stack([-hw + ((idx % ncell) + 0.5) * L, -hw + (((idx // ncell) % ncell) + 0.5) * L,
                    -hw + ((idx // ncell ** 2) + 0.5) * L], axis=1)
    vel = np.random.default_rng(1).normal(0, 0.05, (n, 3))
    ref = orc.Reference(run, pos=pos, vel=vel)
    v0 = ref.s["vel"].copy()
    ref.collision_step()
    v1 = ref.s["vel"]
    a, b = v1[0::2], v1[1::2]
    hit = np.any(v0[0::2] != a, axis=1)
    assert hit.sum() > 5000
    d = (b - a)[hit]
    nh = d / np.linalg.norm(d, axis=1)[:, None]
    m = len(nh)
    assert np.all(np.abs(nh.mean(axis=0)) < 5 / math.sqrt(m))
    second = nh.T @ nh / m
    np.testing.assert_allclose(second, np.eye(3) / 3, atol=6 / math.sqrt(m))


def test_homogeneous_collision_rate(orc):
    """homogeneousGasCollisions.py:33: n sigma <v_r> / 2 = 43.5568 s^-1 for
    N_real = 1e10 in (2 mm)^3 at 20 uK (collisionComparison.py:22-33); the
    reference's own runs give 43.28 .. 43.66 s^-1 for 1e4 .. 1e6 test
    particles (homogeneousGasCollisions.py:30-31)."""
    gold = json.loads((GOLDEN / "homogeneous_rates.json").read_text())
    analytic = configs.analytic_homogeneous_rate(1e10 / (2e-3) ** 3, 20e-6)
    assert abs(analytic - gold["analytic"]) < 5e-5
    n, ncell, steps, dt = 50000, 10, 3000, 5e-6
    run = configs.Run(traps=[configs.Trap(kind="box", n_atoms=n, nc=(ncell,) * 3,
                                          half_width=(1e-3,) * 3, T=20e-6, mu=0.0, n_real=1e10)],
                      precision="fp64", dt=dt, seed=5)
    ref = orc.Reference(run)
    ref.step(steps)
    rate = ref.collisions / (n * steps * dt)
    expected_err = 1.0 / math.sqrt(ref.collisions)
    assert abs(rate / analytic - 1) < 4 * expected_err + 0.005, (rate, analytic)
    lo = min(gold["computed"]["10"][2:] + gold["computed"]["5"][2:])
    hi = max(gold["computed"]["10"][2:] + gold["computed"]["5"][2:])
    assert lo * (1 - 4 * expected_err) < rate < hi * (1 + 4 * expected_err)


def test_reference_trial_count_truncation(orc):
    """collisions.cu:320-322 computes the trial count as (int)(Mc * (ceil(X) /
    Mc)); for pair counts Mc = n (n - 1) / 2 with n = 23, 34, 59, ... the
    product rounds below ceil(X) and a trial is lost.  In a homogeneous gas
    with about 50 atoms per cell that removes about 3.5 % of the collisions:
    the reference pair mode reproduces it, the unbiased mode does not."""
    bad = [n for n in range(2, 200) if int((0.5 * (n - 1) * n) * (1.0 / (0.5 * (n - 1) * n))) == 0]
    assert bad[:5] == [23, 34, 59, 76, 84]
    rates = {}
    for mode in ("uniform", "reference"):
        n, steps, dt = 50000, 1500, 5e-6
        run = configs.Run(traps=[configs.Trap(kind="box", n_atoms=n, nc=(10,) * 3,
                                              half_width=(1e-3,) * 3, T=20e-6, mu=0.0,
                                              n_real=1e10)],
                          precision="fp64", dt=dt, seed=6, pair_mode=mode)
        ref = orc.Reference(run)
        ref.step(steps)
        rates[mode] = ref.collisions / (n * steps * dt)
    assert rates["reference"] < rates["uniform"] * 0.985


def test_trap_rate_formula_matches_checker():
    """collisionChecker.py:34 at its own parameters (N = 1e6, 10 uK, B' = 2.5,
    gs = 0.5): 18.135 s^-1; our formula takes the moment explicitly."""
    r = configs.analytic_trap_rate(1e6, 10e-6, 2.5, mu=0.5 * configs.GS * configs.MUB)
    assert abs(r - 18.135) < 5e-3


# ---------------------------------------------------------------------------
# Spin
# ---------------------------------------------------------------------------
def test_spinor_rotation_equals_bloch_rotation(orc):
    """The SU(2) update of wavefunction.cu:47-70 maps to a rotation of the
    Bloch vector by phi = 2 theta about B (the form the FP32 path evolves)."""
    run = configs.reference(20, 500)
    p = orc.params_from(run)
    st = orc.init_distribution(p, 500, 20e-6, spin=True)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(500, 2)); d = rng.normal(size=(500, 2))
    nrm = np.sqrt((u ** 2).sum(1) + (d ** 2).sum(1))[:, None]
    u /= nrm; d /= nrm
    def bloch(u, d):
        return np.stack([2 * (u[:, 0] * d[:, 0] + u[:, 1] * d[:, 1]),
                         2 * (u[:, 0] * d[:, 1] - u[:, 1] * d[:, 0]),
                         (u ** 2).sum(1) - (d ** 2).sum(1)], axis=1)
    s = np.ascontiguousarray(bloch(u, d))
    ids = np.arange(500, dtype=np.int32)
    dt = 3e-7
    uu, dd = np.ascontiguousarray(u), np.ascontiguousarray(d)
    orc.evolve_wavefunction(p, st["pos"], uu, dd, ids, dt)
    orc.rotate_bloch(p, st["pos"], s, ids, dt)
    np.testing.assert_allclose(bloch(uu, dd), s, atol=1e-12)


def test_aligned_spin_and_uniform_precession(orc):
    """getAlignedSpinUp/Dn (setUp.cu:236-252) give a Bloch vector along B; a
    spin perpendicular to a uniform field precesses by exactly
    phi = gs muB |B| t / hbar."""
    run = configs.reference(20, 200)
    p = orc.params_from(run)
    st = orc.init_distribution(p, 200, 20e-6, spin=True)
    u, d = st["psi_up"], st["psi_dn"]
    s = np.stack([2 * (u[:, 0] * d[:, 0] + u[:, 1] * d[:, 1]),
                  2 * (u[:, 0] * d[:, 1] - u[:, 1] * d[:, 0]),
                  (u ** 2).sum(1) - (d ** 2).sum(1)], axis=1)
    # B at the trap centre is B0 z^ (IP field), so compare at the origin
    pos0 = np.zeros((1, 3))
    sx = np.array([[1.0, 0.0, 0.0]])
    t = 7e-8
    orc.rotate_bloch(p, pos0, sx, np.array([0], np.int32), t)
    phi = run.gs * configs.MUB * 0.01 * t / configs.HBAR
    np.testing.assert_allclose(sx[0], [math.cos(phi), math.sin(phi), 0.0], atol=1e-14)
    # the aligned spinors point along B written out independently
    for i in range(0, 200, 17):
        x, y, z = st["pos"][i]
        B = np.array([20 * x - 2e4 * x * z, -20 * y - 2e4 * y * z,
                      0.01 + 2e4 * (z * z - 0.5 * (x * x + y * y))])
        np.testing.assert_allclose(s[i], B / np.linalg.norm(B), atol=1e-12)


# ---------------------------------------------------------------------------
# Initial distribution
# ---------------------------------------------------------------------------
def test_quadrupole_initial_distribution_moments(orc):
    """Positions: in (x, y, 2z) the radius is Gamma(3, r_th) -> <mu B' rho> = 3 kT;
    velocities Maxwell: <m v^2 / 2> = 3 kT / 2."""
    from scipy import stats
    run = configs.quadrupole_1e4()
    t = run.traps[0]
    p = orc.params_from(run)
    n = 40000
    st = orc.init_distribution(p, n, t.T)
    x, y, z = st["pos"].T
    rho = np.sqrt(x * x + y * y + 4 * z * z)
    kT = configs.KB * t.T
    rth = kT / (t.mu * t.dBdz)
    assert stats.kstest(rho / rth, stats.gamma(3).cdf).pvalue > 1e-3
    ek = 0.5 * run.mass * (st["vel"] ** 2).sum(1)
    assert abs(ek.mean() / (1.5 * kT) - 1) < 0.02
    for d in range(3):
        assert stats.kstest(st["vel"][:, d] / math.sqrt(kT / run.mass), "norm").pvalue > 1e-3
    # isotropy of the direction in (x, y, 2z)
    u = np.stack([x, y, 2 * z], 1) / rho[:, None]
    assert np.all(np.abs(u.mean(0)) < 0.02)


# ---------------------------------------------------------------------------
# Field: Maxwell's equations in vacuum (div B = 0, curl B = 0)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["ip", "quadrupole"])
def test_field_is_divergence_and_curl_free(orc, kind):
    """A sign or index slip in either field (moveAtoms.cu:135-142, or the
    quadrupole B' (x, y, -2z)) breaks div B = 0 or curl B = 0; |B| is the
    magnitude the potential uses (U = mu (|B| - B0))."""
    run = configs.reference() if kind == "ip" else configs.quadrupole_1e4()
    p = orc.params_from(run)
    t = run.traps[0]
    rng = np.random.default_rng(7)
    h = 1e-9
    for _ in range(25):
        r = rng.normal(0, 5e-5, 3)
        J = np.zeros((3, 3))  # J[i, j] = dB_i / dx_j
        for j in range(3):
            e = np.zeros(3); e[j] = h
            J[:, j] = (orc.field(p, r + e) - orc.field(p, r - e)) / (2 * h)
        scale = np.abs(J).max()
        assert abs(np.trace(J)) < 1e-6 * scale
        curl = np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])
        assert np.all(np.abs(curl) < 1e-6 * scale)
        B = orc.field(p, r)
        b0 = t.B0 if kind == "ip" else 0.0
        assert orc.potential(p, r) == pytest.approx(t.mu * (np.linalg.norm(B) - b0), rel=1e-12)


# ---------------------------------------------------------------------------
# Evaporation tag, temperature, energy
# ---------------------------------------------------------------------------
def test_evaporation_tag_spin_projection_and_energy_cut(orc):
    """evaporation.cu:90-94: an atom is lost when its spin points against B.
    Spinors aligned with B (setUp.cu:236-252) survive; the orthogonal spinor
    (-conj d, conj u) has the opposite Bloch vector and is lost.  Energy cut:
    lost iff 1/2 m v^2 + U > E_cut, with U from the independent potential;
    lost atoms have their position and velocity recorded."""
    run = configs.reference(20, 400)
    p = orc.params_from(run)
    st = orc.init_distribution(p, 400, 20e-6, spin=True)
    ids = np.arange(400, dtype=np.int32)
    tag, _, _ = orc.evaporation_tag(p, st["pos"], st["vel"], ids, st["psi_up"], st["psi_dn"])
    assert not tag.any()
    u, d = st["psi_up"], st["psi_dn"]
    flip = np.arange(400) % 3 == 0
    u2, d2 = u.copy(), d.copy()
    u2[flip] = np.stack([-d[flip, 0], d[flip, 1]], 1)   # -conj(d)
    d2[flip] = np.stack([u[flip, 0], -u[flip, 1]], 1)   # conj(u)
    tag, epos, evel = orc.evaporation_tag(p, st["pos"], st["vel"], ids, u2, d2)
    np.testing.assert_array_equal(tag.astype(bool), flip)
    np.testing.assert_array_equal(epos[flip], st["pos"][flip])
    np.testing.assert_array_equal(evel[flip], st["vel"][flip])
    assert np.all(epos[~flip] == 0.0)
    # energy cut alone
    E = np.array([0.5 * run.mass * (v @ v) + _potential(run, r) for r, v in zip(st["pos"], st["vel"])])
    cut = float(np.median(E))
    margin = np.abs(E - cut) > 1e-9 * cut
    tag, _, _ = orc.evaporation_tag(p, st["pos"], st["vel"], ids, e_cut=cut)
    np.testing.assert_array_equal(tag.astype(bool)[margin], (E > cut)[margin])


def test_temperature_energy_and_virial(orc):
    """calculateTemp (evaporation.cu:108-128) on equal speeds gives
    m v^2 / (3 kB) exactly; on a thermal cloud T -> T0.  The quadrupole
    potential is homogeneous of degree 1, so in equilibrium <U> = 3 kT
    (= 2 <Ek>) -- a relation between the sampler and the potential."""
    run = configs.quadrupole_1e4()
    t = run.traps[0]
    p = orc.params_from(run)
    n = 1000
    rng = np.random.default_rng(2)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    v = np.ascontiguousarray(0.03 * dirs)
    ids = np.arange(n, dtype=np.int32)
    T = orc.calculate_temp(p, v, ids)
    assert T == pytest.approx(run.mass * 0.03 ** 2 / (3 * configs.KB), rel=1e-13)
    ek, ep, v2 = orc.energy(p, np.zeros((n, 3)), v, ids)
    assert ek == pytest.approx(n * 0.5 * run.mass * 0.03 ** 2, rel=1e-13)
    assert ep == 0.0 and v2 == pytest.approx(n * 0.03 ** 2, rel=1e-13)
    st = orc.init_distribution(p, 100000, t.T)
    ids = np.arange(100000, dtype=np.int32)
    kT = configs.KB * t.T
    assert orc.calculate_temp(p, st["vel"], ids) == pytest.approx(t.T, rel=0.01)
    ek, ep, _ = orc.energy(p, st["pos"], st["vel"], ids)
    assert ep / 100000 == pytest.approx(3 * kT, rel=0.01)
    assert ep / ek == pytest.approx(2.0, rel=0.015)


def test_sigvrmax_init_is_mean_relative_speed(orc):
    """initSigvrmax (setUp.cu:287-292): sqrt(16 kT / (pi m)) sigma is sigma
    times the mean relative speed of two Maxwellian atoms; check against
    sampled pairs of the initial velocities."""
    run = configs.quadrupole_1e4()
    t = run.traps[0]
    p = orc.params_from(run)
    st = orc.init_distribution(p, 200000, t.T)
    v = st["vel"]
    vr = np.linalg.norm(v[:100000] - v[100000:], axis=1).mean()
    assert orc.sigvrmax_init(p, t.T) == pytest.approx(vr * run.sigma, rel=0.01)


def test_ip_initial_distribution_matches_rejection_density(orc):
    """The rejection sampler of setUp.cu:178-205 proposes from a Gaussian of
    width W/3 per axis and accepts with probability exp(-U / kT) (capped at
    1 where U < 0, in the field zeros of the IP expansion), so positions
    follow q(r) min(1, exp(-U / kT)).  Independent check: importance-weighted
    numpy proposals with the independently written IP potential; the
    quantiles of |x|, |y|, |z| must agree."""
    run = configs.reference(20, 20000)
    t = run.traps[0]
    p = orc.params_from(run)
    st = orc.init_distribution(p, 20000, t.T)
    kT = configs.KB * t.T
    rng = np.random.default_rng(11)
    m = 4_000_000
    X, Y, Z = (rng.normal(0.0, hw / 3.0, m) for hw in t.half_width)
    Bx = t.dBdx * X - 0.5 * t.d2Bdx2 * X * Z
    By = -t.dBdx * Y - 0.5 * t.d2Bdx2 * Y * Z
    Bz = t.B0 + 0.5 * t.d2Bdx2 * (Z * Z - 0.5 * (X * X + Y * Y))
    U = t.mu * (np.sqrt(Bx * Bx + By * By + Bz * Bz) - t.B0)
    w = np.minimum(1.0, np.exp(-U / kT))
    for q, Q in zip(st["pos"].T, (X, Y, Z)):
        order = np.argsort(np.abs(Q))
        cdf = np.cumsum(w[order]) / w.sum()
        for f in (0.25, 0.5, 0.75, 0.9, 0.99):
            expect = float(np.abs(Q)[order][np.searchsorted(cdf, f)])
            assert float(np.quantile(np.abs(q), f)) == pytest.approx(expect, rel=0.035)


def test_box_initial_distribution(orc):
    """Periodic box (homogeneous-gas configuration): positions uniform over
    the box, per-axis Maxwell velocities at Tx = 2 Ty = 2 Tz."""
    from scipy import stats
    run = configs.box_thermalise("fp64", n_atoms=50000, nc=16)
    t = run.traps[0]
    p = orc.params_from(run)
    st = orc.init_distribution(p, 50000, t.T, t.T_axis)
    for d in range(3):
        hw = t.half_width[d]
        assert stats.kstest((st["pos"][:, d] + hw) / (2 * hw), "uniform").pvalue > 1e-3
        s = math.sqrt(configs.KB * t.T_axis[d] / run.mass)
        assert stats.kstest(st["vel"][:, d] / s, "norm").pvalue > 1e-3


def test_run_loop_equals_step_composition(orc):
    """orc_run (the CPU baseline loop) is exactly index -> sort -> table ->
    collide -> lpc x move, the composition the Reference class performs."""
    run = configs.quadrupole_1e4().with_(lpc=2)
    t = run.traps[0]
    ref = orc.Reference(run.with_(traps=[configs.quadrupole_trap(3000, 20)]))
    p = ref.p
    st = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in ref.s.items()}
    sv = ref.sigvrmax.copy()
    res = orc.run(p, st, 3, 0, sv)
    ref.step(6)
    np.testing.assert_array_equal(st["pos"], ref.s["pos"])
    np.testing.assert_array_equal(st["vel"], ref.s["vel"])
    np.testing.assert_array_equal(sv, ref.sigvrmax)
    assert res["collisions"] == ref.collisions and res["atom_steps"] == 6 * 3000
    del t


# ---------------------------------------------------------------------------
# Majorana spin-flip model (wavefunction.cu:75-237)
# ---------------------------------------------------------------------------
def _random_spinors(rng, n):
    u = rng.normal(size=(n, 2)); d = rng.normal(size=(n, 2))
    nrm = np.sqrt((u ** 2).sum(1) + (d ** 2).sum(1))[:, None]
    return np.ascontiguousarray(u / nrm), np.ascontiguousarray(d / nrm)


def _field_dirs(orc, p, pos):
    B = np.array([orc.field(p, r) for r in pos])
    return B / np.linalg.norm(B, axis=1)[:, None]


def test_local_populations_are_projections_on_the_field(orc):
    """getLocalPopulations: P_up = |<up_B|psi>|^2 with up_B the spin-up
    eigenvector of B.sigma, computed here with complex 2x2 algebra; P_up +
    P_down = 1 for a normalised spinor."""
    run = configs.reference(20, 300)
    p = orc.params_from(run)
    st = orc.init_distribution(p, 300, 20e-6, spin=True)
    rng = np.random.default_rng(11)
    u, d = _random_spinors(rng, 300)
    ids = np.arange(300, dtype=np.int32)
    pops = orc.local_populations(p, st["pos"], u, d, ids)
    Bn = _field_dirs(orc, p, st["pos"])
    psi = np.stack([u[:, 0] + 1j * u[:, 1], d[:, 0] + 1j * d[:, 1]], axis=1)
    for i in range(300):
        bx, by, bz = Bn[i]
        H = np.array([[bz, bx - 1j * by], [bx + 1j * by, -bz]])   # B.sigma
        w, V = np.linalg.eigh(H)
        up = V[:, np.argmax(w)]
        assert pops[i, 0] == pytest.approx(abs(np.vdot(up, psi[i])) ** 2, abs=1e-13)
    np.testing.assert_allclose(pops.sum(1), 1.0, atol=1e-14)


def test_exponential_decay_is_partial_projection(orc):
    """exponentialDecay: psi <- P+ psi + exp(-dt/tau) P- psi with P+- =
    (1 +- B.sigma)/2 (matrix form here); dt = 0 is the identity and a very
    long dt leaves the spin-up component along B."""
    run = configs.reference(20, 200)
    p = orc.params_from(run)
    st = orc.init_distribution(p, 200, 20e-6, spin=True)
    rng = np.random.default_rng(12)
    ids = np.arange(200, dtype=np.int32)
    Bn = _field_dirs(orc, p, st["pos"])
    for dt in (0.0, 3e-6, 1e3):
        u, d = _random_spinors(rng, 200)
        psi = np.stack([u[:, 0] + 1j * u[:, 1], d[:, 0] + 1j * d[:, 1]], axis=1)
        orc.exponential_decay(p, st["pos"], u, d, ids, dt)
        for i in range(200):
            bx, by, bz = Bn[i]
            H = np.array([[bz, bx - 1j * by], [bx + 1j * by, -bz]])
            Pp, Pm = (np.eye(2) + H) / 2, (np.eye(2) - H) / 2
            e = math.exp(-dt / orc.decoherence_tau(p, st["pos"][i]))
            want = Pp @ psi[i] + e * (Pm @ psi[i])
            got = np.array([u[i, 0] + 1j * u[i, 1], d[i, 0] + 1j * d[i, 1]])
            np.testing.assert_allclose(got, want, atol=1e-13)
            if dt == 0.0:
                np.testing.assert_allclose(got, psi[i], atol=1e-15)


def test_decoherence_time_from_the_field_gradient(orc):
    """calculateTau with getAcc: the relative acceleration 2 |mu grad|B|| / m is
    taken here from a finite difference of |B|; tau is the cube-harmonic mean of
    1.163 sqrt(sigma / a) and sqrt(2 / pi) hbar / (m sigma a)."""
    run = configs.reference(20, 50)
    p = orc.params_from(run)
    st = orc.init_distribution(p, 50, 20e-6)
    h = 1e-9
    for r in st["pos"][:20]:
        g = np.zeros(3)
        for k in range(3):
            e = np.zeros(3); e[k] = h
            g[k] = (np.linalg.norm(orc.field(p, r + e)) - np.linalg.norm(orc.field(p, r - e))) / (2 * h)
        a = 2.0 * p.mu * np.linalg.norm(g) / p.mass
        tp = 1.163 * math.sqrt(p.sigma / a)
        tv = math.sqrt(2.0 / configs.PI_REF) * configs.HBAR / (p.mass * p.sigma * a)
        want = (tp ** -3 + tv ** -3) ** (-1.0 / 3.0)
        assert orc.decoherence_tau(p, r) == pytest.approx(want, rel=1e-6)


def test_flips_conserve_energy_and_follow_the_population_change(orc):
    """flipAtoms: an atom whose spin-down population grew by q since the
    populations were taken flips with probability q (binomial, 5 sigma); a
    flip keeps the velocity direction and conserves E_k + U, with U_down =
    -U_up - 2 mu B0 (a potential step of 2 mu |B|)."""
    run = configs.reference(20, 20000)
    p = orc.params_from(run)
    n = 20000
    st = orc.init_distribution(p, n, 20e-6, spin=True)
    ids = np.arange(n, dtype=np.int32)
    pops = orc.local_populations(p, st["pos"], st["psi_up"], st["psi_dn"], ids)
    np.testing.assert_allclose(pops[:, 0], 1.0, atol=1e-12)   # aligned spin up
    # rotate every spinor so that its spin-down population becomes q:
    # psi = sqrt(1-q) up_B + sqrt(q) down_B
    q = 0.3
    Bn = _field_dirs(orc, p, st["pos"])
    u = np.zeros((n, 2)); d = np.zeros((n, 2))
    for i in range(n):
        bx, by, bz = Bn[i]
        H = np.array([[bz, bx - 1j * by], [bx + 1j * by, -bz]])
        w, V = np.linalg.eigh(H)
        psi = math.sqrt(1 - q) * V[:, 1] + math.sqrt(q) * V[:, 0]
        u[i] = psi[0].real, psi[0].imag
        d[i] = psi[1].real, psi[1].imag
    vel = st["vel"] * 10.0   # kinetic energy far above the potential step
    v0 = vel.copy()
    up = np.ones(n, np.uint8)
    e0 = orc.energy(p, st["pos"], vel, ids, up)
    k = orc.flip_atoms(p, st["pos"], vel, u, d, pops, up, ids, 7)
    assert abs(k - q * n) < 5 * math.sqrt(n * q * (1 - q))
    flipped = up == 0
    assert flipped.sum() == k
    e1 = orc.energy(p, st["pos"], vel, ids, up)
    assert (e1[0] + e1[1]) == pytest.approx(e0[0] + e0[1], rel=1e-12)
    cosang = (vel * v0).sum(1) / np.linalg.norm(vel, axis=1) / np.linalg.norm(v0, axis=1)
    np.testing.assert_allclose(cosang, 1.0, atol=1e-12)
    np.testing.assert_array_equal(vel[~flipped], v0[~flipped])
    # the same draws again flip exactly the same atoms
    up2 = np.ones(n, np.uint8)
    vel2 = v0.copy()
    assert orc.flip_atoms(p, st["pos"], vel2, u, d, pops, up2, ids, 7) == k
    np.testing.assert_array_equal(up2, up)

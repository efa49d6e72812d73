"""CPU checks of the HDF5 output (include/dsmc_h5.h): the library exports
what the header declares; datasets have the reference's layout (rank 3,
(rows, cols, t), t growing by one per append: hdf5Helpers.cu:13-163) and read
back exactly; the energy checker (energyChecker.py:118-124, 170-185) agrees
with closed forms on the trap axis."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from cudadsmc_amd import configs, output

ROOT = Path(__file__).resolve().parent.parent

pytestmark = pytest.mark.skipif(not output.available(),
                                reason="libdsmc_h5.so not built (no libhdf5 with headers)")


def test_h5_library_exports_every_declared_function():
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "dsmc_h5.h").read_text(), flags=re.S)
    names = sorted(set(re.findall(r"\b(dsmc_h5_[a-z_0-9]+)\s*\(", text)))
    lib = C.CDLL(str(output.H5_LIB_PATH))
    assert len(names) == 8
    assert not [n for n in names if not hasattr(lib, n)]


def test_datasets_round_trip(tmp_path):
    f = tmp_path / "out.h5"
    rng = np.random.default_rng(5)
    pos = [rng.normal(size=(7, 3)) for _ in range(3)]
    ids = [rng.integers(0, 1000, size=(7, 1)).astype(np.int32) for _ in range(3)]
    up = [rng.integers(0, 2, size=(7, 1)).astype(np.uint8) for _ in range(3)]
    with output.H5Writer(str(f)) as w:
        w.dataset("/atomData/positions", 7, 3, output.DOUBLE)
        w.dataset("/atomData/atomID", 7, 1, output.INT)
        w.dataset("/atomData/atomIsSpinUp", 7, 1, output.BOOL)
        for i in range(3):
            w.append("/atomData/positions", pos[i], output.DOUBLE)
            w.append("/atomData/atomID", ids[i], output.INT)
            w.append("/atomData/atomIsSpinUp", up[i], output.BOOL)
    p = output.read(str(f), "/atomData/positions", output.DOUBLE)
    assert p.shape == (7, 3, 3)
    for i in range(3):
        assert np.array_equal(p[:, :, i], pos[i])
        assert np.array_equal(output.read(str(f), "/atomData/atomID", output.INT)[:, :, i], ids[i])
        assert np.array_equal(output.read(str(f), "/atomData/atomIsSpinUp", output.BOOL)[:, :, i], up[i])


def test_errors_are_reported(tmp_path):
    with output.H5Writer(str(tmp_path / "e.h5")) as w:
        w.dataset("/atomData/x", 2, 2)
        with pytest.raises(RuntimeError, match="exists"):
            w.dataset("/atomData/x", 2, 2)
        with pytest.raises(RuntimeError, match="no dataset"):
            w.append("/atomData/y", np.zeros((2, 2)))
        with pytest.raises(RuntimeError, match=">= 1"):
            w.dataset("/atomData/z", 0, 2)
    with pytest.raises(RuntimeError, match="cannot open"):
        output.read(str(tmp_path / "missing.h5"), "/atomData/x")


def test_energy_checker_on_the_trap_axis():
    """On the z axis of the Ioffe-Pritchard trap B = (0, 0, B0 + B'' z^2 / 2),
    so E_p = gs muB B'' z^2 / 4; at rest E_k = 0; a velocity v adds m v^2 / 2."""
    B0, Bp, Bpp = 0.01, 20.0, 40000.0
    z = np.array([0.0, 1e-4, -2e-4, 3e-4])
    n = z.size
    frames = {
        "positions": np.zeros((n, 3, 2)), "velocities": np.zeros((n, 3, 2)),
        "atomIsSpinUp": np.ones((n, 1, 2), dtype=np.uint8),
        "atomNumber": np.full((1, 1, 2), n, dtype=np.int32),
    }
    frames["positions"][:, 2, :] = z[:, None]
    frames["velocities"][:, 0, 1] = 0.01
    Ek, Ep, Et = output.energies(frames, B0=B0, dBdx=Bp, d2Bdx2=Bpp)
    ep = np.mean(configs.GS * configs.MUB * Bpp * z * z / 4.0) / configs.KB * 1e6
    ek = 0.5 * configs.MRB * 0.01 ** 2 / configs.KB * 1e6
    assert Ek[0] == 0.0 and Ep[0] == pytest.approx(ep, rel=1e-12)
    assert Ek[1] == pytest.approx(ek, rel=1e-12) and Ep[1] == pytest.approx(ep, rel=1e-12)
    dE, ok = output.energy_check(frames, B0=B0, dBdx=Bp, d2Bdx2=Bpp)
    assert dE == pytest.approx(abs(ek) / ep * 100.0, rel=1e-9) and not ok


def _rate_file(path, per_frame, time, n_atoms, cells=64):
    """A file of the reference's layout holding only the datasets the
    collision checkers read."""
    T = len(time)
    with output.H5Writer(str(path)) as w:
        w.dataset("/atomData/collisionCount", cells + 1, 1, output.INT)
        w.dataset("/atomData/simuatedTime", 1, 1, output.DOUBLE)
        w.dataset("/atomData/atomNumber", 1, 1, output.INT)
        for i in range(T):
            w.append("/atomData/collisionCount", per_frame[:, i:i + 1], output.INT)
            w.append("/atomData/simuatedTime", np.array([[time[i]]]), output.DOUBLE)
            w.append("/atomData/atomNumber", np.array([[n_atoms[i]]], np.int32), output.INT)
    return {k: output.read(str(path), f"/atomData/{k}", t)
            for k, t in (("collisionCount", output.INT), ("simuatedTime", output.DOUBLE),
                         ("atomNumber", output.INT))}


def test_collision_rate_readers(tmp_path):
    """The two collision-rate checkers on files of a Poisson process of known
    rate: collisionChecker.py reads per-frame counts, collisionComparison.py
    cumulative ones; both must return the rate (events per atom per second)
    within the Poisson error, and the per-frame reader on cumulative counts
    (or the other way round) must not."""
    rng = np.random.default_rng(9)
    R, N, dt, T, cells = 250.0, 20000, 1e-4, 400, 64
    time = np.arange(T) * dt
    n = np.full(T, N)
    lam = R * N * dt                      # events per frame
    counts = rng.poisson(lam / cells, size=(cells + 1, T)).astype(np.int32)
    counts[cells] = 0                     # the overflow cell holds none
    per = _rate_file(tmp_path / "per.h5", counts, time, n, cells)
    cum = _rate_file(tmp_path / "cum.h5", np.cumsum(counts, axis=1).astype(np.int32), time, n, cells)
    m1, s1 = output.rate_summary(output.collision_rate(per))
    m2, s2 = output.rate_summary(output.collision_rate_cumulative(cum))
    err = R / np.sqrt(lam * (T - 11))
    assert abs(m1 - R) < 5 * err and abs(m2 - R) < 5 * err
    assert s1 == pytest.approx(R / np.sqrt(lam), rel=0.15)
    assert output.rate_summary(output.collision_rate(cum))[0] > 10 * R
    assert output.rate_summary(output.collision_rate_cumulative(per))[0] < 0.1 * R

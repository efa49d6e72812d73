"""GPU checks of the HDF5 output of a running simulation: the thirteen
datasets of main.cu:202-383 with one frame per write (main.cu:526-576),
frames equal to what dsmc_copy_out returns at that time, and the C++ driver's
file in the same layout."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs, output

ROOT = Path(__file__).resolve().parent.parent

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not output.available(), reason="libdsmc_h5.so not built")]


def test_frames_of_the_reference_run(tmp_path):
    f = str(tmp_path / "ref.h5")
    run = configs.reference(Nc=10, N=2000)
    print("h5 frames: reference run Nc=10 N=2000", flush=True)
    snaps = []
    with Simulation(run) as sim, output.H5Writer(f) as w:
        sim.init_distribution()
        t = 0.0
        for k in range(3):
            if k:
                sim.step(20)
                t += 20 * run.dt
            w.write_frame(sim, t)
            snaps.append((sim.copy_out(_abi.FIELD_POS), sim.copy_out(_abi.FIELD_ATOM_ID), t))
    fr = output.read_frames(f)
    N = 2000
    assert fr["positions"].shape == (N, 3, 3)
    assert fr["psiUp"].shape == (N, 2, 3)
    assert fr["collisionCount"].shape == (10 ** 3 + 1, 1, 3)
    assert fr["simuatedTime"].shape == (1, 1, 3)
    for k, (pos, ids, t) in enumerate(snaps):
        n = pos.shape[0]
        assert int(fr["atomNumber"][0, 0, k]) == n
        assert np.array_equal(fr["positions"][:n, :, k], pos)
        assert np.array_equal(fr["atomID"][:n, 0, k], ids)
        assert np.all(fr["positions"][n:, :, k] == 0.0)
        assert fr["simuatedTime"][0, 0, k] == t
        assert np.all(fr["atomIsSpinUp"][:n, 0, k] == 1)
        # normalised spinors
        s = (fr["psiUp"][:n, :, k] ** 2).sum(1) + (fr["psiDn"][:n, :, k] ** 2).sum(1)
        assert np.allclose(s, 1.0, atol=1e-12)
    # the per-cell counts of the last cell table hold every atom alive then
    # (spin-flip losses since only remove atoms)
    assert N >= int(fr["atomCount"][:, 0, 2].sum()) >= int(fr["atomNumber"][0, 0, 2])
    dE, _ = output.energy_check(fr)
    print(f"energy change over the frames {dE:.3g} %", flush=True)
    assert np.isfinite(dE) and dE < 1.0


def test_frame_of_a_batched_fp32_run(tmp_path):
    f = str(tmp_path / "b.h5")
    run = configs.batched_traps(n_traps=4, seed=3)
    with Simulation(run) as sim, output.H5Writer(f) as w:
        sim.init_distribution()
        sim.step(5)
        w.write_frame(sim, 5 * run.dt)
        pos = sim.copy_out(_abi.FIELD_POS)
        cap = sim.info().n_capacity
    fr = output.read_frames(f)
    n = pos.shape[0]
    assert fr["positions"].shape == (cap, 3, 1)
    assert np.array_equal(fr["positions"][:n, :, 0], pos)
    assert fr["collisionCount"].shape == (1, 1, 1)   # per-cell counts: FP64 form only
    assert int(fr["atomNumber"][0, 0, 0]) == n


def test_driver_writes_the_reference_file(tmp_path):
    out = tmp_path / "outputData.h5"
    exe = ROOT / "cudadsmc_amd" / "bin" / "cudadsmc"
    print("driver: cudadsmc 8 500 --prints 2", flush=True)
    res = subprocess.run([str(exe), "8", "500", "--prints", "2", "--out", str(out)],
                         capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    fr = output.read_frames(str(out))
    assert fr["positions"].shape == (500, 3, 3)
    assert fr["atomCount"].shape == (8 ** 3 + 1, 1, 3)
    t = fr["simuatedTime"][0, 0, :]
    assert t[0] == 0.0 and np.all(np.diff(t) > 0)


def test_driver_file_collision_rate_against_analytic(tmp_path):
    """collisionChecker.py's rate read from the C++ driver's HDF5 file of a
    quadrupole run (N = 1e4 test particles for 2e8 real atoms at 20 uK,
    B' = 2.5 T/m, 35^3 cells, FP64 form, a frame after every step) against
    the analytic trap rate (collisionChecker.py:34) within 4 sigma + 3 %
    (the 35^3 grid lowers the rate by about 1 %).  The gas is dilute there
    (about 5e3 events per atom per second: a mean free time of 200 steps and
    a mean free path above the cell size), as in the reference's own checker
    run (N = 1e6, 100^3 cells)."""
    out = tmp_path / "quad.h5"
    exe = ROOT / "cudadsmc_amd" / "bin" / "cudadsmc"
    n, nc, steps, n_real, skip = 10000, 35, 120, 2e8, 20
    cmd = [str(exe), "--quadrupole", str(n), str(nc), str(steps), "fp64", "--prints", str(steps),
           "--nreal", str(n_real), "--out", str(out)]
    print("driver: " + " ".join(cmd[1:]), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    fr = output.read_frames(str(out))
    assert fr["collisionCount"].shape == (nc ** 3 + 1, 1, steps + 1)
    rate = output.collision_rate(fr)
    mean, std = output.rate_summary(rate, skip=skip)
    events = int(fr["collisionCount"][:, 0, skip:-1].sum())
    expect = configs.analytic_trap_rate(n_real, 20e-6, 2.5)
    half = len(rate[skip:-1]) // 2
    print(f"rate {mean:.1f} +- {std:.1f} /s vs analytic {expect:.1f} /s ({events} events); "
          f"halves {np.mean(rate[skip:skip + half]):.1f}, {np.mean(rate[skip + half:-1]):.1f}",
          flush=True)
    assert events > 3000
    assert abs(mean / expect - 1.0) < 4.0 / np.sqrt(events) + 0.03

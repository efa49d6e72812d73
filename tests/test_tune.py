"""Cells-per-trap autotuner of the batched form: the regridding on the CPU,
the timed choice through libdsmc.so on the GPU."""
import math

import pytest

from cudadsmc_amd import configs, tune


def _occupancy(n_atoms, nc):
    # test particles in the central cell of an nc^3 grid over +-5 thermal radii
    return 1000.0 * n_atoms / (4.0 * math.pi * nc ** 3)


def test_regrid_hits_the_target_occupancy():
    run = configs.batched_traps(n_traps=32, steps=10)
    for p in tune.PEAKS:
        r = tune.regrid(run, p)
        for t0, t in zip(run.traps, r.traps):
            nc = t.nc[0]
            assert t.nc == (nc, nc, nc)
            assert t.half_width == t0.half_width and t.n_atoms == t0.n_atoms
            if 32 < nc < 64:   # not clamped: one grid step brackets the target
                assert _occupancy(t.n_atoms, nc + 1) <= p <= _occupancy(t.n_atoms, nc - 1)
            else:
                assert nc in (32, 64)
    # finer grids for lower targets, never coarser
    grids = [[t.nc[0] for t in tune.regrid(run, p).traps] for p in tune.PEAKS]
    for lo, hi in zip(grids, grids[1:]):
        assert all(a >= b for a, b in zip(lo, hi))
    # the default workload is the 8-per-central-cell grid
    assert [t.nc for t in tune.regrid(run, 8.0).traps] == [t.nc for t in run.traps]


def test_grid_record_reports_clamped_traps():
    """The tuner's record says how many traps reached the target: with the
    workload's 32..64 range many traps sit at its ends; with a range wide
    enough for every trap none does, and every central cell is within one
    grid step of the target."""
    run = configs.batched_traps(n_traps=64, steps=10)
    r = tune.regrid(run, 8.0)
    rec = tune.grid_record(r, 8.0)
    ncs = [t.nc[0] for t in r.traps]
    assert rec["traps_unclamped"] == sum(32 < n < 64 for n in ncs)
    assert rec["traps_unclamped"] < len(ncs)
    assert rec["cells"] == sum(n ** 3 for n in ncs)
    lo, med, hi = rec["central_peak"]
    assert lo <= med <= hi
    wide = tune.regrid(run, 8.0, nc_range=(4, 256))
    rec_w = tune.grid_record(wide, 8.0, nc_range=(4, 256))
    assert rec_w["traps_unclamped"] == len(ncs)
    for t in wide.traps:
        n = t.nc[0]
        assert _occupancy(t.n_atoms, n + 1) <= 8.0 <= _occupancy(t.n_atoms, n - 1)
    with pytest.raises(ValueError):
        tune.regrid(run, 8.0, nc_range=(64, 32))


def test_regrid_rejects_what_it_does_not_tune():
    run = configs.batched_traps(n_traps=4, steps=10)
    with pytest.raises(ValueError):
        tune.regrid(run, 2.0)
    with pytest.raises(ValueError):
        tune.regrid(configs.reference(), 8.0)   # Ioffe-Pritchard grid
    with pytest.raises(ValueError):
        tune.autotune_cells(run.with_(precision="fp64"))


@pytest.mark.gpu
def test_cells_autotuner_keeps_the_fastest_candidate():
    run = configs.batched_traps(n_traps=16, steps=100).with_(obs_capacity=256)
    print("cells autotuner: 16 traps, targets 4 and 16", flush=True)
    best, report = tune.autotune_cells(run, peaks=(4.0, 16.0), warm=4, steps=10)
    assert [r["target_peak"] for r in report] == [4.0, 16.0]
    assert all(r["ms_per_step"] > 0 and r["block"] > 0 for r in report)
    assert report[0]["cells"] > report[1]["cells"]
    pick = min(report, key=lambda r: r["ms_per_step"])["target_peak"]
    assert [t.nc for t in best.traps] == [t.nc for t in tune.regrid(run, pick).traps]

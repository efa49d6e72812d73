"""Cells-per-trap autotuner of the batched FP32 Kahan form.

The reference takes one grid from its command line (`cudadsmc [Nc N]`,
main.cu:45-69).  The batched form grids each quadrupole trap over +-5 thermal
radii with Nc^3 cells, Nc chosen from the trap's atom count so that the
central cell holds a target number of test particles
(`configs.cells_for_trap`).  This tuner times the step at several targets and
keeps the fastest.  The targets stay within 4..16 test particles in the
central cell: the densest cells then hold several particles to pair, and the
grid is not so fine that empty and single-atom cells dominate the collide.
The move's block size is tuned for each candidate (`dsmc_autotune`), since
the two interact.  Every timing runs the whole step on the device through
libdsmc.so.
"""
from __future__ import annotations

import copy
from typing import Dict, List, Sequence, Tuple

import math

from .configs import Run, cells_for_trap
from .simulation import Simulation

# central-cell occupancies tried, a factor sqrt(2) apart
PEAKS: Tuple[float, ...] = (4.0, 5.66, 8.0, 11.3, 16.0)
PEAK_RANGE = (4.0, 16.0)
# cells per dimension of a trap: the batched workload's 32^3..64^3
NC_RANGE = (32, 64)


def central_peak(n_atoms: int, nc: int) -> float:
    """Test particles in the central cell of a quadrupole trap gridded over
    +-5 thermal radii with nc^3 cells (the estimate cells_for_trap inverts)."""
    return 1000.0 * n_atoms / (4.0 * math.pi * nc ** 3)


def regrid(run: Run, target_peak: float, nc_range: Tuple[int, int] = NC_RANGE) -> Run:
    """`run` with every trap's grid chosen for `target_peak` test particles in
    its central cell, Nc clamped to `nc_range`; everything else (half widths,
    atoms, schedules) kept."""
    lo, hi = PEAK_RANGE
    if not lo <= target_peak <= hi:
        raise ValueError(f"target_peak {target_peak} outside {lo}..{hi} test particles per central cell")
    if not 1 <= nc_range[0] <= nc_range[1]:
        raise ValueError(f"bad cell range {nc_range}")
    out = copy.deepcopy(run)
    for t in out.traps:
        if t.kind != "quadrupole":
            raise ValueError("cells per trap are tuned for quadrupole traps only")
        nc = cells_for_trap(t.n_atoms, target_peak, lo=nc_range[0], hi=nc_range[1])
        t.nc = (nc, nc, nc)
    return out


def grid_record(run: Run, target_peak: float, nc_range: Tuple[int, int] = NC_RANGE) -> Dict:
    """How a regridded run meets its target: total cells, the traps whose Nc
    was not clamped, and the central-cell occupancy (min, median, max) over
    the traps."""
    import numpy as np

    ncs = [int(t.nc[0]) for t in run.traps]
    peaks = np.array([central_peak(t.n_atoms, int(t.nc[0])) for t in run.traps])
    return dict(cells=sum(int(t.nc[0]) * int(t.nc[1]) * int(t.nc[2]) for t in run.traps),
                traps_unclamped=int(sum(nc_range[0] < n < nc_range[1] for n in ncs)),
                central_peak=(float(peaks.min()), float(np.median(peaks)), float(peaks.max())))


def autotune_cells(run: Run, peaks: Sequence[float] = PEAKS, warm: int = 20,
                   steps: int = 100, lib=None,
                   nc_range: Tuple[int, int] = NC_RANGE) -> Tuple[Run, List[Dict]]:
    """Time `steps` graph-replayed steps (after `warm` untimed ones and the
    block-size tuning) of `run` regridded for each target in `peaks` (Nc
    clamped to `nc_range`); return the fastest regridded run and one record
    per candidate {target_peak, cells, traps_unclamped, central_peak, block,
    ms_per_step}: with a narrow range many traps sit at its ends, and the
    record shows how many reached the target."""
    if run.precision != "fp32":
        raise ValueError("the cells autotuner tunes the batched FP32 Kahan form")
    if not peaks:
        raise ValueError("no candidate grids")
    report: List[Dict] = []
    best = None
    for p in peaks:
        cand = regrid(run, p, nc_range)
        with Simulation(cand, lib=lib) as sim:
            sim.init_distribution()
            sim.step(4)
            block = sim.autotune()
            sim.step(warm)
            sim.sync()
            ms = sim.time_steps(steps) / steps
        report.append(dict(target_peak=float(p), **grid_record(cand, p, nc_range), block=int(block),
                           ms_per_step=ms))
        if best is None or ms < best[0]:
            best = (ms, cand)
    return best[1], report

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

from .configs import Run, cells_for_trap
from .simulation import Simulation

# central-cell occupancies tried, a factor sqrt(2) apart
PEAKS: Tuple[float, ...] = (4.0, 5.66, 8.0, 11.3, 16.0)
PEAK_RANGE = (4.0, 16.0)


def regrid(run: Run, target_peak: float) -> Run:
    """`run` with every trap's grid chosen for `target_peak` test particles in
    its central cell; everything else (half widths, atoms, schedules) kept."""
    lo, hi = PEAK_RANGE
    if not lo <= target_peak <= hi:
        raise ValueError(f"target_peak {target_peak} outside {lo}..{hi} test particles per central cell")
    out = copy.deepcopy(run)
    for t in out.traps:
        if t.kind != "quadrupole":
            raise ValueError("cells per trap are tuned for quadrupole traps only")
        nc = cells_for_trap(t.n_atoms, target_peak)
        t.nc = (nc, nc, nc)
    return out


def autotune_cells(run: Run, peaks: Sequence[float] = PEAKS, warm: int = 20,
                   steps: int = 100, lib=None) -> Tuple[Run, List[Dict]]:
    """Time `steps` graph-replayed steps (after `warm` untimed ones and the
    block-size tuning) of `run` regridded for each target in `peaks`; return
    the fastest regridded run and one record per candidate
    {target_peak, cells, block, ms_per_step}."""
    if run.precision != "fp32":
        raise ValueError("the cells autotuner tunes the batched FP32 Kahan form")
    if not peaks:
        raise ValueError("no candidate grids")
    report: List[Dict] = []
    best = None
    for p in peaks:
        cand = regrid(run, p)
        with Simulation(cand, lib=lib) as sim:
            sim.init_distribution()
            sim.step(4)
            block = sim.autotune()
            sim.step(warm)
            sim.sync()
            ms = sim.time_steps(steps) / steps
        cells = sum(int(t.nc[0]) * int(t.nc[1]) * int(t.nc[2]) for t in cand.traps)
        report.append(dict(target_peak=float(p), cells=cells, block=int(block), ms_per_step=ms))
        if best is None or ms < best[0]:
            best = (ms, cand)
    return best[1], report

"""Writes the golden fixtures of tests/golden/.

philox_kat.json        Random123 philox4x32_10 known-answer vectors (published
                       with the Random123 library, Salmon et al. SC'11).
homogeneous_rates.json the reference's own published collision-rate table and
                       analytic value (CUDADSMC/homogeneousGasCollisions.py:27-33),
                       copied as data.
index_fixture.npz      4096 positions in the quadrupole-1e4 grid (35^3 cells over
                       +-5 thermal radii) with hand-placed face, edge, corner,
                       outside and NaN cases, plus the oracle's cell ids, stable
                       sort permutation, cell table and the survivors of a
                       seeded stencil.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))


def philox():
    vec = [
        dict(ctr=[0, 0, 0, 0], key=[0, 0], out=[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        dict(ctr=[0xFFFFFFFF] * 4, key=[0xFFFFFFFF] * 2,
             out=[0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        dict(ctr=[0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], key=[0xA4093822, 0x299F31D0],
             out=[0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    (HERE / "philox_kat.json").write_text(json.dumps({"source": "Random123 philox4x32_10 KAT",
                                                      "vectors": vec}, indent=1))


def homogeneous():
    table = {
        "source": "CUDADSMC/homogeneousGasCollisions.py:27-33",
        "cells_per_dim": [5, 10],
        "n_test": [1e3, 5e3, 1e4, 5e4, 1e5, 5e5, 1e6],
        "computed": {"5": [42.2150, 43.1860, 43.3775, 43.6589, 43.6556, 43.5523, 43.5400],
                     "10": [23.7750, 43.0250, 43.2810, 43.4815, 43.5560, 43.5623, 43.5569]},
        "analytic": 43.5568,
        "setup": "N_real = 1e10 in (2 mm)^3, T = 20 uK, dt = 5e-6 s (collisionComparison.py:22-33)",
    }
    (HERE / "homogeneous_rates.json").write_text(json.dumps(table, indent=1))


def index_fixture():
    from cudadsmc_amd import configs
    from oracle import oracle as orc

    run = configs.quadrupole_1e4()
    p = orc.params_from(run)
    nc = p.nc[0]
    hw = p.half_width[0]
    L = (2.0 * hw) / nc
    rng = np.random.default_rng(2024)
    pos = rng.normal(0, 0.4 * hw, (4096, 3))
    k = 0
    # lower faces, one ulp below, upper grid edge, corners, outside, NaN
    for i in range(0, nc + 1, 5):
        f = -hw + i * L
        pos[k] = [f, 0.0, 0.0]; k += 1
        pos[k] = [np.nextafter(f, -1.0), 0.0, 0.0]; k += 1
        pos[k] = [0.0, f, np.nextafter(f, 1.0)]; k += 1
    pos[k] = [-hw, -hw, -hw]; k += 1
    pos[k] = [np.nextafter(hw, 0), np.nextafter(hw, 0), np.nextafter(hw, 0)]; k += 1
    pos[k] = [hw, hw, hw]; k += 1
    pos[k] = [3 * hw, 0, 0]; k += 1
    pos[k] = [0, -3 * hw, 0]; k += 1
    pos[k] = [np.nan, 0, 0]; k += 1
    pos[k] = [0, 0, np.inf]; k += 1
    ids = np.arange(len(pos), dtype=np.int32)
    cell = orc.index_atoms(p, pos, ids)
    skey, sid = orc.sort_arrays(cell, ids, nc ** 3 + 1)
    se, count, prefix = orc.cell_table(skey, nc ** 3)
    stencil = (rng.random(len(pos)) < 0.3).astype(np.int32)
    survivors = orc.remove_if(sid, stencil)
    vel = rng.normal(0, 0.04, pos.shape)
    np.savez_compressed(HERE / "index_fixture.npz", pos=pos, vel=vel, cell=cell, sorted_cell=skey,
                        sorted_id=sid, start_end=se, count=count, prefix=prefix, stencil=stencil,
                        survivors=survivors)


if __name__ == "__main__":
    philox()
    homogeneous()
    index_fixture()
    print("golden fixtures written to", HERE)

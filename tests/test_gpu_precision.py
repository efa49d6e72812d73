"""The FP32 Kahan form against the FP64 plain form on the precision workload
(N = 1e6 in the B' = 2.5 T/m quadrupole at 20 uK, 100^3 cells, dt = 1e-6 s,
collisions on), per result:

* cell ids and the sort permutation of the step-0 state: bit-exact against
  the oracle's FP64 index of the stored positions (collisions.cu:163-226) and
  numpy's stable argsort;
* over 1e4 steps from one sampled state: the energy drift of each form
  <= 1e-4; N identical (no evaporation); the temperature T(t) = m <v^2> /
  (3 kB) (evaporation.cu:108-128) within 5e-3 of the FP64 run at every step;
  the cumulative collision count within 5 sigma (Poisson) + 0.5 % of the FP64
  run's.  The two forms draw the same Philox numbers, but their acceptance
  tests round differently, so after the first differing trial they are two
  samples of the same process: T and the collision count are compared as
  such.
"""
import math

import numpy as np
import pytest

from cudadsmc_amd import Simulation, _abi, configs

pytestmark = pytest.mark.gpu

N = 1000000


def sampled_state():
    with Simulation(configs.quadrupole_1e6("fp64").with_(obs_capacity=16)) as s64:
        s64.init_distribution()
        ids = s64.copy_out(_abi.FIELD_ATOM_ID)
        pos = np.empty((N, 3))
        vel = np.empty((N, 3))
        pos[ids] = s64.copy_out(_abi.FIELD_POS)
        vel[ids] = s64.copy_out(_abi.FIELD_VEL)
    return pos, vel


def test_cell_ids_and_sort_bit_exact_at_1e6(orc):
    pos, vel = sampled_state()
    run = configs.quadrupole_1e6("fp32").with_(obs_capacity=16)
    p = orc.params_from(run)
    print("cell ids and sort, N = 1e6, 100^3 cells", flush=True)
    with Simulation(run) as sim:
        sim.set_state(pos, vel)
        stored = sim.copy_out(_abi.FIELD_POS)   # float + remainder, as FP64
        keys = sim.copy_out(_abi.FIELD_CELL_ID)
        expect = orc.index_atoms(p, np.ascontiguousarray(stored), np.arange(N, dtype=np.int32))
        np.testing.assert_array_equal(keys, expect)
        sim.sort_arrays()
        order = np.argsort(expect, kind="stable")
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_PERM), order)
        np.testing.assert_array_equal(sim.copy_out(_abi.FIELD_CELL_ID), expect[order])
    assert len(np.unique(expect)) > 100000


def test_fp32_against_fp64_per_result():
    steps = 10000
    pos, vel = sampled_state()
    logs = {}
    for prec in ("fp64", "fp32"):
        print(f"precision workload, {prec}, {steps} steps", flush=True)
        run = configs.quadrupole_1e6(prec).with_(obs_capacity=steps + 16)
        with Simulation(run) as sim:
            sim.set_state(pos, vel)
            sim.step(steps)
            logs[prec] = sim.obs_log()[:, 0]
    out = {}
    for prec, lg in logs.items():
        E = lg["ek"] + lg["ep"]
        out[prec] = dict(drift=float(np.max(np.abs(E - E[0]) / abs(E[0]))),
                         coll=int(lg["collisions"].sum()),
                         T=configs.MRB * lg["v2"].sum(axis=1) / (3 * configs.KB * lg["n_alive"]),
                         N=lg["n_alive"].copy())
    a, b = out["fp32"], out["fp64"]
    dT = float(np.max(np.abs(a["T"] - b["T"]) / b["T"]))
    dc = abs(a["coll"] - b["coll"]) / b["coll"]
    print(f"drift fp32 {a['drift']:.2e} fp64 {b['drift']:.2e}; max |dT/T| {dT:.2e}; "
          f"collisions {a['coll']} vs {b['coll']} ({dc:.2e}); N {a['N'][-1]} vs {b['N'][-1]}",
          flush=True)
    assert a["drift"] <= 1e-4 and b["drift"] <= 1e-4
    np.testing.assert_array_equal(a["N"], b["N"])
    assert dT < 5e-3
    assert b["coll"] > 10000
    assert dc < 5.0 / math.sqrt(b["coll"]) + 0.005

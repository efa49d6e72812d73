"""Independent traps sharded over a group of contexts on one GPU
(configs.split_traps, SimulationGroup, dsmc_step_group): the split on the
CPU; on the GPU, every trap of the group bitwise equal to the same trap in
one context holding all of them."""
import ctypes as C

import numpy as np
import pytest

from cudadsmc_amd import Simulation, SimulationGroup, _abi, configs


def test_split_traps_balances_atoms_and_keeps_streams():
    run = configs.batched_traps(n_traps=64, steps=10)
    for g in (1, 2, 3, 7):
        parts = configs.split_traps(run, g)
        assert len(parts) == g
        traps = [t for p in parts for t in p.traps]
        assert [t.n_atoms for t in traps] == [t.n_atoms for t in run.traps]
        # trap i of the whole run keeps Philox stream i (index in its part + offset)
        streams = [i + t.philox_stream for p in parts for i, t in enumerate(p.traps)]
        assert streams == list(range(len(run.traps)))
        sizes = [p.n_atoms for p in parts]
        biggest = max(t.n_atoms for t in run.traps)
        assert max(sizes) - min(sizes) <= 2 * biggest
    with pytest.raises(ValueError):
        configs.split_traps(run, 0)
    with pytest.raises(ValueError):
        configs.split_traps(run, 65)


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [2, 3])
def test_group_equals_one_context(groups):
    steps = 300
    run = configs.batched_traps(n_traps=24, steps=steps).with_(obs_capacity=steps + 64)
    print(f"group of {groups} contexts vs one: 24 traps, {run.n_atoms} atoms, {steps} steps", flush=True)
    with Simulation(run) as one:
        one.init_distribution()
        one.step(steps)
        lg1 = one.obs_log()
        ts1 = one.copy_out(_abi.FIELD_TRAP_START)
        ref = [one.copy_out(f) for f in (_abi.FIELD_POS, _abi.FIELD_VEL, _abi.FIELD_ATOM_ID)]
    parts = configs.split_traps(run, groups)
    with SimulationGroup(parts) as grp:
        grp.init_distribution()
        grp.step(steps, chunk=50)
        grp.sync()
        lg2 = grp.obs_log()
        got = []
        for s in grp.sims:
            got.append([s.copy_out(f) for f in (_abi.FIELD_POS, _abi.FIELD_VEL, _abi.FIELD_ATOM_ID)])
    assert lg2.shape == lg1.shape
    for f in ("collisions", "n_evap", "n_alive"):
        np.testing.assert_array_equal(lg2[f], lg1[f])
    np.testing.assert_allclose(lg2["ek"], lg1["ek"], rtol=1e-12)
    np.testing.assert_allclose(lg2["ep"], lg1["ep"], rtol=1e-12)
    # the state, trap by trap (labels are numbered per context)
    t0 = 0
    label0 = 0
    for p, (pos, vel, ids) in zip(parts, got):
        nt = len(p.traps)
        sl = slice(ts1[t0], ts1[t0 + nt])
        np.testing.assert_array_equal(pos, ref[0][sl])
        np.testing.assert_array_equal(vel, ref[1][sl])
        np.testing.assert_array_equal(ids + label0, ref[2][sl])
        t0 += nt
        label0 += p.n_atoms
    assert int(lg1["n_evap"][1:].sum()) > 0   # the cuts removed atoms on the way


@pytest.mark.gpu
def test_group_rejects_bad_arguments():
    run = configs.batched_traps(n_traps=4, steps=10).with_(obs_capacity=64)
    a, b = configs.split_traps(run, 2)
    print("group argument checks", flush=True)
    with SimulationGroup([a, b]) as grp:
        grp.init_distribution()
        lib = grp.lib
        h = grp.sims[0].h
        twice = (C.c_void_p * 2)(h, h)
        assert lib.dsmc_step_group(twice, 2, 4, 2) == _abi.DSMC_EINVAL
        assert "twice" in lib.dsmc_last_error(h).decode()
        one = (C.c_void_p * 1)(h)
        assert lib.dsmc_step_group(one, 1, 4, 0) == _abi.DSMC_EINVAL
        assert lib.dsmc_step_group(one, 0, 4, 2) == _abi.DSMC_EINVAL
        assert lib.dsmc_step_group(None, 1, 4, 2) == _abi.DSMC_EINVAL
        grp.step(4, chunk=3)   # a chunk that does not divide the steps
        grp.sync()
        assert all(s.info().step == 4 for s in grp.sims)

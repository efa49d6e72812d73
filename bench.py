"""Benchmark of the DSMC step on one MI355X; prints ONE JSON line.

Headline (`value`): atom-steps/s of the FP32 Kahan form on the batched
workload -- 256 independent quadrupole traps, N log-uniform in [1e3, 1e5],
B' ~ U[1, 5] T/m, T ~ U[5, 50] uK, per-trap 32^3..64^3 cell grids and energy
cuts, collisions on, dt = 1e-6 s -- replayed through a HIP graph.
atom-steps/s = (sum over timed steps and traps of atoms moved) / (hipEvent
time of exactly --steps steps on the context's stream), state resident in HBM.

Also reported: the FP64 plain (drop-in) form and the FP32 form on the single
quadrupole trap N = 1e6 / 100^3 cells (the precision-form workload), their
ratio, the FP64 energy drift of the FP32 runs, a per-kernel roofline from
hipEvents around every launch of a profiled pass of the same workload, the
HBM traffic of each kernel of the step from two rocprofv3 counter runs
(FETCH_SIZE, then WRITE_SIZE) scaled by what the same counters read for a
record gather of known size in the move's access pattern, the per-result
error of the FP32 form against the FP64 form at N = 1e6, the spin and
evaporation workload at N = 4e6, and the CPU oracle (OpenMP FP64 restatement
of the reference kernels) on N = 1e4 / 35^3.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import shutil
import sqlite3
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per atom and launch of each kernel, from the reference
# path's own compulsory traffic (FP32 Kahan form):
#   move: the state read and written once, x, v and their Kahan remainders,
#         12 floats = 48 B each way (96 B); with spin the Bloch vector adds
#         12 B each way (120 B)
#   collide: the velocity record of each atom read and written once (24 B)
#   sort (one logical pass over (key, slot) pairs): 16 B
#   cell table: the sorted key read once (4 B)
# The headline (whole-step) roofline prices one atom-step at the move's bytes.
ALG_BYTES = {
    "move_fused": 96.0,
    "step_fused": 96.0,
    "collide_fused": 24.0,
    "resort_merge": 16.0,
    "radix_downsweep": 16.0,
    "table_fused": 4.0,
}
SPIN_BYTES = 24.0
# Streams the fused move (and the fused step) carry beyond that state: the
# sort permutation (4), the last sort's key of the slot (4) and the next cell
# key (4).  The atom label travels inside the state records (in the low bits
# of the position remainders), so it adds no bytes.
FUSED_EXTRA_BYTES = 12.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def atom_steps(sim, steps):
    """Atoms moved in the last `steps` steps, over all traps (the survivors of
    each step plus the atoms it removed)."""
    lg = sim.obs_log()[-steps:]
    return int(lg["n_alive"].sum() + lg["n_evap"].sum())


def run_gpu_leg(run, steps, warmup, autotune=True, repeats=3):
    """Init, autotune, warm up, then time exactly `steps` steps `repeats`
    times (hipEvents on the context's stream, state resident in HBM); the
    median repetition is reported."""
    from cudadsmc_amd import Simulation

    sim = Simulation(run)
    sim.init_distribution()
    bs = 256
    if autotune and run.precision == "fp32":
        sim.step(4)   # the move is tuned on sorted state, as the step loop runs it
        bs = sim.autotune()
    sim.step(warmup)
    sim.sync()
    reps = []
    for _ in range(repeats):
        ms = sim.time_steps(steps)
        n = atom_steps(sim, steps)
        reps.append(dict(ms=ms, atom_steps=n, value=n / (ms * 1e-3)))
    med = sorted(reps, key=lambda r: r["value"])[len(reps) // 2]
    return sim, dict(med, block=bs, repeats=[round(r["value"], 1) for r in reps])


def run_group_leg(runs, steps, warmup, chunk=50, repeats=3):
    """run_gpu_leg for a group of contexts (one per group of traps, each on
    its own stream), stepped in turns of `chunk` steps (dsmc_step_group) and
    timed from one start event to the last stream's end."""
    from cudadsmc_amd import SimulationGroup

    grp = SimulationGroup(runs)
    grp.init_distribution()
    grp.step(4)
    blocks = grp.autotune()
    grp.step(warmup)
    grp.sync()
    reps = []
    for _ in range(repeats):
        ms = grp.time_steps(steps, chunk)
        n = sum(atom_steps(s, steps) for s in grp.sims)
        reps.append(dict(ms=ms, atom_steps=n, value=n / (ms * 1e-3)))
    med = sorted(reps, key=lambda r: r["value"])[len(reps) // 2]
    return grp, dict(med, blocks=blocks, repeats=[round(r["value"], 1) for r in reps])


def energy_drift(sim, steps):
    """max over traps of max_t |E(t) - E(0)| / E(0), E = alive energy plus the
    energy carried away by the atoms removed so far."""
    lg = sim.obs_log()[-steps:]
    E = lg["ek"] + lg["ep"] + np.cumsum(lg["e_evap"], axis=0)
    return float(np.max(np.abs(E - E[0]) / np.abs(E[0])))


def profile_leg(sim, steps, spin=False):
    """Per-kernel hipEvent timing over `steps` eager steps."""
    sim.profile(True)
    sim.step(steps)
    stats = sim.profile_read()
    sim.profile(False)
    lg = sim.obs_log()[-steps:]
    n_mean = float((lg["n_alive"] + lg["n_evap"]).sum(axis=1).mean())
    out = {}
    for name, st in stats.items():
        avg_ms = st["total_ms"] / max(st["launches"], 1)
        rec = dict(launches=st["launches"], total_ms=round(st["total_ms"], 4),
                   avg_us=round(avg_ms * 1e3, 3))
        if name in ALG_BYTES:
            per = ALG_BYTES[name] + (SPIN_BYTES if spin and name in ("move_fused", "step_fused") else 0.0)
            b = per * n_mean
            if name == "radix_downsweep":
                passes = sim.info().radix_passes
                b = n_mean * (16.0 * passes - 4.0) / passes   # first pass reads no values
            rec["alg_bytes_per_atom"] = per
            rec["alg_bytes"] = b
            rec["achieved_GBs"] = round(b / (avg_ms * 1e-3) / 1e9, 1)
            if name in ("move_fused", "step_fused"):
                rec["extra_bytes_per_atom"] = FUSED_EXTRA_BYTES
                rec["with_extra_GBs"] = round((per + FUSED_EXTRA_BYTES) * n_mean / (avg_ms * 1e-3) / 1e9, 1)
        out[name] = rec
    return out, n_mean


def copy_calibration(mbytes=1024, repeats=20):
    """A streaming copy (read + write) of `mbytes` MB: a kernel of known bytes
    for scaling the FETCH_SIZE / WRITE_SIZE counters (its rate is reported,
    not used as a ceiling: the roofline's peak is the HBM's 8 TB/s)."""
    import ctypes as C
    from cudadsmc_amd import load_library

    lib = load_library()
    nbytes = mbytes << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert lib.dsmc_device_alloc(C.byref(a), nbytes) == 0
    assert lib.dsmc_device_alloc(C.byref(b), nbytes) == 0
    ms = C.c_double()
    st = lib.dsmc_stream_copy(b, a, nbytes, repeats, C.byref(ms))
    lib.dsmc_device_free(a)
    lib.dsmc_device_free(b)
    assert st == 0
    return dict(bytes=nbytes, ms=ms.value, GBs=2.0 * nbytes / (ms.value * 1e-3) / 1e9)


GATHER_SLOTS = 1 << 24


def gather_map(n, displaced=0.0, seed=7):
    """Slot map of the record-gather calibration.  The default is the identity,
    so that the bytes it moves are known exactly (52 read and 48 written per
    slot, in the fused move's record layout and access pattern); `displaced`
    > 0 takes that fraction of the slots from up to 4096 slots away, as the
    fused move does for the atoms that changed cell."""
    rng = np.random.default_rng(seed)
    m = np.arange(n, dtype=np.int64)
    k = rng.random(n) < displaced
    m[k] = np.clip(m[k] + rng.integers(-4096, 4097, int(k.sum())), 0, n - 1)
    return m.astype(np.uint32)


def gather_calibration(n=GATHER_SLOTS, repeats=3):
    """The record gather (dsmc_gather_copy) of n slots through gather_map(n):
    52 B read and 48 B written per slot."""
    import ctypes as C
    from cudadsmc_amd import load_library

    lib = load_library()
    rec_bytes = 48 * n
    a, b, m = C.c_void_p(), C.c_void_p(), C.c_void_p()
    assert lib.dsmc_device_alloc(C.byref(a), rec_bytes) == 0
    assert lib.dsmc_device_alloc(C.byref(b), rec_bytes) == 0
    assert lib.dsmc_device_alloc(C.byref(m), 4 * n) == 0
    mp = gather_map(n)
    assert lib.dsmc_memcpy_htod(m, mp.ctypes.data_as(C.c_void_p), 4 * n) == 0
    ms = C.c_double()
    st = lib.dsmc_gather_copy(b, a, m, n, repeats, C.byref(ms))
    for p in (a, b, m):
        lib.dsmc_device_free(p)
    assert st == 0
    return dict(read_bytes=52 * n, write_bytes=48 * n, ms=ms.value,
                GBs=100.0 * n / (ms.value * 1e-3) / 1e9)


def pmc_child(steps):
    """Body of the counter runs: a streaming copy and a record gather of
    known sizes, then `steps` eager steps of the batched workload."""
    from cudadsmc_amd import Simulation, configs

    cal = copy_calibration(mbytes=1024, repeats=3)
    gat = gather_calibration()
    run = configs.batched_traps().with_(use_graph=False, obs_capacity=steps + 64)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(steps)
        sim.sync()
        lg = sim.obs_log()
    # atoms each step's move read: alive after it plus removed by it
    moved = (lg["n_alive"].sum(axis=1) + lg["n_evap"].sum(axis=1)).astype(int).tolist()
    print(json.dumps({"copy_bytes": cal["bytes"], "gather": gat, "atoms_moved": moved}), flush=True)


def _pmc_values(db_glob):
    """Median counter value per dispatch, by kernel name, from a rocpd database
    (the median drops the first step after initialisation, whose gather runs
    in random order)."""
    out = {}
    for db in glob.glob(db_glob, recursive=True):
        con = sqlite3.connect(db)
        rows = con.execute("select kernel_name, dispatch_id, sum(value) from counters_collection "
                           "group by kernel_name, dispatch_id").fetchall()
        con.close()
        for name, _, v in rows:
            out.setdefault(name, []).append(float(v))
    return {k: float(np.median(v)) for k, v in out.items()}


def pmc_traffic(steps=10, timeout_s=150):
    """HBM bytes per launch of move_fused: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) of two rocprofv3 runs, each scaled by the ratio of the known
    bytes of the streaming copy to what the same counter reads for it.
    Runs before this process touches the GPU; None when rocprofv3 is absent
    or a run fails."""
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3"
                                          if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if prof is None:
        return None
    tmp = tempfile.mkdtemp(prefix="dsmc_pmc_")
    vals = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter)
            cmd = ["timeout", "-s", "KILL", str(timeout_s), prof, "--pmc", counter,
                   "-d", d, "-o", "pmc", "--",
                   sys.executable, str(ROOT / "bench.py"), "--pmc-child", str(steps)]
            log(f"[bench] counter run {counter}")
            res = subprocess.run(cmd, capture_output=True, text=True)
            v = _pmc_values(os.path.join(d, "**", "*.db"))
            if not v:
                log(f"[bench] counter run {counter} gave no data (rc {res.returncode})")
                return None
            vals[counter] = v
    finally:
        shutil.rmtree(tmp, ignore_errors=True)

    def pick(v, key):
        names = [k for k in v if key in k]
        return v[names[0]] if names else None

    copy_bytes = 1024 << 20
    F, W = vals["FETCH_SIZE"], vals["WRITE_SIZE"]
    f_copy, w_copy = pick(F, "stream_copy"), pick(W, "stream_copy")
    f_gat, w_gat = pick(F, "gather_copy"), pick(W, "gather_copy")
    if None in (f_copy, w_copy, f_gat, w_gat) or min(f_copy, w_copy, f_gat, w_gat) <= 0:
        return None
    # counter scales: known bytes / counter bytes, for the streaming copy and
    # for the record gather (the move's access pattern); the latter is used
    scales = dict(copy_fetch=copy_bytes / (f_copy * 1024.0), copy_write=copy_bytes / (w_copy * 1024.0),
                  gather_fetch=52.0 * GATHER_SLOTS / (f_gat * 1024.0),
                  gather_write=48.0 * GATHER_SLOTS / (w_gat * 1024.0))
    out = dict(scales={k: round(v, 4) for k, v in scales.items()}, steps=steps, kernels={})
    # profiler names -> kernel names
    names = {"move_fused": "move_fused", "collide_fused": "collide_fused",
             "resort_merge": "ins_merge", "resort_locate": "ins_locate",
             "obs_finalize": "obs_finalize", "collide_edges": "collide_edges"}
    for key, kname in names.items():
        f, w = pick(F, kname), pick(W, kname)
        if f is None or w is None:
            continue
        fetch = f * 1024.0 * scales["gather_fetch"]
        write = w * 1024.0 * scales["gather_write"]
        out["kernels"][key] = dict(fetch_bytes=round(fetch), write_bytes=round(write),
                                   bytes=round(fetch + write), fetch_kib_raw=f, write_kib_raw=w)
    return out


def cpu_baseline(target_s=12.0, repeats=3):
    """The OpenMP oracle (FP64 restatement of the reference kernels) on the
    CPU configuration, on this host's cores: `repeats` runs of about
    target_s / repeats seconds each, the median reported."""
    from cudadsmc_amd import configs
    from oracle import oracle as orc

    run = configs.quadrupole_1e4()
    p = orc.params_from(run)
    t = run.traps[0]
    st = orc.init_distribution(p, t.n_atoms, t.T)
    sv = np.full(orc.ncells(p) + 1, orc.sigvrmax_init(p, t.T))
    probe = orc.run(p, st, 50, 0, sv)
    per_step = probe["seconds"] / 50
    n = int(max(50, min(200000, target_s / repeats / max(per_step, 1e-6))))
    res = []
    step0 = 50
    for _ in range(repeats):
        r = orc.run(p, st, n, step0, sv)
        step0 += n
        res.append(r)
    vals = sorted(r["atom_steps"] / r["seconds"] for r in res)
    secs = sum(r["seconds"] for r in res)
    return dict(value=vals[len(vals) // 2], unit="atom-steps/s", cores=res[0]["threads"], kind="port",
                runs=[round(v, 1) for v in vals],
                sample=f"N=1e4 quadrupole, 35^3 cells, dt=1e-6 s, collisions on, FP64, "
                       f"{repeats} runs of {n} steps ({secs:.1f} s in all) of the OpenMP oracle; median")


def drift_leg(run, steps=10000):
    """max over traps of |dE/E| (alive + removed energy) over `steps` graph
    steps from the initial state (the metric's energy condition)."""
    from cudadsmc_amd import Simulation

    with Simulation(run.with_(obs_capacity=steps + 16)) as sim:
        sim.init_distribution()
        sim.step(steps)
        lg = sim.obs_log()
    E = lg["ek"] + lg["ep"] + np.cumsum(lg["e_evap"], axis=0)
    per_trap = np.max(np.abs(E - E[0]) / np.abs(E[0]), axis=0)
    return dict(steps=steps, max_abs_dE_over_E=float(per_trap.max()),
                median_trap=float(np.median(per_trap)), removed=int(lg["n_evap"].sum()))


def precision_leg(steps=10000):
    """The precision-form workload (N = 1e6, 100^3 cells, collisions on) run
    for `steps` steps in the FP32 Kahan form and in the FP64 plain form from
    one sampled state (the FP64 form's initial distribution, given to both).
    Per result: the energy drift of each, and the FP32 form's relative
    difference from the FP64 run in T(t) (max over steps), the cumulative
    collision count and N.  The two runs draw the same Philox numbers, but
    the acceptance tests round differently, so they are two samples of the
    same process: the T and collision differences are statistical."""
    from cudadsmc_amd import Simulation, _abi, configs

    out = {}
    logs = {}
    with Simulation(configs.quadrupole_1e6("fp64").with_(obs_capacity=steps + 16)) as s64:
        s64.init_distribution()
        ids = s64.copy_out(_abi.FIELD_ATOM_ID)
        pos = np.empty((1000000, 3))
        vel = np.empty((1000000, 3))
        pos[ids] = s64.copy_out(_abi.FIELD_POS)
        vel[ids] = s64.copy_out(_abi.FIELD_VEL)
    for prec in ("fp64", "fp32"):
        run = configs.quadrupole_1e6(prec).with_(obs_capacity=steps + 16)
        with Simulation(run) as sim:
            sim.set_state(pos, vel)
            sim.step(steps)
            lg = sim.obs_log()[:, 0]
        logs[prec] = lg
        E = lg["ek"] + lg["ep"]
        out[prec] = dict(energy_drift=float(np.max(np.abs(E - E[0]) / abs(E[0]))),
                         collisions=int(lg["collisions"].sum()), n_final=int(lg["n_alive"][-1]))
    T = {k: configs.MRB * v["v2"].sum(axis=1) / (3 * configs.KB * v["n_alive"]) for k, v in logs.items()}
    c64, c32 = out["fp64"]["collisions"], out["fp32"]["collisions"]
    out["fp32_vs_fp64"] = dict(
        steps=steps,
        T_rel_max=float(np.max(np.abs(T["fp32"] - T["fp64"]) / T["fp64"])),
        T_rel_final=float(abs(T["fp32"][-1] - T["fp64"][-1]) / T["fp64"][-1]),
        collisions_rel=float(abs(c32 - c64) / c64),
        collisions_poisson_sigma=float(1.0 / np.sqrt(c64)),
        N_rel=float(abs(out["fp32"]["n_final"] - out["fp64"]["n_final"]) / out["fp64"]["n_final"]),
        energy_drift_fp32=out["fp32"]["energy_drift"], energy_drift_fp64=out["fp64"]["energy_drift"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--single-steps", type=int, default=2000)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="batched,drift,fp64,fp32,precision,spin",
                    help="comma list of: batched (headline), drift (1e4-step energy check of "
                         "the batched workload), fp64 / fp32 (N = 1e6 single trap), precision "
                         "(FP32 vs FP64 per result at N = 1e6, 1e4 steps), spin (N = 4e6 spin "
                         "and evaporation)")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the streaming-copy and record-gather calibration kernels")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter runs (roofline.traffic = null)")
    ap.add_argument("--streams", type=int, default=2,
                    help="trap groups of the batched workload, one context and stream each, "
                         "stepped together (1: one context)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step's kernels one by one instead of replaying a HIP graph "
                         "(every leg)")
    ap.add_argument("--tune-cells", action="store_true",
                    help="also time the batched step at each cells-per-trap target of the "
                         "cells autotuner and report its choice (the measured workload keeps "
                         "8 test particles per central cell)")
    ap.add_argument("--pmc-child", type=int, default=0, help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the DSMC step runs on a single MI355X and "
              "runs over several GPUs are not supported", file=sys.stderr)
        sys.exit(2)

    if args.pmc_child:
        pmc_child(args.pmc_child)
        return
    legs = set(args.legs.split(","))
    # the counter runs are child processes started before this one uses the GPU
    traffic = None
    if not args.no_pmc and not args.copy_only and "batched" in legs:
        traffic = pmc_traffic()

    from cudadsmc_amd import configs, load_library

    load_library()
    if args.copy_only:
        print(json.dumps({"stream_copy": copy_calibration(), "gather_copy": gather_calibration()}),
              flush=True)
        return
    result = {}

    batched = configs.batched_traps(steps=args.steps)
    batched.obs_capacity = 3 * args.steps + args.warmup + args.profile_steps + 64
    if args.no_graph:
        batched.use_graph = False
    roof = None
    kernels = None
    step_roof = None
    if "batched" in legs:
        log(f"[bench] batched: {len(batched.traps)} traps, {batched.n_atoms} atoms")
        sim, b = run_gpu_leg(batched, args.steps, args.warmup)
        b["energy_drift"] = energy_drift(sim, args.steps)
        log(f"[bench] batched FP32, one context: {b['value']:.4e} atom-steps/s (median of "
            f"{b['repeats']}; {b['ms']:.1f} ms, block {b['block']}), max |dE/E| {b['energy_drift']:.2e}")
        # the per-kernel pass runs on the one-context run: each kernel alone
        kernels, n_mean = profile_leg(sim, args.profile_steps)
        sim.close()
        if args.streams > 1 and not args.no_graph:
            one = b
            grp, b = run_group_leg(configs.split_traps(batched, args.streams), args.steps,
                                   args.warmup)
            lg = grp.obs_log()[-args.steps:]
            E = lg["ek"] + lg["ep"] + np.cumsum(lg["e_evap"], axis=0)
            b["energy_drift"] = float(np.max(np.abs(E - E[0]) / np.abs(E[0])))
            grp.close()
            b["one_context"] = dict(value=one["value"], ms=one["ms"], repeats=one["repeats"],
                                    block=one["block"])
            log(f"[bench] batched FP32, {args.streams} trap groups on {args.streams} streams: "
                f"{b['value']:.4e} atom-steps/s (median of {b['repeats']}; {b['ms']:.1f} ms), "
                f"max |dE/E| {b['energy_drift']:.2e}")
        result["batched"] = b
        dom = max((k for k in kernels if "alg_bytes" in kernels[k]),
                  key=lambda k: kernels[k]["total_ms"])
        d = kernels[dom]
        achieved = d["alg_bytes"] / (d["avg_us"] * 1e-6) / 1e9
        tr = (traffic or {}).get("kernels", {}).get(dom)
        cal = copy_calibration()
        roof = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=tr["bytes"] if tr else None,
                    alg_bytes_per_atom=d["alg_bytes_per_atom"],
                    alg_bytes_per_launch=round(d["alg_bytes"]), atoms_per_launch=round(n_mean),
                    avg_launch_us=d["avg_us"],
                    extra_bytes_per_atom=d.get("extra_bytes_per_atom"),
                    with_extra_GBs=d.get("with_extra_GBs"),
                    stream_copy_GBs=round(cal["GBs"], 1))
        if traffic:
            roof["traffic_detail"] = traffic
        step_bw = b["value"] * ALG_BYTES["move_fused"] / 1e9
        step_roof = dict(bytes_per_atom_step=ALG_BYTES["move_fused"], achieved=round(step_bw, 1),
                         peak=HBM_PEAK_GBS, unit="GB/s", frac=round(step_bw / HBM_PEAK_GBS, 4))
        log(f"[bench] dominant kernel {dom}: {achieved:.0f} GB/s; whole step {step_bw:.0f} GB/s")
    if "drift" in legs:
        result["drift"] = drift_leg(configs.batched_traps())
        log(f"[bench] batched 1e4 steps: max |dE/E| {result['drift']['max_abs_dE_over_E']:.2e}")
    if "fp64" in legs:
        run64 = configs.quadrupole_1e6("fp64").with_(obs_capacity=3 * args.single_steps + args.warmup + 64,
                                                     use_graph=not args.no_graph)
        sim, r = run_gpu_leg(run64, args.single_steps, args.warmup)
        sim.close()
        log(f"[bench] FP64 plain N=1e6: {r['value']:.4e} atom-steps/s")
        result["fp64_plain"] = r
    if "fp32" in legs:
        run32 = configs.quadrupole_1e6("fp32").with_(obs_capacity=3 * args.single_steps + args.warmup + 64,
                                                     use_graph=not args.no_graph)
        sim, r = run_gpu_leg(run32, args.single_steps, args.warmup)
        r["energy_drift"] = energy_drift(sim, args.single_steps)
        sim.close()
        log(f"[bench] FP32 Kahan N=1e6: {r['value']:.4e} atom-steps/s, |dE/E| {r['energy_drift']:.2e}")
        result["fp32_single"] = r
    if "precision" in legs:
        result["precision"] = precision_leg()
        log(f"[bench] FP32 vs FP64 at N=1e6: {json.dumps(result['precision']['fp32_vs_fp64'])}")
    if "spin" in legs:
        spin_steps = 200
        run_s = configs.spin_evaporation("fp32").with_(
            obs_capacity=3 * spin_steps + args.warmup + 64 + 50, use_graph=not args.no_graph)
        # the energy cut ramps over 2000 steps here instead of 2e5, so that
        # atoms leave during the timed steps
        run_s.traps[0].eta_steps = 2000
        sim, r = run_gpu_leg(run_s, spin_steps, args.warmup)
        r["energy_drift"] = energy_drift(sim, spin_steps)
        kern, n_mean = profile_leg(sim, 50, spin=True)
        sim.close()
        bw = r["value"] * (ALG_BYTES["move_fused"] + SPIN_BYTES) / 1e9
        r["step_roofline"] = dict(bytes_per_atom_step=ALG_BYTES["move_fused"] + SPIN_BYTES,
                                  achieved=round(bw, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                                  frac=round(bw / HBM_PEAK_GBS, 4))
        r["kernels"] = kern
        log(f"[bench] spin + evaporation N=4e6: {r['value']:.4e} atom-steps/s, whole step {bw:.0f} GB/s")
        result["spin"] = r

    cpu = None if args.no_cpu else cpu_baseline()
    if cpu:
        log(f"[bench] CPU oracle: {cpu['value']:.4e} atom-steps/s on {cpu['cores']} threads")

    cells_tuning = None
    if args.tune_cells:
        from cudadsmc_amd import configs as _configs, tune
        _, report = tune.autotune_cells(_configs.batched_traps())
        best = min(report, key=lambda r: r["ms_per_step"])
        for r in report:
            log(f"[bench] cells target {r['target_peak']:5.2f}/central cell: {r['cells']} cells, "
                f"{r['traps_unclamped']} traps at the target, block {r['block']}, "
                f"{r['ms_per_step'] * 1e3:.1f} us per step")
        cells_tuning = dict(best_target_peak=best["target_peak"],
                            candidates=[{k: (round(v, 5) if isinstance(v, float) else v)
                                         for k, v in r.items()} for r in report])

    head = result.get("batched") or result.get("fp32_single") or result.get("fp64_plain")
    if head is None:
        print(json.dumps({k: v for k, v in result.items()}), flush=True)
        return
    steps = args.steps if "batched" in result else args.single_steps
    line = {
        "metric": "atom-steps/s (collisions on), FP32-batched vs FP64-plain at |dE/E|<=1e-4",
        "value": round(head["value"], 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms"] / steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (Philox-sampled thermal clouds)",
        "config": {
            "workload": "batched_traps" if "batched" in result else "quadrupole_1e6",
            "n_traps": len(batched.traps) if "batched" in result else 1,
            "n_atoms": batched.n_atoms if "batched" in result else 1000000,
            "dt": 1e-6,
            "precision": "fp32-kahan, fp64 reductions",
            "parallelism": "single-gpu",
            "trap_groups": (args.streams if "batched" in result and "one_context" in result["batched"]
                            else 1),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "batched" in result:
        line["repeats"] = result["batched"]["repeats"]
        line["energy_drift"] = result["batched"]["energy_drift"]
        if "one_context" in result["batched"]:
            line["one_context"] = result["batched"]["one_context"]
        line["step_roofline"] = step_roof
    if "drift" in result:
        line["energy_drift_1e4"] = result["drift"]
    if "fp64_plain" in result:
        line["fp64_plain"] = {k: round(v, 3) if isinstance(v, float) else v
                              for k, v in result["fp64_plain"].items()}
        line["ratio_vs_fp64_plain"] = round(head["value"] / result["fp64_plain"]["value"], 3)
    if "fp32_single" in result:
        line["fp32_single"] = {k: (round(v, 3) if isinstance(v, float) and k != "energy_drift" else v)
                               for k, v in result["fp32_single"].items()}
    if "precision" in result:
        line["precision"] = result["precision"]
    if "spin" in result:
        line["spin"] = {k: (round(v, 3) if isinstance(v, float) and k != "energy_drift" else v)
                        for k, v in result["spin"].items()}
    if kernels:
        line["kernels"] = kernels
    if cells_tuning:
        line["cells_tuning"] = cells_tuning
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

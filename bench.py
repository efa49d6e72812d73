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
HBM traffic of the dominant kernel from two rocprofv3 counter runs (FETCH_SIZE,
then WRITE_SIZE) scaled by what the same counters read for a streaming copy of
known size, and the CPU oracle (OpenMP FP64 restatement of the reference
kernels) on N = 1e4 / 35^3.
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
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per atom (or per pass) of each kernel of the FP32 form.
#   move_fused: read perm (4) + 12 floats of state (48) + label (4) + the
#               last sort's key of the slot (4); write 12 floats (48) + label
#               (4) + next cell key (4)
#   resort_split: read new and previous key (8), write the unchanged atoms'
#               (key, slot) (8); resort_merge: read those (8), write the sorted
#               (key, slot) (8) -- movers (a few percent) not counted
#   radix_downsweep: read key + value, write key + value (first pass reads no
#               value: identity); radix_upsweep: read key
#   table_fused: read key; collide_fused: read key (per slot) -- per-cell work
#               is counted by the measured traffic, not here.
PER_ATOM_BYTES = {
    "move_fused": 116.0,
    "resort_split": 16.0,
    "resort_merge": 16.0,
    "radix_downsweep": 16.0,
    "radix_upsweep": 4.0,
    "table_fused": 4.0,
    "collide_fused": 4.0,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def atom_steps(sim, steps):
    lg = sim.obs_log()[-steps:]
    return int(lg["n_alive"].sum())


def run_gpu_leg(run, steps, warmup, autotune=True):
    from cudadsmc_amd import Simulation

    sim = Simulation(run)
    sim.init_distribution()
    bs = 256
    if autotune and run.precision == "fp32":
        sim.step(4)   # the move is tuned on sorted state, as the step loop runs it
        bs = sim.autotune()
    sim.step(warmup)
    sim.sync()
    ms = sim.time_steps(steps)
    n = atom_steps(sim, steps)
    return sim, dict(ms=ms, atom_steps=n, value=n / (ms * 1e-3), block=bs)


def energy_drift(sim, steps):
    """max over traps of max_t |E(t) - E(0)| / E(0), E = alive energy plus the
    energy carried away by the atoms removed so far."""
    lg = sim.obs_log()[-steps:]
    E = lg["ek"] + lg["ep"] + np.cumsum(lg["e_evap"], axis=0)
    return float(np.max(np.abs(E - E[0]) / np.abs(E[0])))


def profile_leg(sim, steps):
    """Per-kernel hipEvent timing over `steps` eager steps."""
    sim.profile(True)
    sim.step(steps)
    stats = sim.profile_read()
    sim.profile(False)
    lg = sim.obs_log()[-steps:]
    n_mean = float(lg["n_alive"].sum(axis=1).mean())
    out = {}
    for name, st in stats.items():
        avg_ms = st["total_ms"] / max(st["launches"], 1)
        rec = dict(launches=st["launches"], total_ms=round(st["total_ms"], 4),
                   avg_us=round(avg_ms * 1e3, 3))
        if name in PER_ATOM_BYTES:
            b = PER_ATOM_BYTES[name] * n_mean
            if name == "radix_downsweep":
                # first pass of each sort reads no values
                passes = sim.info().radix_passes
                b = n_mean * (16.0 * passes - 4.0) / passes
            rec["alg_bytes"] = b
            rec["achieved_GBs"] = round(b / (avg_ms * 1e-3) / 1e9, 1)
        out[name] = rec
    return out, n_mean


def copy_calibration(mbytes=1024, repeats=20):
    """Achieved GB/s of a streaming copy (read + write) of `mbytes` MB: the
    practical bandwidth ceiling of this device, and the reference kernel for
    scaling the FETCH_SIZE / WRITE_SIZE counters."""
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


def pmc_child(steps):
    """Body of the counter runs: a streaming copy of known size, then `steps`
    eager steps of the batched workload."""
    from cudadsmc_amd import Simulation, configs

    cal = copy_calibration(mbytes=1024, repeats=3)
    run = configs.batched_traps().with_(use_graph=False, obs_capacity=steps + 64)
    with Simulation(run) as sim:
        sim.init_distribution()
        sim.step(steps)
        sim.sync()
    print(json.dumps({"copy_bytes": cal["bytes"]}), flush=True)


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
    f_copy, w_copy = pick(vals["FETCH_SIZE"], "stream_copy"), pick(vals["WRITE_SIZE"], "stream_copy")
    f_move, w_move = pick(vals["FETCH_SIZE"], "move_fused"), pick(vals["WRITE_SIZE"], "move_fused")
    if None in (f_copy, w_copy, f_move, w_move) or f_copy <= 0 or w_copy <= 0:
        return None
    f_scale = copy_bytes / (f_copy * 1024.0)
    w_scale = copy_bytes / (w_copy * 1024.0)
    fetch = f_move * 1024.0 * f_scale
    write = w_move * 1024.0 * w_scale
    return dict(kernel="move_fused", fetch_bytes=fetch, write_bytes=write, bytes=fetch + write,
                fetch_scale=round(f_scale, 4), write_scale=round(w_scale, 4),
                fetch_kib_raw=f_move, write_kib_raw=w_move, steps=steps)


def cpu_baseline(target_s=12.0):
    from cudadsmc_amd import configs
    from oracle import oracle as orc

    run = configs.quadrupole_1e4()
    p = orc.params_from(run)
    t = run.traps[0]
    st = orc.init_distribution(p, t.n_atoms, t.T)
    sv = np.full(orc.ncells(p) + 1, orc.sigvrmax_init(p, t.T))
    probe = orc.run(p, st, 50, 0, sv)
    per_step = probe["seconds"] / 50
    n = int(max(50, min(200000, target_s / max(per_step, 1e-6))))
    res = orc.run(p, st, n, 50, sv)
    return dict(value=res["atom_steps"] / res["seconds"], unit="atom-steps/s",
                cores=res["threads"], kind="port",
                sample=f"N=1e4 quadrupole, 35^3 cells, dt=1e-6 s, collisions on, FP64, "
                       f"{n} steps ({res['seconds']:.1f} s) of the OpenMP oracle")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--single-steps", type=int, default=2000)
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="batched,fp64,fp32")
    ap.add_argument("--copy-only", action="store_true",
                    help="only run the streaming-copy bandwidth calibration")
    ap.add_argument("--no-pmc", action="store_true",
                    help="skip the rocprofv3 counter runs (roofline.traffic = null)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the step's kernels one by one instead of replaying a HIP graph")
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
        print(json.dumps({"copy_calibration": copy_calibration()}), flush=True)
        return
    result = {}

    batched = configs.batched_traps(steps=args.steps)
    batched.obs_capacity = args.steps + args.warmup + args.profile_steps + 64
    if args.no_graph:
        batched.use_graph = False
    roof = None
    kernels = None
    if "batched" in legs:
        log(f"[bench] batched: {len(batched.traps)} traps, {batched.n_atoms} atoms")
        sim, b = run_gpu_leg(batched, args.steps, args.warmup)
        b["energy_drift"] = energy_drift(sim, args.steps)
        log(f"[bench] batched FP32: {b['value']:.4e} atom-steps/s ({b['ms']:.1f} ms, block {b['block']}), "
            f"max |dE/E| {b['energy_drift']:.2e}")
        result["batched"] = b
        kernels, n_mean = profile_leg(sim, args.profile_steps)
        sim.close()
        dom = max((k for k in kernels if "alg_bytes" in kernels[k]),
                  key=lambda k: kernels[k]["total_ms"])
        d = kernels[dom]
        achieved = d["alg_bytes"] / (d["avg_us"] * 1e-6) / 1e9
        tr = traffic if (traffic and traffic["kernel"] == dom) else None
        cal = copy_calibration()
        roof = dict(bound="hbm", kernel=dom, achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=round(tr["bytes"]) if tr else None,
                    alg_bytes_per_launch=round(d["alg_bytes"]), avg_launch_us=d["avg_us"],
                    copy_calibration_GBs=round(cal["GBs"], 1),
                    frac_of_copy=round(achieved / cal["GBs"], 4))
        if tr:
            roof["traffic_detail"] = {k: (round(v, 1) if isinstance(v, float) else v)
                                      for k, v in tr.items()}
        log(f"[bench] dominant kernel {dom}: {achieved:.0f} GB/s")
    if "fp64" in legs:
        run64 = configs.quadrupole_1e6("fp64").with_(obs_capacity=args.single_steps + args.warmup + 64)
        sim, r = run_gpu_leg(run64, args.single_steps, args.warmup)
        sim.close()
        log(f"[bench] FP64 plain N=1e6: {r['value']:.4e} atom-steps/s")
        result["fp64_plain"] = r
    if "fp32" in legs:
        run32 = configs.quadrupole_1e6("fp32").with_(obs_capacity=args.single_steps + args.warmup + 64)
        sim, r = run_gpu_leg(run32, args.single_steps, args.warmup)
        r["energy_drift"] = energy_drift(sim, args.single_steps)
        sim.close()
        log(f"[bench] FP32 Kahan N=1e6: {r['value']:.4e} atom-steps/s, |dE/E| {r['energy_drift']:.2e}")
        result["fp32_single"] = r

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
                f"block {r['block']}, {r['ms_per_step'] * 1e3:.1f} us per step")
        cells_tuning = dict(best_target_peak=best["target_peak"],
                            candidates=[{k: (round(v, 5) if isinstance(v, float) else v)
                                         for k, v in r.items()} for r in report])

    head = result.get("batched") or result.get("fp32_single") or result.get("fp64_plain")
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
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "batched" in result:
        line["energy_drift"] = result["batched"]["energy_drift"]
    if "fp64_plain" in result:
        line["fp64_plain"] = {k: round(v, 3) if isinstance(v, float) else v
                              for k, v in result["fp64_plain"].items()}
        line["ratio_vs_fp64_plain"] = round(head["value"] / result["fp64_plain"]["value"], 3)
    if "fp32_single" in result:
        line["fp32_single"] = {k: (round(v, 3) if isinstance(v, float) and k != "energy_drift" else v)
                               for k, v in result["fp32_single"].items()}
    if kernels:
        line["kernels"] = kernels
    if cells_tuning:
        line["cells_tuning"] = cells_tuning
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

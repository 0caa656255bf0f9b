#!/usr/bin/env python3
"""Benchmark: FP64 SGH zone-cycles/s on one MI355X, with HBM roofline and a CPU baseline.

Workload: the 3D point-energy blast (Sedov-type) on the [0,1.2]^3 uniform
hex mesh at 256^3 elements (16,777,216 zones, 16,974,593 nodes), Morton
renumbered, ideal gas gamma=5/3, q1=1, q2=1.333, RK2, CFL 0.5.  A "step" is
one full RK2 cycle of the fused, device-resident cycle loop (dt reduction,
2 x {corner force, velocity+BC+position, energy+volume+EOS+divergence}).
Inputs are resident in HBM before the timed region starts.

One JSON line on stdout (rank 0).  --gpus N>1 is not supported: this build
runs one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "FP64 SGH zone-cycles/s & %HBM roofline, plain vs batched; adjoint/fwd time"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """(CPU model, physical cores among the CPUs this process may run on)."""
    model, cores, cur = "unknown", set(), {}
    try:
        allowed = os.sched_getaffinity(0)
    except AttributeError:
        allowed = set(range(os.cpu_count() or 1))
    try:
        with open("/proc/cpuinfo") as f:
            blocks = f.read().strip().split("\n\n")
        for blk in blocks:
            cur = dict(line.split(":", 1) for line in blk.splitlines() if ":" in line)
            cur = {k.strip(): v.strip() for k, v in cur.items()}
            if int(cur.get("processor", -1)) in allowed:
                model = cur.get("model name", model)
                cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
    except (OSError, ValueError, KeyError):
        pass
    return model, max(len(cores), 1) if cores else len(allowed)


def baseline_threads():
    """Physical cores of this process's affinity, capped by the OpenMP thread
    share the machine grants (OMP_NUM_THREADS when set)."""
    model, phys = host_cpu()
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(phys, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else phys
    return model, phys, threads


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # noqa: E402  (test/baseline infrastructure)
    return O


def cpu_baseline(n, cycles, one_n=64, one_cycles=10):
    """The oracle (a C/OpenMP restatement of the reference operators, built
    with gcc -O3) timed on the host cores: ``cycles`` RK2 cycles of the same
    blast at n^3 on every core this process may use (set-up excluded, as
    K/sgh_solve.cpp:122,637-641 times only the cycle loop), and a 1-thread
    rate on one_n^3."""
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    O = _oracle()
    from fierro_amd import problems as P

    model, phys, threads = baseline_threads()
    L = O.lib(fast=True)
    L.orc_set_num_threads(threads)
    t = time.perf_counter()
    pr = O.OracleProblem(P.sedov(n, cycles=cycles), fast=True)
    log(f"cpu baseline: oracle set-up {n}^3 in {time.perf_counter() - t:.1f} s")
    r = pr.solve()
    E = pr.E
    del pr
    zc = E * r["cycles"] / r["seconds"]
    log(f"cpu baseline: {r['cycles']} cycles at {n}^3 on {threads} threads in "
        f"{r['seconds']:.2f} s")
    L.orc_set_num_threads(1)
    p1 = O.OracleProblem(P.sedov(one_n, cycles=one_cycles), fast=True)
    r1 = p1.solve()
    one = p1.E * r1["cycles"] / r1["seconds"]
    L.orc_set_num_threads(threads)
    return {"value": zc, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "cpu_model": model, "physical_cores_in_affinity": phys,
            "one_thread": {"value": one, "unit": "zone-cycles/s",
                           "sample": f"sedov {one_n}^3, {r1['cycles']} cycles, "
                                     f"{r1['seconds']:.2f} s"},
            "sample": f"sedov {n}^3 ({E} zones), {r['cycles']} RK2 cycles, "
                      f"{r['seconds']:.2f} s on {threads} threads, "
                      f"oracle/liboracle_sgh_fast.so (gcc -O3, OpenMP)"}


def cpu_baseline_rz(n, cycles):
    """The RZ oracle on the host cores over a bounded sample of the 2D RZ
    blast."""
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "threads")
    O = _oracle()
    from fierro_amd import problems as P

    model, phys, threads = baseline_threads()
    O.lib(fast=True).orc_set_num_threads(threads)

    spec = P.sedov_rz(n, cycles=cycles)
    pr = O.OracleRZ(spec, fast=True)
    r = pr.solve()
    zc = pr.E * r["cycles"] / r["seconds"]
    return {"value": zc, "unit": "zone-cycles/s", "cores": threads, "kind": "port",
            "cpu_model": model,
            "sample": f"sedov_rz {n}^2 ({pr.E} zones), {r['cycles']} RK2 cycles, "
                      f"{r['seconds']:.2f} s, oracle/liboracle_sgh_fast.so (gcc -O3, OpenMP)"}


def measured_traffic(workload, E, N, kernel):
    """HBM bytes per launch of `kernel` measured with rocprofv3 FETCH_SIZE /
    WRITE_SIZE (scaled by a plain copy's counters) for this workload, from
    profiles/<workload>_traffic.json; None when there is no such record."""
    path = os.path.join(ROOT, "profiles", f"{workload}_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except OSError:
        return None
    if rec.get("elements") != E or rec.get("nodes") != N or kernel not in rec["kernels"]:
        return None
    return rec["kernels"][kernel]["traffic_bytes"]


# kernel keys of the traffic records: workload -> {key: kernel-name prefix}
TRAFFIC_KERNELS = {
    "sedov": {"div_extremes": "k_div_extremes<", "force": "k_force<",
              "vel_bc_pos": "k_vel_bc_pos<", "energy_state_div": "k_energy_state_div<"},
    "sedov-rz": {"rz_div_extremes": "k_div_extremes<", "rz_force": "k_rz_force",
                 "rz_node": "k_rz_node", "rz_energy_state_div": "k_rz_energy_state_div"},
    # the reverse pass of the calibration gradient (every launch covers the batch)
    "calibration": {"adj_energy": "k_adj_energy", "adj_node": "k_adj_node<",
                    "arg_extremes": "k_arg_extremes", "adj_ext": "k_adj_force_pair<true",
                    "adj_ext_node": "k_adj_ext_node",
                    "adj_force_derived": "k_adj_force_pair<false",
                    "adj_gather": "k_adj_gather<"},
}


def traffic_record(workload, n, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, out):
    """Write profiles/<workload>_traffic.json from four rocprofv3 --pmc runs
    (FETCH_SIZE and WRITE_SIZE of a bench run of this workload, and of
    --copy-calib 1024): per kernel the mean counter value per dispatch, in
    KiB, scaled by the copy's bytes over its counter value."""
    import csv

    def per_kernel(d, counter):
        acc = {}
        with open(os.path.join(d, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                name = r["Kernel_Name"]
                name = name[name.find("k_"):] if "::k_" in name else name
                acc.setdefault(name.split("(")[0], []).append(float(r["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}

    copy_bytes = 1024 << 20
    cal = {}
    for counter, d in (("FETCH_SIZE", copy_fetch_dir), ("WRITE_SIZE", copy_write_dir)):
        vals = [v for k, v in per_kernel(d, counter).items() if "copyBuffer" in k]
        cal[counter] = copy_bytes / (1024.0 * vals[0])
    fetch = per_kernel(fetch_dir, "FETCH_SIZE")
    write = per_kernel(write_dir, "WRITE_SIZE")
    from fierro_amd import bytemodel, problems as P
    from fierro_amd.sgh import SGHSolver

    if workload == "calibration":
        specs, _, _, _ = P.blast_ensemble(64, seed=0, cycles=150)
        g = SGHSolver(specs, renumber=True)
        alg = {k: fn(g.E, g.N) for k, fn in bytemodel.ADJ_KERNELS.items()}
        name = "blast_calibration64"
    else:
        spec = P.sedov_rz(n, cycles=1) if workload == "sedov-rz" else P.sedov(n, cycles=1)
        g = SGHSolver(spec, renumber=True)
        if workload == "sedov-rz":
            alg = rz_alg_bytes(g)
        else:
            kernels, _ = kernel_table(g, {}, tune(g, argparse.Namespace(tune="table")))
            alg = {k: v["alg_bytes"] for k, v in kernels.items()}
        name = f"{workload.replace('-', '_')}{n}"
    rec = {"workload": name, "elements": g.E, "nodes": g.N, "calibration": cal, "kernels": {}}
    for key, prefix in TRAFFIC_KERNELS[workload].items():
        names = [k for k in fetch if k.startswith(prefix)]
        if len(names) != 1:
            raise SystemExit(f"traffic: kernel {prefix!r} matched {names}")
        fb = fetch[names[0]] * 1024 * cal["FETCH_SIZE"]
        wb = write[names[0]] * 1024 * cal["WRITE_SIZE"]
        rec["kernels"][key] = {"variant": names[0], "fetch_bytes": fb, "write_bytes": wb,
                               "traffic_bytes": fb + wb, "alg_bytes": alg[key]}
    with open(out, "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec, indent=1), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="elements per edge")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="the plain line only (no batched, adjoint, calibration and "
                         "operator-path records)")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-cycles", type=int, default=2)
    ap.add_argument("--workload",
                    choices=["sedov", "sedov-rz", "sod-ensemble", "anvil-adjoint",
                             "calibration", "calibration-reverse", "sedov-operators"],
                    default="sedov")
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--checkpoint-every", type=int, default=4,
                    help="calibration workload: cycles between checkpoints of the forward")
    ap.add_argument("--tune", choices=["table", "auto", "default"], default="table",
                    help="launch variants of the fused kernels: the measured table "
                         "(fierro_amd/tuning.py), a fresh fsgh_autotune, or the defaults")
    ap.add_argument("--loop-sample", type=int, default=1,
                    help="time the per-instance loop on every k-th instance (1: all)")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    ap.add_argument("--traffic-from", nargs=4, default=None,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="build profiles/<workload>_traffic.json from rocprofv3 --pmc output "
                         "directories (this workload's FETCH_SIZE and WRITE_SIZE runs, and "
                         "the same counters of --copy-calib 1024)")
    args = ap.parse_args()
    if args.traffic_from:
        name = "sedov_rz" if args.workload == "sedov-rz" else "sedov"
        n = args.n if args.workload == "sedov" else (args.n if args.n != 256 else 4096)
        out = (f"{name}{n}_traffic.json" if args.workload != "calibration"
               else "blast_calibration64_traffic.json")
        return traffic_record(args.workload, n, *args.traffic_from,
                              os.path.join(ROOT, "profiles", out))
    if args.copy_calib:
        return copy_calibration(args.copy_calib)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "sod-ensemble":
        return emit(ensemble(args))
    if args.workload == "anvil-adjoint":
        return emit(adjoint(args))
    if args.workload == "calibration":
        return emit(calibration(args))
    if args.workload == "calibration-reverse":
        return emit(calibration_reverse(args))
    if args.workload == "sedov-rz":
        return sedov_rz(args)
    if args.workload == "sedov-operators":
        return emit(sedov_operators(args))

    out = plain(args)
    if not args.no_extras:
        # the metric's other parts, each on its own workload (the 512-tube
        # ensemble, the anvil adjoint, the 64-blast calibration) and the
        # drop-in operator path at the plain mesh
        out["batched"] = part(ensemble(args, steps=500, warmup=5),
                              ("batched_over_loop", "loop_plain", "roofline", "kernels"))
        out["adjoint"] = part(adjoint(args), ("plain_forward_s", "checkpointed_forward_s",
                                              "reverse_s", "plain_zone_cycles_per_s"))
        out["calibration"] = part(calibration(args), ("plain_forward_s", "gradient_s",
                                                      "plain_zone_cycles_per_s"))
        out["operators"] = part(sedov_operators(args), ("l1_den_vs_fused",))
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_cycles)
        except Exception as exc:  # the baseline is reported, never the product
            out["cpu_baseline"] = {"value": None, "error": repr(exc)}
    else:
        out["cpu_baseline"] = None
    emit(out)


def emit(out):
    print(json.dumps(out), flush=True)


def part(rec, extra):
    """The headline figures of one workload's record, for the combined line."""
    keep = ("metric", "value", "unit", "higher_is_better", "steps", "ms_per_step", "config")
    return {k: rec[k] for k in keep + tuple(extra) if k in rec}


def plain(args):
    """The plain 256^3 blast: fused, device-resident cycle loop, HIP events
    per kernel in the timed region."""
    from fierro_amd import bytemodel, problems as P
    from fierro_amd.sgh import SGHSolver

    n = args.n
    spec = P.sedov(n, cycles=10 ** 9)
    t0 = time.perf_counter()
    g = SGHSolver(spec, renumber=True)
    log(f"setup {n}^3: E={g.E} N={g.N} in {time.perf_counter() - t0:.1f} s")
    tuning = tune(g, args)
    g.run_begin()
    g.run(args.warmup, use_graph=True)
    log(f"warmup {args.warmup} cycles done")

    # graph-replayed loop, no per-kernel events
    t = time.perf_counter()
    g.run(args.steps, use_graph=True)
    graph_s = time.perf_counter() - t
    log(f"graph loop: {args.steps} cycles in {graph_s:.3f} s")

    # timed region: per-kernel HIP events on the context stream
    g.profile(True)
    t = time.perf_counter()
    g.run(args.steps, use_graph=False)
    wall = time.perf_counter() - t
    prof = g.profile_read()
    g.profile(False)
    st = g.status(0)
    log(f"timed loop: {args.steps} cycles in {wall:.3f} s; status {st}")
    # every cycle of both loops ran (fsgh_run stops early at time_final or
    # cycle_stop, which would inflate the rate)
    if st["cycles"] != args.warmup + 2 * args.steps:
        raise SystemExit(f"bench: {st['cycles']} cycles ran, expected "
                         f"{args.warmup + 2 * args.steps}")
    ie, ke = g.energy(0)
    if not math.isfinite(ie + ke):
        raise SystemExit(f"bench: non-finite state after the timed cycles (IE {ie}, KE {ke})")

    E, N = g.E, g.N
    kernels, lists = kernel_table(g, prof, tuning)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["calls"])
    d = kernels[dom]
    cyc_bytes = bytemodel.cycle(E, N, lists)
    value = E * args.steps / wall
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated Sedov blast mesh and initial state)",
        "config": {"workload": f"sedov{n}", "elements": E, "nodes": N, "rk_stages": 2,
                   "renumber": "morton", "cfl": 0.5, "q2": 1.333},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d["GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": d["GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"sedov{n}", E, N, dom)},
        "kernels": kernels,
        "cycle_alg_GBs": cyc_bytes * args.steps / wall / 1e9,
        "value_graph": E * args.steps / graph_s,
        "energy": {"ie": ie, "ke": ke, "time": st["time_value"]},
        "tuning": tuning,
    }
    g.close()
    return out


def rz_alg_bytes(g):
    """Algorithmic bytes per launch of the fused RZ kernels of context g; the
    element kernels' count the node lists of their 256-quad blocks."""
    from fierro_amd import bytemodel

    entries = g.stage_info(256)[0]
    out = {}
    for name, fn in bytemodel.RZ_KERNELS.items():
        staged = name in ("rz_force", "rz_energy_state_div") and entries
        out[name] = fn(g.E, g.N, entries) if staged else fn(g.E, g.N)
    return out


def sedov_rz(args):
    """2D axisymmetric blast (the reference's SedovRZ input) on n x n quads
    of [0,1.2]^2 (default 4096^2 = 16.8M zones, the zone count of the 3D
    256^3 workload), Morton renumbered; the fused RZ cycle loop."""
    from fierro_amd import bytemodel, problems as P
    from fierro_amd.sgh import SGHSolver

    n = args.n if args.n != 256 else 4096
    # the reference's dt_min = 1e-8 floor exceeds the deposit's CFL step on
    # meshes this fine (2.6e-9 at 4096^2) and would clamp the run unstable
    dt_min = 1.0e-8 if n <= 2048 else 1.0e-12
    spec = P.sedov_rz(n, cycles=10 ** 9, dt_min=dt_min)
    t0 = time.perf_counter()
    g = SGHSolver(spec, renumber=True)
    log(f"setup rz {n}^2: E={g.E} N={g.N} in {time.perf_counter() - t0:.1f} s")
    g.run_begin()
    g.run(args.warmup, use_graph=True)
    t = time.perf_counter()
    g.run(args.steps, use_graph=True)
    graph_s = time.perf_counter() - t
    g.profile(True)
    t = time.perf_counter()
    g.run(args.steps, use_graph=False)
    wall = time.perf_counter() - t
    prof = g.profile_read()
    g.profile(False)
    st = g.status(0)
    log(f"timed loop: {args.steps} cycles in {wall:.3f} s (graph loop {graph_s:.3f} s); {st}")
    ie, ke = g.energy(0)
    if not math.isfinite(ie + ke):
        raise SystemExit(f"bench: non-finite state after the timed cycles (IE {ie}, KE {ke})")
    E, N = g.E, g.N
    alg = rz_alg_bytes(g)
    kernels = {}
    for name in bytemodel.RZ_KERNELS:
        ms, calls = prof.get(name, (0.0, 0))
        avg = ms / max(calls, 1)
        b = alg[name]
        kernels[name] = {"avg_ms": avg, "calls": calls, "alg_bytes": b,
                         "GBs": b / (avg * 1e-3) / 1e9 if avg > 0 else 0.0}
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["calls"])
    out = {
        "metric": "FP64 SGH zone-cycles/s (2D RZ Sedov blast)",
        "value": E * args.steps / wall, "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated RZ Sedov mesh and initial state)",
        "config": {"workload": f"sedov_rz{n}", "elements": E, "nodes": N, "rk_stages": 2,
                   "renumber": "morton", "cfl": 0.4, "q2": 1.3333, "dt_min": dt_min},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kernels[dom]["GBs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(f"sedov_rz{n}", E, N, dom)},
        "kernels": kernels,
        "value_graph": E * args.steps / graph_s,
    }
    del g
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline_rz(1024, 100)
        except Exception as exc:  # the baseline is reported, never the product
            out["cpu_baseline"] = {"value": None, "error": repr(exc)}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def sedov_operators(args):
    """The drop-in path: the reference's sgh_solve cycle (K/sgh_solve.cpp:
    125-634) driven through the per-stage C ABI operators, as a C++ caller
    replacing the Kokkos calls would drive it: get_timestep with dt returned
    to the host every cycle (the reference's REDUCE_MIN), rk_init, then per
    RK stage divergence, force, velocity, boundary, energy, position, volume
    and state.  Sedov blast n^3, Morton order, the shipped launch table.
    Also reports the density difference against the fused loop after the
    same cycles."""
    import numpy as np
    from fierro_amd import problems as P
    from fierro_amd.sgh import SGHSolver

    n = args.n
    spec = P.sedov(n, cycles=10 ** 9)
    r = spec.run
    g = SGHSolver(spec, renumber=True)
    tuning = tune(g, args)

    def cycles(k, t, dt):
        for _ in range(k):
            dt = g.get_timestep(t, r["graphics_dt_ival"], r["time_final"], r["dt_max"],
                                r["dt_min"], r["dt_cfl"], dt, r["fuzz"])
            g.rk_init()
            for stage in range(2):
                g.stage(dt, 1.0 / (2.0 - stage))
            t += dt
        return t, dt

    t, dt = cycles(args.warmup, 0.0, r["dt_start"])
    g.energy(0)                        # synchronises the context's stream
    w = time.perf_counter()
    t, dt = cycles(args.steps, t, dt)
    g.energy(0)
    wall = time.perf_counter() - w
    den_ops = g.download("elem_den")
    E = g.E
    g.close()
    f = SGHSolver(P.sedov(n, cycles=args.warmup + args.steps), renumber=True)
    tune(f, args)
    f.sgh_solve()
    den_fused = f.download("elem_den")
    f.close()
    diff = float(np.abs(den_ops - den_fused).sum() / np.abs(den_fused).sum())
    log(f"operators: {args.steps} cycles in {wall:.3f} s; L1(den) vs fused loop {diff:.3e}")
    out = {
        "metric": "FP64 SGH zone-cycles/s through the per-stage operator API (Sedov blast)",
        "value": E * args.steps / wall, "unit": "zone-cycles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated Sedov blast mesh and initial state)",
        "config": {"workload": f"sedov{n}-operators", "elements": E, "rk_stages": 2,
                   "renumber": "morton", "dt_to_host": "every cycle"},
        "l1_den_vs_fused": diff, "time": t, "tuning": tuning, "cpu_baseline": None,
    }
    return out


def kernel_table(g, prof, tuning):
    """Per fused 3D kernel: mean launch time, launches, algorithmic bytes and
    GB/s.  The element kernels' bytes count the node lists of their element
    partitions (the library reports their size); the velocity kernel's count
    the staged corner lists (every node block of these hex meshes, at most 8
    corners per node, fits the kernel's LDS capacity of 9 per node)."""
    from fierro_amd import bytemodel

    lists = {}
    for name in ("force", "energy_state_div"):
        bs = int(tuning.get(name, {}).get("block", 256))
        n = g.stage_info(bs)[0]
        if n:
            lists[name] = n
    kernels = {}
    for name in bytemodel.KERNELS:
        ms, calls = prof.get(name, (0.0, 0))
        avg = ms / max(calls, 1)
        b = bytemodel.kernel_bytes(name, g.E, g.N, lists.get(name))
        kernels[name] = {"avg_ms": avg, "calls": calls, "alg_bytes": b,
                         "GBs": b / (avg * 1e-3) / 1e9 if avg > 0 else 0.0}
    return kernels, lists


def tune(g, args):
    """Autotune the fused-cycle kernels' launch variants (outside the timed
    region; the state is unchanged) and return the choice per kernel."""
    if args.tune == "default":
        return g.get_tuning()
    if args.tune == "table":
        from fierro_amd.tuning import TABLE
        g.set_tuning(TABLE)
        return g.get_tuning()
    t = time.perf_counter()
    res = g.autotune(reps=5)
    log(f"autotune in {time.perf_counter() - t:.1f} s: " + ", ".join(
        f"{k} {v['block']}/{v['min_waves']}/{v['xcd_order']} {v['ms']:.3f} ms"
        f" (default {v['ms_default']:.3f})" for k, v in res.items()))
    return res


def copy_calibration(mib, reps=3):
    """A plain device-to-device copy of known size through the HIP runtime's
    copy kernel: FETCH_SIZE / WRITE_SIZE of this kernel divided by the bytes
    copied gives the counters' scale."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n, "reps": reps}), flush=True)


def ensemble(args, steps=500, warmup=5):
    """Batched ensemble: 512 Sod tubes (nx ~ U{64..1024} x 4 x 4, seeded) in
    one ragged context, ``steps`` (500) timed RK2 cycles after ``warmup``
    with per-instance dt; compared with looping the plain solver over the
    instances (every instance with --loop-sample 1, the default)."""
    from fierro_amd import bytemodel, problems as P
    from fierro_amd.sgh import SGHSolver

    args = argparse.Namespace(**{**vars(args), "warmup": warmup})
    specs = [s.with_run(cycle_stop=args.warmup + steps)
             for s in P.sod_ensemble(args.instances, seed=0, cycles=steps)]
    t0 = time.perf_counter()
    g = SGHSolver(specs, renumber=True)
    log(f"ensemble setup: {len(specs)} instances E={g.E} N={g.N} in "
        f"{time.perf_counter() - t0:.1f} s")
    # the shipped table was measured on the 256^3 cube; the ensemble's thin
    # tubes (1.56 nodes per element) get their own launch variants, timed on
    # this batch outside the timed region (1.64e9 against 1.58e9 zone-cycles/s
    # with the cube's table); the per-instance loop reuses the choice
    tuning = tune(g, argparse.Namespace(tune="auto" if args.tune == "table" else args.tune))
    g.run_begin()
    g.run(args.warmup, use_graph=True)
    g.profile(True)
    t = time.perf_counter()
    g.run(steps, use_graph=False)
    wall = time.perf_counter() - t
    prof = g.profile_read()
    g.profile(False)
    cyc = [g.status(i)["cycles"] for i in range(len(specs))]
    assert min(cyc) == max(cyc) == args.warmup + steps, (min(cyc), max(cyc))
    if not all(math.isfinite(sum(g.energy(i))) for i in range(len(specs))):
        raise SystemExit("bench: non-finite instance state after the timed cycles")
    E, N = g.E, g.N
    kernels, _ = kernel_table(g, prof, tuning)
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["calls"])
    value = E * steps / wall
    del g
    # loop of the plain solver, one context per instance
    loop_zc, loop_s = 0, 0.0
    sample = specs[::args.loop_sample]
    for k, s in enumerate(sample):
        if k % 64 == 0:
            log(f"per-instance loop: instance {k * args.loop_sample} of {len(specs)}")
        p = SGHSolver(s, renumber=True)
        p.set_tuning({k: {q: v[q] for q in ("block", "min_waves", "xcd_order")}
                      for k, v in tuning.items()})
        p.run_begin()
        p.run(args.warmup, use_graph=True)
        t = time.perf_counter()
        p.run(steps, use_graph=True)
        loop_s += time.perf_counter() - t
        loop_zc += p.E * steps
        del p
    loop = loop_zc / loop_s
    out = {
        "metric": "FP64 SGH zone-cycles/s, batched ensemble (512 Sod tubes)",
        "value": value, "unit": "zone-cycles/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded Sod-tube ensemble)",
        "config": {"workload": f"sod-ensemble{len(specs)}", "elements": E, "nodes": N,
                   "seed": 0, "rk_stages": 2},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": kernels[dom]["GBs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kernels[dom]["GBs"] / HBM_PEAK_GBS, "traffic": None},
        "kernels": kernels,
        "loop_plain": {"value": loop, "unit": "zone-cycles/s",
                       "sample": (f"all {len(specs)} instances" if args.loop_sample == 1 else
                                  f"every {args.loop_sample}th instance") +
                       f", one context each, {steps} cycles, hipGraph per cycle"},
        "batched_over_loop": value / loop,
        "tuning": tuning,
        "cpu_baseline": None,
    }
    return out


def adjoint(args):
    """Adjoint cost: the Taylor-anvil bar (16 x 16 x 96 hexes, 400 RK2
    cycles).  Gradient of bar length + base spread with respect to every
    initial density, energy, gamma and velocity = one checkpointed forward
    (every 20 cycles) + one reverse pass; compared with one plain forward
    (fused loop, hipGraph per cycle) of the same cycles."""
    from fierro_amd import problems as P
    from fierro_amd.sgh import SGHSolver

    cycles, every = 400, 20
    spec = P.anvil(16, 16, 96, cycles=cycles)
    plain = []
    for _ in range(3):
        g = SGHSolver(spec, renumber=True)
        tune(g, args)
        g.run_begin()
        t = time.perf_counter()
        g.run(cycles, use_graph=True)
        plain.append(time.perf_counter() - t)
        E = g.E
        del g
    t_plain = min(plain)
    fw, bw = [], []
    for _ in range(3):
        g = SGHSolver(spec, renumber=True)
        tune(g, args)
        seed = P.anvil_objective(spec, g.coords0)
        t = time.perf_counter()
        g.forward(cycles, checkpoint_every=every)
        t1 = time.perf_counter()
        grad = g.backward(seed_x=seed)
        t2 = time.perf_counter()
        fw.append(t1 - t)
        bw.append(t2 - t1)
        del g
    t_fw, t_bw = min(fw), min(bw)
    ratio = (t_fw + t_bw) / t_plain
    log(f"plain {t_plain:.3f} s, checkpointed forward {t_fw:.3f} s, reverse {t_bw:.3f} s")
    out = {
        "metric": "adjoint gradient time / plain forward time (anvil 16x16x96, 400 cycles)",
        "value": ratio, "unit": "x", "n_gpus": 1, "steps": cycles, "warmup": 0,
        "ms_per_step": 1e3 * (t_fw + t_bw) / cycles, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (generated anvil bar and initial state)",
        "config": {"workload": "anvil16x16x96-adjoint", "elements": E, "cycles": cycles,
                   "checkpoint_every": every},
        "plain_forward_s": t_plain, "checkpointed_forward_s": t_fw, "reverse_s": t_bw,
        "plain_zone_cycles_per_s": E * cycles / t_plain,
        "grad_norm_den0": float((grad["den0"] ** 2).sum() ** 0.5),
        "cpu_baseline": None,
    }
    return out


def calibration(args):
    """Batched forward + adjoint calibration sweep: 64 point blasts, n^3 with
    n ~ U{48..96}, E0 ~ U(0.1,1), gamma ~ U(1.3,1.8), rho0 = 1, 150 cycles,
    seed 0; gradient of each instance's radial density-profile misfit against
    the target (the next draw of the stream) with respect to (E0, gamma, rho0)
    from one batched checkpointed forward (every --checkpoint-every cycles)
    + one reverse pass; compared with one plain batched forward (fused loop,
    hipGraph)."""
    import numpy as np
    from fierro_amd import calibration as CAL, problems as P
    from fierro_amd.sgh import SGHSolver

    cycles, every = 150, args.checkpoint_every
    specs, params, tspec, tparams = P.blast_ensemble(64, seed=0, cycles=cycles)
    t0 = time.perf_counter()
    target = CAL.target_profile(SGHSolver, tspec, cycles)
    log(f"target profile ({tparams}) in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    g = SGHSolver(specs, renumber=True)
    E = g.E
    log(f"calibration setup: {len(specs)} instances E={E} N={g.N} in "
        f"{time.perf_counter() - t0:.1f} s")
    tune(g, args)
    g.run_begin()
    t = time.perf_counter()
    g.run(cycles, use_graph=True)
    t_plain = time.perf_counter() - t
    g.close()
    g = SGHSolver(specs, renumber=True)
    tune(g, args)
    layout = CAL.Layout(g)          # bins and origin elements of the initial mesh (set-up)
    t = time.perf_counter()
    J, dE0, dgam, drho, _ = CAL.gradient(g, params, target, cycles, every, layout=layout)
    t_grad = time.perf_counter() - t
    g.close()
    log(f"plain forward {t_plain:.3f} s, forward + reverse {t_grad:.3f} s")
    out = {
        "metric": "batched forward+adjoint time / plain batched forward time "
                  "(64 blasts, 150 cycles)",
        "value": t_grad / t_plain, "unit": "x", "n_gpus": 1, "steps": cycles, "warmup": 0,
        "ms_per_step": 1e3 * t_grad / cycles, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded blast ensemble)",
        "config": {"workload": "blast-calibration64", "elements": E, "cycles": cycles,
                   "checkpoint_every": every, "seed": 0, "bins": CAL.NBINS,
                   "rmax": CAL.RMAX},
        "plain_forward_s": t_plain, "gradient_s": t_grad,
        "plain_zone_cycles_per_s": E * cycles / t_plain,
        "misfit_sum": float(np.sum(J)),
        "grad_norms": {"E0": float(np.linalg.norm(dE0)), "gamma": float(np.linalg.norm(dgam)),
                       "rho0": float(np.linalg.norm(drho))},
        "cpu_baseline": None,
    }
    return out


def calibration_reverse(args):
    """The reverse kernels of the calibration gradient (64 blasts, 150
    cycles, --checkpoint-every), each timed alone: with per-kernel HIP events
    on, the reverse pass keeps one checkpoint bank on the context stream (no
    recompute overlapped with it, no captured segments) and finds the
    divergence extremes itself (k_arg_extremes).  Per kernel: average launch
    time, algorithmic bytes per launch (fierro_amd/bytemodel.py) and their
    rate against 8 TB/s."""
    from fierro_amd import bytemodel, calibration as CAL, problems as P
    from fierro_amd.sgh import SGHSolver

    cycles, every = 150, args.checkpoint_every
    specs, params, tspec, _ = P.blast_ensemble(64, seed=0, cycles=cycles)
    target = CAL.target_profile(SGHSolver, tspec, cycles)
    g = SGHSolver(specs, renumber=True)
    E, N = g.E, g.N
    log(f"calibration reverse: {len(specs)} instances E={E} N={N}")
    tune(g, args)
    layout = CAL.Layout(g)
    g.profile(True)
    t = time.perf_counter()
    CAL.gradient(g, params, target, cycles, every, layout=layout)
    wall = time.perf_counter() - t
    prof = g.profile_read()
    g.profile(False)
    g.close()
    kernels = {}
    for name, fn in bytemodel.ADJ_KERNELS.items():
        if name not in prof:
            continue
        ms, calls = prof[name]
        avg = ms / calls
        b = fn(E, N)
        kernels[name] = {"calls": calls, "avg_ms": avg, "alg_bytes": b,
                         "alg_GBps": b / avg / 1e6, "frac": b / avg / 1e6 / HBM_PEAK_GBS}
        log(f"  {name:13s} {calls:5d} x {1e3 * avg:9.1f} us  {b / 1e9:6.2f} GB  "
            f"{b / avg / 1e6:7.0f} GB/s  {100 * b / avg / 1e6 / HBM_PEAK_GBS:5.1f} %")
    stage = sum(k["avg_ms"] for k in kernels.values())
    return {"metric": "reverse-pass kernels of the calibration gradient (each timed alone)",
            "value": bytemodel.adj_stage(E, N) / stage / 1e6, "unit": "GB/s (one reversed "
            "stage: algorithmic bytes / sum of its kernels' average times)", "n_gpus": 1,
            "steps": cycles, "warmup": 0, "ms_per_step": 2 * stage, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded blast ensemble)",
            "config": {"workload": "blast-calibration64-reverse", "elements": E, "nodes": N,
                       "cycles": cycles, "checkpoint_every": every},
            "gradient_s_one_bank": wall, "kernels": kernels, "cpu_baseline": None}


if __name__ == "__main__":
    main()

"""Benchmark of the energy-game fixpoint on one MI355X.

Workload (default): the single large uniform game -- n = 2^22 vertices, owner uniform,
out-degree exactly 8 distinct uniform successors (m = 2^25), weights uniform in
[-100, 100], seed 2.  The game is generated, loaded into HBM and transposed before the
timed region.  One step is one fixpoint round: one pass of the lift kernel over the
current frontier, with in-place value updates and the predecessor worklist rebuild.
`value` is edges relaxed per second over the timed rounds.

The JSON line also carries
  roofline      the dominant kernel (the lift kernel) against MI355X's 8 TB/s HBM peak:
                algorithmic bytes per launch / average launch time (HIP events on the
                solver's stream), plus the device's measured copy bandwidth and, when a
                profile of this workload is committed, the profiler-measured HBM traffic.
  cpu_baseline  the reference's own CPU solver (cpu_solver.c compiled from its sources
                into oracle/_ref; the oracle's C restatement when that build is absent)
                on a bounded sample of the same game, timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "edges relaxed/s to fixpoint; solves/s + p99 latency, batched vs plain"
HBM_PEAK_GBPS = 8000.0


def algorithmic_bytes(st_delta):
    """B = sum_{processed v} (16 + 12 d_out(v)) + sum_{lifted v} (12 + 12 d_in(v)):
    frontier id 4 + row offsets 8 + own value 4; per out-edge successor 4 + weight 4 +
    gathered value 4; per lifted vertex value store 4 + CSC offsets 8; per in-edge of a
    lifted vertex predecessor id 4 + dedupe stamp read-modify-write 8."""
    return (16 * st_delta["processed"] + 12 * st_delta["edges_relaxed"] + 12 * st_delta["lifted"]
            + 12 * st_delta["pred_visits"])


def read_profile_traffic(workload_key):
    """HBM bytes per launch of the lift kernel from the committed rocprofv3 counter
    summary of this workload (profiles/lift_<workload>_pmc.md), or None."""
    path = os.path.join(ROOT, "profiles", f"lift_{workload_key}_pmc.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    mo = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(mo.group(1)) if mo else None


def cpu_baseline(g, budget_s=20.0):
    """Time the reference CPU solver on a bounded sample of the same game: EG0 Jacobi
    sweeps (cpu_solver.c:109-185; every sweep relaxes all m edges) on all the host
    threads OpenMP is given, for a fixed sweep count sized to ~budget_s."""
    from oracle.oracle import Reference, Oracle, game_from_arrays

    og = game_from_arrays(g.n, g.n0, g.ptr, g.succ, g.w)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = min(threads, 16)
    if Reference.available():
        R = Reference()
        kind = "reference"
        run = lambda k: R.solve(og, "eg0", max_loop=k, threads=threads)[2]
    else:
        O = Oracle()
        kind = "port"

        def run(k):
            t = time.perf_counter()
            O.eg0(og, max_loop=k, threads=threads)
            return time.perf_counter() - t
    probe = run(2)
    sweeps = max(2, int(budget_s / max(probe / 2, 1e-3)))
    secs = run(sweeps)
    return {"value": sweeps * g.m / secs, "unit": "edges relaxed/s", "cores": threads, "kind": kind,
            "sample": f"EG0 (--cpu --eg0 {sweeps}): {sweeps} Jacobi sweeps over all {g.m} edges of the same "
                      f"game from f=0, {secs:.2f} s, OpenMP guided(4), {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=400)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink n (testing only)")
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the single-game path shards by vertex range, see the design notes).", file=sys.stderr)
        sys.exit(2)

    import egsolver_amd as E

    t0 = time.time()
    g = E.config_game(2, scale=args.scale)
    gen_s = time.time() - t0
    print(f"[bench] game n={g.n} m={g.m} mg={g.mg} generated in {gen_s:.1f}s", file=sys.stderr, flush=True)

    copy_gbps = _copy_bw(E)

    s = E.Solver(g.n, g.m)
    s.load(g)
    opt = E.options(group=args.group, grid=args.grid)
    s.prepare(opt)
    prep = s.stats()
    print(f"[bench] prepared (transpose+init) in {prep.ms_prepare:.2f} ms; initial frontier "
          f"{prep.initial_active}", file=sys.stderr, flush=True)

    conv, st_w = s.iterate(args.warmup)
    w_stats = st_w.as_dict()
    t_wall = time.perf_counter()
    conv, st_t = s.iterate(args.steps)
    t_wall = time.perf_counter() - t_wall
    t_stats = st_t.as_dict()
    delta = {k: t_stats[k] - w_stats[k] for k in ("processed", "edges_relaxed", "lifted", "pred_visits", "phases")}
    ms_dev = t_stats["ms_iterate"] - w_stats["ms_iterate"]
    rounds = delta["phases"]
    if rounds <= 0:
        raise SystemExit("no rounds ran in the timed region (game converged during warm-up)")
    edges_per_s = delta["edges_relaxed"] / t_wall
    abytes = algorithmic_bytes(delta)
    per_launch_ms = ms_dev / rounds
    achieved = abytes / rounds / (per_launch_ms * 1e-3) / 1e9
    traffic = read_profile_traffic("config2")
    line = {
        "metric": METRIC,
        "value": edges_per_s,
        "unit": "edges relaxed/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "single large uniform game (n=2^22, out-degree 8, weights U[-100,100], seed 2); "
                        "step = one fixpoint round over the frontier",
            "n": g.n, "m": g.m, "mg": g.mg, "rounds_timed": rounds, "rounds_before": w_stats["phases"],
            "group": s_shape_group(args.group, g), "prepare_ms": prep.ms_prepare,
            "edges_per_round": delta["edges_relaxed"] / rounds,
            "lifted_per_round": delta["lifted"] / rounds,
            "device_ms_per_round": per_launch_ms,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": traffic,
            "kernel": f"k_round<{s_shape_group(args.group, g)}>",
            "algorithmic_bytes_per_launch": abytes / rounds,
            "copy_kernel_gbps": copy_gbps,
        },
    }
    s.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g, args.cpu_budget)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def s_shape_group(group, g):
    if group:
        return group
    avg = g.m / g.n
    p = 1
    while p < avg:
        p <<= 1
    return min(16, p)


def _copy_bw(E):
    import ctypes
    from egsolver_amd import _native as N
    out = ctypes.c_double(0)
    rc = N.egs().egs_copy_bandwidth(0, 1 << 30, 20, ctypes.byref(out))
    return out.value if rc == 0 else None


if __name__ == "__main__":
    main()

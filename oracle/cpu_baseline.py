"""CPU baseline runner (TEST INFRASTRUCTURE: run only by bench.py's cpu_baseline leg).

Runs the reference's own CPU solver (src/cpu_solver.c compiled into _ref/libegs_ref.so by
this directory's Makefile; this directory's restatement when that build is absent) on a
bounded sample of one game and prints one JSON line.  It runs as its own process so that
the OpenMP placement (OMP_PLACES, OMP_PROC_BIND, OMP_NUM_THREADS) the caller sets in its
environment is read before the OpenMP runtime starts.

  --algo eg0   Jacobi sweeps (src/cpu_solver.c:109-185, --cpu --eg0): a sweep relaxes all m
               edges; the sample is a number of sweeps from f = 0.
  --algo eg    FIFO worklist (src/cpu_solver.c:190-333, --cpu --eg, the reference default),
               single-threaded: the sample is a number of pops from f = 0; each pop relaxes
               the popped vertex's out-edges (edges = pops x degree on fixed-degree games).
               Its time includes the host transpose, as cpu_solver() times it
               (src/cpu_solver.c:336-357); a one-pop run gives that fixed cost, so the
               iteration-only time is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# run as a script, sys.path[0] is this directory, where oracle.py would shadow the package
sys.path[0] = os.path.dirname(HERE)

from oracle.oracle import Oracle, Reference, game_from_arrays  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--game", required=True, help="game cache written by egsolver_amd.save_csr")
    ap.add_argument("--algo", choices=["eg0", "eg"], required=True)
    ap.add_argument("--threads", type=int, default=1)
    ap.add_argument("--budget", type=float, default=15.0, help="seconds of CPU work for the timed sample")
    args = ap.parse_args()

    import egsolver_amd as E
    g = E.load_csr(args.game)
    og = game_from_arrays(g.n, g.n0, g.ptr, g.succ, g.w)
    deg = np.diff(og.ptr.astype(np.int64))
    fixed_deg = int(deg[0]) if (deg == deg[0]).all() else None
    ref = Reference.available()
    kind = "reference" if ref else "port"
    R = Reference() if ref else None
    O = Oracle()

    def run(cap):
        if ref:
            _, processed, secs = R.solve(og, args.algo, max_loop=cap, threads=args.threads)
            return processed, secs
        t = time.perf_counter()
        if args.algo == "eg0":
            _, sweeps = O.eg0(og, max_loop=cap, threads=args.threads)
            processed = sweeps * og.n
        else:
            _, processed = O.eg(og, max_loop=cap)
        return processed, time.perf_counter() - t

    out = {"algo": args.algo, "threads": args.threads, "kind": kind, "n": og.n, "m": og.m,
           "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PLACES", "OMP_PROC_BIND")}}
    if args.algo == "eg0":
        _, probe = run(2)
        sweeps = max(2, int(args.budget / max(probe / 2, 1e-3)))
        processed, secs = run(sweeps)
        done = processed // og.n
        out.update({"units": done, "unit_name": "sweeps", "edges_relaxed": done * og.m, "solve_s": secs,
                    "iteration_s": secs, "edges_per_s": done * og.m / secs,
                    "sample": f"EG0 (--cpu --eg0 {sweeps}): {done} Jacobi sweeps over all {og.m} edges from f=0, "
                              f"{secs:.2f} s, OpenMP guided(4), {args.threads} threads"})
    else:
        if fixed_deg is None:
            raise SystemExit("the EG sample needs a fixed-degree game (edges = pops x degree)")
        # fixed cost of a call (initial queue and host transpose): a run of one pop
        _, t_fixed = run(1)
        p1 = max(1000, og.n // 8)
        _, t1 = run(p1)
        _, t2 = run(8 * p1)
        per_pop = max((t2 - t1) / (7 * p1), 1e-9)
        pops = max(8 * p1, int(max(args.budget - t_fixed, 1.0) / per_pop / 2))   # later pops cost more
        processed, secs = run(pops)
        it = max(secs - t_fixed, 1e-9)
        out.update({"units": processed, "unit_name": "pops", "edges_relaxed": processed * fixed_deg,
                    "solve_s": secs, "transpose_s": t_fixed, "iteration_s": it,
                    "edges_per_s": processed * fixed_deg / secs, "edges_per_s_iteration_only": processed * fixed_deg / it,
                    "sample": f"EG (--cpu --eg {pops}): {processed} worklist pops x {fixed_deg} out-edges from f=0, "
                              f"{secs:.2f} s including the initial queue and host transpose ({t_fixed:.2f} s, a "
                              f"one-pop run), 1 thread"})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

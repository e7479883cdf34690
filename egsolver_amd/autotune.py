"""Autotuner for launch shapes and thresholds, timed on the device through the C-ABI.

* Single game: what decides its solve time, tuned on the game itself by coordinate descent
  over full solves to the least fixpoint (proofs on, wall time of prepare + iterate):
    - the phase count of the dense rounds (``slices``; 1 = one in-place sweep), which also
      changes how many rounds the solve takes;
    - the degree thresholds that move a vertex from one lane to a wavefront (``big_deg``)
      and to a whole workgroup (``huge_deg``), when the game has vertices above them;
    - the workgroups per CU of the phased kernels (``phase_wgs_per_cu``), or the lanes per
      vertex and grid of the one-sweep round (``group``, ``grid``).
  The values do not depend on any of these (tests check that tuned options stay exact).
* Many games: workgroup size, LDS share per workgroup (which decides which games keep
  their values in LDS), workgroups per CU and lanes per edge, timed by solving a
  size-stratified sample of the batch.
Each tuner returns the best options and the table it measured.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from .batch import Batch, batch_options
from .solver import Solver, options

_TUNED_KEYS = ("slices", "big_deg", "huge_deg", "phase_wgs_per_cu", "group", "grid")


def _solve_ms(s, opt, reps):
    s.prepare(opt)
    s.iterate()                                   # warm
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        s.prepare(opt)
        conv, _ = s.iterate()
        ts.append(time.perf_counter() - t0)
        if not conv:
            return float("inf")
    return 1e3 * float(np.median(ts))


def tune_single(g, solver=None, reps=3, slices=None, big_degs=(32, 64, 128, 255), huge_degs=(1024, 4096, 16384),
                wgs_per_cu=(4, 8, 12), groups=(1, 2, 4, 8, 16), grids=(512, 1024, 2048)):
    """Coordinate descent over the options of ``_TUNED_KEYS`` (starting from the defaults):
    each step times full solves of the game for every value of one option, keeps the best,
    and moves to the next.  Returns (options dict, table of every measurement)."""
    own = solver is None
    s = solver or Solver(g.n, g.m)
    if own:
        s.load(g)
    max_deg = int(np.diff(np.asarray(g.ptr, np.int64)).max())
    phased_cap = g.n >= (1 << 20)            # a solver sized for the game holds the phased buffers
    if slices is None:
        slices = (4, 8, 16) if g.n > (1 << 20) else ((1, 2, 4, 8) if phased_cap else (1,))
    best = {k: 0 for k in _TUNED_KEYS}
    table = []

    def measure(cand):
        ms = _solve_ms(s, options(**cand), reps)
        table.append(dict(cand, solve_ms=ms))
        return ms

    try:
        cur = measure(dict(best))
        steps = [("slices", slices)]
        if max_deg > 32:
            steps.append(("big_deg", [d for d in big_degs if d < max_deg] or [0]))
        if max_deg > 1024:
            steps.append(("huge_deg", [d for d in huge_degs if d < max_deg] or [0]))
        for key, values in steps:
            for v in values:
                cand = dict(best, **{key: v})
                ms = measure(cand)
                if ms < cur:
                    cur, best = ms, cand
        if best["slices"] not in (0, 1) or (best["slices"] == 0 and g.n > (1 << 20)):
            for v in wgs_per_cu:
                cand = dict(best, phase_wgs_per_cu=v)
                ms = measure(cand)
                if ms < cur:
                    cur, best = ms, cand
        else:
            for G, grid in itertools.product(groups, grids):
                cand = dict(best, group=G, grid=grid)
                ms = measure(cand)
                if ms < cur:
                    cur, best = ms, cand
    finally:
        if own:
            s.close()
    return dict(best, solve_ms=cur), table


def stratified_sample(games, k, seed=0):
    """k games spread over the size distribution (sorted by n, every len/k-th)."""
    if len(games) <= k:
        return list(games)
    order = np.argsort([g.n for g in games], kind="stable")
    pick = np.linspace(0, len(games) - 1, k).round().astype(int)
    return [games[order[i]] for i in pick]


def tune_batch(games, sample=48, threads=(512, 1024), lds=(0, 20000), per_cu=(1, 2), lanes_x16=(1, 2, 4), device=0):
    sub = stratified_sample(games, sample)
    table = []
    with Batch(device) as b:
        for g in sub:
            b.add(g)
        b.upload()
        for th, ld, pc, lx in itertools.product(threads, lds, per_cu, lanes_x16):
            if pc > 1 and ld == 0:
                continue          # lds 0 = the whole LDS share: one workgroup per CU
            opt = batch_options(threads=th, lds_ints=ld, blocks_per_cu=pc, lanes_per_edge_x16=lx)
            try:
                st = b.solve(opt)
            except Exception:
                continue          # shape not launchable (LDS over the per-CU budget)
            table.append({"threads": th, "lds_ints": ld, "blocks_per_cu": pc, "lanes_per_edge_x16": lx,
                          "ms": st["ms"]})
    best = min(table, key=lambda t: t["ms"])
    return batch_options(threads=best["threads"], lds_ints=best["lds_ints"], blocks_per_cu=best["blocks_per_cu"],
                         lanes_per_edge_x16=best["lanes_per_edge_x16"]), table

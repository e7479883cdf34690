"""Autotuner for launch shapes, timed on the device through the C-ABI.

* Single game: lanes per low-degree vertex (1..16), workgroups of the round kernel and
  the dense-sweep threshold, timed over plain rounds of the game itself (proofs off, so
  every candidate does the same work).
* Many games: workgroup size, LDS share per workgroup (which decides which games keep
  their values in LDS), workgroups per CU and lanes per edge, timed by solving a
  size-stratified sample of the batch.
Each tuner returns the best options and the table it measured.
"""
from __future__ import annotations

import itertools

import numpy as np

from .batch import Batch, batch_options
from .solver import Solver, options


def tune_single(g, rounds=120, warmup=40, groups=(1, 2, 4, 8, 16), grids=(512, 1024, 2048),
                dense_pcts=(0,), solver=None):
    own = solver is None
    s = solver or Solver(g.n, g.m)
    if own:
        s.load(g)
    table = []
    try:
        for G, grid, dp in itertools.product(groups, grids, dense_pcts):
            opt = options(group=G, grid=grid, dense_pct=dp, certify_every=-1)
            s.prepare(opt)
            _, a = s.iterate(warmup)
            _, b = s.iterate(rounds)
            r = b.phases - a.phases
            if r <= 0:
                continue
            ms = (b.ms_iterate - a.ms_iterate) / r
            table.append({"group": G, "grid": grid, "dense_pct": dp, "ms_per_round": ms,
                          "edges_per_s": (b.edges_relaxed - a.edges_relaxed) / ((b.ms_iterate - a.ms_iterate) * 1e-3)})
    finally:
        if own:
            s.close()
    if not table:
        return {}, table
    best = min(table, key=lambda t: t["ms_per_round"])
    return {"group": best["group"], "grid": best["grid"], "dense_pct": best["dense_pct"]}, table


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

"""Benchmark of the energy-game solver on one MI355X.

Metric (one JSON line on stdout): edges relaxed per second to fixpoint on the single large
uniform game -- n = 2^22 vertices, owner uniform, out-degree exactly 8 distinct uniform
successors (m = 2^25), weights uniform in [-100, 100], seed 2.  The game is generated,
copied into HBM and loaded into the solver before the timed region.  One step is one full
solve to the least fixpoint: initial frontier, rounds and TOP proofs (the transpose is built
only when a round goes sparse; on this game none does).  `value` = edges relaxed by the
solves / their wall time.  After the timed solves the TOP set of the result is certified on
the device and checked on the host (untimed).

The line also carries
  roofline      the lift kernel on plain dense rounds of the same game (proofs off, so every
                round relaxes all m edges): algorithmic bytes per launch / average launch
                time (HIP events on the solver's stream) against the 8 TB/s HBM peak, the
                measured copy-kernel bandwidth, and the profiler's HBM bytes per launch
                from the committed counter summary of this workload.
  cpu_baseline  the reference's own CPU solver (cpu_solver.c compiled from its sources into
                oracle/_ref; the oracle's restatement if that build is absent) on this host:
                EG0 Jacobi sweeps on every physical core available (OMP_PLACES=cores,
                OMP_PROC_BIND=close) and EG, the worklist default, on one core, each on a
                bounded sample of the same game; nproc/lscpu/affinity recorded.
  skewed        the skewed-degree game (n = 2^20, power-law degrees): full solves and the
                plain dense round's roofline.
  batched       4096 games of the batched family (seed 4) in one launch vs one game per
                launch (single-game path, inputs resident) on a stratified sample.
  serving       a Poisson request stream of the serving family (seed 5) through the
                continuous-batching server vs one game per launch, p50/p99 latency.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "edges relaxed/s to fixpoint; solves/s + p99 latency, batched vs plain"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(d):
    """B = sum_{processed v} (16 + 12 d_out(v)) + sum_{risen v} (12 + 12 d_in(v)):
    frontier id 4 + row offsets 8 + own value 4; per out-edge successor 4 + weight 4 +
    gathered value 4; per risen vertex value store 4 + CSC offsets 8; per in-edge of a
    risen vertex predecessor id 4 + dedupe stamp read-modify-write 8."""
    return 16 * d["processed"] + 12 * d["edges_relaxed"] + 12 * d["lifted"] + 12 * d["pred_visits"]


def read_profile_traffic(workload_key):
    """HBM bytes per launch of the lift kernel from profiles/lift_<workload>_pmc.md."""
    path = os.path.join(ROOT, "profiles", f"lift_{workload_key}_pmc.md")
    if not os.path.exists(path):
        return None
    mo = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(mo.group(1)) if mo else None


def host_topology():
    """CPUs this process may run on, the physical cores among them, the cgroup CPU quota,
    and what nproc and lscpu report."""
    import subprocess
    aff = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in aff:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(per)
    except (OSError, ValueError):
        pass
    lscpu = {}
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                lscpu[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=20).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        nproc = None
    return {"nproc": nproc, "affinity_cpus": len(aff), "physical_cores_in_affinity": len(cores),
            "cgroup_cpu_quota": quota, "lscpu": lscpu}


def cpu_baseline(g, budget_s=15.0):
    """The reference's own CPU solver (src/cpu_solver.c compiled into oracle/_ref) on bounded
    samples of the same game, each in its own process so the OpenMP placement applies:
    EG0 Jacobi sweeps (src/cpu_solver.c:109-185) on every physical core this process may use
    (OMP_PLACES=cores OMP_PROC_BIND=close, as script/runner_omp:30-31), and EG, the
    reference's default worklist solver (src/cpu_solver.c:190-333), on one core."""
    import subprocess
    import tempfile
    import egsolver_amd as E
    topo = host_topology()
    threads = topo["physical_cores_in_affinity"]
    if topo["cgroup_cpu_quota"]:
        threads = max(1, min(threads, int(topo["cgroup_cpu_quota"])))
    script = os.path.join(ROOT, "oracle", "cpu_baseline.py")
    res = {}
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        path = os.path.join(d, "game.egs")
        E.save_csr(g, path)
        for algo, t, budget in (("eg0", threads, budget_s), ("eg", 1, budget_s / 2)):
            env = dict(os.environ, OMP_NUM_THREADS=str(t), OMP_PLACES="cores", OMP_PROC_BIND="close")
            r = subprocess.run([sys.executable, script, "--game", path, "--algo", algo, "--threads", str(t),
                                "--budget", str(budget)], capture_output=True, text=True, env=env, timeout=600)
            if r.returncode != 0:
                raise SystemExit(f"cpu baseline ({algo}) failed: {r.stderr[-2000:]}")
            res[algo] = json.loads(r.stdout.strip().splitlines()[-1])
            log(f"cpu baseline {algo}: {res[algo]['edges_per_s']:.3e} edges/s on {t} thread(s)")
    eg0, eg = res["eg0"], res["eg"]
    return {"value": eg0["edges_per_s"], "unit": "edges relaxed/s", "cores": threads, "kind": eg0["kind"],
            "sample": eg0["sample"] + f" (OMP_PLACES=cores OMP_PROC_BIND=close; {threads} physical cores)",
            "iteration_only_edges_per_s": eg0["edges_relaxed"] / eg0["iteration_s"],
            "eg_one_core": {"value": eg["edges_per_s"], "iteration_only_value": eg["edges_per_s_iteration_only"],
                            "unit": "edges relaxed/s", "cores": 1, "kind": eg["kind"], "sample": eg["sample"],
                            "transpose_s": eg["transpose_s"]},
            "host": topo}


def single_game(E, g, args):
    """Headline: full solves to fixpoint; plus plain dense rounds for the roofline."""
    s = E.Solver(g.n, g.m)
    s.load(g)
    shape = {"group": args.group, "grid": args.grid}
    if args.tune:
        from egsolver_amd.autotune import tune_single, _TUNED_KEYS
        best, table = tune_single(g, solver=s)
        shape = {k: best[k] for k in _TUNED_KEYS}
        log(f"tuned single-game options {best} over {len(table)} candidates")
    # plain rounds (no proofs): the kernel measurement for the roofline
    s.prepare(E.options(certify_every=-1, **shape))
    _, a = s.iterate(args.warmup)
    _, b = s.iterate(args.rounds)
    d = {k: getattr(b, k) - getattr(a, k) for k in ("processed", "edges_relaxed", "lifted", "pred_visits", "phases")}
    ms_dev = b.ms_iterate - a.ms_iterate
    rounds = d["phases"]
    per_launch_ms = ms_dev / max(rounds, 1)
    abytes = algorithmic_bytes(d) / max(rounds, 1)
    achieved = abytes / (per_launch_ms * 1e-3) / 1e9
    plain = {"rounds": rounds, "device_ms_per_round": per_launch_ms,
             "edges_per_s": d["edges_relaxed"] / (ms_dev * 1e-3), "algorithmic_bytes_per_round": abytes,
             "achieved_gbps": achieved}
    log(f"plain rounds: {per_launch_ms:.3f} ms/round, {plain['edges_per_s']:.3e} edges/s, {achieved:.0f} GB/s")
    # full solves to fixpoint (proofs on)
    opt = E.options(certify_every=args.certify_every, **shape)
    for _ in range(args.warmup_solves):
        s.prepare(opt)
        s.iterate()
    solves = []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        s.prepare(opt)
        conv, st = s.iterate()
        solves.append((time.perf_counter() - t0, st))
        if not conv:
            raise SystemExit("a solve did not reach the fixpoint")
    t_all = time.perf_counter() - t_all
    vals = s.values()
    violations = s.violations()
    if violations:
        raise SystemExit(f"fixpoint certificate failed: {violations} vertices")
    # untimed: the TOP set of the last solve certified on the device and checked on the host
    # (host fixpoint check plus the certificate's inequalities), independent of the proofs
    from egsolver_amd import verify
    t_c = time.perf_counter()
    cert = verify.certify(s, g, vals)
    cert["check_s"] = time.perf_counter() - t_c
    csc_ms = transpose_ms(E, s)
    s.close()
    if not cert["ok"]:
        raise SystemExit(f"TOP-set certificate failed: {cert}")
    log(f"certificate of the TOP set: {cert['top_vertices']} vertices, {cert['passes']} passes, host check ok")
    edges = sum(st.edges_relaxed for _, st in solves)
    last = solves[-1][1]
    info = {"solve_ms_mean": 1e3 * t_all / args.steps, "solve_ms_min": 1e3 * min(t for t, _ in solves),
            "rounds_per_solve": last.phases, "proofs": last.certificates, "proof_passes": last.cert_passes,
            "top_marked": last.top_marked, "top_vertices": int((vals == 2147483647).sum()),
            "value_checksum": int(np.bitwise_xor.reduce(vals.astype(np.uint32) * np.uint32(2654435761))),
            "fixpoint_violations": violations, "device_ms_prepare": last.ms_prepare,
            "device_ms_iterate": last.ms_iterate, "transpose_built": bool(last.csc_built),
            "transpose_forced_ms": csc_ms, "top_set_certificate": cert}
    return edges / t_all, t_all, info, plain, shape


def skewed_game(E, args):
    """The skewed-degree game (n = 2^20, power-law out-degrees 1..10^5, exponent 2.1, weights
    U[-100,100], seed 3): full solves to the least fixpoint, certified like the headline game,
    plus the plain dense round (proofs off) of the same game for its roofline."""
    g = E.config_game(3, scale=args.scale)
    s = E.Solver(g.n, g.m)
    s.load(g)
    shape, tuned = {}, None
    if args.tune:
        from egsolver_amd.autotune import tune_single, _TUNED_KEYS
        tuned, table = tune_single(g, solver=s)
        shape = {k: tuned[k] for k in _TUNED_KEYS}
        log(f"skewed game: tuned options {tuned} over {len(table)} candidates")
    s.prepare(E.options(certify_every=-1, **shape))
    _, a = s.iterate(args.warmup)
    _, b = s.iterate(args.rounds // 4)
    d = {k: getattr(b, k) - getattr(a, k) for k in ("processed", "edges_relaxed", "lifted", "pred_visits", "phases")}
    ms_dev = b.ms_iterate - a.ms_iterate
    per_round = ms_dev / max(d["phases"], 1)
    abytes = algorithmic_bytes(d) / max(d["phases"], 1)
    for _ in range(args.warmup_solves):
        s.prepare(E.options(**shape))
        s.iterate()
    times, st = [], None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        s.prepare(E.options(**shape))
        conv, st = s.iterate()
        times.append(time.perf_counter() - t0)
        if not conv:
            raise SystemExit("skewed game: a solve did not reach the fixpoint")
    vals = s.values()
    from egsolver_amd import verify
    cert = verify.certify(s, g, vals)
    csc_ms = transpose_ms(E, s)
    s.close()
    if not cert["ok"]:
        raise SystemExit(f"skewed game: TOP-set certificate failed: {cert}")
    out = {"workload": "skewed-degree game (n=2^20, power-law out-degrees 1..1e5 exponent 2.1, weights U[-100,100], "
                       "seed 3); step = one full solve to the least fixpoint",
           "n": g.n, "m": g.m, "max_outdeg": int(np.diff(g.ptr).max()), "mg": g.mg,
           "solve_ms_mean": 1e3 * sum(times) / len(times), "solve_ms_min": 1e3 * min(times),
           "edges_relaxed_per_s": st.edges_relaxed * len(times) / sum(times), "rounds_per_solve": st.phases,
           "proofs": st.certificates, "transpose_built": bool(st.csc_built), "transpose_forced_ms": csc_ms,
           "device_ms_prepare": st.ms_prepare,
           "device_ms_iterate": st.ms_iterate, "top_vertices": cert["top_vertices"],
           "top_set_certificate_ok": cert["ok"], "options": tuned,
           "plain_rounds": {"rounds": d["phases"], "device_ms_per_round": per_round,
                            "algorithmic_bytes_per_round": abytes,
                            "edges_per_s": d["edges_relaxed"] / (ms_dev * 1e-3),
                            "achieved_gbps": abytes / (per_round * 1e-3) / 1e9,
                            "frac_of_hbm_peak": abytes / (per_round * 1e-3) / 1e9 / HBM_PEAK_GBPS}}
    log(f"skewed game: m {g.m}, solve {out['solve_ms_mean']:.2f} ms, plain round {per_round:.3f} ms "
        f"({out['plain_rounds']['achieved_gbps']:.0f} GB/s)")
    return out


def transpose_ms(E, s, reps=3):
    """Device time of the transpose alone: prepares that build it up front (every round
    sparse-capable, dense_pct > 100) less prepares that do not, medians of `reps`."""
    def prep(opt):
        ts = []
        for _ in range(reps):
            s.prepare(opt)
            ts.append(s.stats().ms_prepare)
        return float(np.median(ts))
    return prep(E.options(dense_pct=101)) - prep(E.options())


def gen_many(E, family, seed, count, scale_n=1.0, workers=16):
    def one(i):
        p = E.family_params(family, seed, i)
        if scale_n != 1.0 and p.deg_kind != 3:
            p.n = max(50, int(p.n * scale_n))
        return E.generate(p, stream=i)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(count)))


def plain_solve_sample(E, games, check_against=None):
    """One game per launch (single-game path, proofs on): per game load (untimed), then
    prepare + iterate + read back (timed)."""
    nmax = max(g.n for g in games)
    mmax = max(g.m for g in games)
    s = E.Solver(nmax, mmax)
    t = 0.0
    try:
        for i, g in enumerate(games):
            s.load(g)
            t0 = time.perf_counter()
            s.prepare()
            s.iterate()
            v = s.values()
            t += time.perf_counter() - t0
            if check_against is not None and not np.array_equal(v, check_against[i]):
                raise SystemExit(f"plain and batched results differ on game {i}")
    finally:
        s.close()
    return t


def worklist_estimate(games, lifted):
    nv = sum(g.n for g in games)
    est = 0.0
    for g in games:
        deg = np.diff(np.asarray(g.ptr, np.int64)).astype(np.float64)
        est += g.m + lifted * (g.n / nv) * float((deg * deg).sum()) / g.n
    return est


def batched(E, args):
    from egsolver_amd.autotune import tune_batch
    t0 = time.perf_counter()
    games = gen_many(E, E.games.FAMILY_BATCH, 4, args.batch_games)
    nv = sum(g.n for g in games)
    ne = sum(g.m for g in games)
    log(f"batched: {len(games)} games, {nv} vertices, {ne} edges generated in {time.perf_counter() - t0:.1f}s")
    opt = E.batch_options()
    tuned = None
    if args.tune:
        opt, table = tune_batch(games, sample=48)
        tuned = {k: getattr(opt, k) for k in ("threads", "lds_ints", "blocks_per_cu", "lanes_per_edge_x16")}
        log(f"tuned batch options {tuned} over {len(table)} candidates")
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        b.solve(opt)                      # warm-up launch
        walls = []
        for _ in range(3):                # median of three timed solves (one is host-noise prone)
            t1 = time.perf_counter()
            st = b.solve(opt)
            vals = b.values_view()        # every value on the host (pinned, owned by the batch)
            walls.append(time.perf_counter() - t1)
        wall = float(np.median(walls))
        vals = [v.copy() for v in vals]   # outlive the batch (untimed)
    if st["converged"] != len(games):
        raise SystemExit("batched: not every game converged")
    # plain sample: spread over the size distribution (sorted by n, every len/k-th)
    order = np.argsort([g.n for g in games], kind="stable")
    idx = order[np.linspace(0, len(games) - 1, min(args.plain_games, len(games))).round().astype(int)]
    sub = [games[i] for i in idx]
    t_plain = plain_solve_sample(E, sub, check_against=[vals[i] for i in idx])
    plain_rate = len(sub) / t_plain
    out = {"workload": f"{len(games)} games of the batched family (n log-uniform 1e3..5e4, out-degree 2..16, "
                       f"W log-uniform 10..1000, every 16th a complete digraph of 200..500), seed 4",
           "vertices": nv, "edges": ne, "solves_per_s": len(games) / wall, "batch_wall_s": wall,
           "batch_device_ms": st["ms"], "sweeps": st["sweeps"], "proofs": st["certificates"],
           "edges_relaxed": st["edges_relaxed"], "lifted": st["lifted"],
           "edges_relaxed_per_s": st["edges_relaxed"] / (st["ms"] * 1e-3),
           # a worklist (re-evaluate the predecessors of every vertex that rose) relaxes every
           # edge once, then, per lift, the out-edges of the risen vertex's predecessors:
           # sum over games of m + lifts * sum_v deg(v)^2 / n with uniform successors; the
           # lifts are split over games in proportion to their vertices
           "worklist_edges_estimate": worklist_estimate(games, st["lifted"]),
           "plain_solves_per_s": plain_rate, "plain_sample_games": len(sub),
           "speedup_vs_plain": (len(games) / wall) / plain_rate, "options": tuned,
           "exact_vs_plain_on_sample": True}
    log(f"batched: {out['solves_per_s']:.1f} solves/s vs plain {plain_rate:.1f} solves/s")
    return out, games


def serving(E, args):
    """Continuous batching against one game per launch on the same Poisson stream."""
    from egsolver_amd.serve import Request, latency_stats, run_plain, run_served, serve_options
    R = args.serve_requests
    t0 = time.perf_counter()
    games = gen_many(E, E.games.FAMILY_SERVE, 5, R)
    log(f"serving: {R} requests generated in {time.perf_counter() - t0:.1f}s")
    arena = int(args.serve_arena_gb * (1 << 30))
    opt = serve_options(slice_kedges=args.slice_kedges)
    nmax = max(g.n for g in games)
    mmax = max(g.m for g in games)
    solver = E.Solver(nmax, mmax)

    def one(g):
        solver.load(g)
        solver.prepare()
        solver.iterate()
        return solver.values()
    # closed-loop capacities (every request present at t = 0) on the first K requests
    K = min(R, args.serve_capacity_requests)
    done, wall, _ = run_served([Request(rid=i, arrival=0.0, game=games[i]) for i in range(K)], arena, opt,
                               large_n=args.serve_large_n)
    cap = K / wall
    P = min(R, args.serve_plain_requests)
    _, pwall = run_plain([Request(rid=i, arrival=0.0, game=games[i]) for i in range(P)], one)
    plain_cap = P / pwall
    log(f"serving: closed-loop capacity served {cap:.1f}/s, plain {plain_cap:.1f}/s")
    rate = args.serve_load * cap
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=R))
    arrivals -= arrivals[0]
    reqs = [Request(rid=i, arrival=float(arrivals[i]), game=games[i]) for i in range(R)]
    done, wall, slices = run_served(reqs, arena, opt, large_n=args.serve_large_n)
    served = latency_stats(done)
    served.update({"solves_per_s": R / wall, "wall_s": wall, "slices": slices, "closed_loop_solves_per_s": cap})
    # the same stream, one game per launch (first P requests: at this rate the plain path
    # falls behind and its queue grows)
    preqs = [Request(rid=i, arrival=float(arrivals[i]), game=games[i]) for i in range(P)]
    pdone, pwall = run_plain(preqs, one)
    solver.close()
    plain = latency_stats(pdone)
    plain.update({"solves_per_s": P / pwall, "requests": P, "closed_loop_solves_per_s": plain_cap})
    log(f"serving: served p50 {served['p50_ms']:.1f} ms p99 {served['p99_ms']:.1f} ms {served['solves_per_s']:.1f}/s; "
        f"plain p50 {plain['p50_ms']:.1f} p99 {plain['p99_ms']:.1f} {plain['solves_per_s']:.1f}/s")
    return {"workload": f"{R} requests of the serving family (n log-uniform 1e3..5e5, otherwise as batched), "
                        f"Poisson arrivals at {rate:.1f}/s = {args.serve_load:.2f} x the served closed-loop "
                        f"capacity, seed 5",
            "served": served, "plain_one_game_per_launch": plain, "arrival_rate_per_s": rate,
            "capacity_speedup_vs_plain": cap / plain_cap, "arena_gb": args.serve_arena_gb}


def ingest(E, args):
    """Host ingest of a game given as .mpg text: the sequential reader, the parallel one,
    the owner partition, and the binary game cache (save and load).  The game is config 2
    at a quarter of its vertices, written to a temporary file first."""
    import ctypes
    import tempfile
    from egsolver_amd import games as G
    from egsolver_amd import _native as N
    p = E.GenParams(n=max(16, int((1 << 22) * args.ingest_scale)), seed=2, deg_kind=0, deg_lo=8, deg_hi=8, wmax=100)
    owner, rptr, rsucc, rw = G.generate_read_order(p)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        path = os.path.join(d, "game.mpg")
        E.write_mpg(path, owner, rptr, rsucc, rw)
        with open(path, "rb") as f:
            text = f.read()
        m = int(rptr[-1])

        def best_of(fn, k=3):
            ts = []
            for _ in range(k):
                t = time.perf_counter()
                r = fn()
                ts.append(time.perf_counter() - t)
            return min(ts), r

        def parse(threads):
            gm = N.EgsMpg()
            err = ctypes.create_string_buffer(512)
            fn = N.host().egs_mpg_read_mem if threads == 1 else N.host().egs_mpg_read_mem_par
            extra = () if threads == 1 else (threads,)
            if fn(text, len(text), 0, -1, 1, ctypes.byref(gm), err, 512, *extra) != 0:
                raise SystemExit(err.value.decode())
            chunks = gm.chunks
            N.host().egs_mpg_free(ctypes.byref(gm))
            return chunks

        t_seq, _ = best_of(lambda: parse(1))
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        t_par, chunks = best_of(lambda: parse(threads))
        t_file, g = best_of(lambda: E.read_mpg(path))          # parallel read + partition + MG
        cache = os.path.join(d, "game.egs")
        t_save, _ = best_of(lambda: E.save_csr(g, cache))
        t_load, h = best_of(lambda: E.load_csr(cache))
        assert np.array_equal(h.succ, g.succ) and np.array_equal(h.ptr, g.ptr)
        cache_bytes = os.path.getsize(cache)
    out = {"workload": f"config-2 distribution at n={p.n} (m={m}) as .mpg text, {len(text) / 1e6:.0f} MB",
           "text_mb": len(text) / 1e6, "edges": m,
           "sequential_read_s": t_seq, "parallel_read_s": t_par, "parallel_chunks": chunks,
           "read_mb_per_s": {"sequential": len(text) / t_seq / 1e6, "parallel": len(text) / t_par / 1e6},
           "file_to_game_s": t_file, "cache_save_s": t_save, "cache_load_s": t_load,
           "cache_gb_per_s": cache_bytes / t_load / 1e9,
           "edges_per_s": {"sequential_read": m / t_seq, "parallel_read": m / t_par, "cache_load": m / t_load}}
    log(f"ingest: {len(text) / 1e6:.0f} MB text, read {t_seq:.2f}s sequential / {t_par:.2f}s in {chunks} chunks, "
        f"file->game {t_file:.2f}s, cache load {t_load:.3f}s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed full solves of the large game")
    ap.add_argument("--warmup", type=int, default=50, help="untimed plain rounds before the roofline rounds")
    ap.add_argument("--warmup-solves", type=int, default=1)
    ap.add_argument("--rounds", type=int, default=600, help="plain rounds timed for the roofline")
    ap.add_argument("--scale", type=float, default=1.0, help="shrink n (testing only)")
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--tune", action="store_true", help="autotune launch shapes before timing")
    ap.add_argument("--certify-every", type=int, default=0, help="rounds between proof attempts (0 = auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-skewed", action="store_true", help="skip the skewed-degree game (n=2^20) leg")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-games", type=int, default=4096)
    ap.add_argument("--plain-games", type=int, default=128)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--serve-capacity-requests", type=int, default=500)
    ap.add_argument("--serve-plain-requests", type=int, default=300)
    ap.add_argument("--serve-load", type=float, default=0.75)
    ap.add_argument("--serve-arena-gb", type=float, default=24.0)
    ap.add_argument("--serve-large-n", type=int, default=0, help="route threshold of the server (0 = its default)")
    ap.add_argument("--slice-kedges", type=int, default=0)
    ap.add_argument("--no-ingest", action="store_true")
    ap.add_argument("--ingest-scale", type=float, default=0.25, help="share of config 2's vertices for the ingest leg")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(how the work would shard across GPUs is described in the design notes).", file=sys.stderr)
        sys.exit(2)

    import egsolver_amd as E

    t0 = time.time()
    g = E.config_game(2, scale=args.scale)
    log(f"game n={g.n} m={g.m} mg={g.mg} generated in {time.time() - t0:.1f}s")
    copy_gbps = _copy_bw()
    value, t_all, info, plain, shape = single_game(E, g, args)
    traffic = read_profile_traffic("config2")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "edges relaxed/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_all / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": "single large uniform game (n=2^22, out-degree 8, weights U[-100,100], seed 2); "
                        "step = one full solve to the least fixpoint (initial frontier, rounds, TOP proofs; "
                        "the transpose is built only if a round goes sparse)",
            "n": g.n, "m": g.m, "mg": g.mg, "shape": shape,
            "warmup_note": "warmup = untimed plain rounds before the roofline rounds; "
                           f"{args.warmup_solves} untimed full solve(s) precede the timed ones",
            "warmup_solves": args.warmup_solves, **info,
            "plain_rounds": plain,
        },
        "roofline": {
            "bound": "hbm", "achieved": plain["achieved_gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": plain["achieved_gbps"] / HBM_PEAK_GBPS, "traffic": traffic,
            "kernel": "dense round: k_round (mode decision) + k_dense_fused (8 successor-range phases in one launch)",
            "algorithmic_bytes_per_launch": plain["algorithmic_bytes_per_round"],
            "avg_launch_ms": plain["device_ms_per_round"],
            "copy_kernel_gbps": copy_gbps,
        },
    }
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline(g, args.cpu_budget)
    del g
    if not args.no_skewed:
        line["skewed"] = skewed_game(E, args)
    if not args.no_batch:
        line["batched"], _ = batched(E, args)
    if not args.no_serve:
        line["serving"] = serving(E, args)
    if not args.no_ingest:
        line["ingest"] = ingest(E, args)
    print(json.dumps(line), flush=True)


def _copy_bw():
    import ctypes
    from egsolver_amd import _native as N
    out = ctypes.c_double(0)
    rc = N.egs().egs_copy_bandwidth(0, 1 << 30, 20, ctypes.byref(out))
    return out.value if rc == 0 else None


if __name__ == "__main__":
    main()

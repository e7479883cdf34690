"""GPU parity of the single-game path (libegs.so) against the oracle and the goldens.

Bar: bit-exact values.  Every case prints its name first so a failing GPU run names it.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, exact_by_certificate, lfp, load_golden_sol, to_oracle_game

pytestmark = pytest.mark.gpu

TOP = 2147483647


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _solve(g, **opt):
    import egsolver_amd as E
    return E.solve(g, **opt)


def test_eg_fig2_family(oracle):
    import egsolver_amd as E
    # Hand solution of script/multi/eg_fig2.txt: V(0)=0, V(1)=3, V(2)=TOP, V(3)=0, V(4)=TOP;
    # the k-copy files repeat it per copy (copy c of vertex i has id i + 5c).
    hand = {0: 0, 1: 3, 2: TOP, 3: 0, 4: TOP}
    for name in ["eg_fig2.txt", "eg_fig2_duecopie.txt", "eg_fig2_quattrocopie.txt",
                 "eg_fig2_duecopie_di_duecopie.txt"]:
        _say(name)
        g = E.read_mpg(os.path.join(GOLDEN, name))
        vals, st = _solve(g)
        assert st["converged"]
        got = dict(zip(g.ext_of_int.tolist(), vals.tolist()))
        for ext, v in got.items():
            assert v == hand[ext % 5], (name, ext, v)
        ref = lfp(oracle, to_oracle_game(g))
        assert np.array_equal(vals, ref)


def test_clique_500_cp4_golden():
    """script/sol.txt: the reference's only recorded run, all 2000 values."""
    import egsolver_amd as E
    _say("clique_500_cp4")
    g = E.read_mpg(os.path.join(GOLDEN, "clique_500_cp4.txt.gz"))
    assert (g.n, g.m, g.mg) == (2000, 1746500, 1992)
    gold = load_golden_sol()
    for group in (1, 8, 64):
        vals, st = _solve(g, group=group)
        assert st["converged"]
        got = dict(zip(g.ext_of_int.tolist(), vals.tolist()))
        assert got == gold, f"group={group}"


def test_config1_vs_oracle(oracle):
    import egsolver_amd as E
    _say("config1 n=2000 deg=4 W=10 seed=1")
    g = E.config_game(1)
    ref = lfp(oracle, to_oracle_game(g))
    for group in (1, 2, 4, 8, 16, 32, 64):
        for dense_pct in (0, 1, 101):
            vals, st = _solve(g, group=group, dense_pct=dense_pct)
            assert st["converged"]
            assert np.array_equal(vals, ref), f"group={group} dense_pct={dense_pct}: {np.count_nonzero(vals != ref)} differ"


@pytest.mark.parametrize("seed", range(6))
def test_random_small_all_paths(oracle, seed):
    """Degrees up to 300 with low thresholds so the lane, wavefront and workgroup
    kernels all run in one round."""
    import egsolver_amd as E
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 600))
    kind = seed % 3
    p = E.GenParams(n=n, seed=100 + seed, deg_kind=[1, 2, 1][kind], deg_lo=1, deg_hi=min(n - 1, [6, 300, 40][kind]),
                    alpha=1.6, wmax=int(rng.integers(1, 50)))
    g = E.generate(p)
    _say(f"random seed={seed} n={g.n} m={g.m} maxdeg={g.max_outdeg}")
    ref = lfp(oracle, to_oracle_game(g))
    for group, big, huge in [(1, 4, 16), (4, 8, 64), (8, 16, 32), (2, 0, 0)]:
        for dense_pct in (0, 1, 30, 101):
            vals, st = _solve(g, group=group, big_deg=big, huge_deg=huge, dense_pct=dense_pct)
            assert st["converged"]
            assert np.array_equal(vals, ref), (group, big, huge, dense_pct, np.count_nonzero(vals != ref))


def _long_rows_game(n=50000, seed=9):
    """Short rows around rows longer than 4,096 edges: adjacent long rows, rows of exactly
    4,096 and 4,097 edges, long rows across 4,096-edge chunk and 16,384-edge tile
    boundaries, one spanning several tiles, and a long last row."""
    import egsolver_amd as E
    rng = np.random.default_rng(seed)
    deg = rng.integers(1, 9, size=n)
    for i, d in [(100, 4096), (101, 4097), (102, 16384), (103, 16385), (104, 3), (105, 4097),
                 (5000, 40000), (5001, 4097), (20000, 70000), (n - 1, 5000)]:
        deg[i] = d
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum(deg)
    m = int(ptr[-1])
    succ = rng.integers(0, n, size=m).astype(np.int32)
    w = rng.integers(-9, 10, size=m).astype(np.int32)
    return E.Game(n=n, n0=n // 2, ptr=ptr, succ=succ, w=w, mg=E.games.compute_mg(n, ptr, w),
                  ext_of_int=np.arange(n, dtype=np.int32))


def test_transpose_matches_oracle(oracle):
    """The device transpose equals the reference's host transpose (src/host_csr2csc.c:15-46,
    restated in the oracle) entry for entry: cptr, and within each column the sources in
    ascending order.  Shapes: a skewed game, games with several destination buckets and edge
    tiles, a vertex count that is not a multiple of the bucket width, complete digraphs
    (every vertex a predecessor of every other; one bucket too large for the LDS placement),
    a single vertex, a ladder, the full-size skewed game (n = 2^20, m = 4,962,742), the
    full-size uniform game (n = 2^22, m = 2^25) and rows longer than 4,096 edges at chunk and
    tile boundaries (their sources come from the per-tile kernel)."""
    import egsolver_amd as E
    games = [E.config_game(3, scale=1 / 256), E.config_game(2, scale=1 / 64), E.config_game(3), E.config_game(2),
             E.generate(E.GenParams(n=9001, seed=5, deg_kind=1, deg_lo=1, deg_hi=40, wmax=9)),
             E.generate(E.GenParams(n=300, seed=6, deg_kind=3, deg_lo=1, deg_hi=1, wmax=5)),
             E.generate(E.GenParams(n=2, seed=7, deg_kind=0, deg_lo=1, deg_hi=1, wmax=5)),
             E.generate(E.GenParams(n=70000, seed=8, deg_kind=2, deg_lo=1, deg_hi=20000, alpha=1.6, wmax=50)),
             _long_rows_game()]
    for g in games:
        _say(f"transpose n={g.n} m={g.m}")
        cptr_ref, cpred_ref = oracle.csr2csc(to_oracle_game(g))
        with E.Solver(g.n, g.m) as s:
            s.load(g)
            s.prepare()
            cptr, cpred = s.csc()
        assert np.array_equal(cptr, cptr_ref)
        assert np.array_equal(cpred, cpred_ref), int((cpred != cpred_ref).sum())


@pytest.mark.parametrize("which,scale", [(2, 1 / 512), (3, 1 / 64)])
def test_scaled_configs_vs_oracle(oracle, which, scale):
    import egsolver_amd as E
    g = E.config_game(which, scale=scale)
    _say(f"config{which} scaled n={g.n} m={g.m}")
    ref = lfp(oracle, to_oracle_game(g))
    for dense_pct in (0, 101):
        vals, st = _solve(g, dense_pct=dense_pct)
        assert st["converged"]
        assert np.array_equal(vals, ref), (dense_pct, np.count_nonzero(vals != ref))
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        s.prepare()
        s.iterate()
        assert s.violations() == 0


def test_partial_run_is_under_approximation(oracle):
    """A phase cap stops early; values stay pointwise <= the fixpoint (the reference's
    --eg N / --timeout semantics) and resuming reaches the fixpoint."""
    import egsolver_amd as E
    g = E.config_game(2, scale=1 / 1024)
    _say(f"partial n={g.n}")
    ref = lfp(oracle, to_oracle_game(g))
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        s.prepare()
        conv, st = s.iterate(max_phases=5)
        assert not conv and st.phases == 5
        part = s.values()
        assert np.all(part <= ref)
        conv, st = s.iterate()
        assert conv
        assert np.array_equal(s.values(), ref)


@pytest.mark.parametrize("seed", range(4))
def test_phased_dense_rounds_vs_oracle(oracle, seed):
    """Dense rounds in successor-range phases (forced on small games through a solver sized
    for them), with proofs on and off and with every round dense: exact."""
    import egsolver_amd as E
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(300, 6000))
    kind = seed % 3
    if kind == 2:
        p = E.GenParams(n=n, seed=700 + seed, deg_kind=2, deg_lo=1, deg_hi=min(n - 1, 3000), alpha=1.8, wmax=50)
    else:
        p = E.GenParams(n=n, seed=700 + seed, deg_kind=kind, deg_lo=2, deg_hi=12, wmax=int(rng.integers(5, 200)))
    g = E.generate(p)
    _say(f"phased seed {seed} n {g.n} m {g.m}")
    ref = lfp(oracle, to_oracle_game(g))
    with E.Solver(1 << 20, max(g.m, 1 << 22)) as s:
        s.load(g)
        for opt in (E.options(slices=4), E.options(slices=4, certify_every=-1), E.options(slices=4, dense_pct=1),
                    E.options(slices=4, group=1, big_deg=4), E.options(slices=16), E.options(slices=3, certify_every=2),
                    E.options(slices=8, certify_every=2, big_deg=300)):
            s.prepare(opt)
            conv, st = s.iterate()
            assert conv
            assert np.array_equal(s.values(), ref), opt
            assert s.violations() == 0


def test_pruned_proofs_and_hubs_vs_oracle(oracle):
    """Power-law games with hub vertices, proofs attempted every 4 rounds: candidates that fail
    the closure check are pruned to their greatest closed subset before and during the
    longest-path passes, and vertices above big_deg go to the workgroup-per-vertex kernels of
    the rounds, the closure and the passes (forced down to degree 8 / 32 in one variant).
    Exact against the oracle, and pruning must have happened."""
    import egsolver_amd as E
    pruned = certs = 0
    for seed in range(6):
        n = 3000 + 2500 * seed
        p = E.GenParams(n=n, seed=900 + seed, deg_kind=2, deg_lo=1, deg_hi=n - 1, alpha=2.1, wmax=100)
        g = E.generate(p)
        _say(f"pruned proofs seed {seed} n {g.n} m {g.m} maxdeg {g.max_outdeg}")
        ref = lfp(oracle, to_oracle_game(g))
        with E.Solver(g.n, g.m) as s:
            s.load(g)
            for opt in (E.options(certify_every=4), E.options(certify_every=4, big_deg=8, huge_deg=32),
                        E.options(certify_every=12, dense_pct=101)):
                s.prepare(opt)
                conv, st = s.iterate()
                assert conv
                assert np.array_equal(s.values(), ref), (seed, opt.big_deg, opt.certify_every)
                assert s.violations() == 0
                pruned += st.cert_pruned
                certs += st.certificates
    _say(f"pruned proofs: {certs} certificates, {pruned} pruned candidates")
    assert certs > 0 and pruned > 0, (certs, pruned)


def test_split_hub_lifts_vs_oracle(oracle):
    """Vertices above huge_deg with thousands of edges are lifted by several workgroups at
    once (chunks of 2048 edges meeting in a per-vertex accumulator), in dense and in sparse
    rounds: exact against the oracle."""
    import egsolver_amd as E
    for seed in range(3):
        g = E.generate(E.GenParams(n=6000, seed=500 + seed, deg_kind=2, deg_lo=1, deg_hi=5999, alpha=1.5, wmax=60))
        deg = np.diff(g.ptr)
        _say(f"split hub lifts seed {seed} n {g.n} m {g.m} maxdeg {deg.max()} above 2048: {(deg > 2048).sum()}")
        assert (deg > 2048).sum() >= 2             # at least two chunks each
        ref = lfp(oracle, to_oracle_game(g))
        with E.Solver(g.n, g.m) as s:
            s.load(g)
            for opt in (E.options(huge_deg=256), E.options(huge_deg=256, dense_pct=1), E.options(huge_deg=256, certify_every=-1),
                        E.options(huge_deg=1024, dense_pct=101)):
                s.prepare(opt)
                conv, st = s.iterate()
                assert conv
                assert np.array_equal(s.values(), ref), (seed, opt.huge_deg, opt.dense_pct)


def test_violations_count_with_hubs(oracle):
    """The device fixpoint check (wide vertices shared by the wave) counts the same
    violations as the oracle on the partial values of a power-law game."""
    import egsolver_amd as E
    g = E.generate(E.GenParams(n=20000, seed=77, deg_kind=2, deg_lo=1, deg_hi=19999, alpha=1.9, wmax=60))
    _say(f"violations with hubs n {g.n} m {g.m} maxdeg {g.max_outdeg}")
    og = to_oracle_game(g)
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        s.prepare(E.options(certify_every=-1))
        for k in (1, 3, 9):
            s.iterate(max_phases=k)
            part = s.values()
            want = oracle.fixpoint_violations(og, part)
            assert want > 0
            assert s.violations() == want, k


def test_winning_set_bitmap(oracle):
    """The device winning-set bitmap is {v : value < TOP} of the exact values, for sizes that
    are and are not multiples of the 32-bit word."""
    import egsolver_amd as E
    for n, seed in ((2000, 1), (2017, 2), (64, 3), (33, 4), (5003, 5)):
        g = E.generate(E.GenParams(n=n, seed=seed, deg_kind=1, deg_lo=1, deg_hi=6, wmax=20))
        _say(f"winning set n {g.n}")
        ref = lfp(oracle, to_oracle_game(g))
        with E.Solver(g.n, g.m) as s:
            s.load(g)
            s.prepare()
            s.iterate()
            assert np.array_equal(s.values(), ref)
            assert np.array_equal(s.winning(), ref != TOP)


def test_config2_full_size_exact(oracle):
    """Config 2 at full size (n = 2^22, m = 2^25), phased and single-sweep dense rounds.
    No oracle reaches this fixpoint from 0 (about 10^6 sweeps), so the result is pinned
    without trusting the device's TOP proofs: host fixpoint check, a device certificate of
    the whole TOP set checked on the host, then the oracle's EG0 from the certified TOP set
    must give the same values bit for bit."""
    import egsolver_amd as E
    g = E.config_game(2)
    _say(f"config 2 full size, phased vs single sweep, n {g.n} m {g.m}")
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        out = []
        for sl in (1, 4):
            s.prepare(E.options(slices=sl))
            conv, st = s.iterate()
            assert conv
            assert st.certificates >= 1
            assert s.violations() == 0
            out.append(s.values())
        assert np.array_equal(out[0], out[1])
        rep = exact_by_certificate(oracle, g, out[1], s)
        _say(f"config 2 certified: {rep}")
        assert rep["top_vertices"] > g.n // 3


def test_phased_unpacked_edges_equal_single_sweep():
    """Weights too wide to share a 32-bit word with a 17-bit successor id keep the slice
    edges as int2; phased and single-sweep rounds end on the same fixpoint.  Proofs run in
    both (every 4 rounds), so the phased closure check over int2 slice edges is compared
    with the one-sweep closure check."""
    import egsolver_amd as E
    g = E.generate(E.GenParams(n=1 << 17, seed=91, deg_kind=1, deg_lo=2, deg_hi=6, wmax=20000))
    _say(f"phased unpacked n {g.n} m {g.m} mg {g.mg}")
    with E.Solver(1 << 20, 1 << 22) as s:
        s.load(g)
        out = []
        for sl in (1, 4):
            s.prepare(E.options(slices=sl, certify_every=4))
            conv, st = s.iterate()
            assert conv and s.violations() == 0
            assert st.cert_attempts > 0 and st.cert_passes > 0, (sl, st.cert_attempts, st.cert_passes)
            out.append(s.values())
        assert np.array_equal(out[0], out[1])


def test_config3_full_size_exact(oracle):
    """The skewed-degree game at full size (n = 2^20, power-law out-degrees up to 10^5):
    phased and single-sweep dense rounds agree, and the result is pinned as for config 2
    (host fixpoint check, certificate of the TOP set checked on the host, oracle EG0 from the
    certified TOP set)."""
    import egsolver_amd as E
    g = E.config_game(3)
    _say(f"config 3 full size n {g.n} m {g.m}")
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        out = []
        for sl in (0, 1):
            s.prepare(E.options(slices=sl))
            conv, st = s.iterate()
            assert conv and s.violations() == 0
            assert st.certificates >= 1
            out.append(s.values())
        assert np.array_equal(out[0], out[1])
        rep = exact_by_certificate(oracle, g, out[0], s)
        _say(f"config 3 certified: {rep}")


def test_top_certificate_small_games(oracle):
    """The certificate of the TOP set on games the oracle solves from 0: found and accepted
    by the host check on the exact values; on values with a finite vertex raised to TOP it
    is not found and the host check rejects whatever potentials come back."""
    import egsolver_amd as E
    from egsolver_amd import verify
    rng = np.random.default_rng(7)
    for i, (n, lo, hi, wm) in enumerate([(2000, 4, 4, 10), (3000, 1, 6, 20), (500, 2, 3, 5), (5000, 1, 12, 100),
                                         (20000, 2, 16, 300)]):
        g = E.generate(E.GenParams(n=n, seed=100 + i, deg_kind=1 if lo != hi else 0, deg_lo=lo, deg_hi=hi, wmax=wm))
        _say(f"certificate small game {i} n {g.n} m {g.m}")
        ref = lfp(oracle, to_oracle_game(g))
        with E.Solver(g.n, g.m) as s:
            s.load(g)
            s.prepare()
            conv, st = s.iterate()
            vals = s.values()
            assert conv and np.array_equal(vals, ref)
            rep = verify.certify(s, g, vals)
            assert rep["ok"], rep
            finite = np.nonzero(vals != TOP)[0]
            if len(finite):
                bad = vals.copy()
                bad[rng.choice(finite)] = TOP
                s.set_values(bad)
                found, h, _ = s.top_certificate(max_passes=4096)
                code, _ = verify.check_top_certificate(g, bad, h)
                assert not found and code != 0


def test_stop_flag_from_another_thread_then_resume(tmp_path):
    """The stop flag (the reference's timeout_expired, polled once per round at
    src/dev_EG_alg.cu:376) set by another host thread 0.5 s into a plain value iteration
    (proofs off) of a ladder of 300,000 rungs (about 300,000 rounds): the run stops within a
    few rounds with values pointwise at most the least fixpoint (even 0, odd TOP,
    script/multi/mkladder.c), then resumes to it."""
    import ctypes
    import threading
    import time
    import egsolver_amd as E
    from test_gpu_cli import _ladder_text
    N = 300_000
    path = str(tmp_path / "ladder.mpg")
    with open(path, "w") as f:
        f.write(_ladder_text(N))
    g = E.read_mpg(path)
    lfp = np.where(g.ext_of_int % 2 == 0, 0, TOP).astype(np.int32)
    _say(f"stop flag, ladder N={N}, n {g.n}")
    flag = ctypes.c_uint32(0)
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        s.prepare(E.options(certify_every=-1, stop_flag=flag))
        t_set = []

        def fire():
            t_set.append(time.perf_counter())
            flag.value = 1
        timer = threading.Timer(0.5, fire)
        t0 = time.perf_counter()
        timer.start()
        conv, st = s.iterate()
        t1 = time.perf_counter()
        timer.join()
        assert not conv and t_set, "the run ended before the stop flag was set"
        lag = t1 - t_set[0]
        _say(f"stopped after {st.phases} rounds, {1e3 * lag:.2f} ms after the flag")
        assert lag < 0.05, lag
        part = s.values()
        assert (part <= lfp).all()
        assert part[g.ext_of_int % 2 == 1].max() > 0
        flag.value = 0
        t2 = time.perf_counter()
        conv, st2 = s.iterate()
        _say(f"resumed: {st2.phases} rounds in total, {time.perf_counter() - t2:.2f} s more")
        assert conv and st2.phases > st.phases
        assert np.array_equal(s.values(), lfp)


@pytest.mark.parametrize("seed", range(4))
def test_graph_replayed_rounds_equal_launched_rounds(oracle, seed):
    """Batches of rounds replayed as captured HIP graphs give the values of the same rounds
    launched kernel by kernel, across repeated solves of one prepared solver (graphs reused)
    and re-prepares with other shapes (graphs re-captured): lane, wavefront and workgroup
    kernels, sparse rounds and proofs.  Round counts and lift counts are not compared: a
    round reads values other workgroups raised in the same round, so they vary from run to
    run (by a few lifts) with either launch form."""
    import egsolver_amd as E
    rng = np.random.default_rng(50 + seed)
    n = int(rng.integers(200, 3000))
    p = E.GenParams(n=n, seed=500 + seed, deg_kind=2, deg_lo=1, deg_hi=min(n - 1, 400), alpha=1.8,
                    wmax=int(rng.integers(2, 60)))
    g = E.generate(p)
    _say(f"graphs seed={seed} n={g.n} m={g.m} maxdeg={g.max_outdeg}")
    ref = lfp(oracle, to_oracle_game(g))
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        for shape in (dict(group=1, big_deg=8, huge_deg=64), dict(group=4, big_deg=16, huge_deg=32, dense_pct=1),
                      dict(group=2, batch_rounds=7), dict(group=8, certify_every=-1, dense_pct=101)):
            for graphs in (-1, 1, 1):
                s.prepare(E.options(graphs=graphs, **shape))
                conv, st = s.iterate()
                assert conv
                vals = s.values()
                assert np.array_equal(vals, ref), (shape, graphs, np.count_nonzero(vals != ref))


def test_graph_replay_on_phased_game(oracle):
    """Phased dense rounds (one launch per round over 8 successor slices) and hub kernels,
    replayed as graphs: the same values as kernel-by-kernel launches, certified exact."""
    import egsolver_amd as E
    p = E.GenParams(n=(1 << 20) + 4097, seed=77, deg_kind=2, deg_lo=1, deg_hi=3000, alpha=2.2, wmax=40)
    g = E.generate(p)
    _say(f"graphs phased n={g.n} m={g.m}")
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        out = {}
        for graphs in (-1, 1):
            s.prepare(E.options(graphs=graphs, slices=8))
            conv, st = s.iterate()
            assert conv
            out[graphs] = s.values()
        assert np.array_equal(out[-1], out[1])
        vals = out[1]
        exact_by_certificate(oracle, g, vals, s)


@pytest.mark.parametrize("group,dense_pct", [(8, 0), (2, 1), (16, 101)])
def test_fused_hub_rounds_many_strides(oracle, group, dense_pct):
    """A one-sweep game with wavefront and workgroup hubs, large enough that the low-degree
    part of the round kernel strides over the vertices many times: with the hub workgroups
    in the same launch (dense rounds) the low-degree workgroups must still cover every
    vertex, and the deferred hubs of sparse rounds stay with their own kernels.  Exact by
    the host fixpoint check, the TOP-set certificate and the oracle's EG0 from it."""
    import egsolver_amd as E
    g = E.generate(E.GenParams(n=150_000, seed=61, deg_kind=2, deg_lo=1, deg_hi=20_000, alpha=2.0, wmax=40))
    _say(f"fused hub rounds n {g.n} m {g.m} maxdeg {g.max_outdeg} group {group} dense_pct {dense_pct}")
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        s.prepare(E.options(group=group, big_deg=64, huge_deg=2048, dense_pct=dense_pct))
        conv, st = s.iterate()
        assert conv and s.violations() == 0
        exact_by_certificate(oracle, g, s.values(), s)

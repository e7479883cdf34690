"""Pins the oracle (CPU, no GPU needed).

Each check is chosen so that a plausible mistake in the oracle fails it: the reference's
golden run, the reference's own compiled solver, hand-solved and closed-form families,
brute-force Kleene iteration, and planted faults.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_golden_sol

from oracle.oracle import TOP, Oracle, Reference, game_from_arrays, kleene, parse_mpg, parse_mpg_text

HAVE_REF = Reference.available()


@pytest.fixture(scope="module")
def clique_cp4():
    return parse_mpg(os.path.join(GOLDEN, "clique_500_cp4.txt.gz"))


def _rand_game(rng, n, dmax, wmax, n0=None):
    deg = rng.integers(1, dmax + 1, size=n)
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum(deg)
    succ = np.concatenate([rng.choice(n, size=int(d), replace=(d > n)) for d in deg]).astype(np.int32)
    w = rng.integers(-wmax, wmax + 1, size=int(ptr[-1])).astype(np.int32)
    if n0 is None:
        n0 = int(rng.integers(0, n + 1))
    return game_from_arrays(n, n0, ptr, succ, w)


# ---- golden run -------------------------------------------------------------------------

def test_golden_clique_cp4_eg0_and_eg(oracle, clique_cp4):
    """script/sol.txt: --eg0 --cpu on clique_500_cp4, 2011 sweeps, MG 1992, all 2000 values."""
    g = clique_cp4
    assert (g.n, g.n0, g.m, g.mg, g.max_abs_w) == (2000, 1000, 1746500, 1992, 1)
    gold = load_golden_sol()
    exp = np.array([gold[int(e)] for e in g.ext_of_int], np.int32)
    v0, sweeps = oracle.eg0(g, threads=4)
    assert sweeps == 2011                      # script/sol.txt:32
    assert np.array_equal(v0, exp)
    v1, _ = oracle.eg(g)
    assert np.array_equal(v1, exp)
    # the reference prints in internal order: first lines of script/sol.txt:43-45
    assert list(g.ext_of_int[:2]) == [0, 500]


def test_golden_counts(clique_cp4):
    gold = load_golden_sol()
    vals = np.array(list(gold.values()))
    assert (vals == 0).sum() == 996 and (vals == 1).sum() == 4 and (vals == TOP).sum() == 1000
    assert sorted(e for e, v in gold.items() if v == 1) == [498, 998, 1498, 1998]


# ---- the reference's own solver ---------------------------------------------------------

@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs the reference tree)")
def test_against_reference_build_random(oracle):
    R = Reference()
    rng = np.random.default_rng(7)
    for trial in range(60):
        g = _rand_game(rng, int(rng.integers(1, 120)), 6, int(rng.integers(1, 30)))
        ref_eg, ref_p, _ = R.solve(g, "eg")
        mine, p = oracle.eg(g)
        assert np.array_equal(mine, ref_eg), trial
        assert p == ref_p, trial                       # same FIFO, same pop count
        ref_eg0, ref_p0, _ = R.solve(g, "eg0", threads=3)
        mine0, sweeps = oracle.eg0(g, threads=3)
        assert np.array_equal(mine0, ref_eg0), trial
        assert sweeps * g.n == ref_p0, trial


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs the reference tree)")
def test_against_reference_build_loop_caps(oracle):
    """--eg N / --eg0 N partial results, including the EG0 odd-cap buffer choice."""
    R = Reference()
    rng = np.random.default_rng(11)
    for trial in range(30):
        g = _rand_game(rng, int(rng.integers(5, 80)), 5, 20)
        for cap in (1, 2, 3, 7, 20):
            a, pa, _ = R.solve(g, "eg", max_loop=cap)
            b, pb = oracle.eg(g, max_loop=cap)
            assert np.array_equal(a, b) and pa == pb, (trial, cap)
            a0, pa0, _ = R.solve(g, "eg0", max_loop=cap, threads=2)
            b0, s0 = oracle.eg0(g, max_loop=cap, threads=2)
            assert np.array_equal(a0, b0) and pa0 == s0 * g.n, (trial, cap)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs the reference tree)")
def test_all_active_queue_wrap_matches_reference(oracle):
    """Every vertex active at start: the reference's queue write index starts at n
    (cpu_solver.c:247,319); the restatement keeps that, so outputs agree."""
    R = Reference()
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(2, 40))
        deg = rng.integers(1, 4, size=n)
        ptr = np.zeros(n + 1, np.int32)
        ptr[1:] = np.cumsum(deg)
        m = int(ptr[-1])
        succ = rng.integers(0, n, size=m).astype(np.int32)
        w = -rng.integers(1, 5, size=m).astype(np.int32)          # all negative: all active
        g = game_from_arrays(n, int(rng.integers(0, n + 1)), ptr, succ, w)
        assert oracle.initial_active(g).all()
        a, pa, _ = R.solve(g, "eg")
        b, pb = oracle.eg(g)
        assert np.array_equal(a, b) and pa == pb


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs the reference tree)")
def test_transpose_against_reference_build(oracle):
    R = Reference()
    rng = np.random.default_rng(5)
    for _ in range(20):
        g = _rand_game(rng, int(rng.integers(1, 200)), 9, 50)
        c1, p1, w1 = R.csr2csc(g.n, g.m, g.ptr, g.succ, g.w)
        c2, p2, w2 = oracle.csr2csc(g, with_weights=True)
        assert np.array_equal(c1, c2) and np.array_equal(p1, p2) and np.array_equal(w1, w2)


# ---- known answers / closed forms ------------------------------------------------------

def test_transpose_known_answer(oracle):
    """The 4x5 example of src/host_csr2csc.c:49-133 (nnz 9)."""
    # rows: 0:{0:1,1:4}  1:{1:2,2:3}  2:{0:5,3:7,4:8}  3:{2:9,4:6}
    ptr = np.array([0, 2, 4, 7, 9], np.int32)
    col = np.array([0, 1, 1, 2, 0, 3, 4, 2, 4], np.int32)
    val = np.array([1, 4, 2, 3, 5, 7, 8, 9, 6], np.int32)
    # as a square n=5 arena (row 4 empty) so the same routine applies
    g = game_from_arrays(5, 5, np.append(ptr, 9).astype(np.int32), col, val)
    cptr, rows, cval = oracle.csr2csc(g, with_weights=True)
    assert cptr.tolist() == [0, 2, 4, 6, 7, 9]
    assert rows.tolist() == [0, 2, 0, 1, 1, 3, 2, 2, 3]
    assert cval.tolist() == [1, 5, 4, 2, 3, 9, 7, 8, 6]


def test_eg_fig2_hand_solution(oracle):
    """script/multi/eg_fig2.txt solved by hand: MG = 7; V(0)=0, V(1)=3, V(2)=TOP,
    V(3)=0, V(4)=TOP (player 0 wins {0,1,3})."""
    hand = {0: 0, 1: 3, 2: TOP, 3: 0, 4: TOP}
    for name in ["eg_fig2.txt", "eg_fig2_duecopie.txt", "eg_fig2_quattrocopie.txt",
                 "eg_fig2_duecopie_di_duecopie.txt"]:
        g = parse_mpg(os.path.join(GOLDEN, name))
        for vals in (oracle.eg(g)[0], oracle.eg0(g)[0], kleene(g)):
            got = dict(zip(g.ext_of_int.tolist(), vals.tolist()))
            assert all(v == hand[e % 5] for e, v in got.items()), name
    assert parse_mpg(os.path.join(GOLDEN, "eg_fig2.txt")).mg == 7


def _ladder(N):
    """script/multi/mkladder.c:42-51: even ids MAX with 0-weight edges to i+1, i+2; odd
    ids MIN with edges to i+1 (0) and i+2 (-1), indices mod 2N."""
    lines = [f"{2 * N - 1};"]
    for i in range(0, 2 * N, 2):
        lines.append(f"{i} MAX {(i + 1) % (2 * N)}:0,{(i + 2) % (2 * N)}:0;")
    for i in range(0, 2 * N, 2):
        lines.append(f"{i + 1} MIN {(i + 2) % (2 * N)}:0,{(i + 3) % (2 * N)}:-1;")
    return "\n".join(lines) + "\n"


def _clique(N):
    """script/multi/mkclique.c:42-65: even ids MAX, odd ids MIN; edge i->j weighs
    (i<j ? 1 : -1) for MAX vertices and the negation for MIN vertices."""
    lines = [f"{N - 1};"]
    for i in range(N):
        own = "MAX" if i % 2 == 0 else "MIN"
        sgn = 1 if i % 2 == 0 else -1
        edges = ",".join(f"{j}:{sgn * (1 if i < j else -1)}" for j in range(N) if j != i)
        lines.append(f"{i} {own} {edges};")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("N", [1, 2, 3, 10, 57])
def test_ladder_closed_form(oracle, N):
    """MIN vertices can cycle on -1 edges forever: odd -> TOP; MAX vertices stay on the
    0-weight even chain: even -> 0.  MG = N."""
    g = parse_mpg_text(_ladder(N))
    assert g.mg == N
    for vals in (oracle.eg(g)[0], oracle.eg0(g)[0]):
        got = dict(zip(g.ext_of_int.tolist(), vals.tolist()))
        assert all(v == (0 if e % 2 == 0 else TOP) for e, v in got.items())


@pytest.mark.parametrize("N", [6, 7, 12, 33, 60])
def test_clique_closed_form(oracle, N):
    """With at least three MIN (odd) vertices the MIN player closes the cycle 1->3->5->1
    of weight -1-1+1 < 0, so odd vertices reach TOP.  An even (MAX) vertex moves to a
    higher even vertex at weight +1 for credit 0; the largest even id has only lower
    targets, at weight -1, so it needs 1 (the script/sol.txt pattern: id 498 -> 1 at
    N=500)."""
    g = parse_mpg_text(_clique(N))
    last_even = N - 1 if (N - 1) % 2 == 0 else N - 2
    for vals in (oracle.eg(g)[0], oracle.eg0(g)[0]):
        got = dict(zip(g.ext_of_int.tolist(), vals.tolist()))
        for e, v in got.items():
            exp = TOP if e % 2 else (1 if e == last_even else 0)
            assert v == exp, (N, e, v)


def test_brute_force_kleene(oracle):
    """EG0 always equals brute force.  EG equals it too, except when every vertex starts
    active (see test_reference_eg_all_active_start_is_not_a_fixpoint)."""
    rng = np.random.default_rng(1)
    for trial in range(150):
        g = _rand_game(rng, int(rng.integers(1, 12)), 4, int(rng.integers(1, 9)))
        exp = kleene(g)
        assert np.array_equal(oracle.eg0(g, threads=2)[0], exp), trial
        assert oracle.fixpoint_violations(g, exp) == 0
        if not oracle.initial_active(g).all():
            assert np.array_equal(oracle.eg(g)[0], exp), trial


def test_reference_eg_all_active_start_is_not_a_fixpoint(oracle):
    """A defect of the reference's worklist solver: with every vertex active at start the
    queue's write index begins at n (cpu_solver.c:247) and the first re-queued vertex is
    written to a slot the reader never visits (cpu_solver.c:293,319), while its queued
    flag stays set; it is never lifted again.  On this 3-vertex game EG stops at 4 for a
    vertex whose least fixpoint is TOP.  The restatement reproduces it (it matches the
    compiled reference, test_all_active_queue_wrap_matches_reference); EG0 and brute force
    give the fixpoint, which is what the GPU path computes."""
    rng = np.random.default_rng(1)
    found = 0
    for trial in range(150):
        g = _rand_game(rng, int(rng.integers(1, 12)), 4, int(rng.integers(1, 9)))
        if not oracle.initial_active(g).all():
            continue
        eg, _ = oracle.eg(g)
        exp = kleene(g)
        if not np.array_equal(eg, exp):
            found += 1
            assert oracle.fixpoint_violations(g, eg) > 0
            assert np.all(eg <= exp)            # still an under-approximation
    assert found >= 1


def test_ominus_cases(oracle):
    """OMINUS (src/common.h:270-272): TOP absorbs, results clamp at 0, and a - w above MG
    becomes TOP while a - w == MG stays finite."""
    O = oracle
    assert O.ominus(TOP, -5, 100) == TOP
    assert O.ominus(5, 7, 100) == 0
    assert O.ominus(5, -7, 100) == 12
    assert O.ominus(95, -5, 100) == 100
    assert O.ominus(95, -6, 100) == TOP
    assert O.ominus(0, 0, 0) == 0


def test_mg_and_max_loop(oracle):
    # MG = sum_v max(0, max_e -w_e)
    g = game_from_arrays(3, 1, np.array([0, 2, 3, 5], np.int32), np.array([1, 2, 0, 0, 1], np.int32),
                         np.array([-3, 5, 4, -1, -7], np.int32))
    assert oracle.mg(g) == 3 + 0 + 7 == g.mg
    big = game_from_arrays(2, 1, np.array([0, 1, 2], np.int32), np.array([1, 0], np.int32),
                           np.array([-1, -1], np.int32))
    big.w = np.array([-2147483647, -1], np.int32)     # 2147483647 + 1 passes INT_MAX
    with pytest.raises(OverflowError):
        oracle.mg(big)
    assert oracle.max_loop(10, 5, 7) == 5 * 10 * 7 + 1
    assert oracle.max_loop(10, 5, 7, 42) == 42
    assert oracle.max_loop(1 << 40, 1 << 20, 1 << 10) == (1 << 63) - 2   # saturates at LONG_MAX-1


def test_fixpoint_certificate_rejects_planted_faults(oracle):
    rng = np.random.default_rng(9)
    g = _rand_game(rng, 300, 5, 20)
    vals, _ = oracle.eg(g)
    assert oracle.fixpoint_violations(g, vals) == 0
    for _ in range(30):
        bad = vals.copy()
        v = int(rng.integers(0, g.n))
        bad[v] = bad[v] + 1 if bad[v] < TOP else 0
        assert oracle.fixpoint_violations(g, bad) >= 1


def test_eg_and_eg0_agree_on_larger_random(oracle):
    rng = np.random.default_rng(21)
    for _ in range(10):
        g = _rand_game(rng, 400, 8, 100)
        a, _ = oracle.eg(g)
        b, _ = oracle.eg0(g, threads=4)
        assert np.array_equal(a, b)
        assert oracle.fixpoint_violations(g, a) == 0


# ---- one lift, initial frontier, seeded EG0 -----------------------------------------------

def _kleene_step(g, f):
    """One step of the brute-force Kleene iteration, straight from the definition."""
    out = []
    for v in range(g.n):
        xs = []
        for e in range(g.ptr[v], g.ptr[v + 1]):
            a, w = int(f[g.succ[e]]), int(g.w[e])
            xs.append(max(0, a - w) if (a < TOP and a - w <= g.mg) else TOP)
        out.append(min(xs) if v < g.n0 else max(xs))
    return np.array(out, np.int32)


def test_lift_all_pinned(oracle):
    """ora_lift_all: equal to a brute-force Kleene step on random values, the identity on the
    least fixpoint, and lift(0) equal to one EG0 sweep from 0 (cpu_solver.c:109-185)."""
    rng = np.random.default_rng(21)
    for trial in range(60):
        g = _rand_game(rng, int(rng.integers(2, 40)), 5, 6)
        f = rng.integers(0, g.mg + 2, size=g.n).astype(np.int32)
        f[rng.random(g.n) < 0.2] = TOP
        assert np.array_equal(oracle.lift_all(g, f), _kleene_step(g, f))
        lfp = kleene(g)
        assert np.array_equal(oracle.lift_all(g, lfp), lfp)
        one, _ = oracle.eg0(g, max_loop=2, threads=1)
        z = np.zeros(g.n, np.int32)
        assert np.array_equal(oracle.lift_all(g, oracle.lift_all(g, z)), one)


def test_initial_active_mixed_signs(oracle):
    """ora_initial_active (cpu_solver.c:217-249) equals lift(0)(v) > 0 on games whose vertices
    have mixed-sign weights, for both owners; swapping the owner rules fails it."""
    rng = np.random.default_rng(22)
    seen = set()
    for trial in range(80):
        g = _rand_game(rng, int(rng.integers(2, 60)), 6, 4)
        act = oracle.initial_active(g).astype(bool)
        want = _kleene_step(g, np.zeros(g.n, np.int32)) > 0
        assert np.array_equal(act, want)
        for v in range(g.n):
            ws = g.w[g.ptr[v]:g.ptr[v + 1]]
            seen.add((v < g.n0, bool((ws < 0).any()), bool((ws >= 0).any()), bool(act[v])))
    # every owner x sign pattern occurred: (owner0, has neg, has non-neg, active)
    assert {(True, True, True, False), (False, True, True, True), (True, True, False, True),
            (False, False, True, False)} <= seen
    # hand-built: owner 0 with mixed weights is idle, owner 1 with mixed weights is active
    g = game_from_arrays(2, 1, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32),
                         np.array([-1, 2, -1, 2], np.int32))
    assert oracle.initial_active(g).tolist() == [0, 1]


def test_eg0_from_seeded(oracle):
    """ora_eg0_from: from 0 it is EG0; from TOP on the least fixpoint's TOP set it returns
    the least fixpoint in no more sweeps; a planted value above the least fixpoint survives
    (the seed is used, not zeroed)."""
    rng = np.random.default_rng(23)
    fewer = 0
    for trial in range(60):
        g = _rand_game(rng, int(rng.integers(2, 80)), 4, 8)
        lfp, sweeps = oracle.eg0(g, threads=2)
        a, s0 = oracle.eg0_from(g, np.zeros(g.n, np.int32), threads=2)
        assert np.array_equal(a, lfp)
        seed = np.where(lfp == TOP, TOP, 0).astype(np.int32)
        b, s1 = oracle.eg0_from(g, seed, threads=2)
        assert np.array_equal(b, lfp) and s1 <= s0
        fewer += s1 < s0
        fin = np.nonzero(lfp != TOP)[0]
        if len(fin):
            v = int(fin[0])
            up = lfp.copy()
            up[v] = TOP
            c, _ = oracle.eg0_from(g, up, threads=2)
            assert c[v] == TOP
    assert fewer > 10


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built (needs the reference tree)")
def test_reference_eg0_race_free_at_bench_threads(oracle):
    """The reference's EG0 declares its inner edge index `idy` at function scope
    (src/cpu_solver.c:111) and uses it inside `#pragma omp parallel for` (:146-149), where it
    is shared between threads: a data race.  The compiled build keeps it in a register, so
    the values still come out right; this checks that at the bench's thread count (16) on a
    game of 2^20 edges, against the restatement, which has no shared index."""
    import egsolver_amd as E
    g = E.config_game(2, scale=1 / 32)                     # n = 131072, m = 2^20
    og = game_from_arrays(g.n, g.n0, g.ptr, g.succ, g.w)
    assert og.m == 1 << 20
    R = Reference()
    for cap in (7, 40):
        a, pa, _ = R.solve(og, "eg0", max_loop=cap, threads=16)
        b, sb = oracle.eg0(og, max_loop=cap, threads=16)
        assert pa == sb * og.n
        assert np.array_equal(a, b), cap

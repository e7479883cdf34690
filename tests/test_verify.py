"""Host-side result checks (libegs_host.so): the fixpoint check and the TOP-set
certificate check, pinned on the CPU against brute force and planted faults.

The certificate check must accept potentials computed for the true TOP set of the least
fixpoint (brute-force Kleene iteration) and reject every planted fault: a finite vertex
put into the TOP set (no potentials can certify it, since the check is sound), a broken
inequality, a broken membership, an edge leaving the set."""
import numpy as np
import pytest

from oracle.oracle import TOP, game_from_arrays, kleene


def _tiny_game(rng, n, wlo=-3, whi=3, dmax=3):
    owner = rng.integers(0, 2, size=n)
    n0 = int((owner == 0).sum())
    deg = rng.integers(1, dmax + 1, size=n)
    ptr = np.zeros(n + 1, np.int32)
    ptr[1:] = np.cumsum(deg)
    succ, w = [], []
    for v in range(n):
        succ += rng.choice(n, size=deg[v], replace=True).tolist()
        w += rng.integers(wlo, whi + 1, size=deg[v]).tolist()
    return game_from_arrays(n, n0, ptr, np.array(succ, np.int32), np.array(w, np.int32))


def _potentials(g, inT, iters=100000):
    """Longest-path game on T from h = 0 (test-side, pure Python): owner 0 the max over its
    edges into T, owner 1 the min over them.  Returns h (-1 outside T) and whether it reached
    a fixpoint."""
    n = g.n
    h = [0 if inT[v] else -1 for v in range(n)]
    ptr, succ, w = g.ptr.tolist(), g.succ.tolist(), g.w.tolist()
    for _ in range(iters):
        changed = False
        for v in range(n):
            if not inT[v]:
                continue
            cands = [(n + 1) * w[e] + 1 + h[succ[e]] for e in range(ptr[v], ptr[v + 1]) if inT[succ[e]]]
            if not cands:
                continue
            c = max(cands) if v < g.n0 else min(cands)
            if c > h[v]:
                h[v] = c
                changed = True
            if h[v] > 10 ** 15:
                return np.array(h, np.int64), False
        if not changed:
            return np.array(h, np.int64), True
    return np.array(h, np.int64), False


def _check(g, vals, h):
    from egsolver_amd import verify
    return verify.check_top_certificate(g, vals, h)


def _fix(g, vals):
    from egsolver_amd import verify
    return verify.check_fixpoint(g, vals)


def test_certificate_accepts_true_top_sets_and_rejects_planted_faults():
    rng = np.random.default_rng(11)
    accepted = planted = 0
    for trial in range(300):
        g = _tiny_game(rng, int(rng.integers(2, 11)))
        lfp = kleene(g)
        inT = lfp == TOP
        h, ok = _potentials(g, inT)
        assert ok, "potentials exist for the true TOP set"
        assert _check(g, lfp, h) == (0, -1)
        accepted += 1
        # 1. a finite vertex raised to TOP: no potentials certify it (try the iteration's and random ones)
        finite = np.nonzero(~inT)[0]
        if len(finite):
            v = int(rng.choice(finite))
            bad = lfp.copy()
            bad[v] = TOP
            for k in (1, 3, 50):
                hb, _ = _potentials(g, bad == TOP, iters=k)
                assert _check(g, bad, hb)[0] != 0
            for _ in range(5):
                hr = np.where(bad == TOP, rng.integers(0, 10 ** 6, size=g.n), -1).astype(np.int64)
                assert _check(g, bad, hr)[0] != 0
            planted += 1
        # 2. a TOP vertex whose inequality is made to fail by one unit
        tops = np.nonzero(inT)[0]
        if len(tops):
            v = int(rng.choice(tops))
            es = range(g.ptr[v], g.ptr[v + 1])
            c = [(g.n + 1) * int(g.w[e]) + 1 + int(h[g.succ[e]]) for e in es if inT[g.succ[e]]]
            need = max(c) if v < g.n0 else min(c)
            if need - 1 >= 0:
                hb = h.copy()
                hb[v] = need - 1
                code, bv = _check(g, lfp, hb)
                assert code in (3, 4) and bv == v
            # 3. membership: a TOP vertex without a potential, a finite vertex with one
            hb = h.copy()
            hb[v] = -1
            code, bv = _check(g, lfp, hb)
            assert code != 0 and (code == 1) == (bv == v)
        if len(finite):
            hb = h.copy()
            hb[int(rng.choice(finite))] = 0
            assert _check(g, lfp, hb)[0] == 1
    assert accepted == 300 and planted > 100


def test_certificate_edge_leaving_top_set():
    # owner-0 vertex 0 -> {2, 1}; owner-0 vertex 1 loops with weight 0 (player 0 wins there),
    # owner-1 vertex 2 loops with weight -1 (player 1 wins there).
    g = game_from_arrays(3, 2, np.array([0, 2, 3, 4], np.int32), np.array([2, 1, 1, 2], np.int32),
                         np.array([0, 0, 0, -1], np.int32))
    lfp = kleene(g)
    assert lfp.tolist() == [0, 0, TOP]
    h, ok = _potentials(g, lfp == TOP)
    assert ok and _check(g, lfp, h) == (0, -1)
    bad = lfp.copy()
    bad[0] = TOP                                 # 0 has an edge to finite vertex 1
    hb = np.array([10 ** 9, -1, h[2]], np.int64)
    assert _check(g, bad, hb) == (2, 0)


def test_certificate_closed_forms():
    """eg_fig2 (hand-solved, script/multi/eg_fig2.txt) and ladder games (mkladder.c): the
    TOP sets of the known solutions are certified."""
    import os
    from conftest import GOLDEN
    import egsolver_amd as E
    g = E.read_mpg(os.path.join(GOLDEN, "eg_fig2.txt"))
    hand = {0: 0, 1: 3, 2: TOP, 3: 0, 4: TOP}
    vals = np.array([hand[e] for e in g.ext_of_int.tolist()], np.int32)
    og = game_from_arrays(g.n, g.n0, g.ptr, g.succ, g.w, g.ext_of_int)
    h, ok = _potentials(og, vals == TOP)
    assert ok and _check(og, vals, h) == (0, -1)
    assert _fix(og, vals) == (0, -1)


def test_fixpoint_check_agrees_with_oracle(oracle):
    """Planted changes of one value: the count equals the oracle's, and the first vertex
    reported is the first whose lift differs (some changes still give a fixpoint, e.g. TOP on
    a zero-weight cycle of owner 1; both checks must then say 0)."""
    rng = np.random.default_rng(5)
    nonzero = 0
    for trial in range(200):
        g = _tiny_game(rng, int(rng.integers(2, 30)), wlo=-5, whi=5)
        lfp = kleene(g)
        assert _fix(g, lfp) == (0, -1)
        bad = lfp.copy()
        v = int(rng.integers(0, g.n))
        bad[v] = TOP if bad[v] != TOP else 0
        cnt, first = _fix(g, bad)
        assert cnt == oracle.fixpoint_violations(g, bad)
        diff = np.nonzero(oracle.lift_all(g, bad) != bad)[0]
        assert first == (int(diff[0]) if len(diff) else -1)
        nonzero += cnt > 0
    assert nonzero > 150

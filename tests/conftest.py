import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle
    return Oracle()


def load_golden_sol(name="clique_500_cp4.sol"):
    vals = {}
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            if line.startswith("#"):
                continue
            e, v = line.split()
            vals[int(e)] = int(v)
    return vals


def to_oracle_game(g):
    from oracle.oracle import game_from_arrays
    return game_from_arrays(g.n, g.n0, g.ptr, g.succ, g.w, g.ext_of_int)


def lfp(oracle, og, threads=4):
    """Least fixpoint from the oracle: EG0 (Jacobi), cross-checked against the EG
    worklist whenever its all-active-start defect cannot apply."""
    import numpy as np
    v0, _ = oracle.eg0(og, threads=threads)
    if not oracle.initial_active(og).all():
        v1, _ = oracle.eg(og)
        assert np.array_equal(v0, v1), "oracle EG and EG0 disagree"
    return v0


def exact_by_certificate(oracle, g, vals, solver):
    """Full-size parity without trusting the device's TOP proofs.

    1. The host fixpoint check (lift(f) == f): f lies at or above the least fixpoint, so the
       least fixpoint's TOP set is inside T = {f == TOP}.
    2. A certificate for T computed on the device (solver holds g, prepared) and checked on
       the host: T is won by player 1, so T is inside the least fixpoint's TOP set.
    3. With T pinned, the oracle's EG0 from (TOP on T, 0 elsewhere) rises to the least
       fixpoint (a start below it), and f must equal it bit for bit.
    Returns the certificate report."""
    import numpy as np
    from egsolver_amd import verify
    rep = verify.certify(solver, g, vals)
    assert rep["fixpoint_violations"] == 0, rep
    assert rep["certificate_found"], rep
    assert rep["certificate_code"] == 0, (verify.CERT_CODES[rep["certificate_code"]], rep)
    og = to_oracle_game(g)
    start = np.where(vals == 2147483647, 2147483647, 0).astype(np.int32)
    ref, sweeps = oracle.eg0_from(og, start)
    assert np.array_equal(vals, ref), f"{int((vals != ref).sum())} values differ from the oracle's seeded EG0"
    rep["oracle_sweeps"] = int(sweeps)
    return rep


def gen_family(family, seed, count, workers=16):
    """Games 0..count-1 of a multi-game family (batched: 4, serving: 5), generated in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    import egsolver_amd as E

    def one(i):
        return E.generate(E.family_params(family, seed, i), stream=i)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(one, range(count)))


def exact_sample(oracle, games, results, idx, solver, unseeded_max_n=12000, threads=16):
    """Bit-exact check of results[i] for i in idx: the oracle's EG0 from 0 (the reference's
    algorithm end to end) on games up to unseeded_max_n vertices, and on every sampled game
    the certificate route of exact_by_certificate (host fixpoint check, device certificate of
    the TOP set checked on the host, oracle EG0 from the certified TOP set).  `solver` must
    be sized for the largest sampled game.  Returns (games checked from 0, games checked
    through the certificate)."""
    import numpy as np
    from egsolver_amd import verify
    unseeded = seeded = 0
    for i in idx:
        g, vals = games[i], results[i]
        og = to_oracle_game(g)
        if g.n <= unseeded_max_n:
            ref, _ = oracle.eg0(og, threads=threads)
            assert np.array_equal(vals, ref), (i, g.n, int((vals != ref).sum()))
            unseeded += 1
        solver.load(g)
        solver.prepare()
        rep = verify.certify(solver, g, vals)
        assert rep["ok"], (i, g.n, rep)
        start = np.where(vals == 2147483647, 2147483647, 0).astype(np.int32)
        ref, _ = oracle.eg0_from(og, start, threads=threads)
        assert np.array_equal(vals, ref), (i, g.n, int((vals != ref).sum()))
        seeded += 1
    return unseeded, seeded

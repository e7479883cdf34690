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

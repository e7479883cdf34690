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

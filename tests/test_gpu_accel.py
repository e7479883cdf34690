"""GPU parity with certified TOP detection: proofs that set climbing vertices to TOP
must leave every value equal to the oracle's least fixpoint."""
import numpy as np
import pytest

from conftest import lfp, to_oracle_game

pytestmark = pytest.mark.gpu

TOP = 2147483647


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.mark.parametrize("which,scale,every", [(2, 1 / 256, 8), (2, 1 / 1024, 4), (3, 1 / 64, 8), (1, 1.0, 4)])
def test_accel_exact_on_configs(oracle, which, scale, every):
    import egsolver_amd as E
    g = E.config_game(which, scale=scale)
    _say(f"accel config{which} n={g.n} every={every}")
    ref = lfp(oracle, to_oracle_game(g))
    vals, st = E.solve(g, certify_every=every)
    assert st["converged"]
    assert np.array_equal(vals, ref), np.count_nonzero(vals != ref)
    if (ref == TOP).any() and which != 1:
        assert st["certificates"] >= 1 and st["top_marked"] > 0
    plain, st0 = E.solve(g, certify_every=-1)
    assert np.array_equal(plain, ref)
    assert st0["certificates"] == 0


@pytest.mark.parametrize("seed", range(8))
def test_accel_exact_random_families(oracle, seed):
    import egsolver_amd as E
    rng = np.random.default_rng(seed)
    kind = [0, 1, 2, 0][seed % 4]
    n = int(rng.integers(200, 3000))
    p = E.GenParams(n=n, seed=500 + seed, deg_kind=kind, deg_lo=int(rng.integers(1, 4)),
                    deg_hi=int(rng.integers(4, 40)) if kind != 2 else n - 1, alpha=2.0,
                    wmax=int(rng.integers(1, 500)))
    if kind == 0:
        p.deg_hi = p.deg_lo
    g = E.generate(p)
    _say(f"accel random seed={seed} n={g.n} m={g.m}")
    ref = lfp(oracle, to_oracle_game(g))
    for every, dense in [(2, 0), (5, 101), (16, 1)]:
        vals, st = E.solve(g, certify_every=every, dense_pct=dense)
        assert np.array_equal(vals, ref), (every, dense, np.count_nonzero(vals != ref))


def test_accel_clique_golden():
    import os
    import egsolver_amd as E
    from conftest import GOLDEN, load_golden_sol
    g = E.read_mpg(os.path.join(GOLDEN, "clique_500_cp4.txt.gz"))
    _say("accel clique_500_cp4")
    vals, st = E.solve(g, certify_every=4)
    got = dict(zip(g.ext_of_int.tolist(), vals.tolist()))
    assert got == load_golden_sol()


def test_accel_no_top_region(oracle):
    """All weights non-negative: every value stays 0, no proof is attempted."""
    import egsolver_amd as E
    p = E.GenParams(n=2000, seed=4, deg_kind=0, deg_lo=3, deg_hi=3, wmax=5)
    g = E.generate(p)
    g.w = np.abs(g.w)
    g.mg = 0
    vals, st = E.solve(g, certify_every=1)
    assert (vals == 0).all()
    assert st["certificates"] == 0

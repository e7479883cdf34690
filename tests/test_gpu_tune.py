"""The autotuner on the device: the options it picks for a game keep the values exact
(they change launch shapes, thresholds and phase counts, never the least fixpoint)."""
import numpy as np
import pytest

from conftest import exact_by_certificate, lfp, to_oracle_game

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def test_tuned_single_game_options_exact(oracle):
    import egsolver_amd as E
    from egsolver_amd.autotune import _TUNED_KEYS, tune_single
    # skewed degrees, one-sweep rounds: the degree thresholds, lanes and grid are tuned
    g = E.config_game(3, scale=1 / 64)
    _say(f"tune skewed n {g.n} m {g.m}")
    best, table = tune_single(g, reps=1)
    assert len(table) >= 10 and np.isfinite(best["solve_ms"])
    vals, st = E.solve(g, **{k: best[k] for k in _TUNED_KEYS})
    assert st["converged"] and np.array_equal(vals, lfp(oracle, to_oracle_game(g)))
    # above 2^20 vertices, phased rounds: the phase count and workgroups per CU are tuned
    g = E.config_game(2, scale=0.3)
    _say(f"tune uniform n {g.n} m {g.m}")
    best, table = tune_single(g, reps=1)
    assert {t["slices"] for t in table} >= {4, 8, 16}
    opts = {k: best[k] for k in _TUNED_KEYS}
    with E.Solver(g.n, g.m) as s:
        s.load(g)
        for o in (opts, dict(opts, slices=16, phase_wgs_per_cu=4)):
            s.prepare(E.options(**o))
            conv, st = s.iterate()
            assert conv
            exact_by_certificate(oracle, g, s.values(), s)

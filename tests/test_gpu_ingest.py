"""The owner partition, internal CSR and MG computed on the device from a game in read order
(egs_solver_load_read_order): equal to the oracle reader's internal game (src/egsolver.c:
146-214, 253-281, 283-304 as restated in oracle/oracle.py) and to the host renaming, and
solving it gives the oracle's values."""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN, lfp, to_oracle_game

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _device_internal(E, owner, rptr, rsucc, rw):
    n, m = len(owner), int(rptr[-1])
    with E.Solver(n, m) as s:
        n0, mg, ext = s.load_read_order(owner, rptr, rsucc, rw)
        ptr, succ, w = s.game()
        s.prepare()
        conv, _ = s.iterate()
        assert conv
        vals = s.values()
    return n0, mg, ext, ptr, succ, w, vals


def test_read_order_generated_games(oracle):
    import egsolver_amd as E
    from egsolver_amd import games as G
    params = [E.GenParams(n=3000, seed=1, deg_kind=0, deg_lo=4, deg_hi=4, wmax=10),
              E.GenParams(n=20000, seed=2, deg_kind=2, deg_lo=1, deg_hi=5000, alpha=1.7, wmax=100),
              E.GenParams(n=300, seed=3, deg_kind=3, deg_lo=1, deg_hi=1, wmax=7),
              E.GenParams(n=1 << 18, seed=4, deg_kind=1, deg_lo=2, deg_hi=16, wmax=1000)]
    for p in params:
        owner, rptr, rsucc, rw = G.generate_read_order(p)
        host = E.generate(p)                      # the host renaming of the same game
        _say(f"read order n {p.n} m {host.m}")
        n0, mg, ext, ptr, succ, w, vals = _device_internal(E, owner, rptr, rsucc, rw)
        assert (n0, mg) == (host.n0, host.mg)
        assert np.array_equal(ext, host.ext_of_int)
        assert np.array_equal(ptr, host.ptr) and np.array_equal(succ, host.succ) and np.array_equal(w, host.w)
        if p.n <= 20000:
            assert np.array_equal(vals, lfp(oracle, to_oracle_game(host)))


def test_read_order_golden_files_match_oracle_reader(oracle):
    """The reference's own inputs through this repository's reader in read order, renamed on
    the device, against the oracle's independent reader and renaming."""
    import egsolver_amd as E
    from oracle.oracle import parse_mpg
    for name in ("eg_fig2.txt", "eg_fig2_quattrocopie.txt", "clique_500_cp4.txt.gz"):
        path = os.path.join(GOLDEN, name)
        raw = open(path, "rb").read()
        text = gzip.decompress(raw) if raw[:2] == b"\x1f\x8b" else raw
        ro = E.parse_mpg(text)                    # read order, successors as external ids
        read_of_ext = np.full(int(ro["ext_of_read"].max()) + 1, -1, np.int64)
        read_of_ext[ro["ext_of_read"]] = np.arange(ro["n"])
        rsucc = read_of_ext[ro["rsucc"]].astype(np.int32)
        _say(f"read order {name}")
        n0, mg, ext, ptr, succ, w, vals = _device_internal(E, ro["owner"], ro["rptr"], rsucc, ro["rw"])
        og = parse_mpg(path)
        assert (n0, mg) == (og.n0, og.mg)
        assert np.array_equal(ro["ext_of_read"][ext], og.ext_of_int)
        assert np.array_equal(ptr, og.ptr) and np.array_equal(succ, og.succ) and np.array_equal(w, og.w)
        assert np.array_equal(vals, lfp(oracle, og))


def test_read_order_refuses_bad_input():
    import egsolver_amd as E
    from egsolver_amd import _native as N
    owner = np.array([0, 1, 0], np.uint8)
    good = (np.array([0, 1, 2, 3], np.int32), np.array([1, 2, 0], np.int32), np.array([-1, 2, 3], np.int32))
    with E.Solver(8, 8) as s:
        s.load_read_order(owner, *good)
        cases = {"no out-edge": (np.array([0, 1, 1, 3], np.int32), good[1], good[2]),
                 "out of range": (good[0], np.array([1, 5, 0], np.int32), good[2]),
                 "offsets": (np.array([0, 1, 2, 2], np.int32), good[1][:2], good[2][:2]),
                 "MG": (good[0], good[1], np.array([-(1 << 30), -(1 << 30), -(1 << 30)], np.int32))}
        for what, arrs in cases.items():
            _say(f"read order refuses: {what}")
            with pytest.raises(N.EgsError):
                s.load_read_order(owner, *arrs)

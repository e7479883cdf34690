"""The no-allocation entry point egs_solve_noalloc (include/egs.h), called through the C-ABI
on device arrays: its values and convergence equal those of egs_solve, the allocating
entry point, and the oracle's least fixpoint, on two calls in a row (the second on a
smaller game, so the shared workspace holds the first call's contents), and games beyond
its fixed workspace are refused with EGS_E_CAPACITY before any device work."""
import ctypes

import numpy as np
import pytest

from conftest import lfp, to_oracle_game

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Dev:
    """Device buffers for the test through the HIP runtime (the test allocates; the entry
    point under test does not)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(arr.nbytes, 4))) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, ctypes.c_void_p(arr.ctypes.data), ctypes.c_size_t(arr.nbytes), H2D) == 0
        return p

    def empty(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p

    def get(self, p, n):
        out = np.empty(n, np.int32)
        assert self.hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), p, ctypes.c_size_t(4 * n), D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc(dev, g, opt=None):
    import egsolver_amd as E
    from egsolver_amd import _native as N
    d_ptr, d_succ, d_w = dev.put(g.ptr.astype(np.int32)), dev.put(g.succ.astype(np.int32)), dev.put(g.w.astype(np.int32))
    d_val = dev.empty(4 * g.n)
    st = N.EgsStats()
    o = opt or E.options()
    rc = N.egs().egs_solve_noalloc(g.n, g.n0, g.m, g.mg, d_ptr, d_succ, d_w, d_val, ctypes.byref(o), ctypes.byref(st),
                                   None)
    return rc, dev.get(d_val, g.n), st


def test_noalloc_matches_allocating_entry_point_twice(oracle):
    import egsolver_amd as E
    first = E.config_game(2, scale=60000 / (1 << 22))              # n 60,000, out-degree 8
    second = E.generate(E.GenParams(n=5000, seed=41, deg_kind=1, deg_lo=2, deg_hi=12, wmax=300))
    assert first.n <= 65536 and first.m <= 1 << 20 and second.n < first.n
    dev = _Dev()
    try:
        for g in (first, second):
            print(f"[gpu-case] noalloc n {g.n} m {g.m}", flush=True)
            rc, vals, st = _noalloc(dev, g)
            assert rc == 1 and st.converged == 1
            want, wst = E.solve(g)
            assert np.array_equal(vals, want)
            assert wst["converged"]
            if g is second:
                assert np.array_equal(vals, lfp(oracle, to_oracle_game(g)))
        # the same small game again, proofs off: still the fixpoint
        rc, vals, st = _noalloc(dev, second, E.options(certify_every=-1))
        assert rc == 1 and np.array_equal(vals, E.solve(second)[0])
    finally:
        dev.free()


def test_noalloc_refuses_games_beyond_its_workspace():
    import egsolver_amd as E
    from egsolver_amd import _native as N
    g = E.generate(E.GenParams(n=70000, seed=3, deg_kind=0, deg_lo=2, deg_hi=2, wmax=10))
    st = N.EgsStats()
    o = E.options()
    # the size check comes first: no device pointer is touched
    fake = ctypes.c_void_p(16)
    rc = N.egs().egs_solve_noalloc(g.n, g.n0, g.m, g.mg, fake, fake, fake, fake, ctypes.byref(o), ctypes.byref(st),
                                   None)
    assert rc == -5
    assert b"exceeds the fixed workspace" in N.egs().egs_last_error()

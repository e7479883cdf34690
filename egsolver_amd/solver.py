"""Single-game solver: Python mirror of the reference's GPU entry (gpu_solver(),
src/dev_solver.cu:81-139) over libegs.so.

``Solver`` keeps one game resident on the device; ``solve()`` is the one-shot form.
The values it returns are the least fixpoint of the energy lifting, bit-identical to the
reference's cpu_solver.c on the same game.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .games import Game


def options(group=0, big_deg=0, huge_deg=0, grid=0, batch_rounds=0, max_phases=-1, stop_flag=None, dense_pct=0,
            certify_every=0, cert_max_passes=0, slices=0, phase_wgs_per_cu=0, graphs=0):
    o = N.EgsOptions(group=group, big_deg=big_deg, huge_deg=huge_deg, grid=grid,
                     batch_rounds=batch_rounds, dense_pct=dense_pct, certify_every=certify_every,
                     cert_max_passes=cert_max_passes, max_phases=max_phases, stop_flag=None, slices=slices,
                     phase_wgs_per_cu=phase_wgs_per_cu, graphs=graphs)
    if stop_flag is not None:
        o.stop_flag = ctypes.cast(ctypes.pointer(stop_flag), ctypes.c_void_p)
    return o


class Solver:
    """One device workspace sized for games up to (n_cap, m_cap)."""

    def __init__(self, n_cap, m_cap, device=0):
        h = ctypes.c_void_p()
        N.check(N.egs().egs_solver_create(device, int(n_cap), int(m_cap), ctypes.byref(h)), "egs_solver_create")
        self.h = h
        self.n = 0
        self.m = 0

    def close(self):
        if self.h:
            N.egs().egs_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def load(self, g: Game):
        ptr = np.ascontiguousarray(g.ptr, np.int32)
        succ = np.ascontiguousarray(g.succ, np.int32)
        w = np.ascontiguousarray(g.w, np.int32)
        N.check(N.egs().egs_solver_load(self.h, 0, g.n, g.n0, g.m, g.mg, ptr.ctypes.data, succ.ctypes.data,
                                        w.ctypes.data), "egs_solver_load")
        self.n, self.m = g.n, g.m

    def load_read_order(self, owner, rptr, rsucc, rw):
        """A game in read order (external id = read index, e.g. games.generate_read_order);
        the owner partition, internal CSR and MG are computed on the device.  Returns
        (n0, mg, ext_of_int)."""
        owner = np.ascontiguousarray(owner, np.uint8)
        rptr, rsucc, rw = (np.ascontiguousarray(a, np.int32) for a in (rptr, rsucc, rw))
        n, m = len(owner), int(rptr[-1])
        n0, mg = ctypes.c_int32(0), ctypes.c_int32(0)
        ext = np.empty(n, np.int32)
        N.check(N.egs().egs_solver_load_read_order(self.h, 0, n, m, owner.ctypes.data, rptr.ctypes.data,
                                                   rsucc.ctypes.data, rw.ctypes.data, ctypes.byref(n0),
                                                   ctypes.byref(mg), ext.ctypes.data), "egs_solver_load_read_order")
        self.n, self.m = n, m
        return int(n0.value), int(mg.value), ext

    def game(self):
        """The loaded game's internal CSR: (ptr, succ, w)."""
        ptr = np.empty(self.n + 1, np.int32)
        succ = np.empty(self.m, np.int32)
        w = np.empty(self.m, np.int32)
        N.check(N.egs().egs_solver_read_game(self.h, ptr.ctypes.data, succ.ctypes.data, w.ctypes.data),
                "egs_solver_read_game")
        return ptr, succ, w

    def load_device(self, n, n0, m, mg, d_ptr, d_succ, d_w):
        """Arrays already resident in device memory (integer device addresses)."""
        N.check(N.egs().egs_solver_load(self.h, 1, n, n0, m, mg, d_ptr, d_succ, d_w), "egs_solver_load")
        self.n, self.m = n, m

    def prepare(self, opt=None):
        opt = opt or options()
        N.check(N.egs().egs_solver_prepare(self.h, ctypes.byref(opt)), "egs_solver_prepare")

    def iterate(self, max_phases=-1):
        st = N.EgsStats()
        rc = N.check(N.egs().egs_solver_iterate(self.h, int(max_phases), ctypes.byref(st)), "egs_solver_iterate")
        return bool(rc), st

    def values(self):
        out = np.empty(self.n, np.int32)
        N.check(N.egs().egs_solver_read_values(self.h, out.ctypes.data), "egs_solver_read_values")
        return out

    def winning(self):
        """Player 0's winning set, {v : value < TOP}, as a bool array (internal numbering)."""
        words = np.empty((self.n + 31) // 32, np.uint32)
        N.check(N.egs().egs_solver_read_winning(self.h, words.ctypes.data), "egs_solver_read_winning")
        return np.unpackbits(words.view(np.uint8), bitorder="little")[: self.n].astype(bool)

    def csc(self):
        cptr = np.empty(self.n + 1, np.int32)
        cpred = np.empty(self.m, np.int32)
        N.check(N.egs().egs_solver_read_csc(self.h, cptr.ctypes.data, cpred.ctypes.data), "egs_solver_read_csc")
        return cptr, cpred

    def violations(self):
        v = ctypes.c_int64(0)
        N.check(N.egs().egs_solver_count_violations(self.h, ctypes.byref(v)), "egs_solver_count_violations")
        return int(v.value)

    def set_values(self, values):
        """Replace the prepared solver's values (for example to certify another path's result)."""
        v = np.ascontiguousarray(values, np.int32)
        if v.shape != (self.n,):
            raise ValueError(f"need {self.n} values, got {v.shape}")
        N.check(N.egs().egs_solver_set_values(self.h, v.ctypes.data), "egs_solver_set_values")

    def top_certificate(self, max_passes=0):
        """Potentials certifying that every TOP vertex of the current values is won by player 1
        (int64[n], >= 0 exactly on the TOP set, -1 elsewhere).  Returns (certified, h, passes);
        check them on the host with egsolver_amd.verify.check_top_certificate."""
        h = np.empty(self.n, np.int64)
        passes = ctypes.c_int32(0)
        rc = N.check(N.egs().egs_solver_top_certificate(self.h, int(max_passes), h.ctypes.data, ctypes.byref(passes)),
                     "egs_solver_top_certificate")
        return bool(rc), h, int(passes.value)

    def stats(self):
        st = N.EgsStats()
        N.check(N.egs().egs_solver_stats(self.h, ctypes.byref(st)), "egs_solver_stats")
        return st


def solve(g: Game, device=0, **opt):
    """Solve one game to its fixpoint; returns (values int32[n], stats dict)."""
    values = np.empty(g.n, np.int32)
    st = N.EgsStats()
    o = options(**opt)
    ptr = np.ascontiguousarray(g.ptr, np.int32)
    succ = np.ascontiguousarray(g.succ, np.int32)
    w = np.ascontiguousarray(g.w, np.int32)
    N.check(N.egs().egs_solve(device, g.n, g.n0, g.m, g.mg, ptr.ctypes.data, succ.ctypes.data, w.ctypes.data,
                              values.ctypes.data, ctypes.byref(o), ctypes.byref(st)), "egs_solve")
    return values, st.as_dict()


def device_info(device=0):
    info = N.EgsDeviceInfo()
    N.check(N.egs().egs_device_query(device, ctypes.byref(info)), "egs_device_query")
    return info

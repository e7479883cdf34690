"""Host-side verification of a result (``libegs_host.so``, no GPU code).

``check_fixpoint`` counts vertices with lift(f)(v) != f(v); zero puts f at or above the
least fixpoint, so the least fixpoint's TOP set lies inside {f == TOP}.
``check_top_certificate`` checks potentials from ``Solver.top_certificate`` that prove
{f == TOP} is won by player 1, so it lies inside the least fixpoint's TOP set.  Both pass
exactly when the TOP set of f is the least fixpoint's.  ``certify`` runs the pair on a result
of any path (single game, batch or server) through a prepared solver holding the same game.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N

CERT_CODES = {0: "ok", 1: "h >= 0 does not match value == TOP", 2: "an owner-0 vertex of T has a successor outside T",
              3: "an owner-0 edge inside T breaks h(v) >= (n+1)w + 1 + h(u)",
              4: "an owner-1 vertex of T has no edge into T with h(v) >= (n+1)w + 1 + h(u)"}


def _arr(x, dt):
    return np.ascontiguousarray(x, dt)


def check_fixpoint(g, values):
    """(violations, first violating vertex or -1)."""
    ptr, succ, w, f = _arr(g.ptr, np.int32), _arr(g.succ, np.int32), _arr(g.w, np.int32), _arr(values, np.int32)
    first = ctypes.c_int64(-1)
    bad = N.host().egs_check_fixpoint(g.n, g.n0, g.mg, ptr.ctypes.data, succ.ctypes.data, w.ctypes.data,
                                      f.ctypes.data, ctypes.byref(first))
    return int(bad), int(first.value)


def check_top_certificate(g, values, h):
    """(code, vertex): code 0 when h certifies that every TOP vertex of `values` is won by
    player 1, else the first failure (see CERT_CODES) and its vertex."""
    ptr, succ, w, f = _arr(g.ptr, np.int32), _arr(g.succ, np.int32), _arr(g.w, np.int32), _arr(values, np.int32)
    h = _arr(h, np.int64)
    if f.shape != (g.n,) or h.shape != (g.n,):
        raise ValueError("values and h need one entry per vertex")
    bad = ctypes.c_int64(-1)
    code = N.host().egs_check_top_certificate(g.n, g.n0, ptr.ctypes.data, succ.ctypes.data, w.ctypes.data,
                                              f.ctypes.data, h.ctypes.data, ctypes.byref(bad))
    return int(code), int(bad.value)


def certify(solver, g, values, max_passes=0):
    """Certificate of `values` (a result of any path for game g, loaded and prepared in
    `solver`) computed on the device and checked on the host, plus the host fixpoint check.
    Returns a dict; ``ok`` is True when both checks pass."""
    solver.set_values(values)
    found, h, passes = solver.top_certificate(max_passes)
    code, vertex = check_top_certificate(g, values, h)
    viol, first = check_fixpoint(g, values)
    return {"ok": found and code == 0 and viol == 0, "certificate_found": found, "certificate_code": code,
            "certificate_vertex": vertex, "passes": passes, "fixpoint_violations": viol, "first_violation": first,
            "top_vertices": int((np.asarray(values) == 2147483647).sum())}

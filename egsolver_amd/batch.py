"""Many games in one persistent launch (libegs.so egs_batch_*).

No reference counterpart: the reference solves one game per process.  Each game's values
are the same least fixpoint the single-game path and the reference's cpu_solver.c give.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N
from .games import Game


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("threads", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("lds_ints", ctypes.c_int32),
        ("lanes_per_edge_x16", ctypes.c_int32),
        ("group_min", ctypes.c_int32),
        ("group_max", ctypes.c_int32),
        ("max_sweeps", ctypes.c_int64),
        ("certify_every", ctypes.c_int32),
        ("cert_max_passes", ctypes.c_int32),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("games", ctypes.c_int64),
        ("sweeps", ctypes.c_int64),
        ("edges_relaxed", ctypes.c_int64),
        ("lifted", ctypes.c_int64),
        ("converged", ctypes.c_int64),
        ("ms", ctypes.c_double),
        ("packed_edges", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("threads", ctypes.c_int32),
        ("lds_ints", ctypes.c_int32),
        ("certificates", ctypes.c_int64),
        ("cert_passes", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = ctypes.c_void_p
N.register_signatures({
    "egs_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "egs_batch_destroy": (ctypes.c_int, [_vp]),
    "egs_batch_clear": (ctypes.c_int, [_vp]),
    "egs_batch_add": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                     _vp, _vp, _vp]),
    "egs_batch_size": (ctypes.c_int, [_vp]),
    "egs_batch_upload": (ctypes.c_int, [_vp]),
    "egs_batch_solve": (ctypes.c_int, [_vp, ctypes.POINTER(BatchOptions), ctypes.POINTER(BatchStats)]),
    "egs_batch_read_values": (ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    "egs_batch_read_all": (ctypes.c_int, [_vp, _vp]),
    "egs_batch_read_all_pinned": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_int32))]),
    "egs_batch_game_stats": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64),
                                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "egs_batch_last_error": (ctypes.c_char_p, []),
})


def _check(rc, what):
    if rc < 0:
        msg = N.egs().egs_batch_last_error()
        raise N.EgsError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def batch_options(threads=0, blocks_per_cu=0, lds_ints=0, lanes_per_edge_x16=0, group_min=0, group_max=0,
                  max_sweeps=0, certify_every=0, cert_max_passes=0):
    return BatchOptions(threads, blocks_per_cu, lds_ints, lanes_per_edge_x16, group_min, group_max, max_sweeps,
                        certify_every, cert_max_passes)


class Batch:
    def __init__(self, device=0):
        h = _vp()
        _check(N.egs().egs_batch_create(device, ctypes.byref(h)), "egs_batch_create")
        self.h = h
        self.sizes = []

    def close(self):
        if self.h:
            N.egs().egs_batch_destroy(self.h)
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

    def add(self, g: Game) -> int:
        ptr = np.ascontiguousarray(g.ptr, np.int32)
        succ = np.ascontiguousarray(g.succ, np.int32)
        w = np.ascontiguousarray(g.w, np.int32)
        i = _check(N.egs().egs_batch_add(self.h, g.n, g.n0, g.m, g.mg, ptr.ctypes.data, succ.ctypes.data,
                                         w.ctypes.data), "egs_batch_add")
        self.sizes.append(g.n)
        return i

    def clear(self):
        _check(N.egs().egs_batch_clear(self.h), "egs_batch_clear")
        self.sizes = []

    def upload(self):
        _check(N.egs().egs_batch_upload(self.h), "egs_batch_upload")

    def solve(self, opt=None):
        opt = opt or batch_options()
        st = BatchStats()
        _check(N.egs().egs_batch_solve(self.h, ctypes.byref(opt), ctypes.byref(st)), "egs_batch_solve")
        return st.as_dict()

    def values(self, i):
        out = np.empty(self.sizes[i], np.int32)
        _check(N.egs().egs_batch_read_values(self.h, i, out.ctypes.data), "egs_batch_read_values")
        return out

    def all_values(self):
        out = np.empty(sum(self.sizes), np.int32)
        _check(N.egs().egs_batch_read_all(self.h, out.ctypes.data), "egs_batch_read_all")
        offs = np.concatenate([[0], np.cumsum(self.sizes)])
        return [out[offs[i]:offs[i + 1]] for i in range(len(self.sizes))]

    def values_view(self):
        """Every game's values as views into a pinned buffer the batch owns (one device copy,
        no host copy); valid until the next solve or read of this batch."""
        ptr = ctypes.POINTER(ctypes.c_int32)()
        _check(N.egs().egs_batch_read_all_pinned(self.h, ctypes.byref(ptr)), "egs_batch_read_all_pinned")
        total = sum(self.sizes)
        out = np.ctypeslib.as_array(ptr, shape=(total,)) if total else np.zeros(0, np.int32)
        offs = np.concatenate([[0], np.cumsum(self.sizes)])
        return [out[offs[i]:offs[i + 1]] for i in range(len(self.sizes))]

    def game_stats(self, i):
        s, e, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _check(N.egs().egs_batch_game_stats(self.h, i, ctypes.byref(s), ctypes.byref(e), ctypes.byref(c)),
               "egs_batch_game_stats")
        return {"sweeps": s.value, "edges": e.value, "converged": bool(c.value)}

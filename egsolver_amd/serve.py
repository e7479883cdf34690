"""Serving loop over libegs.so's egs_server_*: continuous batching of energy-game requests.

Requests arrive over time (a Poisson process in the benchmark).  The loop admits every
arrived request whose device footprint fits the arena, runs one time slice of the
many-games kernel over all resident games, hands back the games that reached their
fixpoint, and repeats.  Uploads of newly admitted games overlap the running slice.
No reference counterpart (the reference solves one game per process).
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .games import Game

_vp = ctypes.c_void_p


class ServeOptions(ctypes.Structure):
    _fields_ = [
        ("threads", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("lds_ints", ctypes.c_int32),
        ("lanes_per_edge_x16", ctypes.c_int32),
        ("group_min", ctypes.c_int32),
        ("group_max", ctypes.c_int32),
        ("slice_sweeps", ctypes.c_int32),
        ("certify_every", ctypes.c_int32),
        ("cert_max_passes", ctypes.c_int32),
        ("slice_kedges", ctypes.c_int32),
        ("large_cus", ctypes.c_int32),
    ]


N.register_signatures({
    "egs_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_vp)]),
    "egs_server_destroy": (ctypes.c_int, [_vp]),
    "egs_server_footprint": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32]),
    "egs_server_free_bytes": (ctypes.c_int64, [_vp]),
    "egs_server_free_slots": (ctypes.c_int, [_vp]),
    "egs_server_set_route": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "egs_server_submit": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                         _vp, _vp, _vp]),
    "egs_server_submit_read_order": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp]),
    "egs_server_launch": (ctypes.c_int, [_vp, ctypes.POINTER(ServeOptions)]),
    "egs_server_wait": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "egs_server_last_slice_ms": (ctypes.c_double, [_vp]),
    "egs_server_ready": (ctypes.c_int, [_vp]),
    "egs_server_status": (ctypes.c_int, [_vp, ctypes.c_int]),
    "egs_server_fetch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64)]),
    "egs_server_last_error": (ctypes.c_char_p, []),
})

EGS_E_CAPACITY = -5


def _check(rc, what):
    if rc < 0:
        msg = N.egs().egs_server_last_error()
        raise N.EgsError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def serve_options(threads=0, blocks_per_cu=0, lds_ints=0, lanes_per_edge_x16=0, group_min=0, group_max=0,
                  slice_sweeps=0, certify_every=0, cert_max_passes=0, slice_kedges=0, large_cus=0):
    return ServeOptions(threads, blocks_per_cu, lds_ints, lanes_per_edge_x16, group_min, group_max, slice_sweeps,
                        certify_every, cert_max_passes, slice_kedges, large_cus)


class Server:
    def __init__(self, arena_bytes, max_games=4096, device=0):
        h = _vp()
        _check(N.egs().egs_server_create(device, int(arena_bytes), int(max_games), ctypes.byref(h)),
               "egs_server_create")
        self.h = h
        self.max_games = max_games
        self.n_of = {}
        self._done = (ctypes.c_int32 * max_games)()

    def close(self):
        if self.h:
            N.egs().egs_server_destroy(self.h)
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

    @staticmethod
    def footprint(g: Game) -> int:
        wmax = int(np.abs(g.w.astype(np.int64)).max()) if g.m else 0
        return int(N.egs().egs_server_footprint(g.n, g.m, wmax))

    def submit(self, g: Game):
        """Ticket, or None when the game does not fit right now."""
        ptr = np.ascontiguousarray(g.ptr, np.int32)
        succ = np.ascontiguousarray(g.succ, np.int32)
        w = np.ascontiguousarray(g.w, np.int32)
        rc = N.egs().egs_server_submit(self.h, g.n, g.n0, g.m, g.mg, ptr.ctypes.data, succ.ctypes.data,
                                       w.ctypes.data)
        if rc == EGS_E_CAPACITY:
            return None
        _check(rc, "egs_server_submit")
        self.n_of[rc] = g.n
        return rc

    def submit_read_order(self, owner, rptr, rsucc, rw):
        """A game in read order (as a .mpg file lists it; successors are read indices): the
        owner partition, internal CSR and MG are made on the device.  Ticket, or None when
        the game does not fit right now; fetch() returns its values in read order."""
        owner = np.ascontiguousarray(owner, np.uint8)
        rptr = np.ascontiguousarray(rptr, np.int32)
        rsucc = np.ascontiguousarray(rsucc, np.int32)
        rw = np.ascontiguousarray(rw, np.int32)
        n, m = len(owner), len(rsucc)
        rc = N.egs().egs_server_submit_read_order(self.h, n, m, owner.ctypes.data, rptr.ctypes.data,
                                                  rsucc.ctypes.data, rw.ctypes.data)
        if rc == EGS_E_CAPACITY:
            return None
        _check(rc, "egs_server_submit_read_order")
        self.n_of[rc] = n
        return rc

    def launch(self, opt=None):
        opt = opt or serve_options()
        return _check(N.egs().egs_server_launch(self.h, ctypes.byref(opt)), "egs_server_launch")

    def wait(self):
        k = _check(N.egs().egs_server_wait(self.h, self._done, self.max_games), "egs_server_wait")
        return [self._done[i] for i in range(min(k, self.max_games))]

    def ready(self):
        """True when wait() would return without blocking."""
        return bool(_check(N.egs().egs_server_ready(self.h), "egs_server_ready"))

    def last_slice_ms(self):
        return N.egs().egs_server_last_slice_ms(self.h)

    def fetch(self, ticket, want_values=True):
        n = self.n_of.pop(ticket)
        out = np.empty(n, np.int32) if want_values else None
        sw, ed = ctypes.c_int64(), ctypes.c_int64()
        _check(N.egs().egs_server_fetch(self.h, ticket, out.ctypes.data if want_values else None, ctypes.byref(sw),
                                        ctypes.byref(ed)), "egs_server_fetch")
        return out, {"sweeps": sw.value, "edges": ed.value}

    def free_bytes(self):
        return int(N.egs().egs_server_free_bytes(self.h))

    def set_route(self, large_n=0, large_workers=0):
        """Games with at least `large_n` vertices go to the device-wide route (0 = default,
        16384), solved `large_workers` at a time (0 = default, 2)."""
        _check(N.egs().egs_server_set_route(self.h, int(large_n), int(large_workers)), "egs_server_set_route")


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the start of the run
    game: Game
    admitted: float = -1.0
    finished: float = -1.0
    values: np.ndarray = field(default=None, repr=False)


def run_served(requests, arena_bytes, opt=None, max_games=4096, keep_values=False, device=0, large_n=0,
               large_workers=0, admit_batch=32):
    """Replay `requests` (sorted by arrival) against a Server; returns (requests,
    wall seconds, slice count).  Latency of a request = finished - arrival.

    Loop: admit up to `admit_batch` arrived requests, start a slice, keep admitting arrived
    requests while the slice runs, then collect what finished."""
    srv = Server(arena_bytes, max_games=max_games, device=device)
    srv.set_route(large_n, large_workers)
    waiting = deque(requests)
    running = {}
    slices = 0
    t0 = time.perf_counter()

    def admit(limit, while_busy=False):
        k = 0
        while waiting and k < limit and waiting[0].arrival <= time.perf_counter() - t0:
            if while_busy and srv.ready():
                break
            r = waiting[0]
            t = srv.submit(r.game)
            if t is None:
                break
            r.admitted = time.perf_counter() - t0
            running[t] = r
            waiting.popleft()
            k += 1
        return k

    try:
        while waiting or running:
            admit(admit_batch)
            if not running:
                # idle until the next arrival
                if waiting:
                    time.sleep(max(0.0, waiting[0].arrival - (time.perf_counter() - t0)))
                continue
            srv.launch(opt)
            slices += 1
            # admissions while the slice runs overlap its execution
            admit(max_games, while_busy=True)
            done = srv.wait()
            for t in done:
                r = running.pop(t)
                vals, _ = srv.fetch(t)           # every result is read back, as the plain path does
                r.finished = time.perf_counter() - t0
                if keep_values:
                    r.values = vals
        wall = time.perf_counter() - t0
    finally:
        srv.close()
    return requests, wall, slices


def run_plain(requests, solve_one, keep_values=False):
    """One game per launch, in arrival order: the baseline the server is compared with.
    `solve_one(game) -> values` runs the single-game path."""
    t0 = time.perf_counter()
    for r in requests:
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        r.admitted = time.perf_counter() - t0
        vals = solve_one(r.game)
        r.finished = time.perf_counter() - t0
        if keep_values:
            r.values = vals
    return requests, time.perf_counter() - t0


def latency_stats(requests):
    lat = np.array([r.finished - r.arrival for r in requests if r.finished >= 0])
    if lat.size == 0:
        return {"count": 0}
    return {"count": int(lat.size), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "max_ms": float(lat.max() * 1e3)}

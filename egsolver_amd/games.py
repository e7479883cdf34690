"""Game containers, the .mpg reader and the seeded generator (host side, numpy arrays).

A ``Game`` is always in the internal numbering of the reference: owner-0 vertices (input
token MAX, min-player) first, both blocks in read order (src/egsolver.c:146-214,253-281).
``ext_of_int`` maps back to external ids for output.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as N

TOP = 2147483647


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Game:
    n: int
    n0: int
    ptr: np.ndarray      # int32[n+1]
    succ: np.ndarray     # int32[m]
    w: np.ndarray        # int32[m]
    mg: int
    ext_of_int: np.ndarray  # int32[n]

    @property
    def m(self) -> int:
        return int(self.ptr[-1])

    @property
    def max_outdeg(self) -> int:
        return int(np.diff(self.ptr).max()) if self.n else 0


def compute_mg(n, ptr, w) -> int:
    out = ctypes.c_int32(0)
    rc = N.host().egs_compute_mg(int(n), _p(ptr), _p(w), ctypes.byref(out))
    if rc != 0:
        raise OverflowError("MG (sum of per-vertex worst negative weights) overflows int32")
    return int(out.value)


def internalise(owner, rptr, rsucc, rw) -> Game:
    """Owner partition of a game whose read order is its external-id order."""
    n = len(owner)
    m = int(rptr[-1])
    ptr = np.empty(n + 1, np.int32)
    succ = np.empty(m, np.int32)
    w = np.empty(m, np.int32)
    ext = np.empty(n, np.int32)
    owner = np.ascontiguousarray(owner, np.uint8)
    rptr = np.ascontiguousarray(rptr, np.int32)
    rsucc = np.ascontiguousarray(rsucc, np.int32)
    rw = np.ascontiguousarray(rw, np.int32)
    n0 = N.host().egs_internalise(n, _p(owner), _p(rptr), _p(rsucc), _p(rw), _p(ptr), _p(succ), _p(w), _p(ext))
    return Game(n=n, n0=int(n0), ptr=ptr, succ=succ, w=w, mg=compute_mg(n, ptr, w), ext_of_int=ext)


@dataclass
class GenParams:
    n: int
    seed: int
    deg_kind: int = 0          # 0 fixed, 1 uniform, 2 power law, 3 complete
    deg_lo: int = 4
    deg_hi: int = 4
    alpha: float = 2.1
    wmax: int = 10

    def c(self):
        return N.EgsGenParams(seed=self.seed, n=self.n, deg_kind=self.deg_kind, deg_lo=self.deg_lo,
                              deg_hi=self.deg_hi, alpha=self.alpha, wmax=self.wmax, reserved=0)


DEG_FIXED, DEG_UNIFORM, DEG_POWERLAW, DEG_COMPLETE = 0, 1, 2, 3
FAMILY_BATCH, FAMILY_SERVE = 4, 5


def generate_read_order(p: GenParams, stream: int = 0):
    """Arrays in external-id order: owner[n], rptr[n+1], rsucc[m], rw[m]."""
    cp = p.c()
    m = N.host().egs_gen_count_edges(ctypes.byref(cp), stream)
    owner = np.empty(p.n, np.uint8)
    rptr = np.empty(p.n + 1, np.int32)
    rsucc = np.empty(m, np.int32)
    rw = np.empty(m, np.int32)
    ro = N.EgsGenReadOrder(owner=owner.ctypes.data_as(N._u8p), ptr=rptr.ctypes.data_as(N._i32p),
                           succ=rsucc.ctypes.data_as(N._i32p), w=rw.ctypes.data_as(N._i32p))
    got = N.host().egs_gen_game(ctypes.byref(cp), stream, ctypes.byref(ro), 0)
    assert got == m
    return owner, rptr, rsucc, rw


def generate(p: GenParams, stream: int = 0) -> Game:
    return internalise(*generate_read_order(p, stream))


def family_params(family: int, seed: int, index: int) -> GenParams:
    cp = N.EgsGenParams()
    N.host().egs_gen_family_params(family, seed, index, ctypes.byref(cp))
    return GenParams(n=cp.n, seed=cp.seed, deg_kind=cp.deg_kind, deg_lo=cp.deg_lo, deg_hi=cp.deg_hi,
                     alpha=cp.alpha, wmax=cp.wmax)


# The workloads the benchmark and the parity tests are quoted on.
def config_game(which: int, scale: float = 1.0) -> Game:
    """Seeded single games: 1 = n 2000, degree 4, W 10, seed 1; 2 = n 2^22, degree 8,
    W 100, seed 2; 3 = n 2^20, power-law degrees (alpha 2.1, 1..1e5), W 100, seed 3.
    `scale` shrinks n for scaled-down parity runs of the same distribution."""
    if which == 1:
        p = GenParams(n=max(8, int(2000 * scale)), seed=1, deg_kind=DEG_FIXED, deg_lo=4, deg_hi=4, wmax=10)
    elif which == 2:
        p = GenParams(n=max(16, int((1 << 22) * scale)), seed=2, deg_kind=DEG_FIXED, deg_lo=8, deg_hi=8, wmax=100)
    elif which == 3:
        p = GenParams(n=max(16, int((1 << 20) * scale)), seed=3, deg_kind=DEG_POWERLAW, deg_lo=1,
                      deg_hi=100000, alpha=2.1, wmax=100)
    else:
        raise ValueError(which)
    return generate(p)


def write_mpg(path, owner, rptr, rsucc, rw, gz=False):
    owner = np.ascontiguousarray(owner, np.uint8)
    rc = N.host().egs_write_mpg(path.encode(), len(owner), _p(owner), _p(np.ascontiguousarray(rptr, np.int32)),
                                _p(np.ascontiguousarray(rsucc, np.int32)), _p(np.ascontiguousarray(rw, np.int32)),
                                1 if gz else 0)
    if rc != 0:
        raise OSError(f"cannot write {path}")


def _read_into(g, err, path, text, weight_mode, rnd_lo, rnd_hi, threads):
    if text is not None:
        data = text.encode() if isinstance(text, str) else bytes(text)
        if threads == 1:
            return N.host().egs_mpg_read_mem(data, len(data), weight_mode, rnd_lo, rnd_hi, ctypes.byref(g), err, 512)
        return N.host().egs_mpg_read_mem_par(data, len(data), weight_mode, rnd_lo, rnd_hi, ctypes.byref(g), err, 512,
                                             threads)
    return N.host().egs_mpg_read(path.encode() if path else None, weight_mode, rnd_lo, rnd_hi, ctypes.byref(g), err, 512)


def parse_mpg(text, weight_mode=0, rnd_lo=-1, rnd_hi=1, threads=1) -> dict:
    """The reader's result in read order, before the owner partition: threads=1 runs the
    sequential reader, anything else the parallel one (0 = OpenMP's default)."""
    g = N.EgsMpg()
    err = ctypes.create_string_buffer(512)
    rc = _read_into(g, err, None, text, weight_mode, rnd_lo, rnd_hi, threads)
    if rc != 0:
        raise ValueError(err.value.decode())
    try:
        n, m = g.n, g.m
        arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, np.int32)
        out = {"n": n, "m": m, "ext_of_read": arr(g.ext_of_read, n), "owner": arr(g.owner, n),
               "rptr": arr(g.rptr, n + 1), "rsucc": arr(g.rsucc, m), "rw": arr(g.rw, m),
               "max_abs_w": g.max_abs_w, "min_outdeg": g.min_outdeg, "max_outdeg": g.max_outdeg,
               "low_outdeg": list(g.low_outdeg),
               "names": [g.names[i].decode() if g.names[i] else None for i in range(n)],
               "chunks": g.chunks}
    finally:
        N.host().egs_mpg_free(ctypes.byref(g))
    return out


def read_mpg(path=None, text=None, weight_mode=0, rnd_lo=-1, rnd_hi=1, threads=1) -> Game:
    """Read a .mpg arena with the package's C reader and internalise it.  A file is read by
    the parallel reader; `text` by the sequential one unless `threads` != 1."""
    g = N.EgsMpg()
    err = ctypes.create_string_buffer(512)
    if text is not None:
        rc = _read_into(g, err, None, text, weight_mode, rnd_lo, rnd_hi, threads)
    else:
        rc = N.host().egs_mpg_read(path.encode() if path else None, weight_mode, rnd_lo, rnd_hi,
                                   ctypes.byref(g), err, 512)
    if rc != 0:
        raise ValueError(err.value.decode())
    try:
        n, m = g.n, g.m
        mapping = np.empty(n, np.int32)
        ptr = np.empty(n + 1, np.int32)
        succ = np.empty(m, np.int32)
        w = np.empty(m, np.int32)
        n0 = N.host().egs_mpg_internalise(ctypes.byref(g), _p(mapping), _p(ptr), _p(succ), _p(w))
        ext_of_read = np.ctypeslib.as_array(g.ext_of_read, shape=(n,)).copy()
        ext = ext_of_read[mapping]
    finally:
        N.host().egs_mpg_free(ctypes.byref(g))
    return Game(n=n, n0=int(n0), ptr=ptr, succ=succ, w=w, mg=compute_mg(n, ptr, w), ext_of_int=ext)


def save_csr(g: Game, path: str):
    """Binary cache of a game in its internal numbering (egs_csr_save)."""
    err = ctypes.create_string_buffer(512)
    arrs = [np.ascontiguousarray(a, np.int32) for a in (g.ptr, g.succ, g.w, g.ext_of_int)]
    rc = N.host().egs_csr_save(path.encode(), g.n, g.n0, g.m, g.mg, *[_p(a) for a in arrs], err, 512)
    if rc != 0:
        raise OSError(err.value.decode())


def load_csr(path: str) -> Game:
    """A game saved by save_csr; raises ValueError on a damaged or foreign file."""
    c = N.EgsCsr()
    err = ctypes.create_string_buffer(512)
    if N.host().egs_csr_load(path.encode(), ctypes.byref(c), err, 512) != 0:
        raise ValueError(err.value.decode())
    try:
        n, m = c.n, c.m
        arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, np.int32)
        g = Game(n=n, n0=c.n0, ptr=arr(c.ptr, n + 1), succ=arr(c.succ, m), w=arr(c.w, m), mg=c.mg,
                 ext_of_int=arr(c.ext_of_int, n))
    finally:
        N.host().egs_csr_free(ctypes.byref(c))
    return g

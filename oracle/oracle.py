"""CPU oracle for the energy-game fixpoint -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` import
this module.  It is the checker, never the thing measured or shipped.

Contents
  * ``parse_mpg``     -- an independent restatement of the reference's ``.mpg`` reader
                         (src/the_lex.l:33-52, src/the_parser.y:63-102,128-295) plus its
                         owner partition and internal CSR build (src/egsolver.c:107-281).
  * ``Oracle``        -- ctypes front end of ``_build/libegs_oracle.so`` (egs_oracle.c, the
                         C restatement of src/cpu_solver.c, src/host_csr2csc.c and the MG /
                         loop-bound arithmetic of src/egsolver.c).
  * ``Reference``     -- ctypes front end of ``_ref/libegs_ref.so``: the reference's own
                         cpu_solver.c compiled from its sources by oracle/Makefile.
  * ``kleene``        -- brute-force synchronous Kleene iteration for tiny games.

Parity status (what pins each function) is listed in the header of egs_oracle.c.
"""
from __future__ import annotations

import ctypes
import gzip
import os
import re
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TOP = 2147483647

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


# ----------------------------------------------------------------------------------------
# .mpg reader (restatement of the flex/bison grammar)
# ----------------------------------------------------------------------------------------

# Token classes of src/the_lex.l:33-52.  At each position the scanner takes the first
# alternative that matches; characters no rule matches are skipped, as the lexer's
# catch-all "." rule discards them.  Blanks and newlines are skipped the same way.
_TOKEN_RE = re.compile(
    r'[0-9]+\.[0-9]+|[0-9]+|"(?:\\.|[^"])*"|arena|MAX|MIN|[;:,\-]|(?:#|%|//)[^\n]*'
)


class MpgError(ValueError):
    pass


@dataclass
class Game:
    """A game in the reference's internal numbering (owner-0 vertices first)."""

    n: int
    n0: int
    ptr: np.ndarray          # int32[n+1]
    succ: np.ndarray         # int32[m]
    w: np.ndarray            # int32[m]
    ext_of_int: np.ndarray   # int32[n]: external id of internal vertex i
    mg: int = 0
    max_abs_w: int = 0
    names: list = field(default_factory=list)

    @property
    def m(self) -> int:
        return int(self.ptr[-1])


def _kind(tok: str) -> str:
    c = tok[0]
    if c.isdigit():
        return "float" if "." in tok else "int"
    if c == '"':
        return "string"
    if c in "#%/":
        return "comment"
    return {";": "semi", ":": "col", ",": "comma", "-": "minus", "arena": "arena",
            "MAX": "max", "MIN": "min"}[tok]


def _a2int(s: str) -> int:
    v = int(s)
    if v > 2147483647:
        raise MpgError("CONVERSION ERROR out-of-range value in input")
    return v


def parse_mpg_text(text: str) -> Game:
    """Parse one arena.  Grammar of src/the_parser.y:121-160:

        header : [arena] INT [INT] ';'
        line   : INT (MAX|MIN) edge (',' edge)* [STRING] ';'
        edge   : INT ':' ['-'] INT

    MAX means owner 0 and MIN owner 1 (the_parser.y:145-146).  Each id in 0..max must
    appear exactly once (the_parser.y:237-247, 283-300); successors must be <= max
    (the_parser.y:317-336)."""
    toks = [t for t in _TOKEN_RE.findall(text) if t[0] not in "#%/"]
    kinds = [_kind(t) for t in toks]
    ntok = len(toks)
    pos = 0

    def fail(what):
        got = kinds[pos] if pos < ntok else "end of input"
        raise MpgError(f"syntax error: expected {what}, got {got}")

    if pos < ntok and kinds[pos] == "arena":
        pos += 1
    if pos >= ntok or kinds[pos] != "int":
        fail("int")
    max_id = _a2int(toks[pos]); pos += 1
    if pos < ntok and kinds[pos] == "int":
        pos += 1
    if pos >= ntok or kinds[pos] != "semi":
        fail(";")
    pos += 1
    n = max_id + 1

    read_ext, read_owner, rptr, rsucc, rw, names = [], [], [0], [], [], []
    int_of_ext = np.full(n + 1, -1, dtype=np.int64)
    while pos < ntok:
        if kinds[pos] != "int":
            fail("vertex id")
        ext = _a2int(toks[pos]); pos += 1
        k = kinds[pos] if pos < ntok else None
        if k == "max":
            owner = 0
        elif k == "min":
            owner = 1
        else:
            fail("MAX or MIN")
        pos += 1
        if ext > max_id:
            raise MpgError(f"Unexpected node identifier: {ext} greater than {max_id}")
        if int_of_ext[ext] != -1:
            raise MpgError(f"Repeated node identifier: {ext}")
        int_of_ext[ext] = len(read_ext)
        read_ext.append(ext)
        read_owner.append(owner)
        while True:
            if pos + 2 >= ntok or kinds[pos] != "int" or kinds[pos + 1] != "col":
                fail("succ:weight")
            s = _a2int(toks[pos])
            if kinds[pos + 2] == "minus":
                if pos + 3 >= ntok or kinds[pos + 3] != "int":
                    fail("weight")
                wv = -_a2int(toks[pos + 3]); pos += 4
            elif kinds[pos + 2] == "int":
                wv = _a2int(toks[pos + 2]); pos += 3
            else:
                fail("weight")
            if s > max_id:
                raise MpgError(f"Unexpected successor identifier: {s}")
            rsucc.append(s)
            rw.append(wv)
            if pos < ntok and kinds[pos] == "comma":
                pos += 1
                continue
            break
        if pos < ntok and kinds[pos] == "string":
            names.append(toks[pos][1:-1]); pos += 1
        else:
            names.append(None)
        if pos >= ntok or kinds[pos] != "semi":
            fail(";")
        pos += 1
        rptr.append(len(rsucc))
    if len(read_ext) != n:
        raise MpgError(f"Expecting {n} nodes, read {len(read_ext)} nodes")

    return _internalise(n, np.asarray(read_ext, np.int64), np.asarray(read_owner, np.uint8),
                        np.asarray(rptr, np.int64), np.asarray(rsucc, np.int64),
                        np.asarray(rw, np.int64), int_of_ext, names)


def _internalise(n, read_ext, read_owner, rptr, rsucc, rw, int_of_ext, names) -> Game:
    """Owner partition and internal CSR: src/egsolver.c:146-214 (qsort with prima0poi1,
    stable under glibc's merge sort) and src/egsolver.c:253-281."""
    order0 = np.nonzero(read_owner == 0)[0]
    order1 = np.nonzero(read_owner != 0)[0]
    mapping = np.concatenate([order0, order1])          # internal -> read index
    revmapping = np.empty(n, np.int64)
    revmapping[mapping] = np.arange(n)
    deg = np.diff(rptr)[mapping]
    ptr = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=ptr[1:])
    m = int(ptr[-1])
    idx = np.arange(m, dtype=np.int64) + np.repeat(rptr[mapping] - ptr[:-1], deg)
    succ = revmapping[int_of_ext[rsucc[idx]]]
    w = rw[idx]
    g = Game(n=n, n0=len(order0), ptr=ptr.astype(np.int32), succ=succ.astype(np.int32),
             w=w.astype(np.int32), ext_of_int=read_ext[mapping].astype(np.int32),
             names=[names[r] for r in mapping] if names else [])
    g.mg = mg_of(g)
    g.max_abs_w = int(np.abs(g.w.astype(np.int64)).max()) if g.m else 0
    return g


def parse_mpg(path: str) -> Game:
    with open(path, "rb") as f:
        head = f.read(2)
    opener = gzip.open if head == b"\x1f\x8b" else open
    with opener(path, "rb") as f:
        text = f.read().decode("latin-1")
    return parse_mpg_text(text)


def mg_of(g: Game) -> int:
    """MG = sum_v max(0, max_e -w_e) (src/egsolver.c:283-304), numpy form."""
    if g.m == 0:
        return 0
    neg = np.maximum(-g.w.astype(np.int64), 0)
    deg = np.diff(g.ptr.astype(np.int64))
    starts = g.ptr[:-1].astype(np.int64)
    worst = np.zeros(g.n, np.int64)
    nz = deg > 0
    worst[nz] = np.maximum.reduceat(neg, starts[nz])
    total = int(worst.sum())
    if total > TOP:
        raise MpgError("Error too high weights --> overflow")
    return total


def game_from_arrays(n, n0, ptr, succ, w, ext_of_int=None) -> Game:
    g = Game(n=int(n), n0=int(n0), ptr=np.ascontiguousarray(ptr, np.int32),
             succ=np.ascontiguousarray(succ, np.int32), w=np.ascontiguousarray(w, np.int32),
             ext_of_int=(np.arange(n, dtype=np.int32) if ext_of_int is None
                         else np.ascontiguousarray(ext_of_int, np.int32)))
    g.mg = mg_of(g)
    g.max_abs_w = int(np.abs(g.w.astype(np.int64)).max()) if g.m else 0
    return g


def format_solution(g: Game, val) -> str:
    """One line per vertex in internal order: "%d)\\t%d\\n" (src/egsolver.c:100-104)."""
    return "".join(f"{int(e)})\t{int(v)}\n" for e, v in zip(g.ext_of_int, val))


# ----------------------------------------------------------------------------------------
# ctypes front ends
# ----------------------------------------------------------------------------------------

def _lib(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run: make -C {HERE})")
    return ctypes.CDLL(path)


class Oracle:
    """C restatement (egs_oracle.c)."""

    def __init__(self, path=os.path.join(HERE, "_build", "libegs_oracle.so")):
        L = self.L = _lib(path)
        L.ora_ominus.argtypes = [ctypes.c_int] * 3
        L.ora_ominus.restype = ctypes.c_int
        L.ora_max_loop.argtypes = [ctypes.c_long] * 4
        L.ora_max_loop.restype = ctypes.c_long
        L.ora_mg.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.POINTER(ctypes.c_int)]
        L.ora_mg.restype = ctypes.c_int
        L.ora_csr2csc.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_void_p,
                                  _i32p, _i32p, ctypes.c_void_p]
        L.ora_initial_active.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _u8p]
        L.ora_initial_active.restype = ctypes.c_int
        L.ora_eg.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                             _i32p, _i32p, ctypes.c_long, ctypes.c_void_p]
        L.ora_eg.restype = ctypes.c_long
        L.ora_eg0.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                              _i32p, _i32p, ctypes.c_long, ctypes.c_int, ctypes.c_void_p]
        L.ora_eg0.restype = ctypes.c_long
        L.ora_fixpoint_violations.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                              _i32p, _i32p]
        L.ora_fixpoint_violations.restype = ctypes.c_long
        L.ora_eg0_from.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p,
                                   ctypes.c_long, ctypes.c_int]
        L.ora_eg0_from.restype = ctypes.c_long
        L.ora_lift_all.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p,
                                   _i32p, _i32p]

    def ominus(self, a, w, mg):
        return self.L.ora_ominus(a, w, mg)

    def max_loop(self, m, n, mg, maxloop=-1):
        return self.L.ora_max_loop(m, n, mg, maxloop)

    def mg(self, g: Game):
        out = ctypes.c_int(0)
        rc = self.L.ora_mg(g.n, g.ptr, g.w, ctypes.byref(out))
        if rc != 0:
            raise OverflowError("MG overflows int32")
        return out.value

    def csr2csc(self, g: Game, with_weights=False):
        cptr = np.zeros(g.n + 1, np.int32)
        cpred = np.zeros(max(g.m, 1), np.int32)
        cw = np.zeros(max(g.m, 1), np.int32) if with_weights else None
        self.L.ora_csr2csc(g.n, g.m, g.ptr, g.succ, g.w.ctypes.data, cptr, cpred,
                           cw.ctypes.data if cw is not None else None)
        return (cptr, cpred[:g.m], cw[:g.m]) if with_weights else (cptr, cpred[:g.m])

    def initial_active(self, g: Game):
        flags = np.zeros(g.n, np.uint8)
        self.L.ora_initial_active(g.n, g.n0, g.ptr, g.w, flags)
        return flags

    def eg(self, g: Game, max_loop=-1, mg=None):
        val = np.zeros(g.n, np.int32)
        processed = self.L.ora_eg(g.n, g.n0, g.m, g.mg if mg is None else mg, g.ptr, g.succ, g.w,
                                  val, max_loop, None)
        return val, processed

    def eg0(self, g: Game, max_loop=-1, threads=0, mg=None):
        val = np.zeros(g.n, np.int32)
        sweeps = self.L.ora_eg0(g.n, g.n0, g.m, g.mg if mg is None else mg, g.ptr, g.succ, g.w,
                                val, max_loop, threads, None)
        return val, sweeps

    def eg0_from(self, g: Game, start, max_sweeps=-1, threads=0):
        """EG0 Jacobi sweeps from `start` (TOP on a set already proved player 1's, 0 elsewhere,
        gives the least fixpoint).  Returns (values, sweeps)."""
        start = np.ascontiguousarray(start, np.int32)
        out = np.zeros(g.n, np.int32)
        sweeps = self.L.ora_eg0_from(g.n, g.n0, g.mg, g.ptr, g.succ, g.w, start, out, max_sweeps, threads)
        return out, sweeps

    def fixpoint_violations(self, g: Game, val, mg=None):
        val = np.ascontiguousarray(val, np.int32)
        return self.L.ora_fixpoint_violations(g.n, g.n0, g.mg if mg is None else mg, g.ptr,
                                              g.succ, g.w, val)

    def lift_all(self, g: Game, val):
        out = np.zeros(g.n, np.int32)
        self.L.ora_lift_all(g.n, g.n0, g.mg, g.ptr, g.succ, g.w,
                            np.ascontiguousarray(val, np.int32), out)
        return out


class Reference:
    """The reference's own cpu_solver.c / host_csr2csc.c, built by oracle/Makefile."""

    PATH = os.path.join(HERE, "_ref", "libegs_ref.so")

    @classmethod
    def available(cls):
        return os.path.exists(cls.PATH)

    def __init__(self, path=None):
        L = self.L = _lib(path or self.PATH)
        L.ref_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _i32p, _i32p, _i32p, _i32p, ctypes.c_long, ctypes.c_int,
                                ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.ref_solve.restype = ctypes.c_long
        L.ref_csr2csc.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p]

    def solve(self, g: Game, algo="eg", max_loop=-1, threads=0, quiet=True):
        val = np.zeros(g.n, np.int32)
        secs = ctypes.c_double(0.0)
        processed = self.L.ref_solve(1 if algo == "eg0" else 0, g.n, g.n0, g.m, g.mg, g.ptr,
                                     g.succ, g.w, val, max_loop, threads, 1 if quiet else 0,
                                     ctypes.byref(secs))
        return val, processed, secs.value

    def csr2csc(self, n, m, ptr, succ, w):
        cptr = np.zeros(n + 1, np.int32)
        cpred = np.zeros(max(m, 1), np.int32)
        cw = np.zeros(max(m, 1), np.int32)
        self.L.ref_csr2csc(n, m, np.ascontiguousarray(ptr, np.int32), np.ascontiguousarray(succ, np.int32),
                           np.ascontiguousarray(w, np.int32), cptr, cpred, cw)
        return cptr, cpred[:m], cw[:m]


# ----------------------------------------------------------------------------------------
# brute force
# ----------------------------------------------------------------------------------------

def kleene(g: Game, mg=None, max_iter=10_000_000):
    """Synchronous Kleene iteration f_{k+1} = lift(f_k) from f_0 = 0 in pure Python,
    straight from the definition of the least fixpoint.  Tiny games only."""
    mg = g.mg if mg is None else mg
    f = [0] * g.n
    ptr, succ, w = g.ptr.tolist(), g.succ.tolist(), g.w.tolist()
    for _ in range(max_iter):
        nf = []
        for v in range(g.n):
            xs = []
            for e in range(ptr[v], ptr[v + 1]):
                a = f[succ[e]]
                xs.append(max(0, a - w[e]) if (a < TOP and a - w[e] <= mg) else TOP)
            nf.append(min(xs) if v < g.n0 else max(xs))
        if nf == f:
            return np.asarray(f, np.int32)
        f = nf
    raise RuntimeError("kleene: no fixpoint within max_iter")

"""Host side without a GPU: the C .mpg reader against the oracle's independent reader,
the owner partition, MG, the generator, the exported C-ABI and the CLI surface."""
import ctypes
import gzip
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

import egsolver_amd as E
from egsolver_amd import _native as N
from oracle.oracle import MpgError, parse_mpg, parse_mpg_text


def _same(a, b):
    assert a.n == b.n and a.n0 == b.n0 and a.mg == b.mg
    assert np.array_equal(a.ptr, b.ptr)
    assert np.array_equal(a.succ, b.succ)
    assert np.array_equal(a.w, b.w)
    assert np.array_equal(a.ext_of_int, b.ext_of_int)


@pytest.mark.parametrize("name", ["eg_fig2.txt", "eg_fig2_duecopie.txt", "eg_fig2_quattrocopie.txt",
                                  "eg_fig2_duecopie_di_duecopie.txt", "clique_500_cp4.txt.gz"])
def test_reader_matches_oracle_reader_on_goldens(name):
    path = os.path.join(GOLDEN, name)
    _same(E.read_mpg(path), parse_mpg(path))


SYNTAX_CASES = [
    # header variants (the_parser.y:131-136), comments (the_lex.l:42-44), names, blanks
    "arena 3 5;\n0 MAX 1:2,2:-3;\n1 MIN 0:1;\n2 MIN 3:-1 ;\n3 MAX 0:0 \"v3\";\n",
    "3;\n# a comment\n0 MAX 1:2, 2:-3; % trailing comment\n1 MIN 0:1 \"one\";\n// another\n2 MIN 3:-1;\n3 MAX 0:0;\n",
    "3 9;\n3 MAX 0:0;\n2 MIN 3:-1;\n1 MIN 0:1;\n0 MAX 1:2,2:-3;\n",
    "1;\n\t1   MIN\t0 : - 4 ;\n0 MAX 1:7,1:-7;\n",
    "arena 0;\n0 MIN 0:-1;\n",
    "2;0 MAX 1:1;1 MAX 2:-2;2 MIN 0:3,1:0;",
]


@pytest.mark.parametrize("text", SYNTAX_CASES)
def test_reader_syntax_variants(text):
    _same(E.read_mpg(text=text), parse_mpg_text(text))


ERROR_CASES = [
    ("2;\n0 MAX 1:1;\n1 MIN 0:1;\n", "Expecting"),                 # node 2 missing
    ("1;\n0 MAX 1:1;\n0 MIN 0:1;\n", "Repeated"),
    ("1;\n0 MAX 1:1;\n5 MIN 0:1;\n", "Unexpected node"),
    ("1;\n0 MAX 3:1;\n1 MIN 0:1;\n", "Unexpected successor"),
    ("1;\n0 MAX 1:1.5;\n1 MIN 0:1;\n", "syntax"),
    ("1;\n0 MAX;\n1 MIN 0:1;\n", "syntax"),
    ("1;\n0 1:1;\n1 MIN 0:1;\n", "syntax"),
    ("1;\n0 MAX 1:99999999999;\n1 MIN 0:1;\n", "CONVERSION"),
]


@pytest.mark.parametrize("text,msg", ERROR_CASES)
def test_reader_rejects_bad_input(text, msg):
    with pytest.raises(ValueError, match=msg):
        E.read_mpg(text=text)
    with pytest.raises(MpgError):
        parse_mpg_text(text)


def test_reader_weight_modes():
    text = "2;\n0 MAX 1:5,2:-6;\n1 MIN 0:-2;\n2 MIN 1:9;\n"
    g = E.read_mpg(text=text, weight_mode=1)
    assert g.w.tolist() == [1, 1, 1, 1] and g.mg == 0
    g = E.read_mpg(text=text, weight_mode=2)               # --exp-weights sets 0
    assert g.w.tolist() == [0, 0, 0, 0]
    # --rnd-weights: rand() per edge in read order after srand(seed) (the_parser.y:326)
    libc = ctypes.CDLL(None)
    libc.srand(12345)
    exp = [-3 + libc.rand() % 7 for _ in range(4)]
    libc.srand(12345)
    g = E.read_mpg(text=text, weight_mode=3, rnd_lo=-3, rnd_hi=3)
    # internal order is owner-0 first: read edges (0->1,0->2) (1->0) (2->1) stay in order here
    assert g.w.tolist() == exp


def test_reader_gzip_and_generated_roundtrip(tmp_path):
    p = E.GenParams(n=300, seed=9, deg_kind=E.games.DEG_UNIFORM, deg_lo=1, deg_hi=12, wmax=40)
    owner, rptr, rsucc, rw = E.games.generate_read_order(p)
    for gz in (False, True):
        path = str(tmp_path / f"g{int(gz)}.mpg")
        E.write_mpg(path, owner, rptr, rsucc, rw, gz=gz)
        with open(path, "rb") as f:
            assert (f.read(2) == b"\x1f\x8b") == gz
        a = E.read_mpg(path)
        b = parse_mpg(path)
        _same(a, b)
        _same(a, E.games.internalise(owner, rptr, rsucc, rw))


def test_generator_properties():
    p = E.GenParams(n=5000, seed=2, deg_kind=0, deg_lo=8, deg_hi=8, wmax=100)
    owner, rptr, rsucc, rw = E.games.generate_read_order(p)
    deg = np.diff(rptr)
    assert (deg == 8).all()
    src = np.repeat(np.arange(p.n), deg)
    assert (rsucc != src).all()                              # no self-loops
    for v in range(0, p.n, 97):                               # distinct successors
        s = rsucc[rptr[v]:rptr[v + 1]]
        assert len(np.unique(s)) == len(s)
    assert rw.min() >= -100 and rw.max() <= 100
    assert abs(owner.mean() - 0.5) < 0.03
    assert abs(rw.mean()) < 2.0
    # determinism and stream separation
    again = E.games.generate_read_order(p)
    assert all(np.array_equal(x, y) for x, y in zip((owner, rptr, rsucc, rw), again))
    other = E.games.generate_read_order(p, stream=1)
    assert not np.array_equal(other[2], rsucc)


def test_generator_powerlaw_and_family():
    p = E.GenParams(n=20000, seed=3, deg_kind=2, deg_lo=1, deg_hi=100000, alpha=2.1, wmax=100)
    owner, rptr, rsucc, rw = E.games.generate_read_order(p)
    deg = np.diff(rptr)
    assert deg.min() >= 1 and deg.max() <= p.n - 1
    assert (deg == 1).mean() > 0.5                            # P(1) = 1/zeta(2.1) ~ 0.64
    assert 3.0 < deg.mean() < 7.0
    for i in (0, 15, 31):
        fp = E.family_params(E.games.FAMILY_BATCH, 4, i)
        if i % 16 == 15:
            assert fp.deg_kind == 3 and 200 <= fp.n <= 500
        else:
            assert fp.deg_kind == 1 and 1000 <= fp.n <= 50000 and (fp.deg_lo, fp.deg_hi) == (2, 16)
        assert 10 <= fp.wmax <= 1000


def test_complete_digraph_generator():
    p = E.GenParams(n=7, seed=1, deg_kind=3, deg_lo=6, deg_hi=6, wmax=5)
    owner, rptr, rsucc, rw = E.games.generate_read_order(p)
    for v in range(7):
        assert rsucc[rptr[v]:rptr[v + 1]].tolist() == [u for u in range(7) if u != v]


def test_mg_overflow_detected():
    with pytest.raises(OverflowError):
        E.games.compute_mg(2, np.array([0, 1, 2], np.int32), np.array([-2147483647, -5], np.int32))


# ---- C-ABI exports ---------------------------------------------------------------------

def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\([^;{]*\)\s*;", txt, flags=re.M)))


@pytest.mark.parametrize("header,lib", [("egs.h", "libegs.so"), ("egs_host.h", "libegs_host.so"),
                                        ("egs_dev_solver.h", "libegs_dev_solver.so")])
def test_libraries_export_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) >= 5
    path = os.path.join(N.LIBDIR, lib)
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    if lib == "libegs_dev_solver.so":
        names = [x for x in names if x != "aggiorna_max_loop"]    # defined by the driver
    missing = [x for x in names if x not in exported]
    assert not missing, missing


def test_libraries_load_and_bind():
    N.egs()
    N.host()
    for name in N.EGS_SIGNATURES:
        assert hasattr(N.egs(), name)


def test_no_cpu_fallback_without_gpu():
    """Without a gfx950 device the product path fails loudly."""
    try:
        E.device_info(0)
    except N.EgsError:
        pass
    else:
        pytest.skip("a GPU is present")
    g = E.config_game(1)
    with pytest.raises(N.EgsError):
        E.solve(g)


# ---- command line ------------------------------------------------------------------------

EGS = os.path.join(N.BINDIR, "egsolver")


def _run(*args, stdin=None):
    return subprocess.run([EGS, *args], capture_output=True, text=True, input=stdin, timeout=60)


def test_cli_help_and_errors():
    r = _run("--help")
    assert r.returncode == 0 and "--input" in r.stderr and "--shuffling" in r.stderr
    r = _run("--cpu")
    assert r.returncode == 1 and "--cpu" in r.stderr
    r = _run("--bogus")
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = _run("--shuffling", "3")
    assert r.returncode == 1 and "illegal parameter" in r.stderr
    r = _run("--tb", "100")
    assert r.returncode == 1
    r = _run("--eg", "x")
    assert r.returncode != 0 and "CONVERSION" in r.stderr
    r = _run("--input", "/nonexistent/file.mpg")
    assert r.returncode == 1 and "opening" in r.stderr


def test_host_check_and_pack():
    """The host-side CSR check and edge packing inside libegs.so (no device needed)."""
    import ctypes
    lib = N.egs()
    i32p = ctypes.POINTER(ctypes.c_int32)
    chk = lib.egs_pk_check
    chk.restype = ctypes.c_int
    chk.argtypes = [ctypes.c_int32, ctypes.c_int64, i32p, i32p, i32p, i32p, i32p, ctypes.POINTER(ctypes.c_int64)]

    def run(ptr, succ, w):
        ptr, succ, w = (np.ascontiguousarray(a, np.int32) for a in (ptr, succ, w))
        wm, dm, bad = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        rc = chk(len(ptr) - 1, len(succ), ptr.ctypes.data_as(i32p), succ.ctypes.data_as(i32p),
                 w.ctypes.data_as(i32p), ctypes.byref(wm), ctypes.byref(dm), ctypes.byref(bad))
        return rc, wm.value, dm.value, bad.value

    g = E.generate(E.GenParams(n=50000, seed=9, deg_kind=1, deg_lo=1, deg_hi=9, wmax=777))
    rc, wm, dm, _ = run(g.ptr, g.succ, g.w)
    assert rc == 0 and wm == int(np.abs(g.w).max()) and dm == int(np.diff(g.ptr).max())
    assert run([0, 1, 2], [1, 0, 1], [1, 2, 3])[0] == -1                 # ptr[n] != m
    rc, _, _, bad = run([0, 2, 2, 3], [1, 0, 1], [1, 2, 3])             # vertex 1 has no out-edge
    assert (rc, bad) == (-2, 1)
    succ = g.succ.copy()
    succ[40000] = g.n
    rc, _, _, bad = run(g.ptr, succ, g.w)
    assert (rc, bad) == (-3, 40000)
    w = g.w.copy()
    w[123] = -(2**31 - 1)
    assert run(g.ptr, g.succ, w)[1] == 2**31 - 1

    p32 = lib.egs_pk_pack32
    p32.restype = None
    p32.argtypes = [ctypes.c_int64, i32p, i32p, ctypes.c_int, ctypes.c_void_p]
    out = np.empty(g.m, np.uint32)
    p32(g.m, g.succ.ctypes.data_as(i32p), g.w.ctypes.data_as(i32p), 16, out.ctypes.data)
    assert np.array_equal(out & 0xFFFF, g.succ.astype(np.uint32))
    assert np.array_equal(out.view(np.int32) >> 16, g.w)
    p64 = lib.egs_pk_pack64
    p64.restype = None
    p64.argtypes = [ctypes.c_int64, i32p, i32p, ctypes.c_void_p]
    out = np.empty(2 * g.m, np.int32)
    p64(g.m, g.succ.ctypes.data_as(i32p), g.w.ctypes.data_as(i32p), out.ctypes.data)
    assert np.array_equal(out[0::2], g.succ) and np.array_equal(out[1::2], g.w)

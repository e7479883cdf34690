"""Host ingest: the parallel .mpg reader against the sequential one, and the binary game
cache.  CPU only.

The sequential reader is pinned by the reference's fixtures (test_host.py); the parallel
reader must return exactly the same result (every array, the names, the degree facts) and
the same error messages, including on text whose statements span lines, where the chunk
cuts fall inside statements and the reader falls back.
"""
import numpy as np
import pytest

import egsolver_amd as E
from egsolver_amd import games as G


def _same(a, b):
    """Equal results; `chunks` tells which reader produced them and is left out."""
    assert a.keys() == b.keys()
    for k in a:
        if k == "chunks":
            continue
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
        else:
            assert a[k] == b[k], k


def _generated_text(tmp_path, n, seed, deg_kind=1, lo=2, hi=12, wmax=50):
    owner, rptr, rsucc, rw = G.generate_read_order(
        E.GenParams(n=n, seed=seed, deg_kind=deg_kind, deg_lo=lo, deg_hi=hi, wmax=wmax))
    path = str(tmp_path / f"g{seed}.mpg")
    E.write_mpg(path, owner, rptr, rsucc, rw)
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("n,seed", [(60000, 1), (150000, 2), (300000, 3)])
def test_parallel_reader_equals_sequential(tmp_path, n, seed):
    text = _generated_text(tmp_path, n, seed)
    assert len(text) > (1 << 20)            # takes the parallel path
    seq = G.parse_mpg(text, threads=1)
    assert seq["chunks"] == 1
    for threads in (2, 3, 8, 16):
        par = G.parse_mpg(text, threads=threads)
        assert par["chunks"] == threads       # parsed in chunks, not by the fallback
        _same(seq, par)


def test_parallel_reader_power_law_and_weight_modes(tmp_path):
    text = _generated_text(tmp_path, 120000, 9, deg_kind=2, lo=1, hi=5000, wmax=1000)
    for mode in (0, 1, 2):                  # keep, unit, exp (random weights read sequentially)
        par = G.parse_mpg(text, weight_mode=mode, threads=8)
        assert par["chunks"] == 8
        _same(G.parse_mpg(text, weight_mode=mode, threads=1), par)
    assert G.parse_mpg(text, weight_mode=3, threads=8)["chunks"] == 1


def _statement_text(n, rng, split_lines=False, names=None, comments=False):
    """A valid game in which every vertex has out-degree 1..6; options make statements span
    lines, add names ("line": with ';' and an escaped newline; "multi": some spanning lines)
    and comment lines containing ';'."""
    out = [f"arena {n - 1} ;\n"]
    for v in rng.permutation(n):
        d = int(rng.integers(1, 7))
        edges = ", ".join(f"{int(s)}:{int(w)}" for s, w in zip(rng.integers(0, n, d), rng.integers(-40, 41, d)))
        owner = "MAX" if rng.random() < 0.5 else "MIN"
        nm = ""
        if names and rng.random() < 0.3:
            nm = f' "v{v};\\n{v}"' if names == "line" or rng.random() < 0.5 else f' "n{v}\n;{v}"'
        if comments and rng.random() < 0.1:
            out.append(f"# a comment; with {v} ; inside\n")
        if split_lines:
            out.append(f"{v} {owner}\n {edges}{nm}\n;\n")
        else:
            out.append(f"{v} {owner} {edges}{nm};\n")
    return "".join(out).encode()


@pytest.mark.parametrize("kind", ["plain", "split", "names", "multiline_names", "comments", "all"])
def test_parallel_reader_statement_shapes(kind):
    rng = np.random.default_rng(len(kind))
    names = {"names": "line", "multiline_names": "multi", "all": "multi"}.get(kind)
    text = _statement_text(90000, rng, split_lines=kind in ("split", "all"), names=names,
                           comments=kind in ("comments", "all"))
    assert len(text) > (1 << 20)
    par = G.parse_mpg(text, threads=8)
    # statements that span lines put cuts inside them and the reader falls back; one statement
    # per line parses in chunks, names and comments included; names spanning lines may go
    # either way, with the same result
    if kind in ("split", "all"):
        assert par["chunks"] == 1
    elif kind != "multiline_names":
        assert par["chunks"] == 8, par["chunks"]
    _same(G.parse_mpg(text, threads=1), par)


@pytest.mark.parametrize("fault", ["repeat_far", "repeat_near", "succ_range", "syntax", "missing", "extra",
                                   "overflow"])
def test_parallel_reader_errors_match(fault):
    rng = np.random.default_rng(5)
    n = 80000
    lines = _statement_text(n, rng).decode().splitlines(keepends=True)
    body = lines[1:]
    if fault == "repeat_far":              # the same id in the first and the last chunk
        v0 = body[0].split()[0]
        body[-1] = f"{v0} MAX 0:1;\n"
    elif fault == "repeat_near":
        v0 = body[40000].split()[0]
        body[40001] = f"{v0} MIN 1:-1;\n"
    elif fault == "succ_range":
        body[50000] = body[50000].replace(";", f", {n + 5}:3;")
    elif fault == "syntax":
        body[30000] = body[30000].replace(";", " : ;")
    elif fault == "missing":
        body = body[:-1]
    elif fault == "extra":
        body.append(f"{n + 7} MAX 0:1;\n")
    elif fault == "overflow":
        body[60000] = body[60000].replace(";", ", 1:99999999999;")
    text = (lines[0] + "".join(body)).encode()
    with pytest.raises(ValueError) as seq:
        G.parse_mpg(text, threads=1)
    with pytest.raises(ValueError) as par:
        G.parse_mpg(text, threads=8)
    assert str(seq.value) == str(par.value)


def test_read_mpg_file_uses_parallel_reader_with_same_game(tmp_path):
    text = _generated_text(tmp_path, 100000, 21)
    path = str(tmp_path / "f.mpg")
    with open(path, "wb") as f:
        f.write(text)
    a = E.read_mpg(path)                   # the file reader parses in parallel
    b = E.read_mpg(text=text)              # the sequential reader
    for k in ("ptr", "succ", "w", "ext_of_int"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert (a.n, a.n0, a.mg) == (b.n, b.n0, b.mg)


def test_csr_cache_round_trip(tmp_path):
    for g in (E.config_game(1), E.config_game(3, scale=1 / 64)):
        p = str(tmp_path / "c.egs")
        E.save_csr(g, p)
        h = E.load_csr(p)
        assert (g.n, g.n0, g.m, g.mg) == (h.n, h.n0, h.m, h.mg)
        for k in ("ptr", "succ", "w", "ext_of_int"):
            assert np.array_equal(getattr(g, k), getattr(h, k)), k


def test_csr_cache_refuses_damaged_files(tmp_path):
    g = E.config_game(1)
    p = str(tmp_path / "c.egs")
    E.save_csr(g, p)
    raw = open(p, "rb").read()
    cases = {
        "flip": raw[:200] + bytes([raw[200] ^ 4]) + raw[201:],
        "truncated": raw[:-4],
        "magic": b"XGSCSR1\0" + raw[8:],
        "longer": raw + b"\0\0\0\0",
        # header fields (64-byte header: n at 16, n0 at 20, m at 24, MG at 32, max |w| at 36)
        "n0": raw[:20] + bytes([raw[20] ^ 1]) + raw[21:],
        "mg": raw[:32] + bytes([raw[32] ^ 1]) + raw[33:],
        "max_abs_w": raw[:36] + bytes([raw[36] ^ 2]) + raw[37:],
    }
    for name, data in cases.items():
        q = str(tmp_path / f"{name}.egs")
        with open(q, "wb") as f:
            f.write(data)
        with pytest.raises(ValueError):
            E.load_csr(q)
    with pytest.raises(ValueError):
        E.load_csr(str(tmp_path / "absent.egs"))


def test_csr_cache_refuses_wrong_mg(tmp_path):
    """A file written with an MG that does not follow from its weights (the checksum then
    covers the wrong value) is refused on load: MG and max |w| are recomputed."""
    g = E.config_game(1)
    bad = G.Game(n=g.n, n0=g.n0, ptr=g.ptr, succ=g.succ, w=g.w, mg=g.mg + 1, ext_of_int=g.ext_of_int)
    p = str(tmp_path / "mg.egs")
    E.save_csr(bad, p)
    with pytest.raises(ValueError, match="MG"):
        E.load_csr(p)
    E.save_csr(g, p)
    h = E.load_csr(p)
    assert (h.n, h.n0, h.m, h.mg) == (g.n, g.n0, g.m, g.mg)


def test_csr_cache_refuses_invalid_graph(tmp_path):
    """A well-formed file whose graph breaks the CSR rules (checksum recomputed by saving it)."""
    g = E.config_game(1)
    bad = G.Game(n=g.n, n0=g.n0, ptr=g.ptr.copy(), succ=g.succ.copy(), w=g.w, mg=g.mg, ext_of_int=g.ext_of_int)
    bad.succ[7] = g.n + 3
    p = str(tmp_path / "bad.egs")
    E.save_csr(bad, p)
    with pytest.raises(ValueError, match="not a valid game"):
        E.load_csr(p)


def test_compute_mg_matches_per_vertex_sum():
    g = E.config_game(3, scale=1 / 16)
    worst = np.zeros(g.n, np.int64)
    np.maximum.at(worst, np.repeat(np.arange(g.n), np.diff(g.ptr)), -g.w.astype(np.int64))
    assert G.compute_mg(g.n, g.ptr, g.w) == int(worst.sum())

"""The command-line driver on the GPU: bin/egsolver over libegs_dev_solver.so (the
reference's seven entry points) and libegs.so.

Its printed solution must match the reference's own golden run (script/sol.txt, kept as
tests/golden/clique_500_cp4.sol) in both output formats; a game cached with --save-csr
and solved again with --load-csr prints the same solution; a round cap (--eg N) prints
values that are pointwise at most the fixpoint (the reference's partial-result semantics).
"""
import os
import re
import subprocess

import pytest

from conftest import GOLDEN, load_golden_sol

pytestmark = pytest.mark.gpu

TOP = 2147483647


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _egs():
    from egsolver_amd import _native as N
    return os.path.join(N.BINDIR, "egsolver")


def _run(*args):
    r = subprocess.run([_egs(), *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def _values(out):
    """{external id: value} from the default "%d)\\t%d" lines."""
    vals = {}
    for line in out.splitlines():
        m = re.fullmatch(r"(\d+)\)\t(-?\d+)", line)
        if m:
            vals[int(m.group(1))] = int(m.group(2))
    return vals


def _oneline_values(out):
    return {int(a): int(b) for a, b in re.findall(r"V\((\d+)\)=(-?\d+)", out)}


def test_cli_clique_golden_both_formats():
    _say("cli clique_500_cp4")
    path = os.path.join(GOLDEN, "clique_500_cp4.txt.gz")
    want = load_golden_sol()
    got = _values(_run("--input", path))
    assert got == want
    one = _oneline_values(_run("--input", path, "--onelineout"))
    assert one == want


def test_cli_eg_fig2_hand_solved():
    """eg_fig2: MG = 7; V(0)=0, V(1)=3, V(2)=TOP, V(3)=0, V(4)=TOP (hand-solved)."""
    _say("cli eg_fig2")
    got = _values(_run("--input", os.path.join(GOLDEN, "eg_fig2.txt")))
    assert got == {0: 0, 1: 3, 2: TOP, 3: 0, 4: TOP}


def test_cli_game_cache_round_trip(tmp_path):
    _say("cli --save-csr / --load-csr")
    path = os.path.join(GOLDEN, "clique_500_cp4.txt.gz")
    cache = str(tmp_path / "clique.egs")
    first = _run("--input", path, "--save-csr", cache)
    assert os.path.getsize(cache) > 64
    again = _run("--load-csr", cache)
    assert "Reading the game cache" in again
    assert _values(again) == _values(first) == load_golden_sol()


def test_cli_round_cap_is_under_approximation():
    _say("cli --eg 3")
    path = os.path.join(GOLDEN, "clique_500_cp4.txt.gz")
    full = load_golden_sol()
    part = _values(_run("--input", path, "--eg", "3"))
    assert part.keys() == full.keys()
    assert all(part[k] <= full[k] for k in full)
    assert any(part[k] < full[k] for k in full)


def _ladder_text(N):
    """script/multi/mkladder.c:42-51: even ids MAX with 0-weight edges to i+1, i+2; odd ids
    MIN with edges to i+1 (0) and i+2 (-1), indices mod 2N.  Least fixpoint: even 0, odd TOP;
    plain value iteration needs about N rounds (the odd cycle climbs 1 per round to MG = N)."""
    M = 2 * N
    out = [f"{M - 1};"]
    out += [f"{i} MAX {(i + 1) % M}:0,{(i + 2) % M}:0;" for i in range(0, M, 2)]
    out += [f"{i + 1} MIN {(i + 2) % M}:0,{(i + 3) % M}:-1;" for i in range(0, M, 2)]
    return "\n".join(out) + "\n"


def test_cli_timeout_stops_with_partial_values(tmp_path):
    """--timeout 2 (SIGALRM sets timeout_expired, src/sig_handling.c:57-64) on a ladder of
    10^6 rungs solved by plain value iteration (--no-proofs, about 10^6 rounds): the run
    stops after about 2 s and prints values pointwise at most the least fixpoint."""
    N = 1_000_000
    path = str(tmp_path / "ladder.mpg")
    with open(path, "w") as f:
        f.write(_ladder_text(N))
    _say(f"cli --timeout 2 --no-proofs ladder N={N}")
    out = _run("--input", path, "--no-proofs", "--timeout", "2")
    assert "Setting 2 seconds timeout" in out
    secs = float(re.search(r"Solving time: ([0-9.]+) sec", out).group(1))
    assert 1.5 < secs < 6.0, secs
    vals = _values(out)
    assert len(vals) == 2 * N
    assert all(vals[e] == 0 for e in range(0, 2 * N, 2))
    odd = [vals[e] for e in range(1, 2 * N, 2)]
    assert max(odd) < TOP and max(odd) > 0          # partial: climbing, not yet TOP
    # with proofs (the default) the same game solves to the least fixpoint at once
    full = _values(_run("--input", path))
    assert all(full[e] == (0 if e % 2 == 0 else TOP) for e in range(2 * N))

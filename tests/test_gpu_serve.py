"""GPU parity of the serving loop: every request's result equals the oracle's least
fixpoint and the single-game solver's result on the same game, whatever mix of other
requests shared its slices (bar: bit-exact)."""
import numpy as np
import pytest

from conftest import lfp, to_oracle_game

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _requests(seed, count, nmax, spacing):
    import egsolver_amd as E
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(count):
        p = E.family_params(E.games.FAMILY_SERVE, seed, i)
        if p.deg_kind != 3:
            p.n = int(min(p.n, nmax))
        reqs.append(E.Request(rid=i, arrival=t, game=E.generate(p, stream=i)))
        t += float(rng.exponential(spacing))
    return reqs


def test_served_equals_alone_and_oracle(oracle):
    import egsolver_amd as E
    reqs = _requests(5, 24, 3000, 0.002)
    _say(f"serve {len(reqs)} requests, sizes {[r.game.n for r in reqs]}")
    # a small arena forces admission waits and slot reuse
    arena = sum(E.Server.footprint(r.game) for r in reqs) // 3
    done, wall, slices = E.run_served(reqs, arena, E.serve_options(slice_sweeps=8), max_games=16, keep_values=True)
    assert all(r.finished >= r.arrival for r in done)
    assert slices >= 2
    for r in done:
        ref = lfp(oracle, to_oracle_game(r.game))
        assert np.array_equal(r.values, ref), (r.rid, r.game.n)
        alone, _ = E.solve(r.game)
        assert np.array_equal(alone, ref)


def test_server_capacity_and_reuse():
    import egsolver_amd as E
    reqs = _requests(6, 6, 1500, 0.0)
    need = max(E.Server.footprint(r.game) for r in reqs)
    _say("serve capacity")
    with E.Server(need + 4096, max_games=4) as srv:
        t0 = srv.submit(reqs[0].game)
        assert t0 is not None
        # a game larger than what is left is refused until the first one leaves
        if E.Server.footprint(reqs[1].game) > srv.free_bytes():
            assert srv.submit(reqs[1].game) is None
        else:
            t_extra = srv.submit(reqs[1].game)
            assert t_extra is not None
        for _ in range(10000):
            srv.launch(E.serve_options(slice_sweeps=4))
            if t0 in srv.wait():
                break
        vals, st = srv.fetch(t0)
        assert st["sweeps"] >= 1
        t1 = srv.submit(reqs[1].game)
        assert t1 is not None


def test_served_partial_sweeps_exact(oracle):
    """A per-workgroup budget of a few thousand edges stops sweeps part-way in nearly every
    slice; the resumed sweeps still end on the exact least fixpoint."""
    import egsolver_amd as E
    reqs = _requests(7, 10, 1500, 0.0)
    _say(f"serve partial sweeps, sizes {[r.game.n for r in reqs]}")
    arena = sum(E.Server.footprint(r.game) for r in reqs) + (1 << 20)
    done, wall, slices = E.run_served(reqs, arena, E.serve_options(slice_kedges=4), max_games=16,
                                      keep_values=True)
    assert slices > 10
    for r in done:
        assert np.array_equal(r.values, lfp(oracle, to_oracle_game(r.game))), (r.rid, r.game.n)


def test_served_large_route_exact(oracle):
    """Games at or above the route threshold are solved by the device-wide solver beside the
    running slices; both routes end on the oracle's least fixpoint and on the single-game
    solver's result."""
    import egsolver_amd as E
    reqs = _requests(8, 16, 3000, 0.001)
    cut = 1500
    sizes = [r.game.n for r in reqs]
    assert any(n >= cut for n in sizes) and any(n < cut for n in sizes)
    _say(f"serve two routes, cut {cut}, sizes {sizes}")
    arena = sum(E.Server.footprint(r.game) for r in reqs) + (1 << 20)
    done, wall, slices = E.run_served(reqs, arena, E.serve_options(slice_kedges=64), max_games=32,
                                      keep_values=True, large_n=cut)
    for r in done:
        ref = lfp(oracle, to_oracle_game(r.game))
        assert np.array_equal(r.values, ref), (r.rid, r.game.n)
    with E.Server(arena, max_games=4) as srv:
        srv.set_route(1)                                  # every game on the large route
        t = srv.submit(reqs[0].game)
        srv.launch()
        assert t in srv.wait()
        vals, st = srv.fetch(t)
        assert np.array_equal(vals, lfp(oracle, to_oracle_game(reqs[0].game)))
        assert st["sweeps"] >= 1


def test_served_large_route_wide_weights(oracle):
    """Large-route games whose weights do not fit beside a packed successor id travel as
    separate successor and weight arrays; small-weight ones as packed 32-bit edges."""
    import egsolver_amd as E
    games = [E.generate(E.GenParams(n=3000, seed=31, deg_kind=1, deg_lo=2, deg_hi=6, wmax=600000)),
             E.generate(E.GenParams(n=2500, seed=32, deg_kind=1, deg_lo=2, deg_hi=6, wmax=50))]
    _say(f"large route wide/narrow weights, mg {[g.mg for g in games]}")
    arena = sum(E.Server.footprint(g) for g in games) + (1 << 20)
    with E.Server(arena, max_games=4) as srv:
        srv.set_route(1, 2)
        tickets = [srv.submit(g) for g in games]
        assert None not in tickets
        got = {}
        while len(got) < len(tickets):
            srv.launch()
            for t in srv.wait():
                got[t], _ = srv.fetch(t)
        for t, g in zip(tickets, games):
            assert np.array_equal(got[t], lfp(oracle, to_oracle_game(g)))


def test_config5_stream_both_routes(oracle):
    """A seed-5 serving stream at its distribution (n log-uniform 1e3..5e5, out-degree 2..16,
    W log-uniform 10..1000, every 16th a complete digraph): 600 requests with Poisson
    arrivals through the server at the default route cut-off (16,384 vertices), so both the
    many-games kernel and the device-wide route serve.  Every result passes the host fixpoint
    check; a sample stratified by size over both routes is bit-exact against the oracle (EG0
    from 0 up to 12,000 vertices, EG0 from the certified TOP set on all) and equals the
    single-game result on the same game."""
    import egsolver_amd as E
    from conftest import exact_sample, gen_family
    from egsolver_amd import verify
    games = gen_family(E.games.FAMILY_SERVE, 5, 600)
    rng = np.random.default_rng(5)
    arrivals = np.cumsum(rng.exponential(1 / 2000.0, size=len(games)))
    reqs = [E.Request(rid=i, arrival=float(arrivals[i] - arrivals[0]), game=g) for i, g in enumerate(games)]
    large = sum(g.n >= 16384 for g in games)
    _say(f"config 5 stream: {len(games)} requests, {large} on the large route, max n {max(g.n for g in games)}")
    assert large > 100 and len(games) - large > 100
    done, wall, slices = E.run_served(reqs, 24 << 30, E.serve_options(), keep_values=True)
    assert len(done) == len(games)
    res = {r.rid: r.values for r in done}
    vals = [res[i] for i in range(len(games))]
    for i, g in enumerate(games):
        viol, first = verify.check_fixpoint(g, vals[i])
        assert viol == 0, (i, g.n, first)
    order = np.argsort([g.n for g in games], kind="stable")
    idx = sorted(set(order[np.linspace(0, len(games) - 1, 24).round().astype(int)].tolist()))
    assert any(games[i].n >= 16384 for i in idx) and any(games[i].n < 16384 for i in idx)
    nmax, mmax = max(games[i].n for i in idx), max(games[i].m for i in idx)
    with E.Solver(nmax, mmax) as s:
        unseeded, seeded = exact_sample(oracle, games, vals, idx, s)
        for i in idx:
            s.load(games[i])
            s.prepare()
            s.iterate()
            assert np.array_equal(s.values(), vals[i]), i
    _say(f"config 5: {wall:.2f} s, {slices} slices; sample {unseeded} exact from 0, {seeded} through the certificate")


def test_served_read_order_games(oracle):
    """Games submitted in read order (owners, read-order CSR with read-index successors, as a
    .mpg file lists them) are partitioned, internalised and given their MG on the device, on
    both routes and in both edge layouts (packed and, for wide weights on the large route,
    successors and weights apart); their values come back in read order and equal the
    oracle's least fixpoint of the host-internalised game, mapped back to read order."""
    import egsolver_amd as E
    params = []
    for i in range(10):
        p = E.family_params(E.games.FAMILY_SERVE, 21, i)
        if p.deg_kind != 3:
            p.n = int(min(p.n, 3000))
        params.append(p)
    params.append(E.GenParams(n=3000, seed=41, deg_kind=1, deg_lo=2, deg_hi=6, wmax=600000))   # wide weights
    params.append(E.GenParams(n=1200, seed=42, deg_kind=1, deg_lo=2, deg_hi=6, wmax=700000))
    games = []
    for i, p in enumerate(params):
        ro = E.games.generate_read_order(p, stream=i)
        g = E.games.internalise(*ro)
        ref = np.empty(g.n, np.int32)
        ref[g.ext_of_int] = lfp(oracle, to_oracle_game(g))
        games.append((ro, g, ref))
    cut = 1500
    _say(f"read-order submits, cut {cut}, sizes {[g.n for _, g, _ in games]}")
    arena = sum(2 * E.Server.footprint(g) + 24 * g.m + 32 * g.n for _, g, _ in games) + (1 << 22)
    with E.Server(arena, max_games=32) as srv:
        srv.set_route(cut, 2)
        tickets = {}
        for k, (ro, g, ref) in enumerate(games):
            t = srv.submit_read_order(*ro)
            assert t is not None
            tickets[t] = k
        got = {}
        while len(got) < len(tickets):
            srv.launch(E.serve_options(slice_kedges=64))
            for t in srv.wait():
                got[t], _ = srv.fetch(t)
        for t, k in tickets.items():
            ro, g, ref = games[k]
            assert np.array_equal(got[t], ref), (k, g.n, int((got[t] != ref).sum()))
        # a read-order game with a vertex that has no out-edge is refused before any upload
        owner, rptr, rsucc, rw = games[0][0]
        bad_ptr = rptr.copy()
        bad_ptr[1] = bad_ptr[0]
        with pytest.raises(RuntimeError):
            srv.submit_read_order(owner, bad_ptr, rsucc, rw)

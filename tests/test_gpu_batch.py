"""GPU parity of the batched path: every game of a mixed batch equals the oracle's least
fixpoint and the single-game solver's result on the same game (alone == in the mix)."""
import numpy as np
import pytest

from conftest import lfp, to_oracle_game

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _mixed_games(seed, count, nmax):
    import egsolver_amd as E
    rng = np.random.default_rng(seed)
    games = []
    for i in range(count):
        kind = i % 5
        n = int(rng.integers(5, nmax))
        if kind == 4:       # complete digraph, as the batched family's every 16th game
            p = E.GenParams(n=min(n, 120), seed=seed * 1000 + i, deg_kind=3, deg_lo=1, deg_hi=1,
                            wmax=int(rng.integers(1, 60)))
        elif kind == 3:     # skewed degrees
            p = E.GenParams(n=n, seed=seed * 1000 + i, deg_kind=2, deg_lo=1, deg_hi=min(n - 1, 400),
                            alpha=1.7, wmax=int(rng.integers(1, 300)))
        else:
            p = E.GenParams(n=n, seed=seed * 1000 + i, deg_kind=1, deg_lo=2, deg_hi=min(n - 1, 16),
                            wmax=int(rng.integers(10, 1000)))
        games.append(E.generate(p))
    return games


def test_batch_matches_oracle_and_alone(oracle):
    import egsolver_amd as E
    games = _mixed_games(1, 40, 900)
    _say(f"batch of {len(games)} games, sizes {min(g.n for g in games)}..{max(g.n for g in games)}")
    refs = [lfp(oracle, to_oracle_game(g)) for g in games]
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        for opt in [E.batch_options(), E.batch_options(threads=256, lds_ints=64),
                    E.batch_options(threads=1024, blocks_per_cu=1, lanes_per_edge_x16=16),
                    E.batch_options(group_max=1), E.batch_options(group_min=64)]:
            st = b.solve(opt)
            assert st["converged"] == len(games)
            vals = b.all_values()
            for i, (g, ref) in enumerate(zip(games, refs)):
                assert np.array_equal(vals[i], ref), (i, g.n, np.count_nonzero(vals[i] != ref))
            view = b.values_view()           # the pinned read-back: same values
            assert all(np.array_equal(view[i], vals[i]) for i in range(len(games)))
    for i in (0, 7, 19, 33):
        alone, _ = E.solve(games[i])
        assert np.array_equal(alone, refs[i])


def test_batch_unpacked_edges(oracle):
    """Weights too wide for 32-bit packing take the 64-bit edge path."""
    import egsolver_amd as E
    games = _mixed_games(2, 12, 400)
    big = E.GenParams(n=300, seed=77, deg_kind=1, deg_lo=2, deg_hi=9, wmax=6_000_000)
    games.append(E.generate(big))
    _say("batch unpacked")
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        st = b.solve()
        assert st["packed_edges"] == 0 and st["converged"] == len(games)
        for i, g in enumerate(games):
            assert np.array_equal(b.values(i), lfp(oracle, to_oracle_game(g))), i


def test_batch_config_family_sample(oracle):
    """A sample of the batched workload's own family (sizes scaled down 10x)."""
    import egsolver_amd as E
    games = []
    for i in range(24):
        p = E.family_params(E.games.FAMILY_BATCH, 4, i)
        p.n = max(50, p.n // 10) if p.deg_kind != 3 else p.n
        games.append(E.generate(p, stream=i))
    _say(f"family sample {[g.n for g in games]}")
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        st = b.solve()
        assert st["converged"] == len(games)
        for i, g in enumerate(games):
            assert np.array_equal(b.values(i), lfp(oracle, to_oracle_game(g))), i


def test_batch_sweep_cap_is_partial(oracle):
    import egsolver_amd as E
    games = _mixed_games(3, 6, 600)
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        st = b.solve(E.batch_options(max_sweeps=2))
        for i, g in enumerate(games):
            ref = lfp(oracle, to_oracle_game(g))
            assert np.all(b.values(i) <= ref)
            assert b.game_stats(i)["sweeps"] <= 2


def test_batch_with_and_without_proofs(oracle):
    """In-workgroup TOP proofs (frequent, and switched off) leave every game exact."""
    import egsolver_amd as E
    games = _mixed_games(5, 30, 2500)
    games.append(E.config_game(2, scale=1 / 1024))
    _say("batch proofs on/off")
    refs = [lfp(oracle, to_oracle_game(g)) for g in games]
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        for every in (2, 7, -1):
            st = b.solve(E.batch_options(certify_every=every))
            assert st["converged"] == len(games)
            if every > 0:
                assert st["certificates"] >= 1
            vals = b.all_values()
            for i, ref in enumerate(refs):
                assert np.array_equal(vals[i], ref), (every, i, np.count_nonzero(vals[i] != ref))


def test_batch_proof_fallback_games(oracle):
    """Batched-family games (seed 4, indices 1651, 737, 1380) on which the lowest gap
    threshold's set holds finite vertices, so a proof attempt has to move on to higher
    thresholds: exact, and the same alone and in the mix."""
    import egsolver_amd as E
    idxs = (1651, 737, 1380)
    games = [E.generate(E.family_params(E.games.FAMILY_BATCH, 4, i), stream=i) for i in idxs]
    _say(f"proof fallback games {[(g.n, g.m) for g in games]}")
    refs = [lfp(oracle, to_oracle_game(g), threads=8) for g in games]
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        for every in (16, 8, 4):
            st = b.solve(E.batch_options(certify_every=every))
            assert st["converged"] == len(games)
            for i, ref in enumerate(refs):
                assert np.array_equal(b.values(i), ref), (every, idxs[i])
    for g, ref in zip(games, refs):
        alone, st = E.solve(g)
        assert np.array_equal(alone, ref)


def test_config4_full_launch(oracle):
    """The batched workload at its size: 4,096 games of the batched family (seed 4; n
    log-uniform 1e3..5e4, out-degree 2..16, W log-uniform 10..1000, every 16th a complete
    digraph of 200..500 vertices) in one launch.  Every game: the host fixpoint check and a
    certificate of its TOP set checked on the host.  A sample of 40 games stratified by size,
    with complete digraphs and games above the 40,448-vertex LDS limit: bit-exact against
    the oracle (EG0 from 0 up to 12,000 vertices, EG0 from the certified TOP set on all)."""
    import egsolver_amd as E
    from conftest import exact_sample, gen_family
    from egsolver_amd import verify
    games = gen_family(E.games.FAMILY_BATCH, 4, 4096)
    _say(f"config 4: {len(games)} games, {sum(g.n for g in games)} vertices, {sum(g.m for g in games)} edges")
    with E.Batch() as b:
        for g in games:
            b.add(g)
        b.upload()
        st = b.solve()
        assert st["converged"] == len(games)
        vals = [v.copy() for v in b.values_view()]
    _say(f"config 4 launch: {st['ms']:.1f} ms, {st['certificates']} proofs")
    nmax, mmax = max(g.n for g in games), max(g.m for g in games)
    with E.Solver(nmax, mmax) as s:
        for i, g in enumerate(games):
            viol, first = verify.check_fixpoint(g, vals[i])
            assert viol == 0, (i, g.n, first)
            s.load(g)
            s.prepare()
            rep = verify.certify(s, g, vals[i])
            assert rep["ok"], (i, g.n, rep)
        order = np.argsort([g.n for g in games], kind="stable")
        idx = set(order[np.linspace(0, len(games) - 1, 32).round().astype(int)].tolist())
        idx |= {i for i in range(15, len(games), 16)[:4]}                 # complete digraphs
        idx |= set(order[-4:].tolist())                                   # the largest games
        assert any(games[i].n > 40448 for i in idx)
        unseeded, seeded = exact_sample(oracle, games, vals, sorted(idx), s)
    _say(f"config 4 sample: {unseeded} games exact against EG0 from 0, {seeded} through the certificate")
    assert unseeded >= 16 and seeded == len(idx)

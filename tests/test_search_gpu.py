"""GPU parity tests of the batched alpha-beta search against the oracle: the
root score must equal the oracle's exact fixed-depth value for every
position, whatever the move order, TT use or number of workers."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
pytestmark = pytest.mark.gpu


def test_terminal_and_draw_positions(zeta, oracle):
    fens = ["6k1/5ppp/8/8/8/8/5PPP/R5K1 w - - 0 1",      # mate in one
            "R5k1/5ppp/8/8/8/8/5PPP/6K1 b - - 0 1",      # mated root
            "7k/5Q2/6K1/8/8/8/8/8 b - - 0 1",            # stalemated root
            "8/8/8/8/3k4/8/3q4/3K4 w - - 0 1",           # forced capture to bare kings
            "7k/8/8/8/8/8/8/R6K w - - 99 80",            # fifty-move rule
            "7k/8/8/8/8/8/8/R6K w - - 0 80"]
    pos = np.concatenate([oracle.setboard(f) for f in fens])
    for d in (1, 2, 3, 4):
        want = [oracle.search(pos[i:i + 1], d)[0] for i in range(len(pos))]
        for tt in (0, 20):
            got = zeta.search(pos, d, tt_log2=tt)
            print("depth", d, "tt", tt, list(got["score"]), want, flush=True)
            assert list(got["score"]) == want
            assert (got["status"] == 0).all()


def test_many_move_positions(zeta, oracle):
    # nodes with more than 64 moves build their move list in the per-lane
    # loop, smaller interior nodes one move per lane: both against the oracle
    fens = ["R6R/3Q4/1Q4Q1/4Q3/2Q4Q/Q4Q2/pp1Q4/kBNN1KB1 w - - 0 1",  # 218 moves
            "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1",
            "3k4/8/8/8/8/8/8/QQQQKQQQ w - - 0 1",  # 101 moves
            "3K4/8/8/8/8/8/8/qqqqkqqq b - - 0 1"]
    pos = np.concatenate([oracle.setboard(f) for f in fens])
    for d in (1, 2, 3):
        want = [oracle.search(pos[i:i + 1], d)[0] for i in range(len(pos))]
        for workers, tt in ((1, 0), (4, 20)):
            got = zeta.search(pos, d, workers=workers, tt_log2=tt)
            print("depth", d, "workers", workers, "tt", tt, list(got["score"]), want, flush=True)
            assert list(got["score"]) == want
            assert (got["status"] == 0).all()


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_playout_positions_small_depth(zeta, oracle, depth):
    pos = oracle.gen_playouts(3, 128, 20, 60)
    want, _ = oracle.search_batch(pos, depth)
    got = zeta.search(pos, depth, tt_log2=0)
    bad = np.nonzero(got["score"] != want)[0]
    print("depth", depth, "mismatches", len(bad), bad[:10], flush=True)
    assert len(bad) == 0
    got = zeta.search(pos, depth, tt_log2=22)
    assert (got["score"] == want).all()


@pytest.mark.parametrize("depth", [2, 3, 4, 5])
def test_single_worker_tree_equals_oracle(zeta, oracle, depth):
    """One worker, no transposition table: the device search orders moves as
    the reference's picker does (MVV-LVA / piece-square deltas, the ply's
    killer at 1400 and the counter move of the previous move at 1300, highest
    first, the earliest generated among equals; src/zeta.cl:2059-2082) and
    keeps killers and counter moves across the iterations 1..depth, so it
    makes exactly the moves the oracle's iterative deepening makes, root by
    root -- the count pins the plain ordering, not only the score."""
    pos = oracle.gen_playouts(3, 128, 20, 60)
    want, wnodes = oracle.search_id_batch(pos, depth)
    got = zeta.search(pos, depth, workers=1, tt_log2=0)
    bad = np.nonzero((got["score"] != want) | (got["nodes"] != wnodes))[0]
    print("depth", depth, "nodes", int(got["nodes"].sum()), int(wnodes.sum()), "mismatches", len(bad), bad[:8],
          flush=True)
    assert (got["status"] == 0).all()
    assert len(bad) == 0


def test_node_budget_single_worker(zeta, oracle):
    """max_nodes stops a root at its node budget (src/zeta.cl:1945-1947):
    status ZETA_EBUDGET, and the score is that of the deepest iteration the
    worker completed -- the oracle's score at that depth."""
    pos = oracle.gen_playouts(3, 32, 20, 60)
    for budget in (1, 300, 5000):
        got = zeta.search(pos, 9, workers=1, tt_log2=0, max_nodes=budget)
        print("budget", budget, list(got["depth"]), list(got["status"]), flush=True)
        assert set(got["status"].tolist()) <= {0, zeta.ZETA_EBUDGET}
        assert (got["status"] == zeta.ZETA_EBUDGET).sum() >= 30
        # a worker polls every 64 nodes: it stops within one poll of the budget
        assert (got["nodes"] <= budget + 64).all()
        for i in np.nonzero(got["depth"] > 0)[0][:8]:
            assert got["score"][i] == oracle.search(pos[i:i + 1], int(got["depth"][i]))[0], i


def test_node_budget_many_workers_returns(zeta, oracle):
    """With helpers and a TT the first worker over its budget gives the root
    up and no worker rejoins it: the batch returns, every root stopped."""
    pos = oracle.gen_playouts(3, 64, 20, 60)
    for workers, tt in ((8, 20), (32, 0)):
        got = zeta.search(pos, 10, workers=workers, tt_log2=tt, max_nodes=2000)
        print("workers", workers, "tt", tt, list(got["status"]), int(got["nodes"].sum()), flush=True)
        assert (got["status"] == zeta.ZETA_EBUDGET).all()
        # each worker stops within a poll (64 nodes) of its budget; a worker
        # that joins as the root is given up stops at its first poll
        assert (got["nodes"] <= 2 * workers * (2000 + 64)).all()
    # an unbudgeted search afterwards is unaffected
    want, _ = oracle.search_batch(pos[:16], 3)
    got = zeta.search(pos[:16], 3, workers=4, tt_log2=20)
    assert (got["status"] == 0).all() and (got["score"] == want).all()


def test_abort_before_a_batch(zeta, oracle):
    pos = oracle.gen_playouts(3, 16, 20, 60)
    zeta.search_abort(True)
    try:
        got = zeta.search(pos, 5, workers=4, tt_log2=20)
    finally:
        zeta.search_abort(False)
    assert (got["status"] == zeta.ZETA_ECANCELED).all() and (got["nodes"] == 0).all()
    want, _ = oracle.search_batch(pos, 3)
    got = zeta.search(pos, 3, workers=4, tt_log2=20)
    assert (got["status"] == 0).all() and (got["score"] == want).all()


def test_abort_during_a_batch(zeta, oracle):
    """zeta_search_abort from another host thread stops a running batch at
    the workers' next poll; the roots it completed keep exact scores."""
    import threading
    import time
    pos = oracle.gen_playouts(7, 512, 20, 60)
    out = {}

    def run():
        t = time.perf_counter()
        out["res"] = zeta.search(pos, 8, workers=8, tt_log2=22)
        out["s"] = time.perf_counter() - t

    th = threading.Thread(target=run)
    th.start()
    time.sleep(1.0)
    t_abort = time.perf_counter()
    zeta.search_abort(True)
    th.join(60)
    stopped_after = time.perf_counter() - t_abort
    zeta.search_abort(False)
    assert not th.is_alive()
    got = out["res"]
    st = got["status"]
    print("finished", int((st == 0).sum()), "cancelled", int((st == zeta.ZETA_ECANCELED).sum()),
          "returned %.2f s after the abort" % stopped_after, flush=True)
    assert set(st.tolist()) <= {0, zeta.ZETA_ECANCELED}
    assert (st == zeta.ZETA_ECANCELED).sum() > 0
    assert stopped_after < 5.0
    done = np.nonzero(st == 0)[0][:4]
    for i in done:
        assert 1 <= got["depth"][i] <= 8


def test_workers_and_tt_do_not_change_scores(zeta, oracle):
    pos = oracle.gen_playouts(3, 64, 20, 60)
    want, _ = oracle.search_batch(pos, 4)
    for workers in (1, 4):
        for tt in (0, 24):
            got = zeta.search(pos, 4, workers=workers, tt_log2=tt)
            assert (got["score"] == want).all(), (workers, tt)


def test_launch_knobs_do_not_change_scores(zeta, oracle):
    # the autotuner's knobs: ABDADA split depth and wavefronts per SIMD
    # (64 roots x 64 workers ask for more wavefronts than one per SIMD)
    pos = oracle.gen_playouts(3, 64, 20, 60)
    want, _ = oracle.search_batch(pos, 4)
    for split_depth, waves in ((1, 1), (3, 2), (5, 0)):
        got = zeta.search(pos, 4, workers=64, tt_log2=24, split_depth=split_depth, waves_per_simd=waves)
        assert (got["score"] == want).all(), (split_depth, waves)
    with pytest.raises(zeta.ZetaError):
        zeta.search(pos[:1], 4, split_depth=-1)


def test_config3_depth6_golden(zeta, oracle):
    path = os.path.join(HERE, "golden", "c3_search6.npz")
    if not os.path.exists(path):
        pytest.skip("golden depth-6 scores not generated")
    gold = np.load(path)
    pos = oracle.gen_playouts(3, len(gold["score6"]), 20, 60)
    assert (pos["b"][:, 5] == gold["hash"]).all()
    got = zeta.search(pos, 4, workers=2, tt_log2=26)
    assert (got["score"] == gold["score4"]).all()
    got = zeta.search(pos, 6, workers=2, tt_log2=26)
    bad = np.nonzero(got["score"] != gold["score6"])[0]
    print("depth 6 mismatches", len(bad), "nodes", int(got["nodes"].sum()), flush=True)
    assert len(bad) == 0
    # the bench's launch shape: up to 128 workers per root, one wavefront per SIMD
    got = zeta.search(pos, 6, workers=128, tt_log2=26, clear_tt=True, waves_per_simd=1)
    bad = np.nonzero(got["score"] != gold["score6"])[0]
    print("depth 6 (128 workers) mismatches", len(bad), "nodes", int(got["nodes"].sum()), flush=True)
    assert len(bad) == 0 and (got["status"] == 0).all()


def _shuffle_games(oracle, n, seed):
    """Random games ending in a there-and-back shuffle by both sides, so the
    root repeats a position of its own game history."""
    rng = np.random.default_rng(seed)
    games = []
    while len(games) < n:
        p = oracle.setboard(oracle.STARTPOS)
        line = [p]
        ok = True
        for _ in range(int(rng.integers(6, 30))):
            k = oracle.count_legal(p)
            if k == 0:
                ok = False
                break
            p = oracle.domove(p, oracle.nth_legal(p, int(rng.integers(0, k))))
            line.append(p)
        if not ok:
            continue
        moves = []
        for _ in range(2):  # a quiet piece move for each side, then back
            cands = [oracle.nth_legal(p, i) for i in range(oracle.count_legal(p))]
            cands = [m for m in cands if (m >> 26) & 15 == 0 and ((m >> 19) & 7) not in (0, 1)
                     and (m >> 22) & 15 == (m >> 18) & 15]
            if not cands:
                ok = False
                break
            m = cands[int(rng.integers(0, len(cands)))]
            moves.append(m)
            p = oracle.domove(p, m)
            line.append(p)
        if not ok:
            continue
        for m in moves:  # undo both in order: same pieces back
            back = [oracle.nth_legal(p, i) for i in range(oracle.count_legal(p))]
            back = [b for b in back if (b & 63) == ((m >> 6) & 63) and ((b >> 6) & 63) == (m & 63)]
            if not back:
                ok = False
                break
            p = oracle.domove(p, back[0])
            line.append(p)
        if ok and oracle.count_legal(p) > 0 and int(p["b"][0][6]) < 100:
            games.append(line)
    return games


def test_repetitions_against_game_history(zeta, oracle):
    games = _shuffle_games(oracle, 48, 17)
    roots = np.concatenate([g[-1] for g in games])
    hashes = [[int(q["b"][0][5]) for q in g] for g in games]
    for depth in (3, 4):
        want = [oracle.search(g[-1], depth, hist=np.array(h, dtype=np.uint64), ply_init=len(h) - 1)[0]
                for g, h in zip(games, hashes)]
        plain = [oracle.search(g[-1], depth)[0] for g in games]
        got = zeta.search(roots, depth, workers=2, tt_log2=20, history=[h[-2::-1] for h in hashes])
        differ = sum(int(a != b) for a, b in zip(want, plain))
        print("depth", depth, "history changes", differ, "of", len(games), "root scores", flush=True)
        assert list(got["score"]) == want
        assert (got["status"] == 0).all()
    assert differ >= 1


def test_principal_variation_is_legal(zeta, oracle):
    pos = oracle.gen_playouts(3, 32, 20, 60)
    res, pv = zeta.search(pos, 5, workers=2, tt_log2=22, pv_len=8)
    for i in range(len(pos)):
        assert int(pv[i][0]) == int(res["bestmove"][i])
        p = pos[i:i + 1]
        for m in pv[i]:
            if not m:
                break
            legal = [oracle.nth_legal(p, k) for k in range(oracle.count_legal(p))]
            assert int(m) in legal
            p = oracle.domove(p, int(m))
    mate = oracle.setboard("6k1/5ppp/8/8/8/8/5PPP/R5K1 w - - 0 1")
    res, pv = zeta.search(mate, 3, tt_log2=20, pv_len=4)
    assert oracle.move2can(int(pv[0][0])) == "a1a8"

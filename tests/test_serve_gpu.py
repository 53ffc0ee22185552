"""GPU tests of the serving loop (zeta_serve_*): each request's root score is
the oracle's exact fixed-depth value, and the same whether the request runs
alone or in a mix of requests of other depths that arrive while it runs;
admission by memory budget, cancellation and the busy guard."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _drain(sv, pending, results, timeout=90.0):
    """Reap until every ticket in pending has a result (results keeps all reaped)."""
    t0 = time.time()
    pending = {t for t in pending if t not in results}
    while pending:
        for d in sv.reap():
            results[int(d["ticket"])] = d
            pending.discard(int(d["ticket"]))
        if time.time() - t0 > timeout:
            raise AssertionError("serving requests did not finish: %s" % sorted(pending)[:8])
        time.sleep(0.0005)


def _run_stream(sv, pos, depths, learned=False, gap=0.0):
    """Submit all requests (a backlog absorbs rejections), collect all results."""
    tick = {}
    results = {}
    pending = set()
    i = 0
    while i < len(pos) or pending:
        while i < len(pos):
            t = sv.submit(pos[i], int(depths[i]), learned=learned)
            if t is None:
                break
            tick[i] = t
            pending.add(t)
            i += 1
            if gap:
                time.sleep(gap)
        for d in sv.reap():
            results[int(d["ticket"])] = d
            pending.discard(int(d["ticket"]))
        time.sleep(0.0002)
    return [results[tick[k]] for k in range(len(pos))]


@pytest.fixture(scope="module")
def mixed(oracle):
    pos = oracle.gen_playouts(5, 24, 20, 60)
    depths = np.array([2, 3, 4, 3, 2, 4] * 4)
    want = np.zeros(len(pos), dtype=np.int32)
    for d in np.unique(depths):
        idx = np.nonzero(depths == d)[0]
        want[idx], _ = oracle.search_batch(pos[idx], int(d))
    return pos, depths, want


def test_serve_mix_equals_oracle_and_alone(zeta, mixed):
    pos, depths, want = mixed
    with zeta.Server(slots=8, workers=4, tt_log2=20) as sv:
        print("serve: mixed stream of", len(pos), "requests", flush=True)
        got = _run_stream(sv, pos, depths, gap=0.0005)
        scores = np.array([g["score"] for g in got])
        status = np.array([g["status"] for g in got])
        print("serve: mixed", list(scores), flush=True)
        assert (status == 0).all()
        assert (scores == want).all()
        # every iteration up to the requested depth, unless a mate within
        # an earlier iteration's depth ended them (src/search.c:177-186)
        assert all(int(g["depth"]) == int(d) or (abs(int(g["score"])) > 30000 and int(g["depth"]) <= int(d))
                   for g, d in zip(got, depths))
        print("serve: each request alone", flush=True)
        for k in range(8):
            t = sv.submit(pos[k], int(depths[k]))
            assert t is not None
            res = {}
            _drain(sv, {t}, res)
            assert int(res[t]["score"]) == int(want[k]) and int(res[t]["status"]) == 0
    # the batch entry point agrees once serving has stopped
    got = zeta.search(pos[:6], 3, workers=2, tt_log2=20)
    assert got["status"].tolist() == [0] * 6


def test_serve_learned_ordering_same_scores(zeta, mixed):
    pos, depths, want = mixed
    zeta.mlp_init(7)
    with zeta.Server(slots=16, workers=4, tt_log2=20, use_mlp=True, mlp_min_depth=2) as sv:
        got = _run_stream(sv, pos, depths, learned=True)
    assert (np.array([g["score"] for g in got]) == want).all()
    assert all(int(g["status"]) == 0 for g in got)


def test_serve_admission_budget(zeta, oracle, mixed):
    pos = mixed[0]
    want4, _ = oracle.search_batch(pos[:3], 4)
    with zeta.Server(slots=8, workers=2, tt_log2=16) as sv:
        c1, c2 = sv.cost(1), sv.cost(2)
    assert c2 > c1 > 0
    # room for exactly two requests of two workers
    with zeta.Server(slots=8, workers=2, tt_log2=16, mem_budget=2 * c2) as sv:
        t0 = sv.submit(pos[0], 4)
        t1 = sv.submit(pos[1], 4)
        assert t0 is not None and t1 is not None
        # the third is not admitted until one of the first two is collected
        assert sv.submit(pos[2], 4) is None
        res = {}
        _drain(sv, {t0}, res)
        t2 = sv.submit(pos[2], 4)
        assert t2 is not None
        _drain(sv, {t1, t2}, res)
        assert [int(res[t]["score"]) for t in (t0, t1, t2)] == [int(x) for x in want4]
    # a request larger than the whole budget is refused outright
    with zeta.Server(slots=8, workers=2, tt_log2=16, mem_budget=c1 // 2) as sv:
        with pytest.raises(zeta.ZetaError):
            sv.submit(pos[0], 3)


def test_serve_cancel_one_request(zeta, oracle, mixed):
    pos, depths, want = mixed
    deep = oracle.setboard("r1bqkb1r/pppp1ppp/2n2n2/4p3/2B1P3/5N2/PPPP1PPP/RNBQK2R w KQkq - 4 4")
    with zeta.Server(slots=8, workers=8, tt_log2=20) as sv:
        # helpers join the deep request; none may join once it is cancelled
        td = sv.submit(deep[0], 14, workers=8)
        assert td is not None
        tickets = [sv.submit(pos[k], int(depths[k])) for k in range(6)]
        assert all(t is not None for t in tickets)
        time.sleep(0.05)
        sv.cancel(td)
        res = {}
        _drain(sv, set(tickets) | {td}, res)
        assert int(res[td]["status"]) == zeta.ZETA_ECANCELED
        for k, t in enumerate(tickets):
            assert int(res[t]["status"]) == 0
            assert int(res[t]["score"]) == int(want[k])


def test_serve_owns_the_device(zeta, mixed):
    pos, depths, want = mixed
    with zeta.Server(slots=4, workers=1, tt_log2=16) as sv:
        with pytest.raises(zeta.ZetaError):
            zeta.search(pos[:2], 2)
        t = sv.submit(pos[0], 2)
        res = {}
        _drain(sv, {t}, res)
        assert int(res[t]["score"]) == int(want[0])
    got = zeta.search(pos[:2], 2, tt_log2=0)
    assert (got["status"] == 0).all()

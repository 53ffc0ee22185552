"""GPU tests of the learned-ordering and serving configurations at their
workloads.

Learned ordering: 1024 held-out playout positions (seed 40, 20-60 plies) at
depth 7, one batch each, up to 128 workers per root on one wavefront per
SIMD (the autotuned launch shape), TT 2^26: with the
committed network (zeta_amd/order_net.npz) every root score equals the plain
ordering's (alpha-beta's value does not depend on move order).

Serving: a Poisson stream (seed 5, 2000 requests/s, depths 4-9) through
zeta_serve_*, plain then learned ordering, each capped in time: every request
both runs complete has the same score in both, and each completed request of
depth <= 6 has the score zeta_search_batch gives the same position."""
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NET = os.path.join(ROOT, "zeta_amd", "order_net.npz")
_cache = {}


def _c4_plain(zeta):
    if "c4_plain" not in _cache:
        pos = zeta.playouts(40, 1024, 20, 60)
        t = time.perf_counter()
        r = zeta.search(pos, 7, workers=128, tt_log2=26, clear_tt=True, waves_per_simd=1)
        _cache["c4_plain"] = (pos, r, time.perf_counter() - t)
    return _cache["c4_plain"]


def test_c4_plain_ordering_completes(zeta):
    pos, r, dt = _c4_plain(zeta)
    print(json.dumps({"plain_s": round(dt, 2), "plain_nodes": int(r["nodes"].sum())}), flush=True)
    assert (r["status"] == 0).all()
    assert ((r["depth"] == 7) | (np.abs(r["score"]) > 30000)).all()


def test_c4_learned_ordering_identical_scores(zeta):
    from zeta_amd import learned as L
    pos, plain, dt_plain = _c4_plain(zeta)
    L.load(NET)
    t = time.perf_counter()
    r = zeta.search(pos, 7, workers=128, tt_log2=26, clear_tt=True, use_mlp=True, mlp_min_depth=3,
                    waves_per_simd=1)
    dt = time.perf_counter() - t
    print(json.dumps({"plain_s": round(dt_plain, 2), "learned_s": round(dt, 2),
                      "plain_nodes": int(plain["nodes"].sum()), "learned_nodes": int(r["nodes"].sum()),
                      "mlp_nodes": zeta.search_counters()["mlp_nodes"]}), flush=True)
    assert (r["status"] == 0).all()
    assert zeta.search_counters()["mlp_nodes"] > 0
    assert (r["score"] == plain["score"]).all()


def _c5_streams(zeta):
    if "c5" not in _cache:
        from zeta_amd import learned as L
        from zeta_amd import serving as S
        L.load(NET)
        pos, depths, arrivals = S.make_stream(2000.0, 24, seed=5, dmin=4, dmax=9)
        runs = {}
        for mode in ("plain", "learned"):
            out, lat, wall, wn = S.run_stream(pos, depths, arrivals, learned=(mode == "learned"), slots=64,
                                              workers=128, tt_log2=24, mlp_min_depth=3, max_seconds=40.0)
            runs[mode] = out
            print(json.dumps({mode: S.summarize(out, lat, wall, depths, wn)}), flush=True)
        _cache["c5"] = (pos, depths, runs)
    return _cache["c5"]


def test_c5_stream_learned_equals_plain(zeta):
    pos, depths, runs = _c5_streams(zeta)
    ok = (runs["plain"]["status"] == 0) & (runs["learned"]["status"] == 0)
    # the cap cancels only what is still running: every request of depth
    # 4-6 completes in both runs
    assert ok[depths <= 6].all()
    assert set(runs["plain"]["status"].tolist()) <= {0, zeta.ZETA_ECANCELED}
    assert (runs["plain"]["score"][ok] == runs["learned"]["score"][ok]).all()


def test_c5_stream_equals_batch_search(zeta):
    pos, depths, runs = _c5_streams(zeta)
    for d in (4, 5, 6):
        idx = np.nonzero((depths == d) & (runs["plain"]["status"] == 0))[0]
        if len(idx) == 0:
            continue
        want = zeta.search(pos[idx], d, workers=8, tt_log2=22, clear_tt=True)
        assert (runs["plain"]["score"][idx] == want["score"]).all(), d

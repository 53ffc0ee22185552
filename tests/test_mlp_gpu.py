"""GPU tests of the learned-ordering network's BF16 MFMA kernels: the
gradient of a batch computed on the GPU (forward GEMMs, loss, backward
GEMMs) against the numpy oracle with the same BF16 rounding points, and
Adam training that lowers the loss."""
import numpy as np
import pytest

from oracle import mlp_ref as M

pytestmark = pytest.mark.gpu


def _pairs(oracle, n, seed):
    rng = np.random.default_rng(seed)
    pos = oracle.gen_playouts(200 + seed, 2 * n, 8, 40)
    out = np.zeros(n, dtype=M.PAIR_DTYPE)
    i = 0
    for j in range(len(pos)):
        p = pos[j:j + 1]
        k = oracle.count_legal(p)
        if k < 2:
            continue
        a, b = rng.choice(k, 2, replace=False)
        out[i]["b"] = p["b"][0, :4]
        out[i]["stm"] = int(p["stm"][0])
        out[i]["mpos"] = oracle.nth_legal(p, int(a))
        out[i]["mneg"] = oracle.nth_legal(p, int(b))
        out[i]["depth"] = M.depth_word(1 + i % 6, int(rng.integers(-400, 400)))
        i += 1
        if i == n:
            break
    assert i == n
    return out


def test_params_roundtrip(zeta):
    w = np.random.default_rng(0).normal(0, 0.05, M.NPARAMS).astype(np.float32)
    zeta.mlp_set(w)
    assert np.array_equal(zeta.mlp_get(), w)


@pytest.mark.parametrize("npairs", [48, 64, 2048])
def test_gradient_matches_oracle(zeta, oracle, npairs):
    # 48 pairs: a row-tile count (3) that leaves waves of a block idle; 64
    # pairs: each K-split product in one block per tile; 2048 pairs (the
    # training batch): several blocks per tile adding into the gradient
    pairs = _pairs(oracle, npairs, 1)
    w = np.random.default_rng(5).normal(0, 0.08, M.NPARAMS).astype(np.float32)
    zeta.mlp_set(w)
    loss, g = zeta.mlp_grad(pairs)
    lref, gref = M.forward_backward(w, pairs, bf16_points=True)
    print("loss", loss, lref, flush=True)
    # tolerance: FP32 sums in a different order can flip single BF16 roundings
    assert abs(loss - lref) <= 1e-3 * max(1.0, abs(lref))
    parts = np.split(np.arange(M.NPARAMS), np.cumsum([908 * 256, 256, 256 * 32, 32, 32]))
    for name, ix in zip(("W1", "b1", "W2", "b2", "W3"), parts[:5]):
        err = np.linalg.norm(g[ix] - gref[ix]) / max(np.linalg.norm(gref[ix]), 1e-12)
        print(name, "relative error", err, flush=True)
        assert err < 2e-2, name
    # s_neg - s_pos does not depend on b3: its gradient is zero up to rounding
    assert abs(float(g[parts[5]][0])) < 1e-6 and abs(float(gref[parts[5]][0])) < 1e-6


def test_training_lowers_loss(zeta, oracle):
    pairs = _pairs(oracle, 512, 2)
    zeta.records_put(pairs)
    zeta.mlp_init(7)
    losses = zeta.mlp_train(200, pairs_per_step=256, lr=2e-3, seed=3)
    print("loss first/last", losses[:3], losses[-3:], flush=True)
    assert losses[-20:].mean() < 0.8 * losses[:5].mean()
    zeta.records_clear()


def test_search_scoring_matches_oracle(zeta, oracle):
    w = np.random.default_rng(9).normal(0, 0.06, M.NPARAMS).astype(np.float32)
    zeta.mlp_set(w)
    pos = oracle.gen_playouts(31, 24, 10, 50)
    moves = [[oracle.nth_legal(pos[i:i + 1], k) for k in range(oracle.count_legal(pos[i:i + 1]))]
             for i in range(len(pos))]
    got = zeta.mlp_score(pos, moves)
    for i in range(len(pos)):
        ref = M.scores(w, pos[i]["b"][:4], int(pos[i]["stm"]), moves[i])
        # 1/256 output resolution; BF16 rounding of h1 summed in another order
        assert np.max(np.abs(got[i] - ref)) <= 2e-2 + 2e-2 * np.max(np.abs(ref)), i


def test_learned_ordering_keeps_scores_and_logging(zeta, oracle):
    pos = oracle.gen_playouts(3, 128, 20, 60)
    want, _ = oracle.search_batch(pos, 4)
    zeta.mlp_init(11)
    for W in (1, 4):
        got = zeta.search(pos, 4, workers=W, tt_log2=22, use_mlp=True, mlp_min_depth=2)
        assert (got["score"] == want).all()
    assert zeta.search_counters()["mlp_nodes"] > 0
    zeta.records_reserve(200000)
    zeta.records_clear()
    zeta.search(pos, 4, workers=1, tt_log2=22, log_records=True)
    n = zeta.records_count()
    print("pairs logged", n, flush=True)
    assert n > 100
    recs = zeta.records_get(300)
    for r in recs:
        p = np.zeros(1, dtype=oracle.POS_DTYPE)
        p["b"][0, :4] = r["b"]
        p["b"][0, 4] = ~np.uint64(0)
        p["stm"][0] = int(r["stm"]) & 1
        legal = {oracle.nth_legal(p, k) & 0x3FFFFFFF for k in range(oracle.count_legal(p))}
        assert int(r["mpos"]) != int(r["mneg"])
        assert int(r["mneg"]) in legal or ((int(r["mneg"]) >> 19) & 7) in (1, 3)  # e.p./castle need flags
    zeta.records_clear()


def test_working_copies_follow_training(zeta, oracle):
    # the Adam kernel also writes the BF16 copies the products read: after
    # some training steps the gradient the GPU computes must still be the
    # oracle's for the parameters the training left
    pairs = _pairs(oracle, 512, 3)
    zeta.records_put(pairs)
    zeta.mlp_init(9)
    zeta.mlp_train(5, pairs_per_step=256, lr=1e-2, seed=5)
    w = zeta.mlp_get()
    check = pairs[:64]
    loss, g = zeta.mlp_grad(check)
    lref, gref = M.forward_backward(w, check, bf16_points=True)
    assert abs(loss - lref) <= 1e-3 * max(1.0, abs(lref))
    parts = np.split(np.arange(M.NPARAMS), np.cumsum([908 * 256, 256, 256 * 32, 32, 32]))
    for name, ix in zip(("W1", "b1", "W2", "b2", "W3"), parts[:5]):
        err = np.linalg.norm(g[ix] - gref[ix]) / max(np.linalg.norm(gref[ix]), 1e-12)
        assert err < 2e-2, name

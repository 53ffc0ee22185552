"""CPU tests pinning the MLP oracle: finite differences of the float64
mathematics, and the BF16-rounded variant staying close to it."""
import numpy as np

from oracle import mlp_ref as M
from oracle import refcpu as R


def make_pairs(n, seed=0):
    rng = np.random.default_rng(seed)
    orng = np.random.default_rng(1000 + seed)
    pos = R.gen_playouts(100 + seed, n, 8, 40)
    out = np.zeros(n, dtype=M.PAIR_DTYPE)
    for i in range(n):
        p = pos[i:i + 1]
        k = R.count_legal(p)
        a, b = rng.choice(k, 2, replace=False) if k > 1 else (0, 0)
        out[i]["b"] = p["b"][0, :4]
        out[i]["stm"] = int(p["stm"][0])
        out[i]["mpos"] = R.nth_legal(p, int(a))
        out[i]["mneg"] = R.nth_legal(p, int(b))
        # the plain ordering's score difference the search logs with a pair
        out[i]["depth"] = M.depth_word(1 + i % 6, int(orng.integers(-400, 400)))
    return out


def random_params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, M.NPARAMS)


def test_feature_layout():
    p = R.setboard(R.STARTPOS)
    m = R.nth_legal(p, 0)
    f = M.features(p["b"][0, :4], 0, m)
    assert len(f) == 32 + 3  # 32 pieces, from, to, piece type; no capture
    assert all(0 <= x < 908 for x in f) and len(set(f)) == len(f)
    # black to move sees the mirrored board: the same features as white
    q = R.setboard("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR b KQkq - 0 1")
    assert sorted(M.features(q["b"][0, :4], 1, 0)[:32]) == sorted(f[:32])


def test_finite_differences_float64():
    pairs = make_pairs(6)
    w = random_params(1)
    loss, g = M.forward_backward(w, pairs, bf16_points=False)
    rng = np.random.default_rng(2)
    W1, b1, W2, b2, W3, b3 = M.unpack(np.arange(M.NPARAMS))
    X = M.design_matrix(pairs)
    active = np.nonzero(X.sum(0))[0]
    picks = [int(W1[f, u]) for f, u in zip(rng.choice(active, 8), rng.integers(0, 256, 8))]
    picks += [int(x) for x in rng.choice(b1, 4)] + [int(x) for x in rng.choice(W2.ravel(), 8)]
    picks += [int(x) for x in rng.choice(b2, 4)] + [int(x) for x in rng.choice(W3, 4)] + [int(b3[0])]
    eps = 1e-6
    for i in picks:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        num = (M.forward_backward(wp, pairs, False)[0] - M.forward_backward(wm, pairs, False)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-5 + 1e-4 * abs(num), (i, num, g[i])


def test_plain_offset_closed_form():
    """With the network constant (all weights zero) every pair's loss is
    softplus(o / PLAIN_PER_LOGIT): the plain ordering's score difference o
    (negative minus positive) counts against the pair in plain-score units."""
    pairs = make_pairs(5, seed=7)
    offs = [int(x) for x in (-250, -1, 0, 37, 900)]
    for i, o in enumerate(offs):
        pairs[i]["depth"] = M.depth_word(3, o)
    assert [int(v) for v in M.pair_offsets(pairs)] == offs
    loss, g = M.forward_backward(np.zeros(M.NPARAMS), pairs, bf16_points=False)
    want = np.mean([np.logaddexp(0, o / 100.0) for o in offs])
    assert abs(loss - want) < 1e-12
    assert np.count_nonzero(g[:-33]) == 0  # relu(0) = 0: only the head moves


def test_bf16_variant_close_to_exact():
    pairs = make_pairs(32, seed=3)
    pairs["depth"] = 3  # no plain offsets: no pair saturates, the BF16 rounding dominates the difference
    w = random_params(4, 0.1)
    l64, g64 = M.forward_backward(w, pairs, bf16_points=False)
    l16, g16 = M.forward_backward(w.astype(np.float32), pairs, bf16_points=True)
    assert abs(l64 - l16) < 2e-2 * max(1.0, abs(l64))
    assert np.linalg.norm(g16 - g64) < 5e-2 * np.linalg.norm(g64)


def test_committed_network_checkpoint():
    """The committed ordering network (zeta_amd/order_net.npz, written by
    zeta_amd.learned.save): BF16 W1 and W2, FP32 biases and head, loadable
    without pickles, NPARAMS parameters in the C-ABI layout; the oracle
    scores moves with it to finite values."""
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "zeta_amd", "order_net.npz")
    with np.load(path, allow_pickle=False) as z:
        w1, w2, rest = z["w1_bf16"], z["w2_bf16"], z["rest_f32"]
    assert w1.dtype == np.uint16 and w1.size == M.NIN * M.NH1
    assert w2.dtype == np.uint16 and w2.size == M.NH1 * M.NH2
    assert rest.dtype == np.float32 and rest.size == M.NH1 + M.NH2 + M.NH2 + 1
    bf = lambda a: (a.astype(np.uint32) << 16).view(np.float32)  # noqa: E731
    params = np.concatenate([bf(w1), rest[:M.NH1], bf(w2), rest[M.NH1:]])
    assert params.size == M.NPARAMS and np.isfinite(params).all()
    p = R.setboard(R.STARTPOS)
    moves = [R.nth_legal(p, i) for i in range(R.count_legal(p))]
    s = M.scores(params, p["b"][0, :4], 0, moves)
    assert len(s) == 20 and np.isfinite(s).all()

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
    pos = oracle.gen_playouts(200 + seed, n, 8, 40)
    out = np.zeros(n, dtype=M.PAIR_DTYPE)
    for i in range(n):
        p = pos[i:i + 1]
        k = oracle.count_legal(p)
        a, b = rng.choice(k, 2, replace=False)
        out[i]["b"] = p["b"][0, :4]
        out[i]["stm"] = int(p["stm"][0])
        out[i]["mpos"] = oracle.nth_legal(p, int(a))
        out[i]["mneg"] = oracle.nth_legal(p, int(b))
    return out


def test_params_roundtrip(zeta):
    w = np.random.default_rng(0).normal(0, 0.05, M.NPARAMS).astype(np.float32)
    zeta.mlp_set(w)
    assert np.array_equal(zeta.mlp_get(), w)


def test_gradient_matches_oracle(zeta, oracle):
    pairs = _pairs(oracle, 64, 1)
    w = np.random.default_rng(5).normal(0, 0.08, M.NPARAMS).astype(np.float32)
    zeta.mlp_set(w)
    loss, g = zeta.mlp_grad(pairs)
    lref, gref = M.forward_backward(w, pairs, bf16_points=True)
    print("loss", loss, lref, flush=True)
    # tolerance: FP32 sums in a different order can flip single BF16 roundings
    assert abs(loss - lref) <= 1e-3 * max(1.0, abs(lref))
    parts = np.split(np.arange(M.NPARAMS), np.cumsum([908 * 256, 256, 256 * 32, 32, 32]))
    for name, ix in zip(("W1", "b1", "W2", "b2", "W3", "b3"), parts):
        err = np.linalg.norm(g[ix] - gref[ix]) / max(np.linalg.norm(gref[ix]), 1e-12)
        print(name, "relative error", err, flush=True)
        assert err < 2e-2, name


def test_training_lowers_loss(zeta, oracle):
    pairs = _pairs(oracle, 512, 2)
    zeta.records_put(pairs)
    zeta.mlp_init(7)
    losses = zeta.mlp_train(200, pairs_per_step=256, lr=2e-3, seed=3)
    print("loss first/last", losses[:3], losses[-3:], flush=True)
    assert losses[-20:].mean() < 0.8 * losses[:5].mean()
    zeta.records_clear()

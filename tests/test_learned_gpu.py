"""GPU test of the learned move ordering end to end: pairs logged by the plain
search, training on the GPU, what the trained network gains over the same
network untrained on pairs from positions it never saw, and the search with
the learned ordering reaching root scores identical to the plain search
(alpha-beta's value does not depend on move order)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_network_gains_over_untrained_at_identical_scores(zeta):
    from zeta_amd import learned as L
    # held-out pairs (seed 41) and training pairs (seed 4) from the plain search
    held = L.heldout_pairs(seed=41, n=256, depth=4, workers=4, maxn=8192)
    rec = L.generate_records(seed=4, n=1024, depth=4, workers=4, capacity=400_000, chunk=1024)
    print(json.dumps(rec), flush=True)
    assert rec["pairs"] > 20_000 and len(held) >= 1024
    zeta.mlp_init(11)
    loss_untrained, _ = zeta.mlp_grad(held)
    tr = L.train(steps=1000, pairs_per_step=2048, lr=2e-4, seed=3, init_seed=11, held=held)
    loss_trained, _ = zeta.mlp_grad(held)
    print(json.dumps({**tr, "heldout_loss_untrained": loss_untrained, "heldout_loss_trained": loss_trained}),
          flush=True)
    # an untrained network ranks a pair at chance (softplus(0) = 0.693);
    # the trained one ranks pairs from positions it never saw better
    assert abs(loss_untrained - np.log(2)) < 0.1
    assert loss_trained < 0.92 * loss_untrained

    held_pos = zeta.playouts(40, 256, 20, 60)
    plain = zeta.search(held_pos, 4, workers=1, tt_log2=24, clear_tt=True)
    learned = zeta.search(held_pos, 4, workers=1, tt_log2=24, clear_tt=True, use_mlp=True, mlp_min_depth=2)
    print(json.dumps({"nodes_plain": int(plain["nodes"].sum()), "nodes_learned": int(learned["nodes"].sum())}),
          flush=True)
    assert (plain["status"] == 0).all() and (learned["status"] == 0).all()
    assert (learned["score"] == plain["score"]).all()

"""Learned move ordering: record generation, training and evaluation.

Pipeline (all on the GPU through the C-ABI):
  1. the plain search (TT move, MVV-LVA / piece-square deltas, killer and
     counter move: the reference's scheme, src/zeta.cl:2059-2104) runs on
     seeded playout positions with pair logging on: at sampled cut nodes of
     the full-width search whose cutoff move is not the TT move, the cutoff
     move is paired with every other move of the node, each pair carrying
     the plain ordering's score difference;
  2. the 908-256-32-1 network is trained on those pairs with the pairwise
     ranking loss softplus(s_neg - s_pos + plain difference / 100) (BF16
     MFMA forward/backward, Adam): it learns a correction of the plain
     ordering, which the search applies as plain score + 100 x network
     output, the TT move still first;
  3. held-out positions are searched with plain and with learned ordering;
     the root scores must be identical (alpha-beta's value does not depend on
     move order) and the node counts / times are compared.
"""
import json
import time

import numpy as np

from . import (MLP_NPARAMS, mlp_get, mlp_grad, mlp_init, mlp_set, mlp_train, playouts, records_clear,
               records_count, records_get, records_reserve, search, search_counters)


def generate_records(seed=4, n=20000, depth=5, workers=16, capacity=3_000_000, chunk=4096, progress=None,
                     log_shift=None):
    """Log training pairs from plain searches of n playout positions (20-60 plies).

    Cut nodes are sampled by hash (1 in 2^shift) so that the pairs spread over
    all n positions instead of filling the buffer with the first ones; the
    rate is log_shift, or set from a probe of the first 256 positions."""
    records_reserve(capacity)
    pos = playouts(seed, n, 20, 60)
    t = time.perf_counter()
    records_clear()
    per_pos = None
    if log_shift is None:
        probe = min(256, n)
        search(pos[:probe], depth, workers=workers, tt_log2=26, clear_tt=True, log_records=True, log_shift=0)
        per_pos = max(records_count(), 1) / probe
        shift = 0
        while shift < 20 and per_pos * n / (1 << shift) > 0.8 * capacity:
            shift += 1
    else:
        shift = int(log_shift)
    records_clear()
    for i in range(0, n, chunk):
        search(pos[i:i + chunk], depth, workers=workers, tt_log2=26, clear_tt=(i == 0), log_records=True,
               log_shift=shift)
        if progress:
            progress({"generated_positions": min(i + chunk, n), "pairs": records_count()})
    return {"positions": n, "depth": depth, "pairs": min(records_count(), capacity), "log_shift": shift,
            "pairs_per_position_unsampled": per_pos, "seconds": time.perf_counter() - t}


def heldout_pairs(seed=41, n=256, depth=4, workers=4, maxn=16384):
    """Pairs from positions the training never sees (validation set)."""
    generate_records(seed=seed, n=n, depth=depth, workers=workers, capacity=max(4 * maxn, 65536), chunk=n)
    held = records_get()
    return held[:(min(len(held), maxn) // 16) * 16]


def train(steps=3000, pairs_per_step=2048, lr=2e-4, seed=1, init_seed=1, held=None, eval_every=100, patience=4):
    """Adam on random batches of the logged pairs.  With validation pairs the
    loss on them is measured every eval_every steps; training stops after
    `patience` evaluations without improvement and keeps the best weights
    (the network overfits the positions it was trained on quickly)."""
    mlp_init(init_seed)
    # the output layer starts at zero: the untrained network adds nothing to
    # the plain ordering, and training (with early stopping on held-out
    # pairs) moves away from it only where that ranks held-out pairs better
    p = mlp_get()
    p[_N_W1 + 256 + _N_W2 + 32:] = 0.0
    mlp_set(p)
    t = time.perf_counter()
    if held is None:
        losses = mlp_train(steps, pairs_per_step, lr, seed)
        return {"steps": steps, "pairs_per_step": pairs_per_step, "lr": lr, "seconds": time.perf_counter() - t,
                "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-50:].mean())}
    best, best_params, best_step, bad = mlp_grad(held)[0], mlp_get(), 0, 0
    first = None
    done = 0
    curve = []
    while done < steps and bad < patience:
        k = min(eval_every, steps - done)
        losses = mlp_train(k, pairs_per_step, lr, seed + done)
        if first is None:
            first = float(losses[:20].mean())
        done += k
        hl = mlp_grad(held)[0]
        curve.append([done, round(float(losses[-20:].mean()), 4), round(hl, 4)])
        if hl < best:
            best, best_params, best_step, bad = hl, mlp_get(), done, 0
        else:
            bad += 1
    mlp_set(best_params)
    return {"steps_run": done, "best_step": best_step, "pairs_per_step": pairs_per_step, "lr": lr,
            "seconds": time.perf_counter() - t, "loss_first": first, "heldout_best": best, "curve": curve}


def evaluate(seed=40, n=1024, depth=7, workers=128, mlp_min_depth=3, tt_log2=26, waves_per_simd=1):
    """Plain vs learned ordering on held-out positions; identical root scores required."""
    pos = playouts(seed, n, 20, 60)
    out = {}
    res = {}
    for mode in ("plain", "learned"):
        t = time.perf_counter()
        r = search(pos, depth, workers=workers, tt_log2=tt_log2, clear_tt=True, use_mlp=(mode == "learned"),
                   mlp_min_depth=mlp_min_depth, waves_per_simd=waves_per_simd)
        dt = time.perf_counter() - t
        res[mode] = r
        out[mode] = {"seconds": dt, "nodes": int(r["nodes"].sum()), **search_counters()}
    same = bool((res["plain"]["score"] == res["learned"]["score"]).all())
    out["identical_root_scores"] = same
    out["node_reduction"] = 1.0 - out["learned"]["nodes"] / max(out["plain"]["nodes"], 1)
    out["time_reduction"] = 1.0 - out["learned"]["seconds"] / max(out["plain"]["seconds"], 1e-9)
    out["config"] = {"positions": n, "seed": seed, "depth": depth, "workers": workers,
                     "waves_per_simd": waves_per_simd, "mlp_min_depth": mlp_min_depth, "tt_log2": tt_log2}
    out["scores"] = res["plain"]["score"].tolist()
    return out


# parameter layout of the C-ABI: W1[908][256], b1[256], W2[256][32], b2[32], W3[32], b3
_N_W1, _N_W2 = 908 * 256, 256 * 32


def save(path):
    """Checkpoint as the search uses it: W1 and W2 as BF16 bit patterns (the
    search reads them as BF16), the biases and the head as FP32."""
    p = mlp_get()
    w1 = p[:_N_W1]
    w2 = p[_N_W1 + 256:_N_W1 + 256 + _N_W2]
    rest = np.concatenate([p[_N_W1:_N_W1 + 256], p[_N_W1 + 256 + _N_W2:]])

    def to_bf16(x):
        u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    np.savez_compressed(path, w1_bf16=to_bf16(w1), w2_bf16=to_bf16(w2), rest_f32=rest.astype(np.float32))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        if "params" in z:
            p = z["params"]
        else:
            bf = lambda a: (a.astype(np.uint32) << 16).view(np.float32)  # noqa: E731
            rest = z["rest_f32"]
            p = np.concatenate([bf(z["w1_bf16"]), rest[:256], bf(z["w2_bf16"]), rest[256:]])
    if p.size != MLP_NPARAMS:
        raise ValueError("checkpoint has %d parameters, expected %d" % (p.size, MLP_NPARAMS))
    mlp_set(p)


def main():
    import argparse
    ap = argparse.ArgumentParser(description="learned move ordering: generate, train, evaluate")
    ap.add_argument("--train-positions", type=int, default=20000)
    ap.add_argument("--train-depth", type=int, default=5)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--eval-positions", type=int, default=1024)
    ap.add_argument("--eval-depth", type=int, default=7)
    ap.add_argument("--eval-workers", type=int, default=128)
    ap.add_argument("--mlp-min-depth", type=int, default=3)
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--load", default=None)
    args = ap.parse_args()
    rep = {}
    if args.load:
        load(args.load)
    else:
        held = heldout_pairs()
        rep["records"] = generate_records(n=args.train_positions, depth=args.train_depth)
        print(json.dumps(rep["records"]), flush=True)
        rep["training"] = train(steps=args.steps, lr=args.lr, held=held)
        print(json.dumps(rep["training"]), flush=True)
        if args.checkpoint:
            save(args.checkpoint)
    ev = evaluate(n=args.eval_positions, depth=args.eval_depth, workers=args.eval_workers,
                  mlp_min_depth=args.mlp_min_depth)
    ev.pop("scores")
    rep["evaluation"] = ev
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()

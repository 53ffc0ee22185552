"""Serving driver: a Poisson stream of search requests through zeta_serve_*.

Each request is a seeded random-playout position (20-60 plies) with a depth
drawn uniformly from [dmin, dmax]; arrivals are Poisson at `rate` requests
per second.  Requests the service cannot admit yet (slots or memory budget)
wait in a host backlog; latency is measured from arrival to the moment the
host collects the result, so it includes that wait.  The same stream is run
with plain and with learned ordering; root scores must be identical.
"""
import collections
import json
import time

import numpy as np

from . import Server, playouts


def make_stream(rate, n, seed=5, dmin=4, dmax=9):
    rng = np.random.default_rng(seed)
    pos = playouts(seed, n, 20, 60)
    depths = rng.integers(dmin, dmax + 1, n)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    return pos, depths, arrivals


def run_stream(pos, depths, arrivals, learned=False, slots=256, workers=16, tt_log2=26, mem_budget=0,
               mlp_min_depth=2, progress=None, max_seconds=None, grid=0):
    """Run the stream; with max_seconds, requests still running (or not yet
    admitted) when it expires are cancelled and reported as such.  grid:
    persistent wavefronts of the service (0 = as many as fit)."""
    n = len(pos)
    finish = np.full(n, np.nan)
    res = {}
    tick = {}
    backlog = collections.deque()
    i = 0
    with Server(slots=slots, workers=workers, tt_log2=tt_log2, use_mlp=learned, mlp_min_depth=mlp_min_depth,
                mem_budget=mem_budget, grid=grid) as sv:
        t0 = time.perf_counter()
        last_report = t0
        expired = False
        while len(res) < n:
            now = time.perf_counter() - t0
            if max_seconds and not expired and now > max_seconds:
                expired = True
                for t in list(tick):
                    sv.cancel(t)
                for k in backlog:
                    res[k] = None
                backlog.clear()
                while i < n:
                    res[i] = None
                    i += 1
            while i < n and arrivals[i] <= now:
                backlog.append(i)
                i += 1
            while backlog:
                k = backlog[0]
                t = sv.submit(pos[k], int(depths[k]), learned=learned)
                if t is None:
                    break
                tick[t] = k
                backlog.popleft()
            for d in sv.reap():
                k = tick.pop(int(d["ticket"]))
                finish[k] = time.perf_counter() - t0
                res[k] = d
            if progress and time.perf_counter() - last_report > 10:
                last_report = time.perf_counter()
                progress({"done": len(res), "of": n, "elapsed_s": round(last_report - t0, 1)})
            if i < n and not backlog:
                time.sleep(max(0.0, min(arrivals[i] - (time.perf_counter() - t0), 0.0005)))
            else:
                time.sleep(0.0002)
        wall = time.perf_counter() - t0
        worker_nodes = sv.stop()
    out = np.zeros(n, dtype=[("score", "<i4"), ("nodes", "<u8"), ("depth", "<i4"), ("status", "<i4")])
    for k, d in res.items():
        out[k] = (0, 0, 0, -100) if d is None else (d["score"], d["nodes"], d["depth"], d["status"])
    lat = (finish - arrivals) * 1e3
    return out, lat, wall, worker_nodes


def summarize(out, lat, wall, depths, worker_nodes):
    """Latency percentiles over the completed requests (status 0)."""
    nodes = float(worker_nodes)
    ok = out["status"] == 0
    per_depth = {}
    for d in np.unique(depths):
        m = (depths == d) & ok
        if m.any():
            per_depth[int(d)] = {"n": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50)),
                                 "p99_ms": float(np.percentile(lat[m], 99))}
    return {"requests": len(out), "completed": int(ok.sum()), "cancelled": int((out["status"] == -7).sum()),
            "not_admitted": int((out["status"] == -100).sum()), "wall_s": round(wall, 3),
            "p50_ms": float(np.percentile(lat[ok], 50)) if ok.any() else None,
            "p99_ms": float(np.percentile(lat[ok], 99)) if ok.any() else None,
            "nodes": int(nodes), "nodes_per_s": nodes / wall, "request_nodes": int(out["nodes"].sum()),
            "by_depth": per_depth}


def compare(rate, n, seed=5, dmin=4, dmax=9, progress=None, **kw):
    """Plain vs learned ordering on one stream; scores must be identical."""
    pos, depths, arrivals = make_stream(rate, n, seed, dmin, dmax)
    rep = {"rate_per_s": rate, "requests": n, "depths": [dmin, dmax], "seed": seed}
    outs = {}
    for mode in ("plain", "learned"):
        out, lat, wall, wn = run_stream(pos, depths, arrivals, learned=(mode == "learned"), progress=progress, **kw)
        outs[mode] = out
        rep[mode] = summarize(out, lat, wall, depths, wn)
        if progress:
            progress({"rate": rate, "mode": mode, **{k: v for k, v in rep[mode].items() if k != "by_depth"}})
    both = (outs["plain"]["status"] == 0) & (outs["learned"]["status"] == 0)
    rep["identical_root_scores"] = bool((outs["plain"]["score"][both] == outs["learned"]["score"][both]).all())
    rep["compared"] = int(both.sum())
    return rep


def main():
    import argparse
    from . import learned as L
    ap = argparse.ArgumentParser(description="Poisson serving benchmark (plain vs learned ordering)")
    ap.add_argument("--rates", default="50,200,2000")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--dmin", type=int, default=4)
    ap.add_argument("--dmax", type=int, default=9)
    ap.add_argument("--slots", type=int, default=256)
    ap.add_argument("--workers", type=int, default=128)
    ap.add_argument("--grid", type=int, default=0, help="persistent wavefronts (0 = as many as fit)")
    ap.add_argument("--mlp-min-depth", type=int, default=3)
    ap.add_argument("--max-seconds", type=float, default=None,
                    help="cancel what is still running after this long (per stream and mode)")
    ap.add_argument("--net", default=None, help="trained network (npz); without it the learned run uses a "
                                                "randomly initialised network")
    args = ap.parse_args()
    if args.net:
        L.load(args.net)
    else:
        from . import mlp_init
        mlp_init(1)
    prog = lambda d: print(json.dumps(d), flush=True)  # noqa: E731
    for r in [float(x) for x in args.rates.split(",")]:
        rep = compare(r, args.requests, dmin=args.dmin, dmax=args.dmax, slots=args.slots, workers=args.workers,
                      progress=prog, max_seconds=args.max_seconds, grid=args.grid, mlp_min_depth=args.mlp_min_depth)
        for mode in ("plain", "learned"):
            rep[mode].pop("by_depth", None)
        print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()

"""Benchmark: batched perft on one MI355X.

Workload (the configuration the headline metric is quoted on): 4096 positions
generated on the GPU from seed 2 by uniformly random legal playouts of 8-40
plies from the initial position, perft(4) of every position in one batched
call; u64 per-position leaf counts.  A step is one such batch.  Inputs are
resident in HBM before the timed region.

Prints ONE JSON line (rank 0, N=1).  --gpus N>1 exits: this build runs on one
GPU; positions shard across GPUs with a host-side sum (no collective).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK = 256 * 4 * 16 * 2.4e9  # VALU lane-instructions/s: 256 CU x 4 SIMD x 16 lanes x 2.4 GHz
HBM_PEAK = 8.0e12


# Work of one k_leaf position (apply the parent's move, analyse, count the
# legal moves), measured once with rocprofv3 --pmc on this workload and kept
# as constants (profiles/perft_c2_pmc.md):
#   SQ_INSTS_VALU = 1,993,137,114 wave instructions per launch of 128,916,248
#   positions (one position per lane) -> 989.5 VALU lane-instructions per
#   position; each wave64 instruction occupies a 16-lane SIMD for 4 cycles,
#   so the issue peak is 256 CU x 4 SIMD x 16 lanes x 2.4 GHz = 39.3 T/s.
#   FETCH_SIZE 777,423.5 KiB x 1.99991 + WRITE_SIZE 63,073.4 KiB x 0.99999
#   per launch, the factors from the library's copy kernel over 1 GiB
#   (FETCH_SIZE 524,311.3 KiB, WRITE_SIZE 1,048,586.0 KiB) -> 12.85 B of HBM
#   traffic per position.
VALU_LANE_INSTS_PER_POSITION = 989.5
HBM_BYTES_PER_POSITION = 12.85
# k_search: one node per wavefront; SQ_INSTS_VALU and SQ_INSTS_SALU / nodes
# over the depth-6 search leg (profiles/search_c3_pmc.md).  A CU's arbiter
# issues at most one scalar instruction per cycle (each SIMD's waves get one
# SALU slot every 4 cycles), so the SALU issue peak is 256 CU x 2.4 GHz.
VALU_WAVE_INSTS_PER_NODE = 459.8
SALU_INSTS_PER_NODE = 870.1
SALU_PEAK = 256 * 2.4e9
# the ordering network: 908 -> 256 -> 32 -> 1, dense FLOP of one scored move
# (2 x (908 x 256 + 256 x 32 + 32)); a training pair scores two moves and its
# step costs about three forward passes (forward, two backward products)
FLOP_PER_SCORED_MOVE = 2 * (908 * 256 + 256 * 32 + 32)
MFMA_BF16_PEAK = 2.5e15
NET_PATH = os.path.join(ROOT, "zeta_amd", "order_net.npz")


def log(msg):
    """Progress of the legs on stderr (the JSON line is the only stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def c4_leg(zeta_amd, depth=7, workers=128, waves_per_simd=1, mlp_min_depth=3):
    """Learned vs plain move ordering: 1024 held-out playout positions (seed
    40, 20-60 plies), depth 7, one batch each, TT 2^26 cleared before each;
    then the same search ordered by the best moves the completed search left
    in the table (order_hint): what a near-perfect ordering would save."""
    from zeta_amd import learned as L
    pos = zeta_amd.playouts(40, 1024, 20, 60)
    out = {"workload": "1024 playout positions seed 40, 20-60 plies, depth %d + quiescence, TT 2^26 x 16 B, "
                       "%d workers max per root, %d wavefront(s) per SIMD" % (depth, workers, waves_per_simd),
           "mlp_min_depth": mlp_min_depth,
           "network": os.path.relpath(NET_PATH, ROOT)}
    ref = None
    for mode in ("plain", "learned", "hinted"):
        if mode == "learned":
            L.load(NET_PATH)
        t = time.perf_counter()
        r = zeta_amd.search(pos, depth, workers=workers, tt_log2=26, clear_tt=True, use_mlp=(mode == "learned"),
                            mlp_min_depth=mlp_min_depth, order_hint=(mode == "hinted"),
                            waves_per_simd=waves_per_simd)
        dt = time.perf_counter() - t
        ctr = zeta_amd.search_counters()
        nodes = int(r["nodes"].sum())
        out[mode] = {"seconds": round(dt, 3), "nodes": nodes, "nodes_per_s": nodes / dt,
                     "completed": int((r["status"] == 0).sum()), "mlp_nodes": ctr["mlp_nodes"]}
        if ref is None:
            ref = r["score"].copy()
            out["score_checksum"] = int(ref.astype(np.int64).sum())
        else:
            out[mode]["identical_scores"] = bool((r["score"] == ref).all())
            out[mode]["node_reduction"] = round(1 - nodes / out["plain"]["nodes"], 4)
            out[mode]["time_reduction"] = round(1 - dt / out["plain"]["seconds"], 4)
        log("C4 %s: %.1f s, %d nodes" % (mode, dt, nodes))
    out["learned_vs_plain"] = {"nodes": [out["plain"]["nodes"], out["learned"]["nodes"]],
                               "ms": [round(out["plain"]["seconds"] * 1e3), round(out["learned"]["seconds"] * 1e3)],
                               "identical_scores": out["learned"]["identical_scores"]}
    return out


def mlp_leg(zeta_amd, steps=400, pairs_per_step=2048):
    """Training throughput of the ordering network's BF16 MFMA kernels on pairs
    logged by a small plain search (256 playout positions, seed 4, depth 4)."""
    zeta_amd.records_reserve(600_000)
    zeta_amd.records_clear()
    pos = zeta_amd.playouts(4, 256, 20, 60)
    zeta_amd.search(pos, 4, workers=4, tt_log2=22, clear_tt=True, log_records=True, log_shift=0)
    npairs = zeta_amd.records_count()
    zeta_amd.mlp_init(1)
    zeta_amd.mlp_train(20, pairs_per_step, 2e-4, 7)
    zeta_amd.synchronize()
    zeta_amd.stats_reset()
    zeta_amd.set_timing(True)
    t = time.perf_counter()
    zeta_amd.mlp_train(steps, pairs_per_step, 2e-4, 11)
    zeta_amd.synchronize()
    dt = time.perf_counter() - t
    kst = zeta_amd.stats()
    zeta_amd.set_timing(False)
    zeta_amd.stats_reset()
    zeta_amd.records_clear()
    flop = 3.0 * 2 * pairs_per_step * FLOP_PER_SCORED_MOVE * steps
    kms = {k: round(v[1], 3) for k, v in kst.items() if v[0] and k.startswith("mlp")}
    gemm_s = (kms.get("mlp_fwd", 0) + kms.get("mlp_bwd", 0)) / 1e3
    log("MLP training: %d steps in %.2f s" % (steps, dt))
    return {"workload": "%d Adam steps of %d pairs (logged by a depth-4 search of 256 seed-4 positions, %d pairs)"
                        % (steps, pairs_per_step, npairs),
            "seconds": round(dt, 3), "pairs_per_s": steps * pairs_per_step / dt, "kernels_ms_total": kms,
            "roofline": {"bound": "mfma", "achieved": round(flop / max(gemm_s, 1e-9) / 1e12, 3),
                         "peak": MFMA_BF16_PEAK / 1e12, "unit": "TFLOP/s",
                         "frac": round(flop / max(gemm_s, 1e-9) / MFMA_BF16_PEAK, 5),
                         "flop_per_step": 3 * 2 * pairs_per_step * FLOP_PER_SCORED_MOVE,
                         "note": "dense FLOP over the summed time of the forward and backward kernels"}}


def c5_leg(zeta_amd, rate=2000.0, n=48, max_seconds=60.0, workers=128, grid=0):
    """Serving: a Poisson stream (seed 5) of n requests at `rate`, depths
    uniform in 4-9, plain then learned ordering, each stream capped at
    max_seconds (requests still running then are cancelled and reported);
    up to `workers` wavefronts per request on a persistent grid of `grid`
    wavefronts (0 = as many as fit)."""
    from zeta_amd import learned as L
    from zeta_amd import serving as S
    L.load(NET_PATH)
    rep = S.compare(rate, n, seed=5, dmin=4, dmax=9, slots=256, workers=workers, max_seconds=max_seconds,
                    mlp_min_depth=3, grid=grid, progress=lambda d: log("C5 " + json.dumps(d)))
    rep["workers_per_request"] = workers
    rep["grid_wavefronts"] = grid
    for mode in ("plain", "learned"):
        rep[mode].pop("by_depth", None)
    return rep


def host_info():
    """The host the CPU baselines ran on."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model}


def cpu_baseline(pos, depth, threads):
    """The oracle's perft of the whole workload on `threads` host cores, with
    bulk counting at depth 1 as the device does (like for like with the
    headline), and the same walk making every leaf move."""
    from oracle import refcpu
    sample = pos  # the whole workload: perft(4) of 4096 positions is a few seconds on 16 threads
    t0 = time.perf_counter()
    c = refcpu.perft_batch(sample, depth, threads, bulk=True)
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    c2 = refcpu.perft_batch(sample, depth, threads)
    dt2 = time.perf_counter() - t1
    assert int(c.sum()) == int(c2.sum())
    return {"value": float(c.sum()) / dt, "unit": "leaves/s", "cores": threads, "kind": "port",
            "bulk_counting": True,
            "sample": "oracle perft(%d) of all %d workload positions with bulk counting (%d leaves, %.1f s, "
                      "%d threads)" % (depth, len(sample), int(c.sum()), dt, threads),
            "leaf_visiting": {"value": float(c2.sum()) / dt2, "unit": "leaves/s", "seconds": round(dt2, 2),
                              "note": "every leaf move made (no bulk counting)"},
            "host": host_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--depth", type=int, default=4)
    ap.add_argument("--positions", type=int, default=4096)
    ap.add_argument("--nobulk", action="store_true", help="visit every leaf (no bulk counting)")
    ap.add_argument("--nobulk-steps", type=int, default=3,
                    help="timed steps of the same workload without bulk counting, reported beside the headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-search", action="store_true", help="skip the fixed-depth search leg")
    ap.add_argument("--search-depth", type=int, default=6)
    ap.add_argument("--search-workers", type=int, default=128)
    ap.add_argument("--search-waves", type=int, default=1, help="persistent search wavefronts per SIMD (0 = all "
                    "that fit)")
    ap.add_argument("--no-c4", action="store_true", help="skip the learned-vs-plain ordering leg (depth 7)")
    ap.add_argument("--no-c5", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-mlp", action="store_true", help="skip the ordering network's training leg")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="BYTES",
                    help="also run the library's streaming copy kernel once over BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (positions shard across GPUs with a "
              "host-side sum; no multi-GPU run is part of it)", file=sys.stderr)
        sys.exit(2)

    # torch is not imported: its wheel carries its own HIP runtime, and two HIP
    # runtimes in one process do not share the device.  Device memory and
    # synchronisation go through the library's C-ABI instead.
    import zeta_amd

    zeta_amd.init()
    pos = zeta_amd.playouts(2, args.positions, 8, 40)
    d_pos = zeta_amd.DeviceBuffer(pos.nbytes)
    d_cnt = zeta_amd.DeviceBuffer(8 * args.positions)
    d_pos.upload(pos)

    def step():
        zeta_amd.perft_device(d_pos.ptr, args.positions, args.depth, d_cnt.ptr, nobulk=args.nobulk)

    for _ in range(args.warmup):
        step()
    zeta_amd.synchronize()
    counts = d_cnt.download(np.zeros(args.positions, dtype=np.uint64))
    leaves = int(counts.sum())
    zeta_amd.set_timing(True)
    step()
    kst = zeta_amd.stats()
    zeta_amd.set_timing(False)
    zeta_amd.stats_reset()

    zeta_amd.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    zeta_amd.synchronize()
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / args.steps
    value = leaves * args.steps / dt

    # roofline of the dominant kernel, from its HIP-event time (separate timed pass)
    leaf = kst.get("perft_leaf", (0, 0.0, 0))
    roof = None
    if leaf[0] and not args.nobulk:
        per_launch_items = leaf[2] / leaf[0]
        avg_s = leaf[1] / leaf[0] / 1e3
        achieved = per_launch_items * VALU_LANE_INSTS_PER_POSITION / avg_s
        roof = {"bound": "valu", "achieved": round(achieved / 1e12, 3), "peak": round(VALU_PEAK / 1e12, 2),
                "unit": "T VALU lane-instructions/s", "frac": round(achieved / VALU_PEAK, 4),
                "traffic": int(per_launch_items * HBM_BYTES_PER_POSITION),
                "kernel": "perft_leaf", "avg_launch_ms": round(leaf[1] / leaf[0], 4),
                "items_per_launch": int(per_launch_items),
                "valu_lane_insts_per_item": VALU_LANE_INSTS_PER_POSITION,
                "hbm_GBps": round(per_launch_items * HBM_BYTES_PER_POSITION / avg_s / 1e9, 1),
                # the HBM roofline of the same kernel, for comparison: far from binding
                "hbm": {"achieved": round(per_launch_items * HBM_BYTES_PER_POSITION / avg_s / 1e9, 1),
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": round(per_launch_items * HBM_BYTES_PER_POSITION / avg_s / HBM_PEAK, 4)}}

    # the same workload without bulk counting: every leaf is generated, as the
    # reference's kernel visits it (src/zetaperft.cl:1316-1326)
    nobulk = None
    if not args.nobulk and args.nobulk_steps > 0:
        def step_nb():
            zeta_amd.perft_device(d_pos.ptr, args.positions, args.depth, d_cnt.ptr, nobulk=True)
        step_nb()
        zeta_amd.synchronize()
        nb_leaves = int(d_cnt.download(np.zeros(args.positions, dtype=np.uint64)).sum())
        assert nb_leaves == leaves
        zeta_amd.stats_reset()
        zeta_amd.set_timing(True)
        step_nb()
        kst_nb = zeta_amd.stats()
        zeta_amd.set_timing(False)
        zeta_amd.stats_reset()
        zeta_amd.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.nobulk_steps):
            step_nb()
        zeta_amd.synchronize()
        dt_nb = time.perf_counter() - t0
        nobulk = {"leaves_per_s": nb_leaves * args.nobulk_steps / dt_nb,
                  "ms_per_step": dt_nb * 1e3 / args.nobulk_steps, "steps": args.nobulk_steps,
                  "kernels_ms_total": {k: round(v[1], 3) for k, v in kst_nb.items() if v[0]},
                  "leaf_launches": kst_nb.get("perft_leaf", (0, 0, 0))[0]}

    calib = None
    if args.calib_copy:
        nb = (args.calib_copy + 15) // 16 * 16
        a, b = zeta_amd.DeviceBuffer(nb), zeta_amd.DeviceBuffer(nb)
        zeta_amd.copy_device(b.ptr, a.ptr, nb)
        a.free()
        b.free()
        calib = {"copy_bytes": nb}

    # config 1: startpos perft(5), one root (the reference's own perft command)
    start = np.zeros(1, dtype=zeta_amd.POS_DTYPE)
    start["b"][0] = [0xFFFF000000000000, 0x91FF00000000FF91, 0x5A0000000000005A, 0xAD000000000000AD,
                     0x6EFFFFFFFFFFFF6E, 0, 0]
    zeta_amd.perft(start, 5)
    t1 = time.perf_counter()
    c1 = int(zeta_amd.perft(start, 5)[0])
    c1_ms = (time.perf_counter() - t1) * 1e3

    # config 3: fixed-depth alpha-beta + quiescence, 1024 positions (seed 3,
    # 20-60 plies), iterative deepening 1..depth, TT 2^26 x 16 B, ABDADA helpers
    search = None
    if not args.no_search:
        spos = zeta_amd.playouts(3, 1024, 20, 60)
        t2 = time.perf_counter()
        res = zeta_amd.search(spos, args.search_depth, workers=args.search_workers, tt_log2=26, clear_tt=True,
                              waves_per_simd=args.search_waves)
        ttd = time.perf_counter() - t2
        ctr = zeta_amd.search_counters()
        nodes = int(res["nodes"].sum())
        tt_bytes = 16 * (ctr["tt_probes"] + ctr["tt_stores"])
        search = {"workload": "1024 playout positions seed 3, 20-60 plies, depth %d + quiescence, TT 2^26 x 16 B, "
                              "%d workers max per root, %d wavefront(s) per SIMD"
                              % (args.search_depth, args.search_workers, args.search_waves),
                  "time_to_depth_s": round(ttd, 3), "nodes": nodes, "nodes_per_s": nodes / ttd,
                  "score_checksum": int(res["score"].astype(np.int64).sum()),
                  "tt_GBps": round(tt_bytes / ttd / 1e9, 3), "tt_frac_of_hbm": round(tt_bytes / ttd / HBM_PEAK, 6),
                  "roofline": {
                      "bound": "salu", "unit": "G SALU instructions/s",
                      "achieved": round(ctr["worker_nodes"] * SALU_INSTS_PER_NODE / ttd / 1e9, 2),
                      "peak": round(SALU_PEAK / 1e9, 1),
                      "frac": round(ctr["worker_nodes"] * SALU_INSTS_PER_NODE / ttd / SALU_PEAK, 4),
                      "salu_insts_per_node": SALU_INSTS_PER_NODE,
                      "valu": {"unit": "T VALU lane-instructions/s",
                               "achieved": round(ctr["worker_nodes"] * VALU_WAVE_INSTS_PER_NODE * 64 / ttd / 1e12, 3),
                               "peak": round(VALU_PEAK / 1e12, 2),
                               "frac": round(ctr["worker_nodes"] * VALU_WAVE_INSTS_PER_NODE * 64 / ttd / VALU_PEAK,
                                             4),
                               "valu_wave_insts_per_node": VALU_WAVE_INSTS_PER_NODE},
                      "note": "issue fractions over the time to depth; one wavefront walks one tree, so the kernel "
                              "is latency-bound and both fractions stay well below 1"},
                  **ctr}

    if search is not None:
        log("C3 search: %.1f s" % search["time_to_depth_s"])
    c4 = None if args.no_c4 else c4_leg(zeta_amd)
    mlp = None if args.no_mlp else mlp_leg(zeta_amd)
    c5 = None if args.no_c5 else c5_leg(zeta_amd)

    cpu = None
    c1_cpu = None
    if not args.no_cpu_baseline:
        # the host cores this run may use: all of them here, the GPU's share
        # (16 cores) on a shared GPU host
        threads = min(16, len(os.sched_getaffinity(0)))
        cpu = cpu_baseline(pos, args.depth, threads)
        # config 1 on one core: the reference runs it as one work-group; the
        # oracle's move generator is that kernel's 64-lane square-wise
        # generator, one lane after the other
        from oracle import refcpu
        best = None
        for _ in range(3):
            t4 = time.perf_counter()
            cc = int(refcpu.perft_batch(start, 5, 1)[0])
            dt4 = time.perf_counter() - t4
            best = dt4 if best is None else min(best, dt4)
        c1_cpu = {"value": cc / best, "unit": "leaves/s", "cores": 1, "kind": "port", "leaves": cc,
                  "sample": "oracle startpos perft(5) on one thread, best of 3 (%.3f s)" % best}
        if search is not None:
            # the oracle's alpha-beta (same node definition: moves made, the
            # same iterative deepening 1..depth) on a bounded sample of the
            # same positions at the same depth
            from oracle import refcpu
            t3 = time.perf_counter()
            cs, cn = refcpu.search_id_batch(spos[:128], args.search_depth, threads)
            dt3 = time.perf_counter() - t3
            search["cpu_baseline"] = {"value": float(cn.sum()) / dt3, "unit": "nodes/s", "cores": threads,
                                      "kind": "port",
                                      "sample": "oracle iterative deepening 1..%d of the first 128 positions "
                                                "(%d nodes, %.1f s, %d threads; no TT)"
                                                % (args.search_depth, int(cn.sum()), dt3, threads),
                                      "scores_equal_device": bool((cs == res["score"][:128]).all())}
        log("CPU baselines done")

    out = {"metric": "perft nodes/s", "value": value, "unit": "leaves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded random playouts, generated on GPU)",
           "config": {"workload": "batched perft(%d), %d playout positions seed 2, 8-40 plies%s"
                                  % (args.depth, args.positions, ", no bulk counting" if args.nobulk else ""),
                      "positions": args.positions, "depth": args.depth, "leaves_per_step": leaves,
                      "bulk_counting": not args.nobulk},
           "kernels_ms": {k: round(v[1] / max(v[0], 1), 4) for k, v in kst.items() if v[0]},
           "roofline": roof, "cpu_baseline": cpu, "nobulk": nobulk,
           "config1": {"workload": "startpos perft(5), one root", "leaves": c1, "ms": round(c1_ms, 3),
                       "leaves_per_s": c1 / (c1_ms / 1e3), "cpu_baseline": c1_cpu},
           "search": search, "learned_ordering": c4, "mlp_training": mlp, "serving": c5}
    if calib:
        out["calibration"] = calib
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Writes the golden fixtures under tests/golden/ with the CPU oracle.

Run from the repo root after `make -C oracle`:
    python tests/golden/make_golden.py [perft] [search]

Fixtures (all data, no code):
  kats.json            the reference's 14 perft known-answer tests
                       (src/test.c:43-73) plus textbook perft values that the
                       reference's semantics must also produce.
  c2_perft4.npz        config 2: 4096 playout positions (seed 2, 8-40 plies),
                       root hash + perft(4) leaf count of each (oracle).
  c3_search6.npz       config 3: 1024 playout positions (seed 3, 20-60 plies),
                       root hash + depth-6 root score of each (oracle).
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import refcpu as R  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

# src/test.c:43-73 (FEN, depth, count); the reference's own known answers.
REFERENCE_KATS = [
    ("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq -", [20, 400, 8902, 197281]),
    ("8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - ", [14, 191, 2812, 43238]),
    ("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - ", [48, 2039, 97862]),
    ("r4rk1/1pp1qppp/p1np1n2/2b1p1B1/2B1P1b1/P1NP1N2/1PP1QPPP/R4RK1 w - -", [46, 2079, 89890]),
]
# Textbook perft values whose trees contain no promotion and no FEN e.p.
# square, so the reference's quirks cannot change them.
TEXTBOOK = [
    ("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", 5, 4865609),
    ("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", 6, 119060324),
    ("8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - 0 1", 5, 674624),
    ("r4rk1/1pp1qppp/p1np1n2/2b1p1B1/2B1P1b1/P1NP1N2/1PP1QPPP/R4RK1 w - - 0 10", 4, 3894594),
]
# Quirk cases: (FEN, depth, reference-semantics count, textbook count, quirk)
QUIRKS = [
    ("4k3/P7/8/8/8/8/8/4K3 w - - 0 1", 1, 6, 9, "Q1 queen-only promotion"),
    ("rnbq1k1r/pp1Pbppp/2p5/8/2B5/8/PPP1NnPP/RNBQK2R w KQ - 1 8", 1, 41, 44, "Q1 queen-only promotion"),
    ("rnbqkbnr/ppp1p1pp/8/3pPp2/8/8/PPPP1PPP/RNBQKBNR w KQkq f6 0 3", 1, 30, 31, "Q3 FEN e.p. ignored"),
]


def write_kats():
    out = {"reference_kats": [], "textbook": [], "quirks": []}
    for fen, counts in REFERENCE_KATS:
        p = R.setboard(fen)
        got = [R.perft(p, d + 1) for d in range(len(counts))]
        assert got == counts, (fen, got, counts)
        out["reference_kats"].append({"fen": fen, "counts": counts})
    for fen, d, c in TEXTBOOK:
        got = R.perft(R.setboard(fen), d)
        assert got == c, (fen, d, got, c)
        out["textbook"].append({"fen": fen, "depth": d, "count": c})
    for fen, d, c, tb, why in QUIRKS:
        got = R.perft(R.setboard(fen), d)
        assert got == c, (fen, got, c)
        out["quirks"].append({"fen": fen, "depth": d, "count": c, "textbook": tb, "quirk": why})
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump(out, f, indent=1)


def write_c2():
    t = time.time()
    pos = R.gen_playouts(2, 4096, 8, 40)
    counts = R.perft_batch(pos, 4)
    c3 = R.perft_batch(pos, 3)
    np.savez_compressed(os.path.join(HERE, "c2_perft4.npz"), hash=pos["b"][:, 5].copy(),
                        perft4=counts, perft3=c3)
    print("c2: %d positions, total perft4 %d, %.1fs" % (len(pos), int(counts.sum()), time.time() - t))


def write_c3(n=1024, depth=6):
    t = time.time()
    pos = R.gen_playouts(3, n, 20, 60)
    scores, nodes = R.search_batch(pos, depth)
    s4, _ = R.search_batch(pos, 4)
    np.savez_compressed(os.path.join(HERE, "c3_search6.npz"), hash=pos["b"][:, 5].copy(),
                        score6=scores, score4=s4, oracle_nodes6=nodes)
    print("c3: %d positions depth %d, %.1fs, oracle nodes %d" % (n, depth, time.time() - t, int(nodes.sum())))


if __name__ == "__main__":
    what = sys.argv[1:] or ["perft"]
    if "perft" in what:
        write_kats()
        write_c2()
    if "search" in what:
        write_c3()

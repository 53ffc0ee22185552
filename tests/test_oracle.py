"""CPU tests pinning the oracle (oracle/refcpu.c) before it is trusted as the
checker: the reference's own known-answer tests, textbook values its quirks
cannot touch, hand-checked table entries and closed forms, and brute force."""
import json
import os

import numpy as np
import pytest

from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kats.json")))


@pytest.mark.parametrize("case", KATS["reference_kats"], ids=lambda c: c["fen"][:20])
def test_reference_kats(case):
    # src/test.c:43-73 -- the reference's only correctness test
    p = R.setboard(case["fen"])
    assert [R.perft(p, d + 1) for d in range(len(case["counts"]))] == case["counts"]


@pytest.mark.parametrize("case", KATS["textbook"], ids=lambda c: "%s d%d" % (c["fen"][:12], c["depth"]))
def test_textbook(case):
    assert R.perft(R.setboard(case["fen"]), case["depth"]) == case["count"]


@pytest.mark.parametrize("case", KATS["quirks"], ids=lambda c: c["quirk"])
def test_quirk_counts(case):
    got = R.perft(R.setboard(case["fen"]), case["depth"])
    assert got == case["count"] and got != case["textbook"]


def _play(pos, *cans):
    for can in cans:
        for k in range(R.count_legal(pos)):
            m = R.nth_legal(pos, k)
            if R.move2can(m) == can:
                pos = R.domove(pos, m)
                break
        else:
            raise AssertionError("illegal move " + can)
    return pos


def test_q3_ep_kept_when_reached_by_moves():
    # same position as the FEN quirk case, reached by moves: e.p. capture exists
    p = _play(R.setboard(R.STARTPOS), "e2e4", "d7d5", "e4e5", "f7f5")
    assert R.perft(p, 1) == 31
    mv, _ = R.divide(p, 1)
    assert "e5f6" in [R.move2can(int(m)) for m in mv]


def test_q1_promotion_is_queen():
    p = R.setboard("4k3/P7/8/8/8/8/8/4K3 w - - 0 1")
    promos = [int(m) for m in R.divide(p, 1)[0] if (int(m) >> 6) & 63 == 56]
    assert len(promos) == 1 and ((promos[0] >> 22) & 15) >> 1 == 6


def test_q2_castle_precondition_uses_d8():
    # SMCRALL = 0x8900000000000091 holds a1,e1,h1,a8,d8,h8 (src/types.h:114).
    # setboard leaves d8 flagged, so castling is generated; clearing every
    # SMCRALL bit (and e8) suppresses it.
    p = R.setboard("r3k2r/8/8/8/8/8/8/R3K2R w KQkq - 0 1")
    assert R.perft(p, 1) == 26
    q = p.copy()
    q["b"][0, 4] &= ~np.uint64(0x8900000000000091 | (1 << 60))
    assert R.perft(q, 1) == 24


def test_tables_known_entries():
    # hand-checkable entries of the reference's tables (src/zetaperft.cl:612-636)
    assert R.lib().rc_attack_table(0) == 0x200               # white pawn a1
    assert R.lib().rc_attack_table(64 + 8) == 0x2            # black pawn a2
    assert R.lib().rc_attack_table(128) == 0x20400           # knight a1
    assert R.lib().rc_attack_table(192) == 0x302             # king a1
    assert R.lib().rc_attack_table(256) == 0x8040201008040200  # bishop a1
    assert R.lib().rc_attack_table(320) == 0x1010101010101FE   # rook a1
    assert R.lib().rc_attack_table(384) == 0x81412111090503FE  # queen a1
    assert R.lib().rc_pawnpush_table(8) == 0x1010000         # white a2
    assert R.lib().rc_pawnpush_table(64 + 48) == 0x10100000000  # black a7
    assert R.lib().rc_line(0, 1) == 0xFF and R.lib().rc_line(0, 9) == 0x8040201008040201
    assert R.lib().rc_line(0, 0) == 0 and R.lib().rc_line(0, 10) == 0


def test_tables_invariants():
    L = R.lib()
    tot = [sum(bin(L.rc_attack_table(t * 64 + s)).count("1") for s in range(64)) for t in range(7)]
    assert tot == [98, 98, 336, 420, 560, 896, 1456]  # pawns: 7 ranks x 14
    for a in range(64):
        for b in range(64):
            assert L.rc_inbetween(a, b) == L.rc_sqinbetween(a, b)  # closed form, src/bit.c:49-66
            ln = L.rc_line(a, b)
            assert ln == L.rc_line(b, a)
            if ln:
                assert ln >> a & 1 and ln >> b & 1


def test_eval_closed_forms():
    assert R.evaluate(R.setboard(R.STARTPOS)) == 0
    # two kings: +14 each side (10 + PST -1 + control 5), symmetric
    assert R.evaluate(R.setboard("4k3/8/8/8/8/8/8/4K3 w - - 0 1")) == 0
    # + white knight d4: 10 + 400 + PST 20 + control 10
    assert R.evaluate(R.setboard("4k3/8/8/8/3N4/8/8/4K3 w - - 0 1")) == 440
    # bishop pair bonus: two white bishops c1,f1 (each 10+400-10+5) + 25
    assert R.evaluate(R.setboard("4k3/8/8/8/8/8/8/2B1KB2 w - - 0 1")) == 2 * 405 + 25
    # doubled white pawns e2,e3: e2 = 10+100-5+5, e3 = 10+100+2+5 -30 (own pawn behind)
    assert R.evaluate(R.setboard("4k3/8/8/8/8/4P3/4P3/4K3 w - - 0 1")) == 110 + 87


def _mirror(pos):
    """Colour-flipped mirror image (ranks reversed, colours swapped)."""
    b = [int(x) for x in pos["b"][0]]
    def flip(x):
        return int.from_bytes(x.to_bytes(8, "little")[::-1], "little")
    occ = b[1] | b[2] | b[3]
    q = pos.copy()
    q["b"][0, 0] = flip(occ ^ b[0])
    for i in (1, 2, 3):
        q["b"][0, i] = flip(b[i])
    q["stm"][0] = 1 - int(pos["stm"][0])
    return q


def test_eval_mirror_antisymmetry():
    pos = R.gen_playouts(11, 200, 8, 60)
    for i in range(len(pos)):
        p = pos[i:i + 1]
        assert R.evaluate(_mirror(p)) == -R.evaluate(p)


def test_hash_incremental_equals_scratch():
    pos = R.gen_playouts(5, 50, 8, 40)
    for i in range(len(pos)):
        p = pos[i:i + 1]
        assert int(p["b"][0, 5]) == R.lib().rc_computehash(p["b"][0].ctypes.data, int(p["stm"][0]))


ENDGAMES = [
    "8/8/4k3/8/8/3K4/3R4/8 w - - 0 1", "8/5k2/8/8/2q5/8/3KR3/8 w - - 0 1",
    "8/8/8/3k4/8/3PK3/8/8 w - - 0 1", "8/1k6/8/2n5/8/3B4/2K1N3/8 b - - 0 1",
    "6k1/5pp1/8/8/8/8/5PP1/3R2K1 w - - 0 1", "8/2k5/2p5/8/1P6/2K5/8/8 w - - 0 1",
    "4k3/8/8/3q4/8/8/3Q4/4K3 w - - 0 1", "2r3k1/8/8/8/8/8/1B6/2R3K1 b - - 0 1",
]


@pytest.mark.parametrize("fen", ENDGAMES)
def test_search_equals_minimax(fen):
    # fail-hard alpha-beta with a full root window returns the minimax value
    p = R.setboard(fen)
    for d in (1, 2, 3):
        assert R.search(p, d)[0] == R.minimax(p, d)[0]


NO_CASTLE = 0x9100000000000091  # moved flags of a1, e1, h1, a8, e8, h8


def _mirror_full(pos):
    """Colour-flipped mirror image with the moved flags mirrored too and the
    hash recomputed; castling rights are removed first (on both sides), since
    the castle precondition masks d8 and not e8 (quirk Q2) and so is not
    colour-symmetric."""
    q = pos.copy()
    q["b"][0, 4] = int(q["b"][0, 4]) | NO_CASTLE
    m = _mirror(q)
    m["b"][0, 4] = int.from_bytes(int(q["b"][0, 4]).to_bytes(8, "little")[::-1], "little")
    for r in (q, m):
        r["b"][0, 5] = R.lib().rc_computehash(r["b"][0].ctypes.data, int(r["stm"][0]))
    return q, m


def test_search_mirror_symmetric():
    # a position and its colour-flipped mirror have the same value for the
    # side to move and the same perft counts: a colour-dependent mistake in
    # pawn directions, promotions, e.p. or the evaluation's sign breaks this
    pos = R.gen_playouts(13, 24, 10, 60)
    for i in range(len(pos)):
        p, m = _mirror_full(pos[i:i + 1])
        assert R.search(p, 4)[0] == R.search(m, 4)[0], i
        assert R.perft(p, 3) == R.perft(m, 3), i


def test_search_order_independent():
    # the root score may not depend on move order (full-width nodes worst first)
    pos = R.gen_playouts(7, 12, 10, 50)
    for i in range(len(pos)):
        p = pos[i:i + 1]
        assert R.search(p, 3)[0] == R.search_reversed(p, 3)[0]


def test_search_terminal_scores():
    INF = 32000
    # mate in one (back rank): root score INF - 1 (child mated at ply 1)
    assert R.search(R.setboard("6k1/5ppp/8/8/8/8/5PPP/R5K1 w - - 0 1"), 2)[0] == INF - 1
    # mate in two (1.Re8+ Rxe8 2.Rxe8#): INF - 3, for either colour, and
    # minimax agrees; the side being mated, one reply before: -(INF - 2)
    for fen in ("r5k1/5ppp/8/8/8/8/4RPPP/4R1K1 w - - 0 1", "4r1k1/4rppp/8/8/8/8/5PPP/R5K1 b - - 0 1"):
        p = R.setboard(fen)
        assert R.search(p, 3)[0] == INF - 3 == R.minimax(p, 3)[0]
    assert R.search(R.setboard("r3R1k1/5ppp/8/8/8/8/5PPP/4R1K1 b - - 1 1"), 3)[0] == -(INF - 2)
    # mated root / stalemate root: PV[0] stays DRAWSCORE
    assert R.search(R.setboard("R5k1/5ppp/8/8/8/8/5PPP/6K1 b - - 0 1"), 3)[0] == 0
    assert R.search(R.setboard("7k/5Q2/6K1/8/8/8/8/8 b - - 0 1"), 3)[0] == 0
    # in check, the only legal move captures the queen: bare kings, draw
    assert R.search(R.setboard("8/8/8/8/3k4/8/3q4/3K4 w - - 0 1"), 3)[0] == 0
    # fifty-move rule: every reply reaches hmc 100 -> draw, despite an extra rook
    assert R.search(R.setboard("7k/8/8/8/8/8/8/R6K w - - 99 80"), 3)[0] == 0
    assert R.search(R.setboard("7k/8/8/8/8/8/8/R6K w - - 0 80"), 3)[0] > 500


def test_playouts_deterministic_and_legal():
    a = R.gen_playouts(2, 64, 8, 40)
    b = R.gen_playouts(2, 64, 8, 40)
    assert (a["b"] == b["b"]).all()
    for i in range(len(a)):
        p = a[i:i + 1]
        assert R.count_legal(p) > 0 and int(p["b"][0, 6]) < 100


def test_bulk_counting_equals_leaf_visiting(R=None):
    """The oracle's perft with bulk counting at depth 1 (the CPU baseline of
    the bulk-counting headline) counts exactly the leaves it visits without."""
    from oracle import refcpu
    pos = refcpu.gen_playouts(2, 64, 8, 40)
    for d in (1, 2, 3):
        assert (refcpu.perft_batch(pos, d, 4, bulk=True) == refcpu.perft_batch(pos, d, 4)).all()

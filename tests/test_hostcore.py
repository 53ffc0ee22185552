"""CPU tests of the device chess core (zeta_amd/csrc/chess.h) through a host
build of the same source (tests/hostcore/chess_host.cpp): the set-wise move
generator must produce exactly the oracle's move lists (codes included)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "hostcore", "libchesshost.so")
SRC = os.path.join(HERE, "hostcore", "chess_host.cpp")
HDR = os.path.join(os.path.dirname(HERE), "zeta_amd", "csrc", "chess.h")


@pytest.fixture(scope="module")
def H():
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(SRC), os.path.getmtime(HDR)):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", SO, SRC])
    L = ctypes.CDLL(SO)
    vp = ctypes.c_void_p
    L.hc_count.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.hc_enumerate.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.hc_enumerate_sets.argtypes = [vp, ctypes.c_int, vp]
    L.hc_nth_bit.argtypes = [ctypes.c_uint64, ctypes.c_int]
    u64 = ctypes.c_uint64
    L.hc_castle.argtypes = [u64, u64, u64, u64, u64, ctypes.c_int, ctypes.c_int, u64, u64, u64]
    L.hc_castle.restype = u64
    L.hc_castle_per_square.argtypes = L.hc_castle.argtypes
    L.hc_castle_per_square.restype = u64
    L.hc_perft.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.hc_perft.restype = ctypes.c_uint64
    L.hc_enumerate_lane.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
    L.hc_eval.argtypes = [vp]
    L.hc_eval.restype = ctypes.c_int32
    L.hc_slider.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    L.hc_slider.restype = ctypes.c_uint64
    L.hc_slider_ks.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int]
    L.hc_slider_ks.restype = ctypes.c_uint64
    L.hc_perft_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
    L.hc_hash.argtypes = [vp, ctypes.c_int]
    L.hc_hash.restype = ctypes.c_uint64
    return L


def _oracle_moves(p, qs):
    _, out, _ = R.genmoves(p, qs)
    b = np.ascontiguousarray(p["b"][0])
    mv = []
    for lid in range(64):
        m = int(out[lid])
        while m:
            t = (m & -m).bit_length() - 1
            m &= m - 1
            mv.append(int(R.lib().rc_makemove_code(b.ctypes.data, int(p["stm"][0]), lid, t)))
    return mv


def test_movelists_match_oracle(H):
    pos = np.concatenate([R.gen_playouts(2, 1500, 8, 40), R.gen_playouts(3, 500, 20, 60)])
    buf = np.zeros(256, dtype=np.uint32)
    for i in range(len(pos)):
        p = pos[i:i + 1]
        b = np.ascontiguousarray(p["b"][0])
        for qs in (0, 1):
            om = _oracle_moves(p, qs)
            n = H.hc_enumerate(b.ctypes.data, int(p["stm"][0]), qs, buf.ctypes.data)
            assert H.hc_count(b.ctypes.data, int(p["stm"][0]), qs) == len(om)
            assert list(buf[:n]) == om  # same codes, same canonical order


def test_perft_matches_oracle(H):
    pos = R.gen_playouts(2, 120, 8, 40)
    want = R.perft_batch(pos, 3)
    for i in range(len(pos)):
        b = np.ascontiguousarray(pos[i]["b"])
        assert H.hc_perft(b.ctypes.data, int(pos[i]["stm"]), 3) == want[i]


def test_special_positions(H):
    fens = ["R6R/3Q4/1Q4Q1/4Q3/2Q4Q/Q4Q2/pp1Q4/kBNN1KB1 w - - 0 1",  # 218 moves
            "r3k2r/8/8/8/8/8/8/R3K2R w KQkq - 0 1", "8/8/8/8/k2Pp2Q/8/8/3K4 b - - 0 1",
            "8/8/8/2k5/3Pp3/8/8/4K2R b K - 0 1", "7k/5Q2/6K1/8/8/8/8/8 b - - 0 1"]
    for fen in fens:
        p = R.setboard(fen)
        b = np.ascontiguousarray(p["b"][0])
        for d in (1, 2, 3):
            assert H.hc_perft(b.ctypes.data, int(p["stm"][0]), d) == R.perft(p, d)


SPECIAL_FENS = ["R6R/3Q4/1Q4Q1/4Q3/2Q4Q/Q4Q2/pp1Q4/kBNN1KB1 w - - 0 1",  # 218 moves
                "r3k2r/8/8/8/8/8/8/R3K2R w KQkq - 0 1", "8/8/8/8/k2Pp2Q/8/8/3K4 b - - 0 1",
                "8/8/8/2k5/3Pp3/8/8/4K2R b K - 0 1", "7k/5Q2/6K1/8/8/8/8/8 b - - 0 1",
                "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1",
                "8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - 0 1",
                "n1n5/PPPk4/8/8/8/8/4Kppp/5N1N b - - 0 1"]


def _grouped_vs_oracle(H, p, buf):
    b = np.ascontiguousarray(p["b"][0])
    n = H.hc_enumerate_sets(b.ctypes.data, int(p["stm"][0]), buf.ctypes.data)
    om = _oracle_moves(p, 0)
    assert n == len(om) == H.hc_count(b.ctypes.data, int(p["stm"][0]), 0)
    assert sorted(buf[:n].tolist()) == sorted(om)  # same codes; the order is by kind


def test_grouped_enumeration_matches_oracle(H):
    # the perft expansion kernel's enumeration (by move kind, origin recovered
    # from each target) gives exactly the oracle's legal moves: playouts, and
    # every position within two plies of positions rich in castling, e.p.,
    # promotions, pins and checks
    buf = np.zeros(256, dtype=np.uint32)
    pos = np.concatenate([R.gen_playouts(5, 1500, 8, 40), R.gen_playouts(6, 500, 20, 60)])
    for i in range(len(pos)):
        _grouped_vs_oracle(H, pos[i:i + 1], buf)
    for fen in SPECIAL_FENS:
        root = R.setboard(fen)
        _grouped_vs_oracle(H, root, buf)
        for m in _oracle_moves(root, 0):
            c = R.domove(root, m)
            _grouped_vs_oracle(H, c, buf)
            for m2 in _oracle_moves(c, 0)[:40]:
                _grouped_vs_oracle(H, R.domove(c, m2), buf)


def test_nth_bit(H):
    # the search's move-list builder takes the k-th target of a lane's set
    rng = np.random.default_rng(11)
    xs = [1, 1 << 63, (1 << 64) - 1, 0x8000000100000001] + [
        int(rng.integers(1, 2**63, dtype=np.int64)) ^ (int(rng.integers(0, 2, dtype=np.int64)) << 63) for _ in range(2000)]
    for x in xs:
        bits = [i for i in range(64) if (x >> i) & 1]
        for k, b in enumerate(bits):
            assert H.hc_nth_bit(x, k) == b, (hex(x), k)


def test_castle_shifted_constants(H):
    # castle_targets' fast path (kings on squares 4-60) against the per-square
    # rule, on random boards, moved flags and attack maps with the castling
    # squares often empty, the rooks often home and the flags often clear
    rng = np.random.default_rng(13)
    r64 = lambda: int(rng.integers(0, 2**63, dtype=np.int64)) | (int(rng.integers(0, 2, dtype=np.int64)) << 63)
    hits = 0
    for _ in range(20000):
        stm = int(rng.integers(0, 2))
        k = int(rng.choice([4, 60, int(rng.integers(0, 64))]))
        rank = 0xFF << (56 if stm else 0)
        occ = r64() & r64() & ~(rank if rng.integers(0, 2) else 0)
        A = r64() & r64() & r64() & ~(rank if rng.integers(0, 2) else 0)
        rooks = (0x81 << (56 if stm else 0)) if rng.integers(0, 2) else r64()
        own = rooks | (r64() & occ)
        b3, b1, b2 = rooks | (r64() & occ), rooks | (r64() & occ), r64() & occ & ~rooks
        cr = r64() & r64() & r64()
        if rng.integers(0, 2):
            cr |= 1 << 59  # a flagged square, so the precondition holds
        args = (0, b1, b2, b3, cr, stm, k, occ | rooks, own, A)
        want = H.hc_castle_per_square(*args)
        hits += want != 0
        assert H.hc_castle(*args) == want, args
    assert hits > 100


def test_castle_every_king_square_modulo_64(H):
    # the castling rule on the squares k-4 .. k+3 taken modulo 64 (shift
    # counts are modulo 64 in the reference, src/zetaperft.cl:1249-1297), for
    # every king square: rooks on k-4 and k+3, the squares between empty,
    # nothing attacked -> targets k-2 and k+2; each broken condition removes
    # exactly its own side
    sq = lambda k, d: 1 << ((k + d) % 64)
    cr = 1 << 59  # d8 flagged: the precondition holds, both sides' squares unflagged
    for stm in (0, 1):
        for k in range(64):
            rooks = sq(k, -4) | sq(k, 3)
            base = dict(occ=rooks | sq(k, 0), rooks=rooks, A=0)

            def targets(occ, rooks, A):
                args = (0, rooks, 0, rooks, cr, stm, k, occ, rooks | sq(k, 0), A)
                got = H.hc_castle(*args)
                assert got == H.hc_castle_per_square(*args), (stm, k)
                return got

            both, q, kk = sq(k, -2) | sq(k, 2), sq(k, -2), sq(k, 2)
            assert targets(**base) == both, (stm, k)
            for d in (-3, -2, -1):  # queen side must be empty
                assert targets(base["occ"] | sq(k, d), rooks, 0) == kk, (stm, k, d)
            for d in (1, 2):  # king side must be empty
                assert targets(base["occ"] | sq(k, d), rooks, 0) == q, (stm, k, d)
            assert targets(base["occ"], rooks, sq(k, 0)) == 0, (stm, k)  # in check
            for d, want in ((-2, kk), (-1, kk), (1, q), (2, q), (-3, both), (-4, both), (3, both)):
                assert targets(base["occ"], rooks, sq(k, d)) == want, (stm, k, d)
            assert targets(base["occ"] ^ sq(k, -4), rooks ^ sq(k, -4), 0) == kk, (stm, k)
            assert targets(base["occ"] ^ sq(k, 3), rooks ^ sq(k, 3), 0) == q, (stm, k)


def test_single_square_sliders_equal_fills(H):
    # hyperbola attacks of one slider = the occluded fills, random occupancies
    rng = np.random.default_rng(7)
    for _ in range(3000):
        occ = int(rng.integers(0, 2**63, dtype=np.int64)) & int(rng.integers(0, 2**63, dtype=np.int64))
        sq = int(rng.integers(0, 64))
        occ |= 1 << sq
        for rook in (0, 1):
            assert H.hc_slider(sq, occ, rook) == H.hc_slider_ks(sq, occ, rook), (sq, hex(occ), rook)


def test_lane_analysis_matches_oracle(H):
    # the search's lane-per-square generator: same legal moves, same order
    pos = np.concatenate([R.gen_playouts(2, 1500, 8, 40), R.gen_playouts(3, 800, 20, 60),
                          R.gen_playouts(9, 500, 60, 120)])
    fens = ["R6R/3Q4/1Q4Q1/4Q3/2Q4Q/Q4Q2/pp1Q4/kBNN1KB1 w - - 0 1",
            "r3k2r/8/8/8/8/8/8/R3K2R w KQkq - 0 1", "8/8/8/8/k2Pp2Q/8/8/3K4 b - - 0 1",
            "8/8/8/2k5/3Pp3/8/8/4K2R b K - 0 1", "7k/5Q2/6K1/8/8/8/8/8 b - - 0 1",
            "4k3/8/8/8/8/8/4r3/R3K2R w KQ - 0 1", "4k3/4r3/8/8/8/8/3B4/R3K2R w KQ - 0 1"]
    pos = np.concatenate([pos] + [R.setboard(f) for f in fens])
    buf = np.zeros(256, dtype=np.uint32)
    ref = np.zeros(256, dtype=np.uint32)
    for i in range(len(pos)):
        b = np.ascontiguousarray(pos[i]["b"])
        stm = int(pos[i]["stm"])
        for qs in (0, 1):
            n = H.hc_enumerate_lane(b.ctypes.data, stm, qs, buf.ctypes.data)
            m = H.hc_enumerate(b.ctypes.data, stm, qs, ref.ctypes.data)
            assert n == m and list(buf[:n]) == list(ref[:m]), (i, qs)


def test_device_eval_matches_oracle(H):
    # the search's per-square evaluation (one table lookup, pawn-file terms by
    # popcount) = the oracle's evaluation of the reference's terms
    pos = np.concatenate([R.gen_playouts(2, 1500, 8, 40), R.gen_playouts(3, 1500, 20, 60),
                          R.gen_playouts(9, 500, 60, 120),
                          R.setboard("4k3/8/3p4/3P4/3P4/3P4/8/4K3 w - - 0 1"),
                          R.setboard("4k3/3p4/3p4/3p4/3P4/8/8/4K3 b - - 0 1")])
    for i in range(len(pos)):
        b = np.ascontiguousarray(pos[i]["b"])
        assert H.hc_eval(b.ctypes.data) == R.evaluate(pos[i:i + 1]), i


def test_device_perft_algorithm_reproduces_golden_workload(H):
    # the whole perft(4) workload of 4096 playout positions (seed 2, 8-40
    # plies) through the device algorithm compiled for the host (grouped
    # expansion, bulk counting at the last level): every per-position count
    # equals the committed golden counts, 4,182,308,444 leaves in all
    gold = np.load(os.path.join(HERE, "golden", "c2_perft4.npz"), allow_pickle=False)
    pos = R.gen_playouts(2, 4096, 8, 40)
    assert (pos["b"][:, 5] == gold["hash"]).all()  # the same workload
    out = np.zeros(len(pos), dtype=np.uint64)
    H.hc_perft_batch(pos.ctypes.data, len(pos), 4, out.ctypes.data, min(8, os.cpu_count() or 1))
    assert (out == gold["perft4"]).all()
    assert int(out.sum()) == 4182308444


def _fen_with(piece, sq, wk, bk):
    board = [["" for _ in range(8)] for _ in range(8)]
    for ch, q in ((piece, sq), ("K", wk), ("k", bk)):
        board[q // 8][q % 8] = ch
    rows = []
    for r in range(7, -1, -1):
        row, run = "", 0
        for f in range(8):
            if board[r][f]:
                row += (str(run) if run else "") + board[r][f]
                run = 0
            else:
                run += 1
        rows.append(row + (str(run) if run else ""))
    return "/".join(rows) + " w - - 0 1"


def test_device_eval_every_table_entry(H):
    # one position per (piece, colour, square) beside the two kings: every
    # entry of the device's evaluation tables is compared with the oracle's
    # separately typed tables
    n = 0
    for piece in "PNBRQpnbrq":
        for sq in range(64):
            if piece in "Pp" and sq // 8 in (0, 7):
                continue
            wk, bk = (0, 63) if sq not in (0, 63) else (7, 56)
            p = R.setboard(_fen_with(piece, sq, wk, bk))
            b = np.ascontiguousarray(p[0]["b"])
            assert H.hc_eval(b.ctypes.data) == R.evaluate(p), (piece, sq)
            n += 1
    for sq in range(64):  # the king tables, the other king kept away
        for wk_side in (True, False):
            other = 63 if sq < 32 else 0
            fen = _fen_with("N", 27 if sq not in (27,) and other != 27 else 36, sq if wk_side else other,
                            other if wk_side else sq)
            p = R.setboard(fen)
            b = np.ascontiguousarray(p[0]["b"])
            assert H.hc_eval(b.ctypes.data) == R.evaluate(p), (sq, wk_side)
            n += 1
    assert n == 2 * (4 * 64 + 48) + 128


def test_device_hash_matches_oracle(H):
    # the device's from-scratch Zobrist hash (its own copy of the keys) = the
    # oracle's computehash, on playout positions and on every position one
    # ply further (double pushes leave an e.p. flag, king and rook moves
    # clear castle rights)
    base = np.concatenate([R.gen_playouts(2, 300, 8, 40), R.gen_playouts(3, 300, 20, 60),
                           R.setboard(R.STARTPOS),
                           R.setboard("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1")])
    n = 0
    for i in range(len(base)):
        p = base[i:i + 1]
        kids = [p] + [R.domove(p, R.nth_legal(p, k)) for k in range(R.count_legal(p))]
        for q in kids:
            b = np.ascontiguousarray(q[0]["b"])
            want = R.lib().rc_computehash(b.ctypes.data, int(q["stm"][0]))
            assert H.hc_hash(b.ctypes.data, int(q["stm"][0])) == want, i
            n += 1
    assert n > 10000

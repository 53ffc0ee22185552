"""CPU tests of the front end's board I/O (zeta_amd/host/zeta_engine.cpp,
setboard / createfen after src/io.c:227-532): FEN round trips, and the FEN
after coordinate moves played on the board, against the standard FEN of the
same game (castle rights, e.p. target, half-move clock, move number); and of
its clock (level / st / time and the update after the engine's move,
src/xboard.c:355-371, 384-455) on worked examples."""
import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SO = os.path.join(HERE, "hostcore", "libfenhost.so")
SRC = os.path.join(HERE, "hostcore", "fen_host.cpp")
DEPS = [SRC, os.path.join(ROOT, "zeta_amd", "host", "zeta_engine.cpp"),
        os.path.join(ROOT, "zeta_amd", "host", "clrun_zeta.cpp"),
        os.path.join(ROOT, "zeta_amd", "csrc", "chess.h")]
LIBDIR = os.path.join(ROOT, "zeta_amd")


@pytest.fixture(scope="module")
def F():
    if not os.path.exists(os.path.join(LIBDIR, "libzetahip.so")):
        pytest.skip("libzetahip.so not built")
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(p) for p in DEPS):
        subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-Wno-unknown-pragmas",
                               "-I" + os.path.join(ROOT, "include"), "-o", SO, SRC,
                               os.path.join(ROOT, "zeta_amd", "host", "clrun_zeta.cpp"),
                               "-L" + LIBDIR, "-lzetahip", "-Wl,-rpath," + LIBDIR])
    L = ctypes.CDLL(SO)
    L.fh_play.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.fh_play.restype = ctypes.c_int
    L.fh_print.argtypes = [ctypes.c_char_p]
    L.fh_print.restype = ctypes.c_int
    L.fh_suite.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.fh_suite.restype = ctypes.c_int
    _clock_types(L)

    def play(fen, moves=""):
        out = ctypes.create_string_buffer(256)
        rc = L.fh_play(fen.encode(), moves.encode(), out, 256)
        assert rc == 0, (fen, moves, rc)
        return out.value.decode()
    play.lib = L
    return play


def _clock_types(L):
    L.fh_clock.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_double,
                           ctypes.POINTER(ctypes.c_double)]
    L.fh_clock.restype = ctypes.c_int


START = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1"


@pytest.mark.parametrize("fen", [
    START,
    "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1",
    "8/2p5/3p4/KP5r/1R3p1k/8/4P1P1/8 w - - 0 1",
    "r3k2r/Pppp1ppp/1b3nbN/nP6/BBP1P3/q4N2/Pp1P2PP/R2Q1RK1 w kq - 0 1",
    "rnbq1k1r/pp1Pbppp/2p5/8/2B5/8/PPP1NnPP/RNBQK2R w KQ - 1 8",
    "r4rk1/1pp1qppp/p1np1n2/2b1p1B1/2B1P1b1/P1NP1N2/1PP1QPPP/R4RK1 w - - 0 10",
    "r3k2r/8/8/8/8/8/8/R3K2R b Qk - 5 40",
])
def test_roundtrip(F, fen):
    assert F(fen) == fen


@pytest.mark.parametrize("fen,moves,want", [
    (START, "e2e4", "rnbqkbnr/pppppppp/8/8/4P3/8/PPPP1PPP/RNBQKBNR b KQkq e3 0 1"),
    (START, "e2e4 d7d5", "rnbqkbnr/ppp1pppp/8/3p4/4P3/8/PPPP1PPP/RNBQKBNR w KQkq d6 0 2"),
    (START, "g1f3", "rnbqkbnr/pppppppp/8/8/8/5N2/PPPPPPPP/RNBQKB1R b KQkq - 1 1"),
    (START, "e2e4 e7e5 e1e2", "rnbqkbnr/pppp1ppp/8/4p3/4P3/8/PPPPKPPP/RNBQ1BNR b kq - 1 2"),
    (START, "g1f3 g8f6 h1g1", "rnbqkb1r/pppppppp/5n2/8/8/5N2/PPPPPPPP/RNBQKBR1 b Qkq - 3 2"),
    (START, "e2e4 d7d5 e4d5", "rnbqkbnr/ppp1pppp/8/3P4/8/8/PPPP1PPP/RNBQKBNR b KQkq - 0 2"),
    # e.p. capture: the captured pawn leaves d5
    (START, "e2e4 a7a6 e4e5 d7d5 e5d6", "rnbqkbnr/1pp1pppp/p2P4/8/8/8/PPPP1PPP/RNBQKBNR b KQkq - 0 3"),
    # castling moves king and rook and ends that side's rights; unlike standard
    # FEN the reference resets the half-move clock on a castle
    # (src/bitboard.c:664,681; src/zeta.cl:2368-2373)
    ("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 3 1", "e1g1",
     "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R4RK1 b kq - 0 1"),
    ("r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R b KQkq - 3 1", "e8c8",
     "2kr3r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQ - 0 2"),
    # a rook captured on its home square takes that side's right with it
    ("r3k2r/8/8/8/8/8/8/R3K2R w KQkq - 0 1", "a1a8", "R3k2r/8/8/8/8/8/8/4K2R b Kk - 0 1"),
    # promotion without a letter is to a queen (the engine's own promotions
    # always are); a letter from the GUI picks the piece (src/io.c:104-113)
    ("8/P7/8/8/8/8/8/k6K w - - 0 1", "a7a8", "Q7/8/8/8/8/8/8/k6K b - - 0 1"),
    ("8/P7/8/8/8/8/8/k6K w - - 0 1", "a7a8q", "Q7/8/8/8/8/8/8/k6K b - - 0 1"),
    ("8/P7/8/8/8/8/8/k6K w - - 0 1", "a7a8n", "N7/8/8/8/8/8/8/k6K b - - 0 1"),
    ("8/P7/8/8/8/8/8/k6K w - - 0 1", "a7a8B", "B7/8/8/8/8/8/8/k6K b - - 0 1"),
    ("8/P7/8/8/8/8/8/k6K w - - 0 1", "a7a8r", "R7/8/8/8/8/8/8/k6K b - - 0 1"),
    ("k7/8/8/8/8/8/1p6/R6K b - - 0 1", "b2a1n", "k7/8/8/8/8/8/8/n6K w - - 0 2"),
])
def test_fen_after_moves(F, fen, moves, want):
    assert F(fen, moves) == want


def test_fen_ep_square_not_kept_from_input(F):
    # the reference's setboard marks a FEN's e.p. square on rank 3/6, where
    # neither its move generator nor its createfen looks (src/io.c:487-496,
    # 292-310): the square reads back as "-" and the e.p. capture is not legal
    fen = "rnbqkbnr/ppp1pppp/8/3pP3/8/8/PPPP1PPP/RNBQKBNR w KQkq d6 0 3"
    assert F(fen) == fen.replace(" d6 ", " - ")
    out = ctypes.create_string_buffer(256)
    L = ctypes.CDLL(SO)
    assert L.fh_play(fen.encode(), b"e5d6", out, 256) == -2


def test_board_print(F, capfd):
    # the board print of src/io.c:164-225: files header, ranks 8..1 with "-"
    # for an empty square, the header again, then the position's FEN
    fen = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1"
    capfd.readouterr()
    assert F.lib.fh_print(fen.encode()) == 0
    lines = capfd.readouterr().out.splitlines()
    assert lines == ["###ABCDEFGH###",
                     "#8 r---k--r", "#7 p-ppqpb-", "#6 bn--pnp-", "#5 ---PN---",
                     "#4 -p--P---", "#3 --N--Q-p", "#2 PPPBBPPP", "#1 R---K--R",
                     "###ABCDEFGH###", "#fen: " + fen]


def test_smp_benchmark_position_sets(F):
    # benchkaufmann's 25 and benchhyatt24's 24 positions (src/xboard.c:986-1013,
    # 1214-1240) all parse, read back to the same board, side to move and
    # castle rights, and have legal moves
    L = F.lib
    for suite, n, first in ((0, 25, "1rbq1rk1/p1b1nppp/1p2p3/8/1B1pN3/P2B4/1P3PPP/2RQ1R1K w - -"),
                            (1, 24, "r2qkbnr/ppp2p1p/2n5/3P4/2BP1pb1/2N2p2/PPPQ2PP/R1B2RK1 b kq -")):
        out = ctypes.create_string_buffer(256)
        nm = ctypes.c_int(0)
        fens = []
        for i in range(n):
            assert L.fh_suite(suite, i, out, 256, ctypes.byref(nm)) == 0, (suite, i)
            assert nm.value > 0, (suite, i)
            fens.append(out.value.decode())
        assert L.fh_suite(suite, n, out, 256, ctypes.byref(nm)) == -2
        assert fens[0].rsplit(" ", 2)[0] == first
        assert len(set(fens)) == n


CLOCK_KEYS = ("base", "left", "inc", "max_time", "moves_left", "max_moves", "mode", "max_nodes")


def clock(L, line="", ply=0, nps=1_000_000, elapsed=-1.0):
    out = (ctypes.c_double * 8)()
    rc = L.fh_clock(line.encode(), ply, nps, elapsed, out)
    assert rc == 0, (line, rc)
    return dict(zip(CLOCK_KEYS, out))


def test_clock_defaults_give_every_move_five_seconds(F, tmp_path):
    # a fresh engine thinks 5 s per move (src/zeta.c:78-85): fixed time per
    # move, and after a move the budget is again 5 s, 5 M nodes at 1 M nodes/s
    # (a fresh copy of the library, so no other test's commands are in it)
    fresh = tmp_path / "libfenhost_fresh.so"
    shutil.copy(SO, fresh)
    L = ctypes.CDLL(str(fresh))
    _clock_types(L)
    c = clock(L)
    assert (c["base"], c["left"], c["inc"], c["max_time"], c["moves_left"], c["mode"]) == (5000, 5000, 0, 5000, 1, 0)
    c = clock(L, ply=1, elapsed=1234.0)
    assert (c["left"], c["max_time"], c["max_nodes"]) == (5000, 5000, 5_000_000)


def test_clock_conventional_sessions(F):
    L = F.lib
    # level 40 5 0 at ply 0: 40 moves in 5 minutes, one move kept spare
    # (moves left = 40 - (ply+1)/2 % 40 + 1 = 41)
    c = clock(L, "level 40 5 0")
    assert (c["base"], c["left"], c["inc"], c["moves_left"], c["max_moves"], c["mode"]) == (300000, 300000, 0, 41, 40, 1)
    assert c["max_time"] == pytest.approx(300000 / 41) and c["max_nodes"] == 7317073
    # the base as minutes:seconds
    c = clock(L, "level 40 2:30 0")
    assert c["base"] == 150000 and c["max_time"] == pytest.approx(150000 / 41)
    # the GUI's clock at ply 10: 36 moves left in the session
    c = clock(L, "time 12000", ply=10)
    assert (c["left"], c["moves_left"]) == (120000, 36) and c["max_time"] == pytest.approx(120000 / 36)
    # two moves per minute: after the engine's first move (ply 1) two moves
    # are left, so the next session's minute is added to the clock
    clock(L, "level 2 1 0")
    c = clock(L, ply=1, elapsed=500.0)
    assert c["moves_left"] == 2 and c["max_time"] == pytest.approx(59500 / 2) and c["left"] == 119500


def test_clock_increment_and_sudden_death(F):
    L = F.lib
    # level 0 2 12: 2 minutes plus 12 s a move, 40 moves assumed
    c = clock(L, "level 0 2 12")
    assert (c["base"], c["inc"], c["moves_left"], c["max_moves"], c["mode"]) == (120000, 12000, 40, 0, 2)
    assert c["max_time"] == 120000 / 40 + 12000 and c["max_nodes"] == 15_000_000
    c = clock(L, ply=1, elapsed=2000.0)
    assert c["left"] == 130000 and c["max_time"] == 130000 / 40 + 12000
    # level 0 5 0 is the same clock without increment: the time runs down,
    # it is not reset every move
    c = clock(L, "level 0 5 0")
    assert (c["mode"], c["moves_left"], c["max_time"]) == (2, 40, 7500)
    c = clock(L, ply=1, elapsed=1000.0)
    assert c["left"] == 299000 and c["max_time"] == 299000 / 40
    # out of time: no node budget (and no conversion of a negative count)
    c = clock(L, "time 0")
    assert (c["left"], c["max_time"], c["max_nodes"]) == (0, 0, 0)
    c = clock(L, ply=3, elapsed=500.0)
    assert c["left"] == -500 and c["max_nodes"] == 0


def test_clock_fixed_time_per_move(F):
    L = F.lib
    # st clears an increment and the session length of an earlier level
    clock(L, "level 0 2 12")
    c = clock(L, "st 10")
    assert (c["base"], c["left"], c["inc"], c["moves_left"], c["max_moves"], c["mode"]) == (10000, 10000, 0, 1, 1, 0)
    assert c["max_time"] == 10000 and c["max_nodes"] == 10_000_000
    c = clock(L, ply=5, elapsed=9000.0)
    assert (c["left"], c["max_time"]) == (10000, 10000)


def test_clock_ignores_malformed_levels(F):
    L = F.lib
    before = clock(L, "st 7")
    for line in ("level 40", "level x 5 0", "level -1 5 0"):
        assert clock(L, line) == before, line
    assert L.fh_clock(b"otim 100", 0, 1, -1.0, (ctypes.c_double * 8)()) == -1

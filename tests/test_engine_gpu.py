"""GPU tests of the CECP front end (zeta_amd/bin/zeta) and of the
reference's six-function device boundary behind it (include/zeta_clrun.h):
the reference's own command flows -- selftest, perft at the `sd` depth, an
xboard game -- run unchanged over libzetahip.so."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = os.path.join(ROOT, "zeta_amd", "bin", "zeta")
CONFIG = os.path.join(ROOT, "zeta_amd", "host", "config.txt")


def run(script, args=(), timeout=100):
    if not os.path.exists(ENGINE):
        pytest.fail("zeta_amd/bin/zeta not built (make -C zeta_amd/csrc)")
    p = subprocess.run([ENGINE, "--config", CONFIG, *args], input=script, capture_output=True, text=True,
                       timeout=timeout)
    print(p.stdout[-3000:], p.stderr[-2000:], flush=True)
    return p


def test_selftest_passes_all_reference_kats():
    p = run("", args=("--selftest",))
    assert p.returncode == 0
    assert p.stdout.count("#> OK, Nodecount Correct") == 14
    assert "14 of 14 tests passed" in p.stdout


def test_perft_command_startpos_depth5():
    p = run("sd 5\nperft\nquit\n")
    assert p.returncode == 0
    assert "nodecount:4865609," in p.stdout


def test_perft_after_setboard_and_usermove():
    # Kiwipete d2 through the FEN path; then a user move and perft 1 of the reply
    p = run("setboard r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1\nsd 2\nperft\n"
            "force\nusermove e1g1\nsd 1\nperft\nquit\n")
    assert p.returncode == 0
    counts = [line for line in p.stdout.splitlines() if line.startswith("nodecount:")]
    assert counts[0].startswith("nodecount:2039,")
    assert len(counts) == 2


def test_xboard_game_plays_mate_in_one():
    p = run("xboard\nprotover 2\naccepted usermove\nnew\nsetboard 6k1/5ppp/8/8/8/8/5PPP/R5K1 w - - 0 1\n"
            "sd 3\ngo\nquit\n")
    assert p.returncode == 0
    assert "feature done=1" in p.stdout
    assert "move a1a8" in p.stdout


def test_think_and_reply_in_a_game():
    p = run("xboard\nprotover 2\nnew\nsd 4\npost\nusermove e2e4\nusermove d2d4\nquit\n")
    assert p.returncode == 0
    moves = [line for line in p.stdout.splitlines() if line.startswith("move ")]
    assert len(moves) == 2


def test_benchsmp_time_to_depth_per_worker_count():
    p = run("sd 5\nbenchsmp\nquit\n")
    assert p.returncode == 0
    assert "# overall results" in p.stdout
    rows = [line for line in p.stdout.splitlines() if line.startswith("### ") and "workers" not in line]
    assert [int(r.split()[1]) for r in rows] == [1, 2, 4, 8, 16]
    scores = {line.split("score: ")[1].split()[0] for line in p.stdout.splitlines() if line.startswith("depth: 5,")}
    assert len(scores) == 1  # the worker count never changes the root score

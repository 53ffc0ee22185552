"""GPU tests of the CECP front end (zeta_amd/bin/zeta) and of the
reference's six-function device boundary behind it (include/zeta_clrun.h):
the reference's own command flows -- selftest, perft at the `sd` depth, an
xboard game -- run unchanged over libzetahip.so."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENGINE = os.path.join(ROOT, "zeta_amd", "bin", "zeta")
CONFIG = os.path.join(ROOT, "zeta_amd", "host", "config.txt")


def run(script, args=(), timeout=100, config=CONFIG):
    if not os.path.exists(ENGINE):
        pytest.fail("zeta_amd/bin/zeta not built (make -C zeta_amd/csrc)")
    p = subprocess.run([ENGINE, "--config", config, *args], input=script, capture_output=True, text=True,
                       timeout=timeout)
    print(p.stdout[-3000:], p.stderr[-2000:], flush=True)
    return p


def test_selftest_passes_all_reference_kats():
    p = run("", args=("--selftest",))
    assert p.returncode == 0
    assert p.stdout.count("#> OK, Nodecount Correct") == 14
    # the reference's summary banner (src/test.c:144-146)
    assert "###############################\n### passed 14 from 14 tests ###\n###############################" \
        in p.stdout


def test_guessconfig_writes_and_reads_back_the_launch_knobs(tmp_path):
    # the autotuner (--guessconfig, src/clconfig.c:878-960): workers per root,
    # ABDADA split depth and wavefronts per SIMD measured on the GPU, written
    # to config.txt with the reference's keys plus this build's two, and read
    # back; then the engine starts from that file and searches with it
    cfg = str(tmp_path / "config.txt")
    p = run("", args=("--guessconfig",), config=cfg, timeout=200)
    assert p.returncode == 0
    vals = {k: int(v) for k, v in re.findall(r"^(\w+): (-?\d+);", open(cfg).read(), re.M)}
    assert vals["threadsX"] in (1, 4, 16, 64, 128) and vals["threadsY"] == 1
    assert vals["split_depth"] in (1, 2, 3) and vals["waves_per_simd"] in (1, 2)
    assert vals["nodes_per_second"] > 0 and vals["tt1_memory"] == 1024
    assert ("#> config read back: threadsX %d, split_depth %d, waves_per_simd %d"
            % (vals["threadsX"], vals["split_depth"], vals["waves_per_simd"])) in p.stdout
    q = run("new\nsd 4\ngo\nquit\n", config=cfg)
    assert q.returncode == 0
    assert re.search(r"^move [a-h][1-8][a-h][1-8]", q.stdout, re.M)


def test_perft_command_startpos_depth5():
    p = run("sd 5\nperft\nquit\n")
    assert p.returncode == 0
    assert "nodecount:4865609," in p.stdout


def config_file(tmp_path, threads=16, nps=100000000, tt_mb=1024):
    path = tmp_path / "config.txt"
    path.write_text("threadsX: %d;\nthreadsY: 1;\nnodes_per_second: %d;\ntt1_memory: %d;\ntt2_memory: 0;\n"
                    "opencl_platform_id: 0;\nopencl_device_id: 0;\n" % (threads, nps, tt_mb))
    return str(path)


def counts_of(p):
    return [int(line.split(":")[1].split(",")[0]) for line in p.stdout.splitlines() if line.startswith("nodecount:")]


KIWIPETE = "r3k2r/p1ppqpb1/bn2pnp1/3PN3/1p2P3/2N2Q1p/PPPBBPPP/R3K2R w KQkq - 0 1"


def test_perft_after_setboard_and_usermove(oracle):
    # Kiwipete d2 through the FEN path; then a user move (castling) and
    # perft 2 of the position after it, against the oracle's host domove
    p = run("setboard %s\nsd 2\nperft\nforce\nusermove e1g1\nperft\nquit\n" % KIWIPETE)
    assert p.returncode == 0
    after = oracle.domove(oracle.setboard(KIWIPETE), oracle.move_from_can(oracle.setboard(KIWIPETE), "e1g1"))
    assert counts_of(p) == [2039, oracle.perft(after, 2)]


def test_undo_and_remove_restore_the_position(oracle):
    """undo takes back one ply and remove two (src/xboard.c:641-664): the
    perft counts after them are those of the positions before the moves."""
    start = oracle.setboard(oracle.STARTPOS)
    e4 = oracle.domove(start, oracle.move_from_can(start, "e2e4"))
    p = run("new\nforce\nusermove e2e4\nusermove e7e5\nsd 3\nperft\nundo\nperft\nusermove c7c5\n"
            "remove\nperft\nquit\n")
    assert p.returncode == 0
    e4e5 = oracle.domove(e4, oracle.move_from_can(e4, "e7e5"))
    assert counts_of(p) == [oracle.perft(e4e5, 3), oracle.perft(e4, 3), oracle.perft(start, 3)]


def test_st_game_stops_at_the_node_budget(tmp_path):
    """Time control by nodes (src/search.c:77-198): with st 1 and a small
    nodes_per_second the device search of every iteration gets the budget
    MaxNodes / workers; the iteration that reaches it reports COUNTERS[1] at
    the budget (src/search.c:133), the engine keeps the last complete
    iteration's move and replies within its time."""
    cfg = config_file(tmp_path, threads=16, nps=20000)
    p = run("xboard\nprotover 2\nnew\nst 1\npost\nusermove e2e4\nusermove d2d4\nquit\n", config=cfg, timeout=120)
    assert p.returncode == 0
    moves = [line for line in p.stdout.splitlines() if line.startswith("move ")]
    assert len(moves) == 2
    depths = [int(line.split()[0]) for line in p.stdout.splitlines() if line[:1].isdigit() and len(line.split()) >= 5]
    assert depths and max(depths) < 20


def test_level_and_time_game(tmp_path):
    cfg = config_file(tmp_path, threads=8, nps=200000)
    p = run("xboard\nprotover 2\nnew\nlevel 40 1 0\ntime 6000\notim 6000\npost\nusermove e2e4\n"
            "time 5900\nusermove g1f3\nquit\n", config=cfg, timeout=120)
    assert p.returncode == 0
    assert len([line for line in p.stdout.splitlines() if line.startswith("move ")]) == 2


@pytest.mark.parametrize("cmd,npos", [("benchkaufmann", 25), ("benchhyatt24", 24)])
def test_smp_bench_suites(tmp_path, cmd, npos):
    """benchkaufmann / benchhyatt24 (src/xboard.c:975-1428): every position
    searched with 1, 2, 4 workers; the root score never depends on it."""
    cfg = config_file(tmp_path, threads=4, tt_mb=64)
    p = run("sd 3\n%s\nquit\n" % cmd, config=cfg, timeout=300)
    assert p.returncode == 0
    assert "# overall results" in p.stdout
    assert p.stdout.count("### setting up board") == npos
    # per position: the last thinking line of each worker count ("depth
    # score time nodes pv"; a found mate ends the iterations early), with
    # one depth and one score whatever the worker count
    blocks = p.stdout.split("### setting up board")[1:]
    for b in blocks:
        runs = b.split("### computing ")[1:]
        assert len(runs) == 3
        last = []
        for r in runs:
            rows = [line.split() for line in r.splitlines()]
            rows = [x for x in rows if len(x) >= 5 and x[0].isdigit() and x[1].lstrip("-").isdigit()]
            assert rows, r[:400]
            last.append((rows[-1][0], rows[-1][1]))
        assert len(set(last)) == 1, last


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
    # benchsmp (src/xboard.c:822-974): the current position searched to depth
    # sd with 1, 2, 4, ... workers, then the table of nps and time-to-depth
    # speedups; the last thinking line of every worker count has the same
    # depth and score
    p = run("sd 5\nbenchsmp\nquit\n")
    assert p.returncode == 0
    assert "### workers\t#nps\t\t#nps speedup\t#time in s\t#ttd speedup \t#relative ttd speedup ###" in p.stdout
    table = [line.split() for line in p.stdout.splitlines() if re.match(r"### \d+\t", line)]
    assert [int(r[1]) for r in table] == [1, 2, 4, 8, 16]
    assert float(table[0][3]) == 1.0 and float(table[0][5]) == 1.0
    runs = p.stdout.split("### computing benchsmp depth 5: ###")[1:]
    assert len(runs) == 5
    last = []
    for r in runs:
        rows = [line.split() for line in r.splitlines()]
        rows = [x for x in rows if len(x) >= 5 and x[0].isdigit() and x[1].lstrip("-").isdigit()]
        assert rows, r[:400]
        last.append((rows[-1][0], rows[-1][1]))
    assert len(set(last)) == 1 and last[0][0] == "5", last

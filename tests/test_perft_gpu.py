"""GPU parity tests of batched perft (libzetahip.so through the C-ABI) against
the oracle and the committed golden fixtures."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kats.json")))
pytestmark = pytest.mark.gpu


def _pos(oracle, fens):
    return np.concatenate([oracle.setboard(f) for f in fens])


def test_reference_kats(zeta, oracle):
    for case in KATS["reference_kats"]:
        p = oracle.setboard(case["fen"])
        got = [int(zeta.perft(p, d + 1)[0]) for d in range(len(case["counts"]))]
        print("KAT", case["fen"][:30], got, flush=True)
        assert got == case["counts"]


def test_textbook_and_quirks(zeta, oracle):
    for case in KATS["textbook"]:
        assert int(zeta.perft(oracle.setboard(case["fen"]), case["depth"])[0]) == case["count"]
    for case in KATS["quirks"]:
        assert int(zeta.perft(oracle.setboard(case["fen"]), case["depth"])[0]) == case["count"]


def test_startpos_perft5_config1(zeta, oracle):
    # config 1: startpos perft 5 = 4,865,609 leaves
    assert int(zeta.perft(oracle.setboard(oracle.STARTPOS), 5)[0]) == 4865609


def test_playouts_match_oracle(zeta, oracle):
    g = zeta.playouts(2, 4096, 8, 40)
    o = oracle.gen_playouts(2, 4096, 8, 40)
    assert (g["b"] == o["b"]).all() and (g["stm"] == o["stm"]).all()
    g3 = zeta.playouts(3, 1024, 20, 60)
    o3 = oracle.gen_playouts(3, 1024, 20, 60)
    assert (g3["b"] == o3["b"]).all()


def test_config2_perft4_golden(zeta, oracle):
    gold = np.load(os.path.join(HERE, "golden", "c2_perft4.npz"))
    pos = oracle.gen_playouts(2, 4096, 8, 40)
    assert (pos["b"][:, 5] == gold["hash"]).all()
    got3 = zeta.perft(pos, 3)
    assert (got3 == gold["perft3"]).all()
    got = zeta.perft(pos, 4)
    print("config2 total", int(got.sum()), flush=True)
    assert (got == gold["perft4"]).all()


def test_nobulk_equals_bulk(zeta, oracle):
    pos = oracle.gen_playouts(2, 256, 8, 40)
    d_pos = zeta.DeviceBuffer(pos.nbytes)
    d_cnt = zeta.DeviceBuffer(8 * len(pos))
    d_pos.upload(pos)
    zeta.perft_device(d_pos.ptr, len(pos), 3, d_cnt.ptr, nobulk=True)
    got = d_cnt.download(np.zeros(len(pos), dtype=np.uint64))
    d_pos.free()
    d_cnt.free()
    assert (got == oracle.perft_batch(pos, 3)).all()


def test_edge_cases(zeta, oracle):
    assert len(zeta.perft(np.zeros(0, dtype=zeta.POS_DTYPE), 3)) == 0
    p = _pos(oracle, [oracle.STARTPOS, "7k/5Q2/6K1/8/8/8/8/8 b - - 0 1",
                      "R5k1/5ppp/8/8/8/8/5PPP/6K1 b - - 0 1",
                      "R6R/3Q4/1Q4Q1/4Q3/2Q4Q/Q4Q2/pp1Q4/kBNN1KB1 w - - 0 1"])
    assert list(zeta.perft(p, 0)) == [1, 1, 1, 1]
    assert list(zeta.perft(p, 1)) == [20, 0, 0, 218]
    for d in (2, 3):
        assert list(zeta.perft(p, d)) == [oracle.perft(p[i:i + 1], d) for i in range(len(p))]


def test_chunked_levels_small_budget(zeta, oracle):
    # a frontier budget far below a level's size splits levels into chunks
    # (cuts computed on the device); counts must not change
    pos = oracle.gen_playouts(2, 64, 8, 40)
    want3 = oracle.perft_batch(pos, 3)
    zeta.shutdown()
    try:
        zeta.init(frontier_budget=4096)
        got = zeta.perft(pos, 3)
        print("chunked perft(3) total", int(got.sum()), flush=True)
        assert (got == want3).all()
        d_pos = zeta.DeviceBuffer(pos.nbytes)
        d_cnt = zeta.DeviceBuffer(8 * len(pos))
        d_pos.upload(pos)
        zeta.perft_device(d_pos.ptr, len(pos), 2, d_cnt.ptr, nobulk=True)
        got2 = d_cnt.download(np.zeros(len(pos), dtype=np.uint64))
        d_pos.free()
        d_cnt.free()
        assert (got2 == oracle.perft_batch(pos, 2)).all()
        zeta.shutdown()
        zeta.init(frontier_budget=100)
        with pytest.raises(zeta.ZetaError):
            zeta.perft(pos, 3)
    finally:
        zeta.shutdown()
        zeta.init()

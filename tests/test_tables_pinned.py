"""The evaluation constants, Zobrist keys and castle masks of both copies --
the oracle's (oracle/refcpu.c) and the device code's (zeta_amd/csrc/evaltab.h
and chess.h, through their host build) -- equal the numbers extracted from
the reference (tests/golden/ref_tables.json, written by
tests/golden/extract_tables.py from src/zeta.cl:258-359, src/zobrist.h:28-35
and src/types.h:114-124), entry by entry.  A typo in either copy fails here."""
import ctypes
import json
import os
import subprocess

import pytest

from oracle import refcpu as R

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "hostcore", "libchesshost.so")
SRC = os.path.join(HERE, "hostcore", "chess_host.cpp")
CSRC = os.path.join(os.path.dirname(HERE), "zeta_amd", "csrc")
FIXTURE = os.path.join(HERE, "golden", "ref_tables.json")


@pytest.fixture(scope="module")
def ref():
    with open(FIXTURE) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def H():
    deps = [SRC, os.path.join(CSRC, "chess.h"), os.path.join(CSRC, "evaltab.h")]
    if not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(d) for d in deps):
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-o", SO, SRC])
    L = ctypes.CDLL(SO)
    L.hc_eval_const.argtypes = [ctypes.c_int, ctypes.c_int]
    L.hc_eval_const.restype = ctypes.c_int32
    L.hc_zobrist_key.argtypes = [ctypes.c_int]
    L.hc_zobrist_key.restype = ctypes.c_uint64
    L.hc_castle_mask.argtypes = [ctypes.c_int]
    L.hc_castle_mask.restype = ctypes.c_uint64
    return L


@pytest.fixture(scope="module")
def O():
    L = R.lib()
    L.rc_eval_const.argtypes = [ctypes.c_int, ctypes.c_int]
    L.rc_eval_const.restype = ctypes.c_int32
    L.rc_zobrist_key.argtypes = [ctypes.c_int]
    L.rc_zobrist_key.restype = ctypes.c_uint64
    L.rc_castle_mask.argtypes = [ctypes.c_int]
    L.rc_castle_mask.restype = ctypes.c_uint64
    return L


MASKS = ("SMCRALL", "SMCRWHITEQ", "SMCRWHITEK", "SMCRBLACKQ", "SMCRBLACKK")


def mismatches(ref, get_const, get_key, get_mask):
    """Every entry where a copy differs from the reference numbers."""
    bad = []
    for which, name, n in ((0, "piece_values", 7), (1, "control", 64), (2, "pst", 7 * 64)):
        assert len(ref[name]) == n
        bad += [(name, i, get_const(which, i), ref[name][i]) for i in range(n) if get_const(which, i) != ref[name][i]]
    keys = [int(z, 16) for z in ref["zobrist"]]
    bad += [("zobrist", i, get_key(i), keys[i]) for i in range(18) if get_key(i) != keys[i]]
    bad += [(m, 0, get_mask(i), int(ref["castle_masks"][m], 16)) for i, m in enumerate(MASKS)
            if get_mask(i) != int(ref["castle_masks"][m], 16)]
    return bad


def test_oracle_tables_equal_reference(ref, O):
    assert mismatches(ref, O.rc_eval_const, O.rc_zobrist_key, O.rc_castle_mask) == []


def test_device_tables_equal_reference(ref, H):
    assert mismatches(ref, H.hc_eval_const, H.hc_zobrist_key, H.hc_castle_mask) == []
    # the folded per-square table the device evaluation reads
    for ty in range(1, 7):
        for sq in range(64):
            want = 10 + ref["piece_values"][ty] + ref["pst"][ty * 64 + sq] + ref["control"][sq]
            assert H.hc_eval_const(3, ty * 64 + sq) == want, (ty, sq)


def test_comparison_catches_a_single_planted_entry(ref, O):
    """One entry changed in a copy of the oracle's tables (planted on the CPU,
    in a wrapper) is reported, at that entry and no other."""
    def planted(which, i):
        v = O.rc_eval_const(which, i)
        return v + 1 if (which, i) == (2, 4 * 64 + 27) else v
    assert mismatches(ref, planted, O.rc_zobrist_key, O.rc_castle_mask) == [("pst", 4 * 64 + 27,
                                                                              ref["pst"][4 * 64 + 27] + 1,
                                                                              ref["pst"][4 * 64 + 27])]
    def planted_key(i):
        return O.rc_zobrist_key(i) ^ (1 << 17) if i == 16 else O.rc_zobrist_key(i)
    assert [b[:2] for b in mismatches(ref, O.rc_eval_const, planted_key, O.rc_castle_mask)] == [("zobrist", 16)]

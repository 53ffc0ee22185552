"""ctypes binding of the CPU oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check the device results, and by bench.py's cpu_baseline leg.  The product
never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

POS_DTYPE = np.dtype([("b", "<u8", (7,)), ("stm", "<i4"), ("pad", "<i4")])
STARTPOS = "rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1"


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "librefcpu.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        u64, i32, vp = ctypes.c_uint64, ctypes.c_int32, ctypes.c_void_p
        L.rc_init.restype = None
        L.rc_setboard.argtypes = [ctypes.c_char_p, vp]
        L.rc_setboard.restype = ctypes.c_int
        L.rc_computehash.argtypes = [vp, ctypes.c_int]
        L.rc_computehash.restype = u64
        for name in ("rc_attack_table", "rc_pawnpush_table"):
            getattr(L, name).argtypes = [ctypes.c_int]
            getattr(L, name).restype = u64
        for name in ("rc_inbetween", "rc_line", "rc_sqinbetween"):
            getattr(L, name).argtypes = [ctypes.c_int, ctypes.c_int]
            getattr(L, name).restype = u64
        L.rc_genmoves.argtypes = [vp, u64, ctypes.c_int, ctypes.c_int, vp, vp]
        L.rc_genmoves.restype = ctypes.c_int
        L.rc_makemove_code.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.rc_makemove_code.restype = ctypes.c_uint32
        L.rc_domove_host.argtypes = [vp, ctypes.c_uint32]
        L.rc_domove_host.restype = None
        L.rc_count_legal.argtypes = [vp]
        L.rc_count_legal.restype = ctypes.c_int
        L.rc_nth_legal.argtypes = [vp, ctypes.c_int]
        L.rc_nth_legal.restype = ctypes.c_uint32
        L.rc_perft.argtypes = [vp, ctypes.c_int]
        L.rc_perft.restype = u64
        L.rc_perft_divide.argtypes = [vp, ctypes.c_int, vp, vp]
        L.rc_perft_divide.restype = ctypes.c_int
        L.rc_perft_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        L.rc_perft_batch.restype = None
        L.rc_perft_bulk_batch.argtypes = L.rc_perft_batch.argtypes
        L.rc_perft_bulk_batch.restype = None
        L.rc_eval.argtypes = [vp]
        L.rc_eval.restype = i32
        L.rc_search.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp]
        L.rc_search.restype = i32
        L.rc_minimax.argtypes = [vp, ctypes.c_int, vp]
        L.rc_minimax.restype = i32
        L.rc_search_reversed.argtypes = [vp, ctypes.c_int, vp]
        L.rc_search_reversed.restype = i32
        L.rc_search_batch.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, vp, ctypes.c_int]
        L.rc_search_batch.restype = None
        L.rc_search_id_batch.argtypes = L.rc_search_batch.argtypes
        L.rc_search_id_batch.restype = None
        L.rc_search_id.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int, vp, vp]
        L.rc_search_id.restype = i32
        L.rc_gen_playouts.argtypes = [u64, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp]
        L.rc_gen_playouts.restype = ctypes.c_int
        L.rc_init()
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def setboard(fen):
    p = np.zeros(1, dtype=POS_DTYPE)
    rc = lib().rc_setboard(fen.encode(), _ptr(p))
    if rc != 0:
        raise ValueError("illegal FEN: " + fen)
    return p


def perft(pos, depth):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    return int(lib().rc_perft(_ptr(pos), depth))


def perft_batch(pos, depth, threads=None, bulk=False):
    """Leaf counts; bulk=True counts the moves of depth-1 positions without
    making them (the same counts)."""
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    out = np.zeros(len(pos), dtype=np.uint64)
    f = lib().rc_perft_bulk_batch if bulk else lib().rc_perft_batch
    f(_ptr(pos), len(pos), depth, _ptr(out), threads or os.cpu_count() or 1)
    return out


def divide(pos, depth):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    mv = np.zeros(256, dtype=np.uint32)
    cnt = np.zeros(256, dtype=np.uint64)
    n = lib().rc_perft_divide(_ptr(pos), depth, _ptr(mv), _ptr(cnt))
    return mv[:n].copy(), cnt[:n].copy()


def genmoves(pos, qs=False):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    out = np.zeros(64, dtype=np.uint64)
    chk = ctypes.c_int(0)
    b = pos["b"][0].copy()
    n = lib().rc_genmoves(_ptr(b), int(b[4]), int(pos["stm"][0]), int(qs), _ptr(out), ctypes.byref(chk))
    return n, out, bool(chk.value)


def evaluate(pos):
    b = np.ascontiguousarray(pos["b"][0], dtype=np.uint64)
    return int(lib().rc_eval(_ptr(b)))


def search(pos, depth, hist=None, ply_init=0):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    nodes = ctypes.c_uint64(0)
    bm = ctypes.c_uint32(0)
    h = None
    if hist is not None:
        h = np.ascontiguousarray(hist, dtype=np.uint64)
    s = lib().rc_search(_ptr(pos), depth, _ptr(h) if h is not None else None, ply_init,
                        ctypes.byref(nodes), ctypes.byref(bm))
    return int(s), int(nodes.value), int(bm.value)


def minimax(pos, depth):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    nodes = ctypes.c_uint64(0)
    s = lib().rc_minimax(_ptr(pos), depth, ctypes.byref(nodes))
    return int(s), int(nodes.value)


def search_reversed(pos, depth):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    nodes = ctypes.c_uint64(0)
    s = lib().rc_search_reversed(_ptr(pos), depth, ctypes.byref(nodes))
    return int(s), int(nodes.value)


def search_batch(pos, depth, threads=None):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    scores = np.zeros(len(pos), dtype=np.int32)
    nodes = np.zeros(len(pos), dtype=np.uint64)
    lib().rc_search_batch(_ptr(pos), len(pos), depth, _ptr(scores), _ptr(nodes),
                          threads or os.cpu_count() or 1)
    return scores, nodes


def search_id_batch(pos, depth, threads=None):
    """Iterative deepening 1..depth per position: (scores, moves made over all
    iterations) -- what a single-worker device search without a TT returns."""
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    scores = np.zeros(len(pos), dtype=np.int32)
    nodes = np.zeros(len(pos), dtype=np.uint64)
    lib().rc_search_id_batch(_ptr(pos), len(pos), depth, _ptr(scores), _ptr(nodes),
                             threads or os.cpu_count() or 1)
    return scores, nodes


def gen_playouts(seed, n, minply, maxply):
    out = np.zeros(n, dtype=POS_DTYPE)
    rc = lib().rc_gen_playouts(seed, n, minply, maxply, _ptr(out))
    if rc != 0:
        raise RuntimeError("playout generation failed")
    return out


def count_legal(pos):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    return int(lib().rc_count_legal(_ptr(pos)))


def nth_legal(pos, k):
    pos = np.ascontiguousarray(pos, dtype=POS_DTYPE)
    return int(lib().rc_nth_legal(_ptr(pos), k))


def domove(pos, move):
    p = np.array(pos, dtype=POS_DTYPE, copy=True)
    lib().rc_domove_host(_ptr(p), move)
    return p


def move2can(move):
    f, t = move & 63, (move >> 6) & 63
    s = "abcdefgh"[f & 7] + str((f >> 3) + 1) + "abcdefgh"[t & 7] + str((t >> 3) + 1)
    pf, pt = (move >> 18) & 15, (move >> 22) & 15
    if (pf >> 1) == 1 and (pt >> 1) != 1:
        s += {6: "q", 5: "r", 4: "b", 2: "n"}.get(pt >> 1, "")
    return s


def move_from_can(pos, can):
    """The legal move of pos written as can (e.g. "e2e4", "e7e8q")."""
    for k in range(count_legal(pos)):
        m = nth_legal(pos, k)
        if move2can(m) == can or move2can(m) == can[:4]:
            return m
    raise ValueError("no legal move %s" % can)

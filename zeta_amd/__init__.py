"""zeta_amd -- Python mirror of libzetahip.so, the MI355X (gfx950) hot path of
the Zeta chess engine.

Everything here goes through the C-ABI declared in include/zeta_hip.h (ctypes,
no torch types).  There is no CPU fallback: if the HIP library is missing or
no gfx950 device is present, the calls raise.

Reference interface mirrored (shintakezou/Zeta 099m):
  perft()        -> perft() + cl_run_perft  (src/search.c:225-271, src/clrun.c:979)
  search()       -> rootsearch() + cl_run_alphabeta (src/search.c:35-223, src/clrun.c:675)
Positions come in as the reference's quad-bitboard records (POS_DTYPE); the
FEN parser with the reference's setboard semantics is the front end's
(zeta_amd/host/zeta_engine.cpp).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libzetahip.so")

# zeta_pos: 7 x u64 quad-bitboard (src/types.h:60-66) + side to move
POS_DTYPE = np.dtype([("b", "<u8", (7,)), ("stm", "<i4"), ("pad", "<i4")])
assert POS_DTYPE.itemsize == 64

ZETA_PERFT_NOBULK = 1

_lib = None
_up = False


class ZetaError(RuntimeError):
    pass


class _KStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64),
                ("ms", ctypes.c_double), ("items", ctypes.c_uint64)]


class _Stats(ctypes.Structure):
    _fields_ = [("nk", ctypes.c_int), ("k", _KStat * 16)]


class SearchOpts(ctypes.Structure):
    _fields_ = [("depth", ctypes.c_int), ("workers", ctypes.c_int), ("tt_log2", ctypes.c_int),
                ("clear_tt", ctypes.c_int), ("max_nodes", ctypes.c_uint64), ("use_mlp", ctypes.c_int),
                ("mlp_min_depth", ctypes.c_int), ("log_records", ctypes.c_int), ("log_shift", ctypes.c_int),
                ("history", ctypes.c_void_p), ("history_len", ctypes.c_void_p),
                ("pv", ctypes.c_void_p), ("pv_len", ctypes.c_int), ("order_hint", ctypes.c_int),
                ("split_depth", ctypes.c_int), ("waves_per_simd", ctypes.c_int)]


ZETA_HISTORY = 128


RESULT_DTYPE = np.dtype([("score", "<i4"), ("bestmove", "<u4"), ("nodes", "<u8"), ("depth", "<i4"),
                         ("status", "<i4")])


class ServeOpts(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("workers", ctypes.c_int), ("tt_log2", ctypes.c_int),
                ("use_mlp", ctypes.c_int), ("mlp_min_depth", ctypes.c_int), ("grid", ctypes.c_int),
                ("watchdog_ms", ctypes.c_int), ("reserved", ctypes.c_int), ("mem_budget", ctypes.c_uint64)]


class _Pos(ctypes.Structure):
    _fields_ = [("b", ctypes.c_uint64 * 7), ("stm", ctypes.c_int32), ("pad", ctypes.c_int32)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("pos", _Pos), ("depth", ctypes.c_int), ("workers", ctypes.c_int), ("flags", ctypes.c_int),
                ("reserved", ctypes.c_int)]


SERVE_DONE_DTYPE = np.dtype([("ticket", "<u8"), ("score", "<i4"), ("bestmove", "<u4"), ("nodes", "<u8"),
                             ("depth", "<i4"), ("status", "<i4")])
ZETA_EBUSY = -6
ZETA_ECANCELED = -7
ZETA_EBUDGET = -8
ZETA_PENDING = 1
ZETA_REQ_LEARNED = 1


class _Config(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("tt_log2", ctypes.c_int),
                ("frontier_budget", ctypes.c_uint64), ("workers", ctypes.c_int),
                ("verbose", ctypes.c_int)]


def lib():
    """Load libzetahip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ZetaError("libzetahip.so not built (run __graft_entry__.build() or make -C zeta_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
        L.zeta_init.argtypes = [ctypes.POINTER(_Config)]
        L.zeta_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.zeta_device_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.zeta_perft_batch.argtypes = [vp, i32, i32, vp]
        L.zeta_perft_batch_device.argtypes = [vp, i32, i32, vp, i32]
        L.zeta_playouts.argtypes = [u64, i32, i32, i32, vp]
        L.zeta_stats_get.argtypes = [ctypes.POINTER(_Stats)]
        L.zeta_set_timing.argtypes = [i32]
        L.zeta_search_batch.argtypes = [vp, i32, ctypes.POINTER(SearchOpts), vp]
        L.zeta_search_counters.argtypes = [vp, i32]
        L.zeta_search_abort.argtypes = [i32]
        L.zeta_mlp_init.argtypes = [u64]
        L.zeta_mlp_set.argtypes = [vp, ctypes.c_size_t]
        L.zeta_mlp_get.argtypes = [vp, ctypes.c_size_t]
        L.zeta_mlp_grad.argtypes = [vp, i32, vp, vp, ctypes.c_size_t]
        L.zeta_mlp_train.argtypes = [i32, i32, ctypes.c_float, u64, vp]
        L.zeta_records_reserve.argtypes = [u64]
        L.zeta_mlp_score.argtypes = [vp, i32, vp, vp, vp]
        L.zeta_records_count.argtypes = [vp]
        L.zeta_records_get.argtypes = [vp, u64, vp]
        L.zeta_records_put.argtypes = [vp, u64]
        L.zeta_dev_alloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.zeta_dev_free.argtypes = [vp]
        L.zeta_memcpy_htod.argtypes = [vp, vp, ctypes.c_size_t]
        L.zeta_memcpy_dtoh.argtypes = [vp, vp, ctypes.c_size_t]
        L.zeta_memcpy_dtod.argtypes = [vp, vp, ctypes.c_size_t]
        L.zeta_serve_start.argtypes = [ctypes.POINTER(ServeOpts)]
        L.zeta_serve_cost.argtypes = [i32, vp]
        L.zeta_serve_submit.argtypes = [ctypes.POINTER(ServeRequest), vp]
        L.zeta_serve_poll.argtypes = [u64, vp]
        L.zeta_serve_reap.argtypes = [vp, i32, vp]
        L.zeta_serve_cancel.argtypes = [u64]
        L.zeta_serve_stop.argtypes = [vp]
        for f in ("zeta_init", "zeta_shutdown", "zeta_last_error", "zeta_device_info",
                  "zeta_perft_batch", "zeta_perft_batch_device", "zeta_playouts",
                  "zeta_stats_get", "zeta_stats_reset", "zeta_set_timing", "zeta_dev_alloc",
                  "zeta_dev_free", "zeta_memcpy_htod", "zeta_memcpy_dtoh", "zeta_memcpy_dtod", "zeta_synchronize",
                  "zeta_search_batch", "zeta_search_counters", "zeta_search_abort", "zeta_mlp_init", "zeta_mlp_set",
                  "zeta_mlp_get", "zeta_mlp_grad", "zeta_mlp_train", "zeta_records_reserve",
                  "zeta_records_count", "zeta_records_clear", "zeta_records_get", "zeta_records_put",
                  "zeta_mlp_score", "zeta_serve_start", "zeta_serve_cost", "zeta_serve_submit",
                  "zeta_serve_poll", "zeta_serve_reap", "zeta_serve_cancel", "zeta_serve_stop"):
            getattr(L, f).restype = ctypes.c_int
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().zeta_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc, what):
    if rc != 0:
        raise ZetaError("%s failed (%d): %s" % (what, rc, last_error()))


def init(device=0, tt_log2=0, frontier_budget=0, workers=0, verbose=0):
    """zeta_init: bind the gfx950 device (raises when there is none)."""
    global _up
    cfg = _Config(device, tt_log2, frontier_budget, workers, verbose)
    _check(lib().zeta_init(ctypes.byref(cfg)), "zeta_init")
    _up = True


def ensure_init(**kw):
    if not _up:
        init(**kw)


def shutdown():
    global _up
    if _lib is not None:
        _check(_lib.zeta_shutdown(), "zeta_shutdown")
    _up = False


def device_info():
    buf = ctypes.create_string_buffer(256)
    _check(lib().zeta_device_info(buf, 256), "zeta_device_info")
    return buf.value.decode()


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def as_positions(pos):
    pos = np.ascontiguousarray(pos)
    if pos.dtype != POS_DTYPE:
        pos = pos.view(POS_DTYPE) if pos.dtype.itemsize == 64 else np.ascontiguousarray(pos, dtype=POS_DTYPE)
    return pos


def perft(pos, depth):
    """Leaf counts of perft(depth) for every position (u64 array)."""
    ensure_init()
    pos = as_positions(pos)
    out = np.zeros(len(pos), dtype=np.uint64)
    _check(lib().zeta_perft_batch(_ptr(pos), len(pos), int(depth), _ptr(out)), "zeta_perft_batch")
    return out


def perft_device(d_pos, n, depth, d_counts, nobulk=False):
    """zeta_perft_batch_device on device pointers (ints)."""
    ensure_init()
    _check(lib().zeta_perft_batch_device(ctypes.c_void_p(d_pos), int(n), int(depth), ctypes.c_void_p(d_counts),
                                         ZETA_PERFT_NOBULK if nobulk else 0), "zeta_perft_batch_device")


def search(pos, depth, workers=1, tt_log2=26, clear_tt=True, max_nodes=0, use_mlp=False,
           mlp_min_depth=3, log_records=False, log_shift=0, history=None, pv_len=0, order_hint=False,
           split_depth=0, waves_per_simd=0):
    """Fixed-depth alpha-beta + quiescence of every position (iterative
    deepening 1..depth).  Returns a structured array (score, bestmove, nodes,
    depth, status) -- the mirror of rootsearch's PV[0]/PV[1]/node count.
    history: optional list (one entry per position) of the hashes of the
    positions played before it, newest first, for repetitions across the root.
    pv_len > 0: also return the principal variations (n x pv_len move codes).
    order_hint: search with the moves of the previous batch's table first
    (zeta_search_opts.order_hint).  split_depth, waves_per_simd: the
    autotuner's ABDADA split depth and persistent wavefronts per SIMD (0 =
    default)."""
    ensure_init()
    pos = as_positions(pos)
    out = np.zeros(len(pos), dtype=RESULT_DTYPE)
    hist = hlen = None
    if history is not None:
        if len(history) != len(pos):
            raise ValueError("history needs one entry per position")
        hist = np.zeros((len(pos), ZETA_HISTORY), dtype=np.uint64)
        hlen = np.zeros(len(pos), dtype=np.int32)
        for i, h in enumerate(history):
            h = np.asarray(h, dtype=np.uint64)[:ZETA_HISTORY]
            hist[i, :len(h)] = h
            hlen[i] = len(h)
    pv = np.zeros((len(pos), max(int(pv_len), 0)), dtype=np.uint32)
    o = SearchOpts(int(depth), int(workers), int(tt_log2), int(bool(clear_tt)), int(max_nodes),
                   int(bool(use_mlp)), int(mlp_min_depth), int(bool(log_records)), int(log_shift),
                   hist.ctypes.data if hist is not None else None, hlen.ctypes.data if hlen is not None else None,
                   pv.ctypes.data if pv_len > 0 else None, int(pv_len), int(bool(order_hint)),
                   int(split_depth), int(waves_per_simd))
    _check(lib().zeta_search_batch(_ptr(pos), len(pos), ctypes.byref(o), _ptr(out)), "zeta_search_batch")
    return (out, pv) if pv_len > 0 else out


def search_abort(on=True):
    """zeta_search_abort: stop a search running on another thread (and every
    search started while on); on=False re-arms."""
    ensure_init()
    _check(lib().zeta_search_abort(1 if on else 0), "zeta_search_abort")


def search_counters():
    """TT probes, TT hits, TT scores used, TT stores, ABDADA deferrals, MAXPLY
    cuts of the last search batch."""
    ensure_init()
    out = np.zeros(9, dtype=np.uint64)
    _check(lib().zeta_search_counters(_ptr(out), 9), "zeta_search_counters")
    keys = ("tt_probes", "tt_hits", "tt_cuts", "tt_stores", "abdada_defers", "maxply_cuts", "mlp_nodes",
            "records_logged", "worker_nodes")
    return {k: int(out[i]) for i, k in enumerate(keys)}


def playouts(seed, n, minply, maxply):
    """Seeded random-playout positions, generated on the GPU."""
    ensure_init()
    out = np.zeros(n, dtype=POS_DTYPE)
    _check(lib().zeta_playouts(int(seed), int(n), int(minply), int(maxply), _ptr(out)), "zeta_playouts")
    return out


class DeviceBuffer:
    """Device memory owned through the C-ABI (zeta_dev_alloc / zeta_dev_free)."""

    def __init__(self, nbytes):
        ensure_init()
        p = ctypes.c_void_p()
        _check(lib().zeta_dev_alloc(ctypes.byref(p), int(nbytes)), "zeta_dev_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().zeta_memcpy_htod(ctypes.c_void_p(self.ptr), _ptr(arr), arr.nbytes), "zeta_memcpy_htod")

    def download(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        _check(lib().zeta_memcpy_dtoh(_ptr(arr), ctypes.c_void_p(self.ptr), arr.nbytes), "zeta_memcpy_dtoh")
        return arr

    def free(self):
        if self.ptr:
            _check(lib().zeta_dev_free(ctypes.c_void_p(self.ptr)), "zeta_dev_free")
            self.ptr = None


def copy_device(dst_ptr, src_ptr, nbytes):
    """Device-to-device copy by the library's streaming kernel (16-byte multiples)."""
    _check(lib().zeta_memcpy_dtod(ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr), int(nbytes)),
           "zeta_memcpy_dtod")


def synchronize():
    ensure_init()
    _check(lib().zeta_synchronize(), "zeta_synchronize")


def set_timing(on=True):
    ensure_init()
    _check(lib().zeta_set_timing(1 if on else 0), "zeta_set_timing")


def stats():
    """Per-kernel HIP-event totals: {name: (launches, ms, items)}."""
    ensure_init()
    st = _Stats()
    _check(lib().zeta_stats_get(ctypes.byref(st)), "zeta_stats_get")
    return {st.k[i].name.decode(): (st.k[i].launches, st.k[i].ms, st.k[i].items) for i in range(st.nk)}


def stats_reset():
    ensure_init()
    _check(lib().zeta_stats_reset(), "zeta_stats_reset")


# ------------------------------------------------------------ learned ordering
MLP_NPARAMS = 240961
PAIR_DTYPE = np.dtype([("b", "<u8", (4,)), ("stm", "<u4"), ("mpos", "<u4"), ("mneg", "<u4"), ("depth", "<u4")])


def mlp_init(seed=1):
    ensure_init()
    _check(lib().zeta_mlp_init(int(seed)), "zeta_mlp_init")


def mlp_set(params):
    ensure_init()
    p = np.ascontiguousarray(params, dtype=np.float32)
    _check(lib().zeta_mlp_set(_ptr(p), p.size), "zeta_mlp_set")


def mlp_get():
    ensure_init()
    p = np.zeros(MLP_NPARAMS, dtype=np.float32)
    _check(lib().zeta_mlp_get(_ptr(p), p.size), "zeta_mlp_get")
    return p


def mlp_grad(pairs):
    ensure_init()
    pairs = np.ascontiguousarray(pairs, dtype=PAIR_DTYPE)
    g = np.zeros(MLP_NPARAMS, dtype=np.float32)
    loss = ctypes.c_float(0)
    _check(lib().zeta_mlp_grad(_ptr(pairs), len(pairs), ctypes.byref(loss), _ptr(g), g.size), "zeta_mlp_grad")
    return float(loss.value), g


def mlp_train(steps, pairs_per_step=2048, lr=1e-3, seed=1):
    ensure_init()
    losses = np.zeros(max(steps, 1), dtype=np.float32)
    _check(lib().zeta_mlp_train(int(steps), int(pairs_per_step), float(lr), int(seed), _ptr(losses)),
           "zeta_mlp_train")
    return losses[:steps]


def records_reserve(capacity):
    ensure_init()
    _check(lib().zeta_records_reserve(int(capacity)), "zeta_records_reserve")


def records_count():
    ensure_init()
    n = ctypes.c_uint64(0)
    _check(lib().zeta_records_count(ctypes.byref(n)), "zeta_records_count")
    return int(n.value)


def records_clear():
    ensure_init()
    _check(lib().zeta_records_clear(), "zeta_records_clear")


def records_get(maxn=None):
    ensure_init()
    n = records_count() if maxn is None else int(maxn)
    out = np.zeros(n, dtype=PAIR_DTYPE)
    got = ctypes.c_uint64(0)
    _check(lib().zeta_records_get(_ptr(out), n, ctypes.byref(got)), "zeta_records_get")
    return out[:got.value]


def records_put(pairs):
    ensure_init()
    pairs = np.ascontiguousarray(pairs, dtype=PAIR_DTYPE)
    _check(lib().zeta_records_put(_ptr(pairs), len(pairs)), "zeta_records_put")


def mlp_score(pos, moves_per_pos):
    """Scores the search's network assigns to the given moves of each position."""
    ensure_init()
    pos = as_positions(pos)
    offs = np.zeros(len(pos) + 1, dtype=np.int32)
    offs[1:] = np.cumsum([len(m) for m in moves_per_pos])
    mv = np.ascontiguousarray(np.concatenate([np.asarray(m, dtype=np.uint32) for m in moves_per_pos]))
    out = np.zeros(len(mv), dtype=np.float32)
    _check(lib().zeta_mlp_score(_ptr(pos), len(pos), _ptr(mv), _ptr(offs), _ptr(out)), "zeta_mlp_score")
    return [out[offs[i]:offs[i + 1]] for i in range(len(pos))]


# ---------------------------------------------------------------- serving
class Server:
    """Continuous-batching search service on the device (zeta_serve_*).

    A persistent grid of wavefront workers takes requests from slots in
    pinned host memory as they are submitted; idle workers help running
    requests.  Each request is a fixed-depth root search with the same
    result as search() on that position (a shared TT cannot change a root
    score).  While a Server is running the other compute calls raise
    (ZETA_EBUSY).
    """

    def __init__(self, slots=256, workers=8, tt_log2=24, use_mlp=False, mlp_min_depth=2, grid=0,
                 watchdog_ms=10000, mem_budget=0):
        opts = ServeOpts(slots, workers, tt_log2, int(use_mlp), mlp_min_depth, grid, watchdog_ms, 0, mem_budget)
        _check(lib().zeta_serve_start(ctypes.byref(opts)), "zeta_serve_start")
        self.slots = slots
        self._done = np.zeros(slots, dtype=SERVE_DONE_DTYPE)
        self.running = True

    def cost(self, workers=0):
        out = ctypes.c_uint64()
        _check(lib().zeta_serve_cost(workers, ctypes.byref(out)), "zeta_serve_cost")
        return out.value

    def submit(self, pos, depth, workers=0, learned=False):
        """Ticket of the admitted request, or None when it cannot be admitted now."""
        p = as_positions(pos)
        req = ServeRequest()
        ctypes.memmove(ctypes.byref(req.pos), p.ctypes.data, 64)
        req.depth = depth
        req.workers = workers
        req.flags = ZETA_REQ_LEARNED if learned else 0
        t = ctypes.c_uint64()
        rc = lib().zeta_serve_submit(ctypes.byref(req), ctypes.byref(t))
        if rc == ZETA_EBUSY:
            return None
        _check(rc, "zeta_serve_submit")
        return t.value

    def poll(self, ticket):
        """Result record of a finished request, or None while it runs."""
        out = np.zeros(1, dtype=RESULT_DTYPE)
        rc = lib().zeta_serve_poll(ticket, _ptr(out))
        if rc == ZETA_PENDING:
            return None
        _check(rc, "zeta_serve_poll")
        return out[0]

    def reap(self):
        """All finished requests not yet collected (SERVE_DONE_DTYPE array)."""
        n = ctypes.c_int()
        _check(lib().zeta_serve_reap(_ptr(self._done), self.slots, ctypes.byref(n)), "zeta_serve_reap")
        return self._done[:n.value].copy()

    def cancel(self, ticket):
        _check(lib().zeta_serve_cancel(ticket), "zeta_serve_cancel")

    def stop(self):
        """Stop the grid; returns the nodes made by all workers (helpers included)."""
        if not self.running:
            return 0
        tot = ctypes.c_uint64()
        self.running = False
        _check(lib().zeta_serve_stop(ctypes.byref(tot)), "zeta_serve_stop")
        return tot.value

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.stop()

/*
 * zeta_hip.h -- C-ABI of libzetahip.so, the MI355X (gfx950) hot path of the
 * Zeta chess engine: batched perft, batched fixed-depth alpha-beta search,
 * learned move ordering and a continuous-batching search service.
 *
 * Plain C: pointers and sizes, no HIP or torch types.  Every function returns
 * 0 on success or a negative ZETA_E* code; zeta_last_error() gives the text.
 * One library context per process (the reference is single-threaded too:
 * src/clrun.c keeps one device, one queue and process-global buffers).
 *
 * Positions cross the boundary as the reference's 7 x u64 quad-bitboard
 * (src/types.h:60-66; src/zeta.h GLOBAL_BOARD) plus the side to move.
 */
#ifndef ZETA_HIP_H
#define ZETA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ZETA_OK 0
#define ZETA_EINVAL (-1)   /* bad argument */
#define ZETA_ENODEV (-2)   /* no usable GPU / HIP runtime error at init */
#define ZETA_EHIP (-3)     /* HIP runtime error */
#define ZETA_ENOMEM (-4)   /* device allocation failed or budget exceeded */
#define ZETA_ESTATE (-5)   /* library not initialised / wrong state */
#define ZETA_EBUSY (-6)    /* not admitted now (serving slots / memory budget), or the device is serving */
#define ZETA_ECANCELED (-7) /* serving request cancelled (zeta_serve_cancel / stop / watchdog), or a search root stopped by zeta_search_abort */
#define ZETA_EBUDGET (-8)   /* search root stopped at its node budget (zeta_search_opts.max_nodes) */
#define ZETA_PENDING 1      /* zeta_serve_poll: not finished yet */

/* quad-bitboard slots (src/types.h:60-66) */
#define ZETA_QBBBLACK 0 /* black pieces */
#define ZETA_QBBP1 1    /* piece type bit 0 */
#define ZETA_QBBP2 2    /* piece type bit 1 */
#define ZETA_QBBP3 3    /* piece type bit 2 */
#define ZETA_QBBPMVD 4  /* moved flags: castle rights, e.p. pawn */
#define ZETA_QBBHASH 5  /* Zobrist hash (rotate-left scheme) */
#define ZETA_QBBHMC 6   /* half-move clock */

typedef struct {
  uint64_t board[7];
  int32_t stm; /* 0 white, 1 black */
  int32_t pad;
} zeta_pos; /* 64 bytes */

typedef struct {
  int device;               /* HIP device ordinal */
  int tt_log2;              /* search TT entries = 2^tt_log2 (16 B each); 0 = no TT */
  uint64_t frontier_budget; /* perft: max child moves held per level (0 = default 1<<28; at least 512) */
  int workers;              /* search: wavefront workers per position (ABDADA), 0 = default */
  int verbose;
} zeta_config;

/* Library set-up / tear-down (replaces cl_init_device / cl_release_device,
 * src/clrun.h:26,31). */
int zeta_init(const zeta_config *cfg);
int zeta_shutdown(void);
int zeta_last_error(char *buf, size_t n);
/* Name of the device in use (hipDeviceProp_t::gcnArchName + name). */
int zeta_device_info(char *buf, size_t n);

/* ------------------------------------------------------------------ perft
 * Leaf counts of perft(depth) for n root positions, one u64 per position.
 * Replaces cl_run_perft (src/clrun.h:29; src/clrun.c:979-1147) driven by
 * perft() (src/search.c:225-271), which counts one root per launch.
 * Node semantics: the reference's move set (queen-only promotion, its
 * castling and e.p. rules); the count is the number of leaves (Q7).
 * zeta_perft_batch: host buffers.  zeta_perft_batch_device: d_pos / d_counts
 * are device pointers (n zeta_pos in, n u64 out); flags below. */
#define ZETA_PERFT_NOBULK 1 /* visit every leaf (generate its moves) instead of counting at depth-1 */
int zeta_perft_batch(const zeta_pos *pos, int n, int depth, uint64_t *counts);
int zeta_perft_batch_device(const void *d_pos, int n, int depth, void *d_counts, int flags);
/* zeta_perft_batch_device without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * frontier lives in a fixed 256 MiB array the library holds on every device
 * it runs on until the process ends; a level too large for it is walked in
 * chunks.  n <= ZETA_NOALLOC_MAX_ROOTS, depth <= ZETA_NOALLOC_MAX_DEPTH, else
 * ZETA_EINVAL.  Every call on a device shares that array: calls that overlap
 * on one device race.  Same counts as zeta_perft_batch_device. */
#define ZETA_NOALLOC_MAX_ROOTS 65536
#define ZETA_NOALLOC_MAX_DEPTH 8
int zeta_perft_batch_device_noalloc(const void *d_pos, int n, int depth, void *d_counts, int flags);

/* Seeded random-playout positions generated on the GPU: position i tries
 * attempts a = 0,1,..; rng = splitmix64 seeded with
 * seed ^ i*0x9E3779B97F4A7C15 ^ a*0xD1B54A32D192ED03; plies =
 * minply + r % (maxply-minply+1); each ply plays legal move r % nlegal in
 * canonical order (origin square ascending, then target ascending) with the
 * host domove rules (src/bitboard.c:583-693); an attempt is rejected when a
 * side has no legal move or the half-move clock reaches 100. */
int zeta_playouts(uint64_t seed, int n, int minply, int maxply, zeta_pos *out);

/* ----------------------------------------------------------------- search
 * Fixed-depth alpha-beta + quiescence of n root positions (iterative
 * deepening 1..depth per root, as rootsearch, src/search.c:77-198).
 * Replaces cl_run_alphabeta (src/clrun.h:28; src/clrun.c:675-977).
 * Node semantics: alphabeta_gpu (src/zeta.cl:1319-1939) with null move, LMR
 * and IID off.  score = PV[0] as rootsearch reports it (src/search.c:137):
 * the exact minimax value of the depth-limited tree, independent of move
 * order, TT and worker count.  nodes = moves made (src/zeta.cl:2283-2285). */
typedef struct {
  int depth;          /* target depth, 1..40 */
  int workers;        /* wavefront workers per root at most (ABDADA helpers), 1..1024 */
  int tt_log2;        /* TT size 2^tt_log2 x 16 B; -1 = zeta_config.tt_log2; 0 = none */
  int clear_tt;       /* 1 = clear the TT before the batch */
  uint64_t max_nodes; /* per-worker node budget per root, 0 = none: the first worker over it
                         stops the root (src/zeta.cl:1945-1947); status ZETA_EBUDGET */
  int use_mlp;        /* learned move ordering at interior nodes (needs zeta_mlp_load) */
  int mlp_min_depth;  /* nodes with depth >= this use the learned ordering */
  int log_records;    /* record cutoff examples for training (zeta_records_*) */
  int log_shift;      /* log the cut nodes whose hash has its low log_shift bits (of bits 40..) zero */
  /* Game history for repetitions across the root (the reference's
   * GLOBAL_HASHHISTORY / PLY, src/zeta.cl:1560-1577): NULL, or n blocks of
   * ZETA_HISTORY hashes, block i holding root i's predecessors newest first
   * (entry 0 = the position one ply before the root), history_len[i] of them
   * valid.  Without it only the searched path is checked. */
  const uint64_t *history;
  const int32_t *history_len;
  /* Principal variation (NULL = none): n rows of pv_len move codes, the best
   * move then the moves the transposition table holds along it, 0-padded
   * (the reference's PV[1..], src/zeta.cl:2450-2512). */
  uint32_t *pv;
  int pv_len;
  /* 1 = order by hints: before the batch the transposition table as the
   * previous batch left it (same size) is copied aside, and at every
   * full-width node the move it holds goes first, above the TT move.  After
   * a completed search of the same roots at the same depth that table holds
   * the best and refuting moves the search found, so the hinted search walks
   * a near-minimal tree: its node count measures what any move ordering
   * could save.  Scores are unchanged (any order gives them). */
  int order_hint;
  /* Launch and split knobs (set by the front end's autotuner, config.txt
   * keys split_depth / waves_per_simd; 0 = default):
   * split_depth: ABDADA -- a worker defers a sibling that another worker is
   *   searching, and takes a busy-table slot for the one it searches, at
   *   nodes with at least this much depth left below the child (default 2);
   *   the reference gates ABDADA on any full-width node below the root
   *   (src/zeta.cl:1648-1660, 1743-1750), here the depth is a knob.
   * waves_per_simd: persistent search wavefronts per SIMD, 1..2 (default:
   *   as many as fit, 2 on gfx950); replaces the work-group count the
   *   reference tunes through threadsY (src/clconfig.c:878-960). */
  int split_depth;
  int waves_per_simd;
} zeta_search_opts;
#define ZETA_HISTORY 128

typedef struct {
  int32_t score;     /* centipawns, side to move; mate = +-(32000 - plies) */
  uint32_t bestmove; /* reference move code (src/types.h:67-74) */
  uint64_t nodes;    /* moves made by all workers on this root */
  int32_t depth;     /* deepest completed iteration */
  int32_t status;    /* 0 ok; ZETA_EBUDGET: stopped at the node budget, score and bestmove are
                        those of the deepest iteration the stopping worker completed (0 if none;
                        the reference's host discards such a search, src/search.c:133);
                        ZETA_ECANCELED: stopped by zeta_search_abort; ZETA_ENOMEM: move
                        stack overflow */
} zeta_result;

int zeta_search_batch(const zeta_pos *pos, int n, const zeta_search_opts *opts, zeta_result *res);
/* Abort switch of the search batches.  No counterpart at the reference's
 * boundary (its helpers stop on the kernel's own finito flag and the node
 * budget, src/zeta.cl:1940-1953).  on = 1: a zeta_search_batch running on
 * another host thread stops at its workers' next poll (every 64 nodes), and
 * every batch started while the switch is on returns at once; the roots left
 * unfinished get status ZETA_ECANCELED.  on = 0 re-arms.  The one call that
 * may be made while zeta_search_batch runs; it touches a word in pinned host
 * memory only. */
int zeta_search_abort(int on);
/* Counters of the last zeta_search_batch: [0] TT probes, [1] TT key hits,
 * [2] TT scores used, [3] TT stores, [4] ABDADA deferrals, [5] MAXPLY cuts,
 * [6] nodes ordered by the network, [7] training pairs logged, [8] nodes
 * made by all workers (after zeta_serve_stop: of the whole serving session).
 * Each probe / store moves one 16-byte entry. */
int zeta_search_counters(uint64_t *out, int n);

/* ---------------------------------------------------------------- serving
 * Continuous batching of search requests.  No single counterpart in the
 * reference, whose host runs one root per cl_run_alphabeta launch
 * (src/search.c:77-198); each request is a rootsearch (fixed depth,
 * iterative deepening, node semantics and scores exactly as
 * zeta_search_batch).  zeta_serve_start launches a persistent grid of
 * wavefront workers that owns the device until zeta_serve_stop (the other
 * compute entry points return ZETA_EBUSY meanwhile).  Requests live in
 * request slots in pinned host memory; an idle worker claims the next
 * submitted request, and workers with nothing to claim join running requests
 * (ABDADA helpers, up to the request's worker limit), so requests that
 * arrive while others run start as soon as a worker frees up and share the
 * device with them.  All requests share one transposition table, whose
 * entries cannot change a root score: each result is the same alone and in
 * any mix.
 * Admission: a request costs workers x (one worker's LDS stack, MAXPLY plies
 * + move stack) + the TT share (TT bytes / slots); zeta_serve_submit returns
 * ZETA_EBUSY while the slots or mem_budget cannot take it (retry after
 * zeta_serve_reap frees some), ZETA_ENOMEM if it could never be admitted.
 * Every host call renews a heartbeat; if the host makes no serving call for
 * watchdog_ms the grid cancels its requests and exits. */
typedef struct {
  int slots;            /* request slots (requests in flight), 1..4096 */
  int workers;          /* default and maximum workers per request, 1..255 */
  int tt_log2;          /* shared TT size 2^tt_log2 x 16 B; -1 = zeta_config.tt_log2; 0 = none */
  int use_mlp;          /* requests with ZETA_REQ_LEARNED use the loaded network */
  int mlp_min_depth;
  int grid;             /* persistent workers, 0 = fill the device */
  int watchdog_ms;      /* 0 = 10000 */
  int reserved;
  uint64_t mem_budget;  /* admission budget in bytes, 0 = no limit */
} zeta_serve_opts;
#define ZETA_REQ_LEARNED 1 /* learned move ordering (plain ordering when no network is loaded) */
typedef struct {
  zeta_pos pos;
  int depth;            /* 1..63 */
  int workers;          /* 0 = zeta_serve_opts.workers */
  int flags;            /* ZETA_REQ_* */
  int reserved;
} zeta_serve_request;
typedef struct {
  uint64_t ticket;
  zeta_result res;      /* status ZETA_OK, ZETA_ENOMEM (move stack overflow) or ZETA_ECANCELED */
} zeta_serve_done;
int zeta_serve_start(const zeta_serve_opts *opts);
/* bytes a request with this many workers charges against mem_budget */
int zeta_serve_cost(int workers, uint64_t *bytes);
int zeta_serve_submit(const zeta_serve_request *req, uint64_t *ticket);
/* ZETA_OK and *res when the request has finished (its slot is freed), else ZETA_PENDING */
int zeta_serve_poll(uint64_t ticket, zeta_result *res);
/* all finished requests not yet collected, at most max; *n = how many */
int zeta_serve_reap(zeta_serve_done *out, int max, int *n);
int zeta_serve_cancel(uint64_t ticket);
/* cancels what is still running, waits for the grid to exit, frees the
 * slots; counters as zeta_search_counters; *total_nodes (may be NULL) = the
 * nodes made by all workers, helpers included */
int zeta_serve_stop(uint64_t *total_nodes);

/* ------------------------------------------------ learned move ordering
 * No counterpart in the reference (its ordering is fixed scores,
 * src/zeta.cl:2059-2104).  A 908 -> 256 -> 32 -> 1 ReLU network with BF16
 * weights and FP32 accumulation scores every legal move at interior nodes
 * when zeta_search_opts.use_mlp is set.  Parameters cross the boundary as
 * FP32 in this order: W1[908][256], b1[256], W2[256][32], b2[32], W3[32], b3.
 * Training pairs (48 B each: 4 x u64 position planes, u32 side to move,
 * u32 cutoff move, u32 failed move, u32 depth) are logged by searches with
 * log_records set, or supplied with zeta_records_put. */
#define ZETA_MLP_NPARAMS 240961
int zeta_mlp_init(uint64_t seed);
int zeta_mlp_set(const float *params, size_t n);
int zeta_mlp_get(float *params, size_t n);
/* loss and gradient (ABI order) of npairs pairs (multiple of 16), no update */
int zeta_mlp_grad(const void *pairs, int npairs, float *loss, float *grads, size_t n);
/* Adam steps on random batches of the logged pairs; losses[steps] may be NULL */
int zeta_mlp_train(int steps, int pairs_per_step, float lr, uint64_t seed, float *losses);
/* network scores (as the search uses them: 1/256 resolution, clamped) of
 * moves[offsets[i] .. offsets[i+1]) from position i */
int zeta_mlp_score(const zeta_pos *pos, int npos, const uint32_t *moves, const int32_t *offsets, float *scores);
int zeta_records_reserve(uint64_t capacity);
int zeta_records_count(uint64_t *n);
int zeta_records_clear(void);
int zeta_records_get(void *out, uint64_t max, uint64_t *n);
int zeta_records_put(const void *in, uint64_t n);

/* ------------------------------------------------------------ device memory
 * Device buffers for callers that keep inputs resident on the GPU (the
 * *_device entry points).  Plain pointers and sizes; all copies are
 * synchronous with respect to the library's stream.  No counterpart at the
 * reference's boundary: its buffers stay inside src/clrun.c
 * (clCreateBuffer / clEnqueueReadBuffer, src/clrun.c:279-1225). */
int zeta_dev_alloc(void **dptr, size_t bytes);
int zeta_dev_free(void *dptr);
int zeta_memcpy_htod(void *dst, const void *src, size_t bytes);
int zeta_memcpy_dtoh(void *dst, const void *src, size_t bytes);
/* Device-to-device copy by the library's own streaming kernel (16-byte
 * aligned pointers, size a multiple of 16).  Besides copying, it is the
 * known-size kernel that calibrates HBM byte counters (FETCH_SIZE,
 * WRITE_SIZE): it reads and writes exactly `bytes` each. */
int zeta_memcpy_dtod(void *dst, const void *src, size_t bytes);
int zeta_synchronize(void);

/* ------------------------------------------------------------------ stats
 * Per-kernel HIP-event times.  No counterpart in the reference, which times
 * whole runs on the host (src/search.c:48-123, 243). */
#define ZETA_MAX_KSTATS 16
typedef struct {
  char name[32];
  uint64_t launches;
  double ms;           /* summed HIP-event time of the kernel's launches */
  uint64_t items;      /* units processed (positions, nodes, records) */
} zeta_kstat;
typedef struct {
  int nk;
  zeta_kstat k[ZETA_MAX_KSTATS];
} zeta_stats;
int zeta_stats_get(zeta_stats *st);
int zeta_stats_reset(void);
/* 1 = record HIP events around every kernel (needed for zeta_stats_get). */
int zeta_set_timing(int on);

#ifdef __cplusplus
}
#endif
#endif

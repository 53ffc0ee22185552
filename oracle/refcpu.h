/*
 * refcpu -- CPU ORACLE for the Zeta hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a clean-room C restatement of the reference engine's semantics
 * (shintakezou/Zeta 099m): quad-bitboard positions, the square-wise ("x64")
 * legal move generator of the perft and alpha-beta kernels, make/unmake,
 * Zobrist hashing, perft leaf counting, the integer evaluation and the
 * fixed-depth alpha-beta + quiescence node semantics.  Every function cites
 * the reference file:line it follows.
 *
 * Pins (tests/test_oracle.py): the reference's 14 perft KATs
 * (src/test.c:43-73), textbook perft counts, quirk positions, perft(d) as
 * the sum over children; table invariants; evaluation closed forms and
 * mirror antisymmetry; incremental hash = from-scratch hash; search =
 * minimax on small trees, order independence, colour symmetry, terminal
 * scores.  The evaluation tables, Zobrist keys and castle masks are pinned
 * entry by entry to numbers extracted from the reference
 * (tests/golden/ref_tables.json, tests/test_tables_pinned.py).  No
 * reference-held vector covers the search itself.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only to check device results or as the timed CPU
 * baseline.  The product (libzetahip.so) never links, loads or calls it.
 */
#ifndef ZETA_REFCPU_H
#define ZETA_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* quad-bitboard slots, src/types.h:60-66 */
enum { RC_QBBBLACK = 0, RC_QBBP1 = 1, RC_QBBP2 = 2, RC_QBBP3 = 3,
       RC_QBBPMVD = 4, RC_QBBHASH = 5, RC_QBBHMC = 6 };

#define RC_INF        32000        /* src/types.h:87 */
#define RC_MATESCORE  30000        /* src/types.h:88 */
#define RC_MAXPLY     64           /* src/types.h:76 */
#define RC_MAXGAMEPLY 1024         /* src/types.h:77 */

/* A position as the reference holds it: 7 x u64 (src/zeta.h GLOBAL_BOARD) + side to move. */
typedef struct {
  uint64_t b[7];
  int32_t  stm;
  int32_t  pad;
} rc_pos;

/* ---- setup / tables ---- */
void     rc_init(void);
/* FEN -> quad bitboard with setboard semantics (src/io.c:335-532, quirk Q3). 0 ok, -1 illegal. */
int      rc_setboard(const char *fen, rc_pos *out);
/* host computehash (src/bitboard.c:461-508) */
uint64_t rc_computehash(const uint64_t *b, int stm);
/* table accessors, for invariant tests */
uint64_t rc_attack_table(int idx);      /* AttackTables[7*64]        src/zetaperft.cl:620-636 */
uint64_t rc_pawnpush_table(int idx);    /* AttackTablesPawnPushes[128] src/zetaperft.cl:612-618 */
uint64_t rc_inbetween(int a, int b);    /* bbInBetween[64*64]        src/bitboard.c:52-118 */
uint64_t rc_line(int a, int b);         /* bbLine[64*64]             src/bitboard.c:119-185 */
uint64_t rc_sqinbetween(int a, int b);  /* closed form, src/bit.c:49-66 */

/* constants, for the test that pins them to the reference's numbers:
 * which 0 material[7], 1 square control[64], 2 piece-square[7*64] */
int32_t  rc_eval_const(int which, int idx);
uint64_t rc_zobrist_key(int i);         /* Zobrist[18]   src/zobrist.h:28-35 */
uint64_t rc_castle_mask(int i);         /* SMCRALL, SMCRWHITEQ/K, SMCRBLACKQ/K  src/types.h:114-124 */

/* ---- move generation ---- */
/* Per-square legal move bitboards exactly as the x64 generator leaves them in
 * globalbbMoves (src/zetaperft.cl:913-1300; src/zeta.cl:1040-1457).
 * qs!=0 suppresses quiet moves and castling.  Returns the legal-move count;
 * *incheck receives rootkic. */
int      rc_genmoves(const uint64_t *board4, uint64_t cr, int stm, int qs,
                     uint64_t out[64], int *incheck);
/* Pack the move lid->sqto as the movepicker does (src/zetaperft.cl:1363-1386). */
uint32_t rc_makemove_code(const uint64_t *board4, int stm, int lid, int sqto);
/* Host domove (src/bitboard.c:583-693): board, moved flags, hmc, hash. */
void     rc_domove_host(rc_pos *p, uint32_t move);
/* Number of legal moves of a position. */
int      rc_count_legal(const rc_pos *p);
/* The k-th legal move in canonical order (from ascending, then to ascending). */
uint32_t rc_nth_legal(const rc_pos *p, int k);

/* ---- perft ---- */
/* Leaf count of perft(depth) (src/zetaperft.cl:801-1477; Q7: leaves only). */
uint64_t rc_perft(const rc_pos *p, int depth);
/* Per-root-move leaf counts; returns number of root moves, fills moves/counts. */
int      rc_perft_divide(const rc_pos *p, int depth, uint32_t *moves, uint64_t *counts);
/* Multi-threaded batch perft over n positions (CPU baseline). */
void     rc_perft_batch(const rc_pos *p, int n, int depth, uint64_t *counts, int threads);
/* The same counts with bulk counting at depth 1 (the moves of a depth-1
 * position are counted, not made), single root and batch. */
uint64_t rc_perft_bulk(const rc_pos *p, int depth);
void     rc_perft_bulk_batch(const rc_pos *p, int n, int depth, uint64_t *counts, int threads);

/* ---- evaluation / search ---- */
/* Static eval, white-positive, as evalscore after the x64 sum (src/zeta.cl:1479-1545). */
int32_t  rc_eval(const uint64_t *board4);
/* Fixed-depth alpha-beta + quiescence root score with the node semantics of
 * alphabeta_gpu (src/zeta.cl:1319-1939) with null move, LMR, IID and the
 * transposition tables switched off.  hist may be NULL (history = root hash).
 * Returns PV[0] as rootsearch reports it (+-INF mapped to 0, src/search.c:137).
 * *nodes receives moves made (src/zeta.cl:2283-2285). */
int32_t  rc_search(const rc_pos *p, int depth, const uint64_t *hist, int ply_init,
                   uint64_t *nodes, uint32_t *bestmove);
/* Iterative deepening 1..depth of rc_search (src/search.c:77-198), the
 * killers and counter moves kept across the iterations and a mate within an
 * iteration's depth ending them (src/search.c:177-186): the score of the
 * last iteration and the moves made over all of them -- the node count a
 * single-worker device search without a transposition table must equal. */
int32_t  rc_search_id(const rc_pos *p, int depth, const uint64_t *hist, int ply_init,
                      uint64_t *nodes, uint32_t *bestmove);
/* Same with plain minimax (no pruning), for small depths. */
int32_t  rc_minimax(const rc_pos *p, int depth, uint64_t *nodes);
/* Same as rc_search with the move ordering reversed (worst first). */
int32_t  rc_search_reversed(const rc_pos *p, int depth, uint64_t *nodes);
/* Multi-threaded batch search (rc_search / rc_search_id per position). */
void     rc_search_batch(const rc_pos *p, int n, int depth, int32_t *scores,
                         uint64_t *nodes, int threads);
void     rc_search_id_batch(const rc_pos *p, int n, int depth, int32_t *scores,
                            uint64_t *nodes, int threads);

/* ---- seeded workload generation ---- */
uint64_t rc_splitmix64(uint64_t *state);
/* Position i: attempt a = 0,1,..: rng state = seed ^ i*0x9E3779B97F4A7C15 ^
 * a*0xD1B54A32D192ED03; plies L = minply + r%(maxply-minply+1); each ply plays
 * move r%nlegal in canonical order; the attempt is rejected if a side has no
 * legal move before or at the end, or if the half-move clock reaches 100. */
int      rc_gen_playouts(uint64_t seed, int n, int minply, int maxply, rc_pos *out);

#ifdef __cplusplus
}
#endif
#endif

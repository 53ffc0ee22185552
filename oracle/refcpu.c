/*
 * refcpu.c -- CPU ORACLE (test infrastructure only; see refcpu.h).
 *
 * Clean-room C restatement of the reference semantics.  Written from the
 * reference's behaviour, not translated from its code: the square-wise
 * generator below is a scalar loop over the 64 "lanes" of the reference
 * work-group, so that every per-lane rule (pins, x-ray king attacks, the
 * e.p. make/test/unmake, castling flags) is applied exactly as the kernels
 * apply it.  Shift counts are masked to 6 bits everywhere (quirk Q4: the
 * OpenCL kernels rely on that masking; in C it would be undefined).
 */
#include "refcpu.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef int32_t s32;
typedef uint8_t u8;

/* Q4: OpenCL C takes shift counts modulo the operand width. */
static inline u64 shl(u64 x, int n) { return x << (n & 63); }
static inline u64 shr(u64 x, int n) { return x >> (n & 63); }
#define SETMASK(sq) shl(1ULL, (int)(sq))

static inline int popc(u64 x) { return __builtin_popcountll(x); }
/* first1 = popcount((x&-x)-1)  (src/zetaperft.cl:330-333) */
static inline int first1(u64 x) { return popc((x & (0 - x)) - 1); }
static inline int popfirst1(u64 *a) { u64 b = *a; *a &= *a - 1; return popc((b & (0 - b)) - 1); }

/* piece types, src/types.h:94-100 */
enum { PNONE = 0, PAWN = 1, KNIGHT = 2, KING = 3, BISHOP = 4, ROOK = 5, QUEEN = 6 };
#define WHITE 0
#define BLACK 1
#define MAKEPIECE(p, c) ((u32)(((p) << 1) | (c)))
#define GETCOLOR(p) ((p) & 1)
#define GETPTYPE(p) (((p) >> 1) & 7)
/* move packing, src/types.h:126-141 */
#define GETSQFROM(m) ((m) & 0x3F)
#define GETSQTO(m) (((m) >> 6) & 0x3F)
#define GETSQCPT(m) (((m) >> 12) & 0x3F)
#define GETPFROM(m) (((m) >> 18) & 0xF)
#define GETPTO(m) (((m) >> 22) & 0xF)
#define GETPCPT(m) (((m) >> 26) & 0xF)
#define MAKEMOVE(f, t, c, pf, pt, pc) \
  ((u32)(f) | ((u32)(t) << 6) | ((u32)(c) << 12) | ((u32)(pf) << 18) | ((u32)(pt) << 22) | ((u32)(pc) << 26))
#define GETRANK(sq) ((sq) >> 3)
#define GETFILE(sq) ((sq) & 7)
#define GETRRANK(sq, c) ((c) ? (((sq) >> 3) ^ 7) : ((sq) >> 3))
#define NULLMOVE 0x41u

/* castle masks on the moved-flags word, src/types.h:113-124 (SMCRALL holds d8, not e8) */
#define SMCRALL 0x8900000000000091ULL
#define SMCRWHITEQ 0x0000000000000011ULL
#define SMCRWHITEK 0x0000000000000090ULL
#define SMCRBLACKQ 0x1100000000000000ULL
#define SMCRBLACKK 0x9000000000000000ULL
#define RANKS45 0x000000FFFF000000ULL
#define NOTA 0xFEFEFEFEFEFEFEFEULL
#define NOTH 0x7F7F7F7F7F7F7F7FULL

/* Zobrist keys of the rotate-left scheme: 12 piece keys, 4 castle keys, 1 e.p.
 * key (src/zobrist.h:28-35 == src/zeta.cl:258-265).  The values must be the
 * reference's so that hashes handed across the boundary stay comparable. */
static const u64 ZK[18] = {
  0x9D39247E33776D41ULL, 0x2AF7398005AAA5C7ULL, 0x44DB015024623547ULL, 0x9C15F73E62A76AE2ULL,
  0x75834465489C0C89ULL, 0x3290AC3A203001BFULL, 0x0FBBAD1F61042279ULL, 0xE83A908FF2FB60CAULL,
  0x0D7E765D58755C10ULL, 0x1A083822CEAFE02DULL, 0x9605D5F0E25EC3B0ULL, 0xD021FF5CD13A2ED5ULL,
  0x40BDF15D4A672E32ULL, 0x011355146FD56395ULL, 0x5DB4832046F3D9E5ULL, 0x239F8B2D7FF719CCULL,
  0x05D1A1AE85B49AA1ULL, 0x679F848F6E8FC971ULL};

/* Evaluation constants (src/zeta.cl:272-359): material, square control and
 * piece-square values, all in "black view" (white indexes sq^56). */
static const s32 EVAL_VALUE[7] = {0, 100, 400, 0, 400, 600, 1200};
static const s32 EVAL_CONTROL[64] = {
  0, 0, 5, 5, 5, 5, 0, 0,    5, 0, 5, 5, 5, 5, 0, 5,
  0, 0, 10, 5, 5, 10, 0, 0,  0, 5, 5, 10, 10, 5, 5, 0,
  0, 5, 5, 10, 10, 5, 5, 0,  0, 0, 10, 5, 5, 10, 0, 0,
  0, 0, 5, 5, 5, 5, 0, 0,    0, 0, 5, 5, 5, 5, 0, 0};
static const s32 EVAL_PST[7 * 64] = {
  /* none */
  0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
  0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
  /* pawn */
  0, 0, 0, 0, 0, 0, 0, 0,        50, 50, 50, 50, 50, 50, 50, 50,
  30, 30, 30, 30, 30, 30, 30, 30, 5, 5, 5, 10, 10, 5, 5, 5,
  3, 3, 3, 8, 8, 3, 3, 3,        2, 2, 2, 2, 2, 2, 2, 2,
  0, 0, 0, -5, -5, 0, 0, 0,      0, 0, 0, 0, 0, 0, 0, 0,
  /* knight */
  -50, -40, -30, -30, -30, -30, -40, -50, -40, -20, 0, 0, 0, 0, -20, -40,
  -30, 0, 10, 15, 15, 10, 0, -30,         -30, 5, 15, 20, 20, 15, 5, -30,
  -30, 0, 15, 20, 20, 15, 0, -30,         -30, 5, 10, 15, 15, 10, 5, -30,
  -40, -20, 0, 5, 5, 0, -20, -40,         -50, -40, -30, -30, -30, -30, -40, -50,
  /* king */
  -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, -1,
  -1, 0, 0, 0, 0, 0, 0, -1,       -1, 0, 0, 0, 0, 0, 0, -1,
  -1, 0, 0, 0, 0, 0, 0, -1,       -1, 0, 0, 0, 0, 0, 0, -1,
  -1, 0, 0, 0, 0, 0, 0, -1,       -1, -1, -1, -1, -1, -1, -1, -1,
  /* bishop */
  -20, -10, -10, -10, -10, -10, -10, -20, -10, 0, 0, 0, 0, 0, 0, -10,
  -10, 0, 5, 10, 10, 5, 0, -10,           -10, 5, 5, 10, 10, 5, 5, -10,
  -10, 0, 10, 10, 10, 10, 0, -10,         -10, 10, 10, 10, 10, 10, 10, -10,
  -10, 5, 0, 0, 0, 0, 5, -10,             -20, -10, -10, -10, -10, -10, -10, -20,
  /* rook */
  0, 0, 0, 0, 0, 0, 0, 0,         5, 10, 10, 10, 10, 10, 10, 5,
  -5, 0, 0, 0, 0, 0, 0, -5,       -5, 0, 0, 0, 0, 0, 0, -5,
  -5, 0, 0, 0, 0, 0, 0, -5,       -5, 0, 0, 0, 0, 0, 0, -5,
  -5, 0, 0, 0, 0, 0, 0, -5,       0, 0, 0, 5, 5, 0, 0, 0,
  /* queen */
  -20, -10, -10, -5, -5, -10, -10, -20, -10, 0, 0, 0, 0, 0, 0, -10,
  -10, 0, 5, 5, 5, 5, 0, -10,             -5, 0, 5, 5, 5, 5, 0, -5,
  -5, 0, 5, 5, 5, 5, 0, -5,               -10, 0, 5, 5, 5, 5, 0, -10,
  -10, 0, 0, 0, 0, 0, 0, -10,             -20, -10, -10, -5, -5, -10, -10, -20};

/* ------------------------------------------------------------------ tables */
/* Generated from their definitions (not transcribed); tests pin them with
 * invariants and hand-checked entries. */
static u64 AT[7 * 64];   /* AttackTables: wpawn, bpawn, knight, king, bishop, rook, queen */
static u64 PP[2 * 64];   /* AttackTablesPawnPushes: single + double from the start rank */
static u64 IB[64 * 64];  /* bbInBetween: squares strictly between two aligned squares */
static u64 LN[64 * 64];  /* bbLine: whole board line through two aligned squares */
static int g_inited = 0;

static int onboard(int f, int r) { return f >= 0 && f < 8 && r >= 0 && r < 8; }

static u64 ray_empty(int sq, int df, int dr) {
  u64 bb = 0;
  int f = GETFILE(sq) + df, r = GETRANK(sq) + dr;
  while (onboard(f, r)) { bb |= 1ULL << (r * 8 + f); f += df; r += dr; }
  return bb;
}

void rc_init(void) {
  if (g_inited) return;
  static const int KN[8][2] = {{1, 2}, {2, 1}, {2, -1}, {1, -2}, {-1, -2}, {-2, -1}, {-2, 1}, {-1, 2}};
  for (int sq = 0; sq < 64; sq++) {
    int f = GETFILE(sq), r = GETRANK(sq);
    u64 wp = 0, bp = 0, kn = 0, kg = 0;
    if (onboard(f - 1, r + 1)) wp |= 1ULL << ((r + 1) * 8 + f - 1);
    if (onboard(f + 1, r + 1)) wp |= 1ULL << ((r + 1) * 8 + f + 1);
    if (onboard(f - 1, r - 1)) bp |= 1ULL << ((r - 1) * 8 + f - 1);
    if (onboard(f + 1, r - 1)) bp |= 1ULL << ((r - 1) * 8 + f + 1);
    for (int k = 0; k < 8; k++)
      if (onboard(f + KN[k][0], r + KN[k][1])) kn |= 1ULL << ((r + KN[k][1]) * 8 + f + KN[k][0]);
    for (int df = -1; df <= 1; df++)
      for (int dr = -1; dr <= 1; dr++)
        if ((df || dr) && onboard(f + df, r + dr)) kg |= 1ULL << ((r + dr) * 8 + f + df);
    u64 bi = ray_empty(sq, 1, 1) | ray_empty(sq, -1, 1) | ray_empty(sq, 1, -1) | ray_empty(sq, -1, -1);
    u64 ro = ray_empty(sq, 1, 0) | ray_empty(sq, -1, 0) | ray_empty(sq, 0, 1) | ray_empty(sq, 0, -1);
    AT[0 * 64 + sq] = wp; AT[1 * 64 + sq] = bp; AT[2 * 64 + sq] = kn; AT[3 * 64 + sq] = kg;
    AT[4 * 64 + sq] = bi; AT[5 * 64 + sq] = ro; AT[6 * 64 + sq] = bi | ro;
    u64 pw = 0, pb = 0;
    if (r < 7) pw |= 1ULL << (sq + 8);
    if (r == 1) pw |= 1ULL << (sq + 16);
    if (r > 0) pb |= 1ULL << (sq - 8);
    if (r == 6) pb |= 1ULL << (sq - 16);
    PP[sq] = pw; PP[64 + sq] = pb;
  }
  static const int DIRS[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
  for (int a = 0; a < 64; a++)
    for (int b = 0; b < 64; b++) {
      IB[a * 64 + b] = 0; LN[a * 64 + b] = 0;
      if (a == b) continue;
      for (int d = 0; d < 8; d++) {
        u64 ray = ray_empty(a, DIRS[d][0], DIRS[d][1]);
        if (ray & (1ULL << b)) {
          /* squares from a toward b, excluding both ends */
          u64 btw = ray & ~ray_empty(b, DIRS[d][0], DIRS[d][1]) & ~(1ULL << b);
          IB[a * 64 + b] = btw;
          LN[a * 64 + b] = ray | ray_empty(a, -DIRS[d][0], -DIRS[d][1]) | (1ULL << a);
        }
      }
    }
  g_inited = 1;
}

/* constants as the oracle holds them, for the test that pins them to numbers
 * extracted from the reference: which 0 EVAL_VALUE, 1 EVAL_CONTROL,
 * 2 EVAL_PST (src/zeta.cl:272-359); keys src/zobrist.h:28-35; masks
 * src/types.h:114-124 in the order SMCRALL, WHITEQ, WHITEK, BLACKQ, BLACKK */
int32_t rc_eval_const(int which, int idx) {
  if (which == 0) return (idx >= 0 && idx < 7) ? EVAL_VALUE[idx] : 0;
  if (which == 1) return (idx >= 0 && idx < 64) ? EVAL_CONTROL[idx] : 0;
  return (idx >= 0 && idx < 7 * 64) ? EVAL_PST[idx] : 0;
}
uint64_t rc_zobrist_key(int i) { return (i >= 0 && i < 18) ? ZK[i] : 0; }
uint64_t rc_castle_mask(int i) {
  static const u64 m[5] = {SMCRALL, SMCRWHITEQ, SMCRWHITEK, SMCRBLACKQ, SMCRBLACKK};
  return (i >= 0 && i < 5) ? m[i] : 0;
}
uint64_t rc_attack_table(int idx) { rc_init(); return AT[idx]; }
uint64_t rc_pawnpush_table(int idx) { rc_init(); return PP[idx]; }
uint64_t rc_inbetween(int a, int b) { rc_init(); return IB[a * 64 + b]; }
uint64_t rc_line(int a, int b) { rc_init(); return LN[a * 64 + b]; }

/* Closed-form squares-in-between (src/bit.c:49-66), kept as an independent
 * check of the generated bbInBetween table. */
uint64_t rc_sqinbetween(int sq1, int sq2) {
  const u64 m1 = 0xFFFFFFFFFFFFFFFFULL, a2a7 = 0x0001010101010100ULL;
  const u64 b2g7 = 0x0040201008040200ULL, h1b7 = 0x0002040810204080ULL;
  u64 btwn = shl(m1, sq1) ^ shl(m1, sq2);
  u64 file = (u64)((sq2 & 7) - (sq1 & 7));
  u64 rank = (u64)(((sq2 | 7) - sq1)) >> 3;
  u64 line = ((file & 7) - 1) & a2a7;
  line += 2 * (((rank & 7) - 1) >> 58);
  line += (((rank - file) & 15) - 1) & b2g7;
  line += (((rank + file) & 15) - 1) & h1b7;
  line *= btwn & (0 - btwn);
  return line & btwn;
}

/* ---------------------------------------------------------- board helpers */
static inline u32 getpiece(const u64 *b, int sq) {
  return (u32)((shr(b[0], sq) & 1) | ((shr(b[1], sq) & 1) << 1) | ((shr(b[2], sq) & 1) << 2) |
               ((shr(b[3], sq) & 1) << 3));
}
static inline void setpiece(u64 *b, int sq, u32 p) {
  b[0] |= shl((u64)(p & 1), sq);
  b[1] |= shl((u64)((p >> 1) & 1), sq);
  b[2] |= shl((u64)((p >> 2) & 1), sq);
  b[3] |= shl((u64)((p >> 3) & 1), sq);
}
static inline void clearmask(u64 *b, u64 keep) { b[0] &= keep; b[1] &= keep; b[2] &= keep; b[3] &= keep; }

/* dumb7fill, one direction, as in src/zetaperft.cl:1068-1160.
 * dir: 0 <<9, 1 <<1, 2 <<7, 3 <<8, 4 >>9, 5 >>1, 6 >>7, 7 >>8 */
static const int DSH[8] = {9, 1, 7, 8, -9, -1, -7, -8};
static const u64 DMASK[8] = {NOTA, NOTA, NOTH, ~0ULL, NOTH, NOTH, NOTA, ~0ULL};
static inline u64 dsh(u64 x, int d) { return DSH[d] > 0 ? (x << DSH[d]) : (x >> -DSH[d]); }
static inline u64 fill(u64 gen, u64 pro, int d) {
  pro &= DMASK[d];
  u64 tmp = gen;
  for (int i = 0; i < 6; i++) { gen = dsh(gen, d) & pro; tmp |= gen; }
  return dsh(tmp, d) & DMASK[d];
}

/* domovequick/undomovequick (src/zetaperft.cl:354-411) */
static void domovequick(u64 *b, u32 move) {
  if (move == 0) return;
  int f = GETSQFROM(move), t = GETSQTO(move), c = GETSQCPT(move);
  clearmask(b, ~SETMASK(f) & ~SETMASK(c) & ~SETMASK(t));
  setpiece(b, t, GETPTO(move));
}
/* kernel domove (src/zetaperft.cl:413-484): pieces only, with the castle rook */
static void domove4(u64 *b, u32 move) {
  if (move == 0 || move == NULLMOVE) return;
  int f = GETSQFROM(move), t = GETSQTO(move), c = GETSQCPT(move);
  u32 pf = GETPFROM(move);
  u64 keep = ~SETMASK(f) & ~SETMASK(c) & ~SETMASK(t);
  int qside = GETPTYPE(pf) == KING && f - t == 2;
  int kside = GETPTYPE(pf) == KING && t - f == 2;
  if (qside) keep &= ~SETMASK(f - 4);
  if (kside) keep &= ~SETMASK(f + 3);
  clearmask(b, keep);
  setpiece(b, t, GETPTO(move));
  if (qside) setpiece(b, t + 1, MAKEPIECE(ROOK, GETCOLOR(pf)));
  if (kside) setpiece(b, t - 1, MAKEPIECE(ROOK, GETCOLOR(pf)));
}

static int getkingsq(const u64 *b, int side) {
  u64 t = side ? b[0] : (b[0] ^ (b[1] | b[2] | b[3]));
  t &= b[1] & b[2] & ~b[3];
  return first1(t);
}

/* squareunderattack, superpiece form (src/zetaperft.cl:649-797): is sq
 * attacked by side `stm`? */
static int squareunderattack(const u64 *b, int stm, int sq) {
  u64 blockers = b[1] | b[2] | b[3];
  u64 me = stm ? b[0] : (b[0] ^ blockers);
  u64 moves = 0;
  const int rd[4] = {1, 3, 5, 7}, bd[4] = {0, 2, 4, 6};
  for (int i = 0; i < 4; i++) moves |= fill(SETMASK(sq), ~blockers, rd[i]);
  if (moves & ((me & (b[1] & ~b[2] & b[3])) | (me & (~b[1] & b[2] & b[3])))) return 1;
  moves = 0;
  for (int i = 0; i < 4; i++) moves |= fill(SETMASK(sq), ~blockers, bd[i]);
  if (moves & ((me & (~b[1] & ~b[2] & b[3])) | (me & (~b[1] & b[2] & b[3])))) return 1;
  if (AT[128 + sq] & me & (~b[1] & b[2] & ~b[3])) return 1;
  if (AT[(!stm) * 64 + sq] & me & (b[1] & ~b[2] & ~b[3])) return 1;
  if (AT[192 + sq] & me & (b[1] & b[2] & ~b[3])) return 1;
  return 0;
}

/* ------------------------------------------------------------- movegen x64 */
int rc_genmoves(const uint64_t *board, uint64_t cr, int stm, int qs, uint64_t out[64], int *incheck) {
  rc_init();
  u64 b[4] = {board[0], board[1], board[2], board[3]};
  const u64 blockers = b[1] | b[2] | b[3];
  const u64 me = stm ? b[0] : (b[0] ^ blockers);
  const u64 opp = (!stm) ? b[0] : (b[0] ^ blockers);
  const u64 kingbb = me & b[1] & b[2] & ~b[3];
  if (!kingbb) { if (incheck) *incheck = 0; memset(out, 0, 64 * sizeof(u64)); return 0; }
  const int sqking = first1(kingbb);

  /* superking rays and pinned pieces (src/zetaperft.cl:936-1061) */
  u64 pinned = 0, work = 0;
  const int rd[4] = {1, 3, 5, 7}, bd[4] = {0, 2, 4, 6};
  for (int i = 0; i < 4; i++) work |= fill(SETMASK(sqking), ~0ULL, rd[i]);
  work &= (opp & (b[1] & ~b[2] & b[3])) | (opp & (~b[1] & b[2] & b[3]));
  while (work) {
    int sq = popfirst1(&work);
    u64 t = IB[sq * 64 + sqking] & blockers;
    if (popc(t) == 1) pinned |= t;
  }
  work = 0;
  for (int i = 0; i < 4; i++) work |= fill(SETMASK(sqking), ~0ULL, bd[i]);
  work &= (opp & (~b[1] & ~b[2] & b[3])) | (opp & (~b[1] & b[2] & b[3]));
  while (work) {
    int sq = popfirst1(&work);
    u64 t = IB[sq * 64 + sqking] & blockers;
    if (popc(t) == 1) pinned |= t;
  }

  /* per-lane targets, opponent attacks, checkers (src/zetaperft.cl:1063-1196) */
  u64 mv[64];
  u64 attacks = 0, checkers = 0;
  int sqchecker = 0;
  for (int lid = 0; lid < 64; lid++) {
    u32 pfrom = getpiece(b, lid);
    int color = GETCOLOR(pfrom), type = GETPTYPE(pfrom);
    mv[lid] = 0;
    if (type == PNONE) continue; /* empty lanes produce nothing: gen is empty */
    u64 w = 0;
    if (type == KNIGHT) w = ~0ULL;
    else if (type == KING) w = AT[192 + lid]; /* the fill always reaches the 8 neighbours */
    else {
      u64 pro = (color == stm) ? ~blockers : ~(blockers ^ SETMASK(sqking));
      u64 gen = blockers & SETMASK(lid);
      for (int d = 0; d < 8; d++) w |= fill(gen, pro, d);
    }
    int n = (type == PAWN) ? color : type;
    u64 mask = AT[n * 64 + lid];
    u64 moves = (color == stm) ? (mask & w & opp) : (mask & w);
    if (color != stm) attacks |= moves;
    if (moves & SETMASK(sqking)) { sqchecker = lid; checkers |= SETMASK(lid); }
    u64 qmask = (type == PAWN) ? PP[stm * 64 + lid] : mask;
    moves |= (color == stm && !qs) ? (qmask & w & ~blockers) : 0;
    mv[lid] = moves;
  }
  const int rootkic = checkers != 0;
  const int nchk = popc(checkers);

  /* e.p. square from the moved flags (src/zetaperft.cl:1224-1230) */
  u64 t = ~cr & RANKS45;
  int sqep = t ? first1(t) : 0;
  u64 eppawns = me & (b[1] & ~b[2] & ~b[3]);
  eppawns &= stm ? 0xFF000000ULL : 0xFF00000000ULL;
  eppawns &= (SETMASK(sqep + 1) | SETMASK(sqep - 1));

  int count = 0;
  for (int lid = 0; lid < 64; lid++) {
    u32 pfrom = getpiece(b, lid);
    int color = GETCOLOR(pfrom), type = GETPTYPE(pfrom);
    u64 moves = (color == stm) ? mv[lid] : 0;
    if (nchk >= 2 && type != KING) moves = 0;
    if (SETMASK(lid) & pinned) moves &= LN[lid * 64 + sqking];
    if (type == KING) moves &= ~attacks;
    if (nchk == 1 && type != KING) moves &= (IB[sqchecker * 64 + sqking] | checkers);
    /* en passant with full make/test/unmake (src/zetaperft.cl:1232-1248) */
    if (eppawns & SETMASK(lid)) {
      int sqto = (u8)(stm ? sqep - 8 : sqep + 8);
      u32 pcpt = getpiece(b, sqep);
      u32 move = MAKEMOVE(lid, sqto, sqep, pfrom, pfrom, pcpt);
      u64 bb[4] = {b[0], b[1], b[2], b[3]};
      domovequick(bb, move);
      int kic = squareunderattack(bb, !stm, getkingsq(bb, stm));
      if (!kic) moves |= SETMASK(sqto);
    }
    /* castling (src/zetaperft.cl:1249-1297; quirk Q2 precondition kept) */
    if (lid == sqking && !qs && (cr & SMCRALL)) {
      int okq = stm ? ((~cr & SMCRBLACKQ) == SMCRBLACKQ) : ((~cr & SMCRWHITEQ) == SMCRWHITEQ);
      int okk = stm ? ((~cr & SMCRBLACKK) == SMCRBLACKK) : ((~cr & SMCRWHITEK) == SMCRWHITEK);
      u32 rook = MAKEPIECE(ROOK, GETCOLOR(pfrom));
      if (okq && getpiece(b, lid - 4) == rook &&
          !(blockers & (SETMASK(lid - 1) | SETMASK(lid - 2) | SETMASK(lid - 3))) &&
          !(attacks & (SETMASK(lid) | SETMASK(lid - 1) | SETMASK(lid - 2))))
        moves |= SETMASK(lid - 2);
      if (okk && getpiece(b, lid + 3) == rook &&
          !(blockers & (SETMASK(lid + 1) | SETMASK(lid + 2))) &&
          !(attacks & (SETMASK(lid) | SETMASK(lid + 1) | SETMASK(lid + 2))))
        moves |= SETMASK(lid + 2);
    }
    out[lid] = moves;
    count += popc(moves);
  }
  if (incheck) *incheck = rootkic;
  return count;
}

/* movepicker packing of lid->sqto (src/zetaperft.cl:1363-1386, src/zeta.cl:2036-2058):
 * e.p. capture square, captured piece, queen-only promotion (quirk Q1). */
uint32_t rc_makemove_code(const uint64_t *b, int stm, int lid, int sqto) {
  u32 pfrom = getpiece(b, lid);
  int sqcpt = sqto;
  u32 pcpt = getpiece(b, sqcpt);
  if (GETPTYPE(pfrom) == PAWN && stm && lid - sqto != 8 && lid - sqto != 16 && pcpt == PNONE) sqcpt = sqto + 8;
  if (GETPTYPE(pfrom) == PAWN && !stm && sqto - lid != 8 && sqto - lid != 16 && pcpt == PNONE) sqcpt = sqto - 8;
  pcpt = getpiece(b, sqcpt);
  u32 pto = (GETPTYPE(pfrom) == PAWN && GETRRANK(sqto, stm) == 7) ? MAKEPIECE(QUEEN, GETCOLOR(pfrom)) : pfrom;
  return MAKEMOVE(lid, sqto, sqcpt, pfrom, pto, pcpt);
}

/* moved-flags update of the perft move-up (src/zetaperft.cl:1447-1465): marks
 * from/to/capture squares and the castle rook origin, re-arms ranks 4/5 and
 * clears the square of a pawn that just double-pushed. */
static u64 cr_after_perft(u64 cr, u32 move) {
  int f = GETSQFROM(move), t = GETSQTO(move);
  u32 pf = GETPFROM(move);
  cr |= SETMASK(f) | SETMASK(t) | SETMASK(GETSQCPT(move));
  if (GETPTYPE(pf) == KING && f - t == 2) cr |= SETMASK(f - 4);
  if (GETPTYPE(pf) == KING && t - f == 2) cr |= SETMASK(f + 3);
  int sqep = (GETPTYPE(pf) == PAWN && (int)GETRRANK(t, GETCOLOR(pf)) - (int)GETRRANK(f, GETCOLOR(pf)) == 2) ? t : 0;
  cr |= RANKS45;
  if (sqep) cr &= ~SETMASK(sqep);
  return cr;
}
/* the alpha-beta kernel's variant (src/zeta.cl:2333-2345): no castle rook flag (quirk Q6) */
static u64 cr_after_ab(u64 cr, u32 move) {
  int f = GETSQFROM(move), t = GETSQTO(move);
  u32 pf = GETPFROM(move);
  cr |= SETMASK(f) | SETMASK(t) | SETMASK(GETSQCPT(move));
  int sqep = (GETPTYPE(pf) == PAWN && (int)GETRRANK(t, GETCOLOR(pf)) - (int)GETRRANK(f, GETCOLOR(pf)) == 2) ? t : 0;
  cr |= RANKS45;
  if (sqep) cr &= ~SETMASK(sqep);
  return cr;
}

/* ------------------------------------------------------------------ hashing */
static inline u64 rotl(u64 z, int sq) { return shl(z, sq) | shr(z, 64 - sq); }

static u64 piece_hash(const u64 *b) {
  u64 h = 0;
  for (int side = WHITE; side <= BLACK; side++) {
    u64 w = (side == BLACK) ? b[0] : (b[0] ^ (b[1] | b[2] | b[3]));
    while (w) {
      int sq = popfirst1(&w);
      u32 p = getpiece(b, sq);
      h ^= rotl(ZK[GETCOLOR(p) * 6 + GETPTYPE(p) - 1], sq);
    }
  }
  return h;
}
static u64 castle_hash(u64 cr) {
  u64 h = 0;
  if (~cr & SMCRWHITEQ) h ^= ZK[12];
  if (~cr & SMCRWHITEK) h ^= ZK[13];
  if (~cr & SMCRBLACKQ) h ^= ZK[14];
  if (~cr & SMCRBLACKK) h ^= ZK[15];
  return h;
}

/* host computehash (src/bitboard.c:461-508) on b[0..4] */
uint64_t rc_computehash(const uint64_t *b, int stm) {
  rc_init();
  u64 h = piece_hash(b) ^ castle_hash(b[4]);
  u64 w = ~b[4] & RANKS45;
  int sqep = w ? first1(w) : 0;
  if (sqep) h ^= rotl(ZK[16], GETFILE(sqep));
  if (stm) h ^= 1ULL;
  return h;
}

/* incremental hash of the alpha-beta move-up (src/zeta.cl:2297-2359) for a
 * child board (pieces already moved), child moved flags and child stm. */
static u64 kernel_hash(const u64 *child, u64 childcr, int childstm, u32 move) {
  u64 h = piece_hash(child);
  u32 pf = GETPFROM(move);
  int f = GETSQFROM(move), t = GETSQTO(move);
  if (GETPTYPE(pf) == PAWN && (int)GETRRANK(t, GETCOLOR(pf)) - (int)GETRRANK(f, GETCOLOR(pf)) == 2)
    h ^= rotl(ZK[16], GETFILE(t));
  if (childstm) h ^= 1ULL;
  h ^= castle_hash(childcr);
  if (move == NULLMOVE) h ^= 1ULL;
  return h;
}

/* host domove (src/bitboard.c:583-693); also flips p->stm. */
void rc_domove_host(rc_pos *p, uint32_t move) {
  rc_init();
  if (move == 0 || move == NULLMOVE) return;
  int f = GETSQFROM(move), t = GETSQTO(move), c = GETSQCPT(move);
  u32 pf = GETPFROM(move), pc = GETPCPT(move);
  u64 hmc = p->b[6] + 1;
  domove4(p->b, move);
  p->b[4] |= SETMASK(f) | SETMASK(t) | SETMASK(c);
  int sqep = (GETPTYPE(pf) == PAWN && (int)GETRRANK(t, GETCOLOR(pf)) - (int)GETRRANK(f, GETCOLOR(pf)) == 2) ? t : 0;
  p->b[4] |= RANKS45;
  if (sqep) p->b[4] &= ~SETMASK(sqep);
  if (GETPTYPE(pf) == KING && f - t == 2) { p->b[4] |= SETMASK(f - 4); hmc = 0; }
  if (GETPTYPE(pf) == KING && t - f == 2) { p->b[4] |= SETMASK(f + 3); hmc = 0; }
  if (GETPTYPE(pf) == PAWN) hmc = 0;
  if (GETPTYPE(pc) != PNONE) hmc = 0;
  p->b[6] = hmc;
  p->b[5] = rc_computehash(p->b, !GETCOLOR(pf));
  p->stm ^= 1;
}

/* ---------------------------------------------------------------- setboard */
/* FEN parsing with the reference's rules (src/io.c:335-532): moved flags start
 * all set, castle letters clear their rook+king squares, the FEN e.p. target
 * (rank 3/6) is cleared but the kernels only read ranks 4/5 (quirk Q3). */
int rc_setboard(const char *fen, rc_pos *out) {
  rc_init();
  char position[1024] = {0}, cstm[1024] = {0}, castle[1024] = {0}, cep[1024] = {0};
  unsigned long long hmc = 0, fendepth = 1;
  sscanf(fen, "%1023s %1023s %1023s %1023s %llu %llu", position, cstm, castle, cep, &hmc, &fendepth);
  static const char map[24] = " PNKBRQ pnkbrq/12345678";
  memset(out, 0, sizeof(*out));
  u64 *b = out->b;
  b[4] = ~0ULL;
  int file = 0, rank = 7;
  size_t i = 0, len = strlen(position);
  while (!(rank <= 0 && file >= 8)) {
    if (i >= len) return -1;
    char ch = position[i++];
    for (int j = 0; j <= 23; j++) {
      if (ch != map[j]) continue;
      if (j == 14) { rank--; file = 0; }
      else if (j >= 15) file += j - 14;
      else if (j >= 7) { setpiece(b, rank * 8 + file, MAKEPIECE(j - 7, BLACK)); file++; }
      else { setpiece(b, rank * 8 + file, MAKEPIECE(j, WHITE)); file++; }
      break;
    }
    if (rank < 0) return -1;
  }
  int stm = (cstm[0] == 'b' || cstm[0] == 'B') ? BLACK : WHITE;
  u64 bbcr = 0;
  if (castle[0] != '-')
    for (const char *q = castle; *q; q++) {
      if (*q == 'Q') bbcr |= SMCRWHITEQ;
      if (*q == 'K') bbcr |= SMCRWHITEK;
      if (*q == 'q') bbcr |= SMCRBLACKQ;
      if (*q == 'k') bbcr |= SMCRBLACKK;
    }
  b[4] ^= bbcr;
  b[6] = hmc;
  int sq = 0;
  if (cep[0] != '-' && cep[0] != '\0' && cep[0] != '\n') sq = (cep[1] - '1') * 8 + (cep[0] - 'a');
  if (sq > 0 && sq < 64) b[4] &= ~SETMASK(sq);
  out->stm = stm;
  b[5] = rc_computehash(b, stm);
  /* isvalid: one king per side (src/bitboard.c:448-459) */
  u64 kings = b[1] & b[2] & ~b[3];
  u64 occ = b[1] | b[2] | b[3];
  if (popc(b[0] & kings) != 1 || popc((b[0] ^ occ) & kings) != 1) return -1;
  return 0;
}

/* ------------------------------------------------------------ legal moves */
int rc_count_legal(const rc_pos *p) {
  u64 mv[64];
  return rc_genmoves(p->b, p->b[4], p->stm, 0, mv, NULL);
}

uint32_t rc_nth_legal(const rc_pos *p, int k) {
  u64 mv[64];
  rc_genmoves(p->b, p->b[4], p->stm, 0, mv, NULL);
  for (int lid = 0; lid < 64; lid++) {
    int c = popc(mv[lid]);
    if (k < c) {
      u64 m = mv[lid];
      while (k-- > 0) m &= m - 1;
      return rc_makemove_code(p->b, p->stm, lid, first1(m));
    }
    k -= c;
  }
  return 0;
}

/* ------------------------------------------------------------------- perft */
/* Recursive restatement of the perft tree walk (src/zetaperft.cl:1313-1476):
 * the order in which moves are taken does not change the leaf count. */
static u64 perft_rec(const u64 *b, u64 cr, int stm, int depth) {
  if (depth == 0) return 1;
  u64 mv[64];
  rc_genmoves(b, cr, stm, 0, mv, NULL);
  u64 sum = 0;
  for (int lid = 0; lid < 64; lid++) {
    u64 m = mv[lid];
    while (m) {
      int sqto = popfirst1(&m);
      u32 move = rc_makemove_code(b, stm, lid, sqto);
      u64 child[4] = {b[0], b[1], b[2], b[3]};
      domove4(child, move);
      sum += perft_rec(child, cr_after_perft(cr, move), !stm, depth - 1);
    }
  }
  return sum;
}

uint64_t rc_perft(const rc_pos *p, int depth) {
  rc_init();
  return perft_rec(p->b, p->b[4], p->stm, depth);
}

/* The same count with bulk counting: a depth-1 position contributes its
 * number of legal moves, which are not made (as the device's leaf level). */
static u64 perft_bulk_rec(const u64 *b, u64 cr, int stm, int depth) {
  if (depth == 0) return 1;
  u64 mv[64];
  const int n = rc_genmoves(b, cr, stm, 0, mv, NULL);
  if (depth == 1) return (u64)n;
  u64 sum = 0;
  for (int lid = 0; lid < 64; lid++) {
    u64 m = mv[lid];
    while (m) {
      int sqto = popfirst1(&m);
      u32 move = rc_makemove_code(b, stm, lid, sqto);
      u64 child[4] = {b[0], b[1], b[2], b[3]};
      domove4(child, move);
      sum += perft_bulk_rec(child, cr_after_perft(cr, move), !stm, depth - 1);
    }
  }
  return sum;
}

uint64_t rc_perft_bulk(const rc_pos *p, int depth) {
  rc_init();
  return perft_bulk_rec(p->b, p->b[4], p->stm, depth);
}

int rc_perft_divide(const rc_pos *p, int depth, uint32_t *moves, uint64_t *counts) {
  rc_init();
  u64 mv[64];
  rc_genmoves(p->b, p->b[4], p->stm, 0, mv, NULL);
  int n = 0;
  for (int lid = 0; lid < 64; lid++) {
    u64 m = mv[lid];
    while (m) {
      int sqto = popfirst1(&m);
      u32 move = rc_makemove_code(p->b, p->stm, lid, sqto);
      u64 child[4] = {p->b[0], p->b[1], p->b[2], p->b[3]};
      domove4(child, move);
      moves[n] = move;
      counts[n] = depth > 1 ? perft_rec(child, cr_after_perft(p->b[4], move), !p->stm, depth - 1) : 1;
      n++;
    }
  }
  return n;
}

/* ------------------------------------------------------------------- eval */
/* x64 evaluation summed over the 64 lanes (src/zeta.cl:1482-1531). */
int32_t rc_eval(const uint64_t *board) {
  rc_init();
  const u64 *b = board;
  const u64 blockers = b[1] | b[2] | b[3];
  const u64 pawns = b[1] & ~b[2] & ~b[3];
  s32 total = 0;
  u64 occ = blockers;
  while (occ) {
    int lid = popfirst1(&occ);
    u32 p = getpiece(b, lid);
    int color = GETCOLOR(p), type = GETPTYPE(p);
    u64 me = color ? b[0] : (b[0] ^ blockers);
    u64 opp = color ? (b[0] ^ blockers) : b[0];
    int sqf = color ? lid : (lid ^ 56);
    s32 s = 10 + EVAL_VALUE[type] + EVAL_PST[type * 64 + sqf] + EVAL_CONTROL[sqf];
    if (type == PAWN && color == WHITE) {
      if (GETRANK(lid) < 7 && (opp & SETMASK(lid + 8))) s -= 15;
      if (GETFILE(lid) < 7 && (pawns & me & SETMASK(lid - 7))) s += 10;
      if (GETFILE(lid) > 0 && (pawns & me & SETMASK(lid - 9))) s += 10;
      for (u8 sq = (u8)(lid - 8); sq > 7; sq = (u8)(sq - 8))
        if (pawns & me & SETMASK(sq)) s -= 30;
    }
    if (type == PAWN && color == BLACK) {
      if (GETRANK(lid) > 0 && (opp & SETMASK(lid - 8))) s -= 15;
      if (GETFILE(lid) > 0 && (pawns & me & SETMASK(lid + 7))) s += 10;
      if (GETFILE(lid) < 7 && (pawns & me & SETMASK(lid + 9))) s += 10;
      for (u8 sq = (u8)(lid + 8); sq < 56; sq = (u8)(sq + 8))
        if (pawns & me & SETMASK(sq)) s -= 30;
    }
    total += color ? -s : s;
  }
  const u64 bishops = ~b[1] & ~b[2] & b[3];
  total -= (popc(b[0] & bishops) == 2) ? 25 : 0;
  total += (popc((b[0] ^ blockers) & bishops) == 2) ? 25 : 0;
  return total;
}

/* ----------------------------------------------------------------- search */
typedef struct {
  u64 hist[RC_MAXGAMEPLY + RC_MAXPLY + 8];
  int ply_init;
  u64 nodes;
  s32 bestscore;
  u32 bestmove;
  int minimax;
  int reverse; /* full-width nodes take moves worst-first (order-independence checks) */
  /* the reference's quiet-move heuristics (src/zeta.cl:1916-1933, 2011-2082):
   * one killer per ply and a counter move per previous move's squares; kept
   * across the iterations of one root */
  u32 killer[RC_MAXPLY + 2];
  u32 counter[64 * 64];
} sctx;

/* move ordering score of the reference picker (src/zeta.cl:2059-2071); any
 * order gives the same root score.  The search below orders by it, then the
 * counter move (1300) and the killer (1400) over it, highest first and the
 * earliest generated among equals -- the device search's order, so a
 * single-worker device search without a transposition table visits the same
 * tree and makes the same number of moves. */
static s32 order_score(const u64 *b, int stm, u32 move) {
  int lid = GETSQFROM(move), sqto = GETSQTO(move);
  u32 pf = GETPFROM(move), pt = GETPTO(move), pc = GETPCPT(move);
  int t = stm ? sqto : (sqto ^ 56), f = stm ? lid : (lid ^ 56);
  s32 s = EVAL_VALUE[GETPTYPE(pt)] + EVAL_PST[GETPTYPE(pt) * 64 + t] + EVAL_CONTROL[t];
  s -= EVAL_VALUE[GETPTYPE(pf)] + EVAL_PST[GETPTYPE(pf) * 64 + f] + EVAL_CONTROL[f];
  if (GETPTYPE(pc) != PNONE) s = EVAL_VALUE[GETPTYPE(pc)] * 16 - EVAL_VALUE[GETPTYPE(pt)];
  (void)b;
  return s;
}

/* One node of alphabeta_gpu (src/zeta.cl:1319-1723) and its backup
 * (src/zeta.cl:1784-1837) with NMP/LMR/IID/TT off.  sd = ply+1. */
static s32 ab(sctx *c, const u64 *b, u64 cr, u64 hash, u8 hmc, int stm, int depth, s32 alpha, s32 beta,
              int sd, int parent_qs, u32 lastmove) {
  const int ply = sd - 1;
  u64 mv[64];
  int rootkic = 0;
  /* in-check first: the extension decides qs before quiets are generated */
  rc_genmoves(b, cr, stm, 1, mv, &rootkic);
  /* check / queen-promotion extension (src/zeta.cl:1330-1345) */
  if (depth >= 0 && !parent_qs &&
      (rootkic || (GETPTYPE(GETPFROM(lastmove)) == PAWN && GETPTYPE(GETPTO(lastmove)) == QUEEN)))
    depth++;
  const int qs = depth <= 0;
  int movecount = qs ? 0 : rc_genmoves(b, cr, stm, 0, mv, NULL);
  if (qs) { for (int i = 0; i < 64; i++) movecount += popc(mv[i]); }
  /* static eval, side-to-move relative (src/zeta.cl:1554) */
  s32 ev = rc_eval(b);
  s32 score = stm ? -ev : ev;
  if (!qs && rootkic && movecount == 0) score = -RC_INF + ply;
  if (!qs && !rootkic && movecount == 0) score = 0;
  /* in-path repetition, fifty moves, bare kings (src/zeta.cl:1560-1589) */
  if (movecount > 0 && !qs && sd > 1) {
    for (int n = ply + c->ply_init - 2; n >= 0 && n >= ply + c->ply_init - (int)hmc; n -= 2)
      if (hash == c->hist[n]) { movecount = 0; score = 0; break; }
  }
  if (sd > 1 && hmc >= 100) { movecount = 0; score = 0; }
  if (sd > 1 && popc(b[1] | b[2] | b[3]) <= 2) { movecount = 0; score = 0; }
  if (sd >= RC_MAXPLY) movecount = 0;
  if (movecount == 0) return score;
  /* stand pat (src/zeta.cl:1601-1609) */
  if (qs && !rootkic && score >= beta) return score;
  if (qs && !rootkic && score > alpha) alpha = score;
  s32 best_standpat = (qs && !rootkic) ? score : -RC_INF - 1;

  u32 moves[256];
  s32 scores[256];
  unsigned char used[256];
  int n = 0;
  const u32 killer = c->killer[sd];
  const u32 counter = c->counter[GETSQFROM(lastmove) * 64 + GETSQTO(lastmove)];
  for (int lid = 0; lid < 64; lid++) {
    u64 m = mv[lid];
    while (m) {
      int sqto = popfirst1(&m);
      u32 move = rc_makemove_code(b, stm, lid, sqto);
      moves[n] = move;
      scores[n] = order_score(b, stm, move);
      if (!c->minimax && !c->reverse && GETPTYPE(GETPCPT(move)) == PNONE) {
        if (move == killer) scores[n] = 1400;        /* src/zeta.cl:2078-2082 */
        else if (move == counter) scores[n] = 1300;  /* src/zeta.cl:2072-2076 */
      }
      used[n] = 0;
      n++;
    }
  }
  s32 mmbest = -RC_INF - 1;
  for (int i = 0; i < n; i++) {
    int bi = -1;
    for (int j = 0; j < n; j++) {
      if (used[j]) continue;
      if (bi < 0 || ((c->reverse && !qs) ? (scores[j] < scores[bi]) : (scores[j] > scores[bi]))) bi = j;
    }
    used[bi] = 1;
    u32 move = moves[bi];
    u64 child[4] = {b[0], b[1], b[2], b[3]};
    domove4(child, move);
    u64 ccr = cr_after_ab(cr, move);
    u64 chash = kernel_hash(child, ccr, !stm, move);
    u32 pf = GETPFROM(move);
    u8 chmc = (u8)(hmc + 1);
    if (GETPTYPE(pf) == PAWN) chmc = 0;
    if (GETPTYPE(GETPCPT(move)) != PNONE) chmc = 0;
    if (GETPTYPE(pf) == KING && (GETSQFROM(move) - GETSQTO(move) == 2 || GETSQTO(move) - GETSQFROM(move) == 2)) chmc = 0;
    c->nodes++;
    c->hist[ply + 1 + c->ply_init] = chash;
    s32 s;
    if (c->minimax) {
      s = -ab(c, child, ccr, chash, chmc, !stm, depth - 1, -RC_INF, RC_INF, sd + 1, qs, move);
      if (s > mmbest) { mmbest = s; if (sd == 1) { c->bestscore = s; c->bestmove = move; } }
      continue;
    }
    s = -ab(c, child, ccr, chash, chmc, !stm, depth - 1, -beta, -alpha, sd + 1, qs, move);
    if (s > alpha) {
      alpha = s;
      if (sd == 1) { c->bestscore = s; c->bestmove = move; }
    }
    if (alpha >= beta) {
      /* a quiet move failing high outside quiescence becomes the killer of
       * the ply and the counter move of the move before (src/zeta.cl:1916-1933) */
      if (!qs && GETPTYPE(GETPCPT(move)) == PNONE) {
        c->killer[sd] = move;
        c->counter[GETSQFROM(lastmove) * 64 + GETSQTO(lastmove)] = move;
      }
      break;
    }
  }
  if (c->minimax) return (best_standpat > mmbest) ? best_standpat : mmbest;
  return alpha;
}

static s32 search_common(const rc_pos *p, int depth, const uint64_t *hist, int ply_init, uint64_t *nodes,
                         uint32_t *bestmove, int minimax, int reverse, int id) {
  rc_init();
  sctx *c = (sctx *)calloc(1, sizeof(sctx));
  if (!c) return 0;
  if (hist) {
    if (ply_init < 0 || ply_init >= RC_MAXGAMEPLY) ply_init = 0;
    memcpy(c->hist, hist, (size_t)(ply_init + 1) * sizeof(u64));
    c->ply_init = ply_init;
  } else {
    c->hist[0] = p->b[5];
    c->ply_init = 0;
  }
  c->minimax = minimax;
  c->reverse = reverse;
  /* iterative deepening 1..depth (src/search.c:77-198) with the killers and
   * counter moves kept across iterations; or the one iteration at depth */
  for (int d = id ? 1 : depth; d <= depth; d++) {
    c->bestscore = 0; /* DRAWSCORE until a root move raises alpha (src/zeta.cl:986) */
    ab(c, p->b, p->b[4], p->b[5], (u8)p->b[6], p->stm, d, -RC_INF, RC_INF, 1, 0, 0);
    /* mate in n: a mate within this iteration's depth ends the iterations
     * (src/search.c:177-186) */
    const s32 bs = c->bestscore;
    if (id && bs > RC_MATESCORE && bs < RC_INF && d >= RC_INF - bs) break;
    if (id && bs < -RC_MATESCORE && bs > -RC_INF && d >= RC_INF + bs) break;
  }
  s32 sc = c->bestscore;
  if (sc == RC_INF || sc == -RC_INF) sc = 0; /* src/search.c:137 */
  if (nodes) *nodes = c->nodes;
  if (bestmove) *bestmove = c->bestmove;
  free(c);
  return sc;
}

int32_t rc_search(const rc_pos *p, int depth, const uint64_t *hist, int ply_init, uint64_t *nodes,
                  uint32_t *bestmove) {
  return search_common(p, depth, hist, ply_init, nodes, bestmove, 0, 0, 0);
}
int32_t rc_search_id(const rc_pos *p, int depth, const uint64_t *hist, int ply_init, uint64_t *nodes,
                     uint32_t *bestmove) {
  return search_common(p, depth, hist, ply_init, nodes, bestmove, 0, 0, 1);
}
int32_t rc_minimax(const rc_pos *p, int depth, uint64_t *nodes) {
  return search_common(p, depth, NULL, 0, nodes, NULL, 1, 0, 0);
}
int32_t rc_search_reversed(const rc_pos *p, int depth, uint64_t *nodes) {
  return search_common(p, depth, NULL, 0, nodes, NULL, 0, 1, 0);
}

/* ------------------------------------------------------------------ batches */
typedef struct {
  const rc_pos *p;
  int n, depth, kind;
  uint64_t *counts;
  int32_t *scores;
  uint64_t *nodes;
  int next;
  pthread_mutex_t mu;
} batch_job;

static void *batch_worker(void *arg) {
  batch_job *j = (batch_job *)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    int i = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (i >= j->n) break;
    if (j->kind == 0) j->counts[i] = rc_perft(&j->p[i], j->depth);
    else if (j->kind == 3) j->counts[i] = rc_perft_bulk(&j->p[i], j->depth);
    else {
      uint64_t nd = 0;
      j->scores[i] = j->kind == 2 ? rc_search_id(&j->p[i], j->depth, NULL, 0, &nd, NULL)
                                  : rc_search(&j->p[i], j->depth, NULL, 0, &nd, NULL);
      if (j->nodes) j->nodes[i] = nd;
    }
  }
  return NULL;
}

static void run_batch(batch_job *j, int threads) {
  rc_init();
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  pthread_mutex_init(&j->mu, NULL);
  j->next = 0;
  for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, batch_worker, j);
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  pthread_mutex_destroy(&j->mu);
}

void rc_perft_batch(const rc_pos *p, int n, int depth, uint64_t *counts, int threads) {
  batch_job j;
  memset(&j, 0, sizeof(j));
  j.p = p; j.n = n; j.depth = depth; j.kind = 0; j.counts = counts;
  run_batch(&j, threads);
}

void rc_perft_bulk_batch(const rc_pos *p, int n, int depth, uint64_t *counts, int threads) {
  batch_job j;
  memset(&j, 0, sizeof(j));
  j.p = p; j.n = n; j.depth = depth; j.kind = 3; j.counts = counts;
  run_batch(&j, threads);
}

void rc_search_batch(const rc_pos *p, int n, int depth, int32_t *scores, uint64_t *nodes, int threads) {
  batch_job j;
  memset(&j, 0, sizeof(j));
  j.p = p; j.n = n; j.depth = depth; j.kind = 1; j.scores = scores; j.nodes = nodes;
  run_batch(&j, threads);
}

void rc_search_id_batch(const rc_pos *p, int n, int depth, int32_t *scores, uint64_t *nodes, int threads) {
  batch_job j;
  memset(&j, 0, sizeof(j));
  j.p = p; j.n = n; j.depth = depth; j.kind = 2; j.scores = scores; j.nodes = nodes;
  run_batch(&j, threads);
}

/* ---------------------------------------------------------------- playouts */
uint64_t rc_splitmix64(uint64_t *s) {
  u64 z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int rc_gen_playouts(uint64_t seed, int n, int minply, int maxply, rc_pos *out) {
  rc_init();
  rc_pos start;
  rc_setboard("rnbqkbnr/pppppppp/8/8/8/8/PPPPPPPP/RNBQKBNR w KQkq - 0 1", &start);
  for (int i = 0; i < n; i++) {
    for (u64 a = 0;; a++) {
      u64 st = seed ^ ((u64)i * 0x9E3779B97F4A7C15ULL) ^ (a * 0xD1B54A32D192ED03ULL);
      int plies = minply + (int)(rc_splitmix64(&st) % (u64)(maxply - minply + 1));
      rc_pos p = start;
      int ok = 1;
      for (int k = 0; k < plies && ok; k++) {
        int nl = rc_count_legal(&p);
        if (nl == 0 || p.b[6] >= 100) { ok = 0; break; }
        int pick = (int)(rc_splitmix64(&st) % (u64)nl);
        rc_domove_host(&p, rc_nth_legal(&p, pick));
      }
      if (ok && (rc_count_legal(&p) == 0 || p.b[6] >= 100)) ok = 0;
      if (ok) { out[i] = p; break; }
      if (a > 1000) return -1;
    }
  }
  return 0;
}

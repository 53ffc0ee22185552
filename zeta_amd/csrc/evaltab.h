// evaltab.h -- the reference evaluation's constants (src/zeta.cl:272-359):
// material {P 100, N 400, K 0, B 400, R 600, Q 1200}, square-control bonus and
// the piece-square tables (Michniewski's "simplified evaluation"), all in
// black view (white indexes sq^56).  They are data the root scores depend on,
// so they must be the reference's values exactly.
#pragma once
#include <stdint.h>

namespace zc {

static constexpr int32_t EVAL_VALUE[7] = {0, 100, 400, 0, 400, 600, 1200};

static constexpr int32_t EVAL_CONTROL[64] = {
    0, 0, 5, 5, 5, 5, 0, 0,   5, 0, 5, 5, 5, 5, 0, 5,   0, 0, 10, 5, 5, 10, 0, 0, 0, 5, 5, 10, 10, 5, 5, 0,
    0, 5, 5, 10, 10, 5, 5, 0, 0, 0, 10, 5, 5, 10, 0, 0, 0, 0, 5, 5, 5, 5, 0, 0,   0, 0, 5, 5, 5, 5, 0, 0};

static constexpr int32_t EVAL_PST[7 * 64] = {
    // none
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    // pawn
    0, 0, 0, 0, 0, 0, 0, 0, 50, 50, 50, 50, 50, 50, 50, 50, 30, 30, 30, 30, 30, 30, 30, 30,
    5, 5, 5, 10, 10, 5, 5, 5, 3, 3, 3, 8, 8, 3, 3, 3, 2, 2, 2, 2, 2, 2, 2, 2,
    0, 0, 0, -5, -5, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    // knight
    -50, -40, -30, -30, -30, -30, -40, -50, -40, -20, 0, 0, 0, 0, -20, -40,
    -30, 0, 10, 15, 15, 10, 0, -30, -30, 5, 15, 20, 20, 15, 5, -30,
    -30, 0, 15, 20, 20, 15, 0, -30, -30, 5, 10, 15, 15, 10, 5, -30,
    -40, -20, 0, 5, 5, 0, -20, -40, -50, -40, -30, -30, -30, -30, -40, -50,
    // king
    -1, -1, -1, -1, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, -1, -1, 0, 0, 0, 0, 0, 0, -1,
    -1, 0, 0, 0, 0, 0, 0, -1, -1, 0, 0, 0, 0, 0, 0, -1, -1, 0, 0, 0, 0, 0, 0, -1,
    -1, 0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1,
    // bishop
    -20, -10, -10, -10, -10, -10, -10, -20, -10, 0, 0, 0, 0, 0, 0, -10,
    -10, 0, 5, 10, 10, 5, 0, -10, -10, 5, 5, 10, 10, 5, 5, -10,
    -10, 0, 10, 10, 10, 10, 0, -10, -10, 10, 10, 10, 10, 10, 10, -10,
    -10, 5, 0, 0, 0, 0, 5, -10, -20, -10, -10, -10, -10, -10, -10, -20,
    // rook
    0, 0, 0, 0, 0, 0, 0, 0, 5, 10, 10, 10, 10, 10, 10, 5, -5, 0, 0, 0, 0, 0, 0, -5,
    -5, 0, 0, 0, 0, 0, 0, -5, -5, 0, 0, 0, 0, 0, 0, -5, -5, 0, 0, 0, 0, 0, 0, -5,
    -5, 0, 0, 0, 0, 0, 0, -5, 0, 0, 0, 5, 5, 0, 0, 0,
    // queen
    -20, -10, -10, -5, -5, -10, -10, -20, -10, 0, 0, 0, 0, 0, 0, -10,
    -10, 0, 5, 5, 5, 5, 0, -10, -5, 0, 5, 5, 5, 5, 0, -5,
    -5, 0, 5, 5, 5, 5, 0, -5, -10, 0, 5, 5, 5, 5, 0, -10,
    -10, 0, 0, 0, 0, 0, 0, -10, -20, -10, -10, -5, -5, -10, -10, -20};

// One square's share of the x64 evaluation (src/zeta.cl:1482-1521), white
// positive: +10 per piece, material, piece-square, square control, and the
// pawn terms (blocked -15, chain +10 per diagonal-behind own pawn, -30 per
// own pawn behind on the file).  The bishop pair (src/zeta.cl:1522-1531) is
// added once per position by the caller.
// 10 + material + piece-square + square control of a piece type on a square
// (black view), folded into one table so a square costs one lookup
struct EvalSq {
  int32_t v[7 * 64];
};
constexpr EvalSq make_eval_sq() {
  EvalSq t{};
  for (int ty = 1; ty < 7; ty++)
    for (int sq = 0; sq < 64; sq++) t.v[ty * 64 + sq] = 10 + EVAL_VALUE[ty] + EVAL_PST[ty * 64 + sq] + EVAL_CONTROL[sq];
  return t;
}
static ZC_CONST const EvalSq EVAL_SQ = make_eval_sq();

ZD int32_t eval_square(const Pos &p, int sq) {
  const u32 pc = piece_at(p, sq);
  const int type = (int)((pc >> 1) & 7);
  if (type == PNONE) return 0;
  const int color = (int)(pc & 1);
  const int sqf = color ? sq : (sq ^ 56);
  int32_t s = EVAL_SQ.v[type * 64 + sqf];
  if (type == PAWN) {
    const u64 occ = p.b1 | p.b2 | p.b3;
    const u64 pawns = p.b1 & ~p.b2 & ~p.b3;
    const u64 me = color ? p.b0 : (p.b0 ^ occ);
    const u64 opp = color ? (p.b0 ^ occ) : p.b0;
    const u64 mp = pawns & me;
    const u64 file = 0x0101010101010101ULL << (sq & 7);
    if (color == 0) {
      if ((sq >> 3) < 7 && (opp & bit(sq + 8))) s -= 15;
      if ((sq & 7) < 7 && (mp & bit(sq - 7))) s += 10;
      if ((sq & 7) > 0 && (mp & bit(sq - 9))) s += 10;
      // own pawns behind on the file, ranks 2 .. rank-1 (the reference's loop
      // runs t = sq-8, sq-16, .. while t > 7)
      if (sq >= 8) s -= 30 * popc(mp & file & (bit(sq) - 1) & ~0xFFULL);
      else
        for (uint8_t t = (uint8_t)(sq - 8); t > 7; t = (uint8_t)(t - 8))
          if (mp & bit(t)) s -= 30;
    } else {
      if ((sq >> 3) > 0 && (opp & bit(sq - 8))) s -= 15;
      if ((sq & 7) > 0 && (mp & bit(sq + 7))) s += 10;
      if ((sq & 7) < 7 && (mp & bit(sq + 9))) s += 10;
      // ranks rank+1 .. 7 (t = sq+8, sq+16, .. while t < 56)
      s -= 30 * popc(mp & file & ~((bit(sq) << 1) - 1) & ~(0xFFULL << 56));
    }
  }
  return color ? -s : s;
}

ZD int32_t eval_bishop_pair(const Pos &p) {
  const u64 occ = p.b1 | p.b2 | p.b3;
  const u64 bishops = ~p.b1 & ~p.b2 & p.b3;
  return (popc((p.b0 ^ occ) & bishops) == 2 ? 25 : 0) - (popc(p.b0 & bishops) == 2 ? 25 : 0);
}

// Move-ordering score of the reference picker (src/zeta.cl:2059-2071):
// captures MVV-LVA (16*V[captured] - V[mover]), quiet moves by the change of
// material + piece-square + control value.
ZD int32_t order_score(u32 stm, u32 m) {
  const int from = (int)mv_from(m), to = (int)mv_to(m);
  const int tf = (int)((mv_pto(m) >> 1) & 7), ff = (int)((mv_pfrom(m) >> 1) & 7), cf = (int)((mv_pcpt(m) >> 1) & 7);
  if (cf != PNONE) return EVAL_VALUE[cf] * 16 - EVAL_VALUE[tf];
  const int t = stm ? to : (to ^ 56), f = stm ? from : (from ^ 56);
  return EVAL_VALUE[tf] + EVAL_PST[tf * 64 + t] + EVAL_CONTROL[t] - EVAL_VALUE[ff] - EVAL_PST[ff * 64 + f] -
         EVAL_CONTROL[f];
}

}  // namespace zc

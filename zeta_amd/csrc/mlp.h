// mlp.h -- the learned move-ordering network (no counterpart in the
// reference; its plain MVV-LVA/killer ordering, src/zeta.cl:2059-2104, stays
// the fallback).  908 -> 256 -> 32 -> 1 with ReLU, BF16 weights, FP32
// accumulation.  Inputs are side-to-move relative one-hots:
//   [0, 768)   piece-square: ((own ? 0 : 1) * 6 + type - 1) * 64 + rsq
//   [768, 832) origin square rsq      [832, 896) target square rsq
//   [896, 902) moving piece type - 1   [902, 908) captured piece type - 1
// where rsq = sq for white to move and sq ^ 56 for black to move.  The
// first layer is padded to 928 inputs so every GEMM dimension is a multiple
// of the 32x32 MFMA tile.
#pragma once
#include "chess.h"

namespace zm {

using zc::u16;
using zc::u32;
using zc::u64;

constexpr int NIN = 908;
constexpr int NINP = 928;  // padded
constexpr int NH1 = 256;
constexpr int NH2 = 32;
constexpr int NPARAMS = NIN * NH1 + NH1 + NH1 * NH2 + NH2 + NH2 + 1;  // 240961

// flat FP32 parameter layout of the C-ABI (zeta_mlp_get / zeta_mlp_set)
constexpr int P_W1 = 0;                   // W1[f][u], 908 x 256
constexpr int P_B1 = P_W1 + NIN * NH1;    // 256
constexpr int P_W2 = P_B1 + NH1;          // W2[u][o], 256 x 32
constexpr int P_B2 = P_W2 + NH1 * NH2;    // 32
constexpr int P_W3 = P_B2 + NH2;          // 32
constexpr int P_B3 = P_W3 + NH2;          // 1

// The search orders moves by plain score + PLAIN_PER_LOGIT x network output:
// the network learns a correction of the plain ordering.  A training pair
// carries the plain score difference (negative minus positive) in the upper
// 24 bits of its depth word; the loss adds it / PLAIN_PER_LOGIT.
constexpr int PLAIN_PER_LOGIT = 100;
constexpr int PAIR_OFFSET_MAX = (1 << 23) - 1;
ZD u32 pair_depth_word(int depth, int offset) { return ((u32)offset << 8) | ((u32)depth & 0xFFu); }
ZD int pair_offset(u32 w) { return (int)w >> 8; }

// device weight set used by the search and by training
struct Weights {
  const u16 *w1f;   // BF16 [928][256]: W1, unit-contiguous (row gathers in the search)
  const u16 *w1t;   // BF16 [256][928]: W1 transposed, input-contiguous
  const u16 *w2t;   // BF16 [32][256]
  const float *b1;  // [256]
  const float *b2;  // [32]
  const float *w3;  // [32]
  const float *b3;  // [1]
};

ZD int rel_sq(int sq, u32 stm) { return stm ? (sq ^ 56) : sq; }
ZD int feat_ps(u32 pc, int sq, u32 stm) {
  const int own = ((pc & 1u) == stm) ? 0 : 1;
  return (own * 6 + (int)((pc >> 1) & 7) - 1) * 64 + rel_sq(sq, stm);
}
ZD int feat_from(u32 m, u32 stm) { return 768 + rel_sq((int)zc::mv_from(m), stm); }
ZD int feat_to(u32 m, u32 stm) { return 832 + rel_sq((int)zc::mv_to(m), stm); }
ZD int feat_piece(u32 m) { return 896 + (int)((zc::mv_pfrom(m) >> 1) & 7) - 1; }
// -1 when the move captures nothing
ZD int feat_capt(u32 m) {
  const int t = (int)((zc::mv_pcpt(m) >> 1) & 7);
  return t ? 902 + t - 1 : -1;
}

// FP32 -> BF16, round to nearest even (finite inputs)
ZD u16 f2bf(float f) {
  u32 x;
  __builtin_memcpy(&x, &f, 4);
  x += 0x7FFFu + ((x >> 16) & 1u);
  return (u16)(x >> 16);
}
ZD float bf2f(u16 h) {
  const u32 x = (u32)h << 16;
  float f;
  __builtin_memcpy(&f, &x, 4);
  return f;
}

}  // namespace zm

// Host build of the device chess core (zeta_amd/csrc/chess.h) for CPU unit
// tests of the kernel logic.  Test-only: the product never loads this.
#include "../../zeta_amd/csrc/chess.h"
#include "../../zeta_amd/csrc/evaltab.h"

#include <atomic>
#include <thread>
#include <vector>

using namespace zc;

static Pos load(const uint64_t *b, int stm) {
  Pos p;
  p.b0 = b[0]; p.b1 = b[1]; p.b2 = b[2]; p.b3 = b[3]; p.cr = b[4]; p.stm = (u32)stm;
  return p;
}

extern "C" int hc_count(const uint64_t *b, int stm, int qs) {
  Pos p = load(b, stm);
  Legal L;
  analyse(p, L);
  return count_legal(p, L, qs != 0);
}

extern "C" int hc_enumerate(const uint64_t *b, int stm, int qs, uint32_t *moves) {
  Pos p = load(b, stm);
  Legal L;
  analyse(p, L);
  int n = 0;
  u64 pcs = L.own;
  while (pcs) {
    int f = lsb(pcs);
    pcs &= pcs - 1;
    u64 t = piece_targets(p, L, f, qs != 0);
    while (t) {
      int to = lsb(t);
      t &= t - 1;
      moves[n++] = make_code(p, f, to);
    }
  }
  return n;
}

// the grouped enumeration the perft expansion kernel writes pairs with
extern "C" int hc_enumerate_sets(const uint64_t *b, int stm, uint32_t *moves) {
  Pos p = load(b, stm);
  Legal L;
  analyse(p, L);
  int n = 0;
  for_each_legal(p, L, [&](u32 code) { moves[n++] = code; });
  return n;
}

extern "C" int hc_nth_bit(uint64_t x, int k) { return nth_bit(x, k); }

// castle_targets against the per-square form of the same rule (king-relative
// squares k-4..k+3, each its own shift), for arbitrary king squares, moved
// flags, occupancy, attack map and rooks: the fast path (shifted constants for
// kings on squares 4-60) must give the same targets
extern "C" uint64_t hc_castle(uint64_t b0, uint64_t b1, uint64_t b2, uint64_t b3, uint64_t cr, int stm, int ksq,
                              uint64_t occ, uint64_t own, uint64_t A) {
  Pos p;
  p.b0 = b0; p.b1 = b1; p.b2 = b2; p.b3 = b3; p.cr = cr; p.stm = (u32)stm;
  Legal L{};
  L.ksq = ksq; L.occ = occ; L.own = own; L.A = A;
  L.rq = b3 & (b1 ^ b2);
  return castle_targets(p, L);
}

extern "C" uint64_t hc_castle_per_square(uint64_t b0, uint64_t b1, uint64_t b2, uint64_t b3, uint64_t cr, int stm,
                                         int k, uint64_t occ, uint64_t own, uint64_t A) {
  (void)b0;
  if (!(cr & SMCRALL)) return 0;
  const bool okq = stm ? ((~cr & SMCRBLACKQ) == SMCRBLACKQ) : ((~cr & SMCRWHITEQ) == SMCRWHITEQ);
  const bool okk = stm ? ((~cr & SMCRBLACKK) == SMCRBLACKK) : ((~cr & SMCRWHITEK) == SMCRWHITEK);
  const u64 ownrooks = own & (b3 & (b1 ^ b2)) & ~b2;
  u64 t = 0;
  if (okq && (ownrooks & bit(k - 4)) && !(occ & (bit(k - 1) | bit(k - 2) | bit(k - 3))) &&
      !(A & (bit(k) | bit(k - 1) | bit(k - 2))))
    t |= bit(k - 2);
  if (okk && (ownrooks & bit(k + 3)) && !(occ & (bit(k + 1) | bit(k + 2))) && !(A & (bit(k) | bit(k + 1) | bit(k + 2))))
    t |= bit(k + 2);
  return t;
}

extern "C" void hc_apply(const uint64_t *b, int stm, uint32_t m, int rook_flag, uint64_t *out) {
  Pos p = load(b, stm);
  Pos c = apply(p, m, rook_flag != 0);
  out[0] = c.b0; out[1] = c.b1; out[2] = c.b2; out[3] = c.b3; out[4] = c.cr; out[5] = c.stm;
}

static uint64_t perft_rec(const Pos &p, int d) {
  Legal L;
  analyse(p, L);
  if (d == 1) return (uint64_t)count_legal(p, L, false);
  uint64_t s = 0;
  u64 pcs = L.own;
  while (pcs) {
    int f = lsb(pcs);
    pcs &= pcs - 1;
    u64 t = piece_targets(p, L, f, false);
    while (t) {
      int to = lsb(t);
      t &= t - 1;
      s += perft_rec(apply(p, make_code(p, f, to), true), d - 1);
    }
  }
  return s;
}

extern "C" uint64_t hc_perft(const uint64_t *b, int stm, int depth) {
  if (depth == 0) return 1;
  return perft_rec(load(b, stm), depth);
}

// the search's lane-per-square analysis, lanes emulated by a loop: attacked
// squares combined as a ballot would, targets per origin square
extern "C" int hc_enumerate_lane(const uint64_t *b, int stm, int qs, uint32_t *moves) {
  Pos p = load(b, stm);
  Node0 B;
  node_basics(p, B);
  u64 A = 0;
  for (int sq = 0; sq < 64; sq++)
    if (attacked_nk(B, sq)) A |= bit(sq);
  int n = 0;
  for (int sq = 0; sq < 64; sq++) {
    if (!((B.own >> sq) & 1)) continue;
    u64 t = lane_targets(p, B, A, sq, qs != 0);
    while (t) {
      int to = lsb(t);
      t &= t - 1;
      moves[n++] = make_code(p, sq, to);
    }
  }
  return n;
}

extern "C" uint64_t hc_checkers(const uint64_t *b, int stm) {
  Pos p = load(b, stm);
  Node0 B;
  node_basics(p, B);
  return B.checkers;
}

extern "C" uint64_t hc_slider(int sq, uint64_t occ, int rook) {
  return rook ? rook_sq(sq, occ) : bishop_sq(sq, occ);
}

extern "C" uint64_t hc_slider_ks(int sq, uint64_t occ, int rook) {
  return rook ? rook_att(bit(sq), ~occ) : bishop_att(bit(sq), ~occ);
}

// the search's static evaluation, lanes emulated by a loop: white positive
extern "C" int32_t hc_eval(const uint64_t *b) {
  Pos p = load(b, 0);
  int32_t e = 0;
  for (int sq = 0; sq < 64; sq++) e += eval_square(p, sq);
  return e + eval_bishop_pair(p);
}

// The device perft algorithm on the host: interior levels expand every legal
// move in the grouped order the expansion kernel writes (for_each_legal,
// apply with the castle rook), the last level counts legal moves (bulk
// counting, as the leaf kernel does).  One position per thread at a time.
static uint64_t walk_bulk(const Pos &p, int d) {
  Legal L;
  analyse(p, L);
  if (d == 1) return (uint64_t)count_legal(p, L, false);
  uint64_t s = 0;
  for_each_legal(p, L, [&](u32 code) { s += walk_bulk(apply(p, code, true), d - 1); });
  return s;
}

extern "C" void hc_perft_batch(const uint64_t *pos, int n, int depth, uint64_t *out, int threads) {
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < (threads > 0 ? threads : 1); t++)
    th.emplace_back([&] {
      for (int i; (i = next++) < n;) {
        const uint64_t *b = pos + (size_t)i * 8;  // 7 x u64 + stm (int32) + pad
        const Pos p = load(b, (int)(uint32_t)b[7]);
        out[i] = depth <= 0 ? 1 : walk_bulk(p, depth);
      }
    });
  for (auto &t : th) t.join();
}

// Zobrist hash of a position from scratch, as the device computes it
extern "C" uint64_t hc_hash(const uint64_t *b, int stm) { return hash_full(load(b, stm)); }

// The device code's constants, for the test that pins them to the numbers
// extracted from the reference: which 0 EVAL_VALUE, 1 EVAL_CONTROL, 2 EVAL_PST,
// 3 the folded per-square table the evaluation reads (10 + material +
// piece-square + control)
extern "C" int32_t hc_eval_const(int which, int idx) {
  if (which == 0) return (idx >= 0 && idx < 7) ? EVAL_VALUE[idx] : 0;
  if (which == 1) return (idx >= 0 && idx < 64) ? EVAL_CONTROL[idx] : 0;
  if (which == 2) return (idx >= 0 && idx < 7 * 64) ? EVAL_PST[idx] : 0;
  return (idx >= 0 && idx < 7 * 64) ? EVAL_SQ.v[idx] : 0;
}
extern "C" uint64_t hc_zobrist_key(int i) { return (i >= 0 && i < 18) ? ZOBRIST[i] : 0; }
extern "C" uint64_t hc_castle_mask(int i) {
  const u64 m[5] = {SMCRALL, SMCRWHITEQ, SMCRWHITEK, SMCRBLACKQ, SMCRBLACKK};
  return (i >= 0 && i < 5) ? m[i] : 0;
}

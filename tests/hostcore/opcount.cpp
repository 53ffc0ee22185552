// Op-counting host build of the device chess core: every 64-bit integer
// operation (and/or/xor/not/shift/add/sub/compare/popcount/bit-scan) on a
// bitboard counts as one op.  Used once to derive the algorithmic work per
// leaf-kernel position (apply + analyse + count) for the perft roofline.
// Test/measurement tool only.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

static uint64_t g_ops = 0;  // counted 64-bit operations
static uint64_t g_ins = 0;  // their minimal gfx950 VALU instruction cost

// cost model (one wave64 VALU instruction = 64 int32 lane-ops):
//   64-bit and/or/xor: 2 (v_and_b32 x2), 1 if one operand is a constant whose
//   high or low half is 0 / all-ones; not/add/sub: 2; shift: 1 (v_lshlrev_b64);
//   compare / test against zero: 1 (v_cmp_*_u64); popcount: 2 (v_bcnt x2);
//   bit scan: 3 (v_ffbl x2 + select)
struct OC {
  uint64_t v;
  bool k;  // compile-time constant
  constexpr OC(uint64_t x = 0) : v(x), k(true) {}
  static OC r(uint64_t x) { OC o(x); o.k = false; return o; }
  explicit operator bool() const { ++g_ops; ++g_ins; return v != 0; }
  explicit operator uint32_t() const { return (uint32_t)v; }
  OC operator~() const { ++g_ops; g_ins += 2; return r(~v); }
  OC &operator&=(OC b);
  OC &operator|=(OC b);
  OC &operator^=(OC b);
};
static inline bool trivial_half(OC a) {
  uint32_t lo = (uint32_t)a.v, hi = (uint32_t)(a.v >> 32);
  return a.k && (lo == 0 || lo == 0xFFFFFFFFu || hi == 0 || hi == 0xFFFFFFFFu);
}
static inline void bitop(OC a, OC b) { ++g_ops; g_ins += (trivial_half(a) || trivial_half(b)) ? 1 : 2; }
inline OC operator&(OC a, OC b) { bitop(a, b); return OC::r(a.v & b.v); }
inline OC operator|(OC a, OC b) { bitop(a, b); return OC::r(a.v | b.v); }
inline OC operator^(OC a, OC b) { bitop(a, b); return OC::r(a.v ^ b.v); }
inline OC &OC::operator&=(OC b) { bitop(*this, b); v &= b.v; k = false; return *this; }
inline OC &OC::operator|=(OC b) { bitop(*this, b); v |= b.v; k = false; return *this; }
inline OC &OC::operator^=(OC b) { bitop(*this, b); v ^= b.v; k = false; return *this; }
inline OC operator+(OC a, OC b) { ++g_ops; g_ins += 2; return OC::r(a.v + b.v); }
inline OC operator-(OC a, OC b) { ++g_ops; g_ins += 2; return OC::r(a.v - b.v); }
inline OC operator*(OC a, OC b) { ++g_ops; g_ins += 4; return OC::r(a.v * b.v); }
inline OC operator<<(OC a, int n) { ++g_ops; g_ins += 1; return OC::r(a.v << n); }
inline OC operator>>(OC a, int n) { ++g_ops; g_ins += 1; return OC::r(a.v >> n); }
inline bool operator==(OC a, OC b) { ++g_ops; ++g_ins; return a.v == b.v; }
inline bool operator!=(OC a, OC b) { ++g_ops; ++g_ins; return a.v != b.v; }
// popcount / bit scans go through zc_raw; cost by caller
static int g_rawcost = 2;
inline uint64_t zc_raw(OC a) { ++g_ops; g_ins += g_rawcost; return a.v; }

#define ZC_U64 OC
#include "../../zeta_amd/csrc/chess.h"

using namespace zc;

static Pos load(const uint64_t *b, int stm) {
  Pos p;
  p.b0 = b[0]; p.b1 = b[1]; p.b2 = b[2]; p.b3 = b[3]; p.cr = b[4]; p.stm = (u32)stm;
  return p;
}

// walk to the leaf-kernel level (depth-1 nodes of perft(depth)) and count the
// ops of apply + analyse + count at every such node
static void walk(const Pos &p, int d, bool qs_leaf, double &ops, double &n) {
  Legal L;
  analyse(p, L);
  uint64_t pcs = ZC_RAW(L.own);
  while (pcs) {
    int f = __builtin_ctzll(pcs);
    pcs &= pcs - 1;
    uint64_t t = ZC_RAW(piece_targets(p, L, f, false));
    while (t) {
      int to = __builtin_ctzll(t);
      t &= t - 1;
      u32 m = make_code(p, f, to);
      if (d == 1) {
        uint64_t before = g_ins;
        Pos c = apply(p, m, true);
        Legal Lc;
        analyse(c, Lc);
        volatile int k = count_legal(c, Lc, qs_leaf);
        (void)k;
        ops += (double)(g_ins - before);
        n += 1;
      } else {
        walk(apply(p, m, true), d - 1, qs_leaf, ops, n);
      }
    }
  }
}

int main(int argc, char **argv) {
  // stdin: lines "b0 b1 b2 b3 cr stm" (hex words) of root positions
  int depth = argc > 1 ? atoi(argv[1]) : 3;
  std::vector<uint64_t> buf(6);
  double ob = 0, nb = 0, oq = 0, nq = 0;
  unsigned long long w[5];
  int stm;
  while (scanf("%llx %llx %llx %llx %llx %d", &w[0], &w[1], &w[2], &w[3], &w[4], &stm) == 6) {
    uint64_t b[5] = {w[0], w[1], w[2], w[3], w[4]};
    Pos p = load(b, stm);
    walk(p, depth, false, ob, nb);
    walk(p, depth, true, oq, nq);
  }
  printf("{\"bulk\": %.1f, \"nobulk\": %.1f, \"items\": %.0f}  (minimal VALU instructions per position)\n",
         ob / nb, oq / nq, nb);
  return 0;
}

// Op-counting host build of the device chess core: every 64-bit integer
// operation (and/or/xor/not/shift/add/sub/compare/popcount/bit-scan) on a
// bitboard counts as one op.  Used once to derive the algorithmic work per
// leaf-kernel position (apply + analyse + count) for the perft roofline.
// Test/measurement tool only.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

static uint64_t g_ops = 0;

struct OC {
  uint64_t v;
  constexpr OC(uint64_t x = 0) : v(x) {}
  explicit operator bool() const { ++g_ops; return v != 0; }
  explicit operator uint32_t() const { return (uint32_t)v; }
  OC operator~() const { ++g_ops; return OC(~v); }
  OC &operator&=(OC b) { ++g_ops; v &= b.v; return *this; }
  OC &operator|=(OC b) { ++g_ops; v |= b.v; return *this; }
  OC &operator^=(OC b) { ++g_ops; v ^= b.v; return *this; }
};
inline OC operator&(OC a, OC b) { ++g_ops; return OC(a.v & b.v); }
inline OC operator|(OC a, OC b) { ++g_ops; return OC(a.v | b.v); }
inline OC operator^(OC a, OC b) { ++g_ops; return OC(a.v ^ b.v); }
inline OC operator+(OC a, OC b) { ++g_ops; return OC(a.v + b.v); }
inline OC operator-(OC a, OC b) { ++g_ops; return OC(a.v - b.v); }
inline OC operator*(OC a, OC b) { ++g_ops; return OC(a.v * b.v); }
inline OC operator<<(OC a, int n) { ++g_ops; return OC(a.v << n); }
inline OC operator>>(OC a, int n) { ++g_ops; return OC(a.v >> n); }
inline bool operator==(OC a, OC b) { ++g_ops; return a.v == b.v; }
inline bool operator!=(OC a, OC b) { ++g_ops; return a.v != b.v; }
inline uint64_t zc_raw(OC a) { ++g_ops; return a.v; }

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
        uint64_t before = g_ops;
        Pos c = apply(p, m, true);
        Legal Lc;
        analyse(c, Lc);
        volatile int k = count_legal(c, Lc, qs_leaf);
        (void)k;
        ops += (double)(g_ops - before);
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
  printf("{\"bulk\": %.1f, \"nobulk\": %.1f, \"items\": %.0f}\n", ob / nb, oq / nq, nb);
  return 0;
}

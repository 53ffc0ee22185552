// chess.h -- device chess core for gfx950 (also compiles for the host so the
// same source can be unit-tested without a GPU; the library itself only runs
// it on the device).
//
// Positions use the reference's quad-bitboard (src/types.h:60-66):
//   b0 = black pieces, b1..b3 = piece-type bit planes (type = b1 + 2*b2 + 4*b3:
//   1 pawn, 2 knight, 3 king, 4 bishop, 5 rook, 6 queen), cr = "moved" flags
//   (castle rights live on a1/e1/h1/a8/e8/h8, the e.p. pawn is the one cleared
//   bit of ranks 4/5).
//
// The generator here is set-wise, not the reference's 64-lane square-wise one:
// one lane analyses a whole position with Kogge-Stone occluded fills (the
// opponent's slider rays give the attack map, and against the king's own rays
// the checkers, check rays, pins and the squares attacked through the king),
// then counts legal moves direction by direction (every target of a
// direction/offset has exactly one origin, so popcounts count moves).  It
// generates exactly the reference's legal move set, including its quirks:
// queen-only promotion (Q1, src/zetaperft.cl:1380-1384), the SMCRALL castle
// precondition (Q2, src/zetaperft.cl:1254), e.p. only from the moved flags
// with a full make/test legality check (src/zetaperft.cl:1222-1248), and
// 6-bit-masked shift counts (Q4).
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ZD __host__ __device__ __forceinline__
#define ZC_CONST __constant__
#else
// host-only build of this header (CPU unit tests of the device logic)
#define ZD static inline
#define ZC_CONST
#endif

namespace zc {

typedef uint64_t u64;
typedef uint32_t u32;
typedef int32_t s32;
typedef uint16_t u16;
typedef uint8_t u8;

constexpr u64 NOTA = 0xFEFEFEFEFEFEFEFEULL;
constexpr u64 NOTH = 0x7F7F7F7F7F7F7F7FULL;
constexpr u64 FULL = ~0ULL;
constexpr u64 RANKS45 = 0x000000FFFF000000ULL;
constexpr u64 RANK4 = 0x00000000FF000000ULL;
constexpr u64 RANK5 = 0x000000FF00000000ULL;
constexpr u64 RANK3 = 0x0000000000FF0000ULL;
constexpr u64 RANK6 = 0x0000FF0000000000ULL;
constexpr u64 SMCRALL = 0x8900000000000091ULL;     // src/types.h:114 (d8, not e8)
constexpr u64 SMCRWHITEQ = 0x0000000000000011ULL;  // src/types.h:119-124
constexpr u64 SMCRWHITEK = 0x0000000000000090ULL;
constexpr u64 SMCRBLACKQ = 0x1100000000000000ULL;
constexpr u64 SMCRBLACKK = 0x9000000000000000ULL;

enum { PNONE = 0, PAWN = 1, KNIGHT = 2, KING = 3, BISHOP = 4, ROOK = 5, QUEEN = 6 };

ZD int popc(u64 x) { return __builtin_popcountll(x); }
ZD int lsb(u64 x) { return __builtin_ctzll(x); }
ZD int msb(u64 x) { return 63 - __builtin_clzll(x); }
// Q4: shift counts are taken modulo 64, as OpenCL C does.
ZD u64 bit(int sq) { return u64(1) << (sq & 63); }
ZD u64 shlm(u64 x, int n) { return x << (n & 63); }
ZD u64 shrm(u64 x, int n) { return x >> (n & 63); }
ZD u64 sel(bool c, u64 x) { return x & (0ULL - (u64)c); }

// A 64-bit value the device compiler must keep whole in a register pair (an
// empty asm statement): it then shifts it with one v_lshlrev_b64 instead of
// splitting the shift into 32-bit halves for consumers that use halves.
ZD u64 whole(u64 x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(x));
#endif
  return x;
}
// (a & b) | c.  The compiler splits 64-bit AND and OR into 32-bit halves
// before it looks for the fused v_and_or_b32, so plain C costs four VALU
// instructions here; written out per half it is two.  Per-lane values only
// (the operands are forced into vector registers).
ZD u64 and_or(u64 a, u64 b, u64 c) {
#if defined(__HIP_DEVICE_COMPILE__)
  u32 lo, hi;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(lo) : "v"((u32)a), "v"((u32)b), "v"((u32)c));
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(hi) : "v"((u32)(a >> 32)), "v"((u32)(b >> 32)), "v"((u32)(c >> 32)));
  return whole(((u64)hi << 32) | lo);
#else
  return (a & b) | c;
#endif
}

// ---- move code: the reference's 32-bit packing (src/types.h:67-74,126-141)
ZD u32 mv_from(u32 m) { return m & 63; }
ZD u32 mv_to(u32 m) { return (m >> 6) & 63; }
ZD u32 mv_cpt(u32 m) { return (m >> 12) & 63; }
ZD u32 mv_pfrom(u32 m) { return (m >> 18) & 15; }
ZD u32 mv_pto(u32 m) { return (m >> 22) & 15; }
ZD u32 mv_pcpt(u32 m) { return (m >> 26) & 15; }
ZD u32 mv_make(u32 f, u32 t, u32 c, u32 pf, u32 pt, u32 pc) {
  return f | (t << 6) | (c << 12) | (pf << 18) | (pt << 22) | (pc << 26);
}
constexpr u32 NULLMOVE = 0x41u;

struct Pos {
  u64 b0, b1, b2, b3, cr;
  u32 stm;
};

ZD u32 piece_at(const Pos &p, int sq) {
  return (u32)((shrm(p.b0, sq) & 1) | ((shrm(p.b1, sq) & 1) << 1) | ((shrm(p.b2, sq) & 1) << 2) |
               ((shrm(p.b3, sq) & 1) << 3));
}

// ---- one-step shifts and Kogge-Stone occluded fills for the 8 directions
// dir: 0 N, 1 S, 2 E, 3 W, 4 NE, 5 NW, 6 SE, 7 SW
ZD u64 step(u64 x, int d) {
  switch (d) {
    case 0: return x << 8;
    case 1: return x >> 8;
    case 2: return (x << 1) & NOTA;
    case 3: return (x >> 1) & NOTH;
    case 4: return (x << 9) & NOTA;
    case 5: return (x << 7) & NOTH;
    case 6: return (x >> 7) & NOTA;
    default: return (x >> 9) & NOTH;
  }
}
// one doubling step of an occluded fill: g | (e & (g shifted by SH)).  A
// shift by 32 or more moves one half into the other, which plain C already
// does without instructions; shorter shifts use the fused AND-OR.
template <int SH, bool UP>
ZD u64 fill_step(u64 g, u64 e) {
  const u64 t = UP ? g << SH : g >> SH;
  if (SH >= 32) return g | (e & t);
  return and_or(e, whole(t), g);
}
// occluded fill: generators plus every square reachable through `empty`
template <int D>
ZD u64 ks_fill(u64 g, u64 e) {
  constexpr int s = (D == 0 || D == 1) ? 8 : (D == 2 || D == 3) ? 1 : (D == 4 || D == 7) ? 9 : 7;
  constexpr bool up = (D == 0 || D == 2 || D == 4 || D == 5);
  constexpr u64 m = (D == 2 || D == 4 || D == 6) ? NOTA : (D == 3 || D == 5 || D == 7) ? NOTH : FULL;
  e &= m;
  if (up) {
    g = fill_step<s, true>(g, e);
    e &= e << s;
    g = fill_step<2 * s, true>(g, e);
    e &= e << (2 * s);
    g = fill_step<4 * s, true>(g, e);
  } else {
    g = fill_step<s, false>(g, e);
    e &= e >> s;
    g = fill_step<2 * s, false>(g, e);
    e &= e >> (2 * s);
    g = fill_step<4 * s, false>(g, e);
  }
  return g;
}
// the one-step shift of direction D without its edge-file mask (NOTA for
// E, NE, SE; NOTH for W, NW, SW; none for N, S): directions that share a
// mask can be masked once, after their union or on the target set
template <int D>
ZD u64 shift_raw(u64 x) {
  return D == 0 ? x << 8 : D == 1 ? x >> 8 : D == 2 ? x << 1 : D == 3 ? x >> 1
       : D == 4 ? x << 9 : D == 5 ? x << 7 : D == 6 ? x >> 7 : x >> 9;
}
// squares a slider set attacks in direction D (through empties, incl. first blocker)
template <int D>
ZD u64 slide(u64 g, u64 empty) {
  return step(ks_fill<D>(g, empty), D);
}
// the same without the edge-file mask
template <int D>
ZD u64 slide_raw(u64 g, u64 empty) {
  return shift_raw<D>(ks_fill<D>(g, empty));
}
ZD u64 rook_att(u64 g, u64 empty) {
  return slide<0>(g, empty) | slide<1>(g, empty) | slide<2>(g, empty) | slide<3>(g, empty);
}
ZD u64 bishop_att(u64 g, u64 empty) {
  return slide<4>(g, empty) | slide<5>(g, empty) | slide<6>(g, empty) | slide<7>(g, empty);
}
ZD u64 knight_att(u64 n) {
  u64 l1 = (n >> 1) & NOTH, l2 = (n >> 2) & 0x3F3F3F3F3F3F3F3FULL;
  u64 r1 = (n << 1) & NOTA, r2 = (n << 2) & 0xFCFCFCFCFCFCFCFCULL;
  u64 h1 = l1 | r1, h2 = l2 | r2;
  return (h1 << 16) | (h1 >> 16) | (h2 << 8) | (h2 >> 8);
}
ZD u64 king_att(u64 k) {
  u64 a = k | ((k << 1) & NOTA) | ((k >> 1) & NOTH);
  a |= (a << 8) | (a >> 8);
  return a ^ k;
}
ZD u64 pawn_att(u64 p, u32 color) {
  return color ? (((p >> 7) & NOTA) | ((p >> 9) & NOTH)) : (((p << 9) & NOTA) | ((p << 7) & NOTH));
}
// full board lines through square k
ZD u64 rank_line(int k) { return 0xFFULL << (k & 56); }
ZD u64 file_line(int k) { return 0x0101010101010101ULL << (k & 7); }
ZD u64 diag_line(int k) {  // a1-h8 direction
  int d = (k & 7) - (k >> 3);
  return d >= 0 ? (0x8040201008040201ULL >> (8 * d)) : (0x8040201008040201ULL << (-8 * d));
}
ZD u64 anti_line(int k) {  // h1-a8 direction
  int s = (k & 7) + (k >> 3);
  return s <= 7 ? (0x0102040810204080ULL >> (8 * (7 - s))) : (0x0102040810204080ULL << (8 * (s - 7)));
}

// ---- single-square slider attacks (hyperbola quintessence with a full
// 64-bit bit reversal, which orders every line -- ranks included -- so the
// borrow of a subtraction runs along it in both directions).  One slider
// per lane: the search analyses one position with 64 lanes, lane = square.
ZD u64 rev64(u64 x) {
#if defined(__clang__)
  return __builtin_bitreverse64(x);
#else
  u64 r = x;
  r = ((r >> 1) & 0x5555555555555555ULL) | ((r & 0x5555555555555555ULL) << 1);
  r = ((r >> 2) & 0x3333333333333333ULL) | ((r & 0x3333333333333333ULL) << 2);
  r = ((r >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((r & 0x0F0F0F0F0F0F0F0FULL) << 4);
  return __builtin_bswap64(r);
#endif
}
// attacks of a slider on sq along the line `mask` (mask excludes sq)
ZD u64 line_att(int sq, u64 occ, u64 mask) {
  const u64 b = bit(sq);
  const u64 o = occ & mask;
  const u64 f = o - b;
  const u64 r = rev64(rev64(o) - rev64(b));
  return (f ^ r) & mask;
}
ZD u64 rook_sq(int sq, u64 occ) {
  const u64 b = bit(sq);
  return line_att(sq, occ, rank_line(sq) ^ b) | line_att(sq, occ, file_line(sq) ^ b);
}
ZD u64 bishop_sq(int sq, u64 occ) {
  const u64 b = bit(sq);
  return line_att(sq, occ, diag_line(sq) ^ b) | line_att(sq, occ, anti_line(sq) ^ b);
}

// ---- per-position legality context
struct Legal {
  u64 own, opp, occ, kbb;
  u64 pawns, knights, kings, rq, bq;
  u64 A;         // opponent attacks as if the own king were not on the board
  u64 checkers;  // opponent pieces giving check
  u64 cmask;     // allowed targets of non-king moves: ~0, block/capture set, or 0 (double check)
  u64 pinR, pinF, pinD, pinA;  // own pieces pinned on the king's rank, file, diagonal, anti-diagonal
  int ksq;
  u32 stm;
};

// Checkers, check ray, x-ray and pins along one line through the king, from
// occluded fills in its two directions U and V: R = the rays of the opponent
// sliders S that move along the line, K = the king's own rays (each up to and
// including the first blocker).  The first piece on the king's U ray is a
// checker if it is in S; an own piece there is pinned if an S ray in
// direction V reaches it from the far side.  A checker's attack continues
// through the king along the king's opposite ray (`xray`: squares the king
// cannot step to).  Returns the pinned pieces; adds the rays to the attack
// map A.
template <int U, int V>
ZD u64 king_line(u64 S, u64 kbb, u64 own, u64 e, u64 &A, u64 &checkers, u64 &chkray, u64 &xray) {
  const u64 Ru = slide<U>(S, e), Rv = slide<V>(S, e);
  const u64 Ku = slide<U>(kbb, e), Kv = slide<V>(kbb, e);
  A |= Ru | Rv;
  const u64 cu = Ku & S, cv = Kv & S;
  checkers |= cu | cv;
  const bool hu = cu != 0, hv = cv != 0;
  chkray |= sel(hu, Ku) | sel(hv, Kv);
  xray |= sel(hu, Kv) | sel(hv, Ku);
  return ((Ku & Rv) | (Kv & Ru)) & own;
}

ZD void analyse(const Pos &p, Legal &L) {
  const u64 occ = p.b1 | p.b2 | p.b3;
  const u64 white = p.b0 ^ occ;
  L.stm = p.stm;
  L.occ = occ;
  L.own = p.stm ? p.b0 : white;
  L.opp = p.stm ? white : p.b0;
  L.pawns = p.b1 & ~p.b2 & ~p.b3;
  L.knights = ~p.b1 & p.b2 & ~p.b3;
  L.kings = p.b1 & p.b2 & ~p.b3;
  L.rq = p.b3 & (p.b1 ^ p.b2);  // rook 101, queen 110
  L.bq = p.b3 & ~p.b1;          // bishop 100, queen 110
  L.kbb = L.own & L.kings;
  L.ksq = L.kbb ? lsb(L.kbb) : 0;
  const u64 orq = L.opp & L.rq, obq = L.opp & L.bq;
  // the opponent's attacks with the own king as a blocker (every fill shares
  // the empty set, and with it the doubling masks, with the move counts),
  // completed by the squares the checking sliders attack through the king
  const u64 e = ~occ;
  u64 A = pawn_att(L.opp & L.pawns, p.stm ^ 1u) | knight_att(L.opp & L.knights) | king_att(L.opp & L.kings);
  u64 checkers = (pawn_att(L.kbb, p.stm) & L.opp & L.pawns) | (knight_att(L.kbb) & L.opp & L.knights);
  u64 chkray = 0, xray = 0;
  L.pinF = king_line<0, 1>(orq, L.kbb, L.own, e, A, checkers, chkray, xray);
  L.pinR = king_line<2, 3>(orq, L.kbb, L.own, e, A, checkers, chkray, xray);
  L.pinD = king_line<4, 7>(obq, L.kbb, L.own, e, A, checkers, chkray, xray);
  L.pinA = king_line<5, 6>(obq, L.kbb, L.own, e, A, checkers, chkray, xray);
  L.A = A | xray;
  L.checkers = checkers;
  const int n = popc(checkers);
  L.cmask = n == 0 ? FULL : (n == 1 ? (checkers | chkray) : 0ULL);
}

// is the side to move's king attacked after the e.p. capture from->to taking
// sqep?  A knight or pawn checks it after the capture exactly when it checks
// it before (unless it is the captured pawn); sliders by the line attacks of
// the king's square on the new occupancy (two single-square slider lookups,
// not eight occluded fills); a king never checks
ZD bool ep_legal(const Legal &L, int from, int to, int sqep) {
  if (!L.kbb) return true;
  const u64 occ2 = (L.occ ^ bit(from) ^ bit(sqep)) | bit(to);
  const u64 opp2 = L.opp & ~bit(sqep);
  const u64 att = (rook_sq(L.ksq, occ2) & opp2 & L.rq) | (bishop_sq(L.ksq, occ2) & opp2 & L.bq) |
                  (L.checkers & ~(L.rq | L.bq) & ~bit(sqep));
  return att == 0;
}

// e.p.: the pawn that just double-pushed is the lowest cleared bit of ranks 4/5
// (src/zetaperft.cl:1224-1230); returns the own pawns that may capture it.
ZD u64 ep_candidates(const Pos &p, const Legal &L, int &sqep) {
  const u64 w = ~p.cr & RANKS45;
  sqep = w ? lsb(w) : 0;
  u64 c = L.own & L.pawns & (p.stm ? RANK4 : RANK5);
  return c & (bit(sqep + 1) | bit(sqep - 1));
}

// castling targets of the king (src/zetaperft.cl:1249-1297).  The rule
// looks at the eight squares k-4 .. k+3 around the king (the reference's
// bit() takes them modulo 64), so the rook, occupancy and attack sets are
// rotated right by k-4: bit i is square k-4+i for every king square, with no
// branch on where the king stands.  A set bit of `bad` is a failed condition:
// queen side bits 0-4 (rook on k-4, k-3..k-1 empty, k-2..k not attacked),
// king side bits 4-7 (k+1, k+2 empty, k..k+2 not attacked, rook on k+3).
ZD u64 rotr64(u64 x, int n) { return shrm(x, n) | shlm(x, 64 - n); }
ZD u64 castle_targets(const Pos &p, const Legal &L) {
  const u64 cr = p.cr;
  if (!(cr & SMCRALL)) return 0;
  const bool okq = p.stm ? ((~cr & SMCRBLACKQ) == SMCRBLACKQ) : ((~cr & SMCRWHITEQ) == SMCRWHITEQ);
  const bool okk = p.stm ? ((~cr & SMCRBLACKK) == SMCRBLACKK) : ((~cr & SMCRWHITEK) == SMCRWHITEK);
  if (!okq && !okk) return 0;
  const u64 ownrooks = L.own & L.rq & ~p.b2;  // rook = 101 (queen has b2 set)
  const int q = (L.ksq - 4) & 63;
  const u32 rk = (u32)rotr64(ownrooks, q), oc = (u32)rotr64(L.occ, q), at = (u32)rotr64(L.A, q);
  const u32 bad = (~rk & 0x81u) | (oc & 0x6Eu) | (at & 0x7Cu);
  const u32 t = ((okq && !(bad & 0x1Fu)) ? 0x04u : 0u) | ((okk && !(bad & 0xF0u)) ? 0x40u : 0u);
  return rotr64((u64)t, (64 - q) & 63);  // back: bit 2 -> k-2, bit 6 -> k+2
}

// Number of legal moves (qs: captures and e.p. only, as the kernels' quiescence).
ZD int count_legal(const Pos &p, const Legal &L, bool qs) {
  const u64 own = L.own, occ = L.occ;
  const u64 T = (qs ? L.opp : ~own) & L.cmask;
  // the target set with each edge-file mask applied once (a shifted set
  // masked and then intersected with T is the shifted set intersected with
  // the masked T)
  const u64 TA = T & NOTA, TH = T & NOTH;
  const u64 pinned = L.pinR | L.pinF | L.pinD | L.pinA;
  int cnt = popc(king_att(L.kbb) & (qs ? L.opp : ~own) & ~L.A);
  const u64 kn = own & L.knights & ~pinned;
  const u64 TAB = T & 0xFCFCFCFCFCFCFCFCULL, TGH = T & 0x3F3F3F3F3F3F3F3FULL;
  cnt += popc((kn << 17) & TA) + popc((kn << 15) & TH) + popc((kn << 10) & TAB) + popc((kn << 6) & TGH);
  cnt += popc((kn >> 6) & TAB) + popc((kn >> 10) & TGH) + popc((kn >> 15) & TA) + popc((kn >> 17) & TH);
  const u64 e = ~occ;
  const u64 orq = own & L.rq, obq = own & L.bq;
  cnt += popc(slide_raw<0>(orq & (~pinned | L.pinF), e) & T);
  cnt += popc(slide_raw<1>(orq & (~pinned | L.pinF), e) & T);
  cnt += popc(slide_raw<2>(orq & (~pinned | L.pinR), e) & TA);
  cnt += popc(slide_raw<3>(orq & (~pinned | L.pinR), e) & TH);
  cnt += popc(slide_raw<4>(obq & (~pinned | L.pinD), e) & TA);
  cnt += popc(slide_raw<7>(obq & (~pinned | L.pinD), e) & TH);
  cnt += popc(slide_raw<5>(obq & (~pinned | L.pinA), e) & TH);
  cnt += popc(slide_raw<6>(obq & (~pinned | L.pinA), e) & TA);
  const u64 P = own & L.pawns;
  const u64 capD = P & (~pinned | L.pinD), capA = P & (~pinned | L.pinA);
  const u64 tgt = L.opp & L.cmask;
  if (p.stm) {
    cnt += popc(((capD >> 9) & NOTH) & tgt) + popc(((capA >> 7) & NOTA) & tgt);
  } else {
    cnt += popc(((capD << 9) & NOTA) & tgt) + popc(((capA << 7) & NOTH) & tgt);
  }
  if (!qs) {
    const u64 pp = P & (~pinned | L.pinF);
    u64 s1, s2;
    if (p.stm) {
      s1 = (pp >> 8) & e;
      s2 = ((s1 & RANK6) >> 8) & e;
    } else {
      s1 = (pp << 8) & e;
      s2 = ((s1 & RANK3) << 8) & e;
    }
    cnt += popc(s1 & L.cmask) + popc(s2 & L.cmask);
    cnt += popc(castle_targets(p, L));
  }
  int sqep;
  u64 cand = ep_candidates(p, L, sqep);
  while (cand) {
    const int f = lsb(cand);
    cand &= cand - 1;
    const int to = (p.stm ? sqep - 8 : sqep + 8) & 63;
    cnt += ep_legal(L, f, to, sqep) ? 1 : 0;
  }
  return cnt;
}

// Legal targets of the own piece on `sq` (used to enumerate moves in
// canonical order: origin ascending, then target ascending).
ZD u64 piece_targets(const Pos &p, const Legal &L, int sq, bool qs) {
  const u64 b = bit(sq);
  const u64 own = L.own, occ = L.occ;
  const u32 type = (u32)((shrm(p.b1, sq) & 1) | ((shrm(p.b2, sq) & 1) << 1) | ((shrm(p.b3, sq) & 1) << 2));
  if (type == KING) {
    u64 t = king_att(b) & (qs ? L.opp : ~own) & ~L.A;
    if (!qs) t |= castle_targets(p, L);
    return t;
  }
  u64 pinmask = FULL;
  if (L.pinR & b) pinmask = rank_line(L.ksq);
  if (L.pinF & b) pinmask = file_line(L.ksq);
  if (L.pinD & b) pinmask = diag_line(L.ksq);
  if (L.pinA & b) pinmask = anti_line(L.ksq);
  const u64 T = (qs ? L.opp : ~own) & L.cmask & pinmask;
  const u64 e = ~occ;
  u64 t = 0;
  switch (type) {
    case PAWN: {
      t = pawn_att(b, p.stm) & L.opp & L.cmask & pinmask;
      if (!qs) {
        u64 s1 = (p.stm ? (b >> 8) : (b << 8)) & e;
        u64 s2 = (p.stm ? ((s1 & RANK6) >> 8) : ((s1 & RANK3) << 8)) & e;
        t |= (s1 | s2) & L.cmask & pinmask;
      }
      int sqep;
      const u64 cand = ep_candidates(p, L, sqep);
      if (cand & b) {
        const int to = (p.stm ? sqep - 8 : sqep + 8) & 63;
        if (ep_legal(L, sq, to, sqep)) t |= bit(to);
      }
      return t;
    }
    case KNIGHT: return knight_att(b) & T;
    case BISHOP: return bishop_att(b, e) & T;
    case ROOK: return rook_att(b, e) & T;
    case QUEEN: return (rook_att(b, e) | bishop_att(b, e)) & T;
    default: return 0;
  }
}

// Index of the k-th set bit (k from 0) of x, which has more than k set bits:
// a binary search on popcounts, with no loop over the bits.
ZD int nth_bit(u64 x, int k) {
  u32 w = (u32)x;
  int base = 0;
  const int c32 = popc((u64)w);
  if (k >= c32) {
    k -= c32;
    w = (u32)(x >> 32);
    base = 32;
  }
#pragma unroll
  for (int s = 16; s >= 1; s >>= 1) {
    const int c = popc((u64)(w & ((1u << s) - 1u)));
    if (k >= c) {
      k -= c;
      w >>= s;
      base += s;
    }
  }
  return base;
}

// ---- node analysis with one lane per square (the search's wave-per-node
// layout).  The uniform part is cheap (king-square attackers); everything
// set-wise is replaced by per-square questions answered by all 64 lanes at
// once and combined with a ballot:
//   attacked(sq): is sq attacked by the opponent with the own king removed?
//   the legal targets of the own piece on sq (pin line, check mask, king
//   safety from the attacked-square mask).
struct Node0 {
  u64 own, opp, occ, kbb;
  u64 pawns, knights, kings, rq, bq;
  u64 checkers, cmask;
  int ksq;
  u32 stm;
};

ZD void node_basics(const Pos &p, Node0 &B) {
  const u64 occ = p.b1 | p.b2 | p.b3;
  const u64 white = p.b0 ^ occ;
  B.stm = p.stm;
  B.occ = occ;
  B.own = p.stm ? p.b0 : white;
  B.opp = p.stm ? white : p.b0;
  B.pawns = p.b1 & ~p.b2 & ~p.b3;
  B.knights = ~p.b1 & p.b2 & ~p.b3;
  B.kings = p.b1 & p.b2 & ~p.b3;
  B.rq = p.b3 & (p.b1 ^ p.b2);
  B.bq = p.b3 & ~p.b1;
  B.kbb = B.own & B.kings;
  B.ksq = B.kbb ? lsb(B.kbb) : 0;
  const u64 orq = B.opp & B.rq, obq = B.opp & B.bq;
  u64 ch = (rook_sq(B.ksq, occ) & orq) | (bishop_sq(B.ksq, occ) & obq);
  ch |= knight_att(B.kbb) & B.opp & B.knights;
  ch |= pawn_att(B.kbb, p.stm) & B.opp & B.pawns;
  ch = sel(B.kbb != 0, ch);
  B.checkers = ch;
  const int n = popc(ch);
  u64 cm = 0;
  if (n == 0) cm = FULL;
  if (n == 1) {
    // the checker, plus the squares between it and the king for a slider
    const int c = lsb(ch);
    const u64 r = rook_sq(B.ksq, occ) & rook_sq(c, occ);
    const u64 d = bishop_sq(B.ksq, occ) & bishop_sq(c, occ);
    const bool rl = (rank_line(B.ksq) & ch) || (file_line(B.ksq) & ch);
    const bool slider = ((B.rq | B.bq) & ch) != 0;
    cm = ch | sel(slider, rl ? r : d);
  }
  B.cmask = cm;
}

// is sq attacked by the opponent, the own king taken off the board?
ZD bool attacked_nk(const Node0 &B, int sq) {
  const u64 b = bit(sq);
  const u64 occ = B.occ ^ B.kbb;
  u64 a = (rook_sq(sq, occ) & B.rq) | (bishop_sq(sq, occ) & B.bq);
  a |= knight_att(b) & B.knights;
  a |= king_att(b) & B.kings;
  a |= pawn_att(b, B.stm) & B.pawns;
  return (a & B.opp) != 0;
}

// the line through the king and sq when the own piece on sq is pinned, else ~0
ZD u64 pin_line(const Node0 &B, int sq) {
  const int k = B.ksq;
  const u64 b = bit(sq);
  u64 m = 0, sl = 0;
  if ((sq >> 3) == (k >> 3)) { m = rank_line(k); sl = B.rq; }
  else if ((sq & 7) == (k & 7)) { m = file_line(k); sl = B.rq; }
  else if (diag_line(k) & b) { m = diag_line(k); sl = B.bq; }
  else if (anti_line(k) & b) { m = anti_line(k); sl = B.bq; }
  if (!m || !B.kbb) return FULL;
  const u64 me = m ^ bit(k);
  const u64 seen = line_att(k, B.occ, me);
  if (!(seen & b)) return FULL;  // not the first piece seen from the king
  const u64 behind = line_att(k, B.occ ^ b, me) & ~seen;
  return (behind & B.opp & sl) ? m : FULL;
}

ZD bool ep_legal0(const Node0 &B, int from, int to, int sqep) {
  const u64 occ2 = (B.occ ^ bit(from) ^ bit(sqep)) | bit(to);
  const u64 opp2 = B.opp & ~bit(sqep);
  u64 att = (rook_sq(B.ksq, occ2) & opp2 & B.rq) | (bishop_sq(B.ksq, occ2) & opp2 & B.bq);
  att |= knight_att(B.kbb) & opp2 & B.knights;
  att |= pawn_att(B.kbb, B.stm) & opp2 & B.pawns;
  att |= king_att(B.kbb) & opp2 & B.kings;
  return att == 0;
}

// legal targets of the own piece on sq; A = squares attacked with the own
// king transparent (only read for the king)
ZD u64 lane_targets(const Pos &p, const Node0 &B, u64 A, int sq, bool qs) {
  const u64 b = bit(sq);
  const u32 type = (u32)((shrm(p.b1, sq) & 1) | ((shrm(p.b2, sq) & 1) << 1) | ((shrm(p.b3, sq) & 1) << 2));
  const u64 own = B.own, occ = B.occ;
  if (type == KING) {
    u64 t = king_att(b) & (qs ? B.opp : ~own) & ~A;
    if (!qs && (p.cr & SMCRALL)) {
      const int k = sq;
      const bool okq = p.stm ? ((~p.cr & SMCRBLACKQ) == SMCRBLACKQ) : ((~p.cr & SMCRWHITEQ) == SMCRWHITEQ);
      const bool okk = p.stm ? ((~p.cr & SMCRBLACKK) == SMCRBLACKK) : ((~p.cr & SMCRWHITEK) == SMCRWHITEK);
      const u64 ownrooks = own & B.rq & ~p.b2;
      t |= sel(okq && (ownrooks & bit(k - 4)) && !(occ & (bit(k - 1) | bit(k - 2) | bit(k - 3))) &&
                   !(A & (bit(k) | bit(k - 1) | bit(k - 2))),
               bit(k - 2));
      t |= sel(okk && (ownrooks & bit(k + 3)) && !(occ & (bit(k + 1) | bit(k + 2))) &&
                   !(A & (bit(k) | bit(k + 1) | bit(k + 2))),
               bit(k + 2));
    }
    return t;
  }
  const u64 pm = pin_line(B, sq);
  const u64 T = (qs ? B.opp : ~own) & B.cmask & pm;
  u64 t;
  if (type == PAWN) {
    t = pawn_att(b, p.stm) & B.opp & B.cmask & pm;
    if (!qs) {
      const u64 e = ~occ;
      const u64 s1 = (p.stm ? (b >> 8) : (b << 8)) & e;
      const u64 s2 = (p.stm ? ((s1 & RANK6) >> 8) : ((s1 & RANK3) << 8)) & e;
      t |= (s1 | s2) & B.cmask & pm;
    }
    const u64 w = ~p.cr & RANKS45;
    const int sqep = w ? lsb(w) : 0;
    const bool cand = (p.stm ? (RANK4 & b) : (RANK5 & b)) && ((bit(sqep + 1) | bit(sqep - 1)) & b) && w;
    if (cand) {
      const int to = (p.stm ? sqep - 8 : sqep + 8) & 63;
      if (ep_legal0(B, sq, to, sqep)) t |= bit(to);
    }
    return t;
  }
  // pieces: branch-free over the type (no divergence across lanes)
  const bool rk = type == ROOK || type == QUEEN, bs = type == BISHOP || type == QUEEN;
  t = sel(rk, rook_sq(sq, occ)) | sel(bs, bishop_sq(sq, occ)) | sel(type == KNIGHT, knight_att(b));
  return t & T;
}

// Pack from->to with the reference movepicker rules (src/zetaperft.cl:1363-1386):
// e.p. capture square, captured piece, queen-only promotion (Q1).
ZD u32 make_code(const Pos &p, int from, int to) {
  const u32 pfrom = piece_at(p, from);
  int sqcpt = to;
  u32 pcpt = piece_at(p, to);
  const bool pawn = ((pfrom >> 1) & 7) == PAWN;
  if (pawn && p.stm && from - to != 8 && from - to != 16 && pcpt == PNONE) sqcpt = to + 8;
  if (pawn && !p.stm && to - from != 8 && to - from != 16 && pcpt == PNONE) sqcpt = to - 8;
  pcpt = piece_at(p, sqcpt);
  const int rrank = p.stm ? ((to >> 3) ^ 7) : (to >> 3);
  const u32 pto = (pawn && rrank == 7) ? ((QUEEN << 1) | (pfrom & 1)) : pfrom;
  return mv_make((u32)from, (u32)to, (u32)sqcpt, pfrom, pto, pcpt);
}

ZD void put_piece(Pos &c, int sq, u32 pc) {
  c.b0 |= shlm((u64)(pc & 1), sq);
  c.b1 |= shlm((u64)((pc >> 1) & 1), sq);
  c.b2 |= shlm((u64)((pc >> 2) & 1), sq);
  c.b3 |= shlm((u64)((pc >> 3) & 1), sq);
}

// Apply a move: pieces as the kernels' domove (src/zetaperft.cl:413-484), then
// the moved flags.  rook_flag selects the perft/host variant that also flags
// the castle rook's origin (src/zetaperft.cl:1452-1455) or the alpha-beta
// variant that does not (src/zeta.cl:2333-2345, quirk Q6).
ZD Pos apply(const Pos &p, u32 m, bool rook_flag) {
  Pos c = p;
  const int f = (int)mv_from(m), t = (int)mv_to(m), x = (int)mv_cpt(m);
  const u32 pf = mv_pfrom(m);
  const bool king = ((pf >> 1) & 7) == KING;
  const bool qside = king && f - t == 2, kside = king && t - f == 2;
  u64 keep = ~bit(f) & ~bit(x) & ~bit(t);
  if (qside) keep &= ~bit(f - 4);
  if (kside) keep &= ~bit(f + 3);
  c.b0 &= keep;
  c.b1 &= keep;
  c.b2 &= keep;
  c.b3 &= keep;
  put_piece(c, t, mv_pto(m));
  const u32 rook = (ROOK << 1) | (pf & 1);
  if (qside) put_piece(c, t + 1, rook);
  if (kside) put_piece(c, t - 1, rook);
  u64 cr = p.cr | bit(f) | bit(t) | bit(x);
  if (rook_flag && qside) cr |= bit(f - 4);
  if (rook_flag && kside) cr |= bit(f + 3);
  const u32 col = pf & 1;
  const int rf = col ? ((f >> 3) ^ 7) : (f >> 3), rt = col ? ((t >> 3) ^ 7) : (t >> 3);
  const bool dbl = ((pf >> 1) & 7) == PAWN && rt - rf == 2;
  cr |= RANKS45;
  if (dbl) cr &= ~bit(t);
  c.cr = cr;
  c.stm = p.stm ^ 1u;
  return c;
}

// Inverse of apply(): the parent board from the child board and the move;
// the parent's moved flags come from the caller (they are not invertible).
ZD Pos undo(const Pos &c, u32 m, u64 parent_cr) {
  Pos p = c;
  const int f = (int)mv_from(m), t = (int)mv_to(m), x = (int)mv_cpt(m);
  const u32 pf = mv_pfrom(m), pc = mv_pcpt(m);
  const bool king = ((pf >> 1) & 7) == KING;
  const bool qside = king && f - t == 2, kside = king && t - f == 2;
  u64 keep = ~bit(t);
  if (qside) keep &= ~bit(t + 1);
  if (kside) keep &= ~bit(t - 1);
  p.b0 &= keep;
  p.b1 &= keep;
  p.b2 &= keep;
  p.b3 &= keep;
  put_piece(p, f, pf);
  if (((pc >> 1) & 7) != PNONE) put_piece(p, x, pc);
  const u32 rook = (ROOK << 1) | (pf & 1);
  if (qside) put_piece(p, f - 4, rook);
  if (kside) put_piece(p, f + 3, rook);
  p.cr = parent_cr;
  p.stm = c.stm ^ 1u;
  return p;
}

// ---- Zobrist hashing, rotate-left scheme: key(piece) rotated left by the
// square (src/bitboard.c:461-508, src/zeta.cl:2297-2359).  12 piece keys
// (colour*6 + type-1), 4 castle keys, 1 e.p. key rotated by the file; the
// values are the reference's (src/zobrist.h:28-35) so hashes crossing the
// boundary stay comparable.
static ZC_CONST const u64 ZOBRIST[18] = {
    0x9D39247E33776D41ULL, 0x2AF7398005AAA5C7ULL, 0x44DB015024623547ULL, 0x9C15F73E62A76AE2ULL,
    0x75834465489C0C89ULL, 0x3290AC3A203001BFULL, 0x0FBBAD1F61042279ULL, 0xE83A908FF2FB60CAULL,
    0x0D7E765D58755C10ULL, 0x1A083822CEAFE02DULL, 0x9605D5F0E25EC3B0ULL, 0xD021FF5CD13A2ED5ULL,
    0x40BDF15D4A672E32ULL, 0x011355146FD56395ULL, 0x5DB4832046F3D9E5ULL, 0x239F8B2D7FF719CCULL,
    0x05D1A1AE85B49AA1ULL, 0x679F848F6E8FC971ULL};

ZD u64 rotl(u64 z, int s) { return shlm(z, s) | shrm(z, 64 - s); }
ZD u64 piece_key(u32 pc, int sq) { return rotl(ZOBRIST[(pc & 1) * 6 + ((pc >> 1) & 7) - 1], sq); }
ZD u64 hash_pieces(const Pos &p) {
  u64 h = 0;
  u64 occ = p.b1 | p.b2 | p.b3;
  while (occ) {
    int sq = lsb(occ);
    occ &= occ - 1;
    h ^= piece_key(piece_at(p, sq), sq);
  }
  return h;
}
ZD u64 hash_castle(u64 cr) {
  u64 h = 0;
  h ^= sel((~cr & SMCRWHITEQ) != 0, ZOBRIST[12]);
  h ^= sel((~cr & SMCRWHITEK) != 0, ZOBRIST[13]);
  h ^= sel((~cr & SMCRBLACKQ) != 0, ZOBRIST[14]);
  h ^= sel((~cr & SMCRBLACKK) != 0, ZOBRIST[15]);
  return h;
}
// host computehash (src/bitboard.c:461-508)
ZD u64 hash_full(const Pos &p) {
  u64 h = hash_pieces(p) ^ hash_castle(p.cr);
  u64 w = ~p.cr & RANKS45;
  int sqep = w ? lsb(w) : 0;
  if (sqep) h ^= rotl(ZOBRIST[16], sqep & 7);
  if (p.stm) h ^= 1ULL;
  return h;
}

// k-th legal move in canonical order (origin ascending, then target ascending)
ZD u32 nth_legal(const Pos &p, const Legal &L, int k) {
  u64 pcs = L.own;
  while (pcs) {
    const int sq = lsb(pcs);
    pcs &= pcs - 1;
    u64 t = piece_targets(p, L, sq, false);
    const int c = popc(t);
    if (k < c) {
      while (k-- > 0) t &= t - 1;
      return make_code(p, sq, lsb(t));
    }
    k -= c;
  }
  return 0;
}

// Origin of a slider move to `to` in direction d: the path from the origin is
// empty, so it is the nearest occupied square back along the direction.
ZD int slider_from(int d, int to, u64 occ) {
  const u64 b = bit(to), below = b - 1, above = ~(below | b);
  switch (d) {
    case 0: return msb(occ & file_line(to) & below);  // moved north
    case 1: return lsb(occ & file_line(to) & above);  // south
    case 2: return msb(occ & rank_line(to) & below);  // east
    case 3: return lsb(occ & rank_line(to) & above);  // west
    case 4: return msb(occ & diag_line(to) & below);  // north-east
    case 5: return msb(occ & anti_line(to) & below);  // north-west
    case 6: return lsb(occ & anti_line(to) & above);  // south-east
    default: return lsb(occ & diag_line(to) & above);  // south-west
  }
}

// Every legal move (the set count_legal(qs = false) counts, so exactly that
// many calls of emit(code)), grouped by kind instead of by origin: each
// kind's target set is built for all its pieces at once, as in count_legal,
// and each target gives back its origin -- a fixed offset for king, knight and
// pawn moves, slider_from() for sliders.  Codes equal make_code(p, from, to);
// the order is not the canonical one (nth_legal).
template <class Emit>
ZD void for_each_legal(const Pos &p, const Legal &L, Emit &&emit) {
  const u32 stm = p.stm;
  const u64 own = L.own, occ = L.occ, e = ~occ;
  const u64 T = ~own & L.cmask;
  const u64 TA = T & NOTA, TH = T & NOTH;
  const u64 pinned = L.pinR | L.pinF | L.pinD | L.pinA;
  // one kind with a fixed origin offset (from = to - delta) and moving piece pf
  auto fixed = [&](u64 t, int delta, u32 pf) {
    while (t) {
      const int to = lsb(t);
      t &= t - 1;
      emit(mv_make((u32)(to - delta), (u32)to, (u32)to, pf, pf, piece_at(p, to)));
    }
  };
  // king (castling included: the rook's move is implied by the code)
  {
    const u32 pk = (KING << 1) | stm;
    u64 t = (king_att(L.kbb) & ~own & ~L.A) | castle_targets(p, L);
    while (t) {
      const int to = lsb(t);
      t &= t - 1;
      emit(mv_make((u32)L.ksq, (u32)to, (u32)to, pk, pk, piece_at(p, to)));
    }
  }
  // knights, one offset at a time
  {
    const u64 kn = own & L.knights & ~pinned;
    const u64 TAB = T & 0xFCFCFCFCFCFCFCFCULL, TGH = T & 0x3F3F3F3F3F3F3F3FULL;
    const u32 pn = (KNIGHT << 1) | stm;
    fixed((kn << 17) & TA, 17, pn);
    fixed((kn << 15) & TH, 15, pn);
    fixed((kn << 10) & TAB, 10, pn);
    fixed((kn << 6) & TGH, 6, pn);
    fixed((kn >> 6) & TAB, -6, pn);
    fixed((kn >> 10) & TGH, -10, pn);
    fixed((kn >> 15) & TA, -15, pn);
    fixed((kn >> 17) & TH, -17, pn);
  }
  // sliders, one direction at a time
  {
    const u64 orq = own & L.rq, obq = own & L.bq;
    auto dir = [&](int d, u64 t) {
      while (t) {
        const int to = lsb(t);
        t &= t - 1;
        const int from = slider_from(d, to, occ);
        const u32 pf = piece_at(p, from);
        emit(mv_make((u32)from, (u32)to, (u32)to, pf, pf, piece_at(p, to)));
      }
    };
    dir(0, slide_raw<0>(orq & (~pinned | L.pinF), e) & T);
    dir(1, slide_raw<1>(orq & (~pinned | L.pinF), e) & T);
    dir(2, slide_raw<2>(orq & (~pinned | L.pinR), e) & TA);
    dir(3, slide_raw<3>(orq & (~pinned | L.pinR), e) & TH);
    dir(4, slide_raw<4>(obq & (~pinned | L.pinD), e) & TA);
    dir(7, slide_raw<7>(obq & (~pinned | L.pinD), e) & TH);
    dir(5, slide_raw<5>(obq & (~pinned | L.pinA), e) & TH);
    dir(6, slide_raw<6>(obq & (~pinned | L.pinA), e) & TA);
  }
  // pawns: pushes, double pushes, captures (queen-only promotion, Q1), e.p.
  {
    const u64 P = own & L.pawns;
    const u32 pp_ = (PAWN << 1) | stm, pq = (QUEEN << 1) | stm;
    const int last = stm ? 0 : 7;
    auto pawn = [&](u64 t, int delta) {
      while (t) {
        const int to = lsb(t);
        t &= t - 1;
        const u32 pt = (to >> 3) == last ? pq : pp_;
        emit(mv_make((u32)(to - delta), (u32)to, (u32)to, pp_, pt, piece_at(p, to)));
      }
    };
    const u64 pp = P & (~pinned | L.pinF);
    const u64 capD = P & (~pinned | L.pinD), capA = P & (~pinned | L.pinA);
    const u64 tgt = L.opp & L.cmask;
    if (stm) {
      const u64 s1 = (pp >> 8) & e, s2 = ((s1 & RANK6) >> 8) & e;
      pawn(s1 & L.cmask, -8);
      pawn(s2 & L.cmask, -16);
      pawn(((capD >> 9) & NOTH) & tgt, -9);
      pawn(((capA >> 7) & NOTA) & tgt, -7);
    } else {
      const u64 s1 = (pp << 8) & e, s2 = ((s1 & RANK3) << 8) & e;
      pawn(s1 & L.cmask, 8);
      pawn(s2 & L.cmask, 16);
      pawn(((capD << 9) & NOTA) & tgt, 9);
      pawn(((capA << 7) & NOTH) & tgt, 7);
    }
    int sqep;
    u64 cand = ep_candidates(p, L, sqep);
    while (cand) {
      const int f = lsb(cand);
      cand &= cand - 1;
      const int to = (stm ? sqep - 8 : sqep + 8) & 63;
      if (ep_legal(L, f, to, sqep)) emit(make_code(p, f, to));
    }
  }
}

// host domove half-move clock rule (src/bitboard.c:604-689): +1, reset by
// pawn moves, captures and castling
ZD u32 next_hmc(u32 hmc, u32 m) {
  const u32 pf = mv_pfrom(m);
  const int f = (int)mv_from(m), t = (int)mv_to(m);
  const u32 ty = (pf >> 1) & 7;
  if (ty == PAWN || ((mv_pcpt(m) >> 1) & 7) != PNONE) return 0;
  if (ty == KING && (f - t == 2 || t - f == 2)) return 0;
  return hmc + 1;
}

}  // namespace zc

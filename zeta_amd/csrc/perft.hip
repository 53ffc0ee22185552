// perft.hip -- batched perft on gfx950.
//
// Replaces perft_gpu (src/zetaperft.cl:801-1477), which walks ONE root with
// one 64-lane work-group, and its host driver (src/search.c:225-271,
// src/clrun.c:979-1147).  Design: a level-synchronous frontier over all roots
// of the batch, one lane per position:
//   level 0 = roots (SoA in HBM) with their legal-move counts;
//   count -> exclusive scan -> expand writes (parent, move) pairs of level L+1
//   -> apply materialises level L+1 (SoA) and counts its moves;
//   the last level is never materialised: perft_leaf applies each move to its
//   parent and adds the child's legal-move count (bulk counting at depth-1)
//   to its root's u64 counter with one wavefront-aggregated atomic per root.
// Every leaf-level lane does the same branch-free set-wise counting, so the
// uneven subtree sizes of a batch cost no SIMD divergence.  Levels that would
// exceed the frontier budget are split into chunks (depth-first over chunks).
#include <string.h>

#include <vector>

#include "chess.h"
#include "runtime.h"
#include "wave.h"

using namespace zc;

namespace {

struct Front {  // a frontier level, structure of arrays (coalesced per field)
  u64 *b0, *b1, *b2, *b3, *cr;
  u32 *meta;  // root index << 1 | side to move
  u32 *cnt;   // number of legal moves
};

__device__ __forceinline__ Pos load_pos(const Front &f, size_t i) {
  Pos p;
  p.b0 = f.b0[i];
  p.b1 = f.b1[i];
  p.b2 = f.b2[i];
  p.b3 = f.b3[i];
  p.cr = f.cr[i];
  p.stm = f.meta[i] & 1u;
  return p;
}

__device__ __forceinline__ void store_pos(const Front &f, size_t i, const Pos &p, u32 root) {
  f.b0[i] = p.b0;
  f.b1[i] = p.b1;
  f.b2[i] = p.b2;
  f.b3[i] = p.b3;
  f.cr[i] = p.cr;
  f.meta[i] = (root << 1) | p.stm;
}

constexpr int TPB = 256;

__global__ __launch_bounds__(TPB) void k_roots(const u64 *__restrict__ in, int n, Front f) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const u64 *q = in + (size_t)i * 8;  // zeta_pos = 7 x u64 + stm + pad
  Pos p;
  p.b0 = q[0];
  p.b1 = q[1];
  p.b2 = q[2];
  p.b3 = q[3];
  p.cr = q[4];
  p.stm = (u32)(q[7] & 1u);
  store_pos(f, i, p, (u32)i);
  Legal L;
  analyse(p, L);
  f.cnt[i] = (u32)count_legal(p, L, false);
}

// ---- exclusive scan of u32 counts into u64 offsets (off[n] = total)
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = TPB * SCAN_ITEMS;

__device__ __forceinline__ u64 block_exscan(u64 v, u64 *lds, u64 &total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  u64 x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u64 y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  u64 wo = 0;
  for (int k = 0; k < w; k++) wo += lds[k];
  total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return wo + x - v;
}

__global__ __launch_bounds__(TPB) void k_scan_reduce(const u32 *__restrict__ cnt, size_t n, u64 *bsum) {
  __shared__ u64 lds[4];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  u64 s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
    if (base + k < n) s += cnt[base + k];
  u64 tot;
  block_exscan(s, lds, tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(TPB) void k_scan_top(u64 *bsum, size_t nb, u64 *total_out) {
  __shared__ u64 lds[4];
  u64 carry = 0;
  for (size_t t0 = 0; t0 < nb; t0 += TPB) {
    const size_t i = t0 + threadIdx.x;
    u64 v = i < nb ? bsum[i] : 0;
    u64 tot;
    u64 ex = block_exscan(v, lds, tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(TPB) void k_scan_down(const u32 *__restrict__ cnt, size_t n, const u64 *bsum,
                                                   u64 *off) {
  __shared__ u64 lds[4];
  const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
  u32 v[SCAN_ITEMS];
  u64 s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    v[k] = base + k < n ? cnt[base + k] : 0u;
    s += v[k];
  }
  u64 tot;
  u64 ex = block_exscan(s, lds, tot) + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    if (base + k < n) off[base + k] = ex;
    ex += v[k];
  }
  // the thread covering element n-1 also writes the total into off[n]
  if (base < n && base + SCAN_ITEMS >= n) off[n] = ex;
}

// ---- expand: (parent, move) pairs of every legal move, at the scanned offsets.
// A block's parents own one contiguous range of pairs.  Each lane enumerates
// its parent's moves into LDS (move code + the lane that owns the parent),
// then the block writes the range out with coalesced 8-byte stores: stored
// straight from the enumeration loop, the pairs of 64 lanes land on 64
// different cache lines per store instruction.  Moves beyond the LDS capacity
// (a block whose parents have more than XCAP moves in all) are stored directly.
constexpr int XTPB = 128;
constexpr int XCAP = 4096;  // 16 KiB of codes + 4 KiB of lane indices per block
__global__ __launch_bounds__(XTPB) void k_expand(Front f, size_t i0, size_t i1, const u64 *__restrict__ off,
                                                 u64 base, u64 *__restrict__ pm, const u64 *__restrict__ gtotal,
                                                 u64 gcap) {
  __shared__ u32 lcode[XCAP];
  __shared__ u8 llane[XCAP];
  // launched before the host knows the level's pair total: the whole grid
  // leaves when the total does not fit the pair buffer (the host relaunches)
  if (gtotal && *gtotal > gcap) return;
  const size_t bi0 = i0 + (size_t)blockIdx.x * XTPB;
  const size_t bi1 = bi0 + XTPB < i1 ? bi0 + XTPB : i1;
  const u64 blo = off[bi0], total = off[bi1] - blo;
  u64 *const out = pm + (blo - base);
  const size_t i = bi0 + threadIdx.x;
  if (i < bi1) {
    const Pos p = load_pos(f, i);
    Legal L;
    analyse(p, L);
    u64 o = off[i] - blo;
    const u64 parent = (u64)i << 32;
    const u8 lane = (u8)threadIdx.x;
    for_each_legal(p, L, [&](u32 code) {
      if (o < (u64)XCAP) {
        lcode[o] = code;
        llane[o] = lane;
      } else {
        out[o] = parent | code;
      }
      o++;
    });
  }
  __syncthreads();
  const u32 n = total < (u64)XCAP ? (u32)total : (u32)XCAP;
  const u32 last = (u32)(bi1 - bi0) - 1;  // every slot below n has an owner; the clamp only bounds the index
  for (u32 j = threadIdx.x; j < n; j += XTPB) {
    const u32 l = llane[j] < last ? llane[j] : last;
    out[j] = ((u64)(bi0 + l) << 32) | lcode[j];
  }
}

// ---- chunk cuts of an oversized level: bucket b holds the parents whose pairs
// start in [b * bp, (b + 1) * bp), so a chunk has at most bp + 218 pairs (the
// most legal moves of a position).  cut_idx[b] = first parent of bucket b
// (relative to the level range), cut_off[b] = its offset; buckets past the
// last parent get (n, total).  Every b in [0, nb] is written exactly once.
__global__ __launch_bounds__(TPB) void k_cuts(const u64 *__restrict__ off, size_t n, u64 bp, u64 nb,
                                              u64 *__restrict__ cut_idx, u64 *__restrict__ cut_off) {
  const size_t i = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const u64 b = off[i] / bp;
  const u64 first = i ? off[i - 1] / bp + 1 : 0;
  for (u64 j = first; j <= b && j <= nb; j++) {
    cut_idx[j] = i;
    cut_off[j] = off[i];
  }
  if (i + 1 == n)
    for (u64 j = b + 1; j <= nb; j++) {
      cut_idx[j] = n;
      cut_off[j] = off[n];
    }
}

// ---- apply: materialise the next level and count its moves
__global__ __launch_bounds__(TPB) void k_apply(Front f, const u64 *__restrict__ pm,
                                               size_t m, Front g) {
  const size_t j = (size_t)blockIdx.x * TPB + threadIdx.x;
  if (j >= m) return;
  const u64 e = pm[j];
  const u32 pi = (u32)(e >> 32);
  const Pos p = load_pos(f, pi);
  const Pos c = apply(p, (u32)e, true);
  store_pos(g, j, c, f.meta[pi] >> 1);
  Legal L;
  analyse(c, L);
  g.cnt[j] = (u32)count_legal(c, L, false);
}

// sum over the wave (all lanes active); a wave's per-root share is at most
// 64 x 218 legal moves, so 32-bit DPP sums suffice
__device__ __forceinline__ u64 wave_sum(u32 v) {
  return (u64)(u32)__builtin_amdgcn_readlane(zw::wave_scan_add((int)v), 63);
}

// ---- leaf level: apply + count, accumulate per root
__global__ __launch_bounds__(TPB) void k_leaf(Front f, const u64 *__restrict__ pm,
                                              size_t m, u64 *__restrict__ rootcnt, int nobulk) {
  const size_t j = (size_t)blockIdx.x * TPB + threadIdx.x;
  bool active = j < m;
  u64 v = 0;
  u32 root = 0;
  if (active) {
    const u64 e = pm[j];
    const u32 pi = (u32)(e >> 32);
    const Pos p = load_pos(f, pi);
    const Pos c = apply(p, (u32)e, true);
    Legal L;
    analyse(c, L);
    if (nobulk) {
      // the reference visits every leaf: it generates the leaf's (quiescence)
      // moves and counts one node (src/zetaperft.cl:1316-1326)
      const int n = count_legal(c, L, true);
      v = 1ULL + (u64)(n > 4096);
    } else {
      v = (u64)count_legal(c, L, false);
    }
    root = f.meta[pi] >> 1;
  }
  // wavefront-aggregated accumulation: one atomic per distinct root in the wave
  const int lane = threadIdx.x & 63;
  u64 pending = __ballot(active);
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    const u32 rl = __shfl(root, leader, 64);
    const bool mine = active && root == rl;
    const u64 mask = __ballot(mine);
    const u64 s = wave_sum(mine ? (u32)v : 0u);
    if (lane == leader) atomicAdd((unsigned long long *)&rootcnt[rl], (unsigned long long)s);
    pending &= ~mask;
    active = active && !mine;
  }
}

__global__ void k_u32_to_u64(const u32 *a, u64 *b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill_u64(u64 *b, int n, u64 v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = v;
}

// ---- playouts (workload generator), one lane per position
__global__ __launch_bounds__(TPB) void k_playout(u64 seed, int n, int minply, int maxply, u64 *out,
                                                 u32 *fails) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  Pos s0;  // the initial position, setboard of the standard FEN with KQkq
  s0.b0 = 0xFFFF000000000000ULL;
  s0.b1 = 0x91FF00000000FF91ULL;
  s0.b2 = 0x5A0000000000005AULL;
  s0.b3 = 0xAD000000000000ADULL;
  s0.cr = 0x6EFFFFFFFFFFFF6EULL;
  s0.stm = 0;
  for (u64 a = 0; a <= 1000; a++) {
    u64 st = seed ^ ((u64)i * 0x9E3779B97F4A7C15ULL) ^ (a * 0xD1B54A32D192ED03ULL);
    auto rnd = [&st]() {
      u64 z = (st += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return z ^ (z >> 31);
    };
    const int plies = minply + (int)(rnd() % (u64)(maxply - minply + 1));
    Pos p = s0;
    u32 hmc = 0;
    bool ok = true;
    for (int k = 0; k < plies; k++) {
      Legal L;
      analyse(p, L);
      const int nl = count_legal(p, L, false);
      if (nl == 0 || hmc >= 100) {
        ok = false;
        break;
      }
      const int pick = (int)(rnd() % (u64)nl);
      const u32 m = nth_legal(p, L, pick);
      hmc = next_hmc(hmc, m);
      p = apply(p, m, true);
    }
    if (ok) {
      Legal L;
      analyse(p, L);
      if (count_legal(p, L, false) == 0 || hmc >= 100) ok = false;
    }
    if (ok) {
      u64 *q = out + (size_t)i * 8;
      q[0] = p.b0;
      q[1] = p.b1;
      q[2] = p.b2;
      q[3] = p.b3;
      q[4] = p.cr;
      q[5] = hash_full(p);
      q[6] = hmc;
      q[7] = p.stm;
      return;
    }
  }
  atomicAdd(fails, 1u);
}

// ---- host driver
constexpr int MAXLEV = 16;

struct LevelBufs {
  zr::DevBuf b0, b1, b2, b3, cr, meta, cnt;  // SoA positions
  zr::DevBuf off, bsum;                      // scan
  zr::DevBuf pm;                             // (parent index << 32 | move) pairs of the next level
  size_t cap = 0, paircap = 0;
};

struct PerftState {
  LevelBufs lev[MAXLEV];
  zr::DevBuf total, roots, counts, fails, cuts;
};
PerftState g_ps;

// ---- the allocation-free entry point's scratch array: the frontier of a
// call to zeta_perft_batch_device_noalloc lives here instead of in buffers
// the library allocates.  One array per device (the whole code object is
// loaded per device), held until the process ends; every call on a device
// reuses it, so calls that overlap on one device race.
constexpr size_t SCRATCH_BYTES = 256ULL << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
unsigned char *g_scratch_dev[SCRATCH_MAX_DEVICES];  // hipGetSymbolAddress of scratch, per device
hipError_t g_scratch_err = hipSuccess;
constexpr int NCUT_FIXED = 511;  // chunk cuts a level can take in the scratch array

// The buffers one perft call works in: the library's growable ones, or views
// into the scratch array with fixed capacities (never grown, never freed).
struct PerftBufs {
  LevelBufs *lev;
  zr::DevBuf *total, *cuts;
  u64 budget;  // most pairs a chunk may hold
  bool fixed;
};
PerftBufs g_pb;
LevelBufs g_fixed_lev[MAXLEV];
zr::DevBuf g_fixed_total, g_fixed_cuts;

void use_growable_bufs() {
  g_pb.lev = g_ps.lev;
  g_pb.total = &g_ps.total;
  g_pb.cuts = &g_ps.cuts;
  g_pb.budget = zr::ctx().cfg.frontier_budget;
  g_pb.fixed = false;
}

int ensure_level(int L, size_t n) {
  LevelBufs &b = g_pb.lev[L];
  if (b.cap >= n && b.cap) return ZETA_OK;
  if (g_pb.fixed)
    return zr::fail(ZETA_ENOMEM, "perft scratch: level %d needs %zu positions, holds %zu", L, n, b.cap);
  size_t cap = n + n / 4 + 1024;
  int r;
  if ((r = zr::ensure(b.b0, cap * 8)) || (r = zr::ensure(b.b1, cap * 8)) || (r = zr::ensure(b.b2, cap * 8)) ||
      (r = zr::ensure(b.b3, cap * 8)) || (r = zr::ensure(b.cr, cap * 8)) || (r = zr::ensure(b.meta, cap * 4)) ||
      (r = zr::ensure(b.cnt, cap * 4)) || (r = zr::ensure(b.off, (cap + 1) * 8)) ||
      (r = zr::ensure(b.bsum, (cap / SCAN_TILE + 2) * 8)))
    return r;
  b.cap = cap;
  return ZETA_OK;
}

int ensure_pairs(int L, size_t m) {
  LevelBufs &b = g_pb.lev[L];
  if (b.paircap >= m && b.paircap) return ZETA_OK;
  if (g_pb.fixed)
    return zr::fail(ZETA_ENOMEM, "perft scratch: level %d needs %zu pairs, holds %zu", L, m, b.paircap);
  size_t cap = m + m / 8 + 1024;
  int r;
  if ((r = zr::ensure(b.pm, cap * 8))) return r;
  b.paircap = cap;
  return ZETA_OK;
}

Front front_of(int L) {
  LevelBufs &b = g_pb.lev[L];
  return Front{(u64 *)b.b0.p, (u64 *)b.b1.p, (u64 *)b.b2.p, (u64 *)b.b3.p,
               (u64 *)b.cr.p, (u32 *)b.meta.p, (u32 *)b.cnt.p};
}

inline unsigned blocks(size_t n, int per) { return (unsigned)((n + per - 1) / per); }

// exclusive scan of level L's counts [i0, i1) into lev[L].off[0..n]; returns total
int scan_counts(int L, size_t i0, size_t i1, u64 *total) {
  LevelBufs &b = g_pb.lev[L];
  const size_t n = i1 - i0;
  const size_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  const u32 *cnt = (const u32 *)b.cnt.p + i0;
  if (n == 0) {
    *total = 0;
    return ZETA_OK;
  }
  ZLAUNCH(zr::K_PERFT_SCAN, n, k_scan_reduce, (unsigned)nb, TPB, 0, cnt, n, (u64 *)b.bsum.p);
  ZLAUNCH(zr::K_PERFT_SCAN, n, k_scan_top, 1, TPB, 0, (u64 *)b.bsum.p, nb, (u64 *)g_pb.total->p);
  ZLAUNCH(zr::K_PERFT_SCAN, n, k_scan_down, (unsigned)nb, TPB, 0, cnt, n, (const u64 *)b.bsum.p, (u64 *)b.off.p);
  // into the pinned word zeta_init set aside for it
  ZCHECK(hipMemcpyAsync(zr::ctx().perft_total_host, g_pb.total->p, 8, hipMemcpyDeviceToHost, zr::ctx().stream));
  ZCHECK(hipEventRecord(zr::ctx().ev_total, zr::ctx().stream));
  *total = ~0ULL;  // known after wait_total()
  return ZETA_OK;
}

// the total copied by scan_counts (the stream may run on: only the copy is waited for)
int wait_total(u64 *total) {
  ZCHECK(hipEventSynchronize(zr::ctx().ev_total));
  *total = *zr::ctx().perft_total_host;
  return ZETA_OK;
}

int process(int L, size_t i0, size_t i1, int D, u64 *d_counts, int nobulk) {
  u64 total = 0;
  int r = scan_counts(L, i0, i1, &total);
  if (r) return r;
  if (i1 == i0) return ZETA_OK;
  const u64 budget = g_pb.budget;
  LevelBufs &lv = g_pb.lev[L];
  // the expansion is queued before the total is known and writes into the
  // current pair buffer when the total fits it (k_expand checks), so the
  // host's wait for the total overlaps the expansion instead of preceding it
  const u64 spec_cap = lv.paircap < budget ? lv.paircap : budget;
  if (spec_cap)
    ZLAUNCH(zr::K_PERFT_EXPAND, i1 - i0, k_expand, blocks(i1 - i0, XTPB), XTPB, 0, front_of(L), i0, i1,
            (const u64 *)lv.off.p - i0, (u64)0, (u64 *)lv.pm.p, (const u64 *)g_pb.total->p, spec_cap);
  if ((r = wait_total(&total))) return r;
  if (total == 0) return ZETA_OK;
  // a level with more pairs than the budget runs in chunks of parents, each
  // with at most `budget` pairs; the cuts come from the device (k_cuts) so the
  // host reads a few words, not the level's offsets
  if (total > budget) {
    if (budget < 512) return zr::fail(ZETA_EINVAL, "frontier budget %llu is below 512 moves", (unsigned long long)budget);
    const u64 bp = budget - 256;
    const u64 nb = total / bp + 1;
    const size_t n = i1 - i0;
    if (g_pb.fixed) {
      if (nb > (u64)NCUT_FIXED)
        return zr::fail(ZETA_ENOMEM, "perft scratch: level %d needs %llu chunks, holds %d", L,
                        (unsigned long long)nb, NCUT_FIXED);
    } else if ((r = zr::ensure(g_ps.cuts, (size_t)(nb + 1) * 16))) {
      return r;
    }
    u64 *d_idx = (u64 *)g_pb.cuts->p, *d_off = d_idx + nb + 1;
    ZLAUNCH(zr::K_PERFT_SCAN, n, k_cuts, blocks(n, TPB), TPB, 0, (const u64 *)lv.off.p, n, bp, nb, d_idx, d_off);
    // the cuts come back into a vector, or into a fixed array on the stack
    // when the call may not allocate
    std::vector<u64> cutv;
    u64 cutf[2 * (NCUT_FIXED + 1)];
    u64 *cut = cutf;
    if (!g_pb.fixed) {
      cutv.resize(2 * (nb + 1));
      cut = cutv.data();
    }
    ZCHECK(hipMemcpyAsync(cut, d_idx, 2 * (nb + 1) * 8, hipMemcpyDeviceToHost, zr::ctx().stream));
    ZCHECK(hipStreamSynchronize(zr::ctx().stream));
    const u64 *cidx = cut, *coff = cut + nb + 1;
    for (u64 c = 0; c < nb; c++) {
      const size_t a0 = (size_t)cidx[c], a1 = (size_t)cidx[c + 1];
      if (a1 <= a0) continue;
      const u64 m = coff[c + 1] - coff[c];
      if (m == 0) continue;
      if ((r = ensure_pairs(L, m))) return r;
      ZLAUNCH(zr::K_PERFT_EXPAND, a1 - a0, k_expand, blocks(a1 - a0, XTPB), XTPB, 0, front_of(L), i0 + a0, i0 + a1,
              (const u64 *)lv.off.p - i0, coff[c], (u64 *)lv.pm.p, (const u64 *)nullptr, (u64)0);
      if (L + 1 == D - 1) {
        ZLAUNCH(zr::K_PERFT_LEAF, m, k_leaf, blocks(m, TPB), TPB, 0, front_of(L), (const u64 *)lv.pm.p,
                (size_t)m, d_counts, nobulk);
      } else {
        if ((r = ensure_level(L + 1, m))) return r;
        ZLAUNCH(zr::K_PERFT_APPLY, m, k_apply, blocks(m, TPB), TPB, 0, front_of(L),
                (const u64 *)lv.pm.p, (size_t)m, front_of(L + 1));
        if ((r = process(L + 1, 0, m, D, d_counts, nobulk))) return r;
      }
    }
    return ZETA_OK;
  }
  if (!spec_cap || total > spec_cap) {
    if ((r = ensure_pairs(L, total))) return r;
    ZLAUNCH(zr::K_PERFT_EXPAND, i1 - i0, k_expand, blocks(i1 - i0, XTPB), XTPB, 0, front_of(L), i0, i1,
            (const u64 *)lv.off.p - i0, (u64)0, (u64 *)lv.pm.p, (const u64 *)nullptr, (u64)0);
  }
  if (L + 1 == D - 1) {
    ZLAUNCH(zr::K_PERFT_LEAF, total, k_leaf, blocks(total, TPB), TPB, 0, front_of(L), (const u64 *)lv.pm.p,
            (size_t)total, d_counts, nobulk);
    return ZETA_OK;
  }
  if ((r = ensure_level(L + 1, total))) return r;
  ZLAUNCH(zr::K_PERFT_APPLY, total, k_apply, blocks(total, TPB), TPB, 0, front_of(L), (const u64 *)lv.pm.p,
          (size_t)total, front_of(L + 1));
  return process(L + 1, 0, total, D, d_counts, nobulk);
}

int perft_device(const void *d_pos, int n, int depth, u64 *d_counts, int flags) {
  const int nobulk = (flags & ZETA_PERFT_NOBULK) ? 1 : 0;
  if (n < 0 || depth < 0) return zr::fail(ZETA_EINVAL, "perft: n=%d depth=%d", n, depth);
  if (n == 0) return ZETA_OK;
  const int D = nobulk ? depth + 1 : depth;
  if (D >= MAXLEV) return zr::fail(ZETA_EINVAL, "perft depth %d exceeds %d", depth, MAXLEV - 2);
  int r;
  if (!g_pb.fixed && (r = zr::ensure(g_ps.total, 64))) return r;
  if (depth == 0) {
    ZLAUNCH(zr::K_PERFT_ROOTS, n, k_fill_u64, blocks(n, TPB), TPB, 0, d_counts, n, (u64)1);
    return ZETA_OK;
  }
  if ((r = ensure_level(0, (size_t)n))) return r;
  ZLAUNCH(zr::K_PERFT_ROOTS, n, k_roots, blocks(n, TPB), TPB, 0, (const u64 *)d_pos, n, front_of(0));
  if (D == 1) {
    ZLAUNCH(zr::K_PERFT_ROOTS, n, k_u32_to_u64, blocks(n, TPB), TPB, 0, (const u32 *)g_pb.lev[0].cnt.p, d_counts,
            n);
    return ZETA_OK;
  }
  ZLAUNCH(zr::K_PERFT_ROOTS, n, k_fill_u64, blocks(n, TPB), TPB, 0, d_counts, n, (u64)0);
  return process(0, 0, (size_t)n, D, d_counts, nobulk);
}

}  // namespace

void perft_release() {
  for (auto &l : g_ps.lev) {
    zr::release(l.b0); zr::release(l.b1); zr::release(l.b2); zr::release(l.b3);
    zr::release(l.cr); zr::release(l.meta); zr::release(l.cnt); zr::release(l.off);
    zr::release(l.bsum); zr::release(l.pm);
    l.cap = l.paircap = 0;
  }
  zr::release(g_ps.total);
  zr::release(g_ps.roots);
  zr::release(g_ps.counts);
  zr::release(g_ps.fails);
  zr::release(g_ps.cuts);
}

extern "C" int zeta_perft_batch_device(const void *d_pos, int n, int depth, void *d_counts, int flags) {
  int r = zr::require_up();
  if (r) return r;
  if ((!d_pos || !d_counts) && n > 0) return zr::fail(ZETA_EINVAL, "perft: null pointer");
  use_growable_bufs();
  r = perft_device(d_pos, n, depth, (u64 *)d_counts, flags);
  hipError_t e = hipStreamSynchronize(zr::ctx().stream);
  zr::time_collect();
  if (r) return r;
  if (e != hipSuccess) return zr::hip_fail(e, "perft sync", __FILE__, __LINE__);
  return ZETA_OK;
}

// zeta_init, with the device current: where this device holds the scratch
// array (the array's name in host code is only a stand-in for it)
int perft_setup() {
  const int dev = zr::ctx().device;
  if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return zr::fail(ZETA_ENODEV, "perft: device %d beyond %d", dev,
                                                             SCRATCH_MAX_DEVICES);
  if (g_scratch_dev[dev]) return ZETA_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    // the other entry points do not need the array: the allocation-free one
    // reports this error when it is called
    (void)hipGetLastError();
    g_scratch_err = e != hipSuccess ? e : hipErrorInvalidSymbol;
    return ZETA_OK;
  }
  g_scratch_dev[dev] = (unsigned char *)ptr;
  return ZETA_OK;
}

// Views of the scratch array for a call of D - 1 frontier levels: every level
// holds up to C positions (5 x u64 planes, meta, count, offset, scan sums)
// and C pairs; C is the largest that fits, and it is the chunk budget, so no
// level and no chunk ever needs more.  Returns C, 0 when it is below what
// the roots need.
size_t fixed_layout(unsigned char *base, int D, size_t n) {
  const int nl = D - 1 > 1 ? D - 1 : 1;
  const size_t align = 256, head = 2 * align + 2 * (NCUT_FIXED + 1) * 8;
  // per level: 9 arrays, each rounded up to the alignment
  const size_t slack = 9 * align + 8 * 3;
  const size_t per_pos = 5 * 8 + 4 + 4 + 8 + 8;  // planes, meta, cnt, off, pairs
  size_t C = (SCRATCH_BYTES - head) / nl;
  C = C > slack ? (C - slack) / per_pos : 0;
  C -= C / (SCAN_TILE / 8) + 2;  // room for the scan sums (8 B per SCAN_TILE positions)
  if (C < 1024 || C < n) return 0;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    void *p = base + o;
    o += (bytes + align - 1) / align * align;
    return p;
  };
  g_fixed_total.p = take(64);
  g_fixed_total.bytes = 64;
  g_fixed_cuts.p = take(2 * (NCUT_FIXED + 1) * 8);
  g_fixed_cuts.bytes = 2 * (NCUT_FIXED + 1) * 8;
  for (int L = 0; L < MAXLEV; L++) {
    LevelBufs &b = g_fixed_lev[L];
    b = LevelBufs();
    if (L >= nl) continue;
    zr::DevBuf *planes[5] = {&b.b0, &b.b1, &b.b2, &b.b3, &b.cr};
    for (auto *q : planes) {
      q->p = take(C * 8);
      q->bytes = C * 8;
    }
    b.meta.p = take(C * 4);
    b.meta.bytes = C * 4;
    b.cnt.p = take(C * 4);
    b.cnt.bytes = C * 4;
    b.off.p = take((C + 1) * 8);
    b.off.bytes = (C + 1) * 8;
    b.bsum.p = take((C / SCAN_TILE + 2) * 8);
    b.bsum.bytes = (C / SCAN_TILE + 2) * 8;
    b.pm.p = take(C * 8);
    b.pm.bytes = C * 8;
    b.cap = C;
    b.paircap = C;
  }
  if (o > SCRATCH_BYTES) return 0;
  return C;
}

extern "C" int zeta_perft_batch_device_noalloc(const void *d_pos, int n, int depth, void *d_counts, int flags) {
  int r = zr::require_up();
  if (r) return r;
  if ((!d_pos || !d_counts) && n > 0) return zr::fail(ZETA_EINVAL, "perft (no-alloc): null pointer");
  if (n < 0 || n > ZETA_NOALLOC_MAX_ROOTS)
    return zr::fail(ZETA_EINVAL, "perft (no-alloc): %d roots, at most %d", n, ZETA_NOALLOC_MAX_ROOTS);
  if (depth < 0 || depth > ZETA_NOALLOC_MAX_DEPTH)
    return zr::fail(ZETA_EINVAL, "perft (no-alloc): depth %d, at most %d", depth, ZETA_NOALLOC_MAX_DEPTH);
  const int dev = zr::ctx().device;
  unsigned char *base = (dev >= 0 && dev < SCRATCH_MAX_DEVICES) ? g_scratch_dev[dev] : nullptr;
  if (!base)
    return zr::fail(ZETA_ENODEV, "perft (no-alloc): no scratch array on device %d (hipGetSymbolAddress: %s)", dev,
                    hipGetErrorString(g_scratch_err));
  const int D = (flags & ZETA_PERFT_NOBULK) ? depth + 1 : depth;
  const size_t C = fixed_layout(base, D, (size_t)n);
  if (!C) return zr::fail(ZETA_EINVAL, "perft (no-alloc): %d roots at depth %d do not fit the scratch array", n, depth);
  g_pb.lev = g_fixed_lev;
  g_pb.total = &g_fixed_total;
  g_pb.cuts = &g_fixed_cuts;
  g_pb.budget = C;
  g_pb.fixed = true;
  // no kernel timing here: its events come from a pool that may grow
  const bool timing = zr::ctx().timing;
  zr::ctx().timing = false;
  r = perft_device(d_pos, n, depth, (u64 *)d_counts, flags);
  zr::ctx().timing = timing;
  use_growable_bufs();
  const hipError_t e = hipStreamSynchronize(zr::ctx().stream);
  if (r) return r;
  if (e != hipSuccess) return zr::hip_fail(e, "perft (no-alloc) sync", __FILE__, __LINE__);
  return ZETA_OK;
}

extern "C" int zeta_perft_batch(const zeta_pos *pos, int n, int depth, uint64_t *counts) {
  int r = zr::require_up();
  if (r) return r;
  if (n < 0 || (n > 0 && (!pos || !counts))) return zr::fail(ZETA_EINVAL, "perft: bad arguments");
  if (n == 0) return ZETA_OK;
  if ((r = zr::ensure(g_ps.roots, (size_t)n * sizeof(zeta_pos))) || (r = zr::ensure(g_ps.counts, (size_t)n * 8)))
    return r;
  ZCHECK(hipMemcpyAsync(g_ps.roots.p, pos, (size_t)n * sizeof(zeta_pos), hipMemcpyHostToDevice, zr::ctx().stream));
  use_growable_bufs();
  r = perft_device(g_ps.roots.p, n, depth, (u64 *)g_ps.counts.p, 0);
  if (r) {
    (void)hipStreamSynchronize(zr::ctx().stream);
    zr::time_collect();
    return r;
  }
  ZCHECK(hipMemcpyAsync(counts, g_ps.counts.p, (size_t)n * 8, hipMemcpyDeviceToHost, zr::ctx().stream));
  ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  zr::time_collect();
  return ZETA_OK;
}

extern "C" int zeta_playouts(uint64_t seed, int n, int minply, int maxply, zeta_pos *out) {
  int r = zr::require_up();
  if (r) return r;
  if (n < 0 || minply < 0 || maxply < minply || (n > 0 && !out))
    return zr::fail(ZETA_EINVAL, "playouts: bad arguments");
  if (n == 0) return ZETA_OK;
  if ((r = zr::ensure(g_ps.roots, (size_t)n * sizeof(zeta_pos))) || (r = zr::ensure(g_ps.fails, 4))) return r;
  ZCHECK(hipMemsetAsync(g_ps.fails.p, 0, 4, zr::ctx().stream));
  ZLAUNCH(zr::K_PLAYOUT, n, k_playout, blocks(n, TPB), TPB, 0, seed, n, minply, maxply, (u64 *)g_ps.roots.p,
          (u32 *)g_ps.fails.p);
  u32 fails = 0;
  ZCHECK(hipMemcpyAsync(out, g_ps.roots.p, (size_t)n * sizeof(zeta_pos), hipMemcpyDeviceToHost, zr::ctx().stream));
  ZCHECK(hipMemcpyAsync(&fails, g_ps.fails.p, 4, hipMemcpyDeviceToHost, zr::ctx().stream));
  ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  zr::time_collect();
  if (fails) return zr::fail(ZETA_EINVAL, "playouts: %u positions found no acceptable playout", fails);
  return ZETA_OK;
}

// search.hip -- batched fixed-depth alpha-beta + quiescence on gfx950.
//
// Replaces alphabeta_gpu (src/zeta.cl:852-2512) and its host driver
// rootsearch (src/search.c:35-223), for many root positions per launch.
//
// Execution model: one wavefront = one worker; a persistent grid of workers
// takes root positions from a queue and, once the queue is empty, joins
// unfinished roots as ABDADA helpers (busy siblings deferred to a second pass,
// src/zeta.cl:1648-1717).  The first worker to finish a root's last
// iteration publishes its score; the others abandon that root.
// Per worker, the depth-first stack lives in LDS: one Ply record per ply and a
// move stack.  Node control is wave-uniform (every lane holds the same
// values); lanes split the per-node work:
//   * move generation: lane = square, each lane enumerates its piece's legal
//     targets; a wave prefix sum compacts them into the LDS move stack;
//   * move selection: wave arg-max over the node's remaining moves
//     (score, then lowest slot, so runs are deterministic per worker);
//   * evaluation: lane = square (src/zeta.cl:1482-1531), wave sum.
//
// Node semantics follow alphabeta_gpu with null move, LMR and IID off:
// check / queen-promotion extension (src/zeta.cl:1330-1345), quiescence at
// depth <= 0 with captures + e.p. only and stand pat when not in check,
// mate = -INF+ply only outside quiescence, stalemate 0, in-path repetition,
// fifty-move and bare-king draws below the root, MAXPLY cut, fail-hard
// window propagation, root score = PV[0] semantics.
//
// Transposition table (2^k entries x 16 B in HBM, lockless: word0 =
// key ^ word1): every entry supplies the move ordering's TT move; its score
// is used to raise alpha only when it cannot change the root value: the
// entry was stored at the same depth by a node whose half-move clock was 0
// and the probing node's clock is 0 too (then neither subtree can see a
// repetition above itself), and the key includes the exact castle squares.
#include <string.h>

#include "chess.h"
#include "evaltab.h"
#include "mlp.h"
#include "runtime.h"
#include "wave.h"

using namespace zc;

namespace {

constexpr int MAXPLY = 64;  // src/types.h:76
constexpr s32 INF = 32000;
constexpr s32 MATESCORE = 30000;
constexpr int MSTACK = 1024;  // move-stack slots per worker (LDS)
constexpr s32 USED = -32768;  // 16-bit move-ordering scores
constexpr s32 TTMOVE_SCORE = 32000;  // above every other score (src/zeta.cl:2101)
constexpr s32 HINT_SCORE = 32100;    // the hint table's move, above the TT move
// the killer (one per ply) and the counter move (indexed by the previous
// move's squares), quiet moves only: src/zeta.cl:2073-2082
constexpr s32 KILLER_SCORE = 1400;
constexpr s32 COUNTER_SCORE = 1300;
constexpr int NCOUNTER = 64 * 64;  // counter-move table entries per worker (LDS)
constexpr u32 DEFERRED = 0x80000000u;  // move-stack flag: deferred by ABDADA (move codes use bits 0-29)
constexpr int ABDADA_MIN_DEPTH = 2;  // default split depth (zeta_search_opts.split_depth)
constexpr int POLL_EVERY = 64;

typedef u32 u32x4 __attribute__((ext_vector_type(4)));  // one 16-byte TT entry

// ---- serving (zeta_serve_*): the host publishes requests in coherent pinned
// host memory; a persistent grid of workers claims them in ticket order,
// helps running ones (ABDADA) when idle, and writes results back there.
struct ServeReq {  // 80 B
  u64 pos[8];      // zeta_pos
  s32 depth;
  s32 flags;       // bit 0: learned ordering
  s32 workers;     // at most this many workers on the request
  s32 pad;
};
struct ServeSlot {  // 128 B, pinned host memory
  ServeReq req;     // written by the host before the ticket is published
  s32 score;        // results, written by the device
  u32 bestmove;
  u64 nodes;
  u64 ticket;       // ticket + 1 of the result
  s32 depth_done;
  s32 status;       // SV_DONE / SV_FAILED / SV_CANCELLED
  u32 state;        // 0 -> 1 when the result is complete (device store, host clears)
  u32 pad[3];
};
struct ServeHost {  // pinned host memory header, then u64 queue[qcap], ack[qcap], cancel[nslots]
  u64 submitted;    // tickets published; ticket t names slot queue[t % qcap]
  u64 heartbeat;    // bumped by every host call (watchdog)
  u32 shutdown;
  u32 cancel_gen;   // bumped when cancel[] changes
  u32 pad[10];
};
struct ServeDev {  // device memory; mirrors of the host words, written by the scheduler wavefront
  u64 claimed;     // tickets claimed by workers
  u64 sub;         // mirror of ServeHost::submitted
  u32 shut;        // 1 shutdown, 2 watchdog, 3 protocol error
  u32 exit_xcc;    // diagnostics of a watchdog / protocol exit
  u64 exit_now;
  u64 exit_arg;
};
// slot word (device): ticket+1 << 24 | max workers << 16 | workers << 2 | status
enum { SV_RUN = 0, SV_DONE = 1, SV_FAILED = 2, SV_CANCELLED = 3 };

// batch counters (device, summed per worker at the end of each root)
struct SearchArgs {
  const u64 *roots;  // zeta_pos[n]
  int nroots;
  int depth;
  int workers;       // at most this many workers on one root
  int split_depth;   // ABDADA deferral / busy marking at depth - 1 >= split_depth
  u64 *tt;           // 2 words per entry
  u64 ttmask;        // entries - 1 (0 = no TT)
  const u64 *hint;   // read-only table in the TT's format whose moves go first (null = none)
  u64 *rbound;       // per root (slot): best root score so far, ticket << 32 | iteration << 16 | score + 32768
  u32 *busy;
  u32 busymask;
  int *queue;
  u32 *done;
  int *nworkers;
  s32 *score;
  u32 *bestmove;
  u64 *nodes;
  s32 *depth_done;
  u32 *err;
  const u32 *habort;  // abort word in coherent pinned host memory (zeta_search_abort)
  u64 max_nodes;  // per worker and root (0 = none)
  u64 *ctr;       // batch counters, see zeta_search_counters
  // learned ordering
  int use_mlp;
  int mlp_min_depth;
  zm::Weights w;
  // training-pair log (cut nodes): PairRec[cap], counter
  u64 *recs;
  u64 *rec_count;
  u64 rec_cap;
  u32 rec_mask;  // a cut node is logged when (hash >> 40) & rec_mask == 0
  // game history before each root, newest first (ZETA_HISTORY per root)
  const u64 *hist;
  const s32 *hlen;
  // serving mode (slots != nullptr)
  ServeHost *sh;
  const u64 *squeue;  // host: slot of ticket t at [t % qcap]
  u64 *sack;          // host: ticket + 1 once claimed
  const u64 *scancel; // host: ticket + 1 to cancel, per slot
  ServeSlot *slots;   // host
  int nslots;
  int qcap;
  ServeDev *sv;       // device
  u64 *sw;            // device slot words
  ServeReq *dreq;     // device copies of the claimed requests
  u64 *dres;          // per slot: score | best move << 32 of the worker that completed the request
  u64 *dcancel;       // device mirror of scancel
  u64 watchdog_ticks; // wall-clock ticks without a heartbeat before the grid exits (0 = never)
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum { C_TT_PROBE = 0, C_TT_HIT, C_TT_CUT, C_TT_STORE, C_DEFER, C_MAXPLY, C_MLP_NODES, C_RECORDS, C_NODES, C_NCTR = 16 };

// One ply of a worker's stack in LDS (64 B).  The board itself is not kept:
// a parent board is recovered with undo() when a child returns.
struct Ply {
  u64 cr, hash, phash;
  s32 alpha, beta, depth;
  u32 move, best, lastmove;
  u16 mbeg, mcnt;
  u8 stm, hmc, qs, kic, iter, ndefer, busy, searched;
  u32 ttm;  // the node's TT move (written at its analysis, read by the pair log)
};

__device__ __forceinline__ void wsync() { __syncthreads(); }  // one wave per block

// Wave-uniform values: readfirstlane puts them in SGPRs, so the node logic
// (position analysis, window and TT bookkeeping) runs on the scalar unit.
__device__ __forceinline__ u32 uf(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ s32 uf(s32 v) { return (s32)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ u64 uf(u64 v) { return ((u64)uf((u32)(v >> 32)) << 32) | (u64)uf((u32)v); }

// add v to statistics counter k, held in lane k of a vector register
__device__ __forceinline__ void lcount(u32 &cvec, int k, u32 v) { cvec += (threadIdx.x == (unsigned)k) ? v : 0u; }
__device__ __forceinline__ s32 wave_sum_i(s32 v) {
  return (s32)__builtin_amdgcn_readlane(zw::wave_scan_add(v), 63);
}
__device__ __forceinline__ u32 wave_max_u32(u32 v) {
  return (u32)__builtin_amdgcn_readlane((int)zw::wave_scan_max(v), 63);
}
__device__ __forceinline__ int wave_exscan_i(int v, int lane, int &total) {
  const int x = zw::wave_scan_add(v);
  total = __builtin_amdgcn_readlane(x, 63);
  return x - v;
}

// TT / busy-table key: the reference hash plus the exact castle squares
// (a1 e1 h1 a8 e8 h8, bits 0 4 7 56 60 63 of the moved-flags word), folded
// into 16 bits and spread by one multiplication by an odd constant (17
// fewer scalar instructions per node than a full 64-bit mix, about 1 % of a
// node's time)
__device__ __forceinline__ u64 tt_key(u64 hash, u64 cr) {
  const u64 c = cr & 0x9100000000000091ULL;
  const u32 x = (u32)c | (u32)(c >> 48);
  return hash ^ ((u64)x * 0x9E3779B97F4A7C15ULL);
}

__device__ __forceinline__ bool is_mate(s32 v) { return (v > MATESCORE && v < INF) || (v < -MATESCORE && v > -INF); }

__device__ __forceinline__ bool is_promo(u32 m) {
  return m && ((mv_pfrom(m) >> 1) & 7) == PAWN && ((mv_pto(m) >> 1) & 7) == QUEEN;
}

struct Node {  // the current node (wave-uniform)
  Pos p;
  u64 hash, phash;
  s32 alpha, beta, depth;
  u32 move, best, lastmove;
  int mbeg, mcnt;
  u32 hmc;
  bool qs, kic;
  int iter, ndefer, searched;
  bool busy;
};

__device__ __forceinline__ void save(Ply &y, const Node &n) {
  y.cr = n.p.cr;
  y.hash = n.hash;
  y.phash = n.phash;
  y.alpha = n.alpha;
  y.beta = n.beta;
  y.depth = n.depth;
  y.move = n.move;
  y.best = n.best;
  y.lastmove = n.lastmove;
  y.mbeg = (u16)n.mbeg;
  y.mcnt = (u16)n.mcnt;
  y.stm = (u8)n.p.stm;
  y.hmc = (u8)n.hmc;
  y.qs = n.qs;
  y.kic = n.kic;
  y.iter = (u8)n.iter;
  y.ndefer = (u8)n.ndefer;
  y.busy = n.busy;
  y.searched = (u8)n.searched;
}
// parent node from its LDS record and the returning child's board
__device__ __forceinline__ void restore(const Ply &y, const Pos &child, Node &n) {
  n.hash = uf(y.hash);
  n.phash = uf(y.phash);
  n.alpha = uf(y.alpha);
  n.beta = uf(y.beta);
  n.depth = uf(y.depth);
  n.move = uf(y.move);
  n.best = uf(y.best);
  n.lastmove = uf(y.lastmove);
  n.mbeg = (int)uf((u32)y.mbeg);
  n.mcnt = (int)uf((u32)y.mcnt);
  n.hmc = uf((u32)y.hmc);
  n.qs = uf((u32)y.qs) != 0;
  n.kic = uf((u32)y.kic) != 0;
  n.iter = (int)uf((u32)y.iter);
  n.ndefer = (int)uf((u32)y.ndefer);
  n.busy = uf((u32)y.busy) != 0;
  n.searched = (int)uf((u32)y.searched);
  n.p = undo(child, n.move, uf(y.cr));
}

// child hash of the alpha-beta move-up (src/zeta.cl:2297-2359), incremental
__device__ __forceinline__ void child_hash(const Node &n, u32 m, const Pos &c, u64 &hash, u64 &phash) {
  const int f = (int)mv_from(m), t = (int)mv_to(m), x = (int)mv_cpt(m);
  const u32 pf = mv_pfrom(m), pt = mv_pto(m), pc = mv_pcpt(m);
  u64 ph = n.phash ^ piece_key(pf, f) ^ piece_key(pt, t);
  if (((pc >> 1) & 7) != PNONE) ph ^= piece_key(pc, x);
  if (((pf >> 1) & 7) == KING && (f - t == 2 || t - f == 2)) {
    const u32 rook = (ROOK << 1) | (pf & 1);
    const int rf = (f - t == 2) ? f - 4 : f + 3, rt = (f - t == 2) ? t + 1 : t - 1;
    ph ^= piece_key(rook, rf) ^ piece_key(rook, rt);
  }
  u64 h = ph;
  const u32 col = pf & 1;
  const int rfr = col ? ((f >> 3) ^ 7) : (f >> 3), rto = col ? ((t >> 3) ^ 7) : (t >> 3);
  if (((pf >> 1) & 7) == PAWN && rto - rfr == 2) h ^= rotl(ZOBRIST[16], t & 7);
  if (c.stm) h ^= 1ULL;
  h ^= hash_castle(c.cr);
  hash = h;
  phash = ph;
}

__device__ __forceinline__ u32 child_hmc(u32 hmc, u32 m) {
  // src/zeta.cl:2362-2373 (8-bit counter)
  const u32 pf = mv_pfrom(m);
  const int f = (int)mv_from(m), t = (int)mv_to(m);
  if (((pf >> 1) & 7) == PAWN || ((mv_pcpt(m) >> 1) & 7) != PNONE) return 0;
  if (((pf >> 1) & 7) == KING && (f - t == 2 || t - f == 2)) return 0;
  return (hmc + 1) & 0xFF;
}

enum { S_ENTER = 0, S_NEXT = 1, S_RETURN = 2 };

// TT store when a full-width node returns: with a move that raised alpha, a
// lower bound of the node's value and that move; with none (every move
// failed low), an upper bound and no move (bit 58).  The score is kept as
// distance-to-mate from this node.  Either bound is the value of the
// depth-limited subtree, which a node with a zero half-move clock cannot see
// depend on the path above it: entries are used at the same depth and clock
// only, so no root value depends on them.
__device__ __forceinline__ u32 tt_store(const SearchArgs &a, const Node &n, int sd, int lane) {
  if (!a.ttmask || n.qs || n.depth < 1) return 0;
  s32 s = n.alpha;
  const int plyn = sd - 1;
  if (is_mate(s) && s > 0) s += plyn;
  if (is_mate(s) && s < 0) s -= plyn;
  const u64 key = tt_key(n.hash, n.p.cr);
  const u64 w1 = (u64)n.best | ((u64)(u16)(int16_t)s << 32) | ((u64)(n.depth & 0xFF) << 48) |
                 ((u64)(n.hmc == 0 ? 1 : 0) << 56) | (1ULL << 57) | ((u64)(n.best ? 0 : 1) << 58);
  if (lane == 0) {
    const u64 w0 = key ^ w1;
    u32x4 e;
    e.x = (u32)w0;
    e.y = (u32)(w0 >> 32);
    e.z = (u32)w1;
    e.w = (u32)(w1 >> 32);
    *(u32x4 *)(a.tt + 2 * (key & a.ttmask)) = e;  // one 16-byte vector store
  }
  return 1;
}

// A counter move is a quiet move, which its from- and to-square and its
// piece determine (sqcpt = to, no captured piece, pto = pfrom or the queen a
// pawn on the last rank becomes): 16 bits hold it, and a quiet candidate
// equals the stored move exactly when its 16 bits do (a capture never does,
// as in the reference's full-code comparison, src/zeta.cl:2072)
__device__ __forceinline__ u16 counter_code(u32 m) { return (u16)((m & 0xFFFu) | (mv_pfrom(m) << 12)); }
__device__ __forceinline__ int counter_index(u32 prev) { return (int)(mv_from(prev) * 64 + mv_to(prev)); }

__device__ __forceinline__ void busy_release(const SearchArgs &a, u64 hash, u64 cr, int lane) {
  const u32 bidx = (u32)(tt_key(hash, cr) >> 40) & a.busymask;
  if (lane == 0) atomicSub(&a.busy[bidx], 1u);
}

// next root for this worker: the queue first, then the unfinished root with
// the fewest workers (ABDADA helper), at most a.workers per root
// A worker takes the next unstarted root; once all have started it joins,
// as an ABDADA helper below the per-root limit, the unfinished root with the
// fewest completed iterations (the most work left: an iteration costs
// several times the one before it), then the fewest workers.  The wave's
// lanes scan the roots together and agree on the minimum.
__device__ int pick_root(const SearchArgs &a, int hstart, int lane) {
  int r = 0;
  if (lane == 0) r = atomicAdd(a.queue, 1);
  r = uf(r);
  if (r < a.nroots) {
    if (lane == 0) atomicAdd(&a.nworkers[r], 1);
    return r;
  }
  if (a.workers <= 1) return -1;
  for (int attempt = 0; attempt < 8; attempt++) {
    u64 best = ~0ULL;  // (iterations done << 48) | (workers << 32) | root
    for (int k = lane; k < a.nroots; k += 64) {
      const int rr = (hstart + k) % a.nroots;
      if (__atomic_load_n(&a.done[rr], __ATOMIC_RELAXED)) continue;
      const int w = __atomic_load_n(&a.nworkers[rr], __ATOMIC_RELAXED);
      if (w >= a.workers) continue;
      const int dd = __atomic_load_n(&a.depth_done[rr], __ATOMIC_RELAXED);
      const u64 key = ((u64)(dd < 0 ? 0 : (dd > 255 ? 255 : dd)) << 48) | ((u64)(u32)w << 32) | (u32)rr;
      best = key < best ? key : best;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const u64 other = ((u64)(u32)__shfl_xor((int)(u32)(best >> 32), o, 64) << 32) |
                        (u32)__shfl_xor((int)(u32)best, o, 64);
      best = other < best ? other : best;
    }
    best = uf(best);
    if (best == ~0ULL) return -1;
    const int pick = (int)(u32)best;
    int got = 0;
    if (lane == 0) {
      got = atomicAdd(&a.nworkers[pick], 1) < a.workers ? 1 : 0;
      if (!got) atomicSub(&a.nworkers[pick], 1);
    }
    if (uf(got)) return pick;
  }
  return -1;
}

__device__ __forceinline__ void add_row(float x[8], const u16 *row) {
  const uint4 q = *(const uint4 *)row;  // 8 BF16
  const u32 v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    x[2 * j] += zm::bf2f((u16)(v[j] & 0xFFFF));
    x[2 * j + 1] += zm::bf2f((u16)(v[j] >> 16));
  }
}

// Learned ordering: network score of every move of the node (moves
// [mbeg, mbeg+total) in LDS) written into mscore.  Layer 1's position part is
// accumulated once (lane = 4 hidden units); each chunk of 32 moves then runs
// layer 2 as 16 MFMA steps whose A operand (move x 8 hidden units, BF16) is
// built in registers from the position part plus the move's 3-4 feature rows.
__device__ void mlp_score_moves(const SearchArgs &a, const Pos &p, int mbeg, int total, int lane, float *hpos,
                                const u32 *mmove, int16_t *mscore) {
  const zm::Weights &w = a.w;
  float h4[4];
#pragma unroll
  for (int j = 0; j < 4; j++) h4[j] = w.b1[lane * 4 + j];
  u64 occ = p.b1 | p.b2 | p.b3;
  while (occ) {
    const int sq = lsb(occ);
    occ &= occ - 1;
    const int f = zm::feat_ps(piece_at(p, sq), sq, p.stm);
    const uint2 q = *(const uint2 *)(w.w1f + (size_t)f * zm::NH1 + lane * 4);
    h4[0] += zm::bf2f((u16)(q.x & 0xFFFF));
    h4[1] += zm::bf2f((u16)(q.x >> 16));
    h4[2] += zm::bf2f((u16)(q.y & 0xFFFF));
    h4[3] += zm::bf2f((u16)(q.y >> 16));
  }
#pragma unroll
  for (int j = 0; j < 4; j++) hpos[lane * 4 + j] = h4[j];
  wsync();
  const int col = lane & 31, half = lane >> 5;
  const float b2o = w.b2[col], w3o = w.w3[col], b3 = w.b3[0];
  for (int c0 = 0; c0 < total; c0 += 32) {
    const int mi = c0 + col;
    const u32 m = mi < total ? mmove[mbeg + mi] : 0u;
    const int ff = m ? zm::feat_from(m, p.stm) : 0, ft = m ? zm::feat_to(m, p.stm) : 0;
    const int fp = m ? zm::feat_piece(m) : 0, fc = m ? zm::feat_capt(m) : -1;
    f32x16 acc = {};
    for (int ks = 0; ks < zm::NH1 / 16; ks++) {
      const int kb = ks * 16 + half * 8;
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; j++) x[j] = hpos[kb + j];
      if (m) {
        add_row(x, w.w1f + (size_t)ff * zm::NH1 + kb);
        add_row(x, w.w1f + (size_t)ft * zm::NH1 + kb);
        add_row(x, w.w1f + (size_t)fp * zm::NH1 + kb);
        if (fc >= 0) add_row(x, w.w1f + (size_t)fc * zm::NH1 + kb);
      }
      bf16x8 av;
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const u16 hb = zm::f2bf(x[j] > 0.f ? x[j] : 0.f);
        av[j] = __builtin_bit_cast(__bf16, hb);
      }
      const bf16x8 bv = *(const bf16x8 *)(w.w2t + (size_t)col * zm::NH1 + kb);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    // layer 3: register i of lane l holds z2[move 8*(i/4)+4*half+i%4][unit col]
    float t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const float z = acc[i] + b2o;
      t[i] = z > 0.f ? w3o * z : 0.f;
    }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1)
#pragma unroll
      for (int i = 0; i < 16; i++) t[i] += __shfl_xor(t[i], o, 64);
    if (col == 0) {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = 8 * (i >> 2) + 4 * half + (i & 3);
        if (c0 + row < total) {
          float sc = (t[i] + b3) * 256.f;
          sc = sc > 31000.f ? 31000.f : (sc < -31000.f ? -31000.f : sc);
          mscore[mbeg + c0 + row] = (int16_t)__float2int_rn(sc);
        }
      }
    }
  }
  wsync();
}

// the plain ordering's score of a move (src/zeta.cl:2059-2082): MVV-LVA or
// the piece-square delta, the killer's 1400 and the counter move's 1300
__device__ __forceinline__ s32 plain_score(u32 stm, u32 m, u32 killer, u32 counter) {
  s32 sc = order_score(stm, m);
  sc = sc > 31999 ? 31999 : (sc < -31999 ? -31999 : sc);
  if (((mv_pcpt(m) >> 1) & 7) == PNONE) {
    if (m == killer) sc = KILLER_SCORE;
    else if (counter_code(m) == counter) sc = COUNTER_SCORE;
  }
  return sc;
}

// Training pairs at a cut node of the full-width search whose cutoff move is
// not the TT move (the TT move goes first whatever the network says): the
// cutoff move against every other move of the node -- those searched before
// it, which failed (scores marked USED), and those never searched, which the
// plain ordering ranked below it.  Each pair carries the plain ordering's
// score difference (negative minus positive), which the ranking loss adds to
// the network's, so that the network learns what to add to the plain
// ordering.  Wave-compacted.
__device__ void log_pairs(const SearchArgs &a, const Node &n, int lane, const u32 *mmove, const int16_t *mscore,
                          u32 ttm, u32 killer, u32 counter, u32 &nrec) {
  if (n.move == ttm) return;
  const s32 ppos = plain_score(n.p.stm, n.move, killer, counter);
  for (int c0 = 0; c0 < n.mcnt; c0 += 64) {
    const int k = c0 + lane;
    bool neg = false;
    u32 m = 0;
    if (k < n.mcnt) {
      m = mmove[n.mbeg + k] & ~DEFERRED;
      neg = m != n.move && m != ttm;
    }
    const u64 mask = __ballot(neg);
    if (!mask) continue;
    const u32 cnt = (u32)popc(mask);
    u64 base = 0;
    if (lane == 0) base = atomicAdd((unsigned long long *)a.rec_count, (unsigned long long)cnt);
    base = uf(base);
    if (neg) {
      const u64 slot = base + (u64)popc(mask & ((1ULL << lane) - 1));
      if (slot < a.rec_cap) {
        s32 dp = plain_score(n.p.stm, m, killer, counter) - ppos;
        dp = dp > zm::PAIR_OFFSET_MAX ? zm::PAIR_OFFSET_MAX : (dp < -zm::PAIR_OFFSET_MAX ? -zm::PAIR_OFFSET_MAX : dp);
        u64 *rp = a.recs + slot * 6;
        rp[0] = n.p.b0;
        rp[1] = n.p.b1;
        rp[2] = n.p.b2;
        rp[3] = n.p.b3;
        rp[4] = (u64)n.p.stm | ((u64)n.move << 32);
        rp[5] = (u64)m | ((u64)zm::pair_depth_word(n.depth, dp) << 32);
      }
    }
    nrec += cnt;
  }
}

// Serving synchronisation polls with relaxed atomics only: on this chip an
// acquire or release at agent scope costs an L2 invalidate / write-back (the
// L2 is per XCD), which, issued by polling workers, evicts every other
// worker's transposition-table and weight lines.  Relaxed agent- and
// system-scope atomics are coherent without cache maintenance.  Ordering is
// paid only where data is handed over, once per request or join: the release
// that opens a claimed request, the acquire of a helper's join, and the
// release that publishes a finished result to the host.
__device__ __forceinline__ u64 sys_load64(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ u32 sys_load32(const u32 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T> __device__ __forceinline__ T ag_load(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void ag_store(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ bool ag_cas(T *p, T expect, T v) {
  return __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void sys_store(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ u64 sw_tk(u64 v) { return v >> 24; }
__device__ __forceinline__ int sw_max(u64 v) { return (int)((v >> 16) & 0xFF); }
__device__ __forceinline__ int sw_w(u64 v) { return (int)((v >> 2) & 0x3FFF); }
__device__ __forceinline__ int sw_st(u64 v) { return (int)(v & 3); }

// Serving: a slot's node count and deepest iteration carry the low 16 bits of
// their request's ticket (count: tag << 48 | nodes; depth: tag << 16 |
// iteration), and a worker adds to them only while the tag is its request's,
// in one compare-and-swap: a late helper of an earlier request can neither
// write into the slot's next request nor see its own numbers mixed with it.
__device__ __forceinline__ u64 slot_tag(u64 tk) { return tk & 0xFFFFULL; }
__device__ __forceinline__ void slot_add_nodes(u64 *p, u64 tk, u64 add) {
  u64 v = ag_load(p);
  while ((v >> 48) == slot_tag(tk)) {
    const u64 nv = (v & ~0xFFFFFFFFFFFFULL) | (((v & 0xFFFFFFFFFFFFULL) + add) & 0xFFFFFFFFFFFFULL);
    u64 e = v;
    if (__hip_atomic_compare_exchange_strong(p, &e, nv, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
    v = e;
  }
}
__device__ __forceinline__ void slot_max_depth(s32 *p, u64 tk, s32 id) {
  s32 v = ag_load(p);
  while ((u64)((u32)v >> 16) == slot_tag(tk) && (v & 0xFFFF) < id) {
    s32 e = v;
    const s32 nv = (s32)((u32)(slot_tag(tk) << 16) | (u32)id);
    if (__hip_atomic_compare_exchange_strong(p, &e, nv, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      return;
    v = e;
  }
}

__device__ __forceinline__ u32 xcc_id() {
  // HW_REG_XCC_ID (hwreg 20), bits [3:0]
  return (u32)__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) & 15u;
}

// The scheduler wavefront (the grid's last block) never searches: every few
// microseconds it mirrors the host words (ticket count, shutdown, cancel
// generation) into device memory, so that the workers poll device memory
// only, and it runs the watchdog on its own clock.
__device__ void serve_scheduler(const SearchArgs &a, int lane) {
  ServeDev *v = a.sv;
  u64 hb_seen = 0, hb_time = 0;
  u32 cg_seen = 0;
  for (;;) {
    u32 shut = 0, cg = 0;
    if (lane == 0) {
      const u64 now = wall_clock64();
      const u64 sub = sys_load64(&a.sh->submitted);
      const u64 hb = sys_load64(&a.sh->heartbeat);
      shut = sys_load32(&a.sh->shutdown) ? 1u : 0u;
      cg = sys_load32(&a.sh->cancel_gen);
      if (hb != hb_seen || hb_time == 0) {
        hb_seen = hb;
        hb_time = now;
      } else if (a.watchdog_ticks && now - hb_time > a.watchdog_ticks) {
        shut = 2;
        ag_store(&v->exit_now, now);
        ag_store(&v->exit_arg, hb_time);
        ag_store(&v->exit_xcc, xcc_id());
      }
      ag_store(&v->sub, sub);
    }
    cg = uf(cg);
    if (cg != cg_seen) {
      for (int s = lane; s < a.nslots; s += 64) ag_store(&a.dcancel[s], sys_load64(&a.scancel[s]));
      cg_seen = cg;
    }
    shut = uf(shut);
    if (shut) {
      if (lane == 0) ag_store(&v->shut, shut);
      return;
    }
    __builtin_amdgcn_s_sleep(127);
  }
}

// serving: claim the next published ticket, else join the running request
// with the fewest workers; returns its slot and ticket+1, or -1 on shutdown
__device__ int serve_pick(const SearchArgs &a, int lane, u64 &tk_out) {
  int backoff = 1;
  for (;;) {
    int res = -2;
    u64 tk = 0;
    if (lane == 0) {
      if (ag_load(&a.sv->shut)) {
        res = -1;
      } else {
        const u64 c = ag_load(&a.sv->claimed);
        if (c < ag_load(&a.sv->sub) && ag_cas(&a.sv->claimed, c, c + 1)) {
          const int qi = (int)(c % (u64)a.qcap);
          const u64 s = sys_load64(&a.squeue[qi]);
          if (s >= (u64)a.nslots) {
            ag_store(&a.sv->exit_now, c);
            ag_store(&a.sv->exit_arg, s);
            ag_store(&a.sv->shut, 3u);  // the host broke the protocol: stop serving
            res = -1;
          } else {
            ServeReq &d = a.dreq[s];
            const ServeReq &h = a.slots[s].req;
            for (int k = 0; k < 8; k++) ag_store(&d.pos[k], sys_load64(&h.pos[k]));
            const s32 hw = (s32)sys_load32((const u32 *)&h.workers);
            ag_store(&d.depth, (s32)sys_load32((const u32 *)&h.depth));
            ag_store(&d.flags, (s32)sys_load32((const u32 *)&h.flags));
            ag_store(&d.workers, hw);
            sys_store(&a.sack[qi], c + 1);
            ag_store(&a.nodes[s], slot_tag(c + 1) << 48);
            ag_store(&a.depth_done[s], (s32)(u32)(slot_tag(c + 1) << 16));
            const u64 wmax = (u64)(hw < 1 ? 1 : (hw > 255 ? 255 : hw));
            // release: the request copy above is visible to every helper
            // that joins through this word (one write-back per request)
            __hip_atomic_store(&a.sw[s], (u64)(((c + 1) << 24) | (wmax << 16) | (u64)4 | (u64)SV_RUN),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            res = (int)s;
            tk = c + 1;
          }
        }
      }
    }
    res = uf(res);
    if (res != -2) {
      tk_out = uf(tk);
      return res;
    }
    // no new request: the running request with the fewest workers below its maximum
    u32 bestkey = 0;
    for (int base = 0; base < a.nslots; base += 64) {
      const int s = base + lane;
      if (s < a.nslots) {
        const u64 v = ag_load(&a.sw[s]);
        const int w = sw_w(v);
        // a cancelled request takes no new helpers: its workers drain and the
        // last one publishes the cancellation
        if (sw_tk(v) && sw_st(v) == SV_RUN && w >= 1 && w < sw_max(v) && ag_load(&a.dcancel[s]) != sw_tk(v))
          bestkey = max(bestkey, ((u32)(255 - w) << 16) | (u32)(0xFFFF - s));
      }
    }
    bestkey = wave_max_u32(bestkey);
    if (bestkey) {
      const int s = 0xFFFF - (int)(bestkey & 0xFFFF);
      int got = -1;
      if (lane == 0) {
        const u64 v = ag_load(&a.sw[s]);
        u64 e = v;
        if (sw_tk(v) && sw_st(v) == SV_RUN && sw_w(v) >= 1 && sw_w(v) < sw_max(v) &&
            ag_load(&a.dcancel[s]) != sw_tk(v) &&
            __hip_atomic_compare_exchange_strong(&a.sw[s], &e, v + 4, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {  // acquire: once per join
          got = s;
          tk = sw_tk(v);
        }
      }
      got = uf(got);
      if (got >= 0) {
        tk_out = uf(tk);
        return got;
      }
    } else {
      // nothing to do: back off (up to ~50 us) so idle workers stay off the
      // memory system the searching ones use
      for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(127);
      backoff = backoff < 16 ? 2 * backoff : 16;
    }
  }
}

// serving: abort test at the periodic poll
__device__ __forceinline__ bool serve_abort(const SearchArgs &a, int r, u64 tk) {
  const u64 v = ag_load(&a.sw[r]);
  return sw_tk(v) != tk || sw_st(v) != SV_RUN || ag_load(&a.dcancel[r]) == tk || ag_load(&a.sv->shut) != 0;
}

// serving: leave the request.  The worker that completes it (or fails it)
// records the result in device memory and sets the status in the slot word,
// which sends the others away at their next poll; the last worker to leave
// publishes the result to the host slot -- DONE or FAILED as set, or
// CANCELLED when every worker left a running request -- so the published
// node count includes every worker's nodes.
__device__ void serve_leave(const SearchArgs &a, int r, u64 tk, bool aborted, bool failed, s32 score, u32 move) {
  if (!aborted && !failed) ag_store(&a.dres[r], (u64)(u32)score | ((u64)move << 32));
  for (;;) {
    const u64 v = ag_load(&a.sw[r]);
    if (sw_tk(v) != tk) return;  // the slot has moved on
    int st = sw_st(v);
    if (st == SV_RUN) {
      if (failed) st = SV_FAILED;
      else if (!aborted) st = SV_DONE;
    }
    const bool last = sw_w(v) == 1;
    const u64 nv = ((v - 4) & ~3ULL) | (u64)st;
    // release: the result recorded above before the status that announces
    // it; acquire: the last worker sees the result the completing one wrote
    u64 e = v;
    if (!__hip_atomic_compare_exchange_strong(&a.sw[r], &e, nv, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT))
      continue;
    if (last) {
      const int pub = st == SV_RUN ? SV_CANCELLED : st;
      const u64 res = pub == SV_DONE ? ag_load(&a.dres[r]) : 0ULL;
      ServeSlot &sl = a.slots[r];
      sys_store(&sl.score, (s32)(u32)res);
      sys_store(&sl.bestmove, (u32)(res >> 32));
      sys_store(&sl.nodes, (u64)(ag_load(&a.nodes[r]) & 0xFFFFFFFFFFFFULL));
      sys_store(&sl.ticket, tk);
      sys_store(&sl.depth_done, (s32)(ag_load(&a.depth_done[r]) & 0xFFFF));
      sys_store(&sl.status, (s32)pub);
      // the one ordering point: the result fields before the done flag
      __hip_atomic_store(&sl.state, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_search(SearchArgs a) {
  __shared__ Ply ply[MAXPLY + 2];
  __shared__ u32 mmove[MSTACK];
  __shared__ int16_t mscore[MSTACK];
  __shared__ __align__(16) u16 ctab[NCOUNTER];  // counter moves by the previous move's squares (cleared 16 B at a time)
  __shared__ u32 killers[MAXPLY + 2];
  __shared__ float hpos[zm::NH1];
  __shared__ u8 mstart[64];  // move-list start marks (lane + 1 at its offset)
  const int lane = threadIdx.x;
  const int hstart = (int)((blockIdx.x * 7919u) % (unsigned)(a.nroots > 0 ? a.nroots : 1));

  const bool serving = a.slots != nullptr;
  if (serving && blockIdx.x == gridDim.x - 1) {
    serve_scheduler(a, lane);
    return;
  }
  for (;;) {
    int r;
    u64 ticket = 0;
    if (serving) {
      r = serve_pick(a, lane, ticket);
    } else {
      u32 ab = 0;
      if (lane == 0) ab = sys_load32(a.habort);
      r = uf(ab) ? -1 : pick_root(a, hstart, lane);
    }
    if (r < 0) return;
    int depth_r = a.depth;
    bool mlp_r = a.use_mlp != 0;
    if (serving) {
      depth_r = uf(ag_load(&a.dreq[r].depth));
      mlp_r = mlp_r && (uf(ag_load(&a.dreq[r].flags)) & 1);
      if (depth_r < 1) depth_r = 1;
      if (depth_r > MAXPLY - 1) depth_r = MAXPLY - 1;
    }

    // ---------------------------------------------------------------- root set-up
    // (the root position is re-read at every iteration rather than held in
    // scalar registers through the search)
    const u64 *const rootq = serving ? (const u64 *)a.dreq[r].pos : a.roots + (size_t)r * 8;
    int hl = 0;
    if (a.hist) {
      hl = uf(a.hlen[r]);
      hl = hl < 0 ? 0 : (hl > ZETA_HISTORY ? ZETA_HISTORY : hl);
    }
    // killers and counter moves start empty for every root (they persist
    // across its iterations), so a root's search does not depend on which
    // worker searched what before
    for (int k = lane; k < MAXPLY + 2; k += 64) killers[k] = 0;
    for (int k = lane; k < NCOUNTER / 8; k += 64) ((uint4 *)ctab)[k] = make_uint4(0u, 0u, 0u, 0u);
    wsync();

    // per-root statistics in one vector register (counter k in lane k) rather
    // than in scalar registers, which the node state needs
    u32 cvec = 0;
    u64 nodes = 0;
    int poll = 0;
    bool aborted = false, failed = false;
    s32 final_score = 0;
    u32 final_move = 0;

    // a worker joining a root that others are already searching starts at the
    // iteration in progress: the earlier ones would only repeat work whose
    // moves are in the shared TT (the final iteration alone sets the score)
    int id0 = 1;
    if (lane == 0) id0 = (__atomic_load_n(&a.depth_done[r], __ATOMIC_RELAXED) & (serving ? 0xFFFF : ~0)) + 1;
    id0 = uf(id0);
    id0 = id0 < 1 ? 1 : (id0 > depth_r ? depth_r : id0);
    for (int id = id0; id <= depth_r && !aborted && !failed; id++) {
      // the best score any worker has found for a root move in this
      // iteration (refreshed at the poll); -INF until one has
      s32 ra = -INF;
      {
        // a worker joining the iteration starts from the bound the root's
        // other workers have published
        u64 rb = 0;
        if (lane == 0) rb = ag_load(&a.rbound[r]);
        rb = uf(rb);
        if ((rb >> 32) == (ticket & 0xFFFFFFFFULL) && (int)((rb >> 16) & 0xFFFF) == id)
          ra = (s32)(rb & 0xFFFF) - 32768;
      }
      Node n;
      {
        u64 q[8];
        if (serving) {
          for (int k = 0; k < 8; k++) q[k] = uf(ag_load(&rootq[k]));
        } else {
          for (int k = 0; k < 8; k++) q[k] = rootq[k];
        }
        n.p.b0 = q[0];
        n.p.b1 = q[1];
        n.p.b2 = q[2];
        n.p.b3 = q[3];
        n.p.cr = q[4];
        n.p.stm = (u32)(q[7] & 1u);
        n.hash = q[5];
        n.hmc = (u32)(q[6] & 0xFF);
      }
      n.phash = hash_pieces(n.p);
      n.alpha = -INF;
      n.beta = INF;
      n.depth = id;
      n.lastmove = 0;
      n.mbeg = 0;
      n.mcnt = 0;
      n.busy = false;
      n.move = 0;
      n.best = 0;
      n.qs = false;
      n.kic = false;
      n.iter = 1;
      n.ndefer = 0;
      n.searched = 0;
      int sd = 1;
      bool parent_qs = false;
      s32 rootscore = 0;  // DRAWSCORE until a root move raises alpha (src/zeta.cl:986)
      u32 rootmove = 0;
      s32 v = 0;
      int state = S_ENTER;

      for (;;) {
        if (state == S_ENTER) {
          // ---------------- ENTER: analyse, extend, generate, terminal tests
          if (++poll >= POLL_EVERY) {
            poll = 0;
            u32 d = 0;
            if (lane == 0) {
              if (serving) {
                d = serve_abort(a, r, ticket);
              } else {
                // stop when the root is finished or given up, or the host aborts
                d = (__atomic_load_n(&a.done[r], __ATOMIC_RELAXED) != 0u || sys_load32(a.habort) != 0u) ? 1u : 0u;
                if (!d && a.max_nodes && nodes > a.max_nodes) {
                  // over the node budget: give the root up, so that no worker
                  // picks it again (a fresh budget on the same iteration could
                  // repeat forever); its result is this worker's deepest
                  // completed iteration
                  if (atomicCAS(&a.done[r], 0u, 3u) == 0u) {
                    a.score[r] = final_score;
                    a.bestmove[r] = final_move;
                  }
                  d = 1;
                }
              }
            }
            if (uf(d)) {
              aborted = true;
              break;
            }
            u64 rb = 0;
            if (lane == 0) rb = ag_load(&a.rbound[r]);
            rb = uf(rb);
            if ((rb >> 32) == (ticket & 0xFFFFFFFFULL) && (int)((rb >> 16) & 0xFFFF) == id)
              ra = (s32)(rb & 0xFFFF) - 32768;
          }
          const int plyn = sd - 1;
          // issue the TT load early: its latency overlaps the analysis below
          const bool early = a.ttmask && sd > 1 && n.depth >= 1;
          const u64 key = tt_key(n.hash, n.p.cr);
          u32x4 tte = {0u, 0u, 0u, 0u};
          if (early) tte = *(const u32x4 *)(a.tt + 2 * (key & a.ttmask));
          // node analysis, uniform part (king-square attackers by single-square
          // sliders): gives the in-check flag before the quiescence stand pat
          Node0 B;
          node_basics(n.p, B);
          n.kic = B.checkers != 0;
          if (n.depth >= 0 && !parent_qs && (n.kic || is_promo(n.lastmove))) n.depth++;
          n.qs = n.depth <= 0;
          // Another worker's root move reached ra: at the root's side (odd
          // plies from 3 on) a value at or below ra - 1 cannot make this root
          // move the best, so the window starts there.  A root move then
          // reports max(its value, a bound below ra) -- below the best root
          // score, which stays exact, as does its move.
          if ((sd & 1) && sd >= 3 && ra - 1 > n.alpha && ra - 1 < n.beta) n.alpha = ra - 1;
          n.best = 0;
          n.iter = 1;
          n.ndefer = 0;
          n.searched = 0;
          n.mcnt = 0;
          s32 ev0 = 0;
          bool have_ev = false;
          if (n.qs && !n.kic) {
            // quiescence stand pat before any move generation: the reference
            // returns the static eval both on a stand-pat cutoff and when no
            // capture exists (src/zeta.cl:1597-1605), so eval >= beta ends the
            // node with the same value either way
            const bool qdraw = sd > 1 && (n.hmc >= 100 || popc(n.p.b1 | n.p.b2 | n.p.b3) <= 2);
            if (qdraw) {
              v = 0;
              state = S_RETURN;
              wsync();
              continue;
            }
            const s32 e = wave_sum_i(eval_square(n.p, lane)) + eval_bishop_pair(n.p);
            ev0 = n.p.stm ? -e : e;
            have_ev = true;
            if (ev0 >= n.beta || sd >= MAXPLY) {
              if (sd >= MAXPLY) lcount(cvec, C_MAXPLY, 1);
              v = ev0;
              state = S_RETURN;
              wsync();
              continue;
            }
          }
          // move generation, lane = square: the lanes answer "is my square
          // attacked" (a ballot gives the attack map) and "which targets has
          // my piece"; a wave prefix sum compacts the lists
          // the attack map is read only for king moves and castling: skip it
          // (a uniform branch) when the king has no target square to test
          u64 A = 0;
          const u64 kt = king_att(B.kbb) & (n.qs ? B.opp : ~B.own);
          if (kt || (!n.qs && (n.p.cr & SMCRALL))) A = __ballot(attacked_nk(B, lane));
          u64 t = 0;
          if ((B.own >> lane) & 1) t = lane_targets(n.p, B, A, lane, n.qs);
          int total;
          const int off = wave_exscan_i(popc(t), lane, total);
          n.mcnt = total;
          if (n.mbeg + total > MSTACK) {
            if (lane == 0) atomicOr(a.err, 1u);
            failed = true;
            break;
          }
          // killer and counter move (quiescence nodes have no quiet moves)
          const u32 kil = killers[sd], cm = n.qs ? 0u : (u32)ctab[counter_index(n.lastmove)];
          {
            auto put = [&](int k, int from, int to) {
              const u32 m = make_code(n.p, from, to);
              s32 sc = order_score(n.p.stm, m);
              sc = sc > 31999 ? 31999 : (sc < -31999 ? -31999 : sc);
              if (((mv_pcpt(m) >> 1) & 7) == PNONE) {
                if (m == kil) sc = KILLER_SCORE;
                else if (counter_code(m) == cm) sc = COUNTER_SCORE;
              }
              mmove[k] = m;
              mscore[k] = (int16_t)sc;
            };
            if (!n.qs && total <= 64) {
              // lane j writes move j of the list (the same order as the loop
              // below: origin lanes ascending, then targets ascending): its
              // origin is the last lane with targets whose offset is <= j (a
              // prefix max over start marks), its target the (j - offset)-th
              // bit of that lane's set -- one pass instead of as many loop
              // trips as the busiest lane has targets
              mstart[lane] = 0;
              wsync();
              if (t) mstart[off] = (u8)(lane + 1);
              wsync();
              const int own1 = (int)zw::wave_scan_max(lane < total ? (u32)mstart[lane] : 0u);
              const int from = (own1 > 0 ? own1 : 1) - 1;
              const int k = lane - __shfl(off, from, 64);
              const u64 ts = ((u64)(u32)__shfl((int)(u32)(t >> 32), from, 64) << 32) | (u32)__shfl((int)(u32)t, from, 64);
              if (lane < total) put(n.mbeg + lane, from, nth_bit(ts, k));
            } else {
              int k = n.mbeg + off;
              while (t) {
                const int to = lsb(t);
                t &= t - 1;
                put(k, lane, to);
                k++;
              }
            }
          }
          if (mlp_r && !n.qs && n.depth >= a.mlp_min_depth && total > 1) {
            wsync();
            mlp_score_moves(a, n.p, n.mbeg, total, lane, hpos, mmove, mscore);
            // every move's key is its plain score plus the network's, in
            // plain-score units (zm::PLAIN_PER_LOGIT per unit of the
            // network's output, the scale the ranking loss is trained at);
            // the TT move still goes first
            for (int k = lane; k < total; k += 64) {
              const u32 m = mmove[n.mbeg + k];
              s32 sc = plain_score(n.p.stm, m, kil, cm) +
                       (s32)mscore[n.mbeg + k] * zm::PLAIN_PER_LOGIT / 256;
              sc = sc > 31999 ? 31999 : (sc < -31999 ? -31999 : sc);
              mscore[n.mbeg + k] = (int16_t)sc;
            }
            wsync();
            lcount(cvec, C_MLP_NODES, 1);
          }
          // terminal / draw tests in the reference's order (src/zeta.cl:1554-1609)
          bool draw = false;
          if (total > 0 && !n.qs && sd > 1) {
            // repetition against plies plyn-2, plyn-4, .. >= plyn-hmc: the
            // searched path, then the game history before the root
            bool hit = false;
            const int nn = plyn - 2 - 2 * lane;
            if (nn >= plyn - (int)n.hmc) {
              if (nn >= 0) hit = ply[nn + 1].hash == n.hash;
              else if (-nn - 1 < hl) hit = a.hist[(size_t)r * ZETA_HISTORY + (-nn - 1)] == n.hash;
            }
            draw = __ballot(hit) != 0;
          }
          if (sd > 1 && n.hmc >= 100) draw = true;
          if (sd > 1 && popc(B.occ) <= 2) draw = true;
          const bool need_eval = !draw && ((n.qs && !n.kic) || (n.qs && total == 0) || (sd >= MAXPLY && total > 0));
          s32 ev = 0;
          if (need_eval && have_ev) {
            ev = ev0;
          } else if (need_eval) {
            const s32 e = wave_sum_i(eval_square(n.p, lane)) + eval_bishop_pair(n.p);
            ev = n.p.stm ? -e : e;
          }
          state = S_NEXT;
          if (draw) {
            v = 0;
            state = S_RETURN;
          } else if (!n.qs && total == 0) {
            v = n.kic ? -INF + plyn : 0;
            state = S_RETURN;
          } else if (sd >= MAXPLY || total == 0) {
            if (sd >= MAXPLY) lcount(cvec, C_MAXPLY, 1);
            v = ev;
            state = S_RETURN;
          } else if (n.qs && !n.kic && ev >= n.beta) {
            v = ev;  // stand pat, fail soft
            state = S_RETURN;
          } else {
            if (n.qs && !n.kic && ev > n.alpha) n.alpha = ev;
            // TT probe: move for ordering; score under the exact-depth rule
            u32 ttmove = 0;
            if (a.ttmask && !n.qs && sd > 1) {
              if (!early) tte = *(const u32x4 *)(a.tt + 2 * (key & a.ttmask));
              const u64 w0 = uf(((u64)tte.y << 32) | tte.x), w1 = uf(((u64)tte.w << 32) | tte.z);
              lcount(cvec, C_TT_PROBE, 1);
              if ((w0 ^ w1) == key && ((w1 >> 57) & 1)) {
                lcount(cvec, C_TT_HIT, 1);
                ttmove = (u32)w1;
                s32 ts = (s32)(int16_t)(u16)(w1 >> 32);
                const int d = (int)((w1 >> 48) & 0xFF);
                if (is_mate(ts) && ts > 0) ts -= plyn;
                if (is_mate(ts) && ts < 0) ts += plyn;
                if (((w1 >> 56) & 1) && n.hmc == 0 && d == n.depth && ts > -INF && ts < INF) {
                  if ((w1 >> 58) & 1) {
                    // upper bound at or below alpha: the node fails low
                    if (ts <= n.alpha) {
                      lcount(cvec, C_TT_CUT, 1);
                      v = n.alpha;
                      state = S_RETURN;
                    }
                  } else if (ts > n.alpha) {
                    n.alpha = ts;
                    lcount(cvec, C_TT_CUT, 1);
                    if (n.alpha >= n.beta) {
                      v = n.alpha;
                      state = S_RETURN;
                    }
                  }
                }
              }
            }
            if (state == S_NEXT && ttmove) {
              for (int k = lane; k < total; k += 64)
                if (mmove[n.mbeg + k] == ttmove) mscore[n.mbeg + k] = (int16_t)TTMOVE_SCORE;
            }
            if (a.recs && lane == 0) ply[sd].ttm = ttmove;
            if (state == S_NEXT && a.hint && !n.qs) {
              // the hint table (a completed search's table): its move first
              const u32x4 he = *(const u32x4 *)(a.hint + 2 * (key & a.ttmask));
              const u64 h0 = uf(((u64)he.y << 32) | he.x), h1 = uf(((u64)he.w << 32) | he.z);
              if ((h0 ^ h1) == key && ((h1 >> 57) & 1)) {
                const u32 hm = (u32)h1;
                for (int k = lane; k < total; k += 64)
                  if (mmove[n.mbeg + k] == hm) mscore[n.mbeg + k] = (int16_t)HINT_SCORE;
              }
            }
          }
          wsync();
          continue;
        }

        if (state == S_NEXT) {
          // ---------------- NEXT: pick the best remaining move (wave arg-max)
          u32 bestkey = 0;
          for (int k = lane; k < n.mcnt; k += 64) {
            const s32 sc = mscore[n.mbeg + k];
            const bool fl = (mmove[n.mbeg + k] & DEFERRED) != 0u;
            const bool ok = sc != USED && (n.iter == 1 ? !fl : fl);
            if (ok) {
              const u32 key = ((u32)(sc + 32768) << 16) | (0xFFFFu - (u32)k);
              bestkey = key > bestkey ? key : bestkey;
            }
          }
          bestkey = wave_max_u32(bestkey);
          if (bestkey == 0) {
            if (n.iter == 1 && n.ndefer > 0) {
              n.iter = 2;
              continue;
            }
            v = n.alpha;
            lcount(cvec, C_TT_STORE, tt_store(a, n, sd, lane));
            state = S_RETURN;
            continue;
          }
          const int k = n.mbeg + (int)(0xFFFFu - (bestkey & 0xFFFFu));
          const u32 m = uf(mmove[k]) & ~DEFERRED;
          const s32 sk = (s32)(bestkey >> 16) - 32768;
          wsync();
          if (lane == 0) mscore[k] = (int16_t)USED;
          const Pos c = apply(n.p, m, false);
          u64 ch, cph;
          child_hash(n, m, c, ch, cph);
          // ABDADA: defer a sibling that another worker is already searching
          bool cbusy = false;
          if (a.workers > 1 && n.depth - 1 >= a.split_depth) {
            const u32 bidx = (u32)(tt_key(ch, c.cr) >> 40) & a.busymask;
            u32 b = 0;
            if (lane == 0) b = __atomic_load_n(&a.busy[bidx], __ATOMIC_RELAXED);
            b = uf(b);
            // batch searches defer a busy first child too, from ply 3 down:
            // a worker that finds it busy moves on to the siblings and comes
            // back to it in the second pass, which keeps helpers off the
            // path every worker starts on.  At the root and ply 2 the first
            // child is searched by every worker that reaches it: a sibling
            // searched before the first child has set the root's bound runs
            // with the full window (with 128 workers per root the depth-7
            // 1024-root batch took 44-45 s this way against 64-67 s with
            // the deferral at every ply, and 51 s from ply 4 down).  Served
            // requests defer only later siblings (a depth 4-9 stream
            // completed 43 instead of 47 of 48 requests within 120 s with
            // the first-child deferral at every ply)
            if (n.iter == 1 && b && (n.searched > 0 || (!serving && sd >= 3))) {
              if (lane == 0) {
                mscore[k] = (int16_t)sk;
                mmove[k] = m | DEFERRED;
              }
              n.ndefer++;
              lcount(cvec, C_DEFER, 1);
              wsync();
              continue;
            }
            if (lane == 0) atomicAdd(&a.busy[bidx], 1u);
            cbusy = true;
          }
          if (n.iter == 1) n.searched++;
          nodes++;
          n.move = m;
          if (lane == 0) save(ply[sd], n);
          // descend
          parent_qs = n.qs;
          Node cn;
          cn.p = c;
          cn.hash = ch;
          cn.phash = cph;
          cn.alpha = -n.beta;
          cn.beta = -n.alpha;
          cn.depth = n.depth - 1;
          cn.lastmove = m;
          cn.hmc = child_hmc(n.hmc, m);
          cn.mbeg = n.mbeg + n.mcnt;
          cn.mcnt = 0;
          cn.busy = cbusy;
          cn.move = 0;
          cn.best = 0;
          cn.qs = false;
          cn.kic = false;
          cn.iter = 1;
          cn.ndefer = 0;
          cn.searched = 0;
          n = cn;
          sd++;
          wsync();
          state = S_ENTER;
          continue;
        }

        // ---------------- RETURN v from node sd
        if (n.busy) busy_release(a, n.hash, n.p.cr, lane);
        if (sd == 1) break;
        sd--;
        wsync();
        {
          const Pos child = n.p;
          restore(ply[sd], child, n);
        }
        const s32 sc = -v;
        if ((sd & 1) && sd >= 3 && ra - 1 > n.alpha && ra - 1 < n.beta) {
          // the root bound rose while this node's children were searched:
          // its best move so far is below it, so no lower bound stands
          n.alpha = ra - 1;
          n.best = 0;
        }
        if (sc > n.alpha) {
          n.alpha = sc;
          n.best = n.move;
          if (sd == 1) {
            rootscore = sc;
            rootmove = n.move;
            if (lane == 0)
              atomicMax((unsigned long long *)&a.rbound[r],
                        (unsigned long long)(((ticket & 0xFFFFFFFFULL) << 32) | ((u64)id << 16) | (u64)(u16)(sc + 32768)));
          }
        }
        if (n.alpha >= n.beta) {
          if (a.recs && !n.qs && n.depth >= 1 && n.searched >= 1 && n.mcnt >= 2 &&
              ((u32)(n.hash >> 40) & a.rec_mask) == 0) {
            u32 nrec = 0;
            log_pairs(a, n, lane, mmove, mscore, ply[sd].ttm, killers[sd], (u32)ctab[counter_index(n.lastmove)], nrec);
            lcount(cvec, C_RECORDS, nrec);
          }
          // a quiet move that fails high becomes the ply's killer and the
          // counter move of the move before it (src/zeta.cl:1916-1933)
          if (!n.qs && ((mv_pcpt(n.move) >> 1) & 7) == PNONE && lane == 0) {
            killers[sd] = n.move;
            ctab[counter_index(n.lastmove)] = counter_code(n.move);
          }
          v = n.alpha;
          lcount(cvec, C_TT_STORE, tt_store(a, n, sd, lane));
          state = S_RETURN;
        } else {
          state = S_NEXT;
        }
        wsync();
      }

      if (aborted || failed) {
        // release the ABDADA marks still held along the path
        if (n.busy) busy_release(a, n.hash, n.p.cr, lane);
        wsync();
        for (int k = 2; k < sd; k++)
          if (ply[k].busy) busy_release(a, uf(ply[k].hash), uf(ply[k].cr), lane);
        break;
      }
      final_score = (rootscore == INF || rootscore == -INF) ? 0 : rootscore;  // src/search.c:137
      final_move = rootmove;
      if (lane == 0) {
        if (serving) slot_max_depth(&a.depth_done[r], ticket, id);
        else atomicMax(&a.depth_done[r], id);
      }
      // a mate within this iteration's depth is the score of every deeper
      // iteration too: iterative deepening stops there (src/search.c:177-186)
      if (is_mate(final_score) && id >= (final_score > 0 ? INF - final_score : INF + final_score)) break;
    }

    // ---------------------------------------------------------------- publish
    if (lane < C_NODES && cvec) atomicAdd((unsigned long long *)&a.ctr[lane], (unsigned long long)cvec);
    if (lane == 0) {
      // before leaving: the last worker to leave publishes the request's
      // count, which then holds every worker's nodes
      if (serving) slot_add_nodes(&a.nodes[r], ticket, nodes);
      else atomicAdd((unsigned long long *)&a.nodes[r], (unsigned long long)nodes);
      atomicAdd((unsigned long long *)&a.ctr[C_NODES], (unsigned long long)nodes);
      if (serving) {
        serve_leave(a, r, ticket, aborted, failed, final_score, final_move);
      } else {
        if (!aborted && !failed && atomicCAS(&a.done[r], 0u, 1u) == 0u) {
          a.score[r] = final_score;
          a.bestmove[r] = final_move;
        }
        if (failed) atomicCAS(&a.done[r], 0u, 2u);
        atomicSub(&a.nworkers[r], 1);
      }
    }
    wsync();
  }
}

// the search's learned-ordering routine applied to given moves (for tests and
// external callers): one wave per position, up to MSTACK moves each
__global__ __launch_bounds__(64) void k_mlp_score(SearchArgs a, const u64 *pos, const u32 *moves, const int *off,
                                                 float *out) {
  __shared__ u32 mmove[MSTACK];
  __shared__ int16_t mscore[MSTACK];
  __shared__ float hpos[zm::NH1];
  const int lane = threadIdx.x, i = blockIdx.x;
  const u64 *q = pos + (size_t)i * 8;
  Pos p;
  p.b0 = q[0];
  p.b1 = q[1];
  p.b2 = q[2];
  p.b3 = q[3];
  p.cr = q[4];
  p.stm = (u32)(q[7] & 1u);
  const int b = off[i], n = off[i + 1] - off[i];
  for (int k = lane; k < n; k += 64) mmove[k] = moves[b + k];
  wsync();
  mlp_score_moves(a, p, 0, n, lane, hpos, mmove, mscore);
  for (int k = lane; k < n; k += 64) out[b + k] = (float)mscore[k] / 256.f;
}

// Principal variation after a search: the best root move, then the moves the
// transposition table holds for the positions that follow, each checked to
// be legal (the reference walks its TT the same way, src/zeta.cl:2450-2512).
// One thread per root; stops at a missing or mismatching entry.
__global__ void k_pv(const u64 *roots, const u32 *bestmove, int n, int len, const u64 *tt, u64 ttmask,
                     u32 *pv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32 *out = pv + (size_t)i * len;
  for (int k = 0; k < len; k++) out[k] = 0;
  const u64 *q = roots + (size_t)i * 8;
  Pos p;
  p.b0 = q[0];
  p.b1 = q[1];
  p.b2 = q[2];
  p.b3 = q[3];
  p.cr = q[4];
  p.stm = (u32)(q[7] & 1u);
  Node nd;
  nd.hash = q[5];
  nd.phash = hash_pieces(p);
  u32 m = bestmove[i];
  for (int k = 0; k < len && m; k++) {
    const int f = (int)mv_from(m), t = (int)mv_to(m);
    Legal L;
    analyse(p, L);
    if (!((L.own >> f) & 1) || !((piece_targets(p, L, f, false) >> t) & 1) || make_code(p, f, t) != m) break;
    out[k] = m;
    const Pos c = apply(p, m, false);
    u64 h, ph;
    child_hash(nd, m, c, h, ph);
    nd.hash = h;
    nd.phash = ph;
    p = c;
    m = 0;
    if (ttmask) {
      const u64 key = tt_key(h, c.cr);
      const u64 w0 = tt[2 * (key & ttmask)], w1 = tt[2 * (key & ttmask) + 1];
      if ((w0 ^ w1) == key && ((w1 >> 57) & 1)) m = (u32)w1;
    }
  }
}

__global__ void k_clear_u64(u64 *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

struct SearchState {
  zr::DevBuf tt, busy, roots, queue, done, nworkers, score, bestmove, nodes, depth_done, err, ctr, hist, hlen, pv, hint,
      rbound;
  u64 tt_entries = 0;
  u64 last_ctr[C_NCTR] = {0};
};
SearchState g_ss;

inline unsigned nblocks(size_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

const zm::Weights *mlp_weights();
void *records_buffer(u64 *cap, u64 **count);

void search_release() {
  zr::release(g_ss.tt); zr::release(g_ss.busy); zr::release(g_ss.roots); zr::release(g_ss.queue);
  zr::release(g_ss.done); zr::release(g_ss.nworkers); zr::release(g_ss.score); zr::release(g_ss.bestmove);
  zr::release(g_ss.nodes); zr::release(g_ss.depth_done); zr::release(g_ss.err);
  zr::release(g_ss.ctr); zr::release(g_ss.hist); zr::release(g_ss.hlen); zr::release(g_ss.pv);
  zr::release(g_ss.hint); zr::release(g_ss.rbound);
  g_ss.tt_entries = 0;
}

extern "C" int zeta_search_batch(const zeta_pos *pos, int n, const zeta_search_opts *opts, zeta_result *res) {
  int r = zr::require_up();
  if (r) return r;
  if (n < 0 || (n > 0 && (!pos || !res)) || !opts) return zr::fail(ZETA_EINVAL, "search: bad arguments");
  if (opts->depth < 1 || opts->depth > 40) return zr::fail(ZETA_EINVAL, "search: depth %d out of range", opts->depth);
  if (opts->split_depth < 0 || opts->split_depth > 40 || opts->waves_per_simd < 0 || opts->waves_per_simd > 16)
    return zr::fail(ZETA_EINVAL, "search: split_depth %d / waves_per_simd %d out of range", opts->split_depth,
                    opts->waves_per_simd);
  if (n == 0) return ZETA_OK;
  const int workers = opts->workers < 1 ? 1 : (opts->workers > 1024 ? 1024 : opts->workers);
  hipStream_t st = zr::ctx().stream;
  const int tt_log2 = opts->tt_log2 < 0 ? zr::ctx().cfg.tt_log2 : opts->tt_log2;
  const u64 tt_entries = tt_log2 > 0 ? (1ULL << tt_log2) : 0;
  if (tt_log2 > 34) return zr::fail(ZETA_EINVAL, "search: tt_log2 %d too large", tt_log2);
  const u32 busy_entries = 1u << 20;
  if ((r = zr::ensure(g_ss.roots, (size_t)n * sizeof(zeta_pos))) || (r = zr::ensure(g_ss.queue, 64)) ||
      (r = zr::ensure(g_ss.done, (size_t)n * 4)) || (r = zr::ensure(g_ss.nworkers, (size_t)n * 4)) ||
      (r = zr::ensure(g_ss.score, (size_t)n * 4)) || (r = zr::ensure(g_ss.bestmove, (size_t)n * 4)) ||
      (r = zr::ensure(g_ss.nodes, (size_t)n * 8)) || (r = zr::ensure(g_ss.depth_done, (size_t)n * 4)) ||
      (r = zr::ensure(g_ss.rbound, (size_t)n * 8)) ||
      (r = zr::ensure(g_ss.err, 64)) || (r = zr::ensure(g_ss.ctr, C_NCTR * 8)) ||
      (r = zr::ensure(g_ss.busy, (size_t)busy_entries * 4)))
    return r;
  if (opts->order_hint && (!tt_entries || g_ss.tt_entries != tt_entries))
    return zr::fail(ZETA_EINVAL, "search: order_hint needs the previous batch's table of the same size");
  if (tt_entries) {
    if ((r = zr::ensure(g_ss.tt, tt_entries * 16))) return r;
    if (opts->order_hint) {
      // the previous batch's table becomes the hint table before it is cleared
      if ((r = zr::ensure(g_ss.hint, tt_entries * 16))) return r;
      ZCHECK(hipMemcpyAsync(g_ss.hint.p, g_ss.tt.p, tt_entries * 16, hipMemcpyDeviceToDevice, st));
    }
    if (g_ss.tt_entries != tt_entries || opts->clear_tt) {
      ZLAUNCH(zr::K_SEARCH, 0, k_clear_u64, 4096, 256, 0, (u64 *)g_ss.tt.p, (size_t)(tt_entries * 2));
      g_ss.tt_entries = tt_entries;
    }
  }
  ZCHECK(hipMemcpyAsync(g_ss.roots.p, pos, (size_t)n * sizeof(zeta_pos), hipMemcpyHostToDevice, st));
  ZCHECK(hipMemsetAsync(g_ss.queue.p, 0, 64, st));
  ZCHECK(hipMemsetAsync(g_ss.done.p, 0, (size_t)n * 4, st));
  ZCHECK(hipMemsetAsync(g_ss.nworkers.p, 0, (size_t)n * 4, st));
  ZCHECK(hipMemsetAsync(g_ss.score.p, 0, (size_t)n * 4, st));
  ZCHECK(hipMemsetAsync(g_ss.bestmove.p, 0, (size_t)n * 4, st));
  ZCHECK(hipMemsetAsync(g_ss.nodes.p, 0, (size_t)n * 8, st));
  ZCHECK(hipMemsetAsync(g_ss.depth_done.p, 0, (size_t)n * 4, st));
  ZCHECK(hipMemsetAsync(g_ss.rbound.p, 0, (size_t)n * 8, st));
  ZCHECK(hipMemsetAsync(g_ss.err.p, 0, 64, st));
  ZCHECK(hipMemsetAsync(g_ss.ctr.p, 0, C_NCTR * 8, st));
  ZCHECK(hipMemsetAsync(g_ss.busy.p, 0, (size_t)busy_entries * 4, st));

  int per_cu = 0;
  ZCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_search, 64, 0));
  if (opts->waves_per_simd > 0 && 4 * opts->waves_per_simd < per_cu) per_cu = 4 * opts->waves_per_simd;
  if (per_cu < 1) per_cu = 1;
  const u64 cap = (u64)per_cu * (u64)zr::ctx().cu_count;
  u64 want = (u64)n * (u64)workers;
  const unsigned grid = (unsigned)(want < cap ? want : cap);

  SearchArgs sa;
  sa.roots = (const u64 *)g_ss.roots.p;
  sa.nroots = n;
  sa.depth = opts->depth;
  sa.workers = workers;
  sa.split_depth = opts->split_depth > 0 ? opts->split_depth : ABDADA_MIN_DEPTH;
  sa.tt = (u64 *)g_ss.tt.p;
  sa.ttmask = tt_entries ? tt_entries - 1 : 0;
  sa.hint = opts->order_hint ? (const u64 *)g_ss.hint.p : nullptr;
  sa.rbound = (u64 *)g_ss.rbound.p;
  sa.busy = (u32 *)g_ss.busy.p;
  sa.busymask = busy_entries - 1;
  sa.queue = (int *)g_ss.queue.p;
  sa.done = (u32 *)g_ss.done.p;
  sa.nworkers = (int *)g_ss.nworkers.p;
  sa.score = (s32 *)g_ss.score.p;
  sa.bestmove = (u32 *)g_ss.bestmove.p;
  sa.nodes = (u64 *)g_ss.nodes.p;
  sa.depth_done = (s32 *)g_ss.depth_done.p;
  sa.err = (u32 *)g_ss.err.p;
  sa.habort = zr::ctx().abort_dev;
  sa.max_nodes = opts->max_nodes;
  sa.ctr = (u64 *)g_ss.ctr.p;
  sa.use_mlp = 0;
  sa.mlp_min_depth = opts->mlp_min_depth;
  memset(&sa.w, 0, sizeof(sa.w));
  if (opts->use_mlp) {
    const zm::Weights *w = mlp_weights();
    if (!w) return zr::fail(ZETA_ESTATE, "search: use_mlp set but no network loaded (zeta_mlp_init/zeta_mlp_set)");
    sa.w = *w;
    sa.use_mlp = 1;
  }
  sa.recs = nullptr;
  sa.rec_count = nullptr;
  sa.rec_cap = 0;
  sa.rec_mask = (opts->log_shift > 0 && opts->log_shift < 24) ? ((1u << opts->log_shift) - 1u) : 0u;
  sa.hist = nullptr;
  sa.hlen = nullptr;
  if (opts->history) {
    if (!opts->history_len) return zr::fail(ZETA_EINVAL, "search: history without history_len");
    if ((r = zr::ensure(g_ss.hist, (size_t)n * ZETA_HISTORY * 8)) || (r = zr::ensure(g_ss.hlen, (size_t)n * 4)))
      return r;
    ZCHECK(hipMemcpyAsync(g_ss.hist.p, opts->history, (size_t)n * ZETA_HISTORY * 8, hipMemcpyHostToDevice, st));
    ZCHECK(hipMemcpyAsync(g_ss.hlen.p, opts->history_len, (size_t)n * 4, hipMemcpyHostToDevice, st));
    sa.hist = (const u64 *)g_ss.hist.p;
    sa.hlen = (const s32 *)g_ss.hlen.p;
  }
  sa.sh = nullptr;
  sa.squeue = nullptr;
  sa.sack = nullptr;
  sa.scancel = nullptr;
  sa.slots = nullptr;
  sa.nslots = 0;
  sa.qcap = 0;
  sa.sv = nullptr;
  sa.sw = nullptr;
  sa.dreq = nullptr;
  sa.dcancel = nullptr;
  sa.watchdog_ticks = 0;
  if (opts->log_records) {
    u64 cap = 0, *cnt = nullptr;
    void *recs = records_buffer(&cap, &cnt);
    if (!recs || !cap) return zr::fail(ZETA_ESTATE, "search: log_records set but no record buffer (zeta_records_reserve)");
    sa.recs = (u64 *)recs;
    sa.rec_count = cnt;
    sa.rec_cap = cap;
  }
  ZLAUNCH(zr::K_SEARCH, (u64)n, k_search, grid, 64, 0, sa);

  std::vector<s32> score(n), dd(n);
  std::vector<u32> bm(n), done(n);
  std::vector<u64> nodes(n);
  u32 err = 0;
  ZCHECK(hipMemcpyAsync(score.data(), g_ss.score.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(bm.data(), g_ss.bestmove.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(done.data(), g_ss.done.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(nodes.data(), g_ss.nodes.p, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(dd.data(), g_ss.depth_done.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(&err, g_ss.err.p, 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(g_ss.last_ctr, g_ss.ctr.p, sizeof(g_ss.last_ctr), hipMemcpyDeviceToHost, st));
  ZCHECK(hipStreamSynchronize(st));
  zr::time_collect();
  // a root the workers left without a result was stopped by zeta_search_abort
  const bool host_abort = __atomic_load_n(zr::ctx().abort_host, __ATOMIC_ACQUIRE) != 0u;
  int incomplete = -1;
  for (int i = 0; i < n; i++) {
    res[i].score = score[i];
    res[i].bestmove = bm[i];
    res[i].nodes = nodes[i];
    res[i].depth = dd[i];
    res[i].status = done[i] == 1 ? ZETA_OK
                  : done[i] == 2 ? ZETA_ENOMEM
                  : done[i] == 3 ? ZETA_EBUDGET
                  : host_abort ? ZETA_ECANCELED : ZETA_ESTATE;
    if (res[i].status == ZETA_ESTATE && incomplete < 0) incomplete = i;
  }
  if (opts->pv && opts->pv_len > 0) {
    const int len = opts->pv_len > MAXPLY ? MAXPLY : opts->pv_len;
    if ((r = zr::ensure(g_ss.pv, (size_t)n * len * 4))) return r;
    hipLaunchKernelGGL(k_pv, dim3((n + 63) / 64), dim3(64), 0, st, (const u64 *)g_ss.roots.p,
                       (const u32 *)g_ss.bestmove.p, n, len, (const u64 *)g_ss.tt.p, sa.ttmask, (u32 *)g_ss.pv.p);
    ZCHECK(hipGetLastError());
    std::vector<u32> pvh((size_t)n * len);
    ZCHECK(hipMemcpyAsync(pvh.data(), g_ss.pv.p, pvh.size() * 4, hipMemcpyDeviceToHost, st));
    ZCHECK(hipStreamSynchronize(st));
    for (int i = 0; i < n; i++)
      for (int k = 0; k < opts->pv_len; k++) opts->pv[(size_t)i * opts->pv_len + k] = k < len ? pvh[(size_t)i * len + k] : 0u;
  }
  if (err & 1u) return zr::fail(ZETA_ENOMEM, "search: move stack overflow (more than %d moves on one path)", MSTACK);
  if (incomplete >= 0) return zr::fail(ZETA_ESTATE, "search: root %d did not complete", incomplete);
  return ZETA_OK;
}

extern "C" int zeta_search_abort(int on) {
  // may run on another host thread while zeta_search_batch waits: it touches
  // the abort word only
  int r = zr::require_init();
  if (r) return r;
  __atomic_store_n(zr::ctx().abort_host, on ? 1u : 0u, __ATOMIC_RELEASE);
  return ZETA_OK;
}

extern "C" int zeta_search_counters(uint64_t *out, int n) {
  int r = zr::require_up();
  if (r) return r;
  if (!out || n < 0) return zr::fail(ZETA_EINVAL, "zeta_search_counters: bad arguments");
  for (int i = 0; i < n && i < C_NCTR; i++) out[i] = g_ss.last_ctr[i];
  return ZETA_OK;
}

extern "C" int zeta_mlp_score(const zeta_pos *pos, int npos, const uint32_t *moves, const int32_t *offsets,
                              float *scores) {
  int r = zr::require_up();
  if (r) return r;
  const zm::Weights *w = mlp_weights();
  if (!w) return zr::fail(ZETA_ESTATE, "zeta_mlp_score: no network loaded");
  if (npos <= 0 || !pos || !moves || !offsets || !scores) return zr::fail(ZETA_EINVAL, "zeta_mlp_score: bad arguments");
  for (int i = 0; i < npos; i++)
    if (offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] > MSTACK)
      return zr::fail(ZETA_EINVAL, "zeta_mlp_score: bad offsets at %d", i);
  const int nm = offsets[npos];
  zr::DevBuf dp, dm, doff, dout;
  if ((r = zr::ensure(dp, (size_t)npos * sizeof(zeta_pos))) || (r = zr::ensure(dm, (size_t)nm * 4 + 4)) ||
      (r = zr::ensure(doff, (size_t)(npos + 1) * 4)) || (r = zr::ensure(dout, (size_t)nm * 4 + 4))) {
    zr::release(dp); zr::release(dm); zr::release(doff); zr::release(dout);
    return r;
  }
  hipStream_t st = zr::ctx().stream;
  SearchArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.w = *w;
  sa.use_mlp = 1;
  hipError_t e = hipMemcpyAsync(dp.p, pos, (size_t)npos * sizeof(zeta_pos), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dm.p, moves, (size_t)nm * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(doff.p, offsets, (size_t)(npos + 1) * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mlp_score, dim3(npos), dim3(64), 0, st, sa, (const u64 *)dp.p, (const u32 *)dm.p,
                       (const int *)doff.p, (float *)dout.p);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(scores, dout.p, (size_t)nm * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  zr::release(dp); zr::release(dm); zr::release(doff); zr::release(dout);
  if (e != hipSuccess) return zr::hip_fail(e, "zeta_mlp_score", __FILE__, __LINE__);
  return ZETA_OK;
}

// ============================================================== serving
namespace {

struct ServeState {
  bool on = false;
  int nslots = 0, qcap = 0, workers = 1;
  void *pinned = nullptr;  // ServeHost + queue + ack + cancel + slots (hipHostMalloc, coherent)
  ServeHost *sh = nullptr;
  u64 *queue = nullptr, *ack = nullptr, *cancel = nullptr;
  ServeSlot *slots = nullptr;
  zr::DevBuf sv, sw, dreq, dcancel, nodes, depth_done, busy, ctr, err, rbound, dres;
  std::vector<u64> slot_ticket;  // ticket + 1 held by the slot, 0 = free
  std::vector<u64> slot_cost;
  std::vector<int> free_slots;
  u64 next_ticket = 0;
  u64 budget = 0, in_use = 0, tt_bytes = 0, worker_bytes = 0;
  bool use_mlp = false;
};
ServeState g_sv;

inline void sv_heartbeat() {
  __atomic_fetch_add(&g_sv.sh->heartbeat, 1ULL, __ATOMIC_RELEASE);
}

u64 sv_cost(int workers) {
  return (u64)workers * g_sv.worker_bytes + (g_sv.nslots ? g_sv.tt_bytes / (u64)g_sv.nslots : 0);
}

int sv_require() {
  int r = zr::require_init();
  if (r) return r;
  if (!g_sv.on) return zr::fail(ZETA_ESTATE, "serving not started (zeta_serve_start)");
  return ZETA_OK;
}

// the grid has exited on its own (watchdog or protocol error): report it
int sv_check_alive() {
  hipError_t q = hipStreamQuery(zr::ctx().stream);
  if (q == hipErrorNotReady) return ZETA_OK;
  if (q != hipSuccess) return zr::hip_fail(q, "serving grid", __FILE__, __LINE__);
  ServeDev d;
  memset(&d, 0, sizeof(d));
  (void)hipMemcpy(&d, g_sv.sv.p, sizeof(d), hipMemcpyDeviceToHost);
  return zr::fail(ZETA_ESTATE,
                  "serving grid has exited (reason %u: 2 watchdog, 3 host protocol error; now/ticket %llu, "
                  "arg %llu, xcc %#x, claimed %llu, mirror %llu); call zeta_serve_stop",
                  d.shut, (unsigned long long)d.exit_now, (unsigned long long)d.exit_arg, d.exit_xcc,
                  (unsigned long long)d.claimed, (unsigned long long)d.sub);
}

void sv_take(int s, zeta_result *res) {
  ServeSlot &sl = g_sv.slots[s];
  res->score = sl.score;
  res->bestmove = sl.bestmove;
  res->nodes = sl.nodes;
  res->depth = sl.depth_done;
  res->status = sl.status == SV_DONE ? ZETA_OK : (sl.status == SV_FAILED ? ZETA_ENOMEM : ZETA_ECANCELED);
  __atomic_store_n(&sl.state, 0u, __ATOMIC_RELEASE);
  g_sv.in_use -= g_sv.slot_cost[s];
  g_sv.slot_cost[s] = 0;
  g_sv.slot_ticket[s] = 0;
  g_sv.free_slots.push_back(s);
}

void sv_free_all() {
  zr::release(g_sv.sv); zr::release(g_sv.sw); zr::release(g_sv.dreq); zr::release(g_sv.dcancel);
  zr::release(g_sv.nodes); zr::release(g_sv.depth_done); zr::release(g_sv.busy); zr::release(g_sv.ctr);
  zr::release(g_sv.err); zr::release(g_sv.rbound); zr::release(g_sv.dres);
  if (g_sv.pinned) (void)hipHostFree(g_sv.pinned);
  g_sv.pinned = nullptr;
  g_sv.sh = nullptr;
  g_sv.slots = nullptr;
  g_sv.on = false;
  zr::ctx().serving = false;
}

}  // namespace

extern "C" int zeta_serve_start(const zeta_serve_opts *o) {
  int r = zr::require_up();
  if (r) return r;
  if (!o) return zr::fail(ZETA_EINVAL, "zeta_serve_start: null options");
  if (o->slots < 1 || o->slots > 4096) return zr::fail(ZETA_EINVAL, "zeta_serve_start: slots %d not in 1..4096", o->slots);
  if (o->workers < 1 || o->workers > 255)  // 8 bits of the slot word
    return zr::fail(ZETA_EINVAL, "zeta_serve_start: workers %d not in 1..255", o->workers);
  const int tt_log2 = o->tt_log2 < 0 ? zr::ctx().cfg.tt_log2 : o->tt_log2;
  if (tt_log2 > 34) return zr::fail(ZETA_EINVAL, "zeta_serve_start: tt_log2 %d too large", tt_log2);
  const zm::Weights *w = nullptr;
  if (o->use_mlp) w = mlp_weights();  // no network loaded: plain ordering for every request
  hipStream_t st = zr::ctx().stream;
  ZCHECK(hipStreamSynchronize(st));

  ServeState &S = g_sv;
  S = ServeState();
  S.nslots = o->slots;
  S.qcap = 2 * o->slots;
  S.workers = o->workers;
  S.budget = o->mem_budget;
  S.use_mlp = w != nullptr;
  hipFuncAttributes fa;
  ZCHECK(hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(k_search)));
  S.worker_bytes = (u64)fa.sharedSizeBytes;
  const u64 tt_entries = tt_log2 > 0 ? (1ULL << tt_log2) : 0;
  S.tt_bytes = tt_entries * 16;

  const size_t hbytes = sizeof(ServeHost) + (size_t)(2 * S.qcap + S.nslots) * 8 + (size_t)S.nslots * sizeof(ServeSlot);
  hipError_t e = hipHostMalloc(&S.pinned, hbytes, hipHostMallocCoherent | hipHostMallocMapped);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    S.pinned = nullptr;
    return zr::fail(ZETA_ENOMEM, "zeta_serve_start: hipHostMalloc of %zu bytes failed (%s)", hbytes, hipGetErrorString(e));
  }
  memset(S.pinned, 0, hbytes);
  char *hp = (char *)S.pinned;
  S.sh = (ServeHost *)hp;
  S.queue = (u64 *)(hp + sizeof(ServeHost));
  S.ack = S.queue + S.qcap;
  S.cancel = S.ack + S.qcap;
  S.slots = (ServeSlot *)(S.cancel + S.nslots);
  void *dev_pinned = nullptr;
  e = hipHostGetDevicePointer(&dev_pinned, S.pinned, 0);
  if (e != hipSuccess) {
    sv_free_all();
    return zr::hip_fail(e, "hipHostGetDevicePointer", __FILE__, __LINE__);
  }
  const ptrdiff_t dd = (char *)dev_pinned - (char *)S.pinned;
  auto dp = [dd](void *p) { return (void *)((char *)p + dd); };

  const u32 busy_entries = 1u << 20;
  if ((r = zr::ensure(S.sv, sizeof(ServeDev))) || (r = zr::ensure(S.sw, (size_t)S.nslots * 8)) ||
      (r = zr::ensure(S.dreq, (size_t)S.nslots * sizeof(ServeReq))) || (r = zr::ensure(S.dcancel, (size_t)S.nslots * 8)) ||
      (r = zr::ensure(S.nodes, (size_t)S.nslots * 8)) || (r = zr::ensure(S.depth_done, (size_t)S.nslots * 4)) ||
      (r = zr::ensure(S.busy, (size_t)busy_entries * 4)) || (r = zr::ensure(S.ctr, C_NCTR * 8)) || (r = zr::ensure(S.err, 64)) ||
      (r = zr::ensure(S.rbound, (size_t)S.nslots * 8)) || (r = zr::ensure(S.dres, (size_t)S.nslots * 8))) {
    sv_free_all();
    return r;
  }
  if (tt_entries && ((r = zr::ensure(g_ss.tt, tt_entries * 16)))) {
    sv_free_all();
    return r;
  }
#define SV_CHECK(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess) {                                                     \
      int rc_ = zr::hip_fail(e_, #call, __FILE__, __LINE__);                    \
      sv_free_all();                                                            \
      return rc_;                                                               \
    }                                                                           \
  } while (0)
  if (tt_entries && g_ss.tt_entries != tt_entries) {
    hipLaunchKernelGGL(k_clear_u64, dim3(4096), dim3(256), 0, st, (u64 *)g_ss.tt.p, (size_t)(tt_entries * 2));
    SV_CHECK(hipGetLastError());
    g_ss.tt_entries = tt_entries;
  }
  SV_CHECK(hipMemsetAsync(S.sv.p, 0, sizeof(ServeDev), st));
  SV_CHECK(hipMemsetAsync(S.sw.p, 0, (size_t)S.nslots * 8, st));
  SV_CHECK(hipMemsetAsync(S.dcancel.p, 0, (size_t)S.nslots * 8, st));
  SV_CHECK(hipMemsetAsync(S.nodes.p, 0, (size_t)S.nslots * 8, st));
  SV_CHECK(hipMemsetAsync(S.depth_done.p, 0, (size_t)S.nslots * 4, st));
  SV_CHECK(hipMemsetAsync(S.busy.p, 0, (size_t)busy_entries * 4, st));
  SV_CHECK(hipMemsetAsync(S.ctr.p, 0, C_NCTR * 8, st));
  SV_CHECK(hipMemsetAsync(S.err.p, 0, 64, st));
  SV_CHECK(hipMemsetAsync(S.rbound.p, 0, (size_t)S.nslots * 8, st));
  SV_CHECK(hipStreamSynchronize(st));

  int rate_khz = 0;
  SV_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, zr::ctx().device));
  if (rate_khz <= 0) rate_khz = 100000;
  const int wd_ms = o->watchdog_ms > 0 ? o->watchdog_ms : 10000;

  int per_cu = 0;
  SV_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_search, 64, 0));
  if (per_cu < 1) per_cu = 1;
  const int cap = per_cu * zr::ctx().cu_count;
  int grid = o->grid > 0 && o->grid < cap ? o->grid : cap;
  if (grid < 2) grid = 2;  // one scheduler wavefront + at least one worker

  SearchArgs sa;
  memset(&sa, 0, sizeof(sa));
  sa.workers = S.workers;
  sa.split_depth = ABDADA_MIN_DEPTH;
  sa.tt = (u64 *)g_ss.tt.p;
  sa.ttmask = tt_entries ? tt_entries - 1 : 0;
  sa.busy = (u32 *)S.busy.p;
  sa.busymask = busy_entries - 1;
  sa.nodes = (u64 *)S.nodes.p;
  sa.depth_done = (s32 *)S.depth_done.p;
  sa.err = (u32 *)S.err.p;
  sa.ctr = (u64 *)S.ctr.p;
  sa.habort = zr::ctx().abort_dev;
  sa.rbound = (u64 *)S.rbound.p;
  sa.dres = (u64 *)S.dres.p;
  sa.use_mlp = S.use_mlp ? 1 : 0;
  sa.mlp_min_depth = o->mlp_min_depth;
  if (w) sa.w = *w;
  sa.sh = (ServeHost *)dp(S.sh);
  sa.squeue = (const u64 *)dp(S.queue);
  sa.sack = (u64 *)dp(S.ack);
  sa.scancel = (const u64 *)dp(S.cancel);
  sa.slots = (ServeSlot *)dp(S.slots);
  sa.nslots = S.nslots;
  sa.qcap = S.qcap;
  sa.sv = (ServeDev *)S.sv.p;
  sa.sw = (u64 *)S.sw.p;
  sa.dreq = (ServeReq *)S.dreq.p;
  sa.dcancel = (u64 *)S.dcancel.p;
  sa.watchdog_ticks = (u64)wd_ms * (u64)rate_khz;
  S.free_slots.clear();
  for (int s = S.nslots - 1; s >= 0; s--) S.free_slots.push_back(s);
  S.slot_ticket.assign(S.nslots, 0);
  S.slot_cost.assign(S.nslots, 0);
  S.on = true;
  zr::ctx().serving = true;
  hipLaunchKernelGGL(k_search, dim3(grid), dim3(64), 0, st, sa);
  SV_CHECK(hipGetLastError());
#undef SV_CHECK
  return ZETA_OK;
}

extern "C" int zeta_serve_cost(int workers, uint64_t *bytes) {
  int r = sv_require();
  if (r) return r;
  if (!bytes) return zr::fail(ZETA_EINVAL, "zeta_serve_cost: null out pointer");
  if (workers <= 0) workers = g_sv.workers;
  if (workers > g_sv.workers) workers = g_sv.workers;
  *bytes = sv_cost(workers);
  return ZETA_OK;
}

extern "C" int zeta_serve_submit(const zeta_serve_request *req, uint64_t *ticket) {
  int r = sv_require();
  if (r) return r;
  ServeState &S = g_sv;
  sv_heartbeat();
  if (!req || !ticket) return zr::fail(ZETA_EINVAL, "zeta_serve_submit: null argument");
  if (req->depth < 1 || req->depth > MAXPLY - 1)
    return zr::fail(ZETA_EINVAL, "zeta_serve_submit: depth %d not in 1..%d", req->depth, MAXPLY - 1);
  if ((r = sv_check_alive())) return r;
  int workers = req->workers <= 0 ? S.workers : req->workers;
  if (workers > S.workers) workers = S.workers;
  const u64 cost = sv_cost(workers);
  if (S.budget && cost > S.budget)
    return zr::fail(ZETA_ENOMEM, "zeta_serve_submit: request needs %llu bytes, budget is %llu",
                    (unsigned long long)cost, (unsigned long long)S.budget);
  if (S.free_slots.empty()) return zr::fail(ZETA_EBUSY, "zeta_serve_submit: all %d slots in use", S.nslots);
  if (S.budget && S.in_use + cost > S.budget)
    return zr::fail(ZETA_EBUSY, "zeta_serve_submit: memory budget in use (%llu of %llu bytes)",
                    (unsigned long long)S.in_use, (unsigned long long)S.budget);
  const u64 t = S.next_ticket;
  const int qi = (int)(t % (u64)S.qcap);
  // the queue entry's previous ticket must have been claimed
  if (t >= (u64)S.qcap && __atomic_load_n(&S.ack[qi], __ATOMIC_ACQUIRE) != t - (u64)S.qcap + 1)
    return zr::fail(ZETA_EBUSY, "zeta_serve_submit: claim queue full");
  const int s = S.free_slots.back();
  S.free_slots.pop_back();
  ServeSlot &sl = S.slots[s];
  memcpy(sl.req.pos, &req->pos, sizeof(zeta_pos));
  sl.req.depth = req->depth;
  sl.req.flags = req->flags;
  sl.req.workers = workers;
  sl.req.pad = 0;
  __atomic_store_n(&sl.state, 0u, __ATOMIC_RELAXED);
  __atomic_store_n(&S.cancel[s], 0ULL, __ATOMIC_RELAXED);
  __atomic_store_n(&S.queue[qi], (u64)s, __ATOMIC_RELAXED);
  S.next_ticket = t + 1;
  __atomic_store_n(&S.sh->submitted, t + 1, __ATOMIC_RELEASE);
  S.slot_ticket[s] = t + 1;
  S.slot_cost[s] = cost;
  S.in_use += cost;
  *ticket = t;
  return ZETA_OK;
}

extern "C" int zeta_serve_poll(uint64_t ticket, zeta_result *res) {
  int r = sv_require();
  if (r) return r;
  sv_heartbeat();
  if (!res) return zr::fail(ZETA_EINVAL, "zeta_serve_poll: null result");
  for (int s = 0; s < g_sv.nslots; s++) {
    if (g_sv.slot_ticket[s] != ticket + 1) continue;
    if (!__atomic_load_n(&g_sv.slots[s].state, __ATOMIC_ACQUIRE)) {
      if ((r = sv_check_alive())) return r;
      return ZETA_PENDING;
    }
    sv_take(s, res);
    return ZETA_OK;
  }
  return zr::fail(ZETA_EINVAL, "zeta_serve_poll: ticket %llu is not in flight", (unsigned long long)ticket);
}

extern "C" int zeta_serve_reap(zeta_serve_done *out, int max, int *n) {
  int r = sv_require();
  if (r) return r;
  sv_heartbeat();
  if (!out || !n || max < 0) return zr::fail(ZETA_EINVAL, "zeta_serve_reap: bad arguments");
  int k = 0;
  for (int s = 0; s < g_sv.nslots && k < max; s++) {
    if (!g_sv.slot_ticket[s] || !__atomic_load_n(&g_sv.slots[s].state, __ATOMIC_ACQUIRE)) continue;
    out[k].ticket = g_sv.slot_ticket[s] - 1;
    sv_take(s, &out[k].res);
    k++;
  }
  *n = k;
  if (k == 0 && (r = sv_check_alive())) return r;
  return ZETA_OK;
}

extern "C" int zeta_serve_cancel(uint64_t ticket) {
  int r = sv_require();
  if (r) return r;
  sv_heartbeat();
  for (int s = 0; s < g_sv.nslots; s++) {
    if (g_sv.slot_ticket[s] != ticket + 1) continue;
    __atomic_store_n(&g_sv.cancel[s], ticket + 1, __ATOMIC_RELAXED);
    __atomic_fetch_add(&g_sv.sh->cancel_gen, 1u, __ATOMIC_RELEASE);
    return ZETA_OK;
  }
  return zr::fail(ZETA_EINVAL, "zeta_serve_cancel: ticket %llu is not in flight", (unsigned long long)ticket);
}

extern "C" int zeta_serve_stop(uint64_t *total_nodes) {
  int r = sv_require();
  if (r) return r;
  ServeState &S = g_sv;
  __atomic_store_n(&S.sh->shutdown, 1u, __ATOMIC_RELEASE);
  sv_heartbeat();
  hipError_t e = hipStreamSynchronize(zr::ctx().stream);
  u64 ctr[C_NCTR] = {0};
  u32 err = 0;
  if (e == hipSuccess) e = hipMemcpy(ctr, S.ctr.p, sizeof(ctr), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&err, S.err.p, 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) memcpy(g_ss.last_ctr, ctr, sizeof(ctr));
  if (total_nodes) *total_nodes = ctr[C_NODES];
  sv_free_all();
  if (e != hipSuccess) return zr::hip_fail(e, "zeta_serve_stop", __FILE__, __LINE__);
  return ZETA_OK;
}

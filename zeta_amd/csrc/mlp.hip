// mlp.hip -- training of the learned move-ordering network on gfx950:
// BF16 MFMA forward and backward GEMMs with FP32 accumulation, pairwise
// ranking loss, Adam with FP32 master weights.  No reference counterpart
// (the reference orders moves with fixed scores, src/zeta.cl:2059-2104).
//
// Training examples are pairs logged by the search at cut nodes: the move
// that caused the beta cutoff and another move of the node.  A step takes P
// pairs = B = 2P examples (row 2p the cutoff move of pair p, row 2p+1 the
// other) and minimises
//   L = mean_p softplus(s(neg_p) - s(pos_p) + o_p / PLAIN_PER_LOGIT),
// o_p the plain ordering's score difference of the pair: the search ranks
// by plain score + PLAIN_PER_LOGIT x s, so s learns a correction of it.
// One step is a feature pass and five products of one NT GEMM kernel,
// C[M][N] = sum_k A[m][k] Bt[n][k], both operands BF16 with k contiguous,
// one 32x32 C tile per wavefront built from v_mfma_f32_32x32x16_bf16, each
// with its consumer fused in as the epilogue:
//   z1   [B][256]   = X     [B][928]  . W1t  [256][928]  -> h1 = relu(z1 + b1)
//   z2   [B][32]    = H1    [B][256]  . W2t  [32][256]   -> scores, loss, dz2
//   gW2t [32][256]  = dZ2t  [32][B]   . H1t  [256][B]    split over K
//   dH1  [B][256]   = dZ2   [B][32]   . W2   [256][32]   -> dz1, b1 gradient
//   gW1t [256][928] = dZ1t  [256][B]  . Xt   [928][B]    split over K
#include <math.h>
#include <string.h>

#include <vector>

#include "mlp.h"
#include "runtime.h"

using namespace zm;
using zc::u32;
using zc::u64;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// internal FP32 layout (master weights, gradients, Adam moments)
constexpr int I_W1T = 0;                      // [256][928]
constexpr int I_B1 = I_W1T + NH1 * NINP;      // [256]
constexpr int I_W2T = I_B1 + NH1;             // [32][256]
constexpr int I_B2 = I_W2T + NH2 * NH1;       // [32]
constexpr int I_W3 = I_B2 + NH2;              // [32]
constexpr int I_B3 = I_W3 + NH2;              // [1]
constexpr int NINT = I_B3 + 1;

struct PairRec {  // one logged training pair (48 B)
  u64 b0, b1, b2, b3;
  u32 stm, mpos, mneg, depth;
};

constexpr u16 BF_ONE = 0x3F80;

// ---- examples of a batch: pair p gives rows 2p (its cutoff move) and 2p+1
// (the other move), so every 32-row tile of a product holds 16 whole pairs
// and the loss is an epilogue of the layer-2 product.

// One-hot features of 32 rows (16 pairs) per block, written densely in both
// layouts (X [B][928] and Xt [928][B]) so that no clearing pass is needed:
// the block (256 threads) collects a 32-bit row mask per feature in LDS,
// then stores 16-B chunks of X rows and 64-B row segments of Xt.
// record index of pair i of a training step: splitmix64 of (seed, i) mod nrec
__device__ __forceinline__ u32 sample_index(u64 seed, int i, u64 nrec) {
  u64 z = seed + (u64)(i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (u32)(z % nrec);
}

// nrec > 0: the step's pairs are drawn here (sample_index) and their indices
// written to idx for the later kernels; nrec = 0: idx holds them already
__global__ __launch_bounds__(256) void k_features(const PairRec *__restrict__ recs, u32 *__restrict__ idx,
                                                  int B, u16 *X, u16 *Xt, u64 nrec, u64 seed) {
  __shared__ u32 mask[NINP];
  const int e0 = blockIdx.x * 32;
  for (int f = threadIdx.x; f < NINP; f += blockDim.x) mask[f] = 0;
  __syncthreads();
  {
    // eight threads per row, one rank of the board each (the first one also
    // the move's features)
    const int r = threadIdx.x >> 3, rank = threadIdx.x & 7, e = e0 + r;
    u32 ri;
    if (nrec) {
      ri = sample_index(seed, e >> 1, nrec);
      if (rank == 0 && !(e & 1)) idx[e >> 1] = ri;
    } else {
      ri = idx[e >> 1];
    }
    const PairRec rc = recs[ri];
    zc::Pos p;
    p.b0 = rc.b0; p.b1 = rc.b1; p.b2 = rc.b2; p.b3 = rc.b3; p.cr = 0; p.stm = rc.stm & 1u;
    const u32 bit = 1u << r;
    u64 occ = (p.b1 | p.b2 | p.b3) & (0xFFull << (8 * rank));
    while (occ) {
      const int sq = zc::lsb(occ);
      occ &= occ - 1;
      atomicOr(&mask[feat_ps(zc::piece_at(p, sq), sq, p.stm)], bit);
    }
    if (rank == 0) {
      const u32 m = (e & 1) ? rc.mneg : rc.mpos;
      atomicOr(&mask[feat_from(m, p.stm)], bit);
      atomicOr(&mask[feat_to(m, p.stm)], bit);
      atomicOr(&mask[feat_piece(m)], bit);
      const int fc = feat_capt(m);
      if (fc >= 0) atomicOr(&mask[fc], bit);
    }
  }
  __syncthreads();
  constexpr int CH = NINP / 8;  // 16-B chunks per X row
  for (int c = threadIdx.x; c < 32 * CH; c += blockDim.x) {
    const int r = c / CH, f0 = (c % CH) * 8;
    u32 w[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      w[j] = (((mask[f0 + 2 * j] >> r) & 1u) ? BF_ONE : 0u) | (((mask[f0 + 2 * j + 1] >> r) & 1u) ? (u32)BF_ONE << 16 : 0u);
    *(uint4 *)(X + (size_t)(e0 + r) * NINP + f0) = make_uint4(w[0], w[1], w[2], w[3]);
  }
  for (int c = threadIdx.x; c < NINP * 4; c += blockDim.x) {
    const int f = c >> 2, r0 = (c & 3) * 8;
    const u32 mk = mask[f] >> r0;
    u32 w[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
      w[j] = (((mk >> (2 * j)) & 1u) ? BF_ONE : 0u) | (((mk >> (2 * j + 1)) & 1u) ? (u32)BF_ONE << 16 : 0u);
    *(uint4 *)(Xt + (size_t)f * B + e0 + r0) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// ---- NT GEMM on MFMA, C[M][N] = sum_k A[m][k] Bt[n][k], one wavefront per
// 32x32 tile of C (and K slice), with the consumer of each product fused in
// as its epilogue.  Accumulator layout of v_mfma_f32_32x32x16_bf16: register
// i of lane l holds C[8*(i/4) + 4*(l/32) + i%4][l%32].
enum { E_SPLITK, E_L1ACT, E_HEAD, E_L1BACK };
struct Epi {
  float *C;                   // E_SPLITK output, ldc = N
  const float *prm;           // master weights (biases, head)
  float *Zf;                  // E_L1ACT: h1 (FP32) out;  E_L1BACK: h1 in
  u16 *H, *Ht;                // E_L1ACT: h1 BF16 [M][N] and [N][M];  E_L1BACK: Ht = dz1 [N][M];  E_HEAD: dz2, dz2t
  float *grad, *loss;
  const PairRec *recs;
  const u32 *idx;
  int P;
};

__device__ __forceinline__ uint2 pack4(const u16 *v) {
  return make_uint2((u32)v[0] | ((u32)v[1] << 16), (u32)v[2] | ((u32)v[3] << 16));
}

// E_SPLITK: the four wavefronts of a block take four K slices of one tile and
// add their partial tiles in LDS; the block then adds the sum into C (zeroed
// before) with FP32 atomics, ksplit/4 blocks per tile.  E_HEAD: one block per
// tile, its four wavefronts take four K slices and, after the LDS sum, one
// row group (accumulator registers 4w..4w+3) each for the epilogue.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_nt(const u16 *__restrict__ A, int lda, const u16 *__restrict__ Bt,
                                                 int ldb, int M, int N, int K, int ksplit, Epi ep) {
  const int lane = threadIdx.x & 63;
  const int tn_count = N / 32;
  int tile, ks;
  if (EPI == E_SPLITK || EPI == E_HEAD) {
    const int bpt = ksplit / 4;
    tile = blockIdx.x / bpt;
    ks = (blockIdx.x % bpt) * 4 + (threadIdx.x >> 6);
  } else if (EPI == E_L1BACK) {
    // the four wavefronts of a block share a column tile (their b1 column
    // sums meet in LDS)
    tile = ((int)(blockIdx.x / tn_count) * 4 + (int)(threadIdx.x >> 6)) * tn_count + (int)(blockIdx.x % tn_count);
    ks = 0;
  } else {
    tile = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    ks = 0;
  }
  const int tm = tile / tn_count, tn = tile % tn_count;
  if (EPI != E_SPLITK && EPI != E_HEAD && EPI != E_L1BACK && tm * 32 >= M) return;
  // K slices of whole 16-steps; the last ones may be short or empty
  const bool live = tm * 32 < M;  // E_L1BACK: false for the waves past M
  const int kc = ((K / 16 + ksplit - 1) / ksplit) * 16, k0 = ks * kc;
  const int kn = !live ? 0 : (K - k0 < kc ? K - k0 : kc);
  const int r = tm * 32 + (lane & 31), c = tn * 32 + (lane & 31);
  const int kofs = (lane >> 5) * 8;
  const u16 *pa = A + (size_t)r * lda + kofs + k0;
  const u16 *pb = Bt + (size_t)c * ldb + kofs + k0;
  // E_HEAD: the plain score differences of this lane's two pairs, loaded
  // (index, then record) while the products run
  u32 hdw[2] = {0u, 0u};
  if (EPI == E_HEAD) {
    const int rb = tm * 32 + 8 * (threadIdx.x >> 6) + 4 * (lane >> 5);
    hdw[0] = ep.recs[ep.idx[rb >> 1]].depth;
    hdw[1] = ep.recs[ep.idx[(rb + 2) >> 1]].depth;
  }
  f32x16 acc = {};
  // four 16-steps per trip, the next trip's operands loaded before this
  // trip's MFMAs: the loads of one trip overlap the products of the last
  // instead of each product waiting for its own load
  int k = 0;
  if (kn >= 64) {
    bf16x8 a[4], b[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      a[j] = *(const bf16x8 *)(pa + 16 * j);
      b[j] = *(const bf16x8 *)(pb + 16 * j);
    }
    for (; k + 128 <= kn; k += 64) {
      bf16x8 na[4], nb[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        na[j] = *(const bf16x8 *)(pa + k + 64 + 16 * j);
        nb[j] = *(const bf16x8 *)(pb + k + 64 + 16 * j);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], b[j], acc, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        a[j] = na[j];
        b[j] = nb[j];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], b[j], acc, 0, 0, 0);
    k += 64;
  }
  for (; k < kn; k += 16) {
    const bf16x8 a = *(const bf16x8 *)(pa + k);
    const bf16x8 b = *(const bf16x8 *)(pb + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  const int half = lane >> 5, col = tn * 32 + (lane & 31);
  if (EPI == E_SPLITK) {
    __shared__ float red[3][16][64];
    const int w = threadIdx.x >> 6;
    if (w) {
#pragma unroll
      for (int i = 0; i < 16; i++) red[w - 1][i][lane] = acc[i];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int i = 0; i < 16; i++)
        atomicAdd(ep.C + (size_t)(tm * 32 + 8 * (i >> 2) + 4 * half + (i & 3)) * N + col,
                  acc[i] + red[0][i][lane] + red[1][i][lane] + red[2][i][lane]);
    }
  } else if (EPI == E_L1ACT) {
    // h1 = relu(z1 + b1): FP32 (kept for relu' in the backward pass), BF16 in both layouts
    const float bb = ep.prm[I_B1 + col];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int rb = tm * 32 + 8 * g + 4 * half;
      u16 hb[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        float h = acc[4 * g + j] + bb;
        h = h > 0.f ? h : 0.f;
        ep.Zf[(size_t)(rb + j) * N + col] = h;
        hb[j] = f2bf(h);
        ep.H[(size_t)(rb + j) * N + col] = hb[j];
      }
      *(uint2 *)(ep.Ht + (size_t)col * M + rb) = pack4(hb);
    }
  } else if (EPI == E_L1BACK) {
    // dz1 = dh1 * relu'(h1): BF16 transposed for the W1 gradient; b1 gradient
    float cs = 0.f;
    if (live) {
#pragma unroll
      for (int g = 0; g < 4; g++) {
        const int rb = tm * 32 + 8 * g + 4 * half;
        u16 db[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const float dz = ep.Zf[(size_t)(rb + j) * N + col] > 0.f ? acc[4 * g + j] : 0.f;
          cs += dz;
          db[j] = f2bf(dz);
        }
        *(uint2 *)(ep.Ht + (size_t)col * M + rb) = pack4(db);
      }
    }
    cs += __shfl_xor(cs, 32, 64);
    __shared__ float cls[4][32];
    if (lane < 32) cls[threadIdx.x >> 6][lane] = cs;
    __syncthreads();
    if (threadIdx.x < 32)
      atomicAdd(ep.grad + I_B1 + col, cls[0][lane] + cls[1][lane] + cls[2][lane] + cls[3][lane]);
  } else if (EPI == E_HEAD) {
    // N = 32: the tile holds all of layer 2.  The K slices' partial tiles
    // are summed in LDS; wavefront w then takes rows 8w + 4*half + j (j < 4,
    // registers 4w + j): h2 = relu(z2 + b2), s = w3.h2 + b3 (sum over the 32
    // lanes of a half), and for its 2 pairs per half the loss
    // softplus(s_neg - s_pos + o/PLAIN_PER_LOGIT), dz2 and the b2, W3, b3
    // gradients and the loss summed per wavefront
    __shared__ float red[4][16][64];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 16; i++) red[w][i][lane] = acc[i];
    __syncthreads();
    const int o = lane & 31;
    const float b2o = ep.prm[I_B2 + o], w3o = ep.prm[I_W3 + o], b3 = ep.prm[I_B3];
    float h[4], sc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int i = 4 * w + j;
      const float z = red[0][i][lane] + red[1][i][lane] + red[2][i][lane] + red[3][i][lane] + b2o;
      h[j] = z > 0.f ? z : 0.f;
      float t = w3o * h[j];
#pragma unroll
      for (int x = 16; x > 0; x >>= 1) t += __shfl_xor(t, x, 64);
      sc[j] = t + b3;
    }
    const float invP = 1.f / (float)ep.P;
    float gb2 = 0.f, gw3 = 0.f, gb3 = 0.f, gl = 0.f;
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int j = 2 * q;  // rows j, j+1 of this lane's group: an even row (cutoff move) and its pair
      const int row = tm * 32 + 8 * w + 4 * half + j;
      const float d = sc[j + 1] - sc[j] + (float)pair_offset(hdw[q]) / (float)PLAIN_PER_LOGIT;
      const float sig = 1.f / (1.f + expf(-d));
      const float l = d > 0.f ? d + log1pf(expf(-d)) : log1pf(expf(d));
      const float ds0 = -sig * invP, ds1 = sig * invP;
      const float dz0 = h[j] > 0.f ? ds0 * w3o : 0.f, dz1 = h[j + 1] > 0.f ? ds1 * w3o : 0.f;
      const u16 d0 = f2bf(dz0), d1 = f2bf(dz1);
      ep.H[(size_t)row * NH2 + o] = d0;
      ep.H[(size_t)(row + 1) * NH2 + o] = d1;
      *(u32 *)(ep.Ht + (size_t)o * M + row) = (u32)d0 | ((u32)d1 << 16);
      gb2 += dz0 + dz1;
      gw3 += ds0 * h[j] + ds1 * h[j + 1];
      if (o == 0) {
        gb3 += ds0 + ds1;
        gl += l * invP;
      }
    }
    gb2 += __shfl_xor(gb2, 32, 64);
    gw3 += __shfl_xor(gw3, 32, 64);
    gb3 += __shfl_xor(gb3, 32, 64);
    gl += __shfl_xor(gl, 32, 64);
    // the four wavefronts' sums meet in LDS: one atomic per value and block
    // (every block adds into the same 66 words)
    __shared__ float hs[4][2 * NH2 + 2];
    if (lane < 32) {
      hs[w][o] = gb2;
      hs[w][NH2 + o] = gw3;
    }
    if (lane == 0) {
      hs[w][2 * NH2] = gb3;
      hs[w][2 * NH2 + 1] = gl;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 2 * NH2 + 2) {
      const float v = hs[0][t] + hs[1][t] + hs[2][t] + hs[3][t];
      float *dst = t < NH2 ? ep.grad + I_B2 + t
                           : (t < 2 * NH2 ? ep.grad + I_W3 + (t - NH2) : (t == 2 * NH2 ? ep.grad + I_B3 : ep.loss));
      atomicAdd(dst, v);
    }
  }
}

// Adam on the internal layout (padded W1 columns stay zero), and in the
// same pass the BF16 working copies of the updated W1 and W2
__global__ void k_adam(float *prm, const float *__restrict__ grad, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, u16 *w1t, u16 *w1f, u16 *w2t, u16 *w2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float p = prm[i];
  if (!(i < I_B1 && (i % NINP) >= NIN)) {
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    prm[i] = p;
  }
  if (i < I_B1) {
    const u16 b = f2bf(p);
    w1t[i] = b;
    w1f[(i % NINP) * NH1 + i / NINP] = b;
  } else if (i >= I_W2T && i < I_B2) {
    const int j = i - I_W2T;
    const u16 b = f2bf(p);
    w2t[j] = b;
    w2[(j % NH1) * NH2 + j / NH1] = b;
  }
}

// BF16 working copies of the master weights
__global__ void k_refresh(const float *__restrict__ prm, u16 *w1t, u16 *w1f, u16 *w2t, u16 *w2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < NH1 * NINP) {
    const u16 b = f2bf(prm[I_W1T + i]);
    w1t[i] = b;
    const int u = i / NINP, f = i % NINP;
    w1f[f * NH1 + u] = b;
  }
  if (i < NH2 * NH1) {
    const u16 b = f2bf(prm[I_W2T + i]);
    w2t[i] = b;
    const int o = i / NH1, u = i % NH1;
    w2[u * NH2 + o] = b;
  }
}

// deterministic initialisation (He-uniform), splitmix64 stream per element
__device__ __forceinline__ float unif(u64 seed, u64 i) {
  u64 z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (float)((z >> 40) * (1.0 / 16777216.0)) * 2.f - 1.f;
}
__global__ void k_init(float *prm, u64 seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NINT) return;
  float v = 0.f;
  if (i < I_B1) v = ((i % NINP) < NIN) ? unif(seed, i) * sqrtf(6.f / 40.f) : 0.f;  // ~40 active inputs
  else if (i >= I_W2T && i < I_B2) v = unif(seed, i) * sqrtf(6.f / NH1);
  else if (i >= I_W3 && i < I_B3) v = unif(seed, i) * sqrtf(6.f / NH2);
  prm[i] = v;
}

struct MlpState {
  zr::DevBuf prm, grad, m, v, w1t, w1f, w2t, w2;     // weights
  zr::DevBuf X, Xt, Z1, H1, H1t, dZ2, dZ2t, dZ1t, idx, loss, recs_scratch;
  int cap_pairs = 0;
  bool ready = false;
  int step = 0;
};
MlpState g_ms;

int ensure_weights() {
  int r;
  if ((r = zr::ensure(g_ms.prm, NINT * 4)) || (r = zr::ensure(g_ms.grad, NINT * 4)) ||
      (r = zr::ensure(g_ms.m, NINT * 4)) || (r = zr::ensure(g_ms.v, NINT * 4)) ||
      (r = zr::ensure(g_ms.w1t, NH1 * NINP * 2)) || (r = zr::ensure(g_ms.w1f, NH1 * NINP * 2)) ||
      (r = zr::ensure(g_ms.w2t, NH2 * NH1 * 2)) ||
      (r = zr::ensure(g_ms.w2, NH2 * NH1 * 2)))
    return r;
  return ZETA_OK;
}

int ensure_batch(int P) {
  if (g_ms.cap_pairs >= P) return ZETA_OK;
  const size_t B = 2 * (size_t)P;
  int r;
  if ((r = zr::ensure(g_ms.X, B * NINP * 2)) || (r = zr::ensure(g_ms.Xt, B * NINP * 2)) ||
      (r = zr::ensure(g_ms.Z1, B * NH1 * 4)) || (r = zr::ensure(g_ms.H1, B * NH1 * 2)) ||
      (r = zr::ensure(g_ms.H1t, B * NH1 * 2)) || (r = zr::ensure(g_ms.dZ2, B * NH2 * 2)) ||
      (r = zr::ensure(g_ms.dZ2t, B * NH2 * 2)) || (r = zr::ensure(g_ms.dZ1t, B * NH1 * 2)) ||
      (r = zr::ensure(g_ms.idx, (size_t)P * 4)) || (r = zr::ensure(g_ms.loss, 64)))
    return r;
  g_ms.cap_pairs = P;
  return ZETA_OK;
}

inline unsigned nb(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// C = A . Bt^T with epilogue EPI.  E_SPLITK: C was zeroed and K slices add
// into it, so a product with few output tiles splits its K to fill the device
template <int EPI>
int gemm(const u16 *A, int lda, const u16 *Bt, int ldb, int M, int N, int K, int kid, const Epi &ep) {
  const size_t tiles = (size_t)(M / 32) * (N / 32);
  int ksplit = 1;
  if (EPI == E_SPLITK) {
    ksplit = 4;
    while (tiles * ksplit < 1024 && K / (2 * ksplit) >= 128) ksplit *= 2;
  } else if (EPI == E_HEAD) {
    if (N != 32) return zr::fail(ZETA_EINVAL, "mlp gemm: the head product needs N = 32");
    ksplit = 4;
  }
  if (M % 32 || N % 32 || K % 16) return zr::fail(ZETA_EINVAL, "mlp gemm: shape %d x %d x %d", M, N, K);
  // E_L1BACK: blocks of four row tiles in one column tile (the last block's
  // extra waves, past M, only join its barrier)
  const size_t blocks = EPI == E_L1BACK ? (size_t)((M / 32 + 3) / 4) * (N / 32) : nb(tiles * ksplit * 64, 256);
  ZLAUNCH(kid, (u64)M * N * K * 2, k_gemm_nt<EPI>, blocks, 256, 0, A, lda, Bt, ldb, M, N, K, ksplit, ep);
  return ZETA_OK;
}

// forward + backward of one batch of P pairs; gradients into g_ms.grad, mean
// loss into g_ms.loss.  nrec = 0: the record indices are in g_ms.idx already;
// nrec > 0: the feature kernel draws them from nrec records with seed
int fwd_bwd(const PairRec *recs, int P, u64 nrec, u64 seed) {
  hipStream_t st = zr::ctx().stream;
  const int B = 2 * P;
  int r;
  ZCHECK(hipMemsetAsync(g_ms.grad.p, 0, NINT * 4, st));
  ZCHECK(hipMemsetAsync(g_ms.loss.p, 0, 4, st));
  u32 *idx = (u32 *)g_ms.idx.p;
  ZLAUNCH(zr::K_MLP_FWD, B, k_features, B / 32, 256, 0, recs, idx, B, (u16 *)g_ms.X.p, (u16 *)g_ms.Xt.p, nrec,
          seed);
  Epi ep = {};
  ep.prm = (const float *)g_ms.prm.p;
  ep.grad = (float *)g_ms.grad.p;
  ep.loss = (float *)g_ms.loss.p;
  ep.recs = recs;
  ep.idx = idx;
  ep.P = P;
  // z1 = X . W1t^T -> h1
  ep.Zf = (float *)g_ms.Z1.p;
  ep.H = (u16 *)g_ms.H1.p;
  ep.Ht = (u16 *)g_ms.H1t.p;
  if ((r = gemm<E_L1ACT>((const u16 *)g_ms.X.p, NINP, (const u16 *)g_ms.w1t.p, NINP, B, NH1, NINP, zr::K_MLP_FWD,
                         ep)))
    return r;
  // z2 = H1 . W2t^T -> scores, loss, dz2 and the head's gradients
  ep.H = (u16 *)g_ms.dZ2.p;
  ep.Ht = (u16 *)g_ms.dZ2t.p;
  if ((r = gemm<E_HEAD>((const u16 *)g_ms.H1.p, NH1, (const u16 *)g_ms.w2t.p, NH1, B, NH2, NH1, zr::K_MLP_BWD, ep)))
    return r;
  // gW2t = dZ2t . H1t^T
  ep.C = (float *)g_ms.grad.p + I_W2T;
  if ((r = gemm<E_SPLITK>((const u16 *)g_ms.dZ2t.p, B, (const u16 *)g_ms.H1t.p, B, NH2, NH1, B, zr::K_MLP_BWD, ep)))
    return r;
  // dh1 = dZ2 . W2^T -> dz1 (transposed), b1 gradient
  ep.Ht = (u16 *)g_ms.dZ1t.p;
  if ((r = gemm<E_L1BACK>((const u16 *)g_ms.dZ2.p, NH2, (const u16 *)g_ms.w2.p, NH2, B, NH1, NH2, zr::K_MLP_BWD,
                          ep)))
    return r;
  // gW1t = dZ1t . Xt^T
  ep.C = (float *)g_ms.grad.p + I_W1T;
  if ((r = gemm<E_SPLITK>((const u16 *)g_ms.dZ1t.p, B, (const u16 *)g_ms.Xt.p, B, NH1, NINP, B, zr::K_MLP_BWD, ep)))
    return r;
  return ZETA_OK;
}

int refresh() {
  ZLAUNCH(zr::K_MLP_ADAM, NH1 * NINP, k_refresh, nb((size_t)NH1 * NINP, 256), 256, 0, (const float *)g_ms.prm.p,
          (u16 *)g_ms.w1t.p, (u16 *)g_ms.w1f.p, (u16 *)g_ms.w2t.p, (u16 *)g_ms.w2.p);
  return ZETA_OK;
}

// ABI layout <-> internal layout
void to_internal(const float *in, std::vector<float> &out) {
  out.assign(NINT, 0.f);
  for (int f = 0; f < NIN; f++)
    for (int u = 0; u < NH1; u++) out[I_W1T + u * NINP + f] = in[P_W1 + f * NH1 + u];
  for (int u = 0; u < NH1; u++) out[I_B1 + u] = in[P_B1 + u];
  for (int u = 0; u < NH1; u++)
    for (int o = 0; o < NH2; o++) out[I_W2T + o * NH1 + u] = in[P_W2 + u * NH2 + o];
  for (int o = 0; o < NH2; o++) out[I_B2 + o] = in[P_B2 + o];
  for (int o = 0; o < NH2; o++) out[I_W3 + o] = in[P_W3 + o];
  out[I_B3] = in[P_B3];
}
void to_abi(const std::vector<float> &in, float *out) {
  for (int f = 0; f < NIN; f++)
    for (int u = 0; u < NH1; u++) out[P_W1 + f * NH1 + u] = in[I_W1T + u * NINP + f];
  for (int u = 0; u < NH1; u++) out[P_B1 + u] = in[I_B1 + u];
  for (int u = 0; u < NH1; u++)
    for (int o = 0; o < NH2; o++) out[P_W2 + u * NH2 + o] = in[I_W2T + o * NH1 + u];
  for (int o = 0; o < NH2; o++) out[P_B2 + o] = in[I_B2 + o];
  for (int o = 0; o < NH2; o++) out[P_W3 + o] = in[I_W3 + o];
  out[P_B3] = in[I_B3];
}

}  // namespace

// ---- records (training pairs) shared with the search kernel
namespace zrec {
zr::DevBuf g_recs, g_count;
u64 g_cap = 0;
int ensure(u64 cap) {
  int r;
  if ((r = zr::ensure(g_count, 64))) return r;
  if (g_cap >= cap && g_recs.p) return ZETA_OK;
  // keep already logged pairs when growing
  u64 n = 0;
  if (g_recs.p) {
    ZCHECK(hipMemcpyAsync(&n, g_count.p, 8, hipMemcpyDeviceToHost, zr::ctx().stream));
    ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  }
  zr::DevBuf nb2;
  if ((r = zr::ensure(nb2, cap * sizeof(PairRec)))) return r;
  if (g_recs.p && n) {
    // the count runs on past the capacity (the search counts the pairs it
    // could not store): keep what the old buffer holds, and count that
    u64 keep = n < g_cap ? n : g_cap;
    if (keep > cap) keep = cap;
    ZCHECK(hipMemcpyAsync(nb2.p, g_recs.p, keep * sizeof(PairRec), hipMemcpyDeviceToDevice, zr::ctx().stream));
    ZCHECK(hipMemcpyAsync(g_count.p, &keep, 8, hipMemcpyHostToDevice, zr::ctx().stream));
    ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  } else {
    ZCHECK(hipMemsetAsync(g_count.p, 0, 8, zr::ctx().stream));
  }
  zr::release(g_recs);
  g_recs = nb2;
  g_cap = cap;
  return ZETA_OK;
}
}  // namespace zrec

void *records_buffer(u64 *cap, u64 **count) {
  *cap = zrec::g_cap;
  *count = (u64 *)zrec::g_count.p;
  return zrec::g_recs.p;
}

const zm::Weights *mlp_weights() {
  static zm::Weights w;
  if (!g_ms.ready) return nullptr;
  const float *prm = (const float *)g_ms.prm.p;
  w.w1t = (const u16 *)g_ms.w1t.p;
  w.w1f = (const u16 *)g_ms.w1f.p;
  w.w2t = (const u16 *)g_ms.w2t.p;
  w.b1 = prm + I_B1;
  w.b2 = prm + I_B2;
  w.w3 = prm + I_W3;
  w.b3 = prm + I_B3;
  return &w;
}

void mlp_release() {
  zr::DevBuf *bufs[] = {&g_ms.prm, &g_ms.grad, &g_ms.m,   &g_ms.v,    &g_ms.w1t,  &g_ms.w1f, &g_ms.w2t, &g_ms.w2,
                        &g_ms.X,   &g_ms.Xt,   &g_ms.Z1,  &g_ms.H1,   &g_ms.H1t,  &g_ms.dZ2,
                        &g_ms.dZ2t, &g_ms.dZ1t, &g_ms.idx, &g_ms.loss, &g_ms.recs_scratch};
  for (auto *b : bufs) zr::release(*b);
  g_ms.cap_pairs = 0;
  g_ms.ready = false;
  zr::release(zrec::g_recs);
  zr::release(zrec::g_count);
  zrec::g_cap = 0;
}

extern "C" int zeta_mlp_init(uint64_t seed) {
  int r = zr::require_up();
  if (r) return r;
  if ((r = ensure_weights())) return r;
  hipStream_t st = zr::ctx().stream;
  ZLAUNCH(zr::K_MLP_ADAM, NINT, k_init, nb(NINT, 256), 256, 0, (float *)g_ms.prm.p, (u64)seed);
  ZCHECK(hipMemsetAsync(g_ms.m.p, 0, NINT * 4, st));
  ZCHECK(hipMemsetAsync(g_ms.v.p, 0, NINT * 4, st));
  if ((r = refresh())) return r;
  ZCHECK(hipStreamSynchronize(st));
  zr::time_collect();
  g_ms.step = 0;
  g_ms.ready = true;
  return ZETA_OK;
}

extern "C" int zeta_mlp_set(const float *params, size_t n) {
  int r = zr::require_up();
  if (r) return r;
  if (!params || n != (size_t)NPARAMS) return zr::fail(ZETA_EINVAL, "zeta_mlp_set: expected %d floats", NPARAMS);
  if ((r = ensure_weights())) return r;
  std::vector<float> in;
  to_internal(params, in);
  hipStream_t st = zr::ctx().stream;
  ZCHECK(hipMemcpyAsync(g_ms.prm.p, in.data(), NINT * 4, hipMemcpyHostToDevice, st));
  ZCHECK(hipMemsetAsync(g_ms.m.p, 0, NINT * 4, st));
  ZCHECK(hipMemsetAsync(g_ms.v.p, 0, NINT * 4, st));
  if ((r = refresh())) return r;
  ZCHECK(hipStreamSynchronize(st));
  zr::time_collect();
  g_ms.step = 0;
  g_ms.ready = true;
  return ZETA_OK;
}

extern "C" int zeta_mlp_get(float *params, size_t n) {
  int r = zr::require_up();
  if (r) return r;
  if (!g_ms.ready) return zr::fail(ZETA_ESTATE, "zeta_mlp_get: no network (zeta_mlp_init / zeta_mlp_set first)");
  if (!params || n != (size_t)NPARAMS) return zr::fail(ZETA_EINVAL, "zeta_mlp_get: expected %d floats", NPARAMS);
  std::vector<float> in(NINT);
  ZCHECK(hipMemcpyAsync(in.data(), g_ms.prm.p, NINT * 4, hipMemcpyDeviceToHost, zr::ctx().stream));
  ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  to_abi(in, params);
  return ZETA_OK;
}

// gradient of one batch without an update (ABI layout), for gradient checks
extern "C" int zeta_mlp_grad(const void *pairs, int npairs, float *loss, float *grads, size_t n) {
  int r = zr::require_up();
  if (r) return r;
  if (!g_ms.ready) return zr::fail(ZETA_ESTATE, "zeta_mlp_grad: no network");
  if (!pairs || npairs <= 0 || npairs % 16 != 0 || !grads || n != (size_t)NPARAMS)
    return zr::fail(ZETA_EINVAL, "zeta_mlp_grad: npairs must be a positive multiple of 16");
  if ((r = ensure_batch(npairs)) || (r = zr::ensure(g_ms.recs_scratch, (size_t)npairs * sizeof(PairRec))))
    return r;
  hipStream_t st = zr::ctx().stream;
  std::vector<u32> idx(npairs);
  for (int i = 0; i < npairs; i++) idx[i] = (u32)i;
  ZCHECK(hipMemcpyAsync(g_ms.recs_scratch.p, pairs, (size_t)npairs * sizeof(PairRec), hipMemcpyHostToDevice, st));
  ZCHECK(hipMemcpyAsync(g_ms.idx.p, idx.data(), (size_t)npairs * 4, hipMemcpyHostToDevice, st));
  if ((r = fwd_bwd((const PairRec *)g_ms.recs_scratch.p, npairs, 0, 0))) return r;
  std::vector<float> g(NINT);
  float l = 0.f;
  ZCHECK(hipMemcpyAsync(g.data(), g_ms.grad.p, NINT * 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipMemcpyAsync(&l, g_ms.loss.p, 4, hipMemcpyDeviceToHost, st));
  ZCHECK(hipStreamSynchronize(st));
  zr::time_collect();
  to_abi(g, grads);
  if (loss) *loss = l;
  return ZETA_OK;
}

// Adam training on the logged pairs; losses[i] = mean loss of step i
extern "C" int zeta_mlp_train(int steps, int pairs_per_step, float lr, uint64_t seed, float *losses) {
  int r = zr::require_up();
  if (r) return r;
  if (!g_ms.ready) return zr::fail(ZETA_ESTATE, "zeta_mlp_train: no network (zeta_mlp_init first)");
  if (steps < 0 || pairs_per_step <= 0 || pairs_per_step % 16 != 0)
    return zr::fail(ZETA_EINVAL, "zeta_mlp_train: pairs_per_step must be a positive multiple of 16");
  u64 nrec = 0;
  if (zrec::g_count.p) {
    ZCHECK(hipMemcpyAsync(&nrec, zrec::g_count.p, 8, hipMemcpyDeviceToHost, zr::ctx().stream));
    ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  }
  if (nrec > zrec::g_cap) nrec = zrec::g_cap;
  if (nrec == 0) return zr::fail(ZETA_ESTATE, "zeta_mlp_train: no logged training pairs");
  if ((r = ensure_batch(pairs_per_step))) return r;
  hipStream_t st = zr::ctx().stream;
  std::vector<float> lv(steps > 0 ? steps : 1);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int s = 0; s < steps; s++) {
    if ((r = fwd_bwd((const PairRec *)zrec::g_recs.p, pairs_per_step, nrec,
                     (u64)seed + (u64)s * 0xD1B54A32D192ED03ULL)))
      return r;
    g_ms.step++;
    const float c1 = 1.f - powf(b1, (float)g_ms.step), c2 = 1.f - powf(b2, (float)g_ms.step);
    ZLAUNCH(zr::K_MLP_ADAM, NINT, k_adam, nb(NINT, 256), 256, 0, (float *)g_ms.prm.p, (const float *)g_ms.grad.p,
            (float *)g_ms.m.p, (float *)g_ms.v.p, NINT, lr, b1, b2, 1e-8f, c1, c2, (u16 *)g_ms.w1t.p,
            (u16 *)g_ms.w1f.p, (u16 *)g_ms.w2t.p, (u16 *)g_ms.w2.p);
    if (losses) ZCHECK(hipMemcpyAsync(&lv[s], g_ms.loss.p, 4, hipMemcpyDeviceToHost, st));
  }
  ZCHECK(hipStreamSynchronize(st));
  zr::time_collect();
  if (losses)
    for (int s = 0; s < steps; s++) losses[s] = lv[s];
  return ZETA_OK;
}

extern "C" int zeta_records_reserve(uint64_t capacity) {
  int r = zr::require_up();
  if (r) return r;
  if (capacity == 0) return zr::fail(ZETA_EINVAL, "zeta_records_reserve: capacity 0");
  return zrec::ensure(capacity);
}

extern "C" int zeta_records_count(uint64_t *n) {
  int r = zr::require_up();
  if (r) return r;
  if (!n) return ZETA_EINVAL;
  *n = 0;
  if (!zrec::g_count.p) return ZETA_OK;
  ZCHECK(hipMemcpyAsync(n, zrec::g_count.p, 8, hipMemcpyDeviceToHost, zr::ctx().stream));
  ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  if (*n > zrec::g_cap) *n = zrec::g_cap;
  return ZETA_OK;
}

extern "C" int zeta_records_clear(void) {
  int r = zr::require_up();
  if (r) return r;
  if (!zrec::g_count.p) return ZETA_OK;
  ZCHECK(hipMemsetAsync(zrec::g_count.p, 0, 8, zr::ctx().stream));
  ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  return ZETA_OK;
}

extern "C" int zeta_records_get(void *out, uint64_t max, uint64_t *n) {
  int r = zr::require_up();
  if (r) return r;
  u64 have = 0;
  if ((r = zeta_records_count(&have))) return r;
  const u64 k = have < max ? have : max;
  if (n) *n = k;
  if (k && !out) return ZETA_EINVAL;
  if (k) {
    ZCHECK(hipMemcpyAsync(out, zrec::g_recs.p, k * sizeof(PairRec), hipMemcpyDeviceToHost, zr::ctx().stream));
    ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  }
  return ZETA_OK;
}

extern "C" int zeta_records_put(const void *in, uint64_t n) {
  int r = zr::require_up();
  if (r) return r;
  if (n && !in) return ZETA_EINVAL;
  if ((r = zrec::ensure(n ? n : 1))) return r;
  if (n) ZCHECK(hipMemcpyAsync(zrec::g_recs.p, in, n * sizeof(PairRec), hipMemcpyHostToDevice, zr::ctx().stream));
  ZCHECK(hipMemcpyAsync(zrec::g_count.p, &n, 8, hipMemcpyHostToDevice, zr::ctx().stream));
  ZCHECK(hipStreamSynchronize(zr::ctx().stream));
  return ZETA_OK;
}

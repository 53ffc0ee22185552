// mlp.hip -- training of the learned move-ordering network on gfx950:
// BF16 MFMA forward and backward GEMMs with FP32 accumulation, pairwise
// ranking loss, Adam with FP32 master weights.  No reference counterpart
// (the reference orders moves with fixed scores, src/zeta.cl:2059-2104).
//
// Training examples are pairs logged by the plain search at cut nodes: the
// move that caused the beta cutoff and a move searched before it that did
// not (or one never searched).  A step takes P pairs = B = 2P examples (rows
// [0,P) the cutoff moves, rows [P,2P) the others) and minimises
//   L = mean_p softplus(s(neg_p) - s(pos_p) + o_p / PLAIN_PER_LOGIT),
// o_p the plain ordering's score difference of the pair: the search ranks
// by plain score + PLAIN_PER_LOGIT x s, so s learns a correction of it.
// All five products are one NT GEMM kernel, C[M][N] = sum_k A[m][k] Bt[n][k],
// both operands BF16 with k contiguous, one 32x32 C tile per wavefront built
// from v_mfma_f32_32x32x16_bf16:
//   z1   [B][256]   = X     [B][928]  . W1t  [256][928]
//   z2   [B][32]    = H1    [B][256]  . W2t  [32][256]
//   gW2t [32][256]  = dZ2t  [32][B]   . H1t  [256][B]
//   dH1  [B][256]   = dZ2   [B][32]   . W2   [256][32]
//   gW1t [256][928] = dZ1t  [256][B]  . Xt   [928][B]
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

// ---- NT GEMM on MFMA: one wavefront per 32x32 tile of C and slice of K.
// The weight-gradient products have few output tiles and a long K (the
// batch): their K is split over ksplit wavefronts per tile, which add their
// partial tiles into C (zeroed before) with FP32 atomics.
__global__ __launch_bounds__(256) void k_gemm_nt(const u16 *__restrict__ A, int lda, const u16 *__restrict__ Bt,
                                                 int ldb, float *__restrict__ C, int ldc, int M, int N, int K,
                                                 int ksplit) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int tn_count = N / 32;
  const int tile = wave / ksplit, ks = wave % ksplit;
  const int tm = tile / tn_count, tn = tile % tn_count;
  if (tm * 32 >= M) return;
  const int kc = K / ksplit, k0 = ks * kc;
  const int r = tm * 32 + (lane & 31), c = tn * 32 + (lane & 31);
  const int kofs = (lane >> 5) * 8;
  const u16 *pa = A + (size_t)r * lda + kofs + k0;
  const u16 *pb = Bt + (size_t)c * ldb + kofs + k0;
  f32x16 acc = {};
  for (int k = 0; k < kc; k += 16) {
    const bf16x8 a = *(const bf16x8 *)(pa + k);
    const bf16x8 b = *(const bf16x8 *)(pb + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  // accumulator layout: register i of lane l holds C[8*(i/4) + 4*(l/32) + i%4][l%32]
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const int row = tm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    float *dst = C + (size_t)row * ldc + tn * 32 + (lane & 31);
    if (ksplit == 1) *dst = acc[i];
    else atomicAdd(dst, acc[i]);
  }
}

// ---- one-hot features of each example, in both layouts (X and its transpose)
__device__ __forceinline__ void set_feat(u16 *X, u16 *Xt, int row, int f, int B) {
  X[(size_t)row * NINP + f] = BF_ONE;
  Xt[(size_t)f * B + row] = BF_ONE;
}
__global__ void k_features(const PairRec *__restrict__ recs, const u32 *__restrict__ idx, int P, u16 *X, u16 *Xt) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int B = 2 * P;
  if (e >= B) return;
  const PairRec rc = recs[idx[e % P]];
  const u32 m = e < P ? rc.mpos : rc.mneg;
  zc::Pos p;
  p.b0 = rc.b0; p.b1 = rc.b1; p.b2 = rc.b2; p.b3 = rc.b3; p.cr = 0; p.stm = rc.stm & 1u;
  u64 occ = p.b1 | p.b2 | p.b3;
  while (occ) {
    const int sq = zc::lsb(occ);
    occ &= occ - 1;
    set_feat(X, Xt, e, feat_ps(zc::piece_at(p, sq), sq, p.stm), B);
  }
  set_feat(X, Xt, e, feat_from(m, p.stm), B);
  set_feat(X, Xt, e, feat_to(m, p.stm), B);
  set_feat(X, Xt, e, feat_piece(m), B);
  const int fc = feat_capt(m);
  if (fc >= 0) set_feat(X, Xt, e, fc, B);
}

// h1 = relu(z1 + b1): FP32 in place, BF16 copies in both layouts
__global__ void k_l1_act(float *Z1, const float *__restrict__ b1, u16 *H1, u16 *H1t, int B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * NH1) return;
  const int e = i / NH1, u = i % NH1;
  float h = Z1[i] + b1[u];
  h = h > 0.f ? h : 0.f;
  Z1[i] = h;
  const u16 hb = f2bf(h);
  H1[i] = hb;
  H1t[(size_t)u * B + e] = hb;
}

// layer 2 activation, layer 3, pairwise ranking loss and the gradients that
// need no GEMM; one thread per pair (rows p and P+p).  The sums over pairs
// (b2, W3, b3 gradients, the loss) are wave reductions, then one LDS slot per
// wave and one global atomic per block and value.
__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_head(const float *__restrict__ Z2, const float *__restrict__ prm, int P,
                                              const PairRec *__restrict__ recs, const u32 *__restrict__ idx,
                                              u16 *dZ2, u16 *dZ2t, float *grad, float *loss_out) {
  constexpr int NV = NH2 * 2 + 2;  // b2, W3, b3, loss
  __shared__ float red[4][NV];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int B = 2 * P;
  const float *b2 = prm + I_B2, *w3 = prm + I_W3;
  float gb2[NH2], gw3[NH2], gb3 = 0.f, gl = 0.f;
#pragma unroll
  for (int o = 0; o < NH2; o++) gb2[o] = gw3[o] = 0.f;
  if (p < P) {
    float h[2][NH2], s[2];
    for (int k = 0; k < 2; k++) {
      const int e = p + k * P;
      float acc = prm[I_B3];
#pragma unroll
      for (int o = 0; o < NH2; o++) {
        float z = Z2[(size_t)e * NH2 + o] + b2[o];
        z = z > 0.f ? z : 0.f;
        h[k][o] = z;
        acc += w3[o] * z;
      }
      s[k] = acc;
    }
    const float d = s[1] - s[0] + (float)pair_offset(recs[idx[p]].depth) / (float)PLAIN_PER_LOGIT;
    const float sig = 1.f / (1.f + expf(-d));
    const float l = d > 0.f ? d + log1pf(expf(-d)) : log1pf(expf(d));
    const float ds[2] = {-sig / (float)P, sig / (float)P};
    for (int k = 0; k < 2; k++) {
      const int e = p + k * P;
#pragma unroll
      for (int o = 0; o < NH2; o++) {
        const float dz = h[k][o] > 0.f ? ds[k] * w3[o] : 0.f;
        const u16 db = f2bf(dz);
        dZ2[(size_t)e * NH2 + o] = db;
        dZ2t[(size_t)o * B + e] = db;
        gb2[o] += dz;
        gw3[o] += ds[k] * h[k][o];
      }
      gb3 += ds[k];
    }
    gl = l / (float)P;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 0; o < NH2; o++) {
    const float a = wave_sum_f(gb2[o]), b = wave_sum_f(gw3[o]);
    if (lane == 0) {
      red[w][o] = a;
      red[w][NH2 + o] = b;
    }
  }
  gb3 = wave_sum_f(gb3);
  gl = wave_sum_f(gl);
  if (lane == 0) {
    red[w][2 * NH2] = gb3;
    red[w][2 * NH2 + 1] = gl;
  }
  __syncthreads();
  const int nw = (int)(blockDim.x >> 6);
  for (int i = threadIdx.x; i < NV; i += blockDim.x) {
    float t = 0.f;
    for (int k = 0; k < nw; k++) t += red[k][i];
    if (i < NH2) atomicAdd(&grad[I_B2 + i], t);
    else if (i < 2 * NH2) atomicAdd(&grad[I_W3 + i - NH2], t);
    else if (i == 2 * NH2) atomicAdd(&grad[I_B3], t);
    else atomicAdd(loss_out, t);
  }
}

// dz1 = dh1 * relu'(h1): BF16 transposed copy for the W1 gradient; b1 gradient
__global__ __launch_bounds__(256) void k_l1_back(const float *__restrict__ dH1, const float *__restrict__ H1f, int B,
                                                 u16 *dZ1t, float *grad) {
  // one block per hidden unit column strip: thread t handles examples t, t+256, ...
  const int u = blockIdx.x;
  float acc = 0.f;
  for (int e = threadIdx.x; e < B; e += blockDim.x) {
    const size_t i = (size_t)e * NH1 + u;
    const float dz = H1f[i] > 0.f ? dH1[i] : 0.f;
    dZ1t[(size_t)u * B + e] = f2bf(dz);
    acc += dz;
  }
  __shared__ float red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[I_B1 + u] = red[0];
}

// Adam on the internal layout; padded W1 columns stay zero
__global__ void k_adam(float *prm, const float *__restrict__ grad, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i < I_B1 && (i % NINP) >= NIN) return;
  const float g = grad[i];
  const float mi = b1 * m[i] + (1.f - b1) * g;
  const float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  prm[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
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

// random pair indices for a step
__global__ void k_sample(u32 *idx, int P, u64 nrec, u64 seed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  u64 z = seed + (u64)(i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  idx[i] = (u32)(z % nrec);
}

struct MlpState {
  zr::DevBuf prm, grad, m, v, w1t, w1f, w2t, w2;     // weights
  zr::DevBuf X, Xt, Z1, H1, H1t, Z2, dZ2, dZ2t, dH1, dZ1t, idx, loss, recs_scratch;
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
      (r = zr::ensure(g_ms.H1t, B * NH1 * 2)) || (r = zr::ensure(g_ms.Z2, B * NH2 * 4)) ||
      (r = zr::ensure(g_ms.dZ2, B * NH2 * 2)) || (r = zr::ensure(g_ms.dZ2t, B * NH2 * 2)) ||
      (r = zr::ensure(g_ms.dH1, B * NH1 * 4)) || (r = zr::ensure(g_ms.dZ1t, B * NH1 * 2)) ||
      (r = zr::ensure(g_ms.idx, (size_t)P * 4)) || (r = zr::ensure(g_ms.loss, 64)))
    return r;
  g_ms.cap_pairs = P;
  return ZETA_OK;
}

inline unsigned nb(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

// C = A . Bt^T.  accumulate = 1: C was zeroed and the K slices may add into
// it, so a product with few output tiles splits its K to fill the device
int gemm(const u16 *A, int lda, const u16 *Bt, int ldb, float *C, int ldc, int M, int N, int K, int kid,
         int accumulate = 0) {
  const size_t tiles = (size_t)(M / 32) * (N / 32);
  int ksplit = 1;
  while (accumulate && tiles * ksplit < 2048 && K % (32 * ksplit) == 0 && K / (2 * ksplit) >= 128) ksplit *= 2;
  ZLAUNCH(kid, (u64)M * N * K * 2, k_gemm_nt, nb(tiles * ksplit * 64, 256), 256, 0, A, lda, Bt, ldb, C, ldc, M, N,
          K, ksplit);
  return ZETA_OK;
}

// forward + backward of one batch of P pairs (indices already in g_ms.idx);
// gradients into g_ms.grad, mean loss into g_ms.loss
int fwd_bwd(const PairRec *recs, int P) {
  hipStream_t st = zr::ctx().stream;
  const int B = 2 * P;
  int r;
  ZCHECK(hipMemsetAsync(g_ms.X.p, 0, (size_t)B * NINP * 2, st));
  ZCHECK(hipMemsetAsync(g_ms.Xt.p, 0, (size_t)B * NINP * 2, st));
  ZCHECK(hipMemsetAsync(g_ms.grad.p, 0, NINT * 4, st));
  ZCHECK(hipMemsetAsync(g_ms.loss.p, 0, 4, st));
  ZLAUNCH(zr::K_MLP_FWD, B, k_features, nb(B, 256), 256, 0, recs, (const u32 *)g_ms.idx.p, P, (u16 *)g_ms.X.p,
          (u16 *)g_ms.Xt.p);
  const float *prm = (const float *)g_ms.prm.p;
  float *grad = (float *)g_ms.grad.p;
  if ((r = gemm((const u16 *)g_ms.X.p, NINP, (const u16 *)g_ms.w1t.p, NINP, (float *)g_ms.Z1.p, NH1, B, NH1, NINP,
                zr::K_MLP_FWD)))
    return r;
  ZLAUNCH(zr::K_MLP_FWD, B, k_l1_act, nb((size_t)B * NH1, 256), 256, 0, (float *)g_ms.Z1.p, prm + I_B1,
          (u16 *)g_ms.H1.p, (u16 *)g_ms.H1t.p, B);
  if ((r = gemm((const u16 *)g_ms.H1.p, NH1, (const u16 *)g_ms.w2t.p, NH1, (float *)g_ms.Z2.p, NH2, B, NH2, NH1,
                zr::K_MLP_FWD)))
    return r;
  ZLAUNCH(zr::K_MLP_BWD, P, k_head, nb(P, 256), 256, 0, (const float *)g_ms.Z2.p, prm, P, recs,
          (const u32 *)g_ms.idx.p, (u16 *)g_ms.dZ2.p, (u16 *)g_ms.dZ2t.p, grad, (float *)g_ms.loss.p);
  if ((r = gemm((const u16 *)g_ms.dZ2t.p, B, (const u16 *)g_ms.H1t.p, B, grad + I_W2T, NH1, NH2, NH1, B,
                zr::K_MLP_BWD, 1)))
    return r;
  if ((r = gemm((const u16 *)g_ms.dZ2.p, NH2, (const u16 *)g_ms.w2.p, NH2, (float *)g_ms.dH1.p, NH1, B, NH1, NH2,
                zr::K_MLP_BWD)))
    return r;
  ZLAUNCH(zr::K_MLP_BWD, B, k_l1_back, NH1, 256, 0, (const float *)g_ms.dH1.p, (const float *)g_ms.Z1.p, B,
          (u16 *)g_ms.dZ1t.p, grad);
  if ((r = gemm((const u16 *)g_ms.dZ1t.p, B, (const u16 *)g_ms.Xt.p, B, grad + I_W1T, NINP, NH1, NINP, B,
                zr::K_MLP_BWD, 1)))
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
                        &g_ms.X,   &g_ms.Xt,   &g_ms.Z1,  &g_ms.H1,   &g_ms.H1t,  &g_ms.Z2,  &g_ms.dZ2,
                        &g_ms.dZ2t, &g_ms.dH1, &g_ms.dZ1t, &g_ms.idx, &g_ms.loss, &g_ms.recs_scratch};
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
  if ((r = fwd_bwd((const PairRec *)g_ms.recs_scratch.p, npairs))) return r;
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
    ZLAUNCH(zr::K_MLP_FWD, pairs_per_step, k_sample, nb(pairs_per_step, 256), 256, 0, (u32 *)g_ms.idx.p,
            pairs_per_step, nrec, (u64)seed + (u64)s * 0xD1B54A32D192ED03ULL);
    if ((r = fwd_bwd((const PairRec *)zrec::g_recs.p, pairs_per_step))) return r;
    g_ms.step++;
    const float c1 = 1.f - powf(b1, (float)g_ms.step), c2 = 1.f - powf(b2, (float)g_ms.step);
    ZLAUNCH(zr::K_MLP_ADAM, NINT, k_adam, nb(NINT, 256), 256, 0, (float *)g_ms.prm.p, (const float *)g_ms.grad.p,
            (float *)g_ms.m.p, (float *)g_ms.v.p, NINT, lr, b1, b2, 1e-8f, c1, c2);
    if ((r = refresh())) return r;
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

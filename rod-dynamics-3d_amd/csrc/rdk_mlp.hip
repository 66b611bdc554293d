// rdk_mlp.hip -- learned PSOR warm start: a per-manifold MLP 12 -> 64 -> 64 -> 3
// (SiLU) that predicts the converged impulses of the manifolds the
// previous-step cache cannot start (new contacts, and contacts whose previous
// normal impulse was zero), from contact geometry and pre-solve relative
// velocity.  Manifolds with a previous normal impulse keep the cached start:
// their step-to-step change is ~1e-4 of the impulse unit, below what a BF16
// network resolves.  Inference runs on the matrix cores (BF16 weights and
// activations, FP32 accumulation, v_mfma_f32_16x16x16_bf16); training runs
// FP32 forward/backward kernels with AdamW.  Plain PSOR then runs from the
// start to the same residual, so the network only changes the starting point.
// There is no counterpart in the reference (its warm start is the
// previous-step cache, nsc_contact.cpp:305-317).
//
// Features (per manifold, dimensionless):
//   0      phi / (r_a + r_b)                      gap
//   1..3   R_a^T n                                normal in A's body frame
//   4..6   R_b^T n (0 for a wall)                 normal in B's body frame
//   7      v_n / v_s                              pre-solve normal velocity
//   8..9   (t1.v_rel, t2.v_rel) / v_s             pre-solve tangential velocity
//   10     mu
//   11     lambda_n(previous step) / lambda_s     0 when the pair is new
// with v_s = |g| dt (1e-3 if gravity is off) and lambda_s = v_s / g_n (an
// impulse that changes the normal velocity by v_s).  Outputs are the start
// impulses / lambda_s for (n, t1, t2), projected onto the friction cone.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "rdk_internal.hpp"

namespace rdk {

constexpr int MLP_IN = 12, MLP_H = 64, MLP_OUT = 3;
constexpr int OFF_W1 = 0;                           // [12][64]
constexpr int OFF_B1 = OFF_W1 + MLP_IN * MLP_H;     // [64]
constexpr int OFF_W2 = OFF_B1 + MLP_H;              // [64][64]
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;      // [64]
constexpr int OFF_W3 = OFF_B2 + MLP_H;              // [64][3]
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_OUT;    // [3]
constexpr int NPARAM = OFF_B3 + MLP_OUT;            // 5187

// packed BF16 B-operand fragments: layer1 (1 kb x 4 nb), layer2 (4 x 4), layer3 (4 x 1)
constexpr int FRAG = 64 * 4;  // bf16 per 16x16 fragment
constexpr int PK_L1 = 0, PK_L2 = PK_L1 + 4 * FRAG, PK_L3 = PK_L2 + 16 * FRAG, PK_TOTAL = PK_L3 + 4 * FRAG;

struct Mlp {
  DevBuf<float> params, adam_m, adam_v, partial, loss_part;
  DevBuf<uint16_t> packed;
  DevBuf<float> feat, out, dx, dy;
  DevBuf<double> lscale;
  int adam_t = 0;
  bool ready = false;
};

static Mlp* get_mlp(rdk_ctx* ctx) {
  if (!ctx->mlp) ctx->mlp = new Mlp();
  return (Mlp*)ctx->mlp;
}

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                    // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }

// ---- features -----------------------------------------------------------------
__device__ __forceinline__ d3 rot_inv(q4 q, d3 v) {
  // R^T v for the rotation of unit quaternion q (conjugate rotation)
  const d3 u = mk3(-q.x, -q.y, -q.z);
  const d3 t = cross(u, v) * 2.0;
  return v + t * q.w + cross(u, t);
}

struct FeatIn {
  const double* f;
  size_t stride;
  const int *a, *b;
  const uint64_t* prev_keys;
  const double* prev_ln;
  int nprev;
  const uint64_t* prev_wkeys;
  const double* prev_wln;
  int nprev_w;
};

__device__ __forceinline__ int lb64(const uint64_t* a, int n, uint64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_mlp_features(BodyPtrs B, const SysDev* __restrict__ sys, FeatIn F, int nm, double vs,
                               float* __restrict__ X, double* __restrict__ lscale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  auto at = [&](int fld) { return F.f[mf_off(fld, i, F.stride)]; };
  const int a = F.a[i], b = F.b[i];
  const bool wall = b < 0;
  const d3 n = mk3(at(0), at(1), at(2)), t1 = mk3(at(3), at(4), at(5)), t2 = mk3(at(6), at(7), at(8));
  const d3 ra = mk3(at(9), at(10), at(11)), rb = mk3(at(12), at(13), at(14));
  const double gn = at(35), phi = at(38), vnpre = at(39);
  const q4 qa = q4{B.q[0][a], B.q[1][a], B.q[2][a], B.q[3][a]};
  const d3 va = mk3(B.v[0][a], B.v[1][a], B.v[2][a]) + cross(mk3(B.w[0][a], B.w[1][a], B.w[2][a]), ra);
  d3 vrel = -va;
  d3 nb = mk3(0, 0, 0);
  double rsum = B.r[a];
  if (!wall) {
    const q4 qb = q4{B.q[0][b], B.q[1][b], B.q[2][b], B.q[3][b]};
    const d3 vb = mk3(B.v[0][b], B.v[1][b], B.v[2][b]) + cross(mk3(B.w[0][b], B.w[1][b], B.w[2][b]), rb);
    vrel = vb - va;
    nb = rot_inv(qb, n);
    rsum += B.r[b];
  }
  const d3 na = rot_inv(qa, n);
  const double ls = vs / gn;
  double lprev = 0.0;
  if (!wall) {
    const uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint64_t)(uint32_t)b;
    const int j = lb64(F.prev_keys, F.nprev, key);
    if (j < F.nprev && F.prev_keys[j] == key) lprev = F.prev_ln[j];
  } else if (F.nprev_w > 0) {
    const uint64_t key = ((uint64_t)(uint32_t)a << 2) | (uint64_t)(-1 - b);
    const int j = lb64(F.prev_wkeys, F.nprev_w, key);
    if (j < F.nprev_w && F.prev_wkeys[j] == key) lprev = F.prev_wln[j];
  }
  float* x = X + (size_t)i * MLP_IN;
  x[0] = (float)(phi / rsum);
  x[1] = (float)na.x;
  x[2] = (float)na.y;
  x[3] = (float)na.z;
  x[4] = (float)nb.x;
  x[5] = (float)nb.y;
  x[6] = (float)nb.z;
  x[7] = (float)(vnpre / vs);
  x[8] = (float)(dot(t1, vrel) / vs);
  x[9] = (float)(dot(t2, vrel) / vs);
  x[10] = (float)sys[B.sys[a]].mu;
  x[11] = (float)(dmax(lprev, 0.0) / ls);
  lscale[i] = ls;
}

// ---- BF16 weight packing ------------------------------------------------------
// B-operand fragment (16x16, K x N) of v_mfma_f32_16x16x16_bf16: lane l holds
// rows k = 4*(l/16) + r (r = 0..3) of column l%16.
__global__ void k_mlp_pack(const float* __restrict__ P, uint16_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= PK_TOTAL) return;
  int off, K, N, nbk, base;
  int frag = t / FRAG, within = t % FRAG;
  if (t < PK_L2) {
    off = OFF_W1; K = MLP_IN; N = MLP_H; nbk = 4; base = PK_L1;
  } else if (t < PK_L3) {
    off = OFF_W2; K = MLP_H; N = MLP_H; nbk = 4; base = PK_L2;
  } else {
    off = OFF_W3; K = MLP_H; N = MLP_OUT; nbk = 1; base = PK_L3;
  }
  frag = (t - base) / FRAG;
  const int kb = frag / nbk, nbi = frag % nbk;
  const int lane = within / 4, r = within % 4;
  const int k = kb * 16 + 4 * (lane / 16) + r, n = nbi * 16 + lane % 16;
  const float w = (k < K && n < N) ? P[off + k * N + n] : 0.0f;
  out[t] = f2bf(w);
}

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// ---- inference on the matrix cores ----------------------------------------------
// Each wavefront evaluates 16 samples per tile: three layers of 16x16x16 BF16
// MFMAs (4 + 16 + 4 per tile), activations relaid through LDS in BF16.
constexpr int INF_WAVES = 4;
__global__ void __launch_bounds__(64 * INF_WAVES) k_mlp_infer(const float* __restrict__ X, int n,
                                                              const uint16_t* __restrict__ pk,
                                                              const float* __restrict__ P, float* __restrict__ Y) {
  __shared__ uint16_t act[INF_WAVES][16 * MLP_H];  // one 16 x 64 bf16 tile per wavefront
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint16_t* A = act[wv];
  const short4_t* frag = (const short4_t*)pk;
  const int row = lane % 16, quad = lane / 16;
  for (int tile = blockIdx.x * INF_WAVES + wv; tile * 16 < n; tile += gridDim.x * INF_WAVES) {
    const int s0 = tile * 16;
    // layer 1: A operand straight from the features (K = 12 padded to 16)
    short4_t a1;
    {
      const int smp = s0 + row;
      for (int r = 0; r < 4; ++r) {
        const int k = 4 * quad + r;
        const float x = (smp < n && k < MLP_IN) ? X[(size_t)smp * MLP_IN + k] : 0.0f;
        a1[r] = (short)f2bf(x);
      }
    }
    for (int nb = 0; nb < 4; ++nb) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, frag[(PK_L1 / 4) + nb * 64 + lane], acc, 0, 0, 0);
      const int col = nb * 16 + row;
      const float bias = P[OFF_B1 + col];
      for (int r = 0; r < 4; ++r) A[(4 * quad + r) * MLP_H + col] = f2bf(silu(acc[r] + bias));
    }
    __syncthreads();
    // layer 2: 64 -> 64
    float4_t h2[4];
    for (int nb = 0; nb < 4; ++nb) {
      float4_t acc = {0.f, 0.f, 0.f, 0.f};
      for (int kb = 0; kb < 4; ++kb) {
        short4_t a;
        for (int r = 0; r < 4; ++r) a[r] = (short)A[row * MLP_H + kb * 16 + 4 * quad + r];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, frag[(PK_L2 / 4) + (kb * 4 + nb) * 64 + lane], acc, 0, 0,
                                                        0);
      }
      h2[nb] = acc;
    }
    __syncthreads();
    for (int nb = 0; nb < 4; ++nb) {
      const int col = nb * 16 + row;
      const float bias = P[OFF_B2 + col];
      for (int r = 0; r < 4; ++r) A[(4 * quad + r) * MLP_H + col] = f2bf(silu(h2[nb][r] + bias));
    }
    __syncthreads();
    // layer 3: 64 -> 3 (16 output columns, 3 used)
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = 0; kb < 4; ++kb) {
      short4_t a;
      for (int r = 0; r < 4; ++r) a[r] = (short)A[row * MLP_H + kb * 16 + 4 * quad + r];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, frag[(PK_L3 / 4) + kb * 64 + lane], acc, 0, 0, 0);
    }
    if (row < MLP_OUT) {
      for (int r = 0; r < 4; ++r) {
        const int smp = s0 + 4 * quad + r;
        if (smp < n) Y[(size_t)smp * MLP_OUT + row] = acc[r] + P[OFF_B3 + row];
      }
    }
    __syncthreads();
  }
}

// Start impulses: manifolds with a previous normal impulse (already in
// l0..l2 from the cache lookup) keep it; the others take the prediction,
// projected onto lambda_n >= 0 and the friction disc.
__global__ void k_mlp_apply(const float* __restrict__ Y, const double* __restrict__ lscale, const int* __restrict__ ma,
                            const SysDev* __restrict__ sys, const int* __restrict__ bsys, int nm, int all,
                            double* __restrict__ l0, double* __restrict__ l1, double* __restrict__ l2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  if (!all && l0[i] > 0.0) return;  // keep the cached start
  const double ls = lscale[i];
  const double ln = dmax(0.0, (double)Y[3 * i] * ls);
  double a = (double)Y[3 * i + 1] * ls, b = (double)Y[3 * i + 2] * ls;
  const double lim = sys[bsys[ma[i]]].mu * ln;
  const double nt = sqrt(a * a + b * b);
  if (nt > lim) {
    const double sc = nt > 0.0 ? lim / nt : 0.0;
    a *= sc;
    b *= sc;
  }
  l0[i] = ln;
  l1[i] = a;
  l2[i] = b;
}

// ---- training (FP32) ------------------------------------------------------------
constexpr int TR_THREADS = 256, TR_CHUNK = 32;
constexpr int PPT = (NPARAM + TR_THREADS - 1) / TR_THREADS;  // parameters per thread

// Each block accumulates the loss gradient of its share of the mini-batch into
// registers (PPT parameters per thread) and writes one partial row.  Samples
// go through in chunks of 32; every phase spreads its (sample, unit) items
// over all 256 threads, with activations and SiLU derivatives held in LDS.
__global__ void __launch_bounds__(TR_THREADS) k_mlp_train_partial(const float* __restrict__ X,
                                                                  const float* __restrict__ T,
                                                                  const int* __restrict__ idx, int nb_samples,
                                                                  float inv_count, const float* __restrict__ P,
                                                                  float* __restrict__ partial,
                                                                  float* __restrict__ loss_part) {
  __shared__ float w[NPARAM];
  __shared__ float sx[TR_CHUNK][MLP_IN], sh1[TR_CHUNK][MLP_H], sd1[TR_CHUNK][MLP_H], sh2[TR_CHUNK][MLP_H],
      sd2[TR_CHUNK][MLP_H];
  __shared__ float sdy[TR_CHUNK][MLP_OUT], se[TR_CHUNK][MLP_OUT];
  const int tid = threadIdx.x;
  for (int k = tid; k < NPARAM; k += TR_THREADS) w[k] = P[k];
  float g[PPT];
  for (int k = 0; k < PPT; ++k) g[k] = 0.0f;
  float lacc = 0.0f;
  constexpr int HI = TR_CHUNK * MLP_H / TR_THREADS;  // (sample, hidden) items per thread
  const int per_block = (nb_samples + gridDim.x - 1) / gridDim.x;
  const int b0 = blockIdx.x * per_block, b1 = min(nb_samples, b0 + per_block);
  for (int c0 = b0; c0 < b1; c0 += TR_CHUNK) {
    const int cn = min(TR_CHUNK, b1 - c0);
    __syncthreads();  // previous chunk's gradient phase is done with the LDS tiles
    for (int it = tid; it < TR_CHUNK * MLP_IN; it += TR_THREADS) {
      const int t = it / MLP_IN, k = it % MLP_IN;
      float x = 0.0f;
      if (t < cn) x = X[(size_t)(idx ? idx[c0 + t] : c0 + t) * MLP_IN + k];
      sx[t][k] = x;
    }
    __syncthreads();
    // layer 1
    for (int r = 0; r < HI; ++r) {
      const int it = tid + r * TR_THREADS, t = it / MLP_H, j = it % MLP_H;
      float z = w[OFF_B1 + j];
      for (int k = 0; k < MLP_IN; ++k) z += sx[t][k] * w[OFF_W1 + k * MLP_H + j];
      const float sg = 1.0f / (1.0f + __expf(-z));
      const bool live = t < cn;
      sh1[t][j] = live ? z * sg : 0.0f;
      sd1[t][j] = live ? sg * (1.0f + z * (1.0f - sg)) : 0.0f;
    }
    __syncthreads();
    // layer 2
    for (int r = 0; r < HI; ++r) {
      const int it = tid + r * TR_THREADS, t = it / MLP_H, j = it % MLP_H;
      float z = w[OFF_B2 + j];
      for (int k = 0; k < MLP_H; ++k) z += sh1[t][k] * w[OFF_W2 + k * MLP_H + j];
      const float sg = 1.0f / (1.0f + __expf(-z));
      const bool live = t < cn;
      sh2[t][j] = live ? z * sg : 0.0f;
      sd2[t][j] = live ? sg * (1.0f + z * (1.0f - sg)) : 0.0f;
    }
    __syncthreads();
    // output layer and loss
    if (tid < TR_CHUNK * MLP_OUT) {
      const int t = tid / MLP_OUT, o = tid % MLP_OUT;
      float y = w[OFF_B3 + o];
      for (int k = 0; k < MLP_H; ++k) y += sh2[t][k] * w[OFF_W3 + k * MLP_OUT + o];
      float e = 0.0f;
      if (t < cn) e = y - T[(size_t)(idx ? idx[c0 + t] : c0 + t) * MLP_OUT + o];
      se[t][o] = e * e;
      sdy[t][o] = 2.0f * e * inv_count;
    }
    __syncthreads();
    // back through layer 3 (sd2 becomes dL/dz2)
    for (int r = 0; r < HI; ++r) {
      const int it = tid + r * TR_THREADS, t = it / MLP_H, j = it % MLP_H;
      float acc = 0.0f;
      for (int o = 0; o < MLP_OUT; ++o) acc += sdy[t][o] * w[OFF_W3 + j * MLP_OUT + o];
      sd2[t][j] *= acc;
    }
    __syncthreads();
    // back through layer 2 (sd1 becomes dL/dz1)
    for (int r = 0; r < HI; ++r) {
      const int it = tid + r * TR_THREADS, t = it / MLP_H, j = it % MLP_H;
      float acc = 0.0f;
      for (int o = 0; o < MLP_H; ++o) acc += sd2[t][o] * w[OFF_W2 + j * MLP_H + o];
      sd1[t][j] *= acc;
    }
    __syncthreads();
    // parameter gradients
    for (int q = 0; q < PPT; ++q) {
      const int k = tid + q * TR_THREADS;
      if (k >= NPARAM) break;
      float acc = 0.0f;
      if (k < OFF_B1) {
        const int r = k / MLP_H, c = k % MLP_H;
        for (int t = 0; t < TR_CHUNK; ++t) acc += sx[t][r] * sd1[t][c];
      } else if (k < OFF_W2) {
        const int c = k - OFF_B1;
        for (int t = 0; t < TR_CHUNK; ++t) acc += sd1[t][c];
      } else if (k < OFF_B2) {
        const int r = (k - OFF_W2) / MLP_H, c = (k - OFF_W2) % MLP_H;
        for (int t = 0; t < TR_CHUNK; ++t) acc += sh1[t][r] * sd2[t][c];
      } else if (k < OFF_W3) {
        const int c = k - OFF_B2;
        for (int t = 0; t < TR_CHUNK; ++t) acc += sd2[t][c];
      } else if (k < OFF_B3) {
        const int r = (k - OFF_W3) / MLP_OUT, c = (k - OFF_W3) % MLP_OUT;
        for (int t = 0; t < TR_CHUNK; ++t) acc += sh2[t][r] * sdy[t][c];
      } else {
        const int c = k - OFF_B3;
        for (int t = 0; t < TR_CHUNK; ++t) acc += sdy[t][c];
      }
      g[q] += acc;
    }
    if (tid == 0)
      for (int t = 0; t < TR_CHUNK; ++t) lacc += se[t][0] + se[t][1] + se[t][2];
  }
  for (int q = 0; q < PPT; ++q) {
    const int k = tid + q * TR_THREADS;
    if (k < NPARAM) partial[(size_t)blockIdx.x * NPARAM + k] = g[q];
  }
  if (tid == 0) loss_part[blockIdx.x] = lacc;
}

// sum the per-block partials; optionally take an AdamW step
__global__ void k_mlp_reduce_adam(const float* __restrict__ partial, int nblk, float* __restrict__ P,
                                  float* __restrict__ m, float* __restrict__ v, float* __restrict__ grad_out, int adam,
                                  float lr, float wd, float b1, float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NPARAM) return;
  float gsum = 0.0f;
  for (int b = 0; b < nblk; ++b) gsum += partial[(size_t)b * NPARAM + k];
  if (grad_out) grad_out[k] = gsum;
  if (!adam) return;
  const float mk = b1 * m[k] + (1.0f - b1) * gsum;
  const float vk = b2 * v[k] + (1.0f - b2) * gsum * gsum;
  m[k] = mk;
  v[k] = vk;
  const float mh = mk / c1, vh = vk / c2;
  const bool is_weight = (k < OFF_B1) || (k >= OFF_W2 && k < OFF_B2) || (k >= OFF_W3 && k < OFF_B3);
  float p = P[k];
  if (is_weight) p -= lr * wd * p;
  P[k] = p - lr * mh / (sqrtf(vh) + eps);
}

// ---- host side ------------------------------------------------------------------
constexpr int TR_BLOCKS = 256;

static int mlp_alloc(rdk_ctx* ctx, Mlp* M) {
  RDK_TRY(ensure(ctx, M->params, NPARAM));
  RDK_TRY(ensure(ctx, M->adam_m, NPARAM));
  RDK_TRY(ensure(ctx, M->adam_v, NPARAM));
  RDK_TRY(ensure(ctx, M->packed, PK_TOTAL));
  RDK_TRY(ensure(ctx, M->partial, (size_t)TR_BLOCKS * NPARAM));
  RDK_TRY(ensure(ctx, M->loss_part, TR_BLOCKS));
  return RDK_OK;
}

static int mlp_pack(rdk_ctx* ctx, Mlp* M) {
  ProfScope ps(ctx, K_MLP);
  hipLaunchKernelGGL(k_mlp_pack, dim3(grid_for(PK_TOTAL, 256)), dim3(256), 0, ctx->stream, M->params.p, M->packed.p);
  return RDK_OK;
}

static int mlp_infer_dev(rdk_ctx* ctx, Mlp* M, const float* X, int n, float* Y) {
  if (n <= 0) return RDK_OK;
  const int tiles = (n + 15) / 16;
  const int blocks = std::min(4096, (tiles + INF_WAVES - 1) / INF_WAVES);
  ProfScope ps(ctx, K_MLP);
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * INF_WAVES), 0, ctx->stream, X, n, M->packed.p,
                     M->params.p, Y);
  return RDK_OK;
}

// features of the manifolds just built (natural order), into M->feat / M->lscale;
// use_prev = false leaves the previous-step impulse feature at 0
static int mlp_features(rdk_ctx* ctx, Mlp* M, bool use_prev = true) {
  const int nm = ctx->nm;
  RDK_TRY(ensure(ctx, M->feat, (size_t)std::max(nm, 1) * MLP_IN));
  RDK_TRY(ensure(ctx, M->lscale, std::max(nm, 1)));
  if (nm == 0) return RDK_OK;
  FeatIn F;
  F.f = ctx->mf.p;
  F.stride = ctx->mf.cap / RDK_MF_FIELDS;
  F.a = ctx->ma.p;
  F.b = ctx->mb.p;
  F.prev_keys = ctx->prev_keys.p;
  F.prev_ln = ctx->prev_lam[0].p;
  F.nprev = use_prev ? ctx->nprev : 0;
  F.prev_wkeys = ctx->prev_wkeys.p;
  F.prev_wln = ctx->prev_wlam[0].p;
  F.nprev_w = use_prev ? ctx->nprev_w : 0;
  const double g = std::sqrt(ctx->phys.gravity[0] * ctx->phys.gravity[0] + ctx->phys.gravity[1] * ctx->phys.gravity[1] +
                             ctx->phys.gravity[2] * ctx->phys.gravity[2]);
  double vs = g * ctx->phys.dt;
  if (!(vs > 0.0)) vs = 1e-3;
  ProfScope ps(ctx, K_MLP);
  hipLaunchKernelGGL(k_mlp_features, dim3(grid_for(nm, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx),
                     ctx->sys_dev.p, F, nm, vs, M->feat.p, M->lscale.p);
  return RDK_OK;
}

// warm_mode 2: prediction for the manifolds without a cached normal impulse;
// 3: for every manifold, the cached impulse an input; 4: for every manifold
// from the contact alone (the impulse input 0, the cache not used)
int mlp_warm_start(rdk_ctx* ctx, int mode) {
  const bool all = mode >= 3;
  Mlp* M = get_mlp(ctx);
  if (!M->ready) {
    ctx->err = "learned warm start: no network loaded (rdk_mlp_init / rdk_mlp_set_params / rdk_mlp_train)";
    return RDK_ERR_STATE;
  }
  const int nm = ctx->nm;
  if (nm == 0) return RDK_OK;
  RDK_TRY(mlp_features(ctx, M, mode != 4));
  RDK_TRY(ensure(ctx, M->out, (size_t)nm * MLP_OUT));
  RDK_TRY(mlp_infer_dev(ctx, M, M->feat.p, nm, M->out.p));
  ProfScope ps(ctx, K_MLP);
  hipLaunchKernelGGL(k_mlp_apply, dim3(grid_for(nm, 256)), dim3(256), 0, ctx->stream, M->out.p, M->lscale.p,
                     ctx->ma.p, ctx->sys_dev.p, ctx->sys_of.p, nm, all ? 1 : 0, ctx->lam[0].p, ctx->lam[1].p,
                     ctx->lam[2].p);
  return RDK_OK;
}

void mlp_free(rdk_ctx* ctx) {
  Mlp* M = (Mlp*)ctx->mlp;
  if (!M) return;
  release(M->params);
  release(M->adam_m);
  release(M->adam_v);
  release(M->partial);
  release(M->loss_part);
  release(M->packed);
  release(M->feat);
  release(M->out);
  release(M->dx);
  release(M->dy);
  release(M->lscale);
  delete M;
  ctx->mlp = nullptr;
}

static int mlp_set_params_host(rdk_ctx* ctx, Mlp* M, const float* p) {
  RDK_TRY(mlp_alloc(ctx, M));
  RDK_HIP(hipMemcpyAsync(M->params.p, p, NPARAM * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RDK_HIP(hipMemsetAsync(M->adam_m.p, 0, NPARAM * sizeof(float), ctx->stream));
  RDK_HIP(hipMemsetAsync(M->adam_v.p, 0, NPARAM * sizeof(float), ctx->stream));
  M->adam_t = 0;
  RDK_TRY(mlp_pack(ctx, M));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  M->ready = true;
  return RDK_OK;
}

// one pass over (idx or 0..n) samples; adam != 0 takes an AdamW step
static int mlp_grad_step(rdk_ctx* ctx, Mlp* M, const float* X, const float* T, const int* idx, int count, int adam,
                         float lr, float wd, float* loss) {
  const int nblk = std::max(1, std::min(TR_BLOCKS, (count + TR_CHUNK - 1) / TR_CHUNK));
  {
    ProfScope ps(ctx, K_MLP);
    hipLaunchKernelGGL(k_mlp_train_partial, dim3(nblk), dim3(TR_THREADS), 0, ctx->stream, X, T, idx, count,
                       1.0f / (float)(count * MLP_OUT), M->params.p, M->partial.p, M->loss_part.p);
  }
  float c1 = 1.0f, c2 = 1.0f;
  if (adam) {
    M->adam_t += 1;
    c1 = 1.0f - std::pow(0.9f, (float)M->adam_t);
    c2 = 1.0f - std::pow(0.999f, (float)M->adam_t);
  }
  {
    ProfScope ps(ctx, K_MLP);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3(grid_for(NPARAM, 256)), dim3(256), 0, ctx->stream, M->partial.p, nblk,
                       M->params.p, M->adam_m.p, M->adam_v.p, adam ? nullptr : M->out.p, adam, lr, wd, 0.9f, 0.999f,
                       1e-8f, c1, c2);
  }
  if (loss) {
    std::vector<float> lp(nblk);
    RDK_HIP(hipMemcpyAsync(lp.data(), M->loss_part.p, nblk * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = (float)(s / ((double)count * MLP_OUT));
  }
  return RDK_OK;
}

}  // namespace rdk

using namespace rdk;

extern "C" {

int rdk_mlp_num_params(void) { return NPARAM; }

int rdk_mlp_init(rdk_ctx* ctx, uint64_t seed) {
  if (!ctx) return RDK_ERR_ARG;
  RDK_HIP(hipSetDevice(ctx->device));
  std::mt19937_64 rng(seed);
  std::vector<float> p(NPARAM, 0.0f);
  auto fill = [&](int off, int fan_in, int fan_out) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> U(-a, a);
    for (int k = 0; k < fan_in * fan_out; ++k) p[off + k] = (float)U(rng);
  };
  fill(OFF_W1, MLP_IN, MLP_H);
  fill(OFF_W2, MLP_H, MLP_H);
  fill(OFF_W3, MLP_H, MLP_OUT);
  return mlp_set_params_host(ctx, get_mlp(ctx), p.data());
}

int rdk_mlp_set_params(rdk_ctx* ctx, const float* p, int n) {
  if (!ctx || !p || n != NPARAM) return RDK_ERR_ARG;
  RDK_HIP(hipSetDevice(ctx->device));
  return mlp_set_params_host(ctx, get_mlp(ctx), p);
}

int rdk_mlp_get_params(rdk_ctx* ctx, float* p, int n) {
  if (!ctx || !p || n != NPARAM) return RDK_ERR_ARG;
  Mlp* M = get_mlp(ctx);
  if (!M->ready) return RDK_ERR_STATE;
  RDK_HIP(hipMemcpyAsync(p, M->params.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_mlp_infer(rdk_ctx* ctx, const float* X, int n, float* Y) {
  if (!ctx || n < 0 || (n > 0 && (!X || !Y))) return RDK_ERR_ARG;
  RDK_HIP(hipSetDevice(ctx->device));
  Mlp* M = get_mlp(ctx);
  if (!M->ready) return RDK_ERR_STATE;
  if (n == 0) return RDK_OK;
  RDK_TRY(ensure(ctx, M->dx, (size_t)n * MLP_IN));
  RDK_TRY(ensure(ctx, M->dy, (size_t)n * MLP_OUT));
  RDK_HIP(hipMemcpyAsync(M->dx.p, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RDK_TRY(mlp_infer_dev(ctx, M, M->dx.p, n, M->dy.p));
  RDK_HIP(hipMemcpyAsync(Y, M->dy.p, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return prof_flush(ctx);
}

int rdk_mlp_features(rdk_ctx* ctx, int cap, float* X, double* lscale) {
  if (!ctx || !X) return RDK_ERR_ARG;
  RDK_HIP(hipSetDevice(ctx->device));
  if (!ctx->have_manifolds) return RDK_ERR_STATE;
  Mlp* M = get_mlp(ctx);
  RDK_TRY(mlp_features(ctx, M));
  const int n = std::min(cap, ctx->nm);
  if (n > 0) {
    RDK_HIP(hipMemcpyAsync(X, M->feat.p, (size_t)n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (lscale)
      RDK_HIP(hipMemcpyAsync(lscale, M->lscale.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return prof_flush(ctx);
}

int rdk_mlp_grad(rdk_ctx* ctx, const float* X, const float* T, int n, float* grad, float* loss) {
  if (!ctx || !X || !T || !grad || n <= 0) return RDK_ERR_ARG;
  RDK_HIP(hipSetDevice(ctx->device));
  Mlp* M = get_mlp(ctx);
  if (!M->ready) return RDK_ERR_STATE;
  RDK_TRY(ensure(ctx, M->dx, (size_t)n * MLP_IN));
  RDK_TRY(ensure(ctx, M->dy, (size_t)n * MLP_OUT));
  RDK_TRY(ensure(ctx, M->out, std::max<size_t>(NPARAM, M->out.cap)));
  RDK_HIP(hipMemcpyAsync(M->dx.p, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RDK_HIP(hipMemcpyAsync(M->dy.p, T, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RDK_TRY(mlp_grad_step(ctx, M, M->dx.p, M->dy.p, nullptr, n, 0, 0.f, 0.f, loss));
  RDK_HIP(hipMemcpyAsync(grad, M->out.p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return prof_flush(ctx);
}

int rdk_mlp_train(rdk_ctx* ctx, const float* X, const float* T, int n, int epochs, int batch, float lr, float wd,
                  uint64_t seed, float* final_loss) {
  if (!ctx || !X || !T || n <= 0 || epochs < 0 || batch <= 0) return RDK_ERR_ARG;
  RDK_HIP(hipSetDevice(ctx->device));
  Mlp* M = get_mlp(ctx);
  if (!M->ready) return RDK_ERR_STATE;
  RDK_TRY(ensure(ctx, M->dx, (size_t)n * MLP_IN));
  RDK_TRY(ensure(ctx, M->dy, (size_t)n * MLP_OUT));
  RDK_TRY(ensure(ctx, M->out, std::max<size_t>(NPARAM, M->out.cap)));
  DevBuf<int> didx;
  RDK_TRY(ensure(ctx, didx, n));
  RDK_HIP(hipMemcpyAsync(M->dx.p, X, (size_t)n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RDK_HIP(hipMemcpyAsync(M->dy.p, T, (size_t)n * MLP_OUT * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::mt19937_64 rng(seed);
  float last = 0.0f;
  for (int e = 0; e < epochs; ++e) {
    std::shuffle(order.begin(), order.end(), rng);
    RDK_HIP(hipMemcpyAsync(didx.p, order.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    for (int s = 0; s < n; s += batch) {
      const int cnt = std::min(batch, n - s);
      const bool last_step = (e == epochs - 1) && (s + batch >= n);
      RDK_TRY(mlp_grad_step(ctx, M, M->dx.p, M->dy.p, didx.p + s, cnt, 1, lr, wd, last_step ? &last : nullptr));
    }
    RDK_HIP(hipStreamSynchronize(ctx->stream));  // the host order buffer is reused next epoch
  }
  RDK_TRY(mlp_pack(ctx, M));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  release(didx);
  if (final_loss) *final_loss = last;
  return prof_flush(ctx);
}

}  // extern "C"

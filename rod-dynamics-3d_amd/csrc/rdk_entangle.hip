// rdk_entangle.hip -- entanglement observer: per system, the sum of
// |Gauss linking number| over rod pairs and the pair count (replaces
// App::computeEntanglement, main.cpp:4287-4331, and
// pairwise_abs_linking_sum_with_cutoff, entanglement_adapter.cpp:56-117).
//
// The linking number of two straight segments is the signed solid angle of
// the tetrahedron they span over 4 pi (Klenin & Langowski 2000):
//   Omega = asin(n1.n2) + asin(n2.n3) + asin(n3.n4) + asin(n4.n1),
//   n1 = r13 x r14, n2 = r14 x r24, n3 = r24 x r23, n4 = r23 x r13,
// each face normal normalised as n / (|n| + 1e-6), signed by (r34 x r12).r13.
// The 1e-6 is the reference routine's convention (its dependency
// entanglement-cpp is not vendored): with it the formula reproduces the
// reference's published ebar0 of every packing in assets/packings_metadata.csv.
//
// Layout: the rods' axis segments are built once per call in FP64 SoA
// (x -/+ u h; on periodic systems each endpoint is wrapped into the box on
// its own, as the reference does).  Pairs are enumerated by tiles of 64 rods:
// one workgroup per (i-tile, j-tile) pair of the upper triangle, both tiles
// staged in LDS, 16 pairs per lane; a pair counts when the cutoff is off or
// its segments are at most cutoff apart (Lumelsky closest points, after a
// centre-distance lower bound).  Each workgroup writes one partial sum and
// count; a single workgroup adds the partials in a fixed order, so the result
// is deterministic.  The observer is arithmetic-bound (4 asin, 4 divides and
// a sqrt per pair in FP64); nothing here is on the time-stepping path.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "rdk_internal.hpp"

namespace rdk {

constexpr int ENT_TILE = 64;
constexpr int ENT_THREADS = 256;

struct Segs {
  double* a[3];
  double* e[3];
};

__global__ void k_ent_segments(BodyPtrs B, const SysDev* __restrict__ sys, int n, Segs S) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d3 x = mk3(B.x[0][i], B.x[1][i], B.x[2][i]);
  const d3 u = axis_of(q4{B.q[0][i], B.q[1][i], B.q[2][i], B.q[3][i]});
  const double h = B.h[i];
  d3 a = x - u * h, e = x + u * h;
  const SysDev& Y = sys[B.sys[i]];
  if (Y.periodic[0] || Y.periodic[1] || Y.periodic[2]) {
    double* pa[3] = {&a.x, &a.y, &a.z};
    double* pe[3] = {&e.x, &e.y, &e.z};
    for (int k = 0; k < 3; ++k) {
      if (!Y.periodic[k]) continue;
      const double span = Y.hi[k] - Y.lo[k];
      *pa[k] -= span * floor((*pa[k] - Y.lo[k]) / span);
      *pe[k] -= span * floor((*pe[k] - Y.lo[k]) / span);
    }
  }
  S.a[0][i] = a.x;
  S.a[1][i] = a.y;
  S.a[2][i] = a.z;
  S.e[0][i] = e.x;
  S.e[1][i] = e.y;
  S.e[2][i] = e.z;
}

__device__ __forceinline__ d3 ent_unit(d3 v) {
  const double n = sqrt(dot(v, v)) + 1e-6;
  return mk3(v.x / n, v.y / n, v.z / n);
}

__device__ __forceinline__ double ent_asin(double x) { return asin(x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x)); }

__device__ __forceinline__ double linking_number(d3 p1, d3 p2, d3 p3, d3 p4) {
  const d3 r13 = p3 - p1, r14 = p4 - p1, r23 = p3 - p2, r24 = p4 - p2;
  const d3 r12 = p2 - p1, r34 = p4 - p3;
  const d3 n1 = ent_unit(cross(r13, r14)), n2 = ent_unit(cross(r14, r24));
  const d3 n3 = ent_unit(cross(r24, r23)), n4 = ent_unit(cross(r23, r13));
  const double om = ent_asin(dot(n1, n2)) + ent_asin(dot(n2, n3)) + ent_asin(dot(n3, n4)) + ent_asin(dot(n4, n1));
  const double tp = dot(cross(r34, r12), r13);
  const double sg = tp > 0.0 ? 1.0 : (tp < 0.0 ? -1.0 : 0.0);
  return om * sg / (4.0 * M_PI);
}

// Tile pair of workgroup b in the upper triangle of T x T tiles (ti <= tj),
// row-major: row ti holds T - ti tile pairs.
__device__ __forceinline__ void tile_pair(long long b, int T, int& ti, int& tj) {
  const double tt = 2.0 * T + 1.0;
  int r = (int)floor((tt - sqrt(tt * tt - 8.0 * (double)b)) * 0.5);
  if (r < 0) r = 0;
  if (r > T - 1) r = T - 1;
  auto start = [T](long long row) { return row * T - row * (row - 1) / 2; };
  while (r > 0 && start(r) > b) --r;
  while (r < T - 1 && start(r + 1) <= b) ++r;
  ti = r;
  tj = r + (int)(b - start(r));
}

__global__ void __launch_bounds__(ENT_THREADS) k_ent_pairs(Segs S, int i0, int n, int T, double cutoff,
                                                           double* __restrict__ psum,
                                                           unsigned long long* __restrict__ pcnt) {
  __shared__ double sa[2][6][ENT_TILE];
  __shared__ double red[ENT_THREADS / 64];
  __shared__ unsigned long long redc[ENT_THREADS / 64];
  int ti, tj;
  tile_pair(blockIdx.x, T, ti, tj);
  const int bi = i0 + ti * ENT_TILE, bj = i0 + tj * ENT_TILE;
  for (int k = threadIdx.x; k < 2 * ENT_TILE; k += ENT_THREADS) {
    const int t = k / ENT_TILE, l = k % ENT_TILE;
    const int g = (t == 0 ? bi : bj) + l;
    const bool ok = g < i0 + n;
    for (int c = 0; c < 3; ++c) {
      sa[t][c][l] = ok ? S.a[c][g] : 0.0;
      sa[t][3 + c][l] = ok ? S.e[c][g] : 0.0;
    }
  }
  __syncthreads();
  double sum = 0.0;
  unsigned long long cnt = 0;
  for (int k = threadIdx.x; k < ENT_TILE * ENT_TILE; k += ENT_THREADS) {
    const int il = k / ENT_TILE, jl = k % ENT_TILE;
    if (ti == tj && jl <= il) continue;
    if (bi + il >= i0 + n || bj + jl >= i0 + n) continue;
    const d3 a0 = mk3(sa[0][0][il], sa[0][1][il], sa[0][2][il]);
    const d3 a1 = mk3(sa[0][3][il], sa[0][4][il], sa[0][5][il]);
    const d3 b0 = mk3(sa[1][0][jl], sa[1][1][jl], sa[1][2][jl]);
    const d3 b1 = mk3(sa[1][3][jl], sa[1][4][jl], sa[1][5][jl]);
    if (cutoff > 0.0) {
      // lower bound: centre distance minus the two half lengths
      const d3 dc = (a0 + a1) * 0.5 - (b0 + b1) * 0.5;
      const double hb = 0.5 * (len(a1 - a0) + len(b1 - b0));
      if (len(dc) - hb > cutoff + 1e-9 * (1.0 + len(dc) + hb + cutoff)) continue;
      double s, t;
      segseg(a0, a1, b0, b1, s, t);
      const d3 d = (a0 + (a1 - a0) * s) - (b0 + (b1 - b0) * t);
      if (sqrt(dot(d, d)) > cutoff) continue;
    }
    sum += fabs(linking_number(a0, a1, b0, b1));
    ++cnt;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_down(sum, off);
    cnt += __shfl_down(cnt, off);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[wid] = sum;
    redc[wid] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    unsigned long long c = 0;
    for (int w = 0; w < ENT_THREADS / 64; ++w) {
      s += red[w];
      c += redc[w];
    }
    psum[blockIdx.x] = s;
    pcnt[blockIdx.x] = c;
  }
}

// One workgroup adds the partials: lane k sums partials k, k + 1024, ... in
// order, then a fixed tree.
__global__ void __launch_bounds__(1024) k_ent_reduce(const double* __restrict__ psum,
                                                     const unsigned long long* __restrict__ pcnt, long long m,
                                                     double* __restrict__ out_sum,
                                                     unsigned long long* __restrict__ out_cnt) {
  __shared__ double rs[1024];
  __shared__ unsigned long long rc[1024];
  double s = 0.0;
  unsigned long long c = 0;
  for (long long k = threadIdx.x; k < m; k += 1024) {
    s += psum[k];
    c += pcnt[k];
  }
  rs[threadIdx.x] = s;
  rc[threadIdx.x] = c;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rs[threadIdx.x] += rs[threadIdx.x + w];
      rc[threadIdx.x] += rc[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_sum = rs[0];
    *out_cnt = rc[0];
  }
}

int entanglement_run(rdk_ctx* ctx, double cutoff, double* sums, int64_t* pairs) {
  const int n = ctx->n;
  const int nsys = ctx->nsys;
  for (int k = 0; k < 3; ++k) {
    RDK_TRY(ensure(ctx, ctx->ent_a[k], std::max(n, 1)));
    RDK_TRY(ensure(ctx, ctx->ent_e[k], std::max(n, 1)));
  }
  Segs S;
  for (int k = 0; k < 3; ++k) {
    S.a[k] = ctx->ent_a[k].p;
    S.e[k] = ctx->ent_e[k].p;
  }
  {
    ProfScope ps(ctx, K_OBSERVE);
    hipLaunchKernelGGL(k_ent_segments, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx),
                       ctx->sys_dev.p, n, S);
  }
  RDK_TRY(ensure(ctx, ctx->ent_out, 2 * (size_t)std::max(nsys, 1)));
  for (int s = 0; s < nsys; ++s) {
    const int i0 = ctx->sys_off[s], ns = ctx->sys_off[s + 1] - i0;
    if (ns < 2) {
      RDK_HIP(hipMemsetAsync(ctx->ent_out.p + 2 * s, 0, 2 * sizeof(double), ctx->stream));
      continue;
    }
    const int T = (ns + ENT_TILE - 1) / ENT_TILE;
    const long long nb = (long long)T * (T + 1) / 2;
    if (nb > 2147483647LL) {
      ctx->err = "rdk_entanglement: too many rods in one system for the pair grid";
      return RDK_ERR_ARG;
    }
    RDK_TRY(ensure(ctx, ctx->ent_psum, (size_t)nb));
    RDK_TRY(ensure(ctx, ctx->ent_pcnt, (size_t)nb));
    {
      ProfScope ps(ctx, K_OBSERVE);
      hipLaunchKernelGGL(k_ent_pairs, dim3((unsigned)nb), dim3(ENT_THREADS), 0, ctx->stream, S, i0, ns, T, cutoff,
                         ctx->ent_psum.p, ctx->ent_pcnt.p);
      hipLaunchKernelGGL(k_ent_reduce, dim3(1), dim3(1024), 0, ctx->stream, ctx->ent_psum.p, ctx->ent_pcnt.p, nb,
                         ctx->ent_out.p + 2 * s, (unsigned long long*)(ctx->ent_out.p + 2 * s + 1));
    }
  }
  std::vector<double> h(2 * (size_t)nsys);
  RDK_HIP(hipMemcpyAsync(h.data(), ctx->ent_out.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < nsys; ++s) {
    if (sums) sums[s] = h[2 * s];
    unsigned long long c;
    std::memcpy(&c, &h[2 * s + 1], sizeof(c));
    if (pairs) pairs[s] = (int64_t)c;
  }
  return RDK_OK;
}

// ---- allocation-free variant ----------------------------------------------
// Scratch of rdk_entanglement_noalloc: the segments (6 doubles per rod) and
// the per-tile partial sums and counts of one system at a time.  One copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// to a kernel: its device address comes from hipGetSymbolAddress, once per
// device, at context creation (noalloc_scratch_init).
constexpr int NOALLOC_TILES = (RDK_NOALLOC_MAX_RODS + ENT_TILE - 1) / ENT_TILE;
constexpr size_t NOALLOC_PAIRS = (size_t)NOALLOC_TILES * (NOALLOC_TILES + 1) / 2;  // tile pairs of one system
constexpr size_t NOALLOC_SEG_BYTES = 6 * sizeof(double) * (size_t)RDK_NOALLOC_MAX_RODS;
constexpr size_t NOALLOC_NEED = NOALLOC_SEG_BYTES + NOALLOC_PAIRS * (sizeof(double) + sizeof(unsigned long long));
static_assert(NOALLOC_NEED <= RDK_NOALLOC_SCRATCH_BYTES, "scratch too small for its own limits");

static __device__ __align__(256) char scratch[RDK_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static char* g_scratch[NOALLOC_MAX_DEVICES] = {};
static std::mutex g_scratch_mu;

// Called by rdk_create with ctx->device current.
int noalloc_scratch_init(rdk_ctx* ctx) {
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return RDK_OK;  // variant unavailable there
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[ctx->device]) return RDK_OK;
  void* ptr = nullptr;
  RDK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch[ctx->device] = (char*)ptr;
  return RDK_OK;
}

__global__ void k_ent_zero(double* __restrict__ sums, unsigned long long* __restrict__ pairs, int s) {
  if (threadIdx.x == 0) {
    sums[s] = 0.0;
    pairs[s] = 0ull;
  }
}

static int entanglement_noalloc_run(rdk_ctx* ctx, double cutoff, double* d_sums, unsigned long long* d_pairs) {
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) base = g_scratch[ctx->device];
  }
  if (!base) {
    ctx->err = "rdk_entanglement_noalloc: no scratch address for this device";
    return RDK_ERR_STATE;
  }
  const int n = ctx->n;
  if (n > RDK_NOALLOC_MAX_RODS) {
    ctx->err = "rdk_entanglement_noalloc: more than RDK_NOALLOC_MAX_RODS rods";
    return RDK_ERR_CAPACITY;
  }
  Segs S;
  double* seg = (double*)base;
  for (int k = 0; k < 3; ++k) {
    S.a[k] = seg + (size_t)k * RDK_NOALLOC_MAX_RODS;
    S.e[k] = seg + (size_t)(3 + k) * RDK_NOALLOC_MAX_RODS;
  }
  double* psum = (double*)(base + NOALLOC_SEG_BYTES);
  unsigned long long* pcnt = (unsigned long long*)(psum + NOALLOC_PAIRS);
  if (n > 0)
    hipLaunchKernelGGL(k_ent_segments, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx),
                       ctx->sys_dev.p, n, S);
  for (int s = 0; s < ctx->nsys; ++s) {
    const int i0 = ctx->sys_off[s], ns = ctx->sys_off[s + 1] - i0;
    if (ns < 2) {
      hipLaunchKernelGGL(k_ent_zero, dim3(1), dim3(64), 0, ctx->stream, d_sums, d_pairs, s);
      continue;
    }
    const int T = (ns + ENT_TILE - 1) / ENT_TILE;
    const long long nb = (long long)T * (T + 1) / 2;  // <= NOALLOC_PAIRS since ns <= n
    hipLaunchKernelGGL(k_ent_pairs, dim3((unsigned)nb), dim3(ENT_THREADS), 0, ctx->stream, S, i0, ns, T, cutoff,
                       psum, pcnt);
    hipLaunchKernelGGL(k_ent_reduce, dim3(1), dim3(1024), 0, ctx->stream, psum, pcnt, nb, d_sums + s, d_pairs + s);
  }
  RDK_HIP(hipGetLastError());
  return RDK_OK;
}

}  // namespace rdk

extern "C" int rdk_entanglement_noalloc(rdk_ctx* ctx, double cutoff, double* d_sums, uint64_t* d_pairs) {
  if (!ctx || !d_sums || !d_pairs) return RDK_ERR_ARG;
  if (hipSetDevice(ctx->device) != hipSuccess) {
    ctx->err = "hipSetDevice failed";
    return RDK_ERR_HIP;
  }
  if (ctx->n <= 0 || ctx->nsys <= 0) return RDK_ERR_STATE;
  return rdk::entanglement_noalloc_run(ctx, cutoff, d_sums, (unsigned long long*)d_pairs);
}

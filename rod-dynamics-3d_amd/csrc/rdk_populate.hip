// rdk_populate.hip -- the reference's non-overlapping populate (App::resetScene
// with populate.mode "nonoverlap", src/app/main.cpp:1857-2140) as a parallel
// random sequential addition that returns the serial result bit for bit.
//
// The serial algorithm draws a fixed stream of candidates (5 numbers each from
// std::mt19937 through std::uniform_real_distribution<float>) and accepts a
// candidate when it collides with no rod accepted before it; a rod that fails
// max_attempts times in a row ends the populate.  Because the stream does not
// depend on which candidates are accepted, acceptance is a property of the
// stream alone: candidate j is accepted iff no accepted candidate k < j
// collides with it.  This file evaluates that rule in windows of W candidates:
//
//   1. the host draws the window's candidates (centre, unit direction) with
//      the same library generator and FP32 formulas as the reference;
//   2. k_pop_old: one lane per candidate tests it against every rod accepted
//      in earlier windows, walking the cells of its AABB in a hashed cell grid
//      (the reference's grid: cell size cs, each rod listed in every cell its
//      AABB covers), with the reference's FP32 segment distance (27 periodic
//      images, its end-point rule); a collision rejects it for good;
//   3. the survivors are listed in a window grid and k_pop_resolve decides
//      them in rounds: a candidate colliding with an accepted lower candidate
//      is rejected, one whose colliding lower candidates are all rejected is
//      accepted (the lexicographically first maximal independent set, which is
//      what the serial loop computes); each round decides at least the lowest
//      undecided candidate, so the rounds end;
//   4. the host walks the decisions in stream order, applies the max_attempts
//      rule and the rod count, and k_pop_commit lists the accepted rods in the
//      main grid.
//
// Only +, -, *, /, sqrt-free comparisons and floor run on the device, all IEEE
// FP32 without contraction, so the device decisions equal the serial loop's.
// The candidates' transcendental functions run on the host with the C
// library, as in the reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "rdk_internal.hpp"

namespace rdk {

namespace {

constexpr int POP_UND = 0, POP_ACC = 1, POP_REJ = 2;
constexpr int POP_THREADS = 256;

struct PopGeom {
  float bmin[3], box[3];
  float halfL, R, diam2, cs;
  int n[3];
  int pbc;
  unsigned hbits;  // log2 of the hash table size
};

struct Cand {
  float cx, cy, cz, ux, uy, uz;
};

// AABB cell range of a rod (main.cpp:1910-1921): ext = |u| halfL + R,
// i0/i1 = floor((c -/+ ext - bmin) / cs) per axis, FP32.
__host__ __device__ inline void pop_range(const PopGeom& G, const Cand& a, int i0[3], int i1[3]) {
  const float c[3] = {a.cx, a.cy, a.cz};
  const float u[3] = {a.ux, a.uy, a.uz};
  for (int k = 0; k < 3; ++k) {
    const float ext = fabsf(u[k]) * G.halfL + G.R;
    const float mn = c[k] - ext, mx = c[k] + ext;
    i0[k] = (int)floorf((mn - G.bmin[k]) / G.cs);
    i1[k] = (int)floorf((mx - G.bmin[k]) / G.cs);
  }
}

__host__ __device__ inline int pop_wrap(int i, int dim, int pbc) {
  if (pbc) {
    int r = i % dim;
    return r < 0 ? r + dim : r;
  }
  return i < 0 ? 0 : (i > dim - 1 ? dim - 1 : i);
}

__host__ __device__ inline int pop_cell(const PopGeom& G, int ix, int iy, int iz) {
  return pop_wrap(ix, G.n[0], G.pbc) + G.n[0] * (pop_wrap(iy, G.n[1], G.pbc) + G.n[1] * pop_wrap(iz, G.n[2], G.pbc));
}

__host__ __device__ inline long long pop_ncells(const int i0[3], const int i1[3]) {
  return (long long)(i1[0] - i0[0] + 1) * (i1[1] - i0[1] + 1) * (i1[2] - i0[2] + 1);
}

__device__ inline unsigned pop_bucket(const PopGeom& G, int cell) {
  return ((unsigned)cell * 2654435761u) >> (32 - G.hbits);
}

struct f3s {
  float x, y, z;
};
__device__ inline f3s fa(f3s a, f3s b) { return f3s{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline f3s fs(f3s a, f3s b) { return f3s{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline f3s fm(f3s a, float s) { return f3s{a.x * s, a.y * s, a.z * s}; }
__device__ inline f3s fm(float s, f3s a) { return f3s{s * a.x, s * a.y, s * a.z}; }
__device__ inline float fdot(f3s a, f3s b) {
  const float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}

// squared segment distance of one image (main.cpp:1955-1991 with pbc_rule,
// main.cpp:2075-2101 without)
__device__ float pop_seg_d2(f3s p0, f3s p1, f3s q0, f3s q1, bool pbc_rule) {
  const float eps = 1e-12f;
  const f3s u = fs(p1, p0), v = fs(q1, q0), w0 = fs(p0, q0);
  const float uu = fdot(u, u), vv = fdot(v, v), uv = fdot(u, v);
  const float wu = fdot(w0, u), wv = fdot(w0, v);
  const float D = uu * vv - uv * uv;
  float s, t;
  if (fabsf(D) < eps) {
    s = 0.0f;
    t = (vv >= eps) ? (-wv / vv) : 0.0f;
  } else {
    s = (uv * wv - vv * wu) / D;
    t = (uu * wv - uv * wu) / D;
  }
  s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
  t = (s * uv + wv) / (vv >= eps ? vv : 1.0f);
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  const float su = (-wu + t * uv) / (uu >= eps ? uu : 1.0f);
  if (!(t > 1e-6f && t < 1.0f - 1e-6f)) {
    if (pbc_rule)
      s = su < 0.0f ? 0.0f : su;
    else
      s = su < 0.0f ? 0.0f : (su > 1.0f ? 1.0f : su);
  }
  const f3s d = fs(fa(w0, fm(s, u)), fm(t, v));
  return fdot(d, d);
}

// does candidate a (as the new rod) collide with rod b?
__device__ bool pop_collides(const PopGeom& G, const Cand& a, const Cand& b) {
  const f3s ca{a.cx, a.cy, a.cz}, ua{a.ux, a.uy, a.uz};
  const f3s cb{b.cx, b.cy, b.cz}, ub{b.ux, b.uy, b.uz};
  const f3s p0 = fs(ca, fm(ua, G.halfL)), p1 = fa(ca, fm(ua, G.halfL));
  const f3s q0 = fs(cb, fm(ub, G.halfL)), q1 = fa(cb, fm(ub, G.halfL));
  if (!G.pbc) return pop_seg_d2(p0, p1, q0, q1, false) < G.diam2;
  float m = 3.402823466e+38f;
  for (int k = -1; k <= 1; ++k)
    for (int j = -1; j <= 1; ++j)
      for (int i = -1; i <= 1; ++i) {
        const f3s sh{(float)i * G.box[0], (float)j * G.box[1], (float)k * G.box[2]};
        const float d2 = pop_seg_d2(p0, p1, fa(q0, sh), fa(q1, sh), true);
        if (d2 < m) m = d2;
      }
  return m < G.diam2;
}

// 2: test each window candidate against the rods accepted before the window
__global__ void __launch_bounds__(POP_THREADS) k_pop_old(PopGeom G, const Cand* __restrict__ cand, int W,
                                                         const Cand* __restrict__ acc, const int* __restrict__ mhead,
                                                         const int* __restrict__ mcell, const int* __restrict__ mrod,
                                                         const int* __restrict__ mnext, int* __restrict__ status) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= W) return;
  const Cand a = cand[j];
  int i0[3], i1[3];
  pop_range(G, a, i0, i1);
  bool hit = false;
  for (int iz = i0[2]; iz <= i1[2] && !hit; ++iz)
    for (int iy = i0[1]; iy <= i1[1] && !hit; ++iy)
      for (int ix = i0[0]; ix <= i1[0] && !hit; ++ix) {
        const int cell = pop_cell(G, ix, iy, iz);
        for (int e = mhead[pop_bucket(G, cell)]; e >= 0 && !hit; e = mnext[e]) {
          if (mcell[e] != cell) continue;
          hit = pop_collides(G, a, acc[mrod[e]]);
        }
      }
  status[j] = hit ? POP_REJ : POP_UND;
}

// list candidate / rod `id` in every cell of its AABB, at slots off, off + 1,
// ..., end - 1 (the host counted them with the same arithmetic; a mismatch
// sets *err instead of writing past end)
__device__ void pop_insert(const PopGeom& G, const Cand& a, int id, int off, int end, int* head, int* pcell,
                           int* pid, int* pnext, int* err) {
  int i0[3], i1[3];
  pop_range(G, a, i0, i1);
  if (pop_ncells(i0, i1) != (long long)(end - off)) {
    atomicOr(err, 1);
    return;
  }
  int slot = off;
  for (int iz = i0[2]; iz <= i1[2]; ++iz)
    for (int iy = i0[1]; iy <= i1[1]; ++iy)
      for (int ix = i0[0]; ix <= i1[0]; ++ix, ++slot) {
        const int cell = pop_cell(G, ix, iy, iz);
        pcell[slot] = cell;
        pid[slot] = id;
        pnext[slot] = atomicExch(&head[pop_bucket(G, cell)], slot);
      }
}

__global__ void k_pop_fill(int* __restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// 3a: list the window's survivors in the window grid
__global__ void __launch_bounds__(POP_THREADS) k_pop_winsert(PopGeom G, const Cand* __restrict__ cand, int W,
                                                             const int* __restrict__ status,
                                                             const int* __restrict__ woff, int* whead, int* wcell,
                                                             int* wcand, int* wnext, int* err) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= W || status[j] != POP_UND) return;
  pop_insert(G, cand[j], j, woff[j], woff[j + 1], whead, wcell, wcand, wnext, err);
}

// 3b: one round of decisions; `undecided` counts what is left
__global__ void __launch_bounds__(POP_THREADS) k_pop_resolve(PopGeom G, const Cand* __restrict__ cand, int W,
                                                             int* status, const int* __restrict__ whead,
                                                             const int* __restrict__ wcell,
                                                             const int* __restrict__ wcand,
                                                             const int* __restrict__ wnext, int* undecided) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= W) return;
  if (__hip_atomic_load(&status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != POP_UND) return;
  const Cand a = cand[j];
  int i0[3], i1[3];
  pop_range(G, a, i0, i1);
  bool rej = false, blocked = false;
  for (int iz = i0[2]; iz <= i1[2] && !rej; ++iz)
    for (int iy = i0[1]; iy <= i1[1] && !rej; ++iy)
      for (int ix = i0[0]; ix <= i1[0] && !rej; ++ix) {
        const int cell = pop_cell(G, ix, iy, iz);
        for (int e = whead[pop_bucket(G, cell)]; e >= 0 && !rej; e = wnext[e]) {
          const int k = wcand[e];
          if (k >= j || wcell[e] != cell) continue;
          const int sk = __hip_atomic_load(&status[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (sk == POP_REJ) continue;
          if (!pop_collides(G, a, cand[k])) continue;
          if (sk == POP_ACC)
            rej = true;
          else
            blocked = true;
        }
      }
  if (rej) {
    __hip_atomic_store(&status[j], POP_REJ, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (!blocked) {
    __hip_atomic_store(&status[j], POP_ACC, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    atomicAdd(undecided, 1);
  }
}

// 4: accepted candidates become rods base, base + 1, ... and enter the main grid
__global__ void __launch_bounds__(POP_THREADS) k_pop_commit(PopGeom G, const Cand* __restrict__ cand,
                                                            const int* __restrict__ list, int na, int base,
                                                            const int* __restrict__ moff, Cand* acc, int* mhead,
                                                            int* mcell, int* mrod, int* mnext, int* err) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= na) return;
  const Cand c = cand[list[a]];
  acc[base + a] = c;
  pop_insert(G, c, base + a, moff[a], moff[a + 1], mhead, mcell, mrod, mnext, err);
}

template <typename T>
struct Scoped {
  DevBuf<T> b;
  ~Scoped() { release(b); }
};

unsigned log2_ceil(long long v) {
  unsigned k = 0;
  while ((1LL << k) < v) ++k;
  return k;
}

}  // namespace

int populate_nonoverlap_run(rdk_ctx* ctx, int n, float length, float diameter, const float* bmin, const float* bmax,
                            int periodic, float cell_size, uint32_t seed, int max_attempts, int window,
                            float* centres, float* dirs, float* quats, int* placed, int64_t* attempts) {
  PopGeom G{};
  for (int k = 0; k < 3; ++k) {
    G.bmin[k] = bmin[k];
    G.box[k] = bmax[k] - bmin[k];
    if (!(G.box[k] > 0.0f)) {
      ctx->err = "rdk_populate_nonoverlap: box must have positive extent";
      return RDK_ERR_ARG;
    }
  }
  G.halfL = 0.5f * length;
  G.R = 0.5f * diameter;
  G.diam2 = (2.0f * G.R) * (2.0f * G.R);
  G.cs = cell_size > 0.0f ? cell_size : std::max(0.25f * length, 2.5f * G.R);
  G.pbc = periodic ? 1 : 0;
  long long ncell = 1;
  for (int k = 0; k < 3; ++k) {
    const float d = std::floor(G.box[k] / G.cs);
    if (!(d < 2147483647.0f)) {
      ctx->err = "rdk_populate_nonoverlap: too many grid cells";
      return RDK_ERR_ARG;
    }
    G.n[k] = std::max(1, (int)d);
    ncell *= G.n[k];
  }
  if (ncell > 2147483647LL) {
    ctx->err = "rdk_populate_nonoverlap: too many grid cells";
    return RDK_ERR_ARG;
  }
  // cells one rod can cover, per axis at most floor((L + 2R) / cs) + 2
  long long per_rod = 1;
  for (int k = 0; k < 3; ++k) per_rod *= (long long)std::floor((length + diameter) / G.cs) + 2;
  int W = window > 0 ? window : std::min(65536, std::max(1024, 2 * n + 256));
  W = (int)std::max(1LL, std::min<long long>(W, (1LL << 27) / std::max(1LL, per_rod)));

  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> urand(0.0f, 1.0f);
  const float two_pi = 2.0f * float(M_PI);

  Scoped<Cand> d_cand, d_acc;
  Scoped<int> d_status, d_woff, d_whead, d_wcell, d_wcand, d_wnext, d_mhead, d_mcell, d_mrod, d_mnext, d_list,
      d_moff, d_und;
  std::vector<Cand> hc(W), hacc;
  hacc.reserve(n);
  std::vector<int> hstatus(W), hwoff(W + 1), list, moff;
  RDK_TRY(ensure(ctx, d_cand.b, (size_t)W));
  RDK_TRY(ensure(ctx, d_status.b, (size_t)W));
  RDK_TRY(ensure(ctx, d_woff.b, (size_t)W + 1));
  RDK_TRY(ensure(ctx, d_acc.b, (size_t)std::max(n, 1)));
  RDK_TRY(ensure(ctx, d_und.b, 2));  // [0] undecided count, [1] slot-count mismatch flag
  // main grid: hash buckets for about two entries each at the end
  const long long est_main = std::max(1024LL, 2LL * std::max(n, 1) * std::min(per_rod, 64LL));
  PopGeom GM = G, GW = G;
  GM.hbits = std::min(26u, log2_ceil(est_main));
  GW.hbits = std::min(24u, log2_ceil(std::max(1024LL, 2LL * W * std::min(per_rod, 64LL))));
  const int HM = 1 << GM.hbits, HW = 1 << GW.hbits;
  RDK_TRY(ensure(ctx, d_mhead.b, (size_t)HM));
  RDK_TRY(ensure(ctx, d_whead.b, (size_t)HW));
  hipLaunchKernelGGL(k_pop_fill, dim3(grid_for(HM, 256)), dim3(256), 0, ctx->stream, d_mhead.b.p, HM, -1);
  hipLaunchKernelGGL(k_pop_fill, dim3(1), dim3(64), 0, ctx->stream, d_und.b.p, 2, 0);
  long long mused = 0;
  int done = 0, fails = 0;
  long long drawn = 0;
  bool stop = false;
  while (!stop && done < n) {
    // 1. the window's candidates, drawn as the reference draws them
    long long wtotal = 0;
    for (int j = 0; j < W; ++j) {
      Cand& c = hc[j];
      const float r1 = urand(gen), r2 = urand(gen), r3 = urand(gen);
      c.cx = G.bmin[0] + r1 * G.box[0];
      c.cy = G.bmin[1] + r2 * G.box[1];
      c.cz = G.bmin[2] + r3 * G.box[2];
      const float uy = 2.0f * urand(gen) - 1.0f;
      const float phi = two_pi * urand(gen);
      const float s = std::sqrt(std::max(0.0f, 1.0f - uy * uy));
      const float rx = s * std::cos(phi), rz = s * std::sin(phi);
      const float tx = rx * rx, ty = uy * uy, tz = rz * rz;
      const float inv = 1.0f / std::sqrt((tx + ty) + tz);
      c.ux = rx * inv;
      c.uy = uy * inv;
      c.uz = rz * inv;
      int i0[3], i1[3];
      pop_range(G, c, i0, i1);
      hwoff[j] = (int)wtotal;
      wtotal += pop_ncells(i0, i1);
    }
    hwoff[W] = (int)std::min(wtotal, 2147483647LL);
    if (wtotal > 2147483647LL) {
      ctx->err = "rdk_populate_nonoverlap: window grid too large (cell size too small for the rods)";
      return RDK_ERR_ARG;
    }
    RDK_TRY(ensure(ctx, d_wcell.b, (size_t)std::max(wtotal, 1LL)));
    RDK_TRY(ensure(ctx, d_wcand.b, (size_t)std::max(wtotal, 1LL)));
    RDK_TRY(ensure(ctx, d_wnext.b, (size_t)std::max(wtotal, 1LL)));
    RDK_HIP(hipMemcpyAsync(d_cand.b.p, hc.data(), W * sizeof(Cand), hipMemcpyHostToDevice, ctx->stream));
    RDK_HIP(hipMemcpyAsync(d_woff.b.p, hwoff.data(), (W + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    // 2. against the rods of earlier windows
    {
      ProfScope ps(ctx, K_POPULATE);
      hipLaunchKernelGGL(k_pop_old, dim3(grid_for(W, POP_THREADS)), dim3(POP_THREADS), 0, ctx->stream, GM,
                       d_cand.b.p, W, d_acc.b.p, d_mhead.b.p, d_mcell.b.p, d_mrod.b.p, d_mnext.b.p,
                       d_status.b.p);
    }
    // 3. within the window
    hipLaunchKernelGGL(k_pop_fill, dim3(grid_for(HW, 256)), dim3(256), 0, ctx->stream, d_whead.b.p, HW, -1);
    hipLaunchKernelGGL(k_pop_winsert, dim3(grid_for(W, POP_THREADS)), dim3(POP_THREADS), 0, ctx->stream, GW,
                       d_cand.b.p, W, d_status.b.p, d_woff.b.p, d_whead.b.p, d_wcell.b.p, d_wcand.b.p,
                       d_wnext.b.p, d_und.b.p + 1);
    for (int round = 0; round <= W; ++round) {
      int und = 0;
      hipLaunchKernelGGL(k_pop_fill, dim3(1), dim3(64), 0, ctx->stream, d_und.b.p, 1, 0);  // [0] only
      hipLaunchKernelGGL(k_pop_resolve, dim3(grid_for(W, POP_THREADS)), dim3(POP_THREADS), 0, ctx->stream, GW,
                         d_cand.b.p, W, d_status.b.p, d_whead.b.p, d_wcell.b.p, d_wcand.b.p, d_wnext.b.p,
                         d_und.b.p);
      RDK_HIP(hipMemcpyAsync(&und, d_und.b.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      RDK_HIP(hipStreamSynchronize(ctx->stream));
      if (und == 0) break;
      if (round == W) {
        ctx->err = "rdk_populate_nonoverlap: window decisions did not settle";
        return RDK_ERR_STATE;
      }
    }
    int mismatch = 0;
    RDK_HIP(hipMemcpyAsync(hstatus.data(), d_status.b.p, W * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipMemcpyAsync(&mismatch, d_und.b.p + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    if (mismatch) {
      ctx->err = "rdk_populate_nonoverlap: device and host cell ranges differ";
      return RDK_ERR_STATE;
    }
    // 4. stream order: rod count and the max_attempts rule
    list.clear();
    moff.clear();
    long long madd = 0;
    for (int j = 0; j < W && done + (int)list.size() < n; ++j) {
      ++drawn;
      if (hstatus[j] == POP_ACC) {
        list.push_back(j);
        moff.push_back((int)(mused + madd));
        int i0[3], i1[3];
        pop_range(G, hc[j], i0, i1);
        madd += pop_ncells(i0, i1);
        hacc.push_back(hc[j]);
        fails = 0;
      } else if (++fails >= max_attempts) {
        stop = true;
        break;
      }
    }
    if (mused + madd > 2147483647LL) {
      ctx->err = "rdk_populate_nonoverlap: grid too large (cell size too small for the rods)";
      return RDK_ERR_ARG;
    }
    const int na = (int)list.size();
    moff.push_back((int)(mused + madd));
    if (na > 0) {
      // grow the main grid's entry arrays, keeping what they hold
      for (Scoped<int>* s : {&d_mcell, &d_mrod, &d_mnext}) {
        if ((size_t)(mused + madd) > s->b.cap) {
          DevBuf<int> nb;
          RDK_TRY(ensure(ctx, nb, (size_t)(mused + madd) * 2));
          if (mused > 0)
            RDK_HIP(hipMemcpyAsync(nb.p, s->b.p, mused * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
          RDK_HIP(hipStreamSynchronize(ctx->stream));
          release(s->b);
          s->b = nb;
        }
      }
      RDK_TRY(ensure(ctx, d_list.b, (size_t)na));
      RDK_TRY(ensure(ctx, d_moff.b, (size_t)na + 1));
      RDK_HIP(hipMemcpyAsync(d_list.b.p, list.data(), na * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
      RDK_HIP(hipMemcpyAsync(d_moff.b.p, moff.data(), (na + 1) * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
      hipLaunchKernelGGL(k_pop_commit, dim3(grid_for(na, POP_THREADS)), dim3(POP_THREADS), 0, ctx->stream, GM,
                         d_cand.b.p, d_list.b.p, na, done, d_moff.b.p, d_acc.b.p, d_mhead.b.p, d_mcell.b.p,
                         d_mrod.b.p, d_mnext.b.p, d_und.b.p + 1);
      RDK_HIP(hipMemcpyAsync(&mismatch, d_und.b.p + 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      RDK_HIP(hipStreamSynchronize(ctx->stream));
      if (mismatch) {
        ctx->err = "rdk_populate_nonoverlap: device and host cell ranges differ";
        return RDK_ERR_STATE;
      }
      mused += madd;
      done += na;
    }
  }
  RDK_HIP(hipGetLastError());
  // outputs: the accepted candidates (host copies) and their quaternions,
  // built as the reference builds them (main.cpp:1931-1939, 2109-2112)
  for (int i = 0; i < done; ++i) {
    const Cand& c = hacc[i];
    centres[3 * i] = c.cx;
    centres[3 * i + 1] = c.cy;
    centres[3 * i + 2] = c.cz;
    dirs[3 * i] = c.ux;
    dirs[3 * i + 1] = c.uy;
    dirs[3 * i + 2] = c.uz;
    if (quats) {
      const float d = std::min(std::max(c.uy, -1.0f), 1.0f);
      const float ang = std::acos(d);
      float q[4] = {1.0f, 0.0f, 0.0f, 0.0f};
      if (!(ang < 1e-6f)) {
        // axis = normalize(cross(+Y, dir)) = normalize(dir.z, 0, -dir.x)
        const float ax = 1.0f * c.uz - 0.0f * c.uy, ay = 0.0f * c.ux - 0.0f * c.uz, az = 0.0f * c.uy - 1.0f * c.ux;
        const float l2 = (ax * ax + ay * ay) + az * az;
        const float inv = 1.0f / std::sqrt(l2);
        const float sh = std::sin(ang * 0.5f);
        q[0] = std::cos(ang * 0.5f);
        q[1] = (ax * inv) * sh;
        q[2] = (ay * inv) * sh;
        q[3] = (az * inv) * sh;
      }
      std::memcpy(quats + 4 * i, q, sizeof(q));
    }
  }
  *placed = done;
  if (attempts) *attempts = drawn;
  return RDK_OK;
}

}  // namespace rdk

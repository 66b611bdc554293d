// rdk_forces.hip -- soft-penalty and Hertz-Mindlin contact forces on capsules
// with an atomic-free body accumulation.
//
// Per contact (canonical (a,b) order) one thread evaluates the force law
// (SoftContactSolver::computeForces soft_contact.cpp:80-486;
// HertzMindlinSolver hertz_mindlin.cpp:183-393 with capsule geometry) and
// writes the force on B, the torques on A and B and the updated tangential
// history.  A stable radix sort of the 2C (body, contact) entries gives every
// body its contacts in increasing contact index; one thread per body then
// sums them in that order, which is the order the serial reference loop
// applies them, so the sums are deterministic and need no atomics.
//
// Tangential history (the Mindlin spring xi, or the Cundall-Strack spring
// force) is kept as arrays sorted by pair key.  Each evaluation looks the
// current pairs up by binary search and then merges the current pairs with
// the retained old entries (pruned after 4 evaluations without contact for
// Hertz-Mindlin, hertz_mindlin.hpp:143 and .cpp:385-393; never pruned for
// Cundall-Strack, whose prune routine the reference never calls).
#include <algorithm>
#include <cmath>

#include "rdk_internal.hpp"

namespace rdk {

__device__ __forceinline__ d3 f_ld3(double* const* a, int i) { return mk3(a[0][i], a[1][i], a[2][i]); }

__device__ __forceinline__ int lb_u64(const uint64_t* a, int n, uint64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

struct CIn {
  const int *a, *b;
  const double *pa[3], *pb[3], *n[3], *sh[3], *dist, *hs;
  const uint64_t* keys;
};

struct COut {
  double* c[12];  // F on B (0..2), tau A (3..5), tau B (6..8), history vector (9..11)
  double* keep;   // null, or 6 doubles per contact: normal force on B, friction force on A
};

struct Hist {
  const uint64_t* keys;
  const double* xi[3];
  int n;
};

__device__ __forceinline__ d3 hist_lookup(const Hist& H, uint64_t key) {
  if (H.n <= 0) return mk3(0, 0, 0);
  int j = lb_u64(H.keys, H.n, key);
  if (j < H.n && H.keys[j] == key) return mk3(H.xi[0][j], H.xi[1][j], H.xi[2][j]);
  return mk3(0, 0, 0);
}

__device__ __forceinline__ void put(const COut& O, int base, int i, d3 v) {
  O.c[base][i] = v.x;
  O.c[base + 1][i] = v.y;
  O.c[base + 2][i] = v.z;
}

__device__ __forceinline__ void block_add(double v, double* out) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if (__lane_id() == 0 && v != 0.0) atomicAdd(out, v);
}

struct SoftK {
  double delta, k, damping, mu, mu_s, nu, kt, deadband, dt, K1, K2;
  int friction, model;
  double g[3];
};

// soft_contact.cpp:215-253
__device__ __forceinline__ double soft_grad(double d, double h, double delta, double K1) {
  if (d > h - delta) {
    const double arg = K1 * (h - d);
    const double e = exp(arg);
    const double lt = log(1.0 + e);
    return -2.0 / (K1 * K1) * lt * e / (1.0 + e);
  }
  return -2.0 * (h - d);
}
__device__ __forceinline__ double soft_pe(double d, double h, double delta, double K1) {
  if (d > h - delta) {
    const double lt = log(1.0 + exp(K1 * (h - d)));
    const double base = lt / K1;
    return base * base;
  }
  const double pen = h - d;
  return pen * pen;
}

__global__ void k_soft_force(BodyPtrs B, CIn C, int nc, SoftK P, Hist H, COut O, double* __restrict__ pe_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double pe = 0.0;
  if (i < nc) {
    const int a = C.a[i], b = C.b[i];
    const d3 n = mk3(C.n[0][i], C.n[1][i], C.n[2][i]);
    const d3 pa = mk3(C.pa[0][i], C.pa[1][i], C.pa[2][i]);
    const d3 pb = mk3(C.pb[0][i], C.pb[1][i], C.pb[2][i]);
    const d3 sh = mk3(C.sh[0][i], C.sh[1][i], C.sh[2][i]);
    const double dist = C.dist[i], hs = C.hs[i];
    const double grad = soft_grad(dist, hs, P.delta, P.K1);
    const double fm = -P.k * grad;
    d3 fa = n * (-fm), fb = n * fm;
    const d3 ra = pa - f_ld3(B.x, a);
    const d3 rb = pb - (f_ld3(B.x, b) + sh);
    const d3 va = f_ld3(B.v, a) + cross(f_ld3(B.w, a), ra);
    const d3 vb = f_ld3(B.v, b) + cross(f_ld3(B.w, b), rb);
    if (P.damping > 1e-9) {
      const d3 vr = va - vb;
      const double vn = dot(vr, n);
      const d3 fd = n * (-P.damping * vn);
      fa = fa + fd;
      fb = fb - fd;
    }
    d3 fra = mk3(0, 0, 0);
    d3 hist = mk3(0, 0, 0);
    bool has_hist = false;
    if (P.friction) {
      const d3 vr = va - vb;
      const double fn = len(fa);
      const d3 vt = vr - n * dot(vr, n);
      const double vtm = len(vt);
      if (P.model == 1) {
        // Cundall-Strack history lives in the spring force itself
        d3 ft = hist_lookup(H, C.keys[i]);
        has_hist = true;
        if (vtm >= 1e-10) {
          ft = ft - vt * (P.kt * P.dt);
          const double nc2 = dot(ft, n);
          ft = ft - n * nc2;
          const double ftm = len(ft);
          const double lim = P.mu * fn;
          if (ftm > lim) {
            if (ftm > 1e-8) ft = ft * (lim / ftm);
            else ft = mk3(0, 0, 0);
          }
          fra = ft;
        }
        hist = ft;
      } else if (vtm >= 1e-10) {
        if (P.model == 2) {
          if (vtm < P.deadband) {
            const double ma = B.inv_mass[a] > 0 ? B.mass[a] : 0.0, mb = B.inv_mass[b] > 0 ? B.mass[b] : 0.0;
            double me = 0.0;
            if (ma > 0 && mb > 0) me = (ma * mb) / (ma + mb);
            else if (ma > 0) me = ma;
            else if (mb > 0) me = mb;
            const d3 Fstop = vt * (-(me / P.dt));
            const d3 g = mk3(P.g[0], P.g[1], P.g[2]);
            const d3 gt = g - n * dot(g, n);
            const d3 Fdrive = gt * me;
            const d3 Freq = (-Fdrive) + Fstop;
            const double Freqm = len(Freq);
            const double maxs = P.mu_s * fn;
            if (Freqm <= maxs) fra = Freq;
            else if (Freqm > 1e-12) fra = Freq * (maxs / Freqm);
          } else {
            const d3 dir = vt * (-1.0 / vtm);
            fra = dir * (P.mu * fn);
          }
        } else {
          const double gamma = (vtm > P.nu) ? 1.0 : 2.0 / (1.0 + exp(-P.K2 * vtm)) - 1.0;
          double mu_eff = P.mu;
          if (P.mu_s > P.mu) mu_eff = P.mu + (P.mu_s - P.mu) * exp(-vtm / P.nu);
          const d3 dir = vt * (-1.0 / vtm);
          const double fmag = (mu_eff * gamma) * fn;
          fra = dir * fmag;
        }
      }
    }
    const d3 Ta = fa + fra;
    const d3 Tb = fb + (-fra);
    if (O.keep) {
      double* k6 = O.keep + 6 * (size_t)i;
      k6[0] = fb.x; k6[1] = fb.y; k6[2] = fb.z;
      k6[3] = fra.x; k6[4] = fra.y; k6[5] = fra.z;
    }
    put(O, 0, i, Tb);
    put(O, 3, i, cross(ra, Ta));
    put(O, 6, i, cross(rb, Tb));
    if (has_hist) put(O, 9, i, hist);
    pe = P.k * soft_pe(dist, hs, P.delta, P.K1);
  }
  block_add(pe, pe_out);
}

struct HmK {
  double Es, Gs, gn, gt, mu_k, mu_s, v_c, mu_roll, dt;
  int tangential, rolling, law;
};

__global__ void k_hm_force(BodyPtrs B, CIn C, int nc, HmK P, Hist H, COut O, double* __restrict__ pe_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double pe = 0.0;
  if (i < nc) {
    const int a = C.a[i], b = C.b[i];
    const d3 n = mk3(C.n[0][i], C.n[1][i], C.n[2][i]);
    const d3 pa = mk3(C.pa[0][i], C.pa[1][i], C.pa[2][i]);
    const d3 pb = mk3(C.pb[0][i], C.pb[1][i], C.pb[2][i]);
    const d3 sh = mk3(C.sh[0][i], C.sh[1][i], C.sh[2][i]);
    const double delta = C.hs[i] - C.dist[i];
    const double rA = B.r[a], rB = B.r[b];
    const double Rs = (rA * rB) / (rA + rB);
    const double ms = (B.mass[a] * B.mass[b]) / (B.mass[a] + B.mass[b]);
    const d3 ra = pa - f_ld3(B.x, a);
    const d3 rb = pb - (f_ld3(B.x, b) + sh);
    const double sd = dmax(delta, 0.0);
    const double sqd = sqrt(sd);
    const double sqR = sqrt(dmax(Rs, 0.0));
    const double kn = (4.0 / 3.0) * P.Es * sqR * sqd;
    const double Fel = (P.law == 0) ? kn * sd * sqd : kn * sd;
    const d3 wa = f_ld3(B.w, a), wb = f_ld3(B.w, b);
    const d3 va = f_ld3(B.v, a) + cross(wa, ra);
    const d3 vb = f_ld3(B.v, b) + cross(wb, rb);
    const d3 vr = vb - va;
    const double vn = dot(vr, n);
    const double Cn = P.gn * 2.0 * sqrt(ms * kn);
    const double Fn = dmax(0.0, Fel + (-Cn * vn));
    const d3 Fnv = n * Fn;
    d3 Ft = mk3(0, 0, 0);
    if (P.tangential) {
      const double kt = 8.0 * P.Gs * sqR * sqd;
      const d3 vt = vr - n * vn;
      d3 xi = hist_lookup(H, C.keys[i]);
      xi = xi + vt * P.dt;
      const d3 Fte = xi * (-kt);
      const double Ct = P.gt * 2.0 * sqrt(ms * kt);
      const d3 Ftd = vt * (-Ct);
      Ft = Fte + Ftd;
      const double vtm = len(vt);
      double mue = P.mu_k;
      if (P.mu_s > P.mu_k && P.v_c > 0.0) mue += (P.mu_s - P.mu_k) * exp(-vtm / P.v_c);
      const double Fnm = len(Fnv);
      const double Ftm = len(Ft);
      const double Ftmax = mue * Fnm;
      if (Ftm > Ftmax && Ftm > 1e-10) {
        Ft = Ft * (Ftmax / Ftm);
        if (kt > 1e-10) xi = Ft * (-1.0 / kt);
      }
      put(O, 9, i, xi);
    }
    d3 ta = mk3(0, 0, 0), tb = mk3(0, 0, 0);
    if (P.rolling) {
      const d3 wr = wa - wb;
      const double wm = len(wr);
      if (wm >= 1e-10) {
        const d3 wh = wr * (1.0 / wm);
        const double trm = P.mu_roll * Rs * len(Fnv);
        const d3 tr = wh * (-trm);
        ta = tr * 0.5;
        tb = (-tr) * 0.5;
      }
    }
    const d3 F = Fnv + Ft;
    put(O, 0, i, F);
    put(O, 3, i, cross(ra, -F) + ta);
    put(O, 6, i, cross(rb, F) + tb);
    pe = (8.0 / 15.0) * P.Es * sqrt(Rs) * pow(delta, 2.5);
  }
  block_add(pe, pe_out);
}

// Sum the contact terms of every body in increasing contact order.
__global__ void k_body_reduce(BodyPtrs B, int nbody, const int* __restrict__ bstart, const int* __restrict__ csr,
                              const int* __restrict__ ca, COut O) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbody) return;
  const int k0 = bstart[i], k1 = bstart[i + 1];
  if (k0 == k1) return;
  d3 f = f_ld3(B.f, i), t = f_ld3(B.tau, i);
  for (int k = k0; k < k1; ++k) {
    const int j = csr[k];
    const d3 F = mk3(O.c[0][j], O.c[1][j], O.c[2][j]);
    if (ca[j] == i) {
      f = f + (-F);
      t = t + mk3(O.c[3][j], O.c[4][j], O.c[5][j]);
    } else {
      f = f + F;
      t = t + mk3(O.c[6][j], O.c[7][j], O.c[8][j]);
    }
  }
  B.f[0][i] = f.x;
  B.f[1][i] = f.y;
  B.f[2][i] = f.z;
  B.tau[0][i] = t.x;
  B.tau[1][i] = t.y;
  B.tau[2][i] = t.z;
}

__global__ void k_csr_contacts(const int* __restrict__ ca, const int* __restrict__ cb, int nc, uint32_t* __restrict__ key,
                               int* __restrict__ val) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nc) return;
  key[2 * j] = (uint32_t)ca[j];
  val[2 * j] = j;
  key[2 * j + 1] = (uint32_t)cb[j];
  val[2 * j + 1] = j;
}

__global__ void k_body_start_c(const uint32_t* __restrict__ key, int ne, int nbody, int* __restrict__ start) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nbody) return;
  int lo = 0, hi = ne;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (key[mid] < (uint32_t)i) lo = mid + 1;
    else hi = mid;
  }
  start[i] = lo;
}

// ---- history merge ---------------------------------------------------------
__global__ void k_hist_flags(const uint64_t* __restrict__ okeys, const uint64_t* __restrict__ olast, int nold,
                             const uint64_t* __restrict__ ckeys, int nc, uint64_t frame, uint64_t horizon,
                             int* __restrict__ flag) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > nold) return;
  if (j == nold) {
    flag[j] = 0;
    return;
  }
  const uint64_t key = okeys[j];
  const int p = lb_u64(ckeys, nc, key);
  const bool present = p < nc && ckeys[p] == key;
  const bool keep = !present && (frame - olast[j] <= horizon) && olast[j] != 0;
  flag[j] = keep ? 1 : 0;
}

__global__ void k_hist_merge_cur(const uint64_t* __restrict__ ckeys, int nc, COut O, const uint64_t* __restrict__ okeys,
                                 int nold, const int* __restrict__ scan, uint64_t frame, uint64_t* __restrict__ nkeys,
                                 double* __restrict__ n0, double* __restrict__ n1, double* __restrict__ n2,
                                 uint64_t* __restrict__ nlast) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const uint64_t key = ckeys[i];
  const int lb = lb_u64(okeys, nold, key);
  const int pos = i + scan[lb];
  nkeys[pos] = key;
  n0[pos] = O.c[9][i];
  n1[pos] = O.c[10][i];
  n2[pos] = O.c[11][i];
  nlast[pos] = frame;
}

__global__ void k_hist_merge_old(const uint64_t* __restrict__ okeys, const double* __restrict__ o0,
                                 const double* __restrict__ o1, const double* __restrict__ o2,
                                 const uint64_t* __restrict__ olast, int nold, const int* __restrict__ flag,
                                 const int* __restrict__ scan, const uint64_t* __restrict__ ckeys, int nc,
                                 uint64_t* __restrict__ nkeys, double* __restrict__ n0, double* __restrict__ n1,
                                 double* __restrict__ n2, uint64_t* __restrict__ nlast) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nold || !flag[j]) return;
  const uint64_t key = okeys[j];
  const int pos = scan[j] + lb_u64(ckeys, nc, key);
  nkeys[pos] = key;
  n0[pos] = o0[j];
  n1[pos] = o1[j];
  n2[pos] = o2[j];
  nlast[pos] = olast[j];
}

// ---------------------------------------------------------------------------

static int build_contact_csr(rdk_ctx* ctx, const int*& csr) {
  const int nc = ctx->nc, n = ctx->n, nb = 256;
  RDK_TRY(ensure(ctx, ctx->csr_key, 2 * nc));
  RDK_TRY(ensure(ctx, ctx->csr_key_alt, 2 * nc));
  RDK_TRY(ensure(ctx, ctx->csr_val, 2 * nc));
  RDK_TRY(ensure(ctx, ctx->csr_val_alt, 2 * nc));
  RDK_TRY(ensure(ctx, ctx->body_start, n + 1));
  {
    ProfScope ps(ctx, K_CSR);
    hipLaunchKernelGGL(k_csr_contacts, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, ctx->ca.p, ctx->cb.p, nc,
                       ctx->csr_key.p, ctx->csr_val.p);
  }
  bool alt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->csr_key.p, ctx->csr_key_alt.p, ctx->csr_val.p, ctx->csr_val_alt.p, 2 * nc,
                         bits_for((uint64_t)n), alt));
  const uint32_t* skey = alt ? ctx->csr_key_alt.p : ctx->csr_key.p;
  csr = alt ? ctx->csr_val_alt.p : ctx->csr_val.p;
  {
    ProfScope ps(ctx, K_CSR);
    hipLaunchKernelGGL(k_body_start_c, dim3(grid_for(n + 1, nb)), dim3(nb), 0, ctx->stream, skey, 2 * nc, n,
                       ctx->body_start.p);
  }
  return RDK_OK;
}

static int history_update(rdk_ctx* ctx, COut O, uint64_t horizon) {
  const int nc = ctx->nc, nold = ctx->nhist, nb = 256;
  const uint64_t frame = ctx->hist_frame;
  RDK_TRY(ensure(ctx, ctx->hflag, nold + 1));
  RDK_TRY(ensure(ctx, ctx->hscan, nold + 1));
  if (nold > 0 || true) {
    ProfScope ps(ctx, K_HISTORY);
    hipLaunchKernelGGL(k_hist_flags, dim3(grid_for(nold + 1, nb)), dim3(nb), 0, ctx->stream, ctx->hkeys.p,
                       ctx->hlast.p, nold, ctx->pkeys.p, nc, frame, horizon, ctx->hflag.p);
  }
  RDK_TRY(exclusive_scan_int(ctx, ctx->hflag.p, ctx->hscan.p, nold + 1));
  int kept = 0;
  RDK_HIP(hipMemcpyAsync(&kept, ctx->hscan.p + nold, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  const int nnew = nc + kept;
  RDK_TRY(ensure(ctx, ctx->hkeys_new, std::max(nnew, 1)));
  RDK_TRY(ensure(ctx, ctx->hlast_new, std::max(nnew, 1)));
  for (int k = 0; k < 3; ++k) RDK_TRY(ensure(ctx, ctx->hxi_new[k], std::max(nnew, 1)));
  {
    ProfScope ps(ctx, K_HISTORY);
    if (nc > 0)
      hipLaunchKernelGGL(k_hist_merge_cur, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, ctx->pkeys.p, nc, O,
                         ctx->hkeys.p, nold, ctx->hscan.p, frame, ctx->hkeys_new.p, ctx->hxi_new[0].p,
                         ctx->hxi_new[1].p, ctx->hxi_new[2].p, ctx->hlast_new.p);
    if (nold > 0)
      hipLaunchKernelGGL(k_hist_merge_old, dim3(grid_for(nold, nb)), dim3(nb), 0, ctx->stream, ctx->hkeys.p,
                         ctx->hxi[0].p, ctx->hxi[1].p, ctx->hxi[2].p, ctx->hlast.p, nold, ctx->hflag.p, ctx->hscan.p,
                         ctx->pkeys.p, nc, ctx->hkeys_new.p, ctx->hxi_new[0].p, ctx->hxi_new[1].p,
                         ctx->hxi_new[2].p, ctx->hlast_new.p);
  }
  std::swap(ctx->hkeys, ctx->hkeys_new);
  std::swap(ctx->hlast, ctx->hlast_new);
  for (int k = 0; k < 3; ++k) std::swap(ctx->hxi[k], ctx->hxi_new[k]);
  ctx->nhist = nnew;
  return RDK_OK;
}

int forces_run(rdk_ctx* ctx, int model, const void* params, double* pe) {
  const int nc = ctx->nc, nb = 256;
  RDK_TRY(ensure(ctx, ctx->pe, 2));
  RDK_HIP(hipMemsetAsync(ctx->pe.p, 0, sizeof(double), ctx->stream));
  for (int k = 0; k < 12; ++k) RDK_TRY(ensure(ctx, ctx->cf[k], std::max(nc, 1)));
  COut O;
  for (int k = 0; k < 12; ++k) O.c[k] = ctx->cf[k].p;
  O.keep = nullptr;
  if (ctx->keep_soft) {
    ctx->soft_kept_n = 0;  // a Hertz-Mindlin evaluation keeps nothing
    if (model == 0) {
      RDK_TRY(ensure(ctx, ctx->soft_kept, 6 * (size_t)std::max(nc, 1)));
      O.keep = ctx->soft_kept.p;
      ctx->soft_kept_n = nc;
    }
  }
  CIn C;
  C.a = ctx->ca.p;
  C.b = ctx->cb.p;
  for (int k = 0; k < 3; ++k) {
    C.pa[k] = ctx->cpa[k].p;
    C.pb[k] = ctx->cpb[k].p;
    C.n[k] = ctx->cn[k].p;
    C.sh[k] = ctx->csh[k].p;
  }
  C.dist = ctx->cdist.p;
  C.hs = ctx->chs.p;
  C.keys = ctx->pkeys.p;
  Hist H;
  H.keys = ctx->hkeys.p;
  for (int k = 0; k < 3; ++k) H.xi[k] = ctx->hxi[k].p;
  H.n = ctx->nhist;
  BodyPtrs B = body_ptrs(ctx);
  bool uses_hist = false;
  uint64_t horizon = 0;
  if (model == 0) {
    const rdk_soft_params* p = (const rdk_soft_params*)params;
    SoftK K;
    K.delta = p->delta;
    K.k = p->k_scaler;
    K.damping = p->damping;
    K.mu = p->mu;
    K.mu_s = p->mu_static;
    K.nu = p->nu;
    K.kt = p->kt;
    K.deadband = p->vel_deadband;
    K.dt = ctx->phys.dt;
    K.K1 = 15.0 / p->delta;
    K.K2 = 15.0 / p->nu;
    K.friction = p->enable_friction;
    K.model = p->friction_model;
    for (int k = 0; k < 3; ++k) K.g[k] = ctx->phys.gravity[k];
    uses_hist = p->enable_friction && p->friction_model == 1;
    horizon = ~(uint64_t)0;
    if (uses_hist) ctx->hist_frame++;
    if (nc > 0) {
      ProfScope ps(ctx, K_FORCE);
      hipLaunchKernelGGL(k_soft_force, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, B, C, nc, K, H, O, ctx->pe.p);
    }
  } else {
    const rdk_hm_params* p = (const rdk_hm_params*)params;
    HmK K;
    K.gn = 0.0;
    if (p->restitution < 1.0) {
      const double le = std::log(p->restitution);
      K.gn = -le / std::sqrt(M_PI * M_PI + le * le);
    }
    K.gt = K.gn;
    K.Es = p->E / (2.0 * (1.0 - p->nu_poisson * p->nu_poisson));
    K.Gs = p->E / (2.0 * (1.0 + p->nu_poisson));
    K.mu_k = p->mu_k;
    K.mu_s = p->mu_s;
    K.v_c = p->v_c;
    K.mu_roll = p->mu_roll;
    K.dt = ctx->phys.dt;
    K.tangential = p->enable_tangential;
    K.rolling = p->enable_rolling;
    K.law = p->law;
    uses_hist = true;
    horizon = 4;
    ctx->hist_frame++;
    if (nc > 0) {
      if (!p->enable_tangential)
        for (int k = 9; k < 12; ++k) RDK_HIP(hipMemsetAsync(O.c[k], 0, sizeof(double) * nc, ctx->stream));
      ProfScope ps(ctx, K_FORCE);
      hipLaunchKernelGGL(k_hm_force, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, B, C, nc, K, H, O, ctx->pe.p);
    }
  }
  ctx->work[W_FORCE_CONTACTS] += nc;
  if (nc > 0) {
    const int* csr = nullptr;
    RDK_TRY(build_contact_csr(ctx, csr));
    ProfScope ps(ctx, K_BODY_REDUCE);
    hipLaunchKernelGGL(k_body_reduce, dim3(grid_for(ctx->n, nb)), dim3(nb), 0, ctx->stream, B, ctx->n,
                       ctx->body_start.p, csr, ctx->ca.p, O);
  }
  if (uses_hist) RDK_TRY(history_update(ctx, O, horizon));
  if (pe) {
    RDK_HIP(hipMemcpyAsync(pe, ctx->pe.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
  }
  return RDK_OK;
}

}  // namespace rdk

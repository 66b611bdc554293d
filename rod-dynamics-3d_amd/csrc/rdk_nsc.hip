// rdk_nsc.hip -- NSC hard contact: manifold build, plane-wall manifolds,
// contact-graph colouring and the coloured projected Gauss-Seidel (PSOR)
// solve with the reference's friction clamp (NscContactSolver,
// nsc_contact.cpp:78-117,190-321,325-655).
//
// The reference sweeps manifolds serially in sorted (a,b) order.  That
// order cannot run in parallel, so manifolds are coloured such that no two
// manifolds of one colour share a body; the manifolds of a colour are then
// updated concurrently with exactly the serial update rule, which makes a
// coloured sweep identical to a serial sweep in colour order.  Colouring is
// a deterministic speculative greedy (smallest colour free among coloured
// neighbours; ties among same-round picks go to the lower manifold index).
#include <algorithm>
#include <cstring>

#include <hip/hip_ext.h>

#include <climits>

#include "rdk_internal.hpp"

namespace rdk {

// manifold field offsets (field pairs interleaved per manifold: mf_off)
enum MF {
  F_N = 0, F_T1 = 3, F_T2 = 6, F_RA = 9, F_RB = 12,
  F_IAN = 15, F_IAT1 = 18, F_IAT2 = 21, F_IBN = 24, F_IBT1 = 27, F_IBT2 = 30,
  F_IMA = 33, F_IMB = 34, F_GN = 35, F_GT1 = 36, F_GT2 = 37, F_PHI = 38, F_VNPRE = 39, F_E = 40,
  F_GTB = 41, F_COUNT = 42
};

static_assert(F_COUNT % 2 == 0, "manifold fields are stored in pairs");

struct MfPtrs {
  double* f;
  size_t stride;
  int *a, *b;
  double* lam[3];
  __device__ __forceinline__ double& at(int field, int i) const { return f[mf_off(field, i, stride)]; }
  __device__ __forceinline__ d3 get3(int field, int i) const {
    return mk3(at(field, i), at(field + 1, i), at(field + 2, i));
  }
  __device__ __forceinline__ int ia(int i) const { return a[i]; }
  __device__ __forceinline__ int ib(int i) const { return b[i]; }
  __device__ __forceinline__ double& lr(int c, int i) const { return lam[c][i]; }
  __device__ __forceinline__ void put3(int field, int i, d3 v) const {
    at(field, i) = v.x;
    at(field + 1, i) = v.y;
    at(field + 2, i) = v.z;
  }
};

__device__ __forceinline__ d3 ld3(double* const* a, int i) { return mk3(a[0][i], a[1][i], a[2][i]); }
__device__ __forceinline__ void st3(double* const* a, int i, d3 v) {
  a[0][i] = v.x;
  a[1][i] = v.y;
  a[2][i] = v.z;
}
__device__ __forceinline__ q4 ldq(const BodyPtrs& B, int i) { return q4{B.q[0][i], B.q[1][i], B.q[2][i], B.q[3][i]}; }

// I_world^-1 v of body i (zero for static bodies)
__device__ __forceinline__ d3 iw_inv(const BodyPtrs& B, int i, d3 u, d3 v) {
  if (!(B.inv_mass[i] > 0.0)) return mk3(0, 0, 0);
  return inertia_apply(u, B.i_perp_inv[i], B.i_axial_inv[i], v);
}

__device__ __forceinline__ double diag_term(double ima, double imb, d3 rAxd, d3 IArAxd, d3 rBxd, d3 IBrBxd,
                                            double cfm, bool wall) {
  double g;
  if (wall) g = ima + dot(rAxd, IArAxd) + cfm;
  else g = ima + imb + dot(rAxd, IArAxd) + dot(rBxd, IBrBxd) + cfm;
  return dmax(g, 1e-12);
}

__device__ __forceinline__ int lower_bound_u64(const uint64_t* a, int n, uint64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Largest eigenvalue of the 2x2 tangential Delassus block
// [[g_t1, W12], [W12, g_t2]], W12 = (r_a x t1).I_a^-1(r_a x t2) (+ B term).
__device__ __forceinline__ double tangent_block_max(d3 ra, d3 t1, d3 IAt2, d3 rb, d3 IBt2, double g1, double g2,
                                                    bool wall) {
  double w12 = dot(cross(ra, t1), IAt2);
  if (!wall) w12 = w12 + dot(cross(rb, t1), IBt2);
  const double hm = 0.5 * (g1 + g2), hd = 0.5 * (g1 - g2);
  return hm + sqrt(hd * hd + w12 * w12);
}

struct ContactIn {
  const int *a, *b;
  const double *pa[3], *pb[3], *n[3], *sh[3], *dist, *hs;
};

__global__ void k_manifold(BodyPtrs B, ContactIn C, Count ncc, MfPtrs M, double cfm, int warm, const uint64_t* prev_keys,
                           int nprev, const int* __restrict__ nprev_dev, const double* pl0, const double* pl1,
                           const double* pl2) {
  const int nc = ncc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  if (nprev_dev) nprev = *nprev_dev;
  const int a = C.a[i], b = C.b[i];
  const d3 n = mk3(C.n[0][i], C.n[1][i], C.n[2][i]);
  const d3 pa = mk3(C.pa[0][i], C.pa[1][i], C.pa[2][i]);
  const d3 pb = mk3(C.pb[0][i], C.pb[1][i], C.pb[2][i]);
  const d3 sh = mk3(C.sh[0][i], C.sh[1][i], C.sh[2][i]);
  d3 t1, t2;
  tangent_basis(n, t1, t2);
  const d3 xa = ld3(B.x, a), xb = ld3(B.x, b);
  const d3 ra = pa - xa;
  const d3 rb = pb - (xb + sh);
  const d3 ua = axis_of(ldq(B, a)), ub = axis_of(ldq(B, b));
  const double ima = B.inv_mass[a], imb = B.inv_mass[b];
  const d3 dirs[3] = {n, t1, t2};
  double g[3];
  d3 IA[3], IB[3];
  for (int d = 0; d < 3; ++d) {
    const d3 rAxd = cross(ra, dirs[d]);
    const d3 rBxd = cross(rb, dirs[d]);
    IA[d] = iw_inv(B, a, ua, rAxd);
    IB[d] = iw_inv(B, b, ub, rBxd);
    g[d] = diag_term(ima, imb, rAxd, IA[d], rBxd, IB[d], cfm, false);
  }
  M.put3(F_N, i, n);
  M.put3(F_T1, i, t1);
  M.put3(F_T2, i, t2);
  M.put3(F_RA, i, ra);
  M.put3(F_RB, i, rb);
  M.put3(F_IAN, i, IA[0]);
  M.put3(F_IAT1, i, IA[1]);
  M.put3(F_IAT2, i, IA[2]);
  M.put3(F_IBN, i, IB[0]);
  M.put3(F_IBT1, i, IB[1]);
  M.put3(F_IBT2, i, IB[2]);
  M.at(F_IMA, i) = ima;
  M.at(F_IMB, i) = imb;
  M.at(F_GN, i) = g[0];
  M.at(F_GT1, i) = g[1];
  M.at(F_GT2, i) = g[2];
  M.at(F_PHI, i) = C.dist[i] - C.hs[i];
  const d3 vrel = (ld3(B.v, b) + cross(ld3(B.w, b), rb)) - (ld3(B.v, a) + cross(ld3(B.w, a), ra));
  M.at(F_VNPRE, i) = dot(n, vrel);
  M.at(F_E, i) = dmin(B.rest[a], B.rest[b]);
  M.at(F_GTB, i) = tangent_block_max(ra, t1, IA[2], rb, IB[2], g[1], g[2], false);
  M.a[i] = a;
  M.b[i] = b;
  double l0 = 0.0, l1 = 0.0, l2 = 0.0;
  if (warm == 1 && nprev > 0) {
    const uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint64_t)(uint32_t)b;
    const int j = lower_bound_u64(prev_keys, nprev, key);
    if (j < nprev && prev_keys[j] == key) {
      l0 = dmax(0.0, pl0[j]);
      l1 = pl1[j];
      l2 = pl2[j];
    }
  }
  M.lam[0][i] = l0;
  M.lam[1][i] = l1;
  M.lam[2][i] = l2;
}

// Plane-wall candidates: 4 slots per body (floor a1, floor a2, ceiling a1,
// ceiling a2).  A slot is active when that endpoint's capsule cap penetrates
// the plane (collision.cpp:155-195 geometry, endpoint form).
__device__ __forceinline__ bool wall_slot(const BodyPtrs& B, const SysDev& S, int i, int slot, d3& nrm, d3& pt,
                                          double& sd) {
  const int plane = slot >> 1, e = slot & 1;
  if (plane == 0 && !S.floor_wall) return false;
  if (plane == 1 && !S.ceil_wall) return false;
  if (!(B.inv_mass[i] > 0.0)) return false;
  const d3 x = ld3(B.x, i);
  const d3 u = axis_of(ldq(B, i));
  const double h = B.h[i], r = B.r[i];
  const d3 p = e == 0 ? x - u * h : x + u * h;
  if (plane == 0) {
    sd = (p.z - S.lo[2]) - r;
    nrm = mk3(0, 0, -1);
    pt = p - mk3(0, 0, r);
  } else {
    sd = (S.hi[2] - p.z) - r;
    nrm = mk3(0, 0, 1);
    pt = p + mk3(0, 0, r);
  }
  return sd < 0.0;
}

__global__ void k_wall_flags(BodyPtrs B, const SysDev* __restrict__ sys, int n, int* __restrict__ flag) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * n) return;
  const int i = t >> 2, slot = t & 3;
  const SysDev S = sys[B.sys[i]];
  d3 nrm, pt;
  double sd;
  flag[t] = wall_slot(B, S, i, slot, nrm, pt, sd) ? 1 : 0;
}

__global__ void k_wall_manifold(BodyPtrs B, const SysDev* __restrict__ sys, int nbody, const int* __restrict__ flag,
                                const int* __restrict__ off, int wall_cap, Count basec, MfPtrs M, double cfm,
                                int warm, const uint64_t* __restrict__ prev_keys, int nprev,
                                const int* __restrict__ nprev_dev, const double* __restrict__ pl0,
                                const double* __restrict__ pl1, const double* __restrict__ pl2) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * nbody || !flag[t] || off[t] >= wall_cap) return;  // past a fixed step's reserve: the step fails
  const int base = basec.get();
  if (nprev_dev) nprev = *nprev_dev;
  const int i = t >> 2, slot = t & 3;
  const SysDev S = sys[B.sys[i]];
  d3 n, pt;
  double sd;
  wall_slot(B, S, i, slot, n, pt, sd);
  const int m = base + off[t];
  d3 t1, t2;
  tangent_basis(n, t1, t2);
  const d3 ra = pt - ld3(B.x, i);
  const d3 u = axis_of(ldq(B, i));
  const double ima = B.inv_mass[i];
  const d3 dirs[3] = {n, t1, t2};
  d3 IA[3];
  double g[3];
  const d3 z = mk3(0, 0, 0);
  for (int d = 0; d < 3; ++d) {
    const d3 rAxd = cross(ra, dirs[d]);
    IA[d] = iw_inv(B, i, u, rAxd);
    g[d] = diag_term(ima, 0.0, rAxd, IA[d], z, z, cfm, true);
  }
  M.put3(F_N, m, n);
  M.put3(F_T1, m, t1);
  M.put3(F_T2, m, t2);
  M.put3(F_RA, m, ra);
  M.put3(F_RB, m, z);
  M.put3(F_IAN, m, IA[0]);
  M.put3(F_IAT1, m, IA[1]);
  M.put3(F_IAT2, m, IA[2]);
  M.put3(F_IBN, m, z);
  M.put3(F_IBT1, m, z);
  M.put3(F_IBT2, m, z);
  M.at(F_IMA, m) = ima;
  M.at(F_IMB, m) = 0.0;
  M.at(F_GN, m) = g[0];
  M.at(F_GT1, m) = g[1];
  M.at(F_GT2, m) = g[2];
  M.at(F_PHI, m) = sd;
  const d3 vrel = -(ld3(B.v, i) + cross(ld3(B.w, i), ra));
  M.at(F_VNPRE, m) = dot(n, vrel);
  M.at(F_E, m) = B.rest[i];
  M.at(F_GTB, m) = tangent_block_max(ra, t1, IA[2], z, z, g[1], g[2], true);
  M.a[m] = i;
  M.b[m] = -1 - slot;
  double l0 = 0.0, l1 = 0.0, l2 = 0.0;
  if (warm && nprev > 0) {
    const uint64_t key = ((uint64_t)(uint32_t)i << 2) | (uint64_t)slot;
    const int j = lower_bound_u64(prev_keys, nprev, key);
    if (j < nprev && prev_keys[j] == key) {
      l0 = dmax(0.0, pl0[j]);
      l1 = pl1[j];
      l2 = pl2[j];
    }
  }
  M.lam[0][m] = l0;
  M.lam[1][m] = l1;
  M.lam[2][m] = l2;
}

__global__ void k_copy_wall_keys_lambda(MfPtrs nat, Count basec, Count nwc, uint64_t* __restrict__ pk,
                                        double* __restrict__ p0, double* __restrict__ p1, double* __restrict__ p2,
                                        int* __restrict__ nprev_out, const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int base = basec.get(), nw = nwc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && nprev_out) *nprev_out = nw;
  if (i >= nw) return;
  const int m = base + i;
  pk[i] = ((uint64_t)(uint32_t)nat.a[m] << 2) | (uint64_t)(-1 - nat.b[m]);
  p0[i] = nat.lam[0][m];
  p1[i] = nat.lam[1][m];
  p2[i] = nat.lam[2][m];
}

__global__ void k_fill_color(int* __restrict__ a, Count nc, int v) {
  const int n = nc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---- body -> manifold CSR -------------------------------------------------
__global__ void k_csr_entries(const int* __restrict__ ma, const int* __restrict__ mb, Count nmc, int nbody,
                              uint32_t* __restrict__ key, int* __restrict__ val) {
  const int nm = nmc.get();
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nm) return;
  key[2 * j] = (uint32_t)ma[j];
  val[2 * j] = j;
  const int b = mb[j];
  key[2 * j + 1] = (uint32_t)(b >= 0 ? b : nbody);
  val[2 * j + 1] = j;
}

__global__ void k_body_start(const uint32_t* __restrict__ key, Count nec, int nbody, int* __restrict__ start) {
  const int ne = nec.get();
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

// ---- colouring -------------------------------------------------------------
// Captured steps launch a fixed number of rounds: a round whose predecessor
// left nothing uncoloured (prev_left[0] == 0), a manifold count the
// single-workgroup colouring handles (nm <= small_limit), or a failed step
// (guard) makes it return at once.  Eager rounds pass nullptr and -1.
__device__ __forceinline__ bool color_round_skipped(int nm, int small_limit, const unsigned* prev_left,
                                                    const int* guard) {
  return nm <= small_limit || (prev_left && *prev_left == 0u) || (guard && *guard);
}

__global__ void k_color_assign(const int* __restrict__ ma, const int* __restrict__ mb, Count nmc,
                               const int* __restrict__ bstart, const int* __restrict__ csr,
                               const int* __restrict__ color, int* __restrict__ tent, unsigned* __restrict__ overflow,
                               int small_limit, const unsigned* __restrict__ prev_left, const int* __restrict__ guard) {
  const int nm = nmc.get();
  if (color_round_skipped(nm, small_limit, prev_left, guard)) return;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nm) return;
  if (color[j] >= 0) {
    tent[j] = -1;
    return;
  }
  unsigned long long m0 = 0, m1 = 0;
  const int bodies[2] = {ma[j], mb[j]};
  for (int s = 0; s < 2; ++s) {
    const int bd = bodies[s];
    if (bd < 0) continue;
    for (int k = bstart[bd]; k < bstart[bd + 1]; ++k) {
      const int o = csr[k];
      if (o == j) continue;
      const int c = color[o];
      if (c >= 0) {
        if (c < 64) m0 |= 1ull << c;
        else m1 |= 1ull << (c - 64);
      }
    }
  }
  int pick;
  if (~m0) pick = __ffsll((long long)~m0) - 1;
  else if (~m1) pick = 64 + __ffsll((long long)~m1) - 1;
  else {
    pick = 127;
    atomicAdd(overflow, 1u);
  }
  tent[j] = pick;
}

__global__ void k_color_resolve(const int* __restrict__ ma, const int* __restrict__ mb, Count nmc,
                                const int* __restrict__ bstart, const int* __restrict__ csr,
                                const int* __restrict__ tent, int* __restrict__ color, unsigned* __restrict__ left,
                                unsigned* __restrict__ maxcol, int small_limit, const unsigned* __restrict__ prev_left,
                                const int* __restrict__ guard) {
  const int nm = nmc.get();
  if (color_round_skipped(nm, small_limit, prev_left, guard)) return;
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nm) return;
  const int tj = tent[j];
  if (tj < 0) return;
  bool ok = true;
  const int bodies[2] = {ma[j], mb[j]};
  for (int s = 0; s < 2 && ok; ++s) {
    const int bd = bodies[s];
    if (bd < 0) continue;
    for (int k = bstart[bd]; k < bstart[bd + 1]; ++k) {
      const int o = csr[k];
      if (o < j && tent[o] == tj) {
        ok = false;
        break;
      }
    }
  }
  if (ok) {
    color[j] = tj;
    atomicMax(maxcol, (unsigned)tj);
  } else {
    atomicAdd(left, 1u);
  }
}

__global__ void k_color_keys(const int* __restrict__ color, int nm, uint32_t* __restrict__ key, int* __restrict__ val,
                             int* __restrict__ hist) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nm) return;
  key[j] = (uint32_t)color[j];
  val[j] = j;
  atomicAdd(&hist[color[j]], 1);
}

__global__ void k_permute(MfPtrs src, MfPtrs dst, const int* __restrict__ perm, Count nmc) {
  const int nm = nmc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const int j = perm[i];
  for (int f = 0; f < F_COUNT; ++f) dst.at(f, i) = src.at(f, j);
  dst.a[i] = src.a[j];
  dst.b[i] = src.b[j];
  dst.lam[0][i] = src.lam[0][j];
  dst.lam[1][i] = src.lam[1][j];
  dst.lam[2][i] = src.lam[2][j];
}

__global__ void k_unpermute_lambda(MfPtrs col, MfPtrs nat, const int* __restrict__ perm, Count nmc) {
  const int nm = nmc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  const int j = perm[i];
  nat.lam[0][j] = col.lam[0][i];
  nat.lam[1][j] = col.lam[1][i];
  nat.lam[2][j] = col.lam[2][i];
}

// ---- impulses ----------------------------------------------------------------
// Warm-start impulse application of manifold i (nsc_contact.cpp:344-380).
template <class MP>
__device__ __forceinline__ void warm_apply_one(const BodyPtrs& B, MP& M, int i) {
  const int a = M.ia(i), b = M.ib(i);
  const bool wall = b < 0;
  const double ima = M.at(F_IMA, i), imb = M.at(F_IMB, i);
  d3 va = ld3(B.v, a), wa = ld3(B.w, a);
  d3 vb = mk3(0, 0, 0), wb = mk3(0, 0, 0);
  if (!wall) {
    vb = ld3(B.v, b);
    wb = ld3(B.w, b);
  }
  const int dir_f[3] = {F_N, F_T1, F_T2};
  const int ia_f[3] = {F_IAN, F_IAT1, F_IAT2};
  const int ib_f[3] = {F_IBN, F_IBT1, F_IBT2};
  bool touched = false;
  for (int d = 0; d < 3; ++d) {
    const double l = M.lr(d, i);
    if (l != 0.0) {
      touched = true;
      const d3 dir = M.get3(dir_f[d], i);
      va = va - dir * (l * ima);
      wa = wa - M.get3(ia_f[d], i) * l;
      if (!wall) {
        vb = vb + dir * (l * imb);
        wb = wb + M.get3(ib_f[d], i) * l;
      }
    }
  }
  if (!touched) return;
  st3(B.v, a, va);
  st3(B.w, a, wa);
  if (!wall) {
    st3(B.v, b, vb);
    st3(B.w, b, wb);
  }
}

// Warm-start application for one colour.
__global__ void k_warm_apply(BodyPtrs B, MfPtrs M, int i0, int i1) {
  int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= i1) return;
  warm_apply_one(B, M, i);
}

struct SolveConst {
  double omega, cfm, beta, dt;
  double stall_omega;  // omega after the first stall; 0: stop at the first stall
  int proj;
};

// Stall rule (stop_rule 1, stall_sweeps > 0): at sweep counts stall, 2 stall,
// 4 stall, ... the residual must have dropped by at least 10 % since the
// previous such milestone.  The first time it has not, a solve whose omega is
// above stall_omega (> 0) continues with omega = stall_omega: with high
// friction and restitution, over-relaxed PSOR can settle on a limit cycle that
// plain Gauss-Seidel does not have.  The next stall (or the first, when
// stall_omega is 0) stops the solve (reported as not converged):
// wedged, statically indeterminate groups of rods let the impulses drift along
// a null direction of the Delassus matrix while the residual sits on a floor,
// at any omega.  Solves that keep making progress are untouched.
struct StallTrack {
  int next;
  double ref;
  __host__ __device__ void init(int stall) {
    next = stall;
    ref = -1.0;
  }
  __host__ __device__ bool stalled(int sweeps, double res) {
    if (next <= 0 || sweeps < next) return false;
    const bool st = ref >= 0.0 && !(res <= 0.9 * ref);
    ref = res;
    while (next <= sweeps) next *= 2;
    return st;
  }
};

// Projection of y onto the disc |x| <= r in the metric of the SPD tangential
// block G = [[a, b], [b, c]] (|y| > r > 0): x = (G + nu I)^-1 G y with the nu > 0
// at which |x| = r.  In G's eigenbasis (e1 >= e2 > 0, z = y rotated) the
// residual h(nu) = 1/r - 1/|x(nu)| is convex and decreasing, and the root lies
// in [e2 (|z|/r - 1), e1 (|z|/r - 1)]: Newton from the left end increases nu
// monotonically to the root, and stops once a step no longer increases it
// (at most 16 steps); x is then scaled onto the circle.
__device__ __forceinline__ void disc_project_metric(double a, double b, double c, double r, double& y1, double& y2) {
  const double hm = 0.5 * (a + c), hd = 0.5 * (a - c);
  const double s = sqrt(hd * hd + b * b);
  const double e1 = hm + s, e2 = hm - s;
  double cs = 1.0, sn = 0.0;
  if (s > 0.0) {
    const double c2 = hd / s;
    cs = sqrt(0.5 * (1.0 + c2));
    sn = sqrt(0.5 * (1.0 - c2));
    if (b < 0.0) sn = -sn;
  }
  const double z1 = cs * y1 + sn * y2, z2 = cs * y2 - sn * y1;
  double nu = e2 * (sqrt(z1 * z1 + z2 * z2) / r - 1.0);
  if (!(nu > 0.0)) nu = 0.0;
  double x1, x2, nx;
  for (int it = 0; it < 16; ++it) {
    const double i1 = 1.0 / (e1 + nu), i2 = 1.0 / (e2 + nu);
    x1 = e1 * z1 * i1;
    x2 = e2 * z2 * i2;
    nx = sqrt(x1 * x1 + x2 * x2);
    const double d = x1 * x1 * i1 + x2 * x2 * i2;
    const double nn = nu + (nx - r) * (nx * nx) / (r * d);
    if (!(nn > nu)) break;
    nu = nn;
  }
  x1 = e1 * z1 / (e1 + nu);
  x2 = e2 * z2 / (e2 + nu);
  nx = sqrt(x1 * x1 + x2 * x2);
  if (nx > r) {
    x1 = x1 * (r / nx);
    x2 = x2 * (r / nx);
  }
  y1 = cs * x1 - sn * x2;
  y2 = sn * x1 + cs * x2;
}

// The part of an update that depends on the manifold's fields and omega alone,
// not on the body velocities: the normal step omega / g_n, the tangential
// steps (omega / g_t1 and omega / g_t2, or omega / g_tb) and the inverse of the
// 2x2 tangential block.  A manifold that stays in registers over many updates
// (MfRegP: the register island classes and the LDS class's hot set) computes
// them once per solve (again when omega changes); every value is the same IEEE
// result of the same operations on the same operands as inside the update, so
// the update's results are unchanged bit for bit and each update issues about
// 80 FP64 instructions fewer (four divisions, the block's off-diagonal term).
template <class MP>
struct PrepOf {
  static constexpr bool value = false;
};

// One projected Gauss-Seidel update of manifold i (nsc_contact.cpp:385-552).
// Every manifold field the update may need is loaded first, unconditionally
// (the loads are independent, so they are in flight together, and the
// branches below touch registers only); body velocities follow in a second
// round.  The arithmetic is the serial reference update in the same order.
// EARLY = false loads each field where it is first used instead (fewer live
// registers, for kernels capped at 128 VGPRs).
// store = false computes the update and writes nothing (a sweep the
// huge-island solve launched after its stop rule fired).
template <class MP, bool EARLY = true>
__device__ __forceinline__ double psor_update(const BodyPtrs& B, MP& M, int i, const SolveConst& K, double mu,
                                              bool store = true) {
  constexpr bool PREP = PrepOf<MP>::value;
  const int a = M.ia(i), b = M.ib(i);
  const bool wall = b < 0;
  const d3 ra = M.get3(F_RA, i), rb = M.get3(F_RB, i);
  const d3 n = M.get3(F_N, i);
  const double ima = M.at(F_IMA, i), imb = M.at(F_IMB, i);
  const double phi = M.at(F_PHI, i), vnpre = M.at(F_VNPRE, i), rest = M.at(F_E, i);
  const double gn = M.at(F_GN, i);
  double ln = M.lr(0, i), lt1 = M.lr(1, i), lt2 = M.lr(2, i);
  d3 IAn, IBn, t1, t2, IAt1, IBt1, IAt2, IBt2;
  double g1, g2;
  if (EARLY) {
    IAn = M.get3(F_IAN, i);
    IBn = M.get3(F_IBN, i);
    t1 = M.get3(F_T1, i);
    t2 = M.get3(F_T2, i);
    IAt1 = M.get3(F_IAT1, i);
    IBt1 = M.get3(F_IBT1, i);
    IAt2 = M.get3(F_IAT2, i);
    IBt2 = M.get3(F_IBT2, i);
    g1 = K.proj == 1 ? M.at(F_GTB, i) : M.at(F_GT1, i);
    g2 = K.proj == 1 ? 0.0 : M.at(F_GT2, i);
  }
  d3 va = ld3(B.v, a), wa = ld3(B.w, a);
  d3 vb = mk3(0, 0, 0), wb = mk3(0, 0, 0);
  if (!wall) {
    vb = ld3(B.v, b);
    wb = ld3(B.w, b);
  }
  d3 v_rel = wall ? -(va + cross(wa, ra)) : (vb + cross(wb, rb)) - (va + cross(wa, ra));
  const double b_n = (phi < 0.0) ? K.beta * phi / K.dt : 0.0;
  const double b_rest = (vnpre < 0.0) ? rest * vnpre : 0.0;
  const double w_n = dot(n, v_rel) + b_n + b_rest + K.cfm * ln;
  double ogn;
  if constexpr (PREP) ogn = M.ogn;
  else ogn = K.omega / gn;
  const double delta_n = -ogn * w_n;
  const double old_n = ln;
  ln = dmax(0.0, old_n + delta_n);
  const double dn = ln - old_n;
  if (dn != 0.0) {
    if (!EARLY) IAn = M.get3(F_IAN, i);
    const d3 dvA = n * (dn * ima);
    va = va - dvA;
    wa = wa - IAn * dn;
    v_rel = v_rel + (dvA + cross(IAn * dn, ra));
    if (!wall) {
      if (!EARLY) IBn = M.get3(F_IBN, i);
      const d3 dvB = n * (dn * imb);
      vb = vb + dvB;
      wb = wb + IBn * dn;
      v_rel = v_rel + (dvB + cross(IBn * dn, rb));
    }
  }
  const double max_fric = mu * ln;
  if (!EARLY) {
    t1 = M.get3(F_T1, i);
    t2 = M.get3(F_T2, i);
    g1 = K.proj == 1 ? M.at(F_GTB, i) : M.at(F_GT1, i);
    g2 = K.proj == 1 ? 0.0 : M.at(F_GT2, i);
  }
  double dt1, dt2;
  if (K.proj == 0) {
    const double limit_t1 = sqrt(dmax(0.0, max_fric * max_fric - lt2 * lt2));
    const double w_t1 = dot(t1, v_rel) + K.cfm * lt1;
    double og1;
    if constexpr (PREP) og1 = M.og1;
    else og1 = K.omega / g1;
    const double delta_t1 = -og1 * w_t1;
    const double old_t1 = lt1;
    lt1 = dclamp(old_t1 + delta_t1, -limit_t1, limit_t1);
    dt1 = lt1 - old_t1;
    if (dt1 != 0.0) {
      if (!EARLY) IAt1 = M.get3(F_IAT1, i);
      const d3 dvA = t1 * (dt1 * ima);
      va = va - dvA;
      wa = wa - IAt1 * dt1;
      v_rel = v_rel + (dvA + cross(IAt1 * dt1, ra));
      if (!wall) {
        if (!EARLY) IBt1 = M.get3(F_IBT1, i);
        const d3 dvB = t1 * (dt1 * imb);
        vb = vb + dvB;
        wb = wb + IBt1 * dt1;
        v_rel = v_rel + (dvB + cross(IBt1 * dt1, rb));
      }
    }
    const double limit_t2 = sqrt(dmax(0.0, max_fric * max_fric - lt1 * lt1));
    const double w_t2 = dot(t2, v_rel) + K.cfm * lt2;
    double og2;
    if constexpr (PREP) og2 = M.og2;
    else og2 = K.omega / g2;
    const double delta_t2 = -og2 * w_t2;
    const double old_t2 = lt2;
    lt2 = dclamp(old_t2 + delta_t2, -limit_t2, limit_t2);
    dt2 = lt2 - old_t2;
  } else if (K.proj >= 2) {
    // block step lambda_t - omega G^-1 w_t with the 2x2 tangential block G
    // (g1 = G11, g2 = G22, G12 from the cached I^-1 (r x t2)): a sticking
    // contact's slip is solved in both directions at once.  Outside the disc,
    // 2 projects the step in G's metric and 3 takes the scalar step of 1 from
    // the current impulse instead; both keep 1's fixed points (maximal
    // dissipation), and a sliding solution's block step always leaves the disc.
    if (!EARLY) {
      IAt2 = M.get3(F_IAT2, i);
      if (!wall) IBt2 = M.get3(F_IBT2, i);
    }
    auto block_g12 = [&]() {
      double g = dot(cross(ra, t1), IAt2);
      if (!wall) g = g + dot(cross(rb, t1), IBt2);
      return g;
    };
    // G^-1 from the manifold's fields alone (off the chain of velocity updates)
    double h11, h12, h22;
    if constexpr (PREP) {
      h11 = M.h11;
      h12 = M.h12;
      h22 = M.h22;
    } else {
      const double g12 = block_g12();
      const double det = g1 * g2 - g12 * g12;
      h11 = g2 / det;
      h12 = g12 / det;
      h22 = g1 / det;
    }
    const double w_t1 = dot(t1, v_rel) + K.cfm * lt1;
    const double w_t2 = dot(t2, v_rel) + K.cfm * lt2;
    const double s1 = h11 * w_t1 - h12 * w_t2, s2 = h22 * w_t2 - h12 * w_t1;
    double y1 = lt1 - K.omega * s1;
    double y2 = lt2 - K.omega * s2;
    if (y1 * y1 + y2 * y2 > max_fric * max_fric) {
      if (K.proj == 3) {
        double st;
        if constexpr (PREP) st = M.og1;
        else st = K.omega / M.at(F_GTB, i);
        y1 = lt1 - st * w_t1;
        y2 = lt2 - st * w_t2;
        const double ny = sqrt(y1 * y1 + y2 * y2);
        if (ny > max_fric) {
          const double sc = ny > 0.0 ? max_fric / ny : 0.0;
          y1 = y1 * sc;
          y2 = y2 * sc;
        }
      } else if (max_fric > 0.0) {
        disc_project_metric(g1, block_g12(), g2, max_fric, y1, y2);
      } else {
        y1 = y2 = 0.0;
      }
    }
    dt1 = y1 - lt1;
    dt2 = y2 - lt2;
    lt1 = y1;
    lt2 = y2;
    if (dt1 != 0.0) {
      if (!EARLY) IAt1 = M.get3(F_IAT1, i);
      va = va - t1 * (dt1 * ima);
      wa = wa - IAt1 * dt1;
      if (!wall) {
        if (!EARLY) IBt1 = M.get3(F_IBT1, i);
        vb = vb + t1 * (dt1 * imb);
        wb = wb + IBt1 * dt1;
      }
    }
  } else {
    // Euclidean disc: one scalar step omega/g_tb for the tangential pair
    const double w_t1 = dot(t1, v_rel) + K.cfm * lt1;
    const double w_t2 = dot(t2, v_rel) + K.cfm * lt2;
    double st;
    if constexpr (PREP) st = M.og1;
    else st = K.omega / g1;
    double y1 = lt1 - st * w_t1;
    double y2 = lt2 - st * w_t2;
    const double ny = sqrt(y1 * y1 + y2 * y2);
    if (ny > max_fric) {
      const double sc = ny > 0.0 ? max_fric / ny : 0.0;
      y1 = y1 * sc;
      y2 = y2 * sc;
    }
    dt1 = y1 - lt1;
    dt2 = y2 - lt2;
    lt1 = y1;
    lt2 = y2;
    if (dt1 != 0.0) {
      if (!EARLY) IAt1 = M.get3(F_IAT1, i);
      va = va - t1 * (dt1 * ima);
      wa = wa - IAt1 * dt1;
      if (!wall) {
        if (!EARLY) IBt1 = M.get3(F_IBT1, i);
        vb = vb + t1 * (dt1 * imb);
        wb = wb + IBt1 * dt1;
      }
    }
  }
  if (dt2 != 0.0) {
    if (!EARLY) IAt2 = M.get3(F_IAT2, i);
    va = va - t2 * (dt2 * ima);
    wa = wa - IAt2 * dt2;
    if (!wall) {
      if (!EARLY) IBt2 = M.get3(F_IBT2, i);
      vb = vb + t2 * (dt2 * imb);
      wb = wb + IBt2 * dt2;
    }
  }
  if (!store) return fabs(dn) + fabs(dt1) + fabs(dt2);
  M.lr(0, i) = ln;
  M.lr(1, i) = lt1;
  M.lr(2, i) = lt2;
  if (dn != 0.0 || dt1 != 0.0 || dt2 != 0.0) {
    st3(B.v, a, va);
    st3(B.w, a, wa);
    if (!wall) {
      st3(B.v, b, vb);
      st3(B.w, b, wb);
    }
  }
  return fabs(dn) + fabs(dt1) + fabs(dt2);
}

__device__ __forceinline__ double body_mu(const BodyPtrs& B, const SysDev* sys, int a) { return sys[B.sys[a]].mu; }

__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

// skip (may be null): when *skip is set the sweep writes nothing; it is read
// alongside the manifold fields, so checking it adds no round trip.
__global__ void __launch_bounds__(256) k_psor_sweep(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys, int i0,
                                                    int i1, SolveConst K, double* __restrict__ ref_res,
                                                    const int* __restrict__ skip) {
  int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  const bool sk = skip && *skip != 0;
  double d = 0.0;
  if (i < i1) d = psor_update(B, M, i, K, body_mu(B, sys, M.a[i]), !sk);
  if (!ref_res || sk) return;  // the caller does not need this sweep's largest change
  for (int off = 32; off > 0; off >>= 1) d = dmax(d, __shfl_down(d, off));
  if (__lane_id() == 0 && d > 0.0) atomic_max_pos(ref_res, d);
}

// Fixed-point residual (read-only), velocity units.
template <class MP>
__device__ __forceinline__ double residual_one(const BodyPtrs& B, MP& M, int i, const SolveConst& K, double mu) {
  const int a = M.ia(i), b = M.ib(i);
  const bool wall = b < 0;
  const d3 ra = M.get3(F_RA, i), rb = M.get3(F_RB, i);
  const d3 va = ld3(B.v, a), wa = ld3(B.w, a);
  d3 v_rel;
  if (wall) v_rel = -(va + cross(wa, ra));
  else v_rel = (ld3(B.v, b) + cross(ld3(B.w, b), rb)) - (va + cross(wa, ra));
  const double phi = M.at(F_PHI, i), vnpre = M.at(F_VNPRE, i);
  const double b_n = (phi < 0.0) ? K.beta * phi / K.dt : 0.0;
  const double b_rest = (vnpre < 0.0) ? M.at(F_E, i) * vnpre : 0.0;
  const double ln = M.lr(0, i), lt1 = M.lr(1, i), lt2 = M.lr(2, i);
  const double gn = M.at(F_GN, i), g1 = M.at(F_GT1, i), g2 = M.at(F_GT2, i);
  const double w_n = dot(M.get3(F_N, i), v_rel) + b_n + b_rest + K.cfm * ln;
  const double pn = dmax(0.0, ln - w_n / gn);
  double r = gn * fabs(ln - pn);
  const double mf = mu * ln;
  const double w_t1 = dot(M.get3(F_T1, i), v_rel) + K.cfm * lt1;
  const double w_t2 = dot(M.get3(F_T2, i), v_rel) + K.cfm * lt2;
  double p1, p2;
  if (K.proj == 0) {
    const double L1 = sqrt(dmax(0.0, mf * mf - lt2 * lt2));
    const double L2 = sqrt(dmax(0.0, mf * mf - lt1 * lt1));
    p1 = dclamp(lt1 - w_t1 / g1, -L1, L1);
    p2 = dclamp(lt2 - w_t2 / g2, -L2, L2);
    r = dmax(r, g1 * fabs(lt1 - p1));
    r = dmax(r, g2 * fabs(lt2 - p2));
  } else {
    const double gtb = M.at(F_GTB, i);
    p1 = lt1 - w_t1 / gtb;
    p2 = lt2 - w_t2 / gtb;
    const double ny = sqrt(p1 * p1 + p2 * p2);
    if (ny > mf) {
      const double sc = ny > 0.0 ? mf / ny : 0.0;
      p1 = p1 * sc;
      p2 = p2 * sc;
    }
    const double d1 = lt1 - p1, d2 = lt2 - p2;
    r = dmax(r, gtb * sqrt(d1 * d1 + d2 * d2));
  }
  return r;
}

__global__ void __launch_bounds__(256) k_residual(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys, int i0,
                                                  int i1, SolveConst K, double* __restrict__ out) {
  int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (i < i1) {
    r = residual_one(B, M, i, K, body_mu(B, sys, M.a[i]));
    if (r != r) r = __longlong_as_double(0x7ff8000000000000ll);  // keep NaN visible
  }
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(r, off);
    r = (o != o || o > r) ? o : r;
  }
  if (__lane_id() == 0 && (r > 0.0 || r != r)) atomic_max_pos(out, r);
}

__global__ void k_copy_keys_lambda(const uint64_t* __restrict__ keys, MfPtrs nat, Count ncc, uint64_t* __restrict__ pk,
                                   double* __restrict__ p0, double* __restrict__ p1, double* __restrict__ p2,
                                   int* __restrict__ nprev_out, const int* __restrict__ guard) {
  if (guard && *guard) return;
  const int nc = ncc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && nprev_out) *nprev_out = nc;
  if (i >= nc) return;
  pk[i] = keys[i];
  p0[i] = nat.lam[0][i];
  p1[i] = nat.lam[1][i];
  p2[i] = nat.lam[2][i];
}

// ---- single-workgroup paths for small manifold counts ----------------------
// One workgroup (SMALL_BLOCK threads) runs every colouring round, or the
// warm-start application plus all PSOR sweeps with their residual checks, in
// one launch; workgroup barriers separate colours and rounds.
constexpr int SMALL_BLOCK = 1024;

__device__ __forceinline__ double block_max(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_down(v, off);
    v = (o != o || o > v) ? o : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int k = 0; k < nw; ++k) {
    const double o = red[k];
    r = (o != o || o > r) ? o : r;
  }
  __syncthreads();
  return r;
}

__global__ void __launch_bounds__(SMALL_BLOCK) k_color_small(const int* __restrict__ ma, const int* __restrict__ mb,
                                                             Count nmc, const int* __restrict__ bstart,
                                                             const int* __restrict__ csr, int* __restrict__ color,
                                                             int* __restrict__ tent, unsigned* __restrict__ ctr,
                                                             int small_limit, const int* __restrict__ guard) {
  __shared__ unsigned s_left, s_over, s_max;
  const int nm = nmc.get();
  if (nm > small_limit || (guard && *guard)) return;  // the rounds of a captured step colour it
  for (int j = threadIdx.x; j < nm; j += blockDim.x) color[j] = -1;
  if (threadIdx.x == 0) {
    s_over = 0;
    s_max = 0;
  }
  __syncthreads();
  for (int round = 0; round < 4096; ++round) {
    if (threadIdx.x == 0) s_left = 0;
    // assign: smallest colour free among coloured neighbours
    for (int j = threadIdx.x; j < nm; j += blockDim.x) {
      if (color[j] >= 0) {
        tent[j] = -1;
        continue;
      }
      unsigned long long m0 = 0, m1 = 0;
      const int bodies[2] = {ma[j], mb[j]};
      for (int s = 0; s < 2; ++s) {
        const int bd = bodies[s];
        if (bd < 0) continue;
        for (int k = bstart[bd]; k < bstart[bd + 1]; ++k) {
          const int o = csr[k];
          if (o == j) continue;
          const int c = color[o];
          if (c >= 0) {
            if (c < 64) m0 |= 1ull << c;
            else m1 |= 1ull << (c - 64);
          }
        }
      }
      int pick;
      if (~m0) pick = __ffsll((long long)~m0) - 1;
      else if (~m1) pick = 64 + __ffsll((long long)~m1) - 1;
      else {
        pick = 127;
        atomicAdd(&s_over, 1u);
      }
      tent[j] = pick;
    }
    __syncthreads();
    // resolve: a pick survives unless a lower-index neighbour picked the same colour
    for (int j = threadIdx.x; j < nm; j += blockDim.x) {
      const int tj = tent[j];
      if (tj < 0) continue;
      bool ok = true;
      const int bodies[2] = {ma[j], mb[j]};
      for (int s = 0; s < 2 && ok; ++s) {
        const int bd = bodies[s];
        if (bd < 0) continue;
        for (int k = bstart[bd]; k < bstart[bd + 1]; ++k) {
          const int o = csr[k];
          if (o < j && tent[o] == tj) {
            ok = false;
            break;
          }
        }
      }
      if (ok) {
        color[j] = tj;
        atomicMax(&s_max, (unsigned)tj);
      } else {
        atomicAdd(&s_left, 1u);
      }
    }
    __syncthreads();
    const unsigned left = s_left;
    __syncthreads();
    if (left == 0 || s_over) break;
  }
  if (threadIdx.x == 0) {
    ctr[0] = s_left;
    ctr[1] = s_max;
    ctr[2] = s_over;
  }
}

// colour histogram -> offsets (device), ncolor in ctr[3]
__global__ void k_color_offsets(const int* __restrict__ color, int nm, const unsigned* __restrict__ maxcol,
                                int* __restrict__ cstart, unsigned* __restrict__ ctr) {
  __shared__ int hist[128];
  for (int c = threadIdx.x; c < 128; c += blockDim.x) hist[c] = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < nm; j += blockDim.x) atomicAdd(&hist[color[j]], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nc = (int)(*maxcol) + 1;
    int acc = 0;
    for (int c = 0; c < nc; ++c) {
      cstart[c] = acc;
      acc += hist[c];
    }
    cstart[nc] = acc;
    ctr[3] = (unsigned)nc;
  }
}

struct SmallOut {
  int sweeps;
  int ncolor;
  double res, res_ref;
};

__global__ void __launch_bounds__(SMALL_BLOCK) k_psor_small(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys,
                                                            const int* __restrict__ cstart,
                                                            const unsigned* __restrict__ ctr, int nm, SolveConst K,
                                                            int warm, double tol, int max_sweeps, int check_every,
                                                            int stop_rule, int stall, SmallOut* __restrict__ out) {
  __shared__ double red[SMALL_BLOCK / 64];
  __shared__ int cs[129];
  const int ncolor = (int)ctr[3];
  for (int c = threadIdx.x; c <= ncolor; c += blockDim.x) cs[c] = cstart[c];
  __syncthreads();
  if (warm) {
    for (int c = 0; c < ncolor; ++c) {
      for (int i = cs[c] + threadIdx.x; i < cs[c + 1]; i += blockDim.x) warm_apply_one(B, M, i);
      __syncthreads();
    }
  }
  int sweeps = 0;
  double res = 0.0, res_ref = 0.0;
  const int k = check_every > 0 ? check_every : 1;
  StallTrack trk;
  trk.init(stall);
  if (stop_rule == 1) {
    double r = 0.0;
    for (int i = threadIdx.x; i < nm; i += blockDim.x) r = dmax(r, residual_one(B, M, i, K, sys[B.sys[M.a[i]]].mu));
    res = block_max(r, red);
  }
  while (sweeps < max_sweeps && !(stop_rule == 1 && res <= tol)) {
    const int kk = stop_rule == 1 ? k : 1;
    double dref = 0.0;
    for (int s = 0; s < kk; ++s) {
      dref = 0.0;
      for (int c = 0; c < ncolor; ++c) {
        for (int i = cs[c] + threadIdx.x; i < cs[c + 1]; i += blockDim.x)
          dref = dmax(dref, psor_update<MfPtrs, false>(B, M, i, K, sys[B.sys[M.a[i]]].mu));
        __syncthreads();
      }
    }
    sweeps += kk;
    res_ref = block_max(dref, red);
    double r = 0.0;
    for (int i = threadIdx.x; i < nm; i += blockDim.x) r = dmax(r, residual_one(B, M, i, K, sys[B.sys[M.a[i]]].mu));
    res = block_max(r, red);
    if (stop_rule == 0 && res_ref < 1e-6) break;
    if (stop_rule == 1 && !(res <= tol) && trk.stalled(sweeps, res)) {
      if (K.stall_omega > 0.0 && K.omega > K.stall_omega) K.omega = K.stall_omega;
      else break;
    }
  }
  if (threadIdx.x == 0) {
    out->sweeps = sweeps;
    out->ncolor = ncolor;
    out->res = res;
    out->res_ref = res_ref;
  }
}

// ---- island-parallel PSOR ---------------------------------------------------
// Rod-rod manifolds join bodies into islands (connected components; wall
// manifolds join nothing).  Islands share no body, so Gauss-Seidel on each
// island is independent of the others and the per-island sweeps in colour
// order are exactly the global serial sweep restricted to that island.  Each
// island stops at its own residual.  Tiny islands: one lane, serial.  Medium
// islands: one 64-lane workgroup with manifolds and body velocities staged in
// LDS.  Large islands: one 256-lane workgroup on global memory.

__device__ __forceinline__ int uf_load(int* p, int x) {
  return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int uf_find(int* parent, int x) {
  for (;;) {
    const int p = uf_load(parent, x);
    if (p == x) return x;
    const int gp = uf_load(parent, p);
    if (gp != p) atomicCAS(&parent[x], p, gp);  // path halving
    x = gp;
  }
}

__global__ void k_uf_init(int* __restrict__ parent, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = i;
}

// link the higher root under the lower one: every root ends as the smallest
// body index of its island, so labels are deterministic
__global__ void k_uf_link(const int* __restrict__ ma, const int* __restrict__ mb, Count nmc, int* parent) {
  const int nm = nmc.get();
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int a = ma[m], b = mb[m];
  if (b < 0) return;
  for (int guard = 0; guard < (1 << 20); ++guard) {
    const int ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra == rb) return;
    const int hi = ra > rb ? ra : rb, lo = ra > rb ? rb : ra;
    if (atomicCAS(&parent[hi], hi, lo) == hi) return;
  }
}

__global__ void k_uf_flatten(int* parent, int n, int* __restrict__ label) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) label[i] = uf_find(parent, i);
}

__global__ void k_island_keys(const int* __restrict__ ma, const int* __restrict__ color, Count nmc,
                              const int* __restrict__ label, uint32_t* __restrict__ key, int* __restrict__ val) {
  const int nm = nmc.get();
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  key[m] = ((uint32_t)label[ma[m]] << 7) | (uint32_t)color[m];
  val[m] = m;
}

__global__ void k_body_label_keys(const int* __restrict__ label, int n, uint32_t* __restrict__ key,
                                  int* __restrict__ val) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = (uint32_t)label[i];
  val[i] = i;
}

__global__ void k_island_flags(const uint32_t* __restrict__ skey, Count nmc, int* __restrict__ flag) {
  const int nm = nmc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nm) return;
  flag[i] = (i == 0 || (skey[i] >> 7) != (skey[i - 1] >> 7)) ? 1 : 0;
}

__global__ void k_island_starts(const int* __restrict__ flag, const int* __restrict__ pos, Count nmc,
                                int* __restrict__ istart, int* __restrict__ nisl_out) {
  const int nm = nmc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && nisl_out && nm == 0) *nisl_out = 0;
  if (i >= nm) return;
  if (flag[i]) istart[pos[i]] = i;
  if (i == nm - 1) {
    istart[pos[i] + flag[i]] = nm;
    if (nisl_out) *nisl_out = pos[i] + flag[i];
  }
}

__device__ __forceinline__ int lb_u32(const uint32_t* a, int lo, int hi, uint32_t key) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

struct IslandClass {
  int max_m[3], max_b[3];  // register classes: manifolds and bodies per island
  int large_max, lds_bmax;  // workgroup classes: manifolds, and bodies that fit in LDS
};

// island list per class; counts in cnt[0..5] (tiny, small, medium, large with
// bodies in LDS, large on global memory, giant)
__global__ void k_island_classify(const int* __restrict__ istart, Count nislc, const uint32_t* __restrict__ skey,
                                  const uint32_t* __restrict__ blab, int nbody, IslandClass C,
                                  int* __restrict__ lists, int list_stride, int* __restrict__ cnt,
                                  int* __restrict__ bseg) {
  const int nisl = nislc.get();
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nisl) return;
  const int s = istart[k], e = istart[k + 1];
  const uint32_t L = skey[s] >> 7;
  const int bs = lb_u32(blab, 0, nbody, L), be = lb_u32(blab, bs, nbody, L + 1);
  bseg[2 * k] = bs;
  bseg[2 * k + 1] = be;
  const int size = e - s, nb = be - bs;
  int cls = 5;
  if (size <= C.max_m[0] && nb <= C.max_b[0]) cls = 0;
  else if (size <= C.max_m[1] && nb <= C.max_b[1]) cls = 1;
  else if (size <= C.max_m[2] && nb <= C.max_b[2]) cls = 2;
  else if (size <= C.large_max) cls = nb <= C.lds_bmax ? 3 : 4;
  const int slot = atomicAdd(&cnt[cls], 1);
  lists[cls * list_stride + slot] = k;
}

struct IslandStats {
  int max_sweeps, nonconv, pad0, pad1;
  unsigned long long manifold_sweeps;
  double res, res_ref;
  unsigned long long lds_manifold_sweeps;  // the LDS class's share of manifold_sweeps
  unsigned long long hot_sweeps;          // sweeps of hot sets (LDS class), summed over islands
};

__device__ __forceinline__ void island_report(IslandStats* S, int sweeps, unsigned long long updates, double res,
                                              double res_ref, bool nonconv) {
  atomicMax(&S->max_sweeps, sweeps);
  atomicAdd(&S->manifold_sweeps, updates);
  if (res > 0.0 || res != res) atomic_max_pos(&S->res, res != res ? __longlong_as_double(0x7ff8000000000000ll) : res);
  if (res_ref > 0.0) atomic_max_pos(&S->res_ref, res_ref);
  if (nonconv) atomicAdd(&S->nonconv, 1);
}

struct IslandSolve {
  SolveConst K;
  double tol;
  int max_sweeps, check_every, stop_rule, warm, stall;
  int hot;  // hot-set phase of the LDS class (rdk_nsc_params::hot_set)
};


// One manifold held in registers (every field index is a compile-time
// constant after inlining) with its bodies as local LDS slots.
struct MfReg {
  double f[F_COUNT];
  double lam[3];
  int a, b;
  __device__ __forceinline__ double at(int field, int) const { return f[field]; }
  __device__ __forceinline__ d3 get3(int field, int) const { return mk3(f[field], f[field + 1], f[field + 2]); }
  __device__ __forceinline__ int ia(int) const { return a; }
  __device__ __forceinline__ int ib(int) const { return b; }
  __device__ __forceinline__ double& lr(int c, int) { return lam[c]; }
};

// Every field of manifold i into f[0 .. F_COUNT): one 16-byte load per field
// pair.
__device__ __forceinline__ void load_fields(double* f, const MfPtrs& M, int i) {
  const double2* f2 = reinterpret_cast<const double2*>(M.f);
#pragma unroll
  for (int k = 0; k < F_COUNT / 2; ++k) {
    const double2 v = f2[(size_t)k * M.stride + i];
    f[2 * k] = v.x;
    f[2 * k + 1] = v.y;
  }
}

// A register manifold with its velocity-independent update terms (PrepOf).
struct MfRegP : MfReg {
  double ogn, og1, og2, h11, h12, h22;
};
template <>
struct PrepOf<MfRegP> {
  static constexpr bool value = true;
};

// The terms psor_update computes from R's fields and K.omega alone, with its
// operations: omega / g_n; friction_proj 0 omega / g_t1 and omega / g_t2;
// 1 and 3 omega / g_tb (g_tb is g1 under 1); 2 and 3 the inverse of the
// tangential block.  Called after every load of R and every change of omega.
__device__ __forceinline__ void prep_manifold(MfRegP& R, const SolveConst& K) {
  const double* f = R.f;
  R.ogn = K.omega / f[F_GN];
  R.og1 = K.omega / (K.proj == 0 ? f[F_GT1] : f[F_GTB]);
  R.og2 = K.omega / f[F_GT2];
  if (K.proj >= 2) {
    const d3 ra = R.get3(F_RA, 0), t1 = R.get3(F_T1, 0);
    double g12 = dot(cross(ra, t1), R.get3(F_IAT2, 0));
    if (R.b >= 0) g12 = g12 + dot(cross(R.get3(F_RB, 0), t1), R.get3(F_IBT2, 0));
    const double g1 = f[F_GT1], g2 = f[F_GT2];
    const double det = g1 * g2 - g12 * g12;
    R.h11 = g2 / det;
    R.h12 = g12 / det;
    R.h22 = g1 / det;
  }
}

__device__ __forceinline__ double nan_max(double v, double o) { return (o != o || o > v) ? o : v; }

// max over the GROUP lanes of one island (GROUP <= 64: shuffles inside the
// wavefront; larger groups combine the per-wave maxima through LDS)
template <int GROUP>
__device__ __forceinline__ double group_max(double v, double* red) {
  constexpr int W = GROUP < 64 ? GROUP : 64;
  for (int off = W / 2; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, W));
  if constexpr (GROUP > 64) {
    constexpr int NW = GROUP / 64;
    const int wv = threadIdx.x >> 6, base = (threadIdx.x / GROUP) * NW;
    if ((threadIdx.x & 63) == 0) red[wv] = v;
    __syncthreads();
    v = red[base];
    for (int q = 1; q < NW; ++q) v = nan_max(v, red[base + q]);
    __syncthreads();
  }
  return v;
}

__device__ __forceinline__ int lb_int(const int* a, int n, int key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Island solver with the manifolds in registers: GROUP lanes per island (one
// manifold each), WG / GROUP islands per workgroup, the island's body
// velocities in LDS.  Colours are swept in order and the manifolds of one
// colour in parallel: they share no body, so every sweep equals the serial
// sweep in colour order bit for bit (nsc_contact.cpp:385-552 per update).
template <int GROUP, int WG>
__device__ __forceinline__ void island_reg_block(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys,
                                                 const int* __restrict__ istart, const uint32_t* __restrict__ skey,
                                                 const int* __restrict__ bseg, const int* __restrict__ sbody,
                                                 const int* __restrict__ list, int count, IslandSolve P,
                                                 IslandStats* S, int blk) {
  constexpr int IPB = WG / GROUP;     // islands per workgroup
  constexpr int BCAP = GROUP + 8;     // bodies per island (a connected island of m manifolds has <= m + 1)
  __shared__ double v_l[3][IPB * BCAP], w_l[3][IPB * BCAP];
  __shared__ int body_l[IPB * BCAP];
  __shared__ double red[WG / 64 + 1];
  __shared__ int s_ncol[IPB];     // colours per island
  __shared__ int s_live_ncol[IPB];  // colours of the islands still sweeping (0 once done)
  const int tid = threadIdx.x, slot = tid / GROUP, j = tid % GROUP;
  const int t = blk * IPB + slot;
  const bool has = t < count;
  int s = 0, n = 0, bs = 0, nb = 0;
  if (has) {
    const int k = list[t];
    s = istart[k];
    n = istart[k + 1] - s;
    bs = bseg[2 * k];
    nb = bseg[2 * k + 1] - bs;
  }
  int* bl = body_l + slot * BCAP;
  for (int q = j; q < nb; q += GROUP) {
    const int g = sbody[bs + q];
    bl[q] = g;
    for (int c = 0; c < 3; ++c) {
      v_l[c][slot * BCAP + q] = B.v[c][g];
      w_l[c][slot * BCAP + q] = B.w[c][g];
    }
  }
  if (tid < IPB) s_ncol[tid] = 0;
  __syncthreads();
  const bool live = has && j < n;
  MfRegP R;
  int col = -1;
  if (live) {
    const int i = s + j;
    load_fields(R.f, M, i);
    for (int c = 0; c < 3; ++c) R.lam[c] = M.lam[c][i];
    R.a = slot * BCAP + lb_int(bl, nb, M.a[i]);
    const int gb = M.b[i];
    R.b = gb >= 0 ? slot * BCAP + lb_int(bl, nb, gb) : -1;
    col = (int)(skey[i] & 127u);
    atomicMax(&s_ncol[slot], col + 1);
  } else {
#pragma unroll
    for (int f = 0; f < F_COUNT; ++f) R.f[f] = 0.0;
    R.lam[0] = R.lam[1] = R.lam[2] = 0.0;
    R.a = R.b = 0;
  }
  __syncthreads();
  int ncol = 0;  // colours of the workgroup's islands (the warm start)
  for (int k = 0; k < IPB; ++k) ncol = max(ncol, s_ncol[k]);
  BodyPtrs LB = B;
  for (int c = 0; c < 3; ++c) {
    LB.v[c] = v_l[c];
    LB.w[c] = w_l[c];
  }
  const double mu = has ? sys[B.sys[M.a[s]]].mu : 0.0;
  if (P.warm) {
    for (int c = 0; c < ncol; ++c) {
      if (live && col == c) warm_apply_one(LB, R, 0);
      __syncthreads();
    }
  }
  int sweeps = 0;
  double res = 0.0, res_ref = 0.0;
  bool done = !has || P.max_sweeps <= 0;
  StallTrack trk;
  trk.init(P.stall);
  SolveConst K = P.K;  // omega may drop to stall_omega on a stall (per island)
  if (live) prep_manifold(R, K);
  if (P.stop_rule == 1) {
    res = group_max<GROUP>(live ? residual_one(LB, R, 0, K, mu) : 0.0, red);
    done = done || res <= P.tol;
  }
  const int kk = P.stop_rule == 1 ? P.check_every : 1;
  while (__syncthreads_or(!done)) {
    const bool act = live && !done;
    // sweep only as many colours as the islands still sweeping have (several
    // islands share a workgroup in the tiny class; a converged island's
    // colours would otherwise cost a barrier round each)
    if (j == 0) s_live_ncol[slot] = done ? 0 : s_ncol[slot];
    __syncthreads();
    int ncol_live = 0;
    for (int k = 0; k < IPB; ++k) ncol_live = max(ncol_live, s_live_ncol[k]);
    double d = 0.0;
    for (int it = 0; it < kk; ++it) {
      for (int c = 0; c < ncol_live; ++c) {
        if (act && col == c) d = psor_update(LB, R, 0, K, mu);
        __syncthreads();
      }
    }
    const double rr = group_max<GROUP>(act ? d : 0.0, red);
    const double rs = group_max<GROUP>(act ? residual_one(LB, R, 0, K, mu) : 0.0, red);
    if (!done) {
      sweeps += kk;
      res_ref = rr;
      res = rs;
      if (P.stop_rule == 0 && res_ref < 1e-6) done = true;
      if (P.stop_rule == 1 && res <= P.tol) done = true;
      if (sweeps >= P.max_sweeps) done = true;
      if (!done && P.stop_rule == 1 && trk.stalled(sweeps, res)) {
        if (K.stall_omega > 0.0 && K.omega > K.stall_omega) {
          K.omega = K.stall_omega;
          if (live) prep_manifold(R, K);
        } else {
          done = true;
        }
      }
    }
  }
  if (live)
    for (int c = 0; c < 3; ++c) M.lam[c][s + j] = R.lam[c];
  for (int q = j; q < nb; q += GROUP) {
    const int g = bl[q];
    for (int c = 0; c < 3; ++c) {
      B.v[c][g] = v_l[c][slot * BCAP + q];
      B.w[c][g] = w_l[c][slot * BCAP + q];
    }
  }
  if (has && j == 0)
    island_report(S, sweeps, (unsigned long long)sweeps * (unsigned long long)n, res, res_ref,
                  P.stop_rule == 1 && !(res <= P.tol));
}

// One group of IPB islands per workgroup.  A fixed step sizes the grid from
// the island counts seen so far and reads the count on the device (a count
// beyond the grid fails the step before this launch); guard: a failed step,
// nothing is solved.
template <int GROUP, int WG>
__global__ void __launch_bounds__(WG) k_island_reg(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys,
                                                   const int* __restrict__ istart, const uint32_t* __restrict__ skey,
                                                   const int* __restrict__ bseg, const int* __restrict__ sbody,
                                                   const int* __restrict__ list, Count countc, IslandSolve P,
                                                   IslandStats* S, const int* __restrict__ guard) {
  constexpr int IPB = WG / GROUP;
  if (guard && *guard) return;
  const int count = countc.get();
  if ((int)blockIdx.x * IPB >= count) return;
  island_reg_block<GROUP, WG>(B, M, sys, istart, skey, bseg, sbody, list, count, P, S, blockIdx.x);
}

// island classes: manifolds and bodies per island for the register kernels
constexpr int ISL_G0 = 8, ISL_WG0 = 64;     // tiny: 8 islands per wavefront
constexpr int ISL_G1 = 64, ISL_WG1 = 64;    // small: one island per wavefront
constexpr int ISL_G2 = 256, ISL_WG2 = 256;  // medium: one island per 4-wave workgroup

// Large islands whose bodies fit in LDS: 512 lanes, body velocities in LDS,
// manifold fields and impulses in global memory (L2-resident), each
// manifold's local body slots looked up once per solve.
constexpr int LDS_THREADS = 512;
// 6 doubles per body (the island's sorted body list stays in global memory):
// 161 280 B of the 160 KiB LDS, enough for the percolating cluster of the
// settling bed at 1e5 rods (about 3200 bodies)
constexpr int LDS_BCAP = 3360;

// Hot-set phase of the LDS island solve (IslandSolve::hot, stop rule 1).
// Once at least HOT_MIN_SWEEPS sweeps have run and at most
// max(HOT_MIN_COUNT, m / HOT_FRAC) of the island's m manifolds are above the
// tolerance, the solve sweeps only the hot set H: the manifolds that share a
// body with one whose residual exceeds HOT_TARGET * tol, in the island's
// solve order (colour by colour), at most LDS_THREADS of them, one per lane
// with its fields in registers, until the residual over H is at most
// HOT_TARGET * tol.  Then the whole island is checked again, and the phase
// repeats while the condition holds.  This is Gauss-Seidel on a subset of the
// manifolds (the others' impulses stay as they are) and the stop rule is
// still the residual of the whole island.  On the settling bed the last
// decade or two of the residual sits in a few wall contacts at the stick/slip
// boundary whose local mode takes thousands of sweeps; a full sweep updates
// every manifold of the island for each of them.  A hot set larger than
// LDS_THREADS is not used; the next attempt waits until the sweep count has
// doubled.
constexpr int HOT_MIN_SWEEPS = 64, HOT_MIN_COUNT = 16, HOT_FRAC = 50;
constexpr double HOT_TARGET = 0.5;
// Selective sweeps (IslandSolve::hot == 2): when no hot set is taken and at
// least SEL_MIN_SWEEPS sweeps have run, a round of check_every sweeps covers
// only the same set H (every manifold sharing a body with one above
// HOT_TARGET * tol, any number up to half the island) in solve order, its
// fields loaded from global memory as in full sweeps; the whole island is
// checked after every round.  Late in a solve most of an island is already
// below the tolerance and H is a fraction of it.
constexpr int SEL_MIN_SWEEPS = 32;

struct MfSlots {
  MfPtrs M;
  const int2* slot;  // local (LDS) body slots per manifold, b = -1 for a wall
  __device__ __forceinline__ double& at(int field, int i) const { return M.at(field, i); }
  __device__ __forceinline__ d3 get3(int field, int i) const { return M.get3(field, i); }
  __device__ __forceinline__ int ia(int i) const { return slot[i].x; }
  __device__ __forceinline__ int ib(int i) const { return slot[i].y; }
  __device__ __forceinline__ double& lr(int c, int i) const { return M.lam[c][i]; }
};

__device__ __forceinline__ void load_manifold(MfReg& R, const MfSlots& A, int i) {
  load_fields(R.f, A.M, i);
  for (int q = 0; q < 3; ++q) R.lam[q] = A.lr(q, i);
  R.a = A.ia(i);
  R.b = A.ib(i);
}

__device__ __forceinline__ void island_lds_block(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys,
                                                 const int* __restrict__ istart, const uint32_t* __restrict__ skey,
                                                 const int* __restrict__ bseg, const int* __restrict__ sbody,
                                                 const int* __restrict__ list, IslandSolve P, IslandStats* S,
                                                 int2* __restrict__ slot, int* __restrict__ hlist, int idx) {
  constexpr int HOT_WORDS = (LDS_BCAP + 31) / 32;
  constexpr int NWAVE = LDS_THREADS / 64;
  __shared__ double v_l[3][LDS_BCAP], w_l[3][LDS_BCAP];
  __shared__ double red[NWAVE];
  __shared__ int cs[129];
  __shared__ int hcs[129];                // colour starts inside the hot set
  __shared__ unsigned hot_bits[HOT_WORDS];  // bodies (LDS slots) of manifolds above HOT_TARGET * tol
  __shared__ int s_above, s_wcnt[NWAVE];
  const int k = list[idx];
  const int s = istart[k], e = istart[k + 1];
  const int bs = bseg[2 * k], nb = bseg[2 * k + 1] - bs;
  const int* body_l = sbody + bs;  // the island's bodies, ascending: LDS slot q holds body body_l[q]
  const int tid = threadIdx.x;
  for (int q = tid; q < nb; q += LDS_THREADS) {
    const int g = body_l[q];
    for (int c = 0; c < 3; ++c) {
      v_l[c][q] = B.v[c][g];
      w_l[c][q] = B.w[c][g];
    }
  }
  const uint32_t base = skey[s] & ~127u;
  const int ncol = (int)(skey[e - 1] & 127u) + 1;
  for (int c = tid; c <= ncol; c += LDS_THREADS) cs[c] = lb_u32(skey, s, e, base | (uint32_t)c);
  for (int w = tid; w < HOT_WORDS; w += LDS_THREADS) hot_bits[w] = 0u;
  if (tid == 0) s_above = 0;
  __syncthreads();
  for (int i = s + tid; i < e; i += LDS_THREADS) {
    const int gb = M.b[i];
    slot[i] = make_int2(lb_int(body_l, nb, M.a[i]), gb >= 0 ? lb_int(body_l, nb, gb) : -1);
  }
  __syncthreads();
  BodyPtrs LB = B;
  for (int c = 0; c < 3; ++c) {
    LB.v[c] = v_l[c];
    LB.w[c] = w_l[c];
  }
  MfSlots A{M, slot};
  const double mu = sys[B.sys[M.a[s]]].mu;
  const int m = e - s;
  if (P.warm) {
    for (int c = 0; c < ncol; ++c) {
      for (int i = cs[c] + tid; i < cs[c + 1]; i += LDS_THREADS) warm_apply_one(LB, A, i);
      __syncthreads();
    }
  }
  int sweeps = 0, hot_next = HOT_MIN_SWEEPS;
  unsigned long long updates = 0ull, hot_sweeps = 0ull;
  double res = 0.0, res_ref = 0.0;
  StallTrack trk;
  trk.init(P.stall);
  SolveConst K = P.K;  // omega may drop to stall_omega on a stall
  const bool hot_on = P.hot != 0 && P.stop_rule == 1;
  const double hot_thr = HOT_TARGET * P.tol;
  // Residual of the whole island; with the hot-set phase on, also counts the
  // manifolds above tol (s_above) and marks the bodies of those above
  // HOT_TARGET * tol.  Ends with a barrier (block_max).
  auto check = [&]() -> double {
    double r = 0.0;
    int above = 0;
    for (int i = s + tid; i < e; i += LDS_THREADS) {
      const double ri = residual_one(LB, A, i, K, mu);
      r = dmax(r, ri);
      if (hot_on) {
        above += ri > P.tol ? 1 : 0;
        if (ri > hot_thr) {
          const int2 sl = slot[i];
          atomicOr(&hot_bits[sl.x >> 5], 1u << (sl.x & 31));
          if (sl.y >= 0) atomicOr(&hot_bits[sl.y >> 5], 1u << (sl.y & 31));
        }
      }
    }
    if (above) atomicAdd(&s_above, above);
    return block_max(r, red);
  };
  if (P.stop_rule == 1) res = check();
  const int kk = P.stop_rule == 1 ? P.check_every : 1;
  while (sweeps < P.max_sweeps && !(P.stop_rule == 1 && res <= P.tol)) {
    // ---- hot set or selective round (block-uniform decision from the last check)
    int nh = 0;
    bool sel = false;
    if (hot_on) {
      const int above = s_above;
      const bool want = sweeps >= hot_next && above > 0 && above <= max(HOT_MIN_COUNT, m / HOT_FRAC);
      const bool want_sel = P.hot == 2 && sweeps >= SEL_MIN_SWEEPS;
      __syncthreads();  // everyone has read s_above
      if (want || want_sel) {
        // H = manifolds with a marked body, in solve (storage) order
        const int lane = tid & 63, wid = tid >> 6;
        for (int c0 = s; c0 < e; c0 += LDS_THREADS) {
          const int i = c0 + tid;
          bool in = false;
          if (i < e) {
            const int2 sl = slot[i];
            in = ((hot_bits[sl.x >> 5] >> (sl.x & 31)) & 1u) != 0u ||
                 (sl.y >= 0 && ((hot_bits[sl.y >> 5] >> (sl.y & 31)) & 1u) != 0u);
          }
          const unsigned long long bal = __ballot(in);
          if (lane == 0) s_wcnt[wid] = __popcll(bal);
          __syncthreads();
          int woff = 0, tot = 0;
          for (int w = 0; w < NWAVE; ++w) {
            woff += w < wid ? s_wcnt[w] : 0;
            tot += s_wcnt[w];
          }
          const int pos = nh + woff + __popcll(bal & ((1ull << lane) - 1ull));
          if (in) hlist[s + pos] = i;
          nh += tot;
          __syncthreads();  // s_wcnt is reused
        }
        if (!(want && nh <= LDS_THREADS && nh > 0)) {
          // no hot set: a selective round over the same H when it is at most
          // half the island, else full sweeps
          if (want) hot_next = 2 * sweeps;
          sel = want_sel && nh > 0 && 2 * nh <= m;
          if (!sel) nh = 0;
        }
        if (nh > 0) {
          // colour bounds of H as positions 0..nh, empty colours dropped
          for (int c = tid; c <= ncol; c += LDS_THREADS) hcs[c] = lb_int(hlist + s, nh, cs[c]);
          __syncthreads();
          if (tid == 0) {
            int j = 0;
            for (int c = 0; c < ncol; ++c)
              if (hcs[c] < hcs[c + 1]) hcs[j++] = hcs[c];
            hcs[j] = nh;
            s_wcnt[0] = j;
          }
        }
      }
      for (int w = tid; w < HOT_WORDS; w += LDS_THREADS) hot_bits[w] = 0u;
      if (tid == 0) s_above = 0;
      __syncthreads();
    }
    // ---- sweeps: the whole island (positions = manifold indices s..e), the
    // hot set or a selective round (positions 0..nh, manifold hlist[s + position])
    const bool hot = nh > 0 && !sel;
    const int* bnd = nh > 0 ? hcs : cs;
    const int nbnd = nh > 0 ? s_wcnt[0] : ncol;
    const int* hl = hlist + s;
    auto mfi = [&](int q) { return sel ? hl[q] : q; };  // position -> manifold (full sweeps, selective)
    int ls = 0;
    double d = 0.0;
    // Full sweeps and selective rounds: lane tid takes positions bnd[c] + tid,
    // bnd[c] + tid + 512, ... of each colour.  Hot set (at most 512 positions):
    // lane tid keeps position tid, whose manifold stays in its registers for
    // the round with its update terms prepared once (full sweeps load a new
    // manifold for every update and use the plain register form).
    MfRegP R;
    MfReg& RF = R;  // the plain register form of full sweeps and selective rounds
    if (hot && tid < nh) {
      load_manifold(R, A, hl[tid]);
      prep_manifold(R, K);
    }
    while (true) {
      for (int it = 0; it < kk; ++it) {
        d = 0.0;
        // Each lane's first manifold of the next colour is loaded into registers
        // before the barrier that ends the current colour (its fields and
        // impulses are not written by other colours; the body velocities it
        // needs are read from LDS after the barrier), so the global-memory
        // latency overlaps the wait.
        if (!hot && bnd[0] + tid < bnd[1]) load_manifold(R, A, mfi(bnd[0] + tid));
        for (int c = 0; c < nbnd; ++c) {
          int q = hot ? tid : bnd[c] + tid;
          if (hot ? (tid >= bnd[c] && tid < bnd[c + 1]) : q < bnd[c + 1]) {
            d = dmax(d, hot ? psor_update(LB, R, 0, K, mu) : psor_update(LB, RF, 0, K, mu));
            // a hot-set lane keeps its impulses in registers until the round
            // ends: no global store waits at the colour barrier
            if (!hot) {
              const int i = mfi(q);
              for (int t = 0; t < 3; ++t) A.lr(t, i) = R.lam[t];
            }
            q += LDS_THREADS;
          }
          // further manifolds of this lane reuse the same registers: all fields
          // in one round of loads, then the update
          if (!hot) {
            for (; q < bnd[c + 1]; q += LDS_THREADS) {
              const int i = mfi(q);
              load_manifold(R, A, i);
              d = dmax(d, psor_update(LB, RF, 0, K, mu));
              for (int t = 0; t < 3; ++t) A.lr(t, i) = R.lam[t];
            }
            if (c + 1 < nbnd && bnd[c + 1] + tid < bnd[c + 2]) load_manifold(R, A, mfi(bnd[c + 1] + tid));
          }
          __syncthreads();
        }
      }
      ls += kk;
      if (!hot) break;
      // the hot set's residual from the manifolds in registers (one per lane)
      if (block_max(tid < nh ? residual_one(LB, R, 0, K, mu) : 0.0, red) <= hot_thr || sweeps + ls >= P.max_sweeps)
        break;
    }
    if (hot) {
      if (tid < nh)
        for (int t = 0; t < 3; ++t) A.lr(t, hl[tid]) = R.lam[t];
      __syncthreads();  // stored before the island check reads them
    }
    sweeps += ls;
    updates += (unsigned long long)ls * (unsigned long long)(nh > 0 ? nh : m);
    if (hot) {
      hot_sweeps += (unsigned long long)ls;
    } else {
      res_ref = block_max(d, red);
      if (P.stop_rule == 0) {
        if (res_ref < 1e-6) break;
        continue;
      }
    }
    res = check();
    if (!(res <= P.tol) && trk.stalled(sweeps, res)) {
      if (K.stall_omega > 0.0 && K.omega > K.stall_omega) K.omega = K.stall_omega;
      else break;
    }
  }
  if (P.stop_rule == 0) {  // the reference rule reports the residual it ends with
    double r = 0.0;
    for (int i = s + tid; i < e; i += LDS_THREADS) r = dmax(r, residual_one(LB, A, i, K, mu));
    res = block_max(r, red);
  }
  for (int q = tid; q < nb; q += LDS_THREADS) {
    const int g = body_l[q];
    for (int c = 0; c < 3; ++c) {
      B.v[c][g] = v_l[c][q];
      B.w[c][g] = w_l[c][q];
    }
  }
  if (tid == 0) {
    island_report(S, sweeps, updates, res, res_ref, P.stop_rule == 1 && !(res <= P.tol));
    atomicAdd(&S->lds_manifold_sweeps, updates);
    if (hot_sweeps) atomicAdd(&S->hot_sweeps, hot_sweeps);
  }
}

// One island per workgroup (a fixed step sizes the grid as for k_island_reg).
__global__ void __launch_bounds__(LDS_THREADS) k_island_lds(BodyPtrs B, MfPtrs M, const SysDev* __restrict__ sys,
                                                           const int* __restrict__ istart,
                                                           const uint32_t* __restrict__ skey,
                                                           const int* __restrict__ bseg, const int* __restrict__ sbody,
                                                           const int* __restrict__ list, Count countc, IslandSolve P,
                                                           IslandStats* S, int2* __restrict__ slot,
                                                           int* __restrict__ hlist, const int* __restrict__ guard) {
  if (guard && *guard) return;
  if ((int)blockIdx.x >= countc.get()) return;
  island_lds_block(B, M, sys, istart, skey, bseg, sbody, list, P, S, slot, hlist, blockIdx.x);
}

// Colour starts of listed islands (the huge class), for the host-driven
// colour-by-colour solve: out[k * HUGE_CS + c] = first manifold of colour c,
// c = 0 .. ncol <= 128 (the last is the island end), out[k * HUGE_CS + 129] = ncol.
constexpr int HUGE_CS = 130;
__global__ void k_island_colours(const int* __restrict__ istart, const uint32_t* __restrict__ skey,
                                 const int* __restrict__ list, int nlist, int* __restrict__ out) {
  const int k = blockIdx.x;
  if (k >= nlist) return;
  const int isl = list[k];
  const int s = istart[isl], e = istart[isl + 1];
  const uint32_t base = skey[s] & ~127u;
  const int ncol = (int)(skey[e - 1] & 127u) + 1;
  for (int c = threadIdx.x; c <= ncol; c += blockDim.x) out[k * HUGE_CS + c] = lb_u32(skey, s, e, base | (uint32_t)c);
  if (threadIdx.x == 0) out[k * HUGE_CS + 129] = ncol;
}

// ---------------------------------------------------------------------------
// host side

static MfPtrs mf_ptrs(rdk_ctx* ctx, bool colour_order) {
  MfPtrs M;
  if (colour_order) {
    M.f = ctx->mfc.p;
    M.stride = ctx->mfc.cap / F_COUNT;
    M.a = ctx->mac.p;
    M.b = ctx->mbc.p;
    for (int k = 0; k < 3; ++k) M.lam[k] = ctx->lamc[k].p;
  } else {
    M.f = ctx->mf.p;
    M.stride = ctx->mf.cap / F_COUNT;
    M.a = ctx->ma.p;
    M.b = ctx->mb.p;
    for (int k = 0; k < 3; ++k) M.lam[k] = ctx->lam[k].p;
  }
  return M;
}

static int ensure_manifolds(rdk_ctx* ctx, int m) {
  m = std::max(m, 1);
  const size_t need = (size_t)F_COUNT * m;
  if (ctx->mf.cap < need || ctx->mfc.cap < need) {
    // both buffers must share one stride
    size_t cap = (size_t)(m + m / 4 + 16);
    release(ctx->mf);
    release(ctx->mfc);
    RDK_TRY(ensure(ctx, ctx->mf, F_COUNT * cap));
    RDK_TRY(ensure(ctx, ctx->mfc, F_COUNT * cap));
    // ensure() pads; force an exact multiple of F_COUNT for the stride
    ctx->mf.cap = (ctx->mf.cap / F_COUNT) * F_COUNT;
    ctx->mfc.cap = (ctx->mfc.cap / F_COUNT) * F_COUNT;
    if (ctx->mfc.cap != ctx->mf.cap) ctx->mfc.cap = ctx->mf.cap = std::min(ctx->mf.cap, ctx->mfc.cap);
  }
  RDK_TRY(ensure(ctx, ctx->ma, m));
  RDK_TRY(ensure(ctx, ctx->mb, m));
  RDK_TRY(ensure(ctx, ctx->mac, m));
  RDK_TRY(ensure(ctx, ctx->mbc, m));
  for (int k = 0; k < 3; ++k) {
    RDK_TRY(ensure(ctx, ctx->lam[k], m));
    RDK_TRY(ensure(ctx, ctx->lamc[k], m));
  }
  RDK_TRY(ensure(ctx, ctx->color, m));
  RDK_TRY(ensure(ctx, ctx->color_tmp, m));
  RDK_TRY(ensure(ctx, ctx->perm, m));
  RDK_TRY(ensure(ctx, ctx->perm_alt, m));
  RDK_TRY(ensure(ctx, ctx->color_key, m));
  RDK_TRY(ensure(ctx, ctx->color_key_alt, m));
  RDK_TRY(ensure(ctx, ctx->csr_key, 2 * m));
  RDK_TRY(ensure(ctx, ctx->csr_key_alt, 2 * m));
  RDK_TRY(ensure(ctx, ctx->csr_val, 2 * m));
  RDK_TRY(ensure(ctx, ctx->csr_val_alt, 2 * m));
  return RDK_OK;
}

int nsc_build_run(rdk_ctx* ctx, const rdk_nsc_params* p) {
  const int nc = ctx->nc, n = ctx->n;
  const int nb = 256;
  BodyPtrs B = body_ptrs(ctx);
  // wall candidates
  bool any_wall = false;
  for (auto& s : ctx->sys_host) any_wall |= (s.floor_wall || s.ceil_wall);
  int nwall = 0;
  if (any_wall) {
    RDK_TRY(ensure(ctx, ctx->wall_flag, 4 * n + 1));
    RDK_TRY(ensure(ctx, ctx->wall_off, 4 * n + 1));
    {
      ProfScope ps(ctx, K_WALL);
      hipLaunchKernelGGL(k_wall_flags, dim3(grid_for(4 * n, nb)), dim3(nb), 0, ctx->stream, B, ctx->sys_dev.p, n,
                         ctx->wall_flag.p);
    }
    RDK_TRY(exclusive_scan_int(ctx, ctx->wall_flag.p, ctx->wall_off.p, 4 * n));
    int last[2];
    RDK_HIP(hipMemcpyAsync(&last[0], ctx->wall_off.p + 4 * n - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipMemcpyAsync(&last[1], ctx->wall_flag.p + 4 * n - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    nwall = last[0] + last[1];
  }
  const int nm = nc + nwall;
  RDK_TRY(ensure_manifolds(ctx, nm));
  MfPtrs M = mf_ptrs(ctx, false);
  if (nc > 0) {
    ContactIn C;
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
    const int warm = p->warm_mode != 0 ? 1 : 0;  // the learned start corrects the previous-step one
    ProfScope ps(ctx, K_MANIFOLD);
    hipLaunchKernelGGL(k_manifold, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, B, C, Count::host(nc), M, p->cfm,
                       warm, ctx->prev_keys.p, ctx->nprev, (const int*)nullptr, ctx->prev_lam[0].p, ctx->prev_lam[1].p,
                       ctx->prev_lam[2].p);
  }
  if (nwall > 0) {
    ProfScope ps(ctx, K_WALL);
    const int wwarm = (p->warm_mode != 0 && p->warm_walls) ? 1 : 0;
    hipLaunchKernelGGL(k_wall_manifold, dim3(grid_for(4 * n, nb)), dim3(nb), 0, ctx->stream, B, ctx->sys_dev.p, n,
                       ctx->wall_flag.p, ctx->wall_off.p, INT_MAX, Count::host(nc), M, p->cfm, wwarm,
                       ctx->prev_wkeys.p,
                       ctx->nprev_w, (const int*)nullptr, ctx->prev_wlam[0].p, ctx->prev_wlam[1].p,
                       ctx->prev_wlam[2].p);
  }
  ctx->nm = nm;
  ctx->nm_wall = nwall;
  ctx->nc_seen_max = std::max(ctx->nc_seen_max, nc);
  ctx->nwall_seen_max = std::max(ctx->nwall_seen_max, nwall);
  ctx->have_manifolds = true;
  ctx->have_coloring = false;
  ctx->work[W_MANIFOLDS] += nm;
  if (p->warm_mode >= 2 && nm > 0) RDK_TRY(mlp_warm_start(ctx, p->warm_mode));
  return RDK_OK;
}

// Manifold counts up to this are coloured (and, on the global path, solved)
// by one workgroup: rdk_set_tuning(RDK_TUNE_SMALL_M), else RDK_SMALL_M, else 16384.
static int small_m_limit(const rdk_ctx* ctx) {
  if (ctx->small_m >= 0) return ctx->small_m;
  static int lim = -1;
  if (lim < 0) {
    const char* e = getenv("RDK_SMALL_M");
    lim = e ? atoi(e) : 16384;
  }
  return lim;
}

// Colour the manifolds and build the colour-ordered copy.  need_host: also
// bring the colour offsets to the host (the multi-launch sweep path needs
// them); the single-workgroup path keeps them on the device.
// Colour every manifold (colour[] on the device).
static int color_compute(rdk_ctx* ctx) {
  const int nm = ctx->nm, n = ctx->n;
  const int nb = 256;
  if (nm == 0) return RDK_OK;
  RDK_TRY(ensure(ctx, ctx->body_start, n + 1));
  RDK_TRY(ensure(ctx, ctx->uncolored, 256));
  RDK_TRY(ensure(ctx, ctx->d_cstart, 132));
  {
    ProfScope ps(ctx, K_CSR);
    hipLaunchKernelGGL(k_csr_entries, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->ma.p, ctx->mb.p,
                       Count::host(nm), n, ctx->csr_key.p, ctx->csr_val.p);
  }
  bool alt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->csr_key.p, ctx->csr_key_alt.p, ctx->csr_val.p, ctx->csr_val_alt.p, 2 * nm,
                         bits_for((uint64_t)n), alt));
  const uint32_t* skey = alt ? ctx->csr_key_alt.p : ctx->csr_key.p;
  const int* csr = alt ? ctx->csr_val_alt.p : ctx->csr_val.p;
  {
    ProfScope ps(ctx, K_CSR);
    hipLaunchKernelGGL(k_body_start, dim3(grid_for(n + 1, nb)), dim3(nb), 0, ctx->stream, skey, Count::host(2 * nm),
                       n, ctx->body_start.p);
  }
  unsigned* dctr = (unsigned*)ctx->uncolored.p;  // [0] left, [1] maxcol, [2] overflow, [3] ncolor
  RDK_HIP(hipMemsetAsync(dctr, 0, 4 * sizeof(unsigned), ctx->stream));
  const bool small = nm <= small_m_limit(ctx);
  (void)alt;
  if (small) {
    ProfScope ps(ctx, K_COLOR);
    hipLaunchKernelGGL(k_color_small, dim3(1), dim3(SMALL_BLOCK), 0, ctx->stream, ctx->ma.p, ctx->mb.p,
                       Count::host(nm), ctx->body_start.p, csr, ctx->color.p, ctx->color_tmp.p, dctr, INT_MAX,
                       (const int*)nullptr);
  } else {
    {
      ProfScope ps(ctx, K_COLOR);
      hipLaunchKernelGGL(k_fill_color, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->color.p,
                         Count::host(nm), -1);
    }
    int rounds = 0;
    unsigned hc[4];
    for (;;) {
      RDK_HIP(hipMemsetAsync(dctr, 0, sizeof(unsigned), ctx->stream));
      {
        ProfScope ps(ctx, K_COLOR);
        hipLaunchKernelGGL(k_color_assign, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->ma.p, ctx->mb.p,
                           Count::host(nm), ctx->body_start.p, csr, ctx->color.p, ctx->color_tmp.p, dctr + 2, -1,
                           (const unsigned*)nullptr, (const int*)nullptr);
        hipLaunchKernelGGL(k_color_resolve, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->ma.p, ctx->mb.p,
                           Count::host(nm), ctx->body_start.p, csr, ctx->color_tmp.p, ctx->color.p, dctr, dctr + 1,
                           -1, (const unsigned*)nullptr, (const int*)nullptr);
      }
      ++rounds;
      ctx->color_rounds_max = std::max(ctx->color_rounds_max, rounds);
      RDK_HIP(hipMemcpyAsync(hc, dctr, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
      RDK_HIP(hipStreamSynchronize(ctx->stream));
      if (hc[2] != 0) {
        ctx->err = "nsc colouring: more than 127 colours needed";
        return RDK_ERR_CAPACITY;
      }
      if (hc[0] == 0) break;
      if (rounds > 4096) {
        ctx->err = "nsc colouring did not terminate";
        return RDK_ERR_STATE;
      }
    }
  }
  return RDK_OK;
}

// Group manifolds by colour (stable) and build the colour-ordered copy.
static int color_group(rdk_ctx* ctx, bool need_host) {
  const int nm = ctx->nm;
  const int nb = 256;
  const bool small = nm <= small_m_limit(ctx);
  unsigned* dctr = (unsigned*)ctx->uncolored.p;
  bool alt = false;
  RDK_TRY(ensure(ctx, ctx->wall_off, std::max<size_t>(ctx->wall_off.cap, 128)));
  int* hist = ctx->wall_off.p;  // scratch 128-bin histogram
  RDK_HIP(hipMemsetAsync(hist, 0, 128 * sizeof(int), ctx->stream));
  {
    ProfScope ps(ctx, K_COLOR);
    hipLaunchKernelGGL(k_color_keys, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->color.p, nm,
                       ctx->color_key.p, ctx->perm_alt.p, hist);
  }
  RDK_TRY(sort_pairs_u32(ctx, ctx->color_key.p, ctx->color_key_alt.p, ctx->perm_alt.p, ctx->perm.p, nm, 7, alt));
  if (!alt)
    RDK_HIP(hipMemcpyAsync(ctx->perm.p, ctx->perm_alt.p, sizeof(int) * nm, hipMemcpyDeviceToDevice, ctx->stream));
  {
    ProfScope ps(ctx, K_COLOR);
    hipLaunchKernelGGL(k_color_offsets, dim3(1), dim3(256), 0, ctx->stream, ctx->color.p, nm, dctr + 1,
                       ctx->d_cstart.p, dctr);
  }
  {
    ProfScope ps(ctx, K_PERMUTE);
    hipLaunchKernelGGL(k_permute, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, mf_ptrs(ctx, false),
                       mf_ptrs(ctx, true), ctx->perm.p, Count::host(nm));
  }
  ctx->have_coloring = true;
  ctx->color_host_valid = false;
  if (need_host || !small) RDK_TRY(color_to_host(ctx));
  return RDK_OK;
}

int nsc_color_run(rdk_ctx* ctx, bool need_host) {
  ctx->ncolor = 0;
  ctx->color_start.assign(1, 0);
  ctx->color_host_valid = true;
  if (ctx->nm == 0) {
    ctx->have_coloring = true;
    return RDK_OK;
  }
  RDK_TRY(color_compute(ctx));
  return color_group(ctx, need_host);
}

int color_to_host(rdk_ctx* ctx) {
  if (ctx->color_host_valid) return RDK_OK;
  unsigned hc[4];
  RDK_HIP(hipMemcpyAsync(hc, ctx->uncolored.p, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  if (hc[2] != 0) {
    ctx->err = "nsc colouring: more than 127 colours needed";
    return RDK_ERR_CAPACITY;
  }
  if (hc[0] != 0) {
    ctx->err = "nsc colouring did not terminate";
    return RDK_ERR_STATE;
  }
  const int ncolor = (int)hc[3];
  ctx->color_start.assign(ncolor + 1, 0);
  RDK_HIP(hipMemcpyAsync(ctx->color_start.data(), ctx->d_cstart.p, (ncolor + 1) * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->ncolor = ncolor;
  ctx->color_host_valid = true;
  return RDK_OK;
}

static SolveConst solve_const(rdk_ctx* ctx, const rdk_nsc_params* p) {
  SolveConst K;
  K.omega = p->omega;
  K.cfm = p->cfm;
  K.beta = p->position_stabilization ? p->beta * 0.01 : p->beta;
  K.dt = ctx->phys.dt;
  K.stall_omega = p->stall_omega;
  K.proj = p->friction_proj;
  return K;
}

static int run_sweeps(rdk_ctx* ctx, const rdk_nsc_params* p, int nsweeps) {
  BodyPtrs B = body_ptrs(ctx);
  MfPtrs M = mf_ptrs(ctx, true);
  SolveConst K = solve_const(ctx, p);
  const int nb = 64;  // one-wave workgroups spread a colour over more CUs
  double* ref = ctx->resid.p + 1;
  for (int s = 0; s < nsweeps; ++s) {
    RDK_HIP(hipMemsetAsync(ref, 0, sizeof(double), ctx->stream));
    for (int c = 0; c < ctx->ncolor; ++c) {
      const int i0 = ctx->color_start[c], i1 = ctx->color_start[c + 1];
      if (i1 <= i0) continue;
      EventRec rec;
      if (prof_ext_events(ctx, K_PSOR_SWEEP, rec)) {
        hipExtLaunchKernelGGL(k_psor_sweep, dim3(grid_for(i1 - i0, nb)), dim3(nb), 0, ctx->stream, rec.a, rec.b, 0,
                              B, M, ctx->sys_dev.p, i0, i1, K, ref, (const int*)nullptr);
        prof_ext_done(ctx, rec);
        ctx->work[W_PSOR_TIMED] += i1 - i0;
      } else {
        hipLaunchKernelGGL(k_psor_sweep, dim3(grid_for(i1 - i0, nb)), dim3(nb), 0, ctx->stream, B, M,
                           ctx->sys_dev.p, i0, i1, K, ref, (const int*)nullptr);
      }
    }
  }
  ctx->work[W_SWEEP_CONTACTS] += (int64_t)nsweeps * ctx->nm;
  return RDK_OK;
}

int nsc_residual_run(rdk_ctx* ctx, const rdk_nsc_params* p, double* res, double* res_ref) {
  RDK_TRY(ensure(ctx, ctx->resid, 4));
  const int nm = ctx->nm;
  double h[2] = {0.0, 0.0};
  if (nm > 0) {
    RDK_HIP(hipMemsetAsync(ctx->resid.p, 0, sizeof(double), ctx->stream));
    ProfScope ps(ctx, K_RESIDUAL);
    hipLaunchKernelGGL(k_residual, dim3(grid_for(nm, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx),
                       mf_ptrs(ctx, true), ctx->sys_dev.p, 0, nm, solve_const(ctx, p), ctx->resid.p);
  }
  if (nm > 0) {
    RDK_HIP(hipMemcpyAsync(h, ctx->resid.p, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
  }
  ctx->work[W_RESID_CONTACTS] += nm;
  if (res) *res = h[0];
  if (res_ref) *res_ref = h[1];
  return RDK_OK;
}

int nsc_sweeps_run(rdk_ctx* ctx, const rdk_nsc_params* p, int nsweeps) {
  if (!ctx->have_manifolds) {
    ctx->err = "rdk_nsc_sweeps: no manifolds (call rdk_nsc_build)";
    return RDK_ERR_STATE;
  }
  RDK_TRY(ensure(ctx, ctx->resid, 4));
  if (!ctx->have_coloring) RDK_TRY(nsc_color_run(ctx, true));
  RDK_TRY(color_to_host(ctx));
  RDK_TRY(run_sweeps(ctx, p, nsweeps));
  const int nb = 256;
  if (ctx->nm > 0) {
    ProfScope ps(ctx, K_PERMUTE);
    hipLaunchKernelGGL(k_unpermute_lambda, dim3(grid_for(ctx->nm, nb)), dim3(nb), 0, ctx->stream,
                       mf_ptrs(ctx, true), mf_ptrs(ctx, false), ctx->perm.p, Count::host(ctx->nm));
  }
  return RDK_OK;
}


static int store_warm_cache(rdk_ctx* ctx, const rdk_nsc_params* p) {
  const int nb = 256;
  // warm cache for the next step: rod-rod manifolds are the first nc, sorted by key
  const int nc = ctx->nc;
  if (p->warm_mode != 0 && nc > 0) {
    RDK_TRY(ensure(ctx, ctx->prev_keys, nc));
    for (int k = 0; k < 3; ++k) RDK_TRY(ensure(ctx, ctx->prev_lam[k], nc));
    ProfScope ps(ctx, K_MISC);
    hipLaunchKernelGGL(k_copy_keys_lambda, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, ctx->pkeys.p,
                       mf_ptrs(ctx, false), Count::host(nc), ctx->prev_keys.p, ctx->prev_lam[0].p, ctx->prev_lam[1].p,
                       ctx->prev_lam[2].p, (int*)nullptr, (const int*)nullptr);
    ctx->nprev = nc;
  } else {
    ctx->nprev = 0;
  }
  const int nw = ctx->nm_wall;
  if (p->warm_mode != 0 && p->warm_walls && nw > 0) {
    RDK_TRY(ensure(ctx, ctx->prev_wkeys, nw));
    for (int k = 0; k < 3; ++k) RDK_TRY(ensure(ctx, ctx->prev_wlam[k], nw));
    ProfScope ps(ctx, K_MISC);
    hipLaunchKernelGGL(k_copy_wall_keys_lambda, dim3(grid_for(nw, nb)), dim3(nb), 0, ctx->stream, mf_ptrs(ctx, false),
                       Count::host(nc), Count::host(nw), ctx->prev_wkeys.p, ctx->prev_wlam[0].p, ctx->prev_wlam[1].p,
                       ctx->prev_wlam[2].p, (int*)nullptr, (const int*)nullptr);
    ctx->nprev_w = nw;
  } else {
    ctx->nprev_w = 0;
  }
  return RDK_OK;
}

// Island-parallel solve.  Sets *handled = false (and does nothing else after
// colouring) when some island is too large for one workgroup; the caller then
// runs the global coloured path.
// Huge islands (more bodies than fit in LDS): solved colour by colour with
// one launch per colour spread over the whole GPU, in the order the island
// kernels sweep (the same serial colour-order Gauss-Seidel), with the stall
// rule and residual checks of the island kernels evaluated on the host.  A
// single workgroup on such an island is bound by one CU's latency and L2
// bandwidth; one launch per colour uses as many CUs as the colour has
// manifolds / 256.
struct HugeOut {
  int max_sweeps = 0, nonconv = 0;
  int64_t manifold_sweeps = 0;
  double res = 0.0, res_ref = 0.0;
};

// Device-side state of one huge-island solve.  The check after each block of
// sweeps (sweep count, stop rules, stall rule) runs on the GPU, so several
// blocks go out per host round trip; once `done` is set, the sweeps of blocks
// already launched compute but write nothing.  done: 0 running, 1 finished,
// 2 stalled with omega to lower to stall_omega (the host recaptures).
struct HugeState {
  int sweeps, done, pad0, pad1;
  double res, res_ref;
  StallTrack trk;
};

// The host loop's checks after one block, in its order: rule 0 (largest
// change < 1e-6), the stall rule (only while res > tol), then the sweep budget
// and the tolerance.
__global__ void k_huge_check(HugeState* st, const double* __restrict__ r, int kk, int max_sweeps, int stop_rule,
                             double tol, double omega, double stall_omega) {
  if (threadIdx.x != 0 || st->done != 0) return;
  const int sweeps = st->sweeps + kk;
  const double res = r[0], res_ref = r[1];
  st->sweeps = sweeps;
  st->res = res;
  st->res_ref = res_ref;
  if (stop_rule == 0 && res_ref < 1e-6) {
    st->done = 1;
    return;
  }
  if (stop_rule == 1 && !(res <= tol) && st->trk.stalled(sweeps, res)) {
    st->done = (stall_omega > 0.0 && omega > stall_omega && sweeps < max_sweeps) ? 2 : 1;
    return;
  }
  if (sweeps >= max_sweeps || (stop_rule == 1 && res <= tol)) st->done = 1;
}

__global__ void k_zero_doubles(double* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

static int huge_islands_solve(rdk_ctx* ctx, const IslandSolve& P, const uint32_t* skey, const int* list, int nh,
                              hipStream_t st, HugeOut& out) {
  RDK_TRY(ensure(ctx, ctx->hcs, (size_t)nh * HUGE_CS));
  RDK_TRY(ensure(ctx, ctx->hres, 2));
  RDK_TRY(ensure(ctx, ctx->hstate, (sizeof(HugeState) + sizeof(double) - 1) / sizeof(double)));
  {
    ProfScope ps(ctx, K_ISL_HUGE, st);
    hipLaunchKernelGGL(k_island_colours, dim3(nh), dim3(128), 0, st, ctx->istart.p, skey, list, nh, ctx->hcs.p);
  }
  std::vector<int> hcs((size_t)nh * HUGE_CS);
  RDK_HIP(hipMemcpyAsync(hcs.data(), ctx->hcs.p, hcs.size() * sizeof(int), hipMemcpyDeviceToHost, st));
  RDK_HIP(hipStreamSynchronize(st));
  BodyPtrs B = body_ptrs(ctx);
  MfPtrs Mc = mf_ptrs(ctx, true);
  // one-wave workgroups: a colour's ~1000 manifolds spread over 16 CUs (and all
  // XCDs) instead of 4; measured 4.9 us per colour launch against 7.1 us with 256
  const int nb = 64;
  double* dres = ctx->hres.p;
  double* dref = ctx->hres.p + 1;
  HugeState* dst = (HugeState*)ctx->hstate.p;
  const int* skip = &dst->done;
  const int kk = P.stop_rule == 1 ? P.check_every : 1;
  // A block is kk sweeps, the residual and the device-side check.  Blocks are
  // replayed from captured graphs: one block per replay for the first two
  // replays of a solve, then four (what a solve that ends mid-replay computes
  // after its end is discarded).  The counters are cleared by a kernel: no
  // memset nodes in the graphs.  While the colour launches are timed, one
  // replay in 16 runs eagerly as a single block with its launches timed (a
  // sample over the whole run); timing the whole huge path turns the graphs off.
  const bool graphs = graphs_enabled() && !timed(ctx, K_ISL_HUGE);
  const bool sample = timed(ctx, K_PSOR_SWEEP);
  constexpr int LONG_REPLAY = 4;
  for (int h = 0; h < nh; ++h) {
    const int* cs = hcs.data() + (size_t)h * HUGE_CS;
    const int ncol = cs[129];
    const int s = cs[0], e = cs[ncol];
    auto enqueue_block = [&](const SolveConst& Kc, bool sweep, bool time_sweeps) {
      if (sweep) {
        for (int it = 0; it < kk; ++it) {
          if (it == kk - 1) {
            hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(64), 0, st, ctx->hres.p, 2);
          }
          double* ref = it == kk - 1 ? dref : nullptr;  // only the last sweep's change is reported
          for (int c = 0; c < ncol; ++c) {
            if (cs[c + 1] <= cs[c]) continue;
            EventRec rec;
            if (time_sweeps && prof_ext_events(ctx, K_PSOR_SWEEP, rec)) {
              hipExtLaunchKernelGGL(k_psor_sweep, dim3(grid_for(cs[c + 1] - cs[c], nb)), dim3(nb), 0, st, rec.a,
                                    rec.b, 0, B, Mc, ctx->sys_dev.p, cs[c], cs[c + 1], Kc, ref, skip);
              prof_ext_done(ctx, rec);
              ctx->work[W_PSOR_TIMED] += cs[c + 1] - cs[c];
            } else {
              ProfScope ps(ctx, K_ISL_HUGE, st);
              hipLaunchKernelGGL(k_psor_sweep, dim3(grid_for(cs[c + 1] - cs[c], nb)), dim3(nb), 0, st, B, Mc,
                                 ctx->sys_dev.p, cs[c], cs[c + 1], Kc, ref, skip);
            }
          }
        }
      } else {
        hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(64), 0, st, ctx->hres.p, 2);
      }
      ProfScope ps(ctx, K_ISL_HUGE, st);
      hipLaunchKernelGGL(k_residual, dim3(grid_for(e - s, nb)), dim3(nb), 0, st, B, Mc, ctx->sys_dev.p, s, e, Kc,
                         dres);
      if (sweep)
        hipLaunchKernelGGL(k_huge_check, dim3(1), dim3(64), 0, st, dst, (const double*)ctx->hres.p, kk,
                           P.max_sweeps, P.stop_rule, P.tol, Kc.omega, Kc.stall_omega);
    };
    if (P.warm) {
      for (int c = 0; c < ncol; ++c) {
        if (cs[c + 1] <= cs[c]) continue;
        ProfScope ps(ctx, K_ISL_HUGE, st);
        hipLaunchKernelGGL(k_warm_apply, dim3(grid_for(cs[c + 1] - cs[c], nb)), dim3(nb), 0, st, B, Mc, cs[c],
                           cs[c + 1]);
      }
    }
    HugeState hs{};
    hs.trk.init(P.stall);
    SolveConst K = P.K;  // omega may drop to stall_omega on a stall
    if (P.stop_rule == 1) {
      enqueue_block(K, false, false);
      double r0 = 0.0;
      RDK_HIP(hipMemcpyAsync(&r0, dres, sizeof(double), hipMemcpyDeviceToHost, st));
      RDK_HIP(hipStreamSynchronize(st));
      hs.res = r0;
      ctx->work[W_RESID_CONTACTS] += e - s;
    }
    hs.done = (P.max_sweeps <= 0 || (P.stop_rule == 1 && hs.res <= P.tol)) ? 1 : 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    auto drop_graphs = [&]() {
      for (int g = 0; g < 2; ++g) {
        if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
        if (graph[g]) (void)hipGraphDestroy(graph[g]);
        exec[g] = nullptr;
        graph[g] = nullptr;
      }
    };
    int rc = RDK_OK;
    int replays = 0;
    if (hs.done != 1) RDK_HIP(hipMemcpyAsync(dst, &hs, sizeof(HugeState), hipMemcpyHostToDevice, st));
    while (rc == RDK_OK && hs.done != 1) {
      const bool timed_block = sample && (ctx->huge_blocks++ % 16 == 0);
      if (graphs && !timed_block) {
        const int g = replays < 2 ? 0 : 1;
        const int nblk = g == 0 ? 1 : LONG_REPLAY;
        if (!exec[g]) {
          if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = RDK_ERR_HIP;
            break;
          }
          ctx->capturing = true;  // no allocation, no timed launch inside the capture
          for (int b = 0; b < nblk; ++b) enqueue_block(K, true, false);
          ctx->capturing = false;
          if (hipStreamEndCapture(st, &graph[g]) != hipSuccess ||
              hipGraphInstantiate(&exec[g], graph[g], nullptr, nullptr, 0) != hipSuccess) {
            rc = RDK_ERR_HIP;
            break;
          }
        }
        if (hipGraphLaunch(exec[g], st) != hipSuccess) {
          rc = RDK_ERR_HIP;
          break;
        }
      } else {
        enqueue_block(K, true, timed_block);
      }
      ++replays;
      const int before = hs.sweeps;
      RDK_HIP(hipMemcpyAsync(&hs, dst, sizeof(HugeState), hipMemcpyDeviceToHost, st));
      RDK_HIP(hipStreamSynchronize(st));
      ctx->work[W_RESID_CONTACTS] += (int64_t)((hs.sweeps - before) / kk) * (e - s);
      if (hs.done == 2) {
        K.omega = K.stall_omega;
        drop_graphs();  // the sweeps are recaptured with the new omega
        hs.done = 0;
        RDK_HIP(hipMemcpyAsync(dst, &hs, sizeof(HugeState), hipMemcpyHostToDevice, st));
      }
    }
    drop_graphs();
    if (rc != RDK_OK) {
      ctx->err = "huge-island solve: graph capture or launch failed";
      return rc;
    }
    out.max_sweeps = std::max(out.max_sweeps, hs.sweeps);
    out.manifold_sweeps += (int64_t)hs.sweeps * (e - s);
    out.res = std::max(out.res, hs.res);
    out.res_ref = std::max(out.res_ref, hs.res_ref);
    if (P.stop_rule == 1 && !(hs.res <= P.tol)) out.nonconv++;
  }
  return RDK_OK;
}

static int island_solve(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st, bool* handled, int* status) {
  *handled = false;
  const int nm = ctx->nm, n = ctx->n, nb = 256;
  if (n >= (1 << 25)) return RDK_OK;
  RDK_TRY(color_compute(ctx));
  RDK_TRY(ensure(ctx, ctx->uf_parent, n));
  RDK_TRY(ensure(ctx, ctx->uf_label, n));
  RDK_TRY(ensure(ctx, ctx->blab, n));
  RDK_TRY(ensure(ctx, ctx->blab_alt, n));
  RDK_TRY(ensure(ctx, ctx->sbody, n));
  RDK_TRY(ensure(ctx, ctx->sbody_alt, n));
  RDK_TRY(ensure(ctx, ctx->iflag, nm + 1));
  RDK_TRY(ensure(ctx, ctx->ipos, nm + 1));
  RDK_TRY(ensure(ctx, ctx->istart, nm + 2));
  RDK_TRY(ensure(ctx, ctx->ilists, 6 * (size_t)nm + 6));
  RDK_TRY(ensure(ctx, ctx->bseg, 2 * (size_t)nm + 2));
  RDK_TRY(ensure(ctx, ctx->icnt, 8));
  RDK_TRY(ensure(ctx, ctx->istats, sizeof(IslandStats)));
  {
    ProfScope ps(ctx, K_ISLAND);
    hipLaunchKernelGGL(k_uf_init, dim3(grid_for(n, nb)), dim3(nb), 0, ctx->stream, ctx->uf_parent.p, n);
    hipLaunchKernelGGL(k_uf_link, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->ma.p, ctx->mb.p,
                       Count::host(nm), ctx->uf_parent.p);
    hipLaunchKernelGGL(k_uf_flatten, dim3(grid_for(n, nb)), dim3(nb), 0, ctx->stream, ctx->uf_parent.p, n,
                       ctx->uf_label.p);
    hipLaunchKernelGGL(k_body_label_keys, dim3(grid_for(n, nb)), dim3(nb), 0, ctx->stream, ctx->uf_label.p, n,
                       ctx->blab.p, ctx->sbody.p);
  }
  bool balt = false, malt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->blab.p, ctx->blab_alt.p, ctx->sbody.p, ctx->sbody_alt.p, n,
                         bits_for((uint64_t)n), balt));
  const uint32_t* blab = balt ? ctx->blab_alt.p : ctx->blab.p;
  const int* sbody = balt ? ctx->sbody_alt.p : ctx->sbody.p;
  {
    ProfScope ps(ctx, K_ISLAND);
    hipLaunchKernelGGL(k_island_keys, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->ma.p, ctx->color.p,
                       Count::host(nm), ctx->uf_label.p, ctx->color_key.p, ctx->perm_alt.p);
  }
  RDK_TRY(sort_pairs_u32(ctx, ctx->color_key.p, ctx->color_key_alt.p, ctx->perm_alt.p, ctx->perm.p, nm,
                         7 + bits_for((uint64_t)n), malt));
  if (!malt)
    RDK_HIP(hipMemcpyAsync(ctx->perm.p, ctx->perm_alt.p, sizeof(int) * nm, hipMemcpyDeviceToDevice, ctx->stream));
  const uint32_t* skey = malt ? ctx->color_key_alt.p : ctx->color_key.p;
  {
    ProfScope ps(ctx, K_PERMUTE);
    hipLaunchKernelGGL(k_permute, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, mf_ptrs(ctx, false),
                       mf_ptrs(ctx, true), ctx->perm.p, Count::host(nm));
  }
  {
    ProfScope ps(ctx, K_ISLAND);
    hipLaunchKernelGGL(k_island_flags, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, skey, Count::host(nm),
                       ctx->iflag.p);
  }
  RDK_TRY(exclusive_scan_int(ctx, ctx->iflag.p, ctx->ipos.p, nm));
  {
    ProfScope ps(ctx, K_ISLAND);
    hipLaunchKernelGGL(k_island_starts, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, ctx->iflag.p, ctx->ipos.p,
                       Count::host(nm), ctx->istart.p, (int*)nullptr);
  }
  int last[2];
  RDK_HIP(hipMemcpyAsync(last, ctx->ipos.p + nm - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipMemcpyAsync(last + 1, ctx->iflag.p + nm - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipMemsetAsync(ctx->icnt.p, 0, 8 * sizeof(int), ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  const int nisl = last[0] + last[1];
  IslandClass C;
  C.max_m[0] = ISL_G0;
  C.max_m[1] = ISL_G1;
  C.max_m[2] = ISL_G2;
  C.max_b[0] = ISL_G0 + 8;
  C.max_b[1] = ISL_G1 + 8;
  C.max_b[2] = ISL_G2 + 8;
  C.large_max = 8192;
  C.lds_bmax = LDS_BCAP;
  {
    ProfScope ps(ctx, K_ISLAND);
    hipLaunchKernelGGL(k_island_classify, dim3(grid_for(nisl, nb)), dim3(nb), 0, ctx->stream, ctx->istart.p,
                       Count::host(nisl), skey, blab, n, C, ctx->ilists.p, nisl, ctx->icnt.p, ctx->bseg.p);
  }
  int cnt[6];
  RDK_HIP(hipMemcpyAsync(cnt, ctx->icnt.p, 6 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->nislands = nisl;
  for (int c = 0; c < 4; ++c) ctx->cls_seen_max[c] = std::max(ctx->cls_seen_max[c], cnt[c]);
  if (cnt[5] > 0) return RDK_OK;  // a giant island: global coloured path
  IslandSolve P;
  P.K = solve_const(ctx, p);
  P.tol = p->tol;
  P.max_sweeps = p->max_sweeps;
  P.check_every = std::max(1, p->check_every);
  P.stop_rule = p->stop_rule;
  P.warm = p->warm_mode != 0 ? 1 : 0;
  P.stall = p->stall_sweeps;
  P.hot = p->hot_set;
  IslandStats* S = (IslandStats*)ctx->istats.p;
  RDK_HIP(hipMemsetAsync(S, 0, sizeof(IslandStats), ctx->stream));
  BodyPtrs B = body_ptrs(ctx);
  MfPtrs Mc = mf_ptrs(ctx, true);
  // The classes run concurrently (islands share no body): tiny islands and
  // the host-driven solve of huge islands on stream2, small on stream3, medium
  // on stream4, large (LDS) on the main stream, joined before the impulses
  // are unpermuted.
  const bool side_used = cnt[0] > 0 || cnt[4] > 0;
  const int used = (side_used ? 1 : 0) + (cnt[1] > 0) + (cnt[2] > 0) + (cnt[3] > 0);
  const bool multi = used > 1 && ctx->stream2 && ctx->stream3 && ctx->stream4;
  hipStream_t s_side = multi && side_used ? ctx->stream2 : ctx->stream;
  hipStream_t s_small = multi && cnt[1] > 0 ? ctx->stream3 : ctx->stream;
  hipStream_t s_med = multi && cnt[2] > 0 ? ctx->stream4 : ctx->stream;
  if (multi) {
    RDK_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    for (hipStream_t sx : {s_side, s_small, s_med})
      if (sx != ctx->stream) RDK_HIP(hipStreamWaitEvent(sx, ctx->ev_fork, 0));
  }
  // The LDS class first: its workgroup needs all of one CU's LDS, so it is
  // dispatched before the register classes spread over every CU (launched
  // after them it could wait for a CU to drain).  The classes hold disjoint
  // islands, so the order does not change any result.
  if (cnt[3] > 0) {
    RDK_TRY(ensure(ctx, ctx->islot, nm));
    RDK_TRY(ensure(ctx, ctx->ihot, nm));
    ProfScope ps(ctx, K_ISL_LARGE);
    hipLaunchKernelGGL(k_island_lds, dim3(cnt[3]), dim3(LDS_THREADS), 0, ctx->stream, B, Mc, ctx->sys_dev.p,
                       ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p + 3 * nisl, Count::host(cnt[3]), P, S,
                       ctx->islot.p, ctx->ihot.p, (const int*)nullptr);
  }
  if (cnt[0] > 0) {
    ProfScope ps(ctx, K_ISL_TINY, s_side);
    constexpr int ipb = ISL_WG0 / ISL_G0;
    hipLaunchKernelGGL((k_island_reg<ISL_G0, ISL_WG0>), dim3(grid_for(cnt[0], ipb)), dim3(ISL_WG0), 0, s_side, B, Mc,
                       ctx->sys_dev.p, ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p, Count::host(cnt[0]), P, S,
                       (const int*)nullptr);
  }
  if (cnt[1] > 0) {
    ProfScope ps(ctx, K_ISL_SMALL, s_small);
    hipLaunchKernelGGL((k_island_reg<ISL_G1, ISL_WG1>), dim3(cnt[1]), dim3(ISL_WG1), 0, s_small, B, Mc,
                       ctx->sys_dev.p, ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p + nisl,
                       Count::host(cnt[1]), P, S, (const int*)nullptr);
  }
  if (cnt[2] > 0) {
    ProfScope ps(ctx, K_ISL_MEDIUM, s_med);
    hipLaunchKernelGGL((k_island_reg<ISL_G2, ISL_WG2>), dim3(cnt[2]), dim3(ISL_WG2), 0, s_med, B, Mc,
                       ctx->sys_dev.p, ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p + 2 * nisl,
                       Count::host(cnt[2]), P, S, (const int*)nullptr);
  }
  HugeOut hu;
  if (cnt[4] > 0) RDK_TRY(huge_islands_solve(ctx, P, skey, ctx->ilists.p + 4 * nisl, cnt[4], s_side, hu));
  if (multi) {
    const std::pair<hipStream_t, hipEvent_t> joins[3] = {
        {s_side, ctx->ev_join}, {s_small, ctx->ev_join3}, {s_med, ctx->ev_join4}};
    for (const auto& j : joins) {
      if (j.first == ctx->stream) continue;
      RDK_HIP(hipEventRecord(j.second, j.first));
      RDK_HIP(hipStreamWaitEvent(ctx->stream, j.second, 0));
    }
  }
  {
    ProfScope ps(ctx, K_PERMUTE);
    hipLaunchKernelGGL(k_unpermute_lambda, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, Mc, mf_ptrs(ctx, false),
                       ctx->perm.p, Count::host(nm));
  }
  IslandStats hs;
  unsigned hc[4];
  RDK_HIP(hipMemcpyAsync(&hs, S, sizeof(IslandStats), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipMemcpyAsync(hc, ctx->uncolored.p, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  if (hc[2] != 0) {
    ctx->err = "nsc colouring: more than 127 colours needed";
    return RDK_ERR_CAPACITY;
  }
  ctx->ncolor = (int)hc[1] + 1;
  ctx->have_coloring = false;  // the solve-ordered copy is island-ordered
  for (int c = 0; c < 3; ++c) ctx->island_counts[c] = cnt[c];
  ctx->island_counts[3] = cnt[3] + cnt[4];
  ctx->work[W_SWEEP_CONTACTS] += (int64_t)hs.manifold_sweeps + hu.manifold_sweeps;
  ctx->work[W_LDS_SWEEPS] += (int64_t)hs.lds_manifold_sweeps;
  ctx->work[W_HOT_SWEEPS] += (int64_t)hs.hot_sweeps;
  *status = hs.nonconv + hu.nonconv > 0 ? RDK_ERR_NOCONV : RDK_OK;
  if (st) {
    st->sweeps = std::max(hs.max_sweeps, hu.max_sweeps);
    const double rs = hs.res, rh = hu.res;
    st->residual = (rs != rs || rh != rh) ? (rs != rs ? rs : rh) : std::max(rs, rh);
    st->residual_ref = std::max(hs.res_ref, hu.res_ref);
  }
  *handled = true;
  return RDK_OK;
}

static int global_solve(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st, bool colored);

int nsc_solve_run(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st) {
  if (!ctx->have_manifolds) {
    ctx->err = "rdk_nsc_solve: no manifolds (call rdk_nsc_build)";
    return RDK_ERR_STATE;
  }
  if (ctx->nm > 0 && ctx->solver_mode != 1) {
    bool handled = false;
    int status = RDK_OK;
    rdk_solve_stats tmp;
    std::memset(&tmp, 0, sizeof(tmp));
    RDK_TRY(island_solve(ctx, p, &tmp, &handled, &status));
    if (handled) {
      RDK_TRY(store_warm_cache(ctx, p));
      if (st) {
        *st = tmp;
        st->manifolds = ctx->nm;
        st->rod_contacts = ctx->nc;
        st->wall_contacts = ctx->nm_wall;
        st->colors = ctx->ncolor;
      }
      return status;
    }
    RDK_TRY(color_group(ctx, false));
    return global_solve(ctx, p, st, true);
  }
  return global_solve(ctx, p, st, false);
}

// Global coloured solve (all manifolds; used when an island is too large for
// one workgroup, or when forced).
static int global_solve(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st, bool colored) {
  RDK_TRY(ensure(ctx, ctx->resid, 4));
  const int nm = ctx->nm, nb = 256;
  const bool small = nm > 0 && nm <= small_m_limit(ctx);
  if (!colored) RDK_TRY(nsc_color_run(ctx, false));
  BodyPtrs B = body_ptrs(ctx);
  MfPtrs Mc = mf_ptrs(ctx, true);
  int sweeps = 0;
  double res = 0.0, res_ref = 0.0;
  int status = RDK_OK;
  if (small) {
    RDK_TRY(ensure(ctx, ctx->small_out, sizeof(SmallOut)));
    {
      ProfScope ps(ctx, K_PSOR_SMALL);
      hipLaunchKernelGGL(k_psor_small, dim3(1), dim3(SMALL_BLOCK), 0, ctx->stream, B, Mc, ctx->sys_dev.p,
                         ctx->d_cstart.p, (const unsigned*)ctx->uncolored.p, nm, solve_const(ctx, p),
                         p->warm_mode != 0 ? 1 : 0, p->tol, p->max_sweeps, std::max(1, p->check_every), p->stop_rule,
                         p->stall_sweeps, (SmallOut*)ctx->small_out.p);
    }
    SmallOut so;
    unsigned hc[4];
    RDK_HIP(hipMemcpyAsync(&so, ctx->small_out.p, sizeof(SmallOut), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipMemcpyAsync(hc, ctx->uncolored.p, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    if (hc[2] != 0) {
      ctx->err = "nsc colouring: more than 127 colours needed";
      return RDK_ERR_CAPACITY;
    }
    if (hc[0] != 0) {
      ctx->err = "nsc colouring did not terminate";
      return RDK_ERR_STATE;
    }
    ctx->ncolor = so.ncolor;
    sweeps = so.sweeps;
    res = so.res;
    res_ref = so.res_ref;
    ctx->work[W_SWEEP_CONTACTS] += (int64_t)sweeps * nm;
    ctx->work[W_RESID_CONTACTS] += (int64_t)nm * (1 + (p->stop_rule == 1 ? sweeps / std::max(1, p->check_every) : sweeps));
    if (p->stop_rule == 1 && !(res <= p->tol)) status = RDK_ERR_NOCONV;
    ProfScope ps(ctx, K_PERMUTE);
    hipLaunchKernelGGL(k_unpermute_lambda, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, Mc, mf_ptrs(ctx, false),
                       ctx->perm.p, Count::host(nm));
  } else if (nm > 0) {
    RDK_TRY(color_to_host(ctx));
  }
  if (!small && p->warm_mode != 0) {
    for (int c = 0; c < ctx->ncolor; ++c) {
      const int i0 = ctx->color_start[c], i1 = ctx->color_start[c + 1];
      if (i1 <= i0) continue;
      ProfScope ps(ctx, K_WARM_APPLY);
      hipLaunchKernelGGL(k_warm_apply, dim3(grid_for(i1 - i0, nb)), dim3(nb), 0, ctx->stream, B, Mc, i0, i1);
    }
  }
  if (!small && nm > 0) {
    if (p->stop_rule == 0) {
      for (int it = 0; it < p->max_sweeps; ++it) {
        RDK_TRY(run_sweeps(ctx, p, 1));
        ++sweeps;
        RDK_TRY(nsc_residual_run(ctx, p, &res, &res_ref));
        if (res_ref < 1e-6) break;
      }
    } else {
      const int k = std::max(1, p->check_every);
      StallTrack trk;
      trk.init(p->stall_sweeps);
      rdk_nsc_params q = *p;  // omega may drop to stall_omega on a stall
      RDK_TRY(nsc_residual_run(ctx, p, &res, &res_ref));
      while (!(res <= p->tol) && sweeps < p->max_sweeps) {
        RDK_TRY(run_sweeps(ctx, &q, k));
        sweeps += k;
        RDK_TRY(nsc_residual_run(ctx, p, &res, &res_ref));
        if (!(res <= p->tol) && trk.stalled(sweeps, res)) {
          if (p->stall_omega > 0.0 && q.omega > p->stall_omega) q.omega = p->stall_omega;
          else break;
        }
      }
      if (!(res <= p->tol)) status = RDK_ERR_NOCONV;
    }
    ProfScope ps(ctx, K_PERMUTE);
    hipLaunchKernelGGL(k_unpermute_lambda, dim3(grid_for(nm, nb)), dim3(nb), 0, ctx->stream, Mc, mf_ptrs(ctx, false),
                       ctx->perm.p, Count::host(nm));
  }
  RDK_TRY(store_warm_cache(ctx, p));
  if (st) {
    st->manifolds = nm;
    st->rod_contacts = ctx->nc;
    st->wall_contacts = ctx->nm_wall;
    st->colors = ctx->ncolor;
    st->sweeps = sweeps;
    st->residual = res;
    st->residual_ref = res_ref;
  }
  return status;
}

// ---- position stabilisation (nsc_contact.cpp:659-880, multi-body branch) ----
// Normal-only PSOR on position corrections.  The pairs are the contact list of
// a detection at the current positions (cached for all passes); every pass
// recomputes the narrow phase of each cached pair, keeps those penetrating
// deeper than slop, sweeps up to position_psor_iters times (colour order; a
// coloured sweep equals the serial sweep in colour order) on per-body dx, dtheta
// accumulators, then applies x += dx, q = normalize(angleAxis(|dth|, dth/|dth|) q)
// and wraps.  The manifold buffers are reused for the position manifolds.

struct PosPairs {
  const int *a, *b;
  const double* hs;
};

constexpr int PF_COUNT = 17;  // n 3, ra 3, rb 3, IAn 3, IBn 3, phi + slop, g_n

__global__ void k_pos_pairs(BodyPtrs B, const SysDev* __restrict__ sys, PosPairs P, int np, double slop, double cfm,
                            int* __restrict__ flag, double* __restrict__ tmp, size_t stride) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int a = P.a[i], b = P.b[i];
  const SysDev& S = sys[B.sys[a]];
  const d3 xa = ld3(B.x, a), xb = ld3(B.x, b);
  const d3 ua = axis_of(ldq(B, a)), ub = axis_of(ldq(B, b));
  NarrowOut no;
  capsule_narrow(xa, ua, B.h[a], B.r[a], xb, ub, B.h[b], B.r[b], S.span, S.periodic, HUGE_VAL, true, no);
  const double phi = no.dist - P.hs[i];
  if (!(phi < -slop)) {
    flag[i] = 0;
    return;
  }
  flag[i] = 1;
  const d3 n = no.n;
  const d3 ra = no.pa - xa;
  const d3 rb = no.pb - (xb + no.shift);
  const d3 rAxn = cross(ra, n), rBxn = cross(rb, n);
  const d3 IA = iw_inv(B, a, ua, rAxn), IB = iw_inv(B, b, ub, rBxn);
  const double vals[PF_COUNT] = {n.x,  n.y,  n.z,  ra.x, ra.y, ra.z, rb.x, rb.y,       rb.z,
                                 IA.x, IA.y, IA.z, IB.x, IB.y, IB.z, phi + slop,
                                 diag_term(B.inv_mass[a], B.inv_mass[b], rAxn, IA, rBxn, IB, cfm, false)};
  for (int f = 0; f < PF_COUNT; ++f) tmp[(size_t)f * stride + i] = vals[f];
}

__global__ void k_pos_compact(BodyPtrs B, PosPairs P, const int* __restrict__ flag, const int* __restrict__ pos, int np,
                              const double* __restrict__ tmp, size_t stride, MfPtrs M) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np || !flag[i]) return;
  const int j = pos[i];
  auto t = [&](int f) { return tmp[(size_t)f * stride + i]; };
  M.put3(F_N, j, mk3(t(0), t(1), t(2)));
  M.put3(F_RA, j, mk3(t(3), t(4), t(5)));
  M.put3(F_RB, j, mk3(t(6), t(7), t(8)));
  M.put3(F_IAN, j, mk3(t(9), t(10), t(11)));
  M.put3(F_IBN, j, mk3(t(12), t(13), t(14)));
  M.at(F_PHI, j) = t(15);
  M.at(F_GN, j) = t(16);
  M.at(F_IMA, j) = B.inv_mass[P.a[i]];
  M.at(F_IMB, j) = B.inv_mass[P.b[i]];
  M.a[j] = P.a[i];
  M.b[j] = P.b[i];
  M.lam[0][j] = 0.0;
}

struct PosAcc {
  double *dx[3], *dth[3];
};

// one normal-only position update (nsc_contact.cpp:828-851)
__device__ __forceinline__ double pos_update(const PosAcc& D, const MfPtrs& M, int i, double omega, double cfm) {
  const int a = M.a[i], b = M.b[i];
  const d3 n = M.get3(F_N, i), ra = M.get3(F_RA, i), rb = M.get3(F_RB, i);
  const d3 dxa = ld3(D.dx, a), dxb = ld3(D.dx, b), tha = ld3(D.dth, a), thb = ld3(D.dth, b);
  const double corr = dot(n, ((dxb + cross(thb, rb)) - dxa) - cross(tha, ra));
  const double lam = M.lam[0][i];
  const double res = corr + M.at(F_PHI, i) + cfm * lam;
  const double delta = -(omega / M.at(F_GN, i)) * res;
  const double nl = dmax(0.0, lam + delta);
  const double dl = nl - lam;
  M.lam[0][i] = nl;
  if (dl != 0.0) {
    st3(D.dx, a, dxa - n * (dl * M.at(F_IMA, i)));
    st3(D.dth, a, tha - M.get3(F_IAN, i) * dl);
    st3(D.dx, b, dxb + n * (dl * M.at(F_IMB, i)));
    st3(D.dth, b, thb + M.get3(F_IBN, i) * dl);
  }
  return fabs(dl);
}

// all inner sweeps in one workgroup (colour starts on the device)
__global__ void __launch_bounds__(SMALL_BLOCK) k_pos_small(PosAcc D, MfPtrs M, const int* __restrict__ cstart,
                                                           const unsigned* __restrict__ ctr, double omega, double cfm,
                                                           int iters, int* __restrict__ sweeps_out) {
  __shared__ double red[SMALL_BLOCK / 64];
  __shared__ int cs[129];
  const int ncolor = (int)ctr[3];
  for (int c = threadIdx.x; c <= ncolor; c += blockDim.x) cs[c] = cstart[c];
  __syncthreads();
  int it = 0;
  for (; it < iters; ++it) {
    double d = 0.0;
    for (int c = 0; c < ncolor; ++c) {
      for (int i = cs[c] + threadIdx.x; i < cs[c + 1]; i += blockDim.x) d = dmax(d, pos_update(D, M, i, omega, cfm));
      __syncthreads();
    }
    if (block_max(d, red) < 1e-8) {
      ++it;
      break;
    }
  }
  if (threadIdx.x == 0) *sweeps_out = it;
}

__global__ void k_pos_sweep(PosAcc D, MfPtrs M, int i0, int i1, double omega, double cfm, double* __restrict__ dmax_out) {
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < i1) d = pos_update(D, M, i, omega, cfm);
  for (int off = 32; off > 0; off >>= 1) d = dmax(d, __shfl_down(d, off));
  if (__lane_id() == 0 && d > 0.0) atomic_max_pos(dmax_out, d);
}

// x += dx; q = normalize(angleAxis(|dth|, dth / |dth|) q) (nsc_contact.cpp:866-877); wrap
__global__ void k_pos_apply(BodyPtrs B, const SysDev* __restrict__ sys, int n, PosAcc D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(B.inv_mass[i] > 0.0)) return;
  const SysDev& S = sys[B.sys[i]];
  d3 x = ld3(B.x, i) + ld3(D.dx, i);
  const d3 th = ld3(D.dth, i);
  const double angle = len(th);
  if (angle > 1e-8) {
    const d3 ax = mk3(th.x / angle, th.y / angle, th.z / angle);
    const double sh = sin(angle * 0.5);
    const q4 dq = q4{cos(angle * 0.5), ax.x * sh, ax.y * sh, ax.z * sh};
    const q4 q = qnormalize(qmul(dq, ldq(B, i)));
    B.q[0][i] = q.w;
    B.q[1][i] = q.x;
    B.q[2][i] = q.y;
    B.q[3][i] = q.z;
  }
  double* xp[3] = {&x.x, &x.y, &x.z};
  for (int k = 0; k < 3; ++k) {
    if (!S.periodic[k] || !(S.span[k] > 0.0)) continue;
    int g = 0;
    while (*xp[k] < S.lo[k] && g++ < 8) *xp[k] += S.span[k];
    g = 0;
    while (*xp[k] >= S.hi[k] && g++ < 8) *xp[k] -= S.span[k];
  }
  st3(B.x, i, x);
}

int nsc_project_run(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* p, int* passes_out,
                    int* npos_out) {
  const int n = ctx->n, nb = 256;
  if (passes_out) *passes_out = 0;
  if (npos_out) *npos_out = 0;
  if (p->position_iters <= 0) return RDK_OK;
  rdk_detect_params d = *dp;
  d.delta = 0.0;  // the NSC detector admits contacts at strict overlap (nsc_contact.cpp:24)
  RDK_TRY(detect_run(ctx, &d));
  const int np = ctx->nc;
  if (np == 0) return RDK_OK;
  // cached pairs (sorted (a,b), every contact of the detection)
  RDK_TRY(ensure(ctx, ctx->ppa, np));
  RDK_TRY(ensure(ctx, ctx->ppb, np));
  RDK_TRY(ensure(ctx, ctx->pphs, np));
  RDK_TRY(ensure(ctx, ctx->pflag, np + 1));
  RDK_TRY(ensure(ctx, ctx->ppos, np + 1));
  RDK_TRY(ensure(ctx, ctx->ptmp, (size_t)PF_COUNT * np));
  for (int k = 0; k < 3; ++k) {
    RDK_TRY(ensure(ctx, ctx->pdx[k], n));
    RDK_TRY(ensure(ctx, ctx->pdth[k], n));
  }
  RDK_HIP(hipMemcpyAsync(ctx->ppa.p, ctx->ca.p, np * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
  RDK_HIP(hipMemcpyAsync(ctx->ppb.p, ctx->cb.p, np * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
  RDK_HIP(hipMemcpyAsync(ctx->pphs.p, ctx->chs.p, np * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
  PosPairs P{ctx->ppa.p, ctx->ppb.p, ctx->pphs.p};
  PosAcc D;
  for (int k = 0; k < 3; ++k) {
    D.dx[k] = ctx->pdx[k].p;
    D.dth[k] = ctx->pdth[k].p;
  }
  BodyPtrs B = body_ptrs(ctx);
  const size_t stride = (size_t)np;
  int passes = 0;
  for (int outer = 0; outer < p->position_iters; ++outer) {
    {
      ProfScope ps(ctx, K_POSPROJ);
      hipLaunchKernelGGL(k_pos_pairs, dim3(grid_for(np, nb)), dim3(nb), 0, ctx->stream, B, ctx->sys_dev.p, P, np,
                         p->slop, p->cfm, ctx->pflag.p, ctx->ptmp.p, stride);
    }
    RDK_TRY(exclusive_scan_int(ctx, ctx->pflag.p, ctx->ppos.p, np));
    int last[2];
    RDK_HIP(hipMemcpyAsync(&last[0], ctx->ppos.p + np - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipMemcpyAsync(&last[1], ctx->pflag.p + np - 1, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    const int m = last[0] + last[1];
    if (outer == 0 && npos_out) *npos_out = m;
    if (m == 0) break;
    ++passes;
    RDK_TRY(ensure_manifolds(ctx, m));
    {
      ProfScope ps(ctx, K_POSPROJ);
      hipLaunchKernelGGL(k_pos_compact, dim3(grid_for(np, nb)), dim3(nb), 0, ctx->stream, B, P, ctx->pflag.p,
                         ctx->ppos.p, np, ctx->ptmp.p, stride, mf_ptrs(ctx, false));
    }
    ctx->nm = m;
    ctx->nm_wall = 0;
    RDK_TRY(nsc_color_run(ctx, false));
    for (int k = 0; k < 3; ++k) {
      RDK_HIP(hipMemsetAsync(D.dx[k], 0, n * sizeof(double), ctx->stream));
      RDK_HIP(hipMemsetAsync(D.dth[k], 0, n * sizeof(double), ctx->stream));
    }
    MfPtrs Mc = mf_ptrs(ctx, true);
    if (m <= small_m_limit(ctx)) {
      RDK_TRY(ensure(ctx, ctx->small_out, sizeof(int)));
      ProfScope ps(ctx, K_POSPROJ);
      hipLaunchKernelGGL(k_pos_small, dim3(1), dim3(SMALL_BLOCK), 0, ctx->stream, D, Mc, ctx->d_cstart.p,
                         (const unsigned*)ctx->uncolored.p, p->omega, p->cfm, p->position_psor_iters,
                         (int*)ctx->small_out.p);
    } else {
      RDK_TRY(color_to_host(ctx));
      RDK_TRY(ensure(ctx, ctx->resid, 4));
      for (int it = 0; it < p->position_psor_iters; ++it) {
        RDK_HIP(hipMemsetAsync(ctx->resid.p + 2, 0, sizeof(double), ctx->stream));
        for (int c = 0; c < ctx->ncolor; ++c) {
          const int i0 = ctx->color_start[c], i1 = ctx->color_start[c + 1];
          if (i1 <= i0) continue;
          ProfScope ps(ctx, K_POSPROJ);
          hipLaunchKernelGGL(k_pos_sweep, dim3(grid_for(i1 - i0, nb)), dim3(nb), 0, ctx->stream, D, Mc, i0, i1,
                             p->omega, p->cfm, ctx->resid.p + 2);
        }
        double dm = 0.0;
        RDK_HIP(hipMemcpyAsync(&dm, ctx->resid.p + 2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        RDK_HIP(hipStreamSynchronize(ctx->stream));
        if (dm < 1e-8) break;
      }
    }
    {
      ProfScope ps(ctx, K_POSPROJ);
      hipLaunchKernelGGL(k_pos_apply, dim3(grid_for(n, nb)), dim3(nb), 0, ctx->stream, B, ctx->sys_dev.p, n, D);
    }
  }
  // the manifold buffers now hold position manifolds; the contact list is the
  // pre-projection detection
  ctx->have_manifolds = false;
  ctx->have_coloring = false;
  if (passes_out) *passes_out = passes;
  return RDK_OK;
}

// ---- fixed-capacity (capturable) step ----------------------------------------
// The launches of the free-flight update, nsc_build_run, island_solve and
// store_warm_cache with every count left on the device: grids are sized for
// the reserved capacities and each kernel reads its count when it runs.  The
// colouring, the islands and every check that can fail depend only on
// positions and the manifolds' body pairs, so they run first and a failed step
// has changed no state when the free-flight update is reached.  Colouring runs the
// single-workgroup kernel (at most small_m_limit manifolds) or a fixed number
// of rounds that return at once after the last needed one; the island classes
// that fit one workgroup (tiny, small, medium, LDS) run as persistent launches
// over their device-side island lists.  What the fixed path cannot do (more
// contacts than reserved, more rounds or colours than enqueued, an island too
// large for one workgroup) sets fx[FX_ERR]; the caller restores the state and
// runs that step eagerly.

__global__ void k_fx_zero(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Per-step device counts after detection: wall manifolds and manifolds.
__device__ __forceinline__ void k_fx_counts_body(const unsigned* __restrict__ counters, int pair_cap,
                                                 const int* __restrict__ wall_off, const int* __restrict__ wall_flag,
                                                 int n4, int wall_cap, int* __restrict__ fx) {
  // more contacts than reserved: the contact list was cut
  if (counters[0] > (unsigned)pair_cap) atomicCAS(&fx[FX_ERR], 0, (int)FXE_PAIRS);
  const int nc = (int)min(counters[0], (unsigned)pair_cap);
  int nw = n4 > 0 ? wall_off[n4 - 1] + wall_flag[n4 - 1] : 0;
  if (nw > wall_cap) {
    atomicCAS(&fx[FX_ERR], 0, (int)FXE_WALLS);
    nw = wall_cap;
  }
  fx[FX_NWALL] = nw;
  fx[FX_NM] = nc + nw;
}

// k_uf_init, and (thread 0) the colouring check of a fixed step.
__global__ void k_fx_uf_init_check(int* __restrict__ parent, int n, const unsigned* __restrict__ dctr,
                                   int* __restrict__ fx, int small_limit, int rounds) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && !fx[FX_ERR]) {
    const int nm = fx[FX_NM];
    if (dctr[2] != 0u) atomicCAS(&fx[FX_ERR], 0, (int)FXE_COLOURS);
    else if (nm <= small_limit ? dctr[0] != 0u : (rounds <= 0 || (unsigned)fx[FX_ROUND0 + rounds - 1] != 0u))
      atomicCAS(&fx[FX_ERR], 0, (int)FXE_ROUNDS);
  }
  if (i < n) parent[i] = i;
}

// An island the fixed step cannot solve (more bodies than the LDS class holds,
// more than 8192 manifolds), or more islands in a class than its grid holds.
__global__ void k_fx_class_check(const int* __restrict__ cnt, int c0, int c1, int c2, int c3, int* __restrict__ fx) {
  if (threadIdx.x != 0) return;
  if (cnt[4] + cnt[5] > 0 || cnt[0] > c0 || cnt[1] > c1 || cnt[2] > c2 || cnt[3] > c3)
    atomicCAS(&fx[FX_ERR], 0, (int)FXE_ISLANDS);
}

__global__ void k_fx_copy_int(const int* __restrict__ src, int* __restrict__ dst, Count nc) {
  const int n = nc.get();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Statistics of one completed step (the eager step's rdk_step_stats and work
// counters), added to the accumulators; FX_STEPS counts completed steps.
__global__ void k_fx_accumulate(const int* __restrict__ fxc, int* __restrict__ fx, const unsigned* __restrict__ counters,
                                int pair_cap, const IslandStats* __restrict__ S, const unsigned* __restrict__ dctr,
                                const int* __restrict__ cnt, double* __restrict__ acc) {
  if (threadIdx.x != 0 || fxc[FX_ERR]) return;
  for (int c = 0; c < 4; ++c) acc[FA_CLS0 + c] = fmax(acc[FA_CLS0 + c], (double)cnt[c]);
  const int nc = (int)min(counters[0], (unsigned)pair_cap);
  const int nm = fx[FX_NM];
  const unsigned long long cand = ((unsigned long long)counters[3] << 32) | counters[2];
  const unsigned long long narrow = ((unsigned long long)counters[5] << 32) | counters[4];
  fx[FX_STEPS] += 1;
  acc[FA_STEPS] += 1.0;
  acc[FA_CONTACTS] += (double)nc;
  acc[FA_MANIFOLDS] += (double)nm;
  acc[FA_WALLS] += (double)fx[FX_NWALL];
  acc[FA_CANDIDATES] += (double)cand;
  acc[FA_NARROW] += (double)narrow;
  if (nm > 0) {
    acc[FA_SWEEPS] += (double)S->max_sweeps;
    acc[FA_COLORS] += (double)(dctr[1] + 1u);
    if (S->nonconv > 0) acc[FA_NONCONV] += 1.0;
    acc[FA_MSWEEPS] += (double)S->manifold_sweeps;
    acc[FA_LDS_SWEEPS] += (double)S->lds_manifold_sweeps;
    acc[FA_HOT_SWEEPS] += (double)S->hot_sweeps;
    const double r = S->res;  // NaN residuals are not folded in (the eager step's std::max)
    if (r > acc[FA_RESMAX]) acc[FA_RESMAX] = r;
  }
}

// The body pairs of the manifolds alone (rod-rod from the contact list, walls
// from the slot offsets): colouring and islands need nothing else, so a fixed
// step decides whether it can finish before any state changes.
// Threads [0, pair_cap): rod-rod manifolds from the contact list; threads
// [pair_cap, pair_cap + 4 nbody): wall slots (walls = false: none).  Thread 0
// also stores the step's counts (k_fx_counts).
__global__ void k_fx_pairs_ab(const int* __restrict__ ca, const int* __restrict__ cb, Count ncc, int pair_cap,
                              const int* __restrict__ flag, const int* __restrict__ off, int nbody, int wall_cap,
                              int walls, int* __restrict__ ma, int* __restrict__ mb,
                              const unsigned* __restrict__ counters, int* __restrict__ fx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = ncc.get();
  if (i == 0) k_fx_counts_body(counters, pair_cap, off, flag, walls ? 4 * nbody : 0, wall_cap, fx);
  if (i < pair_cap) {
    if (i >= nc) return;
    ma[i] = ca[i];
    mb[i] = cb[i];
    return;
  }
  const int t = i - pair_cap;
  if (!walls || t >= 4 * nbody || !flag[t] || off[t] >= wall_cap) return;
  const int m = nc + off[t];
  ma[m] = t >> 2;
  mb[m] = -1 - (t & 3);
}

// Clears the per-step device counters of a fixed step in one launch: the
// broad-phase counters, the colouring counters, the round counters, the
// island class counts and the island statistics.
__global__ void k_fx_clear(unsigned* __restrict__ counters, unsigned* __restrict__ dctr, int* __restrict__ rounds,
                           int nrounds, int* __restrict__ icnt, int* __restrict__ stats, int nstats) {
  const int i = threadIdx.x;
  if (i < 8) counters[i] = 0u;
  if (i < 4) dctr[i] = 0u;
  if (i < 8) icnt[i] = 0;
  if (i < nstats) stats[i] = 0;
  for (int k = i; k < nrounds; k += blockDim.x) rounds[k] = 0;
}

int fixed_clear_enqueue(rdk_ctx* ctx) {
  hipLaunchKernelGGL(k_fx_clear, dim3(1), dim3(256), 0, ctx->stream, ctx->counters.p, (unsigned*)ctx->uncolored.p,
                     ctx->fx.p + FX_ROUND0, FX_WORDS - FX_ROUND0, ctx->icnt.p, (int*)ctx->istats.p,
                     (int)(sizeof(IslandStats) / sizeof(int)));
  return RDK_OK;
}

static bool fixed_any_wall(rdk_ctx* ctx) {
  for (auto& s : ctx->sys_host)
    if (s.floor_wall || s.ceil_wall) return true;
  return false;
}

int nsc_fixed_reserve(rdk_ctx* ctx, const rdk_nsc_params* p) {
  const int n = ctx->n, cap = ctx->fx_pair_cap, mcap = ctx->fx_nm_cap;
  RDK_TRY(ensure(ctx, ctx->wall_flag, 4 * n + 1));
  RDK_TRY(ensure(ctx, ctx->wall_off, std::max(4 * n + 1, 128)));
  RDK_TRY(ensure_manifolds(ctx, mcap));
  RDK_TRY(ensure(ctx, ctx->body_start, n + 1));
  RDK_TRY(ensure(ctx, ctx->uncolored, 256));
  RDK_TRY(ensure(ctx, ctx->uf_parent, n));
  RDK_TRY(ensure(ctx, ctx->uf_label, n));
  RDK_TRY(ensure(ctx, ctx->blab, n));
  RDK_TRY(ensure(ctx, ctx->blab_alt, n));
  RDK_TRY(ensure(ctx, ctx->sbody, n));
  RDK_TRY(ensure(ctx, ctx->sbody_alt, n));
  RDK_TRY(ensure(ctx, ctx->iflag, mcap + 1));
  RDK_TRY(ensure(ctx, ctx->ipos, mcap + 1));
  RDK_TRY(ensure(ctx, ctx->istart, mcap + 2));
  RDK_TRY(ensure(ctx, ctx->ilists, 6 * (size_t)mcap + 6));
  RDK_TRY(ensure(ctx, ctx->bseg, 2 * (size_t)mcap + 2));
  RDK_TRY(ensure(ctx, ctx->icnt, 8));
  RDK_TRY(ensure(ctx, ctx->istats, sizeof(IslandStats)));
  RDK_TRY(ensure(ctx, ctx->islot, mcap));
  RDK_TRY(ensure(ctx, ctx->ihot, mcap));
  // the warm cache of the last step (eager or fixed) is kept
  RDK_TRY(grow_keep(ctx, ctx->prev_keys, cap, ctx->nprev));
  for (int k = 0; k < 3; ++k) RDK_TRY(grow_keep(ctx, ctx->prev_lam[k], cap, ctx->nprev));
  RDK_TRY(grow_keep(ctx, ctx->prev_wkeys, 4 * n + 1, ctx->nprev_w));
  for (int k = 0; k < 3; ++k) RDK_TRY(grow_keep(ctx, ctx->prev_wlam[k], 4 * n + 1, ctx->nprev_w));
  (void)p;
  return RDK_OK;
}

// Workgroups of the fixed step's class launches: none for a class that has
// held no island yet, else enough for four times the most islands of that
// class seen in a step (at least 16, 64, 16 and 4 workgroups), rounded up to a
// power of two, so that a growing packing needs few recaptures.  A step with
// more islands in a class fails its class check and is redone eagerly, which
// raises the count.  Workgroups past the count return at once.
void fixed_class_grids(const rdk_ctx* ctx, int* g) {
  const int ipb[4] = {ISL_WG0 / ISL_G0, 1, 1, 1};
  const int gmin[4] = {16, 64, 16, 4};
  for (int c = 0; c < 4; ++c) {
    if (ctx->cls_seen_max[c] <= 0) {
      g[c] = 0;
      continue;
    }
    const int64_t want = std::max<int64_t>(gmin[c], grid_for(4 * (int64_t)ctx->cls_seen_max[c], ipb[c]));
    int p2 = 1;
    while (p2 < want && p2 < (1 << 24)) p2 <<= 1;
    g[c] = p2;
  }
}

int fixed_prepare(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* np, int pair_cap, int wall_cap) {
  const int n = ctx->n;
  ctx->fx_pair_cap = pair_cap;
  ctx->fx_wall_cap = fixed_any_wall(ctx) ? std::min(wall_cap, 4 * n) : 0;
  ctx->fx_nm_cap = pair_cap + ctx->fx_wall_cap;
  ctx->fx_rounds = std::min(FX_WORDS - FX_ROUND0, std::max(8, 2 * ctx->color_rounds_max + 4));
  fixed_class_grids(ctx, ctx->fx_grid);
  RDK_TRY(ensure(ctx, ctx->fx, FX_WORDS));
  RDK_TRY(ensure(ctx, ctx->facc, FA_NUM));
  RDK_TRY(ensure(ctx, ctx->freport, (int)((sizeof(FxReport) + sizeof(double) - 1) / sizeof(double))));
  ctx->fx_prev_dev[0] = ctx->fx_prev_dev[1] = -1;  // fx may be new: upload the warm-cache counts
  RDK_TRY(ensure(ctx, ctx->resid, 4));
  RDK_TRY(detect_fixed_reserve(ctx, dp, pair_cap));
  RDK_TRY(nsc_fixed_reserve(ctx, np));
  RDK_TRY(sort_reserve(ctx, std::max(std::max(ctx->n_chunks, 2 * ctx->fx_nm_cap), std::max(4 * n, pair_cap))));
  return RDK_OK;
}

int nsc_fixed_enqueue(rdk_ctx* ctx, const rdk_nsc_params* p) {
  const int n = ctx->n, cap = ctx->fx_pair_cap, mcap = ctx->fx_nm_cap;
  const int nb = 256;
  int* fx = ctx->fx.p;
  const int* guard = fx + FX_ERR;
  hipStream_t st = ctx->stream;
  BodyPtrs B = body_ptrs(ctx);
  const Count nc = Count::dev((const int*)ctx->counters.p, cap);
  const Count nm = Count::dev(fx + FX_NM, mcap);
  const bool any_wall = fixed_any_wall(ctx);
  unsigned* dctr = (unsigned*)ctx->uncolored.p;  // [0] left, [1] maxcol, [2] overflow, [3] ncolor
  IslandStats* S = (IslandStats*)ctx->istats.p;
  const int lim = small_m_limit(ctx);
  const int R = mcap > lim ? ctx->fx_rounds : 0;
  // ---- pairs of every manifold (positions only)
  if (any_wall) {
    hipLaunchKernelGGL(k_wall_flags, dim3(grid_for(4 * n, nb)), dim3(nb), 0, st, B, ctx->sys_dev.p, n,
                       ctx->wall_flag.p);
    RDK_TRY(exclusive_scan_int(ctx, ctx->wall_flag.p, ctx->wall_off.p, Count::host(4 * n), nullptr));
  }
  hipLaunchKernelGGL(k_fx_pairs_ab, dim3(grid_for((int64_t)cap + (any_wall ? 4 * n : 0), nb)), dim3(nb), 0, st,
                     ctx->ca.p, ctx->cb.p, nc, cap, ctx->wall_flag.p, ctx->wall_off.p, n, ctx->fx_wall_cap,
                     any_wall ? 1 : 0, ctx->ma.p, ctx->mb.p, (const unsigned*)ctx->counters.p, fx);
  // ---- colouring (color_compute)
  hipLaunchKernelGGL(k_csr_entries, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->ma.p, ctx->mb.p, nm, n,
                     ctx->csr_key.p, ctx->csr_val.p);
  bool alt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->csr_key.p, ctx->csr_key_alt.p, ctx->csr_val.p, ctx->csr_val_alt.p,
                         Count::dev(fx + FX_NM, 2 * mcap, 2, 0), bits_for((uint64_t)n), alt, nullptr));
  const uint32_t* ckey = alt ? ctx->csr_key_alt.p : ctx->csr_key.p;
  const int* csr = alt ? ctx->csr_val_alt.p : ctx->csr_val.p;
  hipLaunchKernelGGL(k_body_start, dim3(grid_for(n + 1, nb)), dim3(nb), 0, st, ckey,
                     Count::dev(fx + FX_NM, 2 * mcap, 2, 0), n, ctx->body_start.p);
  hipLaunchKernelGGL(k_color_small, dim3(1), dim3(SMALL_BLOCK), 0, st, ctx->ma.p, ctx->mb.p, nm, ctx->body_start.p,
                     csr, ctx->color.p, ctx->color_tmp.p, dctr, lim, guard);
  if (R > 0) {
    Count nmr = nm;
    nmr.skip_below = lim + 1;  // counts the single workgroup coloured read as 0
    hipLaunchKernelGGL(k_fill_color, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->color.p, nmr, -1);
    for (int r = 0; r < R; ++r) {
      const unsigned* prev = r > 0 ? (const unsigned*)(fx + FX_ROUND0 + r - 1) : nullptr;
      hipLaunchKernelGGL(k_color_assign, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->ma.p, ctx->mb.p, nm,
                         ctx->body_start.p, csr, ctx->color.p, ctx->color_tmp.p, dctr + 2, lim, prev, guard);
      hipLaunchKernelGGL(k_color_resolve, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->ma.p, ctx->mb.p, nm,
                         ctx->body_start.p, csr, ctx->color_tmp.p, ctx->color.p, (unsigned*)(fx + FX_ROUND0 + r),
                         dctr + 1, lim, prev, guard);
    }
  }
  // ---- islands (island_solve); the colouring is checked by the first launch
  hipLaunchKernelGGL(k_fx_uf_init_check, dim3(grid_for(n, nb)), dim3(nb), 0, st, ctx->uf_parent.p, n,
                     (const unsigned*)dctr, fx, lim, R);
  hipLaunchKernelGGL(k_uf_link, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->ma.p, ctx->mb.p, nm,
                     ctx->uf_parent.p);
  hipLaunchKernelGGL(k_uf_flatten, dim3(grid_for(n, nb)), dim3(nb), 0, st, ctx->uf_parent.p, n, ctx->uf_label.p);
  hipLaunchKernelGGL(k_body_label_keys, dim3(grid_for(n, nb)), dim3(nb), 0, st, ctx->uf_label.p, n, ctx->blab.p,
                     ctx->sbody.p);
  bool balt = false, malt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->blab.p, ctx->blab_alt.p, ctx->sbody.p, ctx->sbody_alt.p, Count::host(n),
                         bits_for((uint64_t)n), balt, nullptr));
  const uint32_t* blab = balt ? ctx->blab_alt.p : ctx->blab.p;
  const int* sbody = balt ? ctx->sbody_alt.p : ctx->sbody.p;
  hipLaunchKernelGGL(k_island_keys, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->ma.p, ctx->color.p, nm,
                     ctx->uf_label.p, ctx->color_key.p, ctx->perm_alt.p);
  RDK_TRY(sort_pairs_u32(ctx, ctx->color_key.p, ctx->color_key_alt.p, ctx->perm_alt.p, ctx->perm.p, nm,
                         7 + bits_for((uint64_t)n), malt, nullptr));
  if (!malt)
    hipLaunchKernelGGL(k_fx_copy_int, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->perm_alt.p, ctx->perm.p, nm);
  const uint32_t* skey = malt ? ctx->color_key_alt.p : ctx->color_key.p;
  hipLaunchKernelGGL(k_island_flags, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, skey, nm, ctx->iflag.p);
  RDK_TRY(exclusive_scan_int(ctx, ctx->iflag.p, ctx->ipos.p, nm, nullptr));
  hipLaunchKernelGGL(k_island_starts, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->iflag.p, ctx->ipos.p, nm,
                     ctx->istart.p, fx + FX_NISL);
  IslandClass C;
  C.max_m[0] = ISL_G0;
  C.max_m[1] = ISL_G1;
  C.max_m[2] = ISL_G2;
  C.max_b[0] = ISL_G0 + 8;
  C.max_b[1] = ISL_G1 + 8;
  C.max_b[2] = ISL_G2 + 8;
  C.large_max = 8192;
  C.lds_bmax = LDS_BCAP;
  const Count nisl = Count::dev(fx + FX_NISL, mcap);
  hipLaunchKernelGGL(k_island_classify, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, ctx->istart.p, nisl, skey, blab, n,
                     C, ctx->ilists.p, mcap, ctx->icnt.p, ctx->bseg.p);
  const int* g = ctx->fx_grid;  // sized from the island counts seen so far
  hipLaunchKernelGGL(k_fx_class_check, dim3(1), dim3(64), 0, st, (const int*)ctx->icnt.p, g[0] * (ISL_WG0 / ISL_G0),
                     g[1], g[2], g[3], fx);
  // ---- every check is behind: free flight, then the manifolds (nsc_build_run)
  RDK_TRY(integrate_run(ctx, RDK_NSC_FREE, guard));
  MfPtrs M = mf_ptrs(ctx, false);
  {
    ContactIn Ci;
    Ci.a = ctx->ca.p;
    Ci.b = ctx->cb.p;
    for (int k = 0; k < 3; ++k) {
      Ci.pa[k] = ctx->cpa[k].p;
      Ci.pb[k] = ctx->cpb[k].p;
      Ci.n[k] = ctx->cn[k].p;
      Ci.sh[k] = ctx->csh[k].p;
    }
    Ci.dist = ctx->cdist.p;
    Ci.hs = ctx->chs.p;
    const int warm = p->warm_mode != 0 ? 1 : 0;
    hipLaunchKernelGGL(k_manifold, dim3(grid_for(cap, nb)), dim3(nb), 0, st, B, Ci, nc, M, p->cfm, warm,
                       ctx->prev_keys.p, 0, (const int*)(fx + FX_NPREV), ctx->prev_lam[0].p, ctx->prev_lam[1].p,
                       ctx->prev_lam[2].p);
  }
  if (any_wall) {
    const int wwarm = (p->warm_mode != 0 && p->warm_walls) ? 1 : 0;
    hipLaunchKernelGGL(k_wall_manifold, dim3(grid_for(4 * n, nb)), dim3(nb), 0, st, B, ctx->sys_dev.p, n,
                       ctx->wall_flag.p, ctx->wall_off.p, ctx->fx_wall_cap, nc, M, p->cfm, wwarm,
                       ctx->prev_wkeys.p, 0, (const int*)(fx + FX_NPREV_W), ctx->prev_wlam[0].p, ctx->prev_wlam[1].p,
                       ctx->prev_wlam[2].p);
  }
  MfPtrs Mc = mf_ptrs(ctx, true);
  hipLaunchKernelGGL(k_permute, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, M, Mc, ctx->perm.p, nm);
  IslandSolve P;
  P.K = solve_const(ctx, p);
  P.tol = p->tol;
  P.max_sweeps = p->max_sweeps;
  P.check_every = std::max(1, p->check_every);
  P.stop_rule = p->stop_rule;
  P.warm = p->warm_mode != 0 ? 1 : 0;
  P.stall = p->stall_sweeps;
  P.hot = p->hot_set;
  // The classes that have held islands, concurrently on four streams as in
  // the eager solve when there are at least two of them; a class never seen is
  // not launched (an island of it fails the class check above).
  const int* cnt = ctx->icnt.p;
  const bool active[3] = {ctx->cls_seen_max[0] > 0, ctx->cls_seen_max[1] > 0, ctx->cls_seen_max[2] > 0};
  const int nactive = active[0] + active[1] + active[2] + (ctx->cls_seen_max[3] > 0);
  const hipStream_t side[3] = {ctx->stream2, ctx->stream3, ctx->stream4};
  const hipEvent_t joins[3] = {ctx->ev_join, ctx->ev_join3, ctx->ev_join4};
  hipStream_t cs[3];
  for (int c = 0; c < 3; ++c) cs[c] = (nactive >= 2 && active[c]) ? side[c] : st;
  if (nactive >= 2) {
    RDK_HIP(hipEventRecord(ctx->ev_fork, st));
    for (int c = 0; c < 3; ++c)
      if (cs[c] != st) RDK_HIP(hipStreamWaitEvent(cs[c], ctx->ev_fork, 0));
  }
  // the LDS class first, as in island_solve: it needs a whole CU's LDS
  if (g[3] > 0)
    hipLaunchKernelGGL(k_island_lds, dim3(g[3]), dim3(LDS_THREADS), 0, st, B, Mc, ctx->sys_dev.p, ctx->istart.p,
                       skey, ctx->bseg.p, sbody, ctx->ilists.p + 3 * (size_t)mcap, Count::dev(cnt + 3, mcap), P, S,
                       ctx->islot.p, ctx->ihot.p, guard);
  if (g[0] > 0)
    hipLaunchKernelGGL((k_island_reg<ISL_G0, ISL_WG0>), dim3(g[0]), dim3(ISL_WG0), 0, cs[0], B, Mc, ctx->sys_dev.p,
                       ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p, Count::dev(cnt + 0, mcap), P, S,
                       guard);
  if (g[1] > 0)
    hipLaunchKernelGGL((k_island_reg<ISL_G1, ISL_WG1>), dim3(g[1]), dim3(ISL_WG1), 0, cs[1], B, Mc, ctx->sys_dev.p,
                       ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p + mcap, Count::dev(cnt + 1, mcap), P,
                       S, guard);
  if (g[2] > 0)
    hipLaunchKernelGGL((k_island_reg<ISL_G2, ISL_WG2>), dim3(g[2]), dim3(ISL_WG2), 0, cs[2], B, Mc, ctx->sys_dev.p,
                       ctx->istart.p, skey, ctx->bseg.p, sbody, ctx->ilists.p + 2 * (size_t)mcap,
                       Count::dev(cnt + 2, mcap), P, S, guard);
  for (int c = 0; c < 3; ++c) {
    if (cs[c] == st) continue;
    RDK_HIP(hipEventRecord(joins[c], cs[c]));
    RDK_HIP(hipStreamWaitEvent(st, joins[c], 0));
  }
  hipLaunchKernelGGL(k_unpermute_lambda, dim3(grid_for(mcap, nb)), dim3(nb), 0, st, Mc, M, ctx->perm.p, nm);
  // ---- warm cache (store_warm_cache)
  if (p->warm_mode != 0)
    hipLaunchKernelGGL(k_copy_keys_lambda, dim3(grid_for(cap, nb)), dim3(nb), 0, st, ctx->pkeys.p, M, nc,
                       ctx->prev_keys.p, ctx->prev_lam[0].p, ctx->prev_lam[1].p, ctx->prev_lam[2].p, fx + FX_NPREV,
                       guard);
  if (p->warm_mode != 0 && p->warm_walls && any_wall)
    hipLaunchKernelGGL(k_copy_wall_keys_lambda, dim3(grid_for(4 * n, nb)), dim3(nb), 0, st, M, nc,
                       Count::dev(fx + FX_NWALL, ctx->fx_wall_cap), ctx->prev_wkeys.p, ctx->prev_wlam[0].p,
                       ctx->prev_wlam[1].p, ctx->prev_wlam[2].p, fx + FX_NPREV_W, guard);
  hipLaunchKernelGGL(k_fx_accumulate, dim3(1), dim3(64), 0, st, (const int*)fx, fx, (const unsigned*)ctx->counters.p,
                     cap, (const IslandStats*)S, (const unsigned*)dctr, (const int*)ctx->icnt.p, ctx->facc.p);
  return RDK_OK;
}

// Packs the fixed-step words, counters, colouring state, island counts and
// statistics into one FxReport (a single transfer to the host afterwards).
__global__ void k_fx_report(const int* __restrict__ fx, const unsigned* __restrict__ ctr,
                            const unsigned* __restrict__ dc, const int* __restrict__ cnt,
                            const double* __restrict__ acc, FxReport* __restrict__ out) {
  const int t = threadIdx.x;
  if (t < FA_NUM) out->acc[t] = acc[t];
  if (t < FX_ROUND0) out->fx[t] = fx[t];
  if (t < 2) out->ctr[t] = ctr[t];
  if (t < 4) out->dc[t] = dc[t];
  if (t < 6) out->cnt[t] = cnt[t];
}

// Enqueues the report of the fixed steps run so far and its copy into the
// pinned host FxReport (valid after the next synchronisation of the stream).
int fixed_report_enqueue(rdk_ctx* ctx) {
  constexpr int words = (int)((sizeof(FxReport) + sizeof(double) - 1) / sizeof(double));
  if (ctx->freport.cap < (size_t)words) {
    ctx->err = "fixed step: report buffer not reserved";
    return RDK_ERR_STATE;
  }
  if (!ctx->h_report) RDK_HIP(hipHostMalloc((void**)&ctx->h_report, sizeof(FxReport), hipHostMallocDefault));
  hipLaunchKernelGGL(k_fx_report, dim3(1), dim3(64), 0, ctx->stream, (const int*)ctx->fx.p,
                     (const unsigned*)ctx->counters.p, (const unsigned*)ctx->uncolored.p, (const int*)ctx->icnt.p,
                     (const double*)ctx->facc.p, (FxReport*)ctx->freport.p);
  RDK_HIP(hipMemcpyAsync(ctx->h_report, ctx->freport.p, sizeof(FxReport), hipMemcpyDeviceToHost, ctx->stream));
  return RDK_OK;
}

// Host copies of the last fixed step's counts (what the eager step leaves in
// the context), and the warm-cache sizes the eager path reads from the host.
void fixed_apply_counts(rdk_ctx* ctx, const FxReport& R) {
  ctx->nc = (int)std::min<unsigned>(R.ctr[0], (unsigned)ctx->fx_pair_cap);
  ctx->nm = R.fx[FX_NM];
  ctx->nm_wall = R.fx[FX_NWALL];
  ctx->nprev = R.fx[FX_NPREV];
  ctx->nprev_w = R.fx[FX_NPREV_W];
  ctx->fx_prev_dev[0] = ctx->nprev;
  ctx->fx_prev_dev[1] = ctx->nprev_w;
  ctx->nislands = R.fx[FX_NISL];
  ctx->ncolor = ctx->nm > 0 ? (int)R.dc[1] + 1 : 0;
  for (int c = 0; c < 3; ++c) ctx->island_counts[c] = R.cnt[c];
  ctx->island_counts[3] = R.cnt[3] + R.cnt[4];
  ctx->have_manifolds = true;
  ctx->have_coloring = false;
}

int fixed_sync_host_counts(rdk_ctx* ctx) {
  RDK_TRY(fixed_report_enqueue(ctx));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  fixed_apply_counts(ctx, *ctx->h_report);
  return RDK_OK;
}

}  // namespace rdk

// rdk_integrate.hip -- rigid-body integrators and observables.
//
//   RDK_NSC_FREE  free-flight velocity predictor       main.cpp:4397-4408
//   RDK_NSC_POS   position / quaternion update, damping, omega clamp,
//                 periodic wrap and force clear         main.cpp:4469-4483,4508-4527
//   RDK_VV_HALF1  integrateHalfPos                     integrator.cpp:113-135
//   RDK_VV_HALF2  integrateSecondHalf                  integrator.cpp:138-159
// One thread per body, planar FP64 loads (fully coalesced).  The world
// inertia of a solid cylinder is applied in closed form from the rod axis.
#include <cmath>

#include "rdk_internal.hpp"

namespace rdk {

struct PhysK {
  double g[3];
  double dt, lin_damp, ang_damp, w_max;
  double exp_lin, exp_ang;  // exp(-c dt) for the Verlet damping
};

__device__ __forceinline__ d3 i_ld3(double* const* a, int i) { return mk3(a[0][i], a[1][i], a[2][i]); }
__device__ __forceinline__ void i_st3(double* const* a, int i, d3 v) {
  a[0][i] = v.x;
  a[1][i] = v.y;
  a[2][i] = v.z;
}

// I_w^-1 (tau - w x (I_w w))  (integrator.cpp:39-54)
__device__ __forceinline__ d3 gyro_rhs(const BodyPtrs& B, int i, d3 u, d3 w, d3 tau) {
  const d3 Iw_w = inertia_apply(u, B.i_perp[i], B.i_axial[i], w);
  const d3 gyro = cross(w, Iw_w);
  return inertia_apply(u, B.i_perp_inv[i], B.i_axial_inv[i], tau - gyro);
}

// integrator.cpp:23-35 (bounded while loops)
__device__ __forceinline__ double wrap1(double p, double lo, double hi, int periodic) {
  if (!periodic) return p;
  const double span = hi - lo;
  if (!(span > 0.0)) return p;
  int g = 0;
  while (p < lo && g++ < 8) p += span;
  g = 0;
  while (p >= hi && g++ < 8) p -= span;
  return p;
}

__global__ void __launch_bounds__(256) k_integrate(BodyPtrs B, const SysDev* __restrict__ sys, int n, int stage,
                                                   PhysK P, const int* __restrict__ guard) {
  if (guard && *guard) return;  // a failed captured step changes no state
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const d3 zero = mk3(0, 0, 0);
  if (stage == RDK_CLEAR) {
    i_st3(B.f, i, zero);
    i_st3(B.tau, i, zero);
    return;
  }
  if (!(B.inv_mass[i] > 0.0)) {
    if (stage == RDK_NSC_POS || stage == RDK_VV_HALF2) {
      i_st3(B.f, i, zero);
      i_st3(B.tau, i, zero);
    }
    return;
  }
  const double dt = P.dt;
  const d3 g = mk3(P.g[0], P.g[1], P.g[2]);
  q4 q = q4{B.q[0][i], B.q[1][i], B.q[2][i], B.q[3][i]};
  d3 v = i_ld3(B.v, i), w = i_ld3(B.w, i);
  if (stage == RDK_NSC_FREE) {
    const d3 f = i_ld3(B.f, i), tau = i_ld3(B.tau, i);
    v = v + (f * B.inv_mass[i] + g) * dt;
    w = w + gyro_rhs(B, i, axis_of(q), w, tau) * dt;
    i_st3(B.v, i, v);
    i_st3(B.w, i, w);
  } else if (stage == RDK_NSC_POS) {
    d3 x = i_ld3(B.x, i);
    x = x + v * dt;
    const double hdt = 0.5 * dt;
    const q4 sw = q4{0.0 * hdt, w.x * hdt, w.y * hdt, w.z * hdt};
    const q4 dq = qmul(sw, q);
    q = q4{q.w + dq.w, q.x + dq.x, q.y + dq.y, q.z + dq.z};
    q = qnormalize(q);
    v = v * (1.0 - P.lin_damp * dt);
    w = w * (1.0 - P.ang_damp * dt);
    const double wl = len(w);
    if (wl > P.w_max) w = w * (P.w_max / wl);
    const SysDev S = sys[B.sys[i]];
    x.x = wrap1(x.x, S.lo[0], S.hi[0], S.periodic[0]);
    x.y = wrap1(x.y, S.lo[1], S.hi[1], S.periodic[1]);
    x.z = wrap1(x.z, S.lo[2], S.hi[2], S.periodic[2]);
    i_st3(B.x, i, x);
    B.q[0][i] = q.w;
    B.q[1][i] = q.x;
    B.q[2][i] = q.y;
    B.q[3][i] = q.z;
    i_st3(B.v, i, v);
    i_st3(B.w, i, w);
    i_st3(B.f, i, zero);
    i_st3(B.tau, i, zero);
  } else if (stage == RDK_VV_HALF1) {
    const d3 f = i_ld3(B.f, i), tau = i_ld3(B.tau, i);
    const d3 al = f * (1.0 / B.mass[i]) + g;
    const d3 aa = gyro_rhs(B, i, axis_of(q), w, tau);
    v = v + al * (0.5 * dt);
    w = w + aa * (0.5 * dt);
    d3 x = i_ld3(B.x, i);
    x = x + v * dt;
    const SysDev S = sys[B.sys[i]];
    x.x = wrap1(x.x, S.lo[0], S.hi[0], S.periodic[0]);
    x.y = wrap1(x.y, S.lo[1], S.hi[1], S.periodic[1]);
    x.z = wrap1(x.z, S.lo[2], S.hi[2], S.periodic[2]);
    const double wl = len(w);
    if (wl > 1e-8) {
      const double angle = wl * dt;
      const d3 ax = w * (1.0 / wl);
      const double s = sin(angle * 0.5);
      const q4 dq = q4{cos(angle * 0.5), ax.x * s, ax.y * s, ax.z * s};
      q = qnormalize(qmul(dq, q));
    }
    i_st3(B.x, i, x);
    B.q[0][i] = q.w;
    B.q[1][i] = q.x;
    B.q[2][i] = q.y;
    B.q[3][i] = q.z;
    i_st3(B.v, i, v);
    i_st3(B.w, i, w);
  } else if (stage == RDK_VV_HALF2) {
    const d3 f = i_ld3(B.f, i), tau = i_ld3(B.tau, i);
    const d3 al = f * (1.0 / B.mass[i]) + g;
    const d3 aa = gyro_rhs(B, i, axis_of(q), w, tau);
    v = v + al * (0.5 * dt);
    w = w + aa * (0.5 * dt);
    if (P.lin_damp > 0.0) v = v * P.exp_lin;
    if (P.ang_damp > 0.0) w = w * P.exp_ang;
    const double wl = len(w);
    if (wl > P.w_max) w = w * (P.w_max / wl);
    i_st3(B.v, i, v);
    i_st3(B.w, i, w);
    i_st3(B.f, i, zero);
    i_st3(B.tau, i, zero);
  }
}

// KE = sum 1/2 m v^2 + 1/2 w.(I_w w)  (main.cpp:3357-3368)
__global__ void k_kinetic(BodyPtrs B, int n, double* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0.0;
  if (i < n) {
    const d3 v = i_ld3(B.v, i), w = i_ld3(B.w, i);
    const q4 q = q4{B.q[0][i], B.q[1][i], B.q[2][i], B.q[3][i]};
    ke = 0.5 * B.mass[i] * dot(v, v);
    if (B.mass[i] > 0.0) ke += 0.5 * dot(w, inertia_apply(axis_of(q), B.i_perp[i], B.i_axial[i], w));
  }
  for (int off = 32; off > 0; off >>= 1) ke += __shfl_down(ke, off);
  if (__lane_id() == 0 && ke != 0.0) atomicAdd(out, ke);
}

int integrate_run(rdk_ctx* ctx, int stage, const int* guard) {
  if (!ctx->have_phys) {
    ctx->err = "rdk_integrate: physics parameters not set";
    return RDK_ERR_STATE;
  }
  if (stage < 0 || stage > RDK_CLEAR) {
    ctx->err = "rdk_integrate: unknown stage";
    return RDK_ERR_ARG;
  }
  PhysK P;
  for (int k = 0; k < 3; ++k) P.g[k] = ctx->phys.gravity[k];
  P.dt = ctx->phys.dt;
  P.lin_damp = ctx->phys.lin_damp;
  P.ang_damp = ctx->phys.ang_damp;
  P.w_max = ctx->phys.w_max;
  P.exp_lin = std::exp(-ctx->phys.lin_damp * ctx->phys.dt);
  P.exp_ang = std::exp(-ctx->phys.ang_damp * ctx->phys.dt);
  const int kid = stage == RDK_NSC_FREE ? K_NSC_FREE
                  : stage == RDK_NSC_POS ? K_NSC_POS
                  : stage == RDK_VV_HALF1 ? K_VV_HALF1
                  : stage == RDK_VV_HALF2 ? K_VV_HALF2
                                          : K_MISC;
  ProfScope ps(ctx, kid);
  hipLaunchKernelGGL(k_integrate, dim3(grid_for(ctx->n, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx),
                     ctx->sys_dev.p, ctx->n, stage, P, guard);
  if (stage != RDK_CLEAR) ctx->work[W_INTEGRATE_BODIES] += ctx->n;
  return RDK_OK;
}

int observables_run(rdk_ctx* ctx, double* ke, double* phi) {
  RDK_TRY(ensure(ctx, ctx->red, 4));
  RDK_HIP(hipMemsetAsync(ctx->red.p, 0, sizeof(double), ctx->stream));
  {
    ProfScope ps(ctx, K_OBSERVE);
    hipLaunchKernelGGL(k_kinetic, dim3(grid_for(ctx->n, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx), ctx->n,
                       ctx->red.p);
  }
  double h = 0.0;
  RDK_HIP(hipMemcpyAsync(&h, ctx->red.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  if (ke) *ke = h;
  if (phi) {
    double vol = 0.0, box = 0.0;
    for (int i = 0; i < ctx->n; ++i) {
      const double r = ctx->h_r[i], hh = ctx->h_h[i];
      vol += M_PI * r * r * (2.0 * hh) + (4.0 / 3.0) * M_PI * r * r * r;
    }
    for (auto& s : ctx->sys_host) box += s.span[0] * s.span[1] * s.span[2];
    *phi = box > 0.0 ? vol / box : 0.0;
  }
  return RDK_OK;
}

// Affine isotropic scaling about each box's lower corner: x <- lo + s (x - lo).
__global__ void k_affine(BodyPtrs B, const SysDev* __restrict__ sys, int n, double s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SysDev& S = sys[B.sys[i]];
  for (int k = 0; k < 3; ++k) B.x[k][i] = S.lo[k] + s * (B.x[k][i] - S.lo[k]);
}

int affine_run(rdk_ctx* ctx, double s) {
  for (int sy = 0; sy < ctx->nsys; ++sy) {
    SysDev& S = ctx->sys_host[sy];
    for (int k = 0; k < 3; ++k) {
      S.span[k] *= s;
      S.hi[k] = S.lo[k] + S.span[k];
      ctx->sys_desc[sy].hi[k] = S.hi[k];
    }
  }
  // positions scale with the old device constants (lo is unchanged)
  {
    ProfScope ps(ctx, K_NSC_POS);
    hipLaunchKernelGGL(k_affine, dim3(grid_for(ctx->n, 256)), dim3(256), 0, ctx->stream, body_ptrs(ctx),
                       ctx->sys_dev.p, ctx->n, s);
  }
  if (ctx->n_chunks > 0) return regrid(ctx);  // cells, margins and spans for the new boxes
  ctx->grid_dirty = true;
  RDK_TRY(ensure(ctx, ctx->sys_dev, ctx->nsys));
  RDK_HIP(hipMemcpyAsync(ctx->sys_dev.p, ctx->sys_host.data(), ctx->nsys * sizeof(SysDev), hipMemcpyHostToDevice,
                         ctx->stream));
  return RDK_OK;
}

}  // namespace rdk

// rdk_api.hip -- the C ABI of librdk (declared in include/rdk.h).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

#include "rdk_internal.hpp"

using namespace rdk;

namespace rdk {

BodyPtrs body_ptrs(rdk_ctx* ctx) {
  BodyPtrs B;
  for (int k = 0; k < 3; ++k) {
    B.x[k] = ctx->x[k].p;
    B.v[k] = ctx->v[k].p;
    B.w[k] = ctx->w[k].p;
    B.f[k] = ctx->f[k].p;
    B.tau[k] = ctx->tau[k].p;
  }
  for (int k = 0; k < 4; ++k) B.q[k] = ctx->q[k].p;
  B.r = ctx->r.p;
  B.h = ctx->h.p;
  B.mass = ctx->mass.p;
  B.inv_mass = ctx->inv_mass.p;
  B.i_perp = ctx->i_perp.p;
  B.i_axial = ctx->i_axial.p;
  B.i_perp_inv = ctx->i_perp_inv.p;
  B.i_axial_inv = ctx->i_axial_inv.p;
  B.rest = ctx->rest.p;
  B.sys = ctx->sys_of.p;
  return B;
}

bool graphs_enabled() {
  static int on = -1;
  if (on < 0) {
    const char* e = getenv("RDK_GRAPHS");
    on = (e && e[0] == '0') ? 0 : 1;
  }
  return on == 1;
}

void prof_begin(rdk_ctx* ctx, int kernel, EventRec& rec, hipStream_t s) {
  rec.kernel = kernel;
  for (int k = 0; k < 2; ++k) {
    hipEvent_t e;
    if (!ctx->ev_pool.empty()) {
      e = ctx->ev_pool.back();
      ctx->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    (k == 0 ? rec.a : rec.b) = e;
  }
  if (rec.a) (void)hipEventRecord(rec.a, s);
}

bool prof_ext_events(rdk_ctx* ctx, int k, EventRec& rec) {
  if (!timed(ctx, k)) return false;
  rec.kernel = k;
  for (int j = 0; j < 2; ++j) {
    hipEvent_t e;
    if (!ctx->ev_pool.empty()) {
      e = ctx->ev_pool.back();
      ctx->ev_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    (j == 0 ? rec.a : rec.b) = e;
  }
  ctx->k_launches[k] += 1;
  return true;
}

void prof_ext_done(rdk_ctx* ctx, EventRec& rec) { ctx->ev_pending.push_back(rec); }

void prof_end(rdk_ctx* ctx, EventRec& rec, hipStream_t s) {
  if (rec.b) (void)hipEventRecord(rec.b, s);
  ctx->ev_pending.push_back(rec);
}

int prof_flush(rdk_ctx* ctx) {
  if (ctx->ev_pending.empty()) return RDK_OK;
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->ev_pending) {
    float ms = 0.0f;
    if (r.b) (void)hipEventSynchronize(r.b);  // events may sit on the side stream
    if (r.a && r.b && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) ctx->k_ms[r.kernel] += ms;
    if (r.a) ctx->ev_pool.push_back(r.a);
    if (r.b) ctx->ev_pool.push_back(r.b);
  }
  ctx->ev_pending.clear();
  return RDK_OK;
}

static const char* kKernelNames[K_NUM] = {
    "chunk_keys", "cell_ranges", "chunk_gather", "broad_narrow", "contact_record", "sort_scan",
    "manifold", "wall", "csr", "color", "permute", "warm_apply", "psor_sweep", "psor_small", "residual",
    "nsc_free", "nsc_pos", "vv_half1", "vv_half2", "force", "body_reduce", "history", "observe",
    "mlp", "island", "island_tiny", "island_small", "island_medium", "island_large", "island_huge", "pos_project",
    "populate", "misc"};

}  // namespace rdk

#define CTX_CHECK()                  \
  do {                               \
    if (!ctx) return RDK_ERR_ARG;    \
    if (hipSetDevice(ctx->device) != hipSuccess) { \
      ctx->err = "hipSetDevice failed"; \
      return RDK_ERR_HIP;            \
    }                                \
  } while (0)

extern "C" {

int rdk_version(void) { return 1; }

static void destroy_streams(rdk_ctx* ctx) {
  for (hipEvent_t* e : {&ctx->ev_fork, &ctx->ev_join, &ctx->ev_join3, &ctx->ev_join4}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
  for (hipStream_t* s : {&ctx->stream4, &ctx->stream3, &ctx->stream2, &ctx->stream}) {
    if (*s) (void)hipStreamDestroy(*s);
    *s = nullptr;
  }
}

int rdk_create(int device, rdk_ctx** out) {
  if (!out) return RDK_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RDK_ERR_HIP;
  if (device < 0 || device >= ndev) return RDK_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return RDK_ERR_HIP;
  rdk_ctx* ctx = new (std::nothrow) rdk_ctx();
  if (!ctx) return RDK_ERR_HIP;
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream4, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join3, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->ev_join4, hipEventDisableTiming) != hipSuccess) {
    destroy_streams(ctx);
    delete ctx;
    return RDK_ERR_HIP;
  }
  if (noalloc_scratch_init(ctx) != RDK_OK) {  // scratch of rdk_entanglement_noalloc on this device
    destroy_streams(ctx);
    delete ctx;
    return RDK_ERR_HIP;
  }
  *out = ctx;
  return RDK_OK;
}

static void drop_step_graph(rdk_ctx* ctx);

int rdk_destroy(rdk_ctx* ctx) {
  if (!ctx) return RDK_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  mlp_free(ctx);
  for (int k = 0; k < 3; ++k) {
    release(ctx->x[k]); release(ctx->v[k]); release(ctx->w[k]); release(ctx->f[k]); release(ctx->tau[k]);
    release(ctx->cpa[k]); release(ctx->cpb[k]); release(ctx->cn[k]); release(ctx->csh[k]);
    release(ctx->lam[k]); release(ctx->lamc[k]); release(ctx->prev_lam[k]); release(ctx->prev_wlam[k]);
    release(ctx->hxi[k]); release(ctx->hxi_new[k]);
  }
  for (int k = 0; k < 4; ++k) release(ctx->q[k]);
  for (int k = 0; k < 12; ++k) release(ctx->cf[k]);
  release(ctx->r); release(ctx->h); release(ctx->mass); release(ctx->inv_mass); release(ctx->i_perp);
  release(ctx->i_axial); release(ctx->i_perp_inv); release(ctx->i_axial_inv); release(ctx->rest);
  release(ctx->sys_of); release(ctx->sys_dev);
  release(ctx->chunk_rod); release(ctx->chunk_k); release(ctx->rod_nchunk); release(ctx->rodrec); release(ctx->rodinfo);
  release(ctx->ckey); release(ctx->ckey_alt); release(ctx->cid); release(ctx->cid_alt);
  release(ctx->cell_start); release(ctx->scf); release(ctx->sca); release(ctx->ucf); release(ctx->uca); release(ctx->srod); release(ctx->sk);
  release(ctx->pkeys); release(ctx->pkeys_alt); release(ctx->counters);
  release(ctx->ca); release(ctx->cb); release(ctx->ctype); release(ctx->cdist); release(ctx->chs);
  release(ctx->cs); release(ctx->ct);
  release(ctx->mf); release(ctx->mfc); release(ctx->ma); release(ctx->mb); release(ctx->mac); release(ctx->mbc);
  release(ctx->wall_flag); release(ctx->wall_off); release(ctx->color); release(ctx->color_tmp);
  release(ctx->perm); release(ctx->perm_alt); release(ctx->uncolored); release(ctx->color_key);
  release(ctx->color_key_alt); release(ctx->csr_key); release(ctx->csr_key_alt); release(ctx->csr_val);
  release(ctx->csr_val_alt); release(ctx->body_start); release(ctx->d_cstart); release(ctx->small_out);
  release(ctx->uf_parent); release(ctx->uf_label); release(ctx->sbody); release(ctx->sbody_alt);
  release(ctx->iflag); release(ctx->ipos); release(ctx->istart); release(ctx->ilists); release(ctx->bseg);
  release(ctx->icnt); release(ctx->blab); release(ctx->blab_alt); release(ctx->istats); release(ctx->hcs);
  for (int k = 0; k < 3; ++k) { release(ctx->ent_a[k]); release(ctx->ent_e[k]); }
  release(ctx->ent_out); release(ctx->ent_psum); release(ctx->ent_pcnt); release(ctx->hres); release(ctx->hstate); release(ctx->raw_ptrs); release(ctx->raw_out); release(ctx->kept_mf); release(ctx->kept_ab); release(ctx->kept_lam); release(ctx->soft_kept); release(ctx->islot); release(ctx->ihot); release(ctx->prev_keys); release(ctx->prev_wkeys); release(ctx->resid);
  release(ctx->hkeys); release(ctx->hkeys_new); release(ctx->hlast); release(ctx->hlast_new);
  release(ctx->hflag); release(ctx->hscan); release(ctx->pe); release(ctx->temp); release(ctx->scan_tmp); release(ctx->red);
  drop_step_graph(ctx);
  release(ctx->fx); release(ctx->facc); release(ctx->freport);
  if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
  if (ctx->h_report) (void)hipHostFree(ctx->h_report);
  if (ctx->h_fx_up) (void)hipHostFree(ctx->h_fx_up);
  for (auto& r : ctx->ev_pending) {
    if (r.a) (void)hipEventDestroy(r.a);
    if (r.b) (void)hipEventDestroy(r.b);
  }
  for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
  destroy_streams(ctx);
  delete ctx;
  return RDK_OK;
}

const char* rdk_last_error(rdk_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rdk_sync(rdk_ctx* ctx) {
  CTX_CHECK();
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

static int upload_planar(rdk_ctx* ctx, DevBuf<double>* dst, int comps, const double* src, int n) {
  std::vector<double> tmp(n);
  for (int k = 0; k < comps; ++k) {
    for (int i = 0; i < n; ++i) tmp[i] = src[(size_t)comps * i + k];
    RDK_HIP(hipMemcpyAsync(dst[k].p, tmp.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
  }
  return RDK_OK;
}

static int download_planar(rdk_ctx* ctx, DevBuf<double>* src, int comps, double* dst, int n) {
  std::vector<double> tmp(n);
  for (int k = 0; k < comps; ++k) {
    RDK_HIP(hipMemcpyAsync(tmp.data(), src[k].p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) dst[(size_t)comps * i + k] = tmp[i];
  }
  return RDK_OK;
}

static int upload_scalar(rdk_ctx* ctx, DevBuf<double>& dst, const std::vector<double>& src) {
  RDK_HIP(hipMemcpyAsync(dst.p, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_set_bodies(rdk_ctx* ctx, int n, const double* x, const double* q, const double* v, const double* w,
                   const double* r, const double* h, const double* mass, const double* rest) {
  CTX_CHECK();
  if (n <= 0 || !x || !q || !r || !h || !mass) {
    ctx->err = "rdk_set_bodies: n > 0 and x, q, r, h, mass are required";
    return RDK_ERR_ARG;
  }
  ctx->n = n;
  for (int k = 0; k < 3; ++k) {
    RDK_TRY(ensure(ctx, ctx->x[k], n)); RDK_TRY(ensure(ctx, ctx->v[k], n)); RDK_TRY(ensure(ctx, ctx->w[k], n));
    RDK_TRY(ensure(ctx, ctx->f[k], n)); RDK_TRY(ensure(ctx, ctx->tau[k], n));
    RDK_HIP(hipMemsetAsync(ctx->f[k].p, 0, n * sizeof(double), ctx->stream));
    RDK_HIP(hipMemsetAsync(ctx->tau[k].p, 0, n * sizeof(double), ctx->stream));
    RDK_HIP(hipMemsetAsync(ctx->v[k].p, 0, n * sizeof(double), ctx->stream));
    RDK_HIP(hipMemsetAsync(ctx->w[k].p, 0, n * sizeof(double), ctx->stream));
  }
  for (int k = 0; k < 4; ++k) RDK_TRY(ensure(ctx, ctx->q[k], n));
  RDK_TRY(ensure(ctx, ctx->r, n)); RDK_TRY(ensure(ctx, ctx->h, n)); RDK_TRY(ensure(ctx, ctx->mass, n));
  RDK_TRY(ensure(ctx, ctx->inv_mass, n)); RDK_TRY(ensure(ctx, ctx->i_perp, n)); RDK_TRY(ensure(ctx, ctx->i_axial, n));
  RDK_TRY(ensure(ctx, ctx->i_perp_inv, n)); RDK_TRY(ensure(ctx, ctx->i_axial_inv, n));
  RDK_TRY(ensure(ctx, ctx->rest, n)); RDK_TRY(ensure(ctx, ctx->sys_of, n));
  ctx->h_r.assign(r, r + n);
  ctx->h_h.assign(h, h + n);
  ctx->h_mass.assign(mass, mass + n);
  std::vector<double> im(n), ip(n), ia(n), ipi(n), iai(n), rs(n);
  for (int i = 0; i < n; ++i) {
    // rigid_body.cpp:17-38: solid cylinder of length 2h about local +Y
    const double m = mass[i];
    if (m > 0.0) {
      const double L = 2.0 * h[i];
      im[i] = 1.0 / m;
      ip[i] = m * (3.0 * r[i] * r[i] + L * L) / 12.0;
      ia[i] = m * (r[i] * r[i]) / 2.0;
      ipi[i] = ip[i] > 0 ? 1.0 / ip[i] : 0.0;
      iai[i] = ia[i] > 0 ? 1.0 / ia[i] : 0.0;
    } else {
      im[i] = ip[i] = ia[i] = ipi[i] = iai[i] = 0.0;
    }
    rs[i] = rest ? rest[i] : 0.0;
  }
  RDK_TRY(upload_scalar(ctx, ctx->r, ctx->h_r));
  RDK_TRY(upload_scalar(ctx, ctx->h, ctx->h_h));
  RDK_TRY(upload_scalar(ctx, ctx->mass, ctx->h_mass));
  RDK_TRY(upload_scalar(ctx, ctx->inv_mass, im));
  RDK_TRY(upload_scalar(ctx, ctx->i_perp, ip));
  RDK_TRY(upload_scalar(ctx, ctx->i_axial, ia));
  RDK_TRY(upload_scalar(ctx, ctx->i_perp_inv, ipi));
  RDK_TRY(upload_scalar(ctx, ctx->i_axial_inv, iai));
  RDK_TRY(upload_scalar(ctx, ctx->rest, rs));
  RDK_TRY(rdk_upload_state(ctx, x, q, v, w));
  // default: one system holding all bodies, non-periodic unit box
  ctx->h_sys_of.assign(n, 0);
  RDK_HIP(hipMemsetAsync(ctx->sys_of.p, 0, n * sizeof(int), ctx->stream));
  ctx->nsys = 0;
  ctx->built_chunk_len = -2.0;
  ctx->n_chunks = 0;
  ctx->nc = ctx->nm = ctx->nprev = ctx->nprev_w = ctx->nhist = 0;
  ctx->have_manifolds = ctx->have_coloring = false;
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_set_systems(rdk_ctx* ctx, int nsys, const int* off, const rdk_system_desc* desc) {
  CTX_CHECK();
  if (ctx->n <= 0) {
    ctx->err = "rdk_set_systems: call rdk_set_bodies first";
    return RDK_ERR_STATE;
  }
  if (nsys <= 0 || !off || !desc || off[0] != 0 || off[nsys] != ctx->n) {
    ctx->err = "rdk_set_systems: body_offsets must start at 0 and end at n";
    return RDK_ERR_ARG;
  }
  ctx->nsys = nsys;
  ctx->sys_desc.assign(desc, desc + nsys);
  ctx->sys_off.assign(off, off + nsys + 1);
  ctx->sys_host.assign(nsys, SysDev{});
  ctx->h_sys_of.assign(ctx->n, 0);
  for (int s = 0; s < nsys; ++s) {
    if (off[s + 1] < off[s]) {
      ctx->err = "rdk_set_systems: body_offsets must be non-decreasing";
      return RDK_ERR_ARG;
    }
    for (int i = off[s]; i < off[s + 1]; ++i) ctx->h_sys_of[i] = s;
    SysDev& S = ctx->sys_host[s];
    for (int k = 0; k < 3; ++k) {
      S.lo[k] = desc[s].lo[k];
      S.hi[k] = desc[s].hi[k];
      S.span[k] = desc[s].hi[k] - desc[s].lo[k];
      S.periodic[k] = desc[s].periodic[k] ? 1 : 0;
      if (!(S.span[k] > 0.0)) {
        ctx->err = "rdk_set_systems: box extent must be positive";
        return RDK_ERR_ARG;
      }
      S.dims[k] = 1;
      S.cell[k] = S.span[k];
      S.inv_cell[k] = 1.0 / S.span[k];
    }
    if ((desc[s].floor_wall || desc[s].ceil_wall) && desc[s].periodic[2]) {
      ctx->err = "rdk_set_systems: plane walls need a non-periodic z axis";
      return RDK_ERR_ARG;
    }
    S.floor_wall = desc[s].floor_wall ? 1 : 0;
    S.ceil_wall = desc[s].ceil_wall ? 1 : 0;
    S.cell_off = 0;
    S.ytile = 1;
    S.mu = desc[s].mu;
  }
  RDK_HIP(hipMemcpyAsync(ctx->sys_of.p, ctx->h_sys_of.data(), ctx->n * sizeof(int), hipMemcpyHostToDevice,
                         ctx->stream));
  RDK_TRY(ensure(ctx, ctx->sys_dev, nsys));
  RDK_HIP(hipMemcpyAsync(ctx->sys_dev.p, ctx->sys_host.data(), nsys * sizeof(SysDev), hipMemcpyHostToDevice,
                         ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->built_chunk_len = -2.0;
  ctx->n_chunks = 0;
  return RDK_OK;
}

int rdk_set_physics(rdk_ctx* ctx, const rdk_physics* p) {
  CTX_CHECK();
  if (!p || !(p->dt > 0.0)) {
    ctx->err = "rdk_set_physics: dt must be positive";
    return RDK_ERR_ARG;
  }
  ctx->phys = *p;
  ctx->have_phys = true;
  return RDK_OK;
}

int rdk_upload_state(rdk_ctx* ctx, const double* x, const double* q, const double* v, const double* w) {
  CTX_CHECK();
  if (ctx->n <= 0) return RDK_ERR_STATE;
  if (x) RDK_TRY(upload_planar(ctx, ctx->x, 3, x, ctx->n));
  if (q) RDK_TRY(upload_planar(ctx, ctx->q, 4, q, ctx->n));
  if (v) RDK_TRY(upload_planar(ctx, ctx->v, 3, v, ctx->n));
  if (w) RDK_TRY(upload_planar(ctx, ctx->w, 3, w, ctx->n));
  return RDK_OK;
}

int rdk_upload_forces(rdk_ctx* ctx, const double* f, const double* tau) {
  CTX_CHECK();
  if (ctx->n <= 0) return RDK_ERR_STATE;
  if (f) RDK_TRY(upload_planar(ctx, ctx->f, 3, f, ctx->n));
  if (tau) RDK_TRY(upload_planar(ctx, ctx->tau, 3, tau, ctx->n));
  return RDK_OK;
}

int rdk_download_state(rdk_ctx* ctx, double* x, double* q, double* v, double* w, double* f, double* tau) {
  CTX_CHECK();
  if (ctx->n <= 0) return RDK_ERR_STATE;
  if (x) RDK_TRY(download_planar(ctx, ctx->x, 3, x, ctx->n));
  if (q) RDK_TRY(download_planar(ctx, ctx->q, 4, q, ctx->n));
  if (v) RDK_TRY(download_planar(ctx, ctx->v, 3, v, ctx->n));
  if (w) RDK_TRY(download_planar(ctx, ctx->w, 3, w, ctx->n));
  if (f) RDK_TRY(download_planar(ctx, ctx->f, 3, f, ctx->n));
  if (tau) RDK_TRY(download_planar(ctx, ctx->tau, 3, tau, ctx->n));
  return RDK_OK;
}

static int check_ready(rdk_ctx* ctx) {
  if (ctx->n <= 0 || ctx->nsys <= 0) {
    ctx->err = "bodies and systems must be set first";
    return RDK_ERR_STATE;
  }
  return RDK_OK;
}

int rdk_detect(rdk_ctx* ctx, const rdk_detect_params* p, int* n_contacts) {
  CTX_CHECK();
  if (!p || p->delta < 0.0) return RDK_ERR_ARG;
  RDK_TRY(check_ready(ctx));
  RDK_TRY(detect_run(ctx, p));
  ctx->have_manifolds = false;
  if (n_contacts) *n_contacts = ctx->nc;
  return prof_flush(ctx);
}

int rdk_set_pair_capacity(rdk_ctx* ctx, int cap) {
  if (!ctx || cap < 0) return RDK_ERR_ARG;
  ctx->pair_cap_forced = cap;
  if (cap > 0) {
    release(ctx->pkeys);
    release(ctx->pkeys_alt);
  }
  return RDK_OK;
}

int rdk_detect_info(rdk_ctx* ctx, int* n_chunks, int* n_cells, int64_t* cand, int* regrows) {
  if (!ctx) return RDK_ERR_ARG;
  if (n_chunks) *n_chunks = ctx->n_chunks;
  if (n_cells) *n_cells = ctx->n_cells;
  if (cand) *cand = ctx->last_candidates;
  if (regrows) *regrows = ctx->regrows;
  return RDK_OK;
}

}  // extern "C"
template <class T>
static int dl(rdk_ctx* ctx, T* dst, const T* src, int n) {
  if (!dst || n <= 0) return RDK_OK;
  RDK_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  return RDK_OK;
}
extern "C" {

static int dl3(rdk_ctx* ctx, double* dst, DevBuf<double>* src, int n) {
  if (!dst || n <= 0) return RDK_OK;
  std::vector<double> tmp(n);
  for (int k = 0; k < 3; ++k) {
    RDK_HIP(hipMemcpyAsync(tmp.data(), src[k].p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) dst[3 * i + k] = tmp[i];
  }
  return RDK_OK;
}

int rdk_get_contacts(rdk_ctx* ctx, int cap, int* a, int* b, double* pa, double* pb, double* nrm, double* dist,
                     double* shift, double* s, double* t, int* type) {
  CTX_CHECK();
  const int n = std::min(cap, ctx->nc);
  RDK_TRY(dl(ctx, a, ctx->ca.p, n));
  RDK_TRY(dl(ctx, b, ctx->cb.p, n));
  RDK_TRY(dl(ctx, type, ctx->ctype.p, n));
  RDK_TRY(dl(ctx, dist, ctx->cdist.p, n));
  RDK_TRY(dl(ctx, s, ctx->cs.p, n));
  RDK_TRY(dl(ctx, t, ctx->ct.p, n));
  RDK_TRY(dl3(ctx, pa, ctx->cpa, n));
  RDK_TRY(dl3(ctx, pb, ctx->cpb, n));
  RDK_TRY(dl3(ctx, nrm, ctx->cn, n));
  RDK_TRY(dl3(ctx, shift, ctx->csh, n));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

// FP64 contact list -> the plug-in's FP32 records
__global__ void k_pack_raw(int n, const int* __restrict__ a, const int* __restrict__ b, const double* __restrict__ r,
                           const double* const* __restrict__ pa, const double* const* __restrict__ pb,
                           const double* const* __restrict__ nrm, const double* const* __restrict__ sh,
                           const double* __restrict__ dist, const double* __restrict__ s,
                           const double* __restrict__ t, rdk_contact_raw* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  rdk_contact_raw c;
  c.a = a[i];
  c.b = b[i];
  c.px_a = (float)pa[0][i];
  c.py_a = (float)pa[1][i];
  c.pz_a = (float)pa[2][i];
  c.px_b = (float)pb[0][i];
  c.py_b = (float)pb[1][i];
  c.pz_b = (float)pb[2][i];
  c.nx = (float)nrm[0][i];
  c.ny = (float)nrm[1][i];
  c.nz = (float)nrm[2][i];
  c.dist = (float)dist[i];
  c.surface_limit = (float)(r[c.a] + r[c.b]);
  c.s = (float)s[i];
  c.t = (float)t[i];
  c.shift_bx = (float)sh[0][i];
  c.shift_by = (float)sh[1][i];
  c.shift_bz = (float)sh[2][i];
  out[i] = c;
}

int rdk_get_contacts_raw(rdk_ctx* ctx, int cap, rdk_contact_raw* out, int* count) {
  CTX_CHECK();
  if (cap < 0 || (cap > 0 && !out)) return RDK_ERR_ARG;
  const int nc = ctx->nc;
  if (count) *count = nc;
  const int n = std::min(cap, nc);
  if (n > 0) {
    const double* ptrs[12];
    for (int k = 0; k < 3; ++k) {
      ptrs[k] = ctx->cpa[k].p;
      ptrs[3 + k] = ctx->cpb[k].p;
      ptrs[6 + k] = ctx->cn[k].p;
      ptrs[9 + k] = ctx->csh[k].p;
    }
    RDK_TRY(ensure(ctx, ctx->raw_ptrs, 12));
    RDK_TRY(ensure(ctx, ctx->raw_out, (size_t)n * sizeof(rdk_contact_raw)));
    RDK_HIP(hipMemcpyAsync(ctx->raw_ptrs.p, ptrs, sizeof(ptrs), hipMemcpyHostToDevice, ctx->stream));
    const double* const* P = (const double* const*)ctx->raw_ptrs.p;
    rdk_contact_raw* dout = (rdk_contact_raw*)ctx->raw_out.p;
    hipLaunchKernelGGL(k_pack_raw, dim3(grid_for(n, 256)), dim3(256), 0, ctx->stream, n, ctx->ca.p, ctx->cb.p,
                       ctx->r.p, P, P + 3, P + 6, P + 9, ctx->cdist.p, ctx->cs.p, ctx->ct.p, dout);
    RDK_HIP(hipMemcpyAsync(out, dout, n * sizeof(rdk_contact_raw), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (cap < nc) {
    ctx->err = "rdk_get_contacts_raw: capacity below the contact count";
    return RDK_ERR_CAPACITY;
  }
  return RDK_OK;
}

int rdk_detect_capsule_pairs_raw(rdk_ctx* ctx, int n, const float* x, const float* q, const float* r,
                                 const float* h, float margin, int pbc, float sx, float sy, float sz, int cap,
                                 rdk_contact_raw* out, int* count) {
  CTX_CHECK();
  if (n == 0 && count && cap >= 0) {  // no capsules: no contacts
    *count = 0;
    return RDK_OK;
  }
  if (n < 0 || !x || !q || !r || !h || !count || cap < 0 || (cap > 0 && !out) || !(margin >= 0.0f) ||
      (pbc && !(sx > 0.0f && sy > 0.0f && sz > 0.0f))) {
    ctx->err = "rdk_detect_capsule_pairs_raw: invalid argument";
    return RDK_ERR_ARG;
  }
  std::vector<double> X(3 * (size_t)n), Q(4 * (size_t)n), R(n), H(n), M(n, 1.0), E(n, 0.0), Z(3 * (size_t)n, 0.0);
  double reach = 0.0;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) X[3 * i + k] = x[3 * i + k];
    for (int k = 0; k < 4; ++k) Q[4 * i + k] = q[4 * i + k];
    R[i] = r[i];
    H[i] = h[i];
    reach = std::max(reach, H[i] + R[i]);
  }
  RDK_TRY(rdk_set_bodies(ctx, n, X.data(), Q.data(), Z.data(), Z.data(), R.data(), H.data(), M.data(), E.data()));
  rdk_system_desc d{};
  if (pbc) {
    // only the extents matter for the minimum image; cells wrap any position
    const double L[3] = {sx, sy, sz};
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = 0.0;
      d.hi[k] = L[k];
      d.periodic[k] = 1;
    }
  } else {
    // a free box around the rods (positions beyond it fall in its edge cells)
    for (int k = 0; k < 3; ++k) {
      double lo = X[k], hi = X[k];
      for (int i = 1; i < n; ++i) {
        lo = std::min(lo, X[3 * i + k]);
        hi = std::max(hi, X[3 * i + k]);
      }
      d.lo[k] = lo - 2.0 * reach - margin - 1e-6;
      d.hi[k] = hi + 2.0 * reach + margin + 1e-6;
    }
  }
  const int off[2] = {0, n};
  RDK_TRY(rdk_set_systems(ctx, 1, off, &d));
  rdk_detect_params dp{(double)margin, 0.0, 1.0};
  int nc = 0;
  RDK_TRY(rdk_detect(ctx, &dp, &nc));
  return rdk_get_contacts_raw(ctx, cap, out, count);
}

int rdk_forces_soft(rdk_ctx* ctx, const rdk_soft_params* p, double* pe) {
  CTX_CHECK();
  if (!p || !(p->delta > 0.0) || !(p->nu > 0.0)) return RDK_ERR_ARG;
  RDK_TRY(check_ready(ctx));
  if (!ctx->have_phys) return RDK_ERR_STATE;
  RDK_TRY(forces_run(ctx, 0, p, pe));
  return prof_flush(ctx);
}

int rdk_forces_hm(rdk_ctx* ctx, const rdk_hm_params* p, double* pe) {
  CTX_CHECK();
  if (!p || !(p->E > 0.0) || !(p->restitution > 0.0)) return RDK_ERR_ARG;
  RDK_TRY(check_ready(ctx));
  if (!ctx->have_phys) return RDK_ERR_STATE;
  RDK_TRY(forces_run(ctx, 1, p, pe));
  return prof_flush(ctx);
}

int rdk_history_size(rdk_ctx* ctx, int* n) {
  if (!ctx || !n) return RDK_ERR_ARG;
  *n = ctx->nhist;
  return RDK_OK;
}

static int check_nsc(rdk_ctx* ctx, const rdk_nsc_params* p) {
  if (!p || !(p->omega > 0.0) || p->max_sweeps < 0 || p->friction_proj < 0 || p->friction_proj > 3 ||
      p->warm_mode < 0 || p->warm_mode > 4 || p->hot_set < 0 || p->hot_set > 2) {
    ctx->err = "invalid rdk_nsc_params";
    return RDK_ERR_ARG;
  }
  if (!ctx->have_phys) {
    ctx->err = "physics parameters not set";
    return RDK_ERR_STATE;
  }
  return RDK_OK;
}

int rdk_nsc_build(rdk_ctx* ctx, const rdk_nsc_params* p, int* nm) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  RDK_TRY(check_nsc(ctx, p));
  RDK_TRY(nsc_build_run(ctx, p));
  if (nm) *nm = ctx->nm;
  return prof_flush(ctx);
}

// Fields of manifolds 0 .. n-1 into rec (RDK_MF_FIELDS per manifold), from the
// device layout of field pairs (mf_off).
static int download_fields(rdk_ctx* ctx, const double* mf, size_t stride, int n, double* rec) {
  std::vector<double> tmp(2 * (size_t)n);
  for (int k = 0; k < RDK_MF_FIELDS / 2; ++k) {
    RDK_HIP(hipMemcpyAsync(tmp.data(), mf + mf_off(2 * k, 0, stride), 2 * (size_t)n * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
      rec[(size_t)RDK_MF_FIELDS * i + 2 * k] = tmp[2 * (size_t)i];
      rec[(size_t)RDK_MF_FIELDS * i + 2 * k + 1] = tmp[2 * (size_t)i + 1];
    }
  }
  return RDK_OK;
}

int rdk_nsc_get_manifolds(rdk_ctx* ctx, int cap, double* rec, int* ab, double* lambda) {
  CTX_CHECK();
  if (!ctx->have_manifolds) return RDK_ERR_STATE;
  const int n = std::min(cap, ctx->nm);
  if (n <= 0) return RDK_OK;
  const size_t stride = ctx->mf.cap / RDK_MF_FIELDS;
  if (rec) RDK_TRY(download_fields(ctx, ctx->mf.p, stride, n, rec));
  if (ab) {
    std::vector<int> a(n), b(n);
    RDK_TRY(dl(ctx, a.data(), ctx->ma.p, n));
    RDK_TRY(dl(ctx, b.data(), ctx->mb.p, n));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n; ++i) {
      ab[2 * i] = a[i];
      ab[2 * i + 1] = b[i];
    }
  }
  if (lambda) RDK_TRY(dl3(ctx, lambda, ctx->lam, n));
  return RDK_OK;
}

int rdk_nsc_set_lambda(rdk_ctx* ctx, int m, const double* lambda) {
  CTX_CHECK();
  if (!ctx->have_manifolds || m != ctx->nm || !lambda) return RDK_ERR_ARG;
  if (m == 0) return RDK_OK;
  RDK_TRY(upload_planar(ctx, ctx->lam, 3, lambda, m));
  ctx->have_coloring = false;
  return RDK_OK;
}

int rdk_nsc_solve(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st) {
  CTX_CHECK();
  RDK_TRY(check_nsc(ctx, p));
  int s = nsc_solve_run(ctx, p, st);
  int f = prof_flush(ctx);
  return s != RDK_OK ? s : f;
}

int rdk_nsc_sweeps(rdk_ctx* ctx, const rdk_nsc_params* p, int nsweeps) {
  CTX_CHECK();
  RDK_TRY(check_nsc(ctx, p));
  if (nsweeps < 0) return RDK_ERR_ARG;
  RDK_TRY(nsc_sweeps_run(ctx, p, nsweeps));
  return prof_flush(ctx);
}

int rdk_nsc_get_coloring(rdk_ctx* ctx, int cap, int* color, int* order, int* ncol) {
  CTX_CHECK();
  if (!ctx->have_manifolds) return RDK_ERR_STATE;
  if (!ctx->have_coloring) RDK_TRY(nsc_color_run(ctx, true));
  RDK_TRY(color_to_host(ctx));
  const int n = std::min(cap, ctx->nm);
  RDK_TRY(dl(ctx, color, ctx->color.p, n));
  RDK_TRY(dl(ctx, order, ctx->perm.p, n));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  if (ncol) *ncol = ctx->ncolor;
  return RDK_OK;
}

int rdk_nsc_residual(rdk_ctx* ctx, const rdk_nsc_params* p, double* res) {
  CTX_CHECK();
  RDK_TRY(check_nsc(ctx, p));
  if (!ctx->have_manifolds) return RDK_ERR_STATE;
  if (!ctx->have_coloring) RDK_TRY(nsc_color_run(ctx, false));
  // residual runs on the colour-ordered copy: refresh its lambda
  RDK_TRY(nsc_residual_run(ctx, p, res, nullptr));
  return prof_flush(ctx);
}

int rdk_set_solver_mode(rdk_ctx* ctx, int mode) {
  if (!ctx || mode < 0 || mode > 1) return RDK_ERR_ARG;
  ctx->solver_mode = mode;
  return RDK_OK;
}

int rdk_nsc_project(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* p, int* passes, int* n_first) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  if (!dp || !p || p->position_iters < 0 || p->position_psor_iters < 0 || !(p->slop >= 0.0)) return RDK_ERR_ARG;
  RDK_TRY(nsc_project_run(ctx, dp, p, passes, n_first));
  return prof_flush(ctx);
}

int rdk_nsc_island_info(rdk_ctx* ctx, int* n_islands, int* counts4) {
  if (!ctx) return RDK_ERR_ARG;
  if (n_islands) *n_islands = ctx->nislands;
  if (counts4)
    for (int k = 0; k < 4; ++k) counts4[k] = ctx->island_counts[k];
  return RDK_OK;
}

int rdk_nsc_get_solve_order(rdk_ctx* ctx, int cap, int* order) {
  CTX_CHECK();
  if (!order) return RDK_ERR_ARG;
  const int n = std::min(cap, ctx->nm);
  RDK_TRY(dl(ctx, order, ctx->perm.p, n));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_integrate(rdk_ctx* ctx, int stage) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  RDK_TRY(integrate_run(ctx, stage));
  return prof_flush(ctx);
}

// copy of the natural-order manifolds, their bodies and impulses after the
// velocity solve of a fused step
static int keep_manifolds(rdk_ctx* ctx) {
  const int m = ctx->nm;
  ctx->kept_m = m;
  ctx->kept_stride = ctx->mf.cap / RDK_MF_FIELDS;
  if (m <= 0) return RDK_OK;
  RDK_TRY(ensure(ctx, ctx->kept_mf, ctx->mf.cap));
  RDK_TRY(ensure(ctx, ctx->kept_ab, 2 * (size_t)m));
  RDK_TRY(ensure(ctx, ctx->kept_lam, 3 * (size_t)m));
  RDK_HIP(hipMemcpyAsync(ctx->kept_mf.p, ctx->mf.p, ctx->mf.cap * sizeof(double), hipMemcpyDeviceToDevice,
                         ctx->stream));
  RDK_HIP(hipMemcpyAsync(ctx->kept_ab.p, ctx->ma.p, m * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
  RDK_HIP(hipMemcpyAsync(ctx->kept_ab.p + m, ctx->mb.p, m * sizeof(int), hipMemcpyDeviceToDevice, ctx->stream));
  for (int k = 0; k < 3; ++k)
    RDK_HIP(hipMemcpyAsync(ctx->kept_lam.p + (size_t)k * m, ctx->lam[k].p, m * sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream));
  return RDK_OK;
}

int rdk_keep_manifolds(rdk_ctx* ctx, int on) {
  CTX_CHECK();
  ctx->keep_manifolds = on != 0;
  ctx->kept_m = 0;
  return RDK_OK;
}

int rdk_nsc_get_kept_manifolds(rdk_ctx* ctx, int cap, double* rec, int* ab, double* lambda, int* count) {
  CTX_CHECK();
  if (cap < 0 || !count) return RDK_ERR_ARG;
  const int m = ctx->kept_m;
  *count = m;
  const int n = std::min(cap, m);
  if (n > 0) {
    std::vector<double> tmp(n);
    if (rec) RDK_TRY(download_fields(ctx, ctx->kept_mf.p, ctx->kept_stride, n, rec));
    if (ab) {
      std::vector<int> a(n), b(n);
      RDK_HIP(hipMemcpyAsync(a.data(), ctx->kept_ab.p, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      RDK_HIP(hipMemcpyAsync(b.data(), ctx->kept_ab.p + m, n * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
      RDK_HIP(hipStreamSynchronize(ctx->stream));
      for (int i = 0; i < n; ++i) {
        ab[2 * i] = a[i];
        ab[2 * i + 1] = b[i];
      }
    }
    if (lambda) {
      for (int k = 0; k < 3; ++k) {
        RDK_HIP(hipMemcpyAsync(tmp.data(), ctx->kept_lam.p + (size_t)k * m, n * sizeof(double),
                               hipMemcpyDeviceToHost, ctx->stream));
        RDK_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; ++i) lambda[3 * i + k] = tmp[i];
      }
    }
  }
  if (cap < m) {
    ctx->err = "rdk_nsc_get_kept_manifolds: capacity below the manifold count";
    return RDK_ERR_CAPACITY;
  }
  return RDK_OK;
}

int rdk_record_stage_energies(rdk_ctx* ctx, int on) {
  CTX_CHECK();
  ctx->stage_ke = on != 0;
  for (double& e : ctx->ke_stage) e = 0.0;
  return RDK_OK;
}

int rdk_last_potential_energy(rdk_ctx* ctx, double* pe) {
  CTX_CHECK();
  if (!pe) return RDK_ERR_ARG;
  *pe = 0.0;
  if (ctx->pe.p) {
    RDK_HIP(hipMemcpyAsync(pe, ctx->pe.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
  }
  return RDK_OK;
}

int rdk_get_stage_energies(rdk_ctx* ctx, double* out) {
  CTX_CHECK();
  if (!out) return RDK_ERR_ARG;
  for (int k = 0; k < 3; ++k) out[k] = ctx->ke_stage[k];
  return RDK_OK;
}

int rdk_keep_soft_forces(rdk_ctx* ctx, int on) {
  CTX_CHECK();
  ctx->keep_soft = on != 0;
  ctx->soft_kept_n = 0;
  return RDK_OK;
}

int rdk_get_soft_forces(rdk_ctx* ctx, int cap, double* out, int* count) {
  CTX_CHECK();
  if (cap < 0 || !count || (cap > 0 && !out)) return RDK_ERR_ARG;
  const int m = ctx->soft_kept_n;
  *count = m;
  const int n = std::min(cap, m);
  if (n > 0) {
    RDK_HIP(hipMemcpyAsync(out, ctx->soft_kept.p, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (cap < m) {
    ctx->err = "rdk_get_soft_forces: capacity below the contact count";
    return RDK_ERR_CAPACITY;
  }
  return RDK_OK;
}

// One eager hard-contact step (App::physicsStep, main.cpp:4333-4858): every
// stage sized from counts read back to the host.
static int eager_nsc_step(rdk_ctx* ctx, const rdk_detect_params* d, const rdk_nsc_params* np, rdk_step_stats& S) {
  ctx->eager_nsc_steps++;
  RDK_TRY(integrate_run(ctx, RDK_NSC_FREE));
  if (ctx->stage_ke) RDK_TRY(observables_run(ctx, &ctx->ke_stage[0], nullptr));  // main.cpp:4411
  RDK_TRY(detect_run(ctx, d));
  RDK_TRY(nsc_build_run(ctx, np));
  rdk_solve_stats ss;
  std::memset(&ss, 0, sizeof(ss));
  int rc = nsc_solve_run(ctx, np, &ss);
  if (rc == RDK_ERR_NOCONV) S.nonconverged_steps++;
  else if (rc != RDK_OK) return rc;
  if (ctx->keep_manifolds) RDK_TRY(keep_manifolds(ctx));
  if (ctx->stage_ke) RDK_TRY(observables_run(ctx, &ctx->ke_stage[1], nullptr));  // main.cpp:4461
  RDK_TRY(integrate_run(ctx, RDK_NSC_POS));
  if (np->position_stabilization && np->position_iters > 0)
    RDK_TRY(nsc_project_run(ctx, d, np, nullptr, nullptr));  // main.cpp:4497-4503
  if (ctx->stage_ke) RDK_TRY(observables_run(ctx, &ctx->ke_stage[2], nullptr));  // main.cpp:4505,4519
  S.steps++;
  S.contacts_total += ss.rod_contacts;
  S.manifolds_total += ss.manifolds;
  S.sweeps_total += ss.sweeps;
  S.colors_total += ss.colors;
  S.residual_max = std::max(S.residual_max, ss.residual);
  ctx->work[W_STEPS] += 1;
  if (ctx->profiling) RDK_TRY(prof_flush(ctx));
  return RDK_OK;
}

// ---- fixed-capacity step (no reference counterpart) -------------------------
// The step of eager_nsc_step with every launch sized for reserved capacities
// and every count read on the device (detect_fixed_enqueue,
// nsc_fixed_enqueue), so it can be captured once into a graph and replayed.
// A step that fails
// (see FxErr) does so before its first state change and makes every later
// state-changing kernel return; the host then runs that step eagerly (which
// regrows what it must) and goes on with the fixed path.

static bool fixed_supported(rdk_ctx* ctx, const rdk_nsc_params* np) {
  return !(np->position_stabilization && np->position_iters > 0) && np->warm_mode <= 1 && ctx->solver_mode == 0 &&
         !ctx->stage_ke && !ctx->keep_manifolds && ctx->n < (1 << 25);
}

static void drop_step_graph(rdk_ctx* ctx) {
  if (ctx->step_exec) (void)hipGraphExecDestroy(ctx->step_exec);
  if (ctx->step_graph) (void)hipGraphDestroy(ctx->step_graph);
  ctx->step_exec = nullptr;
  ctx->step_graph = nullptr;
  ctx->step_graph_key.clear();
}

// One fixed step: detection (positions only), then nsc_fixed_enqueue (checks,
// free flight, manifolds, solve, warm cache), then the position update.  The
// free-flight update does not move the rods, so detecting first finds the
// contacts of the eager step's order (free flight, then detection).
static int fixed_enqueue_step(rdk_ctx* ctx, const rdk_detect_params* d, const rdk_nsc_params* np) {
  RDK_TRY(fixed_clear_enqueue(ctx));
  RDK_TRY(detect_fixed_enqueue(ctx, d));
  RDK_TRY(nsc_fixed_enqueue(ctx, np));
  RDK_TRY(integrate_run(ctx, RDK_NSC_POS, ctx->fx.p + FX_ERR));
  return RDK_OK;
}

// the parameters a captured step bakes in
static std::vector<char> step_key(rdk_ctx* ctx, const rdk_detect_params* d, const rdk_nsc_params* np) {
  std::vector<char> k(sizeof(*d) + sizeof(*np) + sizeof(ctx->phys) + 12 * sizeof(int));
  char* o = k.data();
  std::memcpy(o, d, sizeof(*d));
  o += sizeof(*d);
  std::memcpy(o, np, sizeof(*np));
  o += sizeof(*np);
  std::memcpy(o, &ctx->phys, sizeof(ctx->phys));
  o += sizeof(ctx->phys);
  // the class grids, and which classes have held islands (the streams they run on)
  int seen = 0;
  for (int c = 0; c < 4; ++c) seen |= (ctx->cls_seen_max[c] > 0 ? 1 : 0) << c;
  const int v[12] = {ctx->n,         ctx->fx_pair_cap, ctx->fx_rounds,  ctx->nsys,       ctx->n_cells,    ctx->n_chunks,
                     ctx->fx_wall_cap, ctx->fx_grid[0],  ctx->fx_grid[1], ctx->fx_grid[2], ctx->fx_grid[3], seen};
  std::memcpy(o, v, sizeof(v));
  return k;
}

static int fixed_capture(rdk_ctx* ctx, const rdk_detect_params* d, const rdk_nsc_params* np) {
  drop_step_graph(ctx);
  int64_t work[W_NUM];
  std::memcpy(work, ctx->work, sizeof(work));
  ctx->capturing = true;
  int rc = RDK_OK;
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    ctx->capturing = false;
    ctx->err = "step capture: hipStreamBeginCapture failed";
    return RDK_ERR_HIP;
  }
  rc = fixed_enqueue_step(ctx, d, np);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->capturing = false;
  std::memcpy(ctx->work, work, sizeof(work));
  if (rc != RDK_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess || !g) {
    ctx->err = std::string("step capture: hipStreamEndCapture: ") + hipGetErrorString(e);
    return RDK_ERR_HIP;
  }
  ctx->step_graph = g;
  if (hipGraphInstantiate(&ctx->step_exec, g, nullptr, nullptr, 0) != hipSuccess) {
    drop_step_graph(ctx);
    ctx->err = "step capture: hipGraphInstantiate failed";
    return RDK_ERR_HIP;
  }
  ctx->step_graph_key = step_key(ctx, d, np);
  ctx->step_graph_gen = ctx->alloc_gen;
  ctx->graph_captures++;
  return RDK_OK;
}

// Host warm-cache sizes -> device words (before fixed steps follow eager ones).
// The warm-cache sizes the fixed step reads from fx: uploaded when the host's
// differ from what fx holds (after an eager step or a new fx), from pinned
// memory the host does not touch again before the next synchronisation.
static int fixed_upload_warm_counts(rdk_ctx* ctx) {
  if (ctx->fx_prev_dev[0] == ctx->nprev && ctx->fx_prev_dev[1] == ctx->nprev_w) return RDK_OK;
  if (!ctx->h_fx_up) RDK_HIP(hipHostMalloc((void**)&ctx->h_fx_up, 2 * sizeof(int), hipHostMallocDefault));
  ctx->h_fx_up[0] = ctx->nprev;
  ctx->h_fx_up[1] = ctx->nprev_w;
  RDK_HIP(hipMemcpyAsync(ctx->fx.p + FX_NPREV, ctx->h_fx_up, 2 * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  ctx->fx_prev_dev[0] = ctx->nprev;
  ctx->fx_prev_dev[1] = ctx->nprev_w;
  return RDK_OK;
}

static int step_nsc_fixed(rdk_ctx* ctx, const rdk_detect_params* d, const rdk_nsc_params* np, int nsteps,
                          rdk_step_stats& S) {
  const bool graph = ctx->step_graph_mode == 2 && graphs_enabled();  // RDK_GRAPHS=0: the same launches, uncaptured
  const int chunk_max = graph ? 16 : 1;
  int done = 0;
  bool uploaded = false;
  int backoff = 0, eager_left = 0;  // steps left to run eagerly after a failed fixed step
  while (done < nsteps) {
    if (eager_left > 0) {  // repeated failures (e.g. a percolating island): stay eager for a while
      RDK_TRY(eager_nsc_step(ctx, d, np, S));
      --eager_left;
      ++done;
      continue;
    }
    if (ctx->eager_nsc_steps == 0) {  // one eager step measures the counts the capacities follow
      RDK_TRY(eager_nsc_step(ctx, d, np, S));
      ++done;
      continue;
    }
    // capacities: 1.25 times the most contacts (wall manifolds) an eager step
    // of this context has built, rounded up to a power of two (at least 1024 /
    // 256), or the requested pair capacity; a step that needs more is redone
    // eagerly, which raises them.  Tight capacities keep the sorts of a step on
    // their single-workgroup path as long as the eager step's are.
    auto pow2 = [](int64_t v) {
      int64_t p = 1;
      while (p < v) p <<= 1;
      return (int)std::min<int64_t>(p, 1 << 30);
    };
    const int cap = ctx->step_graph_pair_cap > 0 ? ctx->step_graph_pair_cap
                                                 : pow2(std::max<int64_t>(1024, ctx->nc_seen_max * 5 / 4 + 1));
    const int wcap = pow2(std::max<int64_t>(256, ctx->nwall_seen_max * 5 / 4 + 1));
    int grids[4];
    fixed_class_grids(ctx, grids);
    if (cap != ctx->fx_pair_cap || wcap != ctx->fx_wall_cap || !ctx->fx.p || ctx->n_chunks <= 0 ||
        ctx->grid_dirty || d->chunk_len != ctx->built_chunk_len || 2 * ctx->color_rounds_max + 4 > ctx->fx_rounds ||
        std::memcmp(grids, ctx->fx_grid, sizeof(grids)) != 0) {
      RDK_TRY(fixed_prepare(ctx, d, np, cap, wcap));
    }
    if (!uploaded) {
      RDK_TRY(fixed_upload_warm_counts(ctx));
      uploaded = true;
    }
    RDK_HIP(hipMemsetAsync(ctx->fx.p, 0, 2 * sizeof(int), ctx->stream));  // FX_ERR, FX_STEPS
    RDK_HIP(hipMemsetAsync(ctx->facc.p, 0, FA_NUM * sizeof(double), ctx->stream));
    const int chunk = std::min(chunk_max, nsteps - done);
    if (graph) {
      if (!ctx->step_exec || ctx->step_graph_gen != ctx->alloc_gen || ctx->step_graph_key != step_key(ctx, d, np))
        RDK_TRY(fixed_capture(ctx, d, np));
      for (int k = 0; k < chunk; ++k) RDK_HIP(hipGraphLaunch(ctx->step_exec, ctx->stream));
    } else {
      int64_t work[W_NUM];
      std::memcpy(work, ctx->work, sizeof(work));
      ctx->capturing = true;  // same rules as a capture: no allocation, no timing
      const int rc = fixed_enqueue_step(ctx, d, np);
      ctx->capturing = false;
      std::memcpy(ctx->work, work, sizeof(work));
      RDK_TRY(rc);
    }
    RDK_TRY(fixed_report_enqueue(ctx));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    const FxReport& R = *ctx->h_report;
    const int* fh = R.fx;
    const double* acc = R.acc;
    fixed_apply_counts(ctx, R);  // the failed step of a chunk changed no state: its counts are the last good ones
    const int ok = fh[FX_STEPS];
    S.steps += ok;
    S.contacts_total += (int64_t)acc[FA_CONTACTS];
    S.manifolds_total += (int64_t)acc[FA_MANIFOLDS];
    S.sweeps_total += (int64_t)acc[FA_SWEEPS];
    S.colors_total += (int64_t)acc[FA_COLORS];
    S.nonconverged_steps += (int64_t)acc[FA_NONCONV];
    if (ok > 0) S.residual_max = std::max(S.residual_max, acc[FA_RESMAX]);
    ctx->work[W_STEPS] += ok;
    ctx->work[W_BODY_PASSES] += ok;
    ctx->work[W_CHUNKS] += (int64_t)ok * ctx->n_chunks;
    ctx->work[W_CANDIDATES] += (int64_t)acc[FA_CANDIDATES];
    ctx->work[W_NARROW] += (int64_t)acc[FA_NARROW];
    ctx->work[W_CONTACTS] += (int64_t)acc[FA_CONTACTS];
    ctx->work[W_MANIFOLDS] += (int64_t)acc[FA_MANIFOLDS];
    ctx->work[W_SWEEP_CONTACTS] += (int64_t)acc[FA_MSWEEPS];
    ctx->work[W_LDS_SWEEPS] += (int64_t)acc[FA_LDS_SWEEPS];
    ctx->work[W_HOT_SWEEPS] += (int64_t)acc[FA_HOT_SWEEPS];
    ctx->work[W_INTEGRATE_BODIES] += (int64_t)2 * ok * ctx->n;
    for (int c = 0; c < 4; ++c) ctx->cls_seen_max[c] = std::max(ctx->cls_seen_max[c], (int)acc[FA_CLS0 + c]);
    ctx->graph_steps += ok;
    done += ok;
    if (fh[FX_ERR] != 0) {
      // the failed step changed no state: redo it eagerly (regrowing what it must)
      RDK_TRY(eager_nsc_step(ctx, d, np, S));
      ctx->graph_fallbacks++;
      done += 1;
      uploaded = false;
      eager_left = backoff;
      backoff = std::min(256, std::max(1, 2 * backoff));
    } else if (ok != chunk) {
      ctx->err = "fixed step: fewer steps completed than launched";
      return RDK_ERR_STATE;
    } else {
      backoff = 0;
    }
  }
  return RDK_OK;
}

int rdk_step_nsc(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* np, int nsteps,
                 rdk_step_stats* st) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  RDK_TRY(check_nsc(ctx, np));
  if (!dp || nsteps < 0) return RDK_ERR_ARG;
  rdk_detect_params d = *dp;
  d.delta = 0.0;  // NSC admits contacts at strict overlap (nsc_contact.cpp:24)
  rdk_step_stats S;
  std::memset(&S, 0, sizeof(S));
  if (ctx->step_graph_mode != 0 && fixed_supported(ctx, np) && nsteps > 0) {
    RDK_TRY(step_nsc_fixed(ctx, &d, np, nsteps, S));
  } else {
    for (int s = 0; s < nsteps; ++s) RDK_TRY(eager_nsc_step(ctx, &d, np, S));
  }
  RDK_TRY(prof_flush(ctx));
  if (st) *st = S;
  return RDK_OK;
}

int rdk_set_tuning(rdk_ctx* ctx, int what, int value) {
  CTX_CHECK();
  if (what == RDK_TUNE_SMALL_M) {
    if (value < -1) return RDK_ERR_ARG;
    ctx->small_m = value;
    drop_step_graph(ctx);
    return RDK_OK;
  }
  if (what == RDK_TUNE_BROAD_BLOCK) {
    if (value != 64 && value != 128 && value != 256 && value != 512) return RDK_ERR_ARG;
    ctx->bn_threads = value;
    drop_step_graph(ctx);
    return RDK_OK;
  }
  if (what == RDK_TUNE_BROAD_STAGE) {
    if (value != 0 && value != 1) return RDK_ERR_ARG;
    ctx->bn_stage = value;
    drop_step_graph(ctx);
    return RDK_OK;
  }
  if (what == RDK_TUNE_CELL_BAND) {
    if (value < 0) return RDK_ERR_ARG;
    ctx->cell_band = value;
    ctx->grid_dirty = true;
    drop_step_graph(ctx);
    return RDK_OK;
  }
  return RDK_ERR_ARG;
}

int rdk_set_step_graph(rdk_ctx* ctx, int mode, int pair_capacity) {
  CTX_CHECK();
  if (mode < 0 || mode > 2 || pair_capacity < 0) return RDK_ERR_ARG;
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  if (mode != 2) drop_step_graph(ctx);
  ctx->step_graph_mode = mode;
  ctx->step_graph_pair_cap = pair_capacity;
  return RDK_OK;
}

int rdk_step_graph_info(rdk_ctx* ctx, int64_t* fixed_steps, int64_t* fallbacks, int64_t* captures) {
  CTX_CHECK();
  if (fixed_steps) *fixed_steps = ctx->graph_steps;
  if (fallbacks) *fallbacks = ctx->graph_fallbacks;
  if (captures) *captures = ctx->graph_captures;
  return RDK_OK;
}

int rdk_step_vv(rdk_ctx* ctx, const rdk_detect_params* dp, int model, const void* params, int nsteps,
                rdk_step_stats* st) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  if (!dp || !params || nsteps < 0 || model < 0 || model > 1) return RDK_ERR_ARG;
  if (!ctx->have_phys) return RDK_ERR_STATE;
  rdk_detect_params d = *dp;
  d.delta = model == 0 ? ((const rdk_soft_params*)params)->delta : 0.0;
  rdk_step_stats S;
  std::memset(&S, 0, sizeof(S));
  for (int s = 0; s < nsteps; ++s) {
    for (int pass = 0; pass < 2; ++pass) {
      RDK_TRY(detect_run(ctx, &d));
      RDK_TRY(forces_run(ctx, model, params, nullptr));
      S.contacts_total += ctx->nc;
      if (pass == 0) {
        RDK_TRY(integrate_run(ctx, RDK_VV_HALF1));
        RDK_TRY(integrate_run(ctx, RDK_CLEAR));
      } else {
        RDK_TRY(integrate_run(ctx, RDK_VV_HALF2));
      }
    }
    S.steps++;
    ctx->work[W_STEPS] += 1;
    if (ctx->profiling) RDK_TRY(prof_flush(ctx));
  }
  RDK_TRY(prof_flush(ctx));
  if (st) *st = S;
  return RDK_OK;
}

int rdk_affine_scale(rdk_ctx* ctx, double s) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  if (!(s > 0.0) || !std::isfinite(s)) {
    ctx->err = "rdk_affine_scale: the factor must be positive and finite";
    return RDK_ERR_ARG;
  }
  RDK_TRY(affine_run(ctx, s));
  return prof_flush(ctx);
}

int rdk_observables(rdk_ctx* ctx, double* ke, double* phi) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  RDK_TRY(observables_run(ctx, ke, phi));
  return prof_flush(ctx);
}

int rdk_entanglement(rdk_ctx* ctx, double cutoff, double* sums, int64_t* pairs) {
  CTX_CHECK();
  RDK_TRY(check_ready(ctx));
  if (!(cutoff == cutoff)) {
    ctx->err = "rdk_entanglement: cutoff is NaN";
    return RDK_ERR_ARG;
  }
  RDK_TRY(entanglement_run(ctx, cutoff, sums, pairs));
  return prof_flush(ctx);
}

int rdk_populate_nonoverlap(rdk_ctx* ctx, int n, float length, float diameter, const float* bmin,
                            const float* bmax, int periodic, float cell_size, uint32_t seed, int max_attempts,
                            int window, float* centres, float* dirs, float* quats, int* placed,
                            int64_t* attempts) {
  CTX_CHECK();
  if (n < 0 || !bmin || !bmax || !placed || (n > 0 && (!centres || !dirs)) || max_attempts < 1 || window < 0 ||
      !(length > 0.0f) || !(diameter > 0.0f) || !(cell_size == cell_size)) {
    ctx->err = "rdk_populate_nonoverlap: invalid argument";
    return RDK_ERR_ARG;
  }
  *placed = 0;
  if (attempts) *attempts = 0;
  RDK_TRY(populate_nonoverlap_run(ctx, n, length, diameter, bmin, bmax, periodic, cell_size, seed, max_attempts,
                                  window, centres, dirs, quats, placed, attempts));
  return prof_flush(ctx);
}

int rdk_set_profiling(rdk_ctx* ctx, int on) {
  if (!ctx) return RDK_ERR_ARG;
  RDK_TRY(prof_flush(ctx));
  ctx->profiling = on != 0;
  for (int k = 0; k < K_NUM; ++k) {
    ctx->k_launches[k] = 0;
    ctx->k_ms[k] = 0.0;
  }
  for (int k = 0; k < W_NUM; ++k) ctx->work[k] = 0;
  return RDK_OK;
}

int rdk_set_profile_mask(rdk_ctx* ctx, uint64_t mask) {
  if (!ctx) return RDK_ERR_ARG;
  RDK_TRY(prof_flush(ctx));
  ctx->prof_mask = mask;
  return RDK_OK;
}

int rdk_kernel_times(rdk_ctx* ctx, int kernel, int64_t* launches, double* total_ms) {
  if (!ctx || kernel < 0 || kernel >= K_NUM) return RDK_ERR_ARG;
  RDK_TRY(prof_flush(ctx));
  if (launches) *launches = ctx->k_launches[kernel];
  if (total_ms) *total_ms = ctx->k_ms[kernel];
  return RDK_OK;
}

const char* rdk_kernel_name(int kernel) {
  if (kernel < 0 || kernel >= K_NUM) return "";
  return kKernelNames[kernel];
}

int rdk_num_kernels(void) { return K_NUM; }

// Device buffers of the sort/scan utilities: allocated per call (a test and
// tooling path, not a step path).
struct TmpDev {
  void* p = nullptr;
  ~TmpDev() {
    if (p) (void)hipFree(p);
  }
};

static int util_count(rdk_ctx* ctx, int n, int cap, TmpDev& dn, Count& c) {
  if (n < 0 || (cap != 0 && cap < n)) return RDK_ERR_ARG;
  if (cap > n) {
    RDK_HIP(hipMalloc(&dn.p, sizeof(int)));
    RDK_HIP(hipMemcpyAsync(dn.p, &n, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    c = Count::dev((const int*)dn.p, cap);
  } else {
    c = Count::host(n);
  }
  return RDK_OK;
}

int rdk_sort_pairs_u32(rdk_ctx* ctx, int n, int cap, uint32_t* keys, int32_t* vals, int key_bits) {
  CTX_CHECK();
  if (!keys || !vals || key_bits < 1 || key_bits > 32) return RDK_ERR_ARG;
  TmpDev dn, buf;
  Count c;
  RDK_TRY(util_count(ctx, n, cap, dn, c));
  const int m = std::max(1, std::max(n, cap));
  RDK_HIP(hipMalloc(&buf.p, (size_t)m * 16));
  uint32_t* k0 = (uint32_t*)buf.p;
  uint32_t* k1 = k0 + m;
  int* v0 = (int*)(k1 + m);
  int* v1 = v0 + m;
  RDK_HIP(hipMemcpyAsync(k0, keys, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  RDK_HIP(hipMemcpyAsync(v0, vals, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  bool alt = false;
  RDK_TRY(sort_pairs_u32(ctx, k0, k1, v0, v1, c, key_bits, alt, nullptr));
  RDK_HIP(hipMemcpyAsync(keys, alt ? k1 : k0, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipMemcpyAsync(vals, alt ? v1 : v0, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_sort_pair_keys_u64(rdk_ctx* ctx, int n, int cap, uint64_t* keys, int index_bits) {
  CTX_CHECK();
  if (!keys || index_bits < 1 || index_bits > 32) return RDK_ERR_ARG;
  TmpDev dn, buf;
  Count c;
  RDK_TRY(util_count(ctx, n, cap, dn, c));
  const int m = std::max(1, std::max(n, cap));
  RDK_HIP(hipMalloc(&buf.p, (size_t)m * 16));
  uint64_t* k0 = (uint64_t*)buf.p;
  uint64_t* k1 = k0 + m;
  RDK_HIP(hipMemcpyAsync(k0, keys, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  bool alt = false;
  RDK_TRY(sort_pair_keys_u64(ctx, k0, k1, c, index_bits, alt, nullptr));
  RDK_HIP(hipMemcpyAsync(keys, alt ? k1 : k0, (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_exclusive_scan_i32(rdk_ctx* ctx, int n, int cap, const int32_t* in, int32_t* out) {
  CTX_CHECK();
  if (!in || !out) return RDK_ERR_ARG;
  TmpDev dn, buf;
  Count c;
  RDK_TRY(util_count(ctx, n, cap, dn, c));
  const int m = std::max(1, std::max(n, cap));
  RDK_HIP(hipMalloc(&buf.p, (size_t)m * 8));
  int* a = (int*)buf.p;
  int* b = a + m;
  RDK_HIP(hipMemcpyAsync(a, in, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
  RDK_TRY(exclusive_scan_int(ctx, a, b, c, nullptr));
  RDK_HIP(hipMemcpyAsync(out, b, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
  RDK_HIP(hipStreamSynchronize(ctx->stream));
  return RDK_OK;
}

int rdk_work_counters(rdk_ctx* ctx, int64_t* counters, int n) {
  if (!ctx || !counters) return RDK_ERR_ARG;
  for (int k = 0; k < n && k < W_NUM; ++k) counters[k] = ctx->work[k];
  return RDK_OK;
}

}  // extern "C"

// rdk_internal.hpp -- device context of librdk (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/rdk.h"
#include "rdk_math.hpp"

namespace rdk {

// Kernel ids for the per-kernel event timers.
enum KernelId {
  K_CHUNK_KEYS = 0,
  K_CELL_RANGES,
  K_CHUNK_GATHER,
  K_BROAD_NARROW,
  K_CONTACT_RECORD,
  K_SORT,
  K_MANIFOLD,
  K_WALL,
  K_CSR,
  K_COLOR,
  K_PERMUTE,
  K_WARM_APPLY,
  K_PSOR_SWEEP,
  K_PSOR_SMALL,
  K_RESIDUAL,
  K_NSC_FREE,
  K_NSC_POS,
  K_VV_HALF1,
  K_VV_HALF2,
  K_FORCE,
  K_BODY_REDUCE,
  K_HISTORY,
  K_OBSERVE,
  K_MLP,
  K_ISLAND,
  K_ISL_TINY,
  K_ISL_SMALL,
  K_ISL_MEDIUM,
  K_ISL_LARGE,
  K_ISL_HUGE,
  K_POSPROJ,
  K_POPULATE,
  K_MISC,
  K_NUM
};

// Work counters (algorithmic units) accumulated by the fused steps.
enum WorkId {
  W_STEPS = 0,
  W_BODY_PASSES,      // detection passes over all bodies
  W_CHUNKS,           // chunk entries binned
  W_CANDIDATES,       // chunk pairs passing the bounding-sphere test
  W_CONTACTS,         // contact records emitted
  W_MANIFOLDS,        // manifolds built
  W_SWEEP_CONTACTS,   // manifold updates in PSOR sweeps
  W_RESID_CONTACTS,   // manifold evaluations in residual passes
  W_INTEGRATE_BODIES, // body updates in integrator passes
  W_FORCE_CONTACTS,   // soft/HM force evaluations
  W_NARROW,           // FP64 narrow-phase calls (pairs passing the separating-axis test)
  W_PSOR_TIMED,       // manifold updates in k_psor_sweep launches whose time was recorded
  W_LDS_SWEEPS,       // manifold updates in sweeps of the LDS island class
  W_HOT_SWEEPS,       // sweeps of hot sets in the LDS island class, summed over islands
  W_NUM
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
};

// Element count of a launch: a host value (p == nullptr: cap), or a device
// value read when the kernel runs, mul * *p + add clamped to [0, cap], so a
// launch sized for cap can be enqueued (or captured) before the count exists.
struct Count {
  const int* p = nullptr;
  int mul = 1, add = 0, cap = 0;
  int skip_below = 0;  // a device count below this reads as 0 (another launch handles it)
  static Count host(int n) {
    Count c;
    c.cap = n;
    return c;
  }
  static Count dev(const int* ptr, int cap, int mul = 1, int add = 0) {
    Count c;
    c.p = ptr;
    c.cap = cap;
    c.mul = mul;
    c.add = add;
    return c;
  }
  __device__ __forceinline__ int get() const {
    if (!p) return cap;
    const long long v = (long long)mul * (long long)(*p) + add;
    const int c = v < 0 ? 0 : (v > cap ? cap : (int)v);
    return c < skip_below ? 0 : c;
  }
};

// Per-system constants on the device.
struct SysDev {
  double lo[3], hi[3], span[3], cell[3], inv_cell[3];
  int periodic[3], dims[3];
  int cell_off, floor_wall, ceil_wall;
  int ytile;      // grid rows (y) per band of the cell order; dims[1] when the order is not banded
  float fmargin;  // FP32 broad-phase prefilter margin
  double mu;
};

struct BodyPtrs {
  double *x[3], *q[4], *v[3], *w[3], *f[3], *tau[3];
  const double *r, *h, *mass, *inv_mass, *i_perp, *i_axial, *i_perp_inv, *i_axial_inv, *rest;
  const int* sys;
};

struct EventRec {
  int kernel;
  hipEvent_t a, b;
};

struct FxReport;  // fixed-step report (below)

// Manifold field storage: fields 2k and 2k + 1 of a manifold side by side,
// pair planes of `stride` manifolds each, so one 16-byte load fetches two
// fields of one manifold and consecutive manifolds are consecutive in memory.
__host__ __device__ __forceinline__ size_t mf_off(int field, size_t i, size_t stride) {
  return (((size_t)(field >> 1) * stride + i) << 1) + (size_t)(field & 1);
}

}  // namespace rdk

struct rdk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // side streams (default priority): island classes run concurrently
  hipStream_t stream3 = nullptr, stream4 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_join3 = nullptr, ev_join4 = nullptr;
  std::string err;

  // ---- bodies (planar FP64) ----
  int n = 0;
  rdk::DevBuf<double> x[3], q[4], v[3], w[3], f[3], tau[3];
  rdk::DevBuf<double> r, h, mass, inv_mass, i_perp, i_axial, i_perp_inv, i_axial_inv, rest;
  rdk::DevBuf<int> sys_of;
  std::vector<double> h_r, h_h, h_mass;
  std::vector<int> h_sys_of;

  // ---- systems ----
  int nsys = 0;
  std::vector<rdk_system_desc> sys_desc;
  std::vector<int> sys_off;
  std::vector<rdk::SysDev> sys_host;
  rdk::DevBuf<rdk::SysDev> sys_dev;
  rdk_physics phys{};
  bool have_phys = false;

  // ---- broad phase ----
  double built_chunk_len = -2.0, built_cell_scale = -1.0, built_delta = -1.0;
  int n_chunks = 0, n_cells = 0;
  std::vector<double> sys_radmax;  // largest chunk bounding radius per system
  std::vector<int64_t> sys_nchunk; // chunks per system
  bool grid_dirty = false;         // a box changed since the grid was sized
  rdk::DevBuf<int> chunk_rod, chunk_k, rod_nchunk;
  rdk::DevBuf<double> rodrec;  // per rod: x, FP64 axis, half length, radius (k_chunk_keys)
  rdk::DevBuf<int> rodinfo;     // per rod: system, chunk count
  rdk::DevBuf<uint32_t> ckey, ckey_alt;
  rdk::DevBuf<int> cid, cid_alt;
  rdk::DevBuf<int> cell_start;  // dense: n_cells + 1 entries
  rdk::DevBuf<float4> scf, sca;
  rdk::DevBuf<float4> ucf, uca;  // the same records in chunk order (k_chunk_keys)
  rdk::DevBuf<int> srod, sk;
  rdk::DevBuf<uint64_t> pkeys, pkeys_alt;
  rdk::DevBuf<unsigned int> counters;  // [0] pair count, [1..] misc
  unsigned int* h_counters = nullptr;  // pinned mirror
  int pair_cap_forced = 0;
  int regrows = 0;
  int64_t last_candidates = 0;
  int64_t last_narrow = 0;

  // ---- contacts (sorted by (a,b)) ----
  int nc = 0;
  rdk::DevBuf<int> ca, cb, ctype;
  rdk::DevBuf<double> cpa[3], cpb[3], cn[3], csh[3], cdist, chs, cs, ct;
  double last_delta = 0.0;

  // ---- NSC manifolds ----
  int nm = 0, nm_wall = 0;
  rdk::DevBuf<double> mf, mfc;  // 41 planar fields (natural / colour order)
  rdk::DevBuf<int> ma, mb, mac, mbc;
  rdk::DevBuf<double> lam[3], lamc[3];
  rdk::DevBuf<int> wall_flag, wall_off;
  rdk::DevBuf<int> color, color_tmp, perm, uncolored;
  rdk::DevBuf<uint32_t> color_key, color_key_alt;
  rdk::DevBuf<int> perm_alt;
  std::vector<int> color_start;  // host, ncolor+1
  int ncolor = 0;
  bool have_manifolds = false, have_coloring = false, color_host_valid = false;
  rdk::DevBuf<int> d_cstart;
  // islands
  int solver_mode = 0;  // 0 islands when possible, 1 always global coloured
  int small_m = -1;     // single-workgroup colouring limit (RDK_TUNE_SMALL_M; -1: default)
  int bn_threads = 256; // broad-phase workgroup size (RDK_TUNE_BROAD_BLOCK)
  int bn_stage = 0;     // broad phase with LDS-staged stencil rows (RDK_TUNE_BROAD_STAGE)
  int cell_band = 0;    // rows per band of the cell order, 0: sized per grid (RDK_TUNE_CELL_BAND)
  int nislands = 0;
  int island_counts[4] = {0, 0, 0, 0};  // tiny, small, medium, large
  rdk::DevBuf<int> uf_parent, uf_label, sbody, sbody_alt, iflag, ipos, istart, ilists, bseg, icnt;
  rdk::DevBuf<uint32_t> blab, blab_alt;
  rdk::DevBuf<char> istats;
  rdk::DevBuf<int> hcs;       // colour starts of the huge islands
  rdk::DevBuf<double> ent_a[3], ent_e[3], ent_out, ent_psum;  // entanglement observer
  rdk::DevBuf<unsigned long long> ent_pcnt;
  rdk::DevBuf<double> hres;   // residual and last-sweep change of the huge-island solve
  rdk::DevBuf<double> hstate;  // device-side state of the huge-island solve
  // velocity-solve manifolds of the last fused hard-contact step, kept when
  // keep_manifolds is set (the position projection reuses the manifold buffers)
  bool keep_manifolds = false;
  rdk::DevBuf<double> kept_mf, kept_lam;
  rdk::DevBuf<int> kept_ab;
  int kept_m = 0;
  size_t kept_stride = 0;
  // per-contact terms of the last soft force evaluation when keep_soft is set:
  // 6 doubles per contact in contact-list order, the normal force on b
  // (spring + damping) then the friction force on a
  bool keep_soft = false;
  // kinetic energy after the free-motion update, the velocity solve and the
  // position update (+ projection) of the last fused hard-contact step, when
  // stage_ke is set
  bool stage_ke = false;
  double ke_stage[3] = {0.0, 0.0, 0.0};
  rdk::DevBuf<double> soft_kept;
  int soft_kept_n = 0;
  rdk::DevBuf<const double*> raw_ptrs;  // contact component arrays for the FP32 record export
  rdk::DevBuf<char> raw_out;            // FP32 contact records
  rdk::DevBuf<int2> islot;  // local body slots of large-island manifolds
  rdk::DevBuf<int> ihot;    // hot sets of large islands (in each island's manifold range)
  // position projection
  rdk::DevBuf<int> ppa, ppb, pflag, ppos;
  rdk::DevBuf<double> pphs, ptmp, pdx[3], pdth[3];
  rdk::DevBuf<char> small_out;
  // body -> manifold CSR
  rdk::DevBuf<uint32_t> csr_key, csr_key_alt;
  rdk::DevBuf<int> csr_val, csr_val_alt, body_start;
  // warm cache
  int nprev = 0, nprev_w = 0;
  rdk::DevBuf<uint64_t> prev_wkeys;
  rdk::DevBuf<double> prev_wlam[3];
  rdk::DevBuf<uint64_t> prev_keys;
  rdk::DevBuf<double> prev_lam[3];
  rdk::DevBuf<double> resid;  // [0] residual bits, [1] ref residual bits
  // MLP warm start (learned)
  void* mlp = nullptr;

  // ---- force history (HM spring / Cundall) ----
  int nhist = 0;
  uint64_t hist_frame = 0;
  rdk::DevBuf<uint64_t> hkeys, hkeys_new;
  rdk::DevBuf<double> hxi[3], hxi_new[3];
  rdk::DevBuf<uint64_t> hlast, hlast_new;
  rdk::DevBuf<int> hflag, hscan;
  rdk::DevBuf<double> cf[12];  // per-contact force terms: Fb(3), Ta(3), Tb(3), xi(3)
  rdk::DevBuf<double> pe;

  // ---- scratch ----
  rdk::DevBuf<char> temp;
  rdk::DevBuf<int> scan_tmp;  // per-tile sums of the exclusive scan
  rdk::DevBuf<double> red;  // reduction scratch

  // ---- profiling ----
  bool profiling = false;
  uint64_t prof_mask = ~(uint64_t)0;
  std::vector<rdk::EventRec> ev_pending;
  std::vector<hipEvent_t> ev_pool;
  int64_t k_launches[rdk::K_NUM] = {};  // timed launches
  int64_t huge_blocks = 0;              // check blocks of the huge-island solve (timing sample)
  double k_ms[rdk::K_NUM] = {};
  int64_t work[rdk::W_NUM] = {};

  // ---- fixed-capacity (capturable) hard-contact step ----
  // Device words fx[FX_*] (error, step counts, per-step counts, colouring
  // round counters), accumulated statistics facc[FA_*], and the graph of one
  // step.
  bool capturing = false;  // inside stream capture: no allocation, no timing
  uint64_t alloc_gen = 0;  // bumped by every reallocation (captured pointers go stale)
  int fx_pair_cap = 0, fx_wall_cap = 0, fx_nm_cap = 0, fx_rounds = 0, color_rounds_max = 0;
  int nc_seen_max = 0, nwall_seen_max = 0;  // most contacts / wall manifolds an eager step has built
  int cls_seen_max[4] = {0, 0, 0, 0};       // most islands per class (tiny, small, medium, LDS) in a step
  int fx_grid[4] = {0, 0, 0, 0};            // workgroups of the fixed step's class launches
  int64_t eager_nsc_steps = 0;
  int step_graph_mode = 0;  // rdk_set_step_graph
  int step_graph_pair_cap = 0;
  rdk::DevBuf<int> fx;
  rdk::DevBuf<double> facc;
  rdk::DevBuf<double> freport;         // FxReport, packed on the device
  rdk::FxReport* h_report = nullptr;   // pinned host copy
  int* h_fx_up = nullptr;              // pinned: warm-cache counts uploaded to fx
  int fx_prev_dev[2] = {-1, -1};       // warm-cache counts fx holds (-1: unknown)
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  uint64_t step_graph_gen = 0;
  std::vector<char> step_graph_key;  // parameters the graph was captured with
  int64_t graph_steps = 0, graph_fallbacks = 0, graph_captures = 0;
};

namespace rdk {

// error handling
#define RDK_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                    \
      return RDK_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

#define RDK_TRY(expr)        \
  do {                       \
    int st_ = (expr);        \
    if (st_ != RDK_OK) return st_; \
  } while (0)

template <class T>
int ensure(rdk_ctx* ctx, DevBuf<T>& b, size_t n) {
  if (n <= b.cap && b.p) return RDK_OK;
  if (ctx->capturing) {  // a captured step must find every buffer at its capacity
    ctx->err = "step capture: a buffer would have to grow";
    return RDK_ERR_STATE;
  }
  ctx->alloc_gen++;
  size_t nc = n < 16 ? 16 : n + n / 4;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) {
      ctx->err = std::string("hipFree: ") + hipGetErrorString(e);
      return RDK_ERR_HIP;
    }
  }
  b.p = nullptr;
  b.cap = 0;
  hipError_t e = hipMalloc((void**)&b.p, nc * sizeof(T));
  if (e != hipSuccess) {
    ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return RDK_ERR_HIP;
  }
  b.cap = nc;
  return RDK_OK;
}

// ensure() that keeps the first `used` elements (state carried between steps,
// such as the warm-start cache).
template <class T>
int grow_keep(rdk_ctx* ctx, DevBuf<T>& b, size_t n, size_t used) {
  if (n <= b.cap && b.p) return RDK_OK;
  DevBuf<T> nb;
  RDK_TRY(ensure(ctx, nb, n));
  if (used > 0 && b.p) {
    hipError_t e = hipMemcpyAsync(nb.p, b.p, used * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      ctx->err = std::string("grow_keep: ") + hipGetErrorString(e);
      (void)hipFree(nb.p);
      return RDK_ERR_HIP;
    }
  }
  if (b.p) (void)hipFree(b.p);
  b = nb;
  return RDK_OK;
}

template <class T>
void release(DevBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

BodyPtrs body_ptrs(rdk_ctx* ctx);

// event timing around one launch
void prof_begin(rdk_ctx* ctx, int kernel, EventRec& rec, hipStream_t s);
void prof_end(rdk_ctx* ctx, EventRec& rec, hipStream_t s);
int prof_flush(rdk_ctx* ctx);

// Times the launches made while it lives (events on the stream they run on).
struct ProfScope {
  rdk_ctx* ctx;
  EventRec rec;
  hipStream_t s;
  bool on;
  ProfScope(rdk_ctx* c, int k, hipStream_t st = nullptr)
      : ctx(c), s(st ? st : c->stream), on(c->profiling && !c->capturing && ((c->prof_mask >> k) & 1)) {
    if (on) {
      c->k_launches[k] += 1;
      prof_begin(c, k, rec, s);
    }
  }
  ~ProfScope() {
    if (on) prof_end(ctx, rec, s);
  }
};

static inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

// Events for a launch made with hipExtLaunchKernelGGL, which stamps them at
// the kernel's own start and end (no dispatch gap in the measurement, which
// matters for launches of a few microseconds).  Returns false when kernel k
// is not being timed; otherwise the caller launches with rec.a / rec.b and
// then calls prof_ext_done.
bool prof_ext_events(rdk_ctx* ctx, int k, EventRec& rec);
void prof_ext_done(rdk_ctx* ctx, EventRec& rec);

// True when kernel k's launches are being timed (timed launches are kept out
// of graphs: a graph replay cannot be split into its kernels by events).
static inline bool timed(const rdk_ctx* ctx, int k) {
  return ctx->profiling && !ctx->capturing && ((ctx->prof_mask >> k) & 1);
}

// Graph replay of repeated launch sequences (the huge-island check blocks, the
// captured fixed-capacity step); RDK_GRAPHS=0 turns it off, launching the same
// kernels one by one (for tools that cannot follow captured graphs).
bool graphs_enabled();

// Fixed-capacity step: device word and accumulator layout.
enum FxWord {
  FX_ERR = 0,     // first failure of the current run of captured steps (0: none)
  FX_STEPS,       // steps completed since the host last cleared it
  FX_NWALL,       // plane-wall manifolds of the current step
  FX_NM,          // manifolds of the current step
  FX_NISL,        // islands of the current step
  FX_NPREV,       // rod-rod warm-cache entries
  FX_NPREV_W,     // wall warm-cache entries
  FX_ROUND0 = 8,  // colouring round r: manifolds left uncoloured (unsigned)
  FX_WORDS = 8 + 256
};
// errors of a captured step (the step is then redone eagerly)
enum FxErr { FXE_PAIRS = 1, FXE_COLOURS = 2, FXE_ROUNDS = 3, FXE_ISLANDS = 4, FXE_WALLS = 5 };
enum FxAcc {
  FA_STEPS = 0, FA_CONTACTS, FA_MANIFOLDS, FA_SWEEPS, FA_COLORS, FA_NONCONV, FA_RESMAX, FA_CANDIDATES, FA_NARROW,
  FA_MSWEEPS, FA_LDS_SWEEPS, FA_HOT_SWEEPS, FA_WALLS, FA_CLS0, FA_CLS1, FA_CLS2, FA_CLS3, FA_NUM
};
// Everything the host reads after a run of fixed steps, packed on the device
// by one kernel and copied back in one transfer: the step words, the pair and
// wall counters, the colouring state, the island class counts and the step
// statistics.
struct FxReport {
  double acc[FA_NUM];
  int fx[FX_ROUND0];
  unsigned ctr[2];
  unsigned dc[4];
  int cnt[6];
};

// One hard-contact step with no host synchronisation (capturable): detection
// into the reserved pair capacity, manifolds, colouring, islands, the island
// classes that fit one workgroup, warm cache, statistics.  Requires
// fixed_prepare (capacities reserved, grid built).  Failures (capacity,
// colouring, an island too large for one workgroup) set fx[FX_ERR] before the
// step's first state change; every later kernel that changes state then
// returns at once.
int fixed_prepare(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* np, int pair_cap, int wall_cap);
int detect_fixed_enqueue(rdk_ctx* ctx, const rdk_detect_params* p);
int nsc_fixed_enqueue(rdk_ctx* ctx, const rdk_nsc_params* p);
int detect_fixed_reserve(rdk_ctx* ctx, const rdk_detect_params* p, int pair_cap);
int nsc_fixed_reserve(rdk_ctx* ctx, const rdk_nsc_params* p);
int fixed_sync_host_counts(rdk_ctx* ctx);
int fixed_report_enqueue(rdk_ctx* ctx);
void fixed_apply_counts(rdk_ctx* ctx, const FxReport& R);
int fixed_clear_enqueue(rdk_ctx* ctx);
void fixed_class_grids(const rdk_ctx* ctx, int* grids);

// stage entry points implemented in the .hip files
int detect_run(rdk_ctx* ctx, const rdk_detect_params* p);
int regrid(rdk_ctx* ctx);
int affine_run(rdk_ctx* ctx, double s);
int nsc_build_run(rdk_ctx* ctx, const rdk_nsc_params* p);
int nsc_solve_run(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st);
int nsc_sweeps_run(rdk_ctx* ctx, const rdk_nsc_params* p, int nsweeps);
int nsc_color_run(rdk_ctx* ctx, bool need_host);
int color_to_host(rdk_ctx* ctx);
int nsc_residual_run(rdk_ctx* ctx, const rdk_nsc_params* p, double* res, double* res_ref);
int integrate_run(rdk_ctx* ctx, int stage, const int* guard = nullptr);
int forces_run(rdk_ctx* ctx, int model, const void* params, double* pe);
int observables_run(rdk_ctx* ctx, double* ke, double* phi);
int entanglement_run(rdk_ctx* ctx, double cutoff, double* sums, int64_t* pairs);
// device address of the allocation-free scratch array for ctx->device (rdk_entangle.hip)
int noalloc_scratch_init(rdk_ctx* ctx);
int populate_nonoverlap_run(rdk_ctx* ctx, int n, float length, float diameter, const float* bmin, const float* bmax,
                            int periodic, float cell_size, uint32_t seed, int max_attempts, int window,
                            float* centres, float* dirs, float* quats, int* placed, int64_t* attempts);
int mlp_warm_start(rdk_ctx* ctx, int mode);
int nsc_project_run(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* p, int* passes, int* npos);
void mlp_free(rdk_ctx* ctx);

// Stable radix sorts and the exclusive scan (rdk_sort.hip).  The result of a
// sort is in the *_alt buffers when result_in_alt is set, else in place.  A
// Count with a device pointer sizes the launches for its capacity and reads
// the element count when they run; st = nullptr is the context stream.
int sort_pairs_u32(rdk_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, int* vals, int* vals_alt, int n,
                   int end_bit, bool& result_in_alt);
int sort_pairs_u32(rdk_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, int* vals, int* vals_alt, Count n,
                   int end_bit, bool& result_in_alt, hipStream_t st);
// 64-bit pair keys (lo << 32) | hi with lo, hi < 2^index_bits
int sort_pair_keys_u64(rdk_ctx* ctx, uint64_t* keys, uint64_t* keys_alt, Count n, int index_bits,
                       bool& result_in_alt, hipStream_t st);
int exclusive_scan_int(rdk_ctx* ctx, const int* in, int* out, int n);
int exclusive_scan_int(rdk_ctx* ctx, const int* in, int* out, Count n, hipStream_t st);
int sort_reserve(rdk_ctx* ctx, int max_count);

static inline int bits_for(uint64_t maxval) {
  int b = 1;
  while (b < 64 && (uint64_t(1) << b) <= maxval) ++b;
  return b;
}

}  // namespace rdk

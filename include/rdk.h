/*
 * rdk.h -- C ABI of librdk, the MI355X (gfx950) rod-dynamics contact/step
 * library.  Plain pointers and sizes only; nothing throws across this
 * boundary.  Every function returns an rdk_status (0 = RDK_OK).
 *
 * Citations name the reference interface each entry point replaces, in
 * yeonsu-jung/rod-dynamics-3d.
 *
 * Host arrays are interleaved per body: x[3N], q[4N] as (w,x,y,z), v[3N],
 * w[3N], f[3N], tau[3N]; scalars r[N] (radius), h[N] (half length of the
 * cylinder part, the reference's Capsule::h), mass[N] (0 = static body),
 * restitution[N].  On the device the library keeps its own planar FP64
 * structure-of-arrays copy; host arrays are read or written only by the
 * upload/download entry points.
 */
#ifndef RDK_H
#define RDK_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct rdk_ctx rdk_ctx;

enum rdk_status {
  RDK_OK = 0,
  RDK_ERR_HIP = 1,      /* a HIP runtime call failed; see rdk_last_error */
  RDK_ERR_ARG = 2,      /* invalid argument */
  RDK_ERR_CAPACITY = 3, /* a fixed capacity was exceeded; *required holds the size */
  RDK_ERR_STATE = 4,    /* call out of order (e.g. solve before build) */
  RDK_ERR_NOCONV = 5    /* PSOR hit max_sweeps before reaching tol (results kept) */
};

/* One independent box.  Axis k is periodic when periodic[k] != 0 (the
 * reference's PBC, main.cpp:4508-4517 / integrator.cpp:23-35).  floor_wall /
 * ceil_wall put plane walls at lo[2] / hi[2] (axis 2 must then be
 * non-periodic).  mu is the Coulomb coefficient used by the NSC solve for
 * this box. */
typedef struct rdk_system_desc {
  double lo[3], hi[3];
  int periodic[3];
  int floor_wall, ceil_wall;
  double mu;
} rdk_system_desc;

typedef struct rdk_physics {
  double gravity[3];
  double dt;
  double lin_damp, ang_damp, w_max; /* integrator.cpp:14-16 globals */
} rdk_physics;

/* Broad-phase tuning.  chunk_len: rods are cut into ceil(2h/chunk_len)
 * sub-segments for binning (<= 0: one chunk per rod).  cell_scale >= 1
 * multiplies the minimal cell size. */
typedef struct rdk_detect_params {
  double delta;     /* activation margin: contact if dist < r_a + r_b + delta */
  double chunk_len;
  double cell_scale;
} rdk_detect_params;

typedef struct rdk_nsc_params {
  double beta, cfm, omega; /* NscContactCfg, config.hpp:66-88 */
  double tol;              /* fixed-point residual tolerance (velocity units) */
  int max_sweeps;
  int check_every;         /* residual evaluated every k sweeps */
  int friction_proj;       /* 0 reference sequential disc clamp, 1 Euclidean disc (one scalar
                              step omega/g_tb), 2 Euclidean disc with the block step
                              omega G^-1 w_t (G the 2x2 tangential Delassus block) projected
                              in G's metric, 3 the block step where it stays in the disc
                              (sticking) and 1's step where it leaves it (sliding); 1, 2 and
                              3 share their fixed points and residual */
  int position_stabilization; /* scales beta by 0.01 (nsc_contact.cpp:339-341) */
  int warm_mode;           /* 0 cold, 1 previous step (reference), 2 learned (MLP) for the
                              manifolds without a previous-step normal impulse, 3 learned for
                              every manifold (the previous-step impulse is an input), 4 learned
                              for every manifold from the contact alone (impulse input 0) */
  int stop_rule;           /* 0 reference (max sum|dlambda| < 1e-6), 1 residual <= tol */
  int warm_walls;          /* 1: plane-wall manifolds also start from the previous
                              step's impulses, keyed by (rod, plane, endpoint); the
                              reference starts wall contacts cold */
  int stall_sweeps;        /* stop_rule 1 only; 0 = off.  > 0: at sweeps = k, 2k,
                              4k, ... an island (or the global solve) whose residual
                              has not dropped by 10 % since the previous milestone
                              continues with omega = stall_omega if its omega is
                              larger (over-relaxed PSOR can cycle), else stops there
                              and counts as not converged */
  int position_iters;      /* position projection (config.hpp:77-80), run by the step
                              after the position update when position_stabilization
                              is set: outer passes (reference default 5) */
  int position_psor_iters; /* inner normal-only sweeps per pass (default 50) */
  double slop;             /* allowed penetration before correction (default 1e-4) */
  double stall_omega;      /* omega after the first stall (1 = plain Gauss-Seidel);
                              0 = stop at the first stall */
  int hot_set;             /* stop_rule 1, islands solved in one workgroup with their
                              bodies in LDS: 1 = once at least 64 sweeps have run and at
                              most max(16, m/50) of the island's m manifolds are above
                              tol, sweep only the manifolds sharing a body with one
                              above tol/2 (at most 512, in the island's colour order)
                              until their residual is at most tol/2, then check the
                              whole island again (Gauss-Seidel on a subset; the stop
                              rule is unchanged); 2 = as 1, and when no such hot set is
                              taken and at least 32 sweeps have run, rounds of
                              check_every sweeps over the same set of any size up to
                              half the island, each followed by the island check;
                              0 = full sweeps only */
} rdk_nsc_params;

typedef struct rdk_soft_params { /* SoftContactCfg, config.hpp:6-27 */
  double delta, k_scaler, damping, mu, mu_static, nu, kt, vel_deadband;
  int enable_friction;
  int friction_model; /* 0 smooth tanh/Stribeck, 1 Cundall-Strack, 2 Karnopp */
} rdk_soft_params;

typedef struct rdk_hm_params { /* HertzMindlinCfg, config.hpp:29-64 */
  double E, nu_poisson, restitution, mu_k, mu_s, v_c, mu_roll;
  int enable_tangential, enable_rolling;
  int law; /* 0 reference (elastic force ~ delta^2), 1 Hertz delta^1.5 */
} rdk_hm_params;

typedef struct rdk_solve_stats {
  int manifolds, rod_contacts, wall_contacts, colors, sweeps;
  double residual;       /* final fixed-point residual */
  double residual_ref;   /* reference rule: max sum|dlambda| of the last sweep */
} rdk_solve_stats;

typedef struct rdk_step_stats {
  int64_t steps, contacts_total, manifolds_total, sweeps_total, colors_total;
  int64_t nonconverged_steps;
  double residual_max;
  double ms_detect, ms_build, ms_color, ms_solve, ms_integrate; /* event timers (profiling on) */
} rdk_step_stats;

/* ---- context --------------------------------------------------------- */
/* replaces the solver objects' value members (nsc_contact.hpp:159-188,
 * contact_detector.hpp:77-93) and the file-static device state of
 * cuda_broadphase.cu:293,658 */
int rdk_create(int device, rdk_ctx** out);
int rdk_destroy(rdk_ctx* ctx);
const char* rdk_last_error(rdk_ctx* ctx);
int rdk_sync(rdk_ctx* ctx);
int rdk_version(void);

/* ---- bodies and systems ---------------------------------------------- */
/* replaces RigidBody::makeCapsule (rigid_body.cpp:11-47): mass, solid-
 * cylinder inertia about local +Y, static when mass == 0 */
int rdk_set_bodies(rdk_ctx* ctx, int n, const double* x, const double* q,
                   const double* v, const double* w, const double* r,
                   const double* h, const double* mass,
                   const double* restitution);
/* body_offsets[nsys+1]: system s owns bodies [off[s], off[s+1]) */
int rdk_set_systems(rdk_ctx* ctx, int nsys, const int* body_offsets,
                    const rdk_system_desc* desc);
int rdk_set_physics(rdk_ctx* ctx, const rdk_physics* p);
int rdk_upload_state(rdk_ctx* ctx, const double* x, const double* q,
                     const double* v, const double* w);
int rdk_upload_forces(rdk_ctx* ctx, const double* f, const double* tau);
int rdk_download_state(rdk_ctx* ctx, double* x, double* q, double* v,
                       double* w, double* f, double* tau);

/* ---- detection (ContactDetector::detectContacts,
 *      contact_detector.cpp:71-211,487-1035; cuda_broadphase.hpp:95-101) --- */
int rdk_detect(rdk_ctx* ctx, const rdk_detect_params* p, int* n_contacts);
/* sorted by (a,b), a<b; any pointer may be NULL.  type: 0 P2P, 1 E2P, 2 E2E */
int rdk_get_contacts(rdk_ctx* ctx, int cap, int* a, int* b, double* pa,
                     double* pb, double* nrm, double* dist, double* shift,
                     double* s, double* t, int* type);
/* test hook: force the pair-buffer capacity (0 = automatic) */
int rdk_set_pair_capacity(rdk_ctx* ctx, int cap);

/* The reference's GPU plug-in record (GpuContactRaw, cuda_broadphase.hpp:29-38),
 * same fields, order and types: closest points (point_b on b's periodic
 * image), normal a -> b, segment distance, r_a + r_b, Lumelsky s and t, the
 * minimum-image shift applied to b. */
typedef struct rdk_contact_raw {
  int a, b;
  float px_a, py_a, pz_a;
  float px_b, py_b, pz_b;
  float nx, ny, nz;
  float dist;
  float surface_limit;
  float s, t;
  float shift_bx, shift_by, shift_bz;
} rdk_contact_raw;
/* The current contact list (after rdk_detect) in that layout, FP32, sorted by
 * (a, b).  RDK_ERR_CAPACITY when cap < the contact count. */
int rdk_get_contacts_raw(rdk_ctx* ctx, int cap, rdk_contact_raw* out, int* count);
/* One call in place of cudaDetectCapsulePairsTwoPass (cuda_broadphase.hpp:95-101):
 * n capsules given in FP32 as the reference holds them (centres x[3n],
 * quaternions q[4n] as (w,x,y,z), radii r[n], half lengths h[n]); contact when
 * the segment distance is below r_a + r_b + margin; periodic images by the
 * box extents sx, sy, sz when pbc is set (any box origin).  Detection runs in
 * FP64 on the device; the records are rounded to FP32.  Replaces the bodies
 * and systems of ctx.  *count = contacts found; RDK_ERR_CAPACITY when cap is
 * smaller (out then holds the first cap). */
int rdk_detect_capsule_pairs_raw(rdk_ctx* ctx, int n, const float* x, const float* q,
                                 const float* r, const float* h, float margin, int pbc,
                                 float sx, float sy, float sz, int cap,
                                 rdk_contact_raw* out, int* count);
int rdk_detect_info(rdk_ctx* ctx, int* n_chunks, int* n_cells,
                    int64_t* candidates, int* regrows);

/* ---- forces (SoftContactSolver::computeForces soft_contact.cpp:80-213;
 *      HertzMindlinSolver::computeForces hertz_mindlin.cpp:183-247) ------
 * accumulate into the device f, tau of the bodies in the last detection */
int rdk_forces_soft(rdk_ctx* ctx, const rdk_soft_params* p, double* pe);
int rdk_forces_hm(rdk_ctx* ctx, const rdk_hm_params* p, double* pe);
int rdk_history_size(rdk_ctx* ctx, int* n);

/* ---- NSC (NscContactSolver, nsc_contact.hpp:71-157) -------------------- */
/* detectAndBuildManifolds + plane-wall addWallContact (nsc_contact.cpp:190-321) */
int rdk_nsc_build(rdk_ctx* ctx, const rdk_nsc_params* p, int* n_manifolds);
/* rec: RDK_MF_FIELDS doubles per manifold (n,t1,t2,r_a,r_b,
 * I_A^-1(r_a x n|t1|t2), I_B^-1(r_b x n|t1|t2), invMA, invMB, g_n, g_t1,
 * g_t2, phi, v_n_pre, e, g_tb = largest eigenvalue of the (t1,t2) block);
 * ab: 2 ints (b < 0 for a wall: b = -1 - slot, slot = 2*plane + endpoint);
 * lambda: 3 doubles; natural order */
#define RDK_MF_FIELDS 42
int rdk_nsc_get_manifolds(rdk_ctx* ctx, int cap, double* rec, int* ab,
                          double* lambda);
int rdk_nsc_set_lambda(rdk_ctx* ctx, int m, const double* lambda);
/* solveVelocities (nsc_contact.cpp:325-655): colours, warm impulses,
 * coloured PSOR to tol, warm-cache update */
int rdk_nsc_solve(rdk_ctx* ctx, const rdk_nsc_params* p, rdk_solve_stats* st);
/* fixed sweeps from the current lambda (no warm application); test hook */
int rdk_nsc_sweeps(rdk_ctx* ctx, const rdk_nsc_params* p, int nsweeps);
/* colour of every manifold and the colour-ordered permutation */
int rdk_nsc_get_coloring(rdk_ctx* ctx, int cap, int* color, int* order,
                         int* n_colors);
int rdk_nsc_residual(rdk_ctx* ctx, const rdk_nsc_params* p, double* res);
/* 0 (default): solve each island (connected component of the rod-rod contact
 * graph) independently on the GPU whenever every island fits one workgroup;
 * 1: always sweep all manifolds globally, colour by colour */
int rdk_set_solver_mode(rdk_ctx* ctx, int mode);
/* after rdk_nsc_solve: islands and their split by solver class (tiny <= 8
 * manifolds, small <= 64, medium <= 256, large <= 8192), and the manifold
 * order the last solve swept in (natural indices) */
int rdk_nsc_island_info(rdk_ctx* ctx, int* n_islands, int* counts4);
/* projectPositions, multi-body branch (nsc_contact.cpp:718-880): detect at the
 * current positions, then up to position_iters passes of normal-only position
 * PSOR (omega, cfm, slop from p) over those pairs, each applying
 * x += dx, q <- normalize(angleAxis(|dth|, dth/|dth|) q) and the periodic wrap.
 * *passes: passes that found pairs deeper than slop; *n_first: their count in
 * the first pass.  Leaves the manifold buffers holding position manifolds
 * (rdk_nsc_get_solve_order gives the last pass's sweep order). */
int rdk_nsc_project(rdk_ctx* ctx, const rdk_detect_params* dp, const rdk_nsc_params* p, int* passes,
                    int* n_first);
int rdk_nsc_get_solve_order(rdk_ctx* ctx, int cap, int* order);

/* ---- integrators (integrator.cpp:56-159, main.cpp:4397-4483) ----------- */
enum rdk_stage {
  RDK_NSC_FREE = 0,  /* v += (f/m + g)dt; w += I^-1(tau - w x Iw)dt */
  RDK_NSC_POS = 1,   /* x += v dt; q += dt/2 (0,w) q; damping; clamp; wrap; clear f,tau */
  RDK_VV_HALF1 = 2,  /* integrateHalfPos */
  RDK_VV_HALF2 = 3,  /* integrateSecondHalf */
  RDK_CLEAR = 4      /* f = tau = 0 */
};
int rdk_integrate(rdk_ctx* ctx, int stage);

/* ---- fused steps (App::physicsStep, main.cpp:4333-4858) ---------------- */
int rdk_step_nsc(rdk_ctx* ctx, const rdk_detect_params* dp,
                 const rdk_nsc_params* np, int nsteps, rdk_step_stats* st);
/* Keep the velocity-solve manifolds of each fused hard-contact step (the
 * position projection reuses the manifold buffers): what the reference's
 * network export reads from NscContactSolver::getManifolds after a step
 * (main.cpp:3699-3720).  rdk_nsc_get_kept_manifolds returns those of the
 * last step in the layout of rdk_nsc_get_manifolds; *count = their number,
 * RDK_ERR_CAPACITY when cap is smaller. */
int rdk_keep_manifolds(rdk_ctx* ctx, int on);
int rdk_nsc_get_kept_manifolds(rdk_ctx* ctx, int cap, double* rec, int* ab,
                               double* lambda, int* count);
/* Keep the per-contact forces of each soft force evaluation: what the
 * reference's network export reads from SoftContactSolver::getContacts
 * (main.cpp:3674-3688; force_a/b and friction_a/b of soft_contact.cpp:115-147).
 * rdk_get_soft_forces returns those of the last evaluation, 6 doubles per
 * contact in the order of rdk_get_contacts_raw: the normal force on b
 * (spring + damping; force_a is its negative) and the friction force on a
 * (friction_b is its negative).  A Hertz-Mindlin evaluation keeps none.
 * *count = their number, RDK_ERR_CAPACITY when cap is smaller. */
int rdk_keep_soft_forces(rdk_ctx* ctx, int on);
/* Record the kinetic energy after the stages of each fused hard-contact step,
 * the profile log's KE_after_integrate, KE_after_solve and
 * KE_after_posCorrect (main.cpp:4411,4461,4505; KE_after_pbcWrap, :4519,
 * equals the last: the wrap moves no velocity).  rdk_get_stage_energies
 * writes those of the last step into out[3]. */
int rdk_record_stage_energies(rdk_ctx* ctx, int on);
int rdk_get_stage_energies(rdk_ctx* ctx, double* out);
/* Contact potential energy of the last soft or Hertz-Mindlin force evaluation
 * (the profile log's soft_PE, main.cpp:3321); 0 before the first. */
int rdk_last_potential_energy(rdk_ctx* ctx, double* pe);
int rdk_get_soft_forces(rdk_ctx* ctx, int cap, double* out, int* count);
/* model 0 soft (params = rdk_soft_params*), 1 Hertz-Mindlin (rdk_hm_params*) */
int rdk_step_vv(rdk_ctx* ctx, const rdk_detect_params* dp, int model,
                const void* params, int nsteps, rdk_step_stats* st);

/* ---- observables (App::totalKE main.cpp:3357-3368) --------------------- */
int rdk_observables(rdk_ctx* ctx, double* kinetic_energy,
                    double* packing_fraction);

/* ---- entanglement observer (App::computeEntanglement, main.cpp:4287-4331;
 *      pairwise_abs_linking_sum_with_cutoff, entanglement_adapter.cpp:56-117)
 * Per system s: sums[s] = sum over rod pairs of |Gauss linking number| of
 * the two axis segments, pairs[s] = number of pairs summed; every pair for
 * cutoff <= 0, else the pairs whose segments are at most cutoff apart.  On
 * periodic systems each segment endpoint is wrapped into the box first, as
 * the reference does.  The reference prunes pairs by centre distance minus
 * 1.0 before the segment distance (exact for its unit-length rods); this
 * test uses the half lengths, so it is exact for any length.  The mean
 * sums[s] / pairs[s] is the reference's normalised entanglement (ebar). */
int rdk_entanglement(rdk_ctx* ctx, double cutoff, double* sums, int64_t* pairs);

/* Allocation-free variant of rdk_entanglement.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  Same sums and counts as
 * rdk_entanglement, from the same kernels, written asynchronously on the
 * context's stream into caller-owned device memory: d_sums[s] and d_pairs[s]
 * for each system s (nsys entries each); the caller synchronises (rdk_sync).
 * The call allocates and frees nothing: its scratch is one static device
 * array of RDK_NOALLOC_SCRATCH_BYTES per GPU, shared by every context on that
 * GPU and held until the process ends, so calls that overlap on one GPU
 * (contexts on different streams) race on it.  RDK_ERR_CAPACITY when the
 * context holds more than RDK_NOALLOC_MAX_RODS rods. */
#define RDK_NOALLOC_MAX_RODS 32768
#define RDK_NOALLOC_SCRATCH_BYTES (4u << 20)
int rdk_entanglement_noalloc(rdk_ctx* ctx, double cutoff, double* d_sums, uint64_t* d_pairs);

/* ---- fixed-capacity, graph-captured hard-contact step ------------------
 * No counterpart in the reference, whose step reads counts back to the host
 * between stages (main.cpp:4333-4858, cuda_broadphase.cu:704-722).  Mode 0
 * (default): rdk_step_nsc sizes every stage from counts read back to the
 * host.  Mode 1: the same step with every launch sized for reserved
 * capacities and every count read on the device (no host synchronisation
 * inside a step).  Mode 2: mode 1 captured once into a HIP graph and replayed
 * for each step (recaptured when the parameters or buffers change).  Results
 * are bit-identical to mode 0.  pair_capacity = rod-rod contacts reserved per
 * step (0: 1.25 times the most that a mode-0 step of this context has
 * found, at least 1024, rounded up to a power of two; the first step of a
 * context runs in mode 0 to measure it); wall manifolds are reserved the same
 * way (at least 256), islands per class for four times the most found.  A step the fixed path cannot finish (more
 * contacts, wall manifolds or islands than reserved, colouring beyond its
 * enqueued rounds, an island with more bodies than one workgroup's LDS holds
 * or more than 8192 manifolds) fails before it changes any state and is
 * redone by the mode-0 step, which raises the capacities; the fixed path then
 * resumes.  Modes 1 and 2 apply when
 * position stabilisation, stage energies, kept manifolds and learned starts
 * are off and the island solver is on; otherwise rdk_step_nsc runs mode 0. */
int rdk_set_step_graph(rdk_ctx* ctx, int mode, int pair_capacity);
/* Launch-shape tuning (no reference counterpart; results do not depend on
 * it).  RDK_TUNE_SMALL_M: manifold counts up to `value` are coloured (and on
 * the global path solved) by one workgroup, larger ones by rounds over the
 * whole GPU; -1 restores the default (16384, or the RDK_SMALL_M environment
 * variable).  RDK_TUNE_BROAD_BLOCK: workgroup size of the broad phase, 64,
 * 128, 256 (default) or 512.  RDK_TUNE_BROAD_STAGE: 1 stages each
 * workgroup's stencil rows in LDS before its candidate walk (three passes, one
 * per z offset), 0 (default) walks them in global memory.
 * RDK_TUNE_CELL_BAND: the cell order of each system's grid runs z over bands
 * of y-rows so that a large dense grid's stencil planes stay in L2; `value`
 * >= 1 fixes the band height in rows (capped at the grid's rows), 0 (default)
 * sizes it from the grid and the chunk count. */
#define RDK_TUNE_SMALL_M 0
#define RDK_TUNE_BROAD_BLOCK 1
#define RDK_TUNE_BROAD_STAGE 2
#define RDK_TUNE_CELL_BAND 3
int rdk_set_tuning(rdk_ctx* ctx, int what, int value);
/* steps run by the fixed path, steps redone by the eager step, graph captures */
int rdk_step_graph_info(rdk_ctx* ctx, int64_t* fixed_steps, int64_t* fallbacks, int64_t* captures);

/* ---- profiling: per-kernel HIP-event timing on the stream each kernel runs on */
int rdk_set_profiling(rdk_ctx* ctx, int on);
/* time only the kernels whose id bit is set (bit k = kernel id k).  The
 * colour launches of the huge-island solve are replayed from captured graphs;
 * while "psor_sweep" is timed, one check block in 16 runs without the graph
 * and its launches are timed (a sample over the whole timed region). */
int rdk_set_profile_mask(rdk_ctx* ctx, uint64_t mask);
/* kernel ids: see rdk_kernel_name.  launches = launches whose time was
 * recorded, total_ms = their summed event time */
int rdk_kernel_times(rdk_ctx* ctx, int kernel, int64_t* launches,
                     double* total_ms);
const char* rdk_kernel_name(int kernel);
int rdk_num_kernels(void);
/* algorithmic-byte counters accumulated by the fused steps, in this order:
 * steps, detection passes, chunks binned, chunk pairs passing the sphere
 * test, contacts, manifolds, manifold updates in sweeps, manifold residual
 * evaluations, body integrations, force evaluations, FP64 narrow-phase
 * calls, manifold updates in timed colour launches, manifold updates in
 * sweeps of the LDS island class, hot-set sweeps (summed over islands) */
int rdk_work_counters(rdk_ctx* ctx, int64_t* counters, int n);

/* ---- the library's device sort and scan on host arrays ---------------------
 * No counterpart in the reference (it sorts with std::sort,
 * contact_detector.cpp:780-793, nsc_contact.cpp:87-102): these run the stable
 * radix sorts and the exclusive scan that every stage of this library uses,
 * so they can be checked on their own.  When cap > n the launches are sized
 * for cap and read n from device memory, as in a captured step. */
/* stable sort of n (key, value) pairs by keys below 2^key_bits, in place */
int rdk_sort_pairs_u32(rdk_ctx* ctx, int n, int cap, uint32_t* keys, int32_t* vals, int key_bits);
/* sort of n pair keys (lo << 32) | hi with lo, hi < 2^index_bits, in place */
int rdk_sort_pair_keys_u64(rdk_ctx* ctx, int n, int cap, uint64_t* keys, int index_bits);
/* out[i] = in[0] + ... + in[i-1] */
int rdk_exclusive_scan_i32(rdk_ctx* ctx, int n, int cap, const int32_t* in, int32_t* out);

/* ---- scene generation (resetScene populate/randomInit,
 *      main.cpp:1857-2140,2243-2263) ------------------------------------- */
/* Random sequential addition of non-overlapping capsules in the box, using
 * all 27 periodic images on periodic axes and a clearance `gap` between
 * surfaces (and from walls).  Orientations uniform on S^2.  Returns the
 * number placed in *placed (may be < n if max_attempts is hit). */
int rdk_gen_rsa(int n, const double* r, const double* h,
                const rdk_system_desc* box, double gap, uint64_t seed,
                int max_attempts, double* x, double* q, int* placed);
/* The reference's populate mode "nonoverlap" (App::resetScene,
 * main.cpp:1857-2140), reproduced bit for bit on the GPU: equal capsules of
 * length `length` and diameter `diameter` drawn from std::mt19937(seed)
 * through std::uniform_real_distribution<float> (5 numbers per attempt:
 * centre uniform in [bmin, bmax), direction uniform on S^2), accepted when no
 * earlier rod lies closer than one diameter (FP32 segment distance, 27
 * images when periodic, the reference's placement grid of cell size
 * cell_size, or max(L/4, 1.25 d) when cell_size <= 0).  A rod that fails
 * max_attempts times in a row ends the populate (the reference uses
 * max(1000, populate.maxAttempts)).  Outputs, in placement order: centres and
 * unit directions (n x 3 floats), quaternions (w,x,y,z) rotating +Y onto each
 * direction (n x 4, may be NULL), *placed, *attempts (candidates drawn).
 * `window` = candidates decided per parallel round (0 = automatic); the
 * result does not depend on it. */
int rdk_populate_nonoverlap(rdk_ctx* ctx, int n, float length, float diameter,
                            const float* bmin, const float* bmax, int periodic,
                            float cell_size, uint32_t seed, int max_attempts,
                            int window, float* centres, float* dirs,
                            float* quats, int* placed, int64_t* attempts);
/* The reference's randomInit (App::resetScene, main.cpp:2213-2294) bit for
 * bit: one std::mt19937(seed) (seed 0: std::random_device) over the rods in
 * order, fresh std::normal_distribution<float> objects per rod, FP32
 * arithmetic.  Rods are in the reference's y-up frame; q = FP32 quaternions
 * (w, x, y, z) rotating +Y onto the rod axis, as the reference holds them
 * (rdk_populate_nonoverlap returns them).  mode RDK_RANDOM_INIT_GAUSSIAN:
 * v ~ N(0, v_sigma), w ~ N(0, w_sigma) per component; THERMAL: sigma_v =
 * sqrt(kbt_trans / mass), sigma_w = sqrt(kbt_rot / i_perp), w in the plane
 * perpendicular to the axis (the reference's `dw(gen) * e1 + dw(gen) * e2`
 * drawn in g++'s order, the e2 term first); UNIFORM: v ~ U(-v_sigma,
 * v_sigma), w = w_speed times a uniform direction.  project != 0 removes the
 * spin about the axis (gaussian, uniform).  skip[i] != 0 (may be NULL): rod i
 * draws nothing and keeps its v, w.  v, w: n x 3 floats. */
#define RDK_RANDOM_INIT_GAUSSIAN 0
#define RDK_RANDOM_INIT_THERMAL 1
#define RDK_RANDOM_INIT_UNIFORM 2
int rdk_random_init(int n, const float* q, const float* mass, const float* i_perp, int mode, float v_sigma,
                    float w_sigma, float w_speed, float kbt_trans, float kbt_rot, uint32_t seed, int project,
                    const unsigned char* skip, float* v, float* w);
/* Thermal velocities: v ~ N(0, kT/m), w ~ N(0, kT/I_perp) on the two axes
 * perpendicular to the rod, zero spin about the axis. */
int rdk_gen_thermal(int n, const double* q, const double* r, const double* h,
                    const double* mass, double kT, uint64_t seed, double* v,
                    double* w);

/* ---- affine strain ----------------------------------------------------
 * No counterpart in the reference (it has no strain or compression); the
 * dense Hertz-Mindlin workload applies x <- lo + s (x - lo) and scales every
 * box by s (isotropic, about the lower corner) once per step, with
 * s = 1 - strain_rate * dt.  Velocities are not changed.  Cell grids are
 * resized for the new boxes. */
int rdk_affine_scale(rdk_ctx* ctx, double s);

/* ---- learned PSOR warm start (warm_mode 2) ----------------------------
 * No counterpart in the reference: its warm start is the previous-step
 * impulse cache (src/physics/nsc_contact.cpp:305-317, warm_mode 1).  A
 * per-manifold MLP 12 -> 64 -> 64 -> 3 (SiLU) predicts the converged
 * impulses; PSOR then iterates from the prediction to the same tolerance.
 * Parameters are FP32, row-major per layer: W1[12][64], b1[64],
 * W2[64][64], b2[64], W3[64][3], b3[3].  Inference runs BF16 on the matrix
 * cores; training runs FP32 with AdamW. */
#define RDK_MLP_IN 12
#define RDK_MLP_OUT 3
int rdk_mlp_num_params(void);
/* Glorot-uniform weights, zero biases, fresh AdamW state. */
int rdk_mlp_init(rdk_ctx* ctx, uint64_t seed);
int rdk_mlp_set_params(rdk_ctx* ctx, const float* params, int n);
int rdk_mlp_get_params(rdk_ctx* ctx, float* params, int n);
/* Features of the current manifolds (natural order, after rdk_nsc_build):
 * X[cap][12]; lscale[i] is the impulse unit of manifold i (outputs and
 * training targets are lambda / lscale). */
int rdk_mlp_features(rdk_ctx* ctx, int cap, float* X, double* lscale);
/* Forward pass on host features: Y[n][3] (BF16 matrix-core path). */
int rdk_mlp_infer(rdk_ctx* ctx, const float* X, int n, float* Y);
/* Gradient of the mean squared error over n samples (FP32 path):
 * grad[num_params], *loss = mean over n*3 outputs. */
int rdk_mlp_grad(rdk_ctx* ctx, const float* X, const float* T, int n,
                 float* grad, float* loss);
/* Mini-batch AdamW (beta 0.9/0.999, eps 1e-8, decoupled weight decay on
 * the weight matrices) for `epochs` shuffled passes over (X, T). */
int rdk_mlp_train(rdk_ctx* ctx, const float* X, const float* T, int n,
                  int epochs, int batch, float lr, float weight_decay,
                  uint64_t seed, float* final_loss);

#ifdef __cplusplus
}
#endif
#endif

/*
 * rd_oracle.h -- FP64 CPU oracle for the rod-dynamics contact/step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (librdk) never links, loads or calls it.
 *
 * It restates, in double precision, the reference algorithms of
 * yeonsu-jung/rod-dynamics-3d (citations are to that repository):
 *   - capsule axis / endpoints      src/physics/rigid_body.cpp:112-139, include/physics/rigid_body.hpp:69-75
 *   - minimum-image shift           src/physics/contact_detector.cpp:827-839
 *   - Lumelsky segment-segment      src/physics/contact_detector.cpp:978-1035
 *   - capsule narrow phase          src/physics/contact_detector.cpp:811-910
 *   - naive O(N^2) detection        src/physics/contact_detector.cpp:213-281
 *   - spatial-hash detection        src/physics/contact_detector.cpp:366-399,473-480,487-809
 *   - NSC manifold build            src/physics/nsc_contact.cpp:78-117,239-321
 *   - wall manifold                 src/physics/nsc_contact.cpp:190-235, src/physics/collision.cpp:155-195
 *   - serial PSOR                   src/physics/nsc_contact.cpp:325-655
 *   - NSC kinematics                src/app/main.cpp:4397-4408,4469-4483,4508-4527
 *   - soft penalty + friction       src/physics/soft_contact.cpp:80-486
 *   - Hertz-Mindlin                 src/physics/hertz_mindlin.cpp:183-393, include/config/config.hpp:54-63
 *   - Velocity-Verlet split         src/physics/integrator.cpp:23-159, src/app/main.cpp:4529-4810
 *   - kinetic energy                src/app/main.cpp:3357-3368
 *
 * Body arrays are interleaved per body: x[3N], q[4N] as (w,x,y,z), v[3N],
 * w[3N], f[3N], tau[3N]; r[N], h[N] (half length of the cylinder part),
 * mass[N] (0 = static body), restitution[N].
 */
#ifndef RD_ORACLE_H
#define RD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

typedef struct orc_nsc_params {
  double mu, beta, cfm, omega;
  double tol;            /* build-defined fixed-point residual tolerance */
  int max_sweeps;
  int check_every;       /* evaluate the residual every k sweeps */
  int friction_proj;     /* 0 = reference sequential disc clamp, 1 = Euclidean disc (scalar step),
                            2 = Euclidean disc, block step G^-1 w_t projected in G's metric,
                            3 = block step where it stays in the disc, else the scalar step of 1 */
  int stop_rule;         /* 0 = reference rule (max sum|dlambda| < 1e-6), 1 = residual <= tol */
  int position_stabilization; /* only scales beta (x0.01) as the reference does */
  int warm_start;        /* 0 = cold, 1 = previous-step key match (reference) */
  int warm_walls;        /* 1 = plane-wall manifolds also warm-started, keyed (rod, slot) */
  int position_iters;    /* position projection: outer passes (config.hpp:78) */
  int position_psor_iters; /* inner normal-only sweeps per pass (config.hpp:79) */
  double slop;           /* allowed penetration before correction (config.hpp:80) */
} orc_nsc_params;

typedef struct orc_soft_params {
  double delta, k_scaler, damping, mu, mu_static, nu, kt, vel_deadband;
  int enable_friction, friction_model; /* 0 smooth (default), 1 Cundall-Strack, 2 Karnopp */
} orc_soft_params;

typedef struct orc_hm_params {
  double E, nu_poisson, restitution, mu_k, mu_s, v_c, mu_roll;
  int enable_tangential, enable_rolling;
  int law; /* 0 = reference (F = k_n*delta^1.5 with k_n ~ sqrt(delta), i.e. delta^2), 1 = Hertz delta^1.5 */
} orc_hm_params;

orc_world* orc_create(int n, const double* x, const double* q, const double* v,
                      const double* w, const double* r, const double* h,
                      const double* mass, const double* restitution);
void orc_destroy(orc_world* W);
void orc_set_box(orc_world* W, const double* lo, const double* hi,
                 const int* periodic, int floor_wall, int ceil_wall);
void orc_set_physics(orc_world* W, const double* gravity, double dt,
                     double lin_damp, double ang_damp, double w_max);
void orc_set_state(orc_world* W, const double* x, const double* q,
                   const double* v, const double* w);
void orc_set_forces(orc_world* W, const double* f, const double* tau);
void orc_get_state(orc_world* W, double* x, double* q, double* v, double* w,
                   double* f, double* tau);

/* geometry primitives (exposed for closed-form tests) */
void orc_axis(const double* q, double* u);
void orc_segseg(const double* a1, const double* a2, const double* b1,
                const double* b2, double* s, double* t);

/* detection: method 0 = naive all-pairs, 1 = reference spatial hash.
 * Result is sorted by (a,b) and unique; returns the contact count. */
int orc_detect(orc_world* W, int method, double delta, double cell_size);
int orc_get_contacts(orc_world* W, int cap, int* a, int* b, double* pa,
                     double* pb, double* nrm, double* dist, double* shift,
                     double* s, double* t, int* type);

/* NSC */
int orc_nsc_build(orc_world* W, const orc_nsc_params* P);
int orc_nsc_num_manifolds(orc_world* W);
/* rec: ORC_MF_FIELDS doubles per manifold; ab: 2 ints (b = -1 - slot for a wall) */
int orc_nsc_get_manifolds(orc_world* W, int cap, double* rec, int* ab,
                          double* lambda);
int orc_nsc_set_manifolds(orc_world* W, int m, const double* rec,
                          const int* ab, const double* lambda);
void orc_nsc_apply_warm(orc_world* W);
/* order: optional permutation of manifolds for the sweep (NULL = natural) */
int orc_nsc_sweeps(orc_world* W, const orc_nsc_params* P, const int* order,
                   int nsweeps);
int orc_nsc_solve(orc_world* W, const orc_nsc_params* P, const int* order,
                  double* residual_out, int* sweeps_out);
/* One island solved as the library's island kernels solve it (stop rule 1):
 * the island's manifolds ix[0..m) in solve order, the residual over them
 * checked every check_every sweeps, the stall rule (stall_sweeps, stall_omega;
 * 0 = off) and, when hot != 0, the hot-set phase of the LDS island class
 * (hot == 2: also its selective rounds).
 * Returns the sweeps run; *res = final residual over the island, *updates =
 * manifold updates, *hot_sweeps = sweeps of hot sets. */
int orc_nsc_solve_island(orc_world* W, const orc_nsc_params* P, const int* ix, int m, int stall_sweeps,
                         double stall_omega, int hot, double* res, long long* updates, int* hot_sweeps);
double orc_nsc_residual(orc_world* W, const orc_nsc_params* P);
/* The metric disc projection of friction_proj 2: y <- argmin over |x| <= r of
 * (x - y)^T G (x - y), G = [[a, b], [b, c]] SPD, for |y| > r > 0. */
void orc_disc_project_metric(double a, double b, double c, double r, double* y);
void orc_nsc_store_warm(orc_world* W);
int orc_nsc_step(orc_world* W, const orc_nsc_params* P, int nsteps,
                 int detect_method, double cell_size, double* sweeps_total,
                 double* residual_max);

/* integrator stages: 0 NSC free flight, 1 NSC position update (+wrap, clear),
 * 2 VV half kick + drift, 3 VV second half kick (+damping, clamp, clear),
 * 4 clear forces */
void orc_integrate(orc_world* W, int stage);
/* NscContactSolver::projectPositions, multi-body branch (nsc_contact.cpp:
 * 718-880): the pairs are the current contact list (detect first, delta 0).
 * order: sweep order of the first pass's position manifolds (NULL = sorted
 * pair order, the reference's).  Returns the number of passes that found
 * penetrating pairs; npos (may be NULL) receives the first pass's count. */
int orc_nsc_project(orc_world* W, const orc_nsc_params* P, const int* order, int* npos);

/* soft penalty / Hertz-Mindlin force accumulation into f, tau */
int orc_soft_forces(orc_world* W, const orc_soft_params* P, double* pe);
int orc_hm_forces(orc_world* W, const orc_hm_params* P, double* pe);
int orc_hm_history_size(orc_world* W);
/* Velocity-Verlet step with double force evaluation. model 0 soft, 1 HM */
int orc_vv_step(orc_world* W, int model, const void* params, int nsteps,
                int detect_method, double cell_size);

double orc_kinetic_energy(orc_world* W);
double orc_packing_fraction(orc_world* W);
/* entanglement observer: Gauss linking number of two segments, and the sum
 * of |Lk| over rod pairs (all pairs for cutoff <= 0) with the pair count */
double orc_linking_number(const double* a0, const double* a1, const double* b0, const double* b1);
double orc_entanglement(orc_world* W, double cutoff, long long* pairs);

/* reference populate "nonoverlap" (main.cpp:1857-2140), FP32, serial:
 * returns rods placed; c, u: n x 3 centres and unit directions */
int orc_populate_nonoverlap(int n, float length, float diameter, const float* bmin, const float* bmax,
                            int pbc, float cell_size, unsigned seed, int max_attempts, float* c,
                            float* u, long long* attempts);
float orc_rsa_seg_d2(const float* p0, const float* p1, const float* q0, const float* q1, int pbc_rule);
void orc_mt_uniform_float(unsigned seed, int k, float* out);
unsigned orc_mt_raw(unsigned seed, long long k);
/* reference randomInit (main.cpp:2213-2294): mode 0 gaussian, 1 thermal, 2 uniform */
void orc_random_init(int n, const float* q, const float* mass, const float* i_perp, int mode, float v_sigma,
                     float w_sigma, float w_speed, float kbt_trans, float kbt_rot, unsigned seed, int project,
                     const unsigned char* skip, float* v, float* w);
void orc_mt_normal_float(unsigned seed, int k, float* out);

#define ORC_MF_FIELDS 42

#ifdef __cplusplus
}
#endif
#endif

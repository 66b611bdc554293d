// facade_check.cpp -- C entry points that drive the C++ classes of
// include/rdk_physics.hpp (ContactDetector, NscContactSolver), so that tests
// can run the reference-shaped interface and compare it with the C ABI used
// directly.  Built into librdk_facade_check.so beside librdk.so.
#include <cstring>
#include <string>
#include <vector>

#include "rdk_physics.hpp"

namespace {

std::vector<rdk::RigidBody> bodies_from(int n, const double* x, const double* q, const double* v, const double* w,
                                        const double* r, const double* h, const double* mass, const double* rest) {
  std::vector<rdk::RigidBody> B(n);
  for (int i = 0; i < n; ++i) {
    B[i].x = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    B[i].v = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
    B[i].w = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
    B[i].q = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    B[i].radius = r[i];
    B[i].half_length = h[i];
    B[i].mass = mass[i];
    B[i].restitution = rest[i];
  }
  return B;
}

int fail(const rdk::Error& e, char* err, int cap) {
  if (err && cap > 0) {
    std::strncpy(err, e.what(), (size_t)cap - 1);
    err[cap - 1] = '\0';
  }
  return e.status();
}

}  // namespace

extern "C" {

// ContactDetector::detectContacts on a periodic box (pbc != 0) or the open
// domain; writes up to cap contacts (a, b, distance) and their count.
int rdk_facade_detect(int n, const double* x, const double* q, const double* r, const double* h, const double* lo,
                      const double* hi, int pbc, double delta, double chunk_len, int cap, int* a, int* b,
                      double* dist, int* count, char* err, int err_cap) {
  try {
    std::vector<double> zero(3 * (size_t)n, 0.0), one(n, 1.0), none(n, 0.0);
    auto B = bodies_from(n, x, q, zero.data(), zero.data(), r, h, one.data(), none.data());
    rdk::ContactDetectionCfg cfg;
    cfg.delta = delta;
    cfg.chunk_len = chunk_len;
    rdk::ContactDetector det(cfg);
    det.setPBC(pbc != 0, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]});
    det.detectContacts(B);
    const auto& C = det.getContacts();
    *count = (int)C.size();
    for (int i = 0; i < (int)C.size() && i < cap; ++i) {
      a[i] = C[i].body_a;
      b[i] = C[i].body_b;
      dist[i] = C[i].distance;
    }
    return RDK_OK;
  } catch (const rdk::Error& e) {
    return fail(e, err, err_cap);
  }
}

// `frames` times NscContactSolver::detectAndBuildManifolds then
// solveVelocities(dt) on the same positions (the second frame starts from the
// first's impulses); the velocities after the last frame, the manifold count
// and the solver's last residual.
int rdk_facade_nsc_frames(int n, const double* x, const double* q, const double* v, const double* w, const double* r,
                          const double* h, const double* mass, const double* rest, const double* lo, const double* hi,
                          const int* periodic, int floor_wall, int ceil_wall, double mu, double dt, int friction_proj,
                          int stop_rule, double tol, int max_sweeps, int check_every, int frames, double* v_out,
                          double* w_out, int* n_manifolds, int* n_impacts, double* residual, char* err,
                          int err_cap) {
  try {
    auto B = bodies_from(n, x, q, v, w, r, h, mass, rest);
    rdk::NscContactCfg cfg;
    cfg.mu = mu;
    cfg.beta = 0.0;
    cfg.cfm = 0.0;
    cfg.omega = 1.3;
    cfg.velocity_iters = max_sweeps;
    cfg.position_stabilization = false;
    cfg.friction_proj = friction_proj;
    cfg.stop_rule = stop_rule;
    cfg.tol = tol;
    cfg.check_every = check_every;
    rdk::NscContactSolver S;
    S.setConfig(cfg);
    S.setBox(true, {lo[0], lo[1], lo[2]}, {hi[0], hi[1], hi[2]}, periodic, floor_wall != 0, ceil_wall != 0);
    for (int f = 0; f < frames; ++f) {
      S.detectAndBuildManifolds(B);
      S.solveVelocities(B, dt);
    }
    for (int i = 0; i < n; ++i) {
      v_out[3 * i] = B[i].v.x, v_out[3 * i + 1] = B[i].v.y, v_out[3 * i + 2] = B[i].v.z;
      w_out[3 * i] = B[i].w.x, w_out[3 * i + 1] = B[i].w.y, w_out[3 * i + 2] = B[i].w.z;
    }
    *n_manifolds = (int)S.getNumContacts();
    *n_impacts = S.countImpacts();
    *residual = S.getLastResidual();
    return RDK_OK;
  } catch (const rdk::Error& e) {
    return fail(e, err, err_cap);
  }
}

}  // extern "C"

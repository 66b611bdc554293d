/* rdk_physics.hpp -- the reference's contact classes over librdk's C ABI.
 *
 * ContactDetector (include/physics/contact_detector.hpp:64-75) and
 * NscContactSolver (include/physics/nsc_contact.hpp:71-157) with their names,
 * call sequence and meaning, for C++ callers of the reference's physics
 * interface.  Bodies are a glm-free FP64 record holding the fields of
 * RigidBody (include/physics/rigid_body.hpp) that the contact path reads and
 * writes; every call uploads them to the context's device arrays and
 * downloads what the reference writes back (velocities after
 * solveVelocities, positions and orientations after projectPositions).  A
 * failing library call throws rdk::Error with the library's status and
 * message.  Header-only; link librdk.so.
 *
 * Differences from the reference, all on the library's side of the boundary:
 * contacts and manifolds are FP64; without setPBC the detection box is the
 * bodies' bounding box, grown by the largest rod extent, closed on every
 * axis; plane walls at the box's z faces come from setBox (the reference adds
 * wall manifolds from the application, nsc_contact.hpp:128-140). */
#pragma once

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rdk.h"

namespace rdk {

struct Vec3 {
  double x = 0.0, y = 0.0, z = 0.0;
};
struct Quat {  // w + x i + y j + z k; the capsule axis is the body's local +Y
  double w = 1.0, x = 0.0, y = 0.0, z = 0.0;
};

// The contact path's part of RigidBody (rigid_body.hpp): state, capsule radius
// and half length, mass (0: static), coefficient of restitution.
struct RigidBody {
  Vec3 x, v, w;
  Quat q;
  double radius = 0.5, half_length = 0.5;
  double mass = 1.0;
  double restitution = 0.0;
};

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// ContactDetectionCfg (contact_detector.hpp:43-50): the activation margin;
// the broad-phase shape is the library's (chunk length, cell scale).
struct ContactDetectionCfg {
  double delta = 0.005;
  double chunk_len = 0.0;  // <= 0: one chunk per rod
  double cell_scale = 1.0;
  int device = 0;
};

enum class ContactType { POINT_TO_POINT = 0, EDGE_TO_POINT = 1, EDGE_TO_EDGE = 2 };

// ContactPrimitive (contact_detector.hpp:28-41), geometry only
struct ContactPrimitive {
  ContactType type = ContactType::POINT_TO_POINT;
  int body_a = 0, body_b = 0;
  Vec3 point_a, point_b, normal;
  double distance = 0.0;
  Vec3 shift_b;  // minimum-image translation applied to body_b
  double s = 0.0, t = 0.0;  // closest-point parameters on the two axes
};

struct ContactBroadphaseStats {
  int chunks = 0, cells = 0, regrows = 0, contacts = 0;
  int64_t candidates = 0;  // sphere-prefilter survivors
};

namespace detail {

inline void check(int st, rdk_ctx* ctx, const char* what) {
  if (st != RDK_OK) throw Error(st, std::string(what) + ": " + (ctx ? rdk_last_error(ctx) : "no context"));
}

// One device context (one HIP stream); not copyable.
class Context {
 public:
  explicit Context(int device) { check(rdk_create(device, &p_), nullptr, "rdk_create"); }
  ~Context() {
    if (p_) rdk_destroy(p_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  rdk_ctx* get() const { return p_; }

 private:
  rdk_ctx* p_ = nullptr;
};

struct Box {
  bool set = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int periodic[3] = {0, 0, 0};
  int floor_wall = 0, ceil_wall = 0;
};

// The bodies' fixed properties last given to the context: while they stay the
// same, a call uploads only the state, so the context keeps the previous
// frame's impulses (rdk_set_bodies starts a new scene).
struct Scene {
  int n = -1;
  std::vector<double> r, h, m, e;
};

// bodies -> the interleaved FP64 arrays of rdk.h, and the system box
inline void upload_bodies(rdk_ctx* ctx, Scene& scene, const std::vector<RigidBody>& B, const Box& box, double mu,
                          double margin) {
  const int n = (int)B.size();
  std::vector<double> x(3 * n), q(4 * n), v(3 * n), w(3 * n), r(n), h(n), m(n), e(n);
  double reach = 0.0;
  for (int i = 0; i < n; ++i) {
    const RigidBody& b = B[i];
    x[3 * i] = b.x.x, x[3 * i + 1] = b.x.y, x[3 * i + 2] = b.x.z;
    v[3 * i] = b.v.x, v[3 * i + 1] = b.v.y, v[3 * i + 2] = b.v.z;
    w[3 * i] = b.w.x, w[3 * i + 1] = b.w.y, w[3 * i + 2] = b.w.z;
    q[4 * i] = b.q.w, q[4 * i + 1] = b.q.x, q[4 * i + 2] = b.q.y, q[4 * i + 3] = b.q.z;
    r[i] = b.radius, h[i] = b.half_length, m[i] = b.mass, e[i] = b.restitution;
    reach = std::max(reach, b.radius + b.half_length);
  }
  if (n != scene.n || r != scene.r || h != scene.h || m != scene.m || e != scene.e) {
    check(rdk_set_bodies(ctx, n, x.data(), q.data(), v.data(), w.data(), r.data(), h.data(), m.data(), e.data()),
          ctx, "rdk_set_bodies");
    scene.n = n;
    scene.r = r, scene.h = h, scene.m = m, scene.e = e;
  } else {
    check(rdk_upload_state(ctx, x.data(), q.data(), v.data(), w.data()), ctx, "rdk_upload_state");
  }
  rdk_system_desc d{};
  if (box.set) {
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = box.lo[k];
      d.hi[k] = box.hi[k];
      d.periodic[k] = box.periodic[k];
    }
    d.floor_wall = box.floor_wall;
    d.ceil_wall = box.ceil_wall;
  } else {  // open domain: the bounding box of the centres, grown by the longest reach
    const double g = 2.0 * reach + margin + 1.0;
    for (int k = 0; k < 3; ++k) {
      double lo = 0.0, hi = 1.0;
      for (int i = 0; i < n; ++i) {
        const double c = x[3 * i + k];
        lo = i == 0 ? c : std::min(lo, c);
        hi = i == 0 ? c : std::max(hi, c);
      }
      d.lo[k] = lo - g;
      d.hi[k] = hi + g;
      d.periodic[k] = 0;
    }
  }
  d.mu = mu;
  const int off[2] = {0, n};
  check(rdk_set_systems(ctx, 1, off, &d), ctx, "rdk_set_systems");
}

inline std::vector<ContactPrimitive> download_contacts(rdk_ctx* ctx, int nc) {
  std::vector<int> a(nc), b(nc), ty(nc);
  std::vector<double> pa(3 * nc), pb(3 * nc), nn(3 * nc), dist(nc), sh(3 * nc), s(nc), t(nc);
  if (nc > 0)
    check(rdk_get_contacts(ctx, nc, a.data(), b.data(), pa.data(), pb.data(), nn.data(), dist.data(), sh.data(),
                           s.data(), t.data(), ty.data()),
          ctx, "rdk_get_contacts");
  std::vector<ContactPrimitive> out(nc);
  for (int i = 0; i < nc; ++i) {
    ContactPrimitive& c = out[i];
    c.type = (ContactType)ty[i];
    c.body_a = a[i];
    c.body_b = b[i];
    c.point_a = {pa[3 * i], pa[3 * i + 1], pa[3 * i + 2]};
    c.point_b = {pb[3 * i], pb[3 * i + 1], pb[3 * i + 2]};
    c.normal = {nn[3 * i], nn[3 * i + 1], nn[3 * i + 2]};
    c.shift_b = {sh[3 * i], sh[3 * i + 1], sh[3 * i + 2]};
    c.distance = dist[i];
    c.s = s[i];
    c.t = t[i];
  }
  return out;
}

inline ContactBroadphaseStats detect_stats(rdk_ctx* ctx, int nc) {
  ContactBroadphaseStats st;
  check(rdk_detect_info(ctx, &st.chunks, &st.cells, &st.candidates, &st.regrows), ctx, "rdk_detect_info");
  st.contacts = nc;
  return st;
}

}  // namespace detail

// ContactDetector (contact_detector.hpp:64-75): detectContacts fills the
// sorted, unique contact list of the bodies (rdk_detect, rdk_get_contacts).
class ContactDetector {
 public:
  explicit ContactDetector(const ContactDetectionCfg& config = ContactDetectionCfg())
      : config_(config), ctx_(config.device) {
    rdk_physics ph{};
    ph.dt = 1e-3;
    ph.w_max = 1e30;
    detail::check(rdk_set_physics(ctx_.get(), &ph), ctx_.get(), "rdk_set_physics");
  }

  void setConfig(const ContactDetectionCfg& config) { config_ = config; }
  void setPBC(bool enabled, const Vec3& min, const Vec3& max) {
    box_.set = enabled;
    box_.lo[0] = min.x, box_.lo[1] = min.y, box_.lo[2] = min.z;
    box_.hi[0] = max.x, box_.hi[1] = max.y, box_.hi[2] = max.z;
    for (int k = 0; k < 3; ++k) box_.periodic[k] = enabled ? 1 : 0;
    box_.floor_wall = box_.ceil_wall = 0;
  }

  void detectContacts(const std::vector<RigidBody>& bodies) {
    rdk_ctx* c = ctx_.get();
    detail::upload_bodies(c, scene_, bodies, box_, 0.0, config_.delta);
    const rdk_detect_params dp{config_.delta, config_.chunk_len, config_.cell_scale};
    int nc = 0;
    detail::check(rdk_detect(c, &dp, &nc), c, "rdk_detect");
    contacts_ = detail::download_contacts(c, nc);
    stats_ = detail::detect_stats(c, nc);
  }

  const std::vector<ContactPrimitive>& getContacts() const { return contacts_; }
  std::vector<ContactPrimitive>& accessContacts() { return contacts_; }
  size_t getNumContacts() const { return contacts_.size(); }
  const ContactBroadphaseStats& getStats() const { return stats_; }
  rdk_ctx* context() const { return ctx_.get(); }

 private:
  ContactDetectionCfg config_;
  detail::Context ctx_;
  detail::Scene scene_;
  detail::Box box_;
  std::vector<ContactPrimitive> contacts_;
  ContactBroadphaseStats stats_;
};

// NscContactCfg (config.hpp:66-88) with the reference's defaults; the fields
// after enable_warm_start select the library's variants (the defaults are the
// reference's algorithm: sequential friction clamp, stop rule max |dlambda|).
struct NscContactCfg {
  double mu = 0.3;
  double beta = 0.2;
  double cfm = 0.0;
  double omega = 1.0;
  int velocity_iters = 40;
  bool position_stabilization = true;
  int position_iters = 5;
  int position_psor_iters = 50;
  double slop = 1e-4;
  bool enable_warm_start = true;
  // library variants (rdk_nsc_params)
  int friction_proj = 0;  // 0 reference clamp; 1-3 Euclidean disc steps
  int stop_rule = 0;      // 0 reference (max sum |dlambda| < 1e-6), 1 residual <= tol
  double tol = 1e-8;
  int check_every = 1;
  int stall_sweeps = 0;
  double stall_omega = 1.0;
  int hot_set = 0;
  int warm_walls = 0;
  double delta = 0.0;  // contacts at strict overlap (nsc_contact.cpp:24)
  double chunk_len = 0.0;
  int device = 0;
};

// NscManifold (nsc_contact.hpp:37-63) in FP64
struct NscManifold {
  int body_a = 0, body_b = -1;  // body_b < 0: plane wall
  Vec3 normal, t1, t2, r_a, r_b;
  double phi = 0.0, v_n_pre = 0.0, restitution = 0.0;
  double g_n = 0.0, g_t1 = 0.0, g_t2 = 0.0;
  Vec3 IinvA_rAxn, IinvA_rAxt1, IinvA_rAxt2, IinvB_rBxn, IinvB_rBxt1, IinvB_rBxt2;
  double invMassA = 0.0, invMassB = 0.0;
  double lambda_n = 0.0, lambda_t1 = 0.0, lambda_t2 = 0.0;
  bool isWall = false;
};

// NscContactSolver (nsc_contact.hpp:71-157): detectAndBuildManifolds
// (rdk_detect + rdk_nsc_build, previous-frame impulses kept by the context),
// solveVelocities (rdk_nsc_solve on the bodies' current velocities, written
// back), projectPositions (rdk_nsc_project, positions written back).
class NscContactSolver {
 public:
  NscContactSolver() = default;

  void setConfig(const NscContactCfg& cfg) {
    cfg_ = cfg;
    ctx_.reset();  // a new context on the configured device at the next call
    scene_ = detail::Scene();
  }
  void setPBC(bool enabled, const Vec3& min, const Vec3& max) {
    const int p[3] = {enabled ? 1 : 0, enabled ? 1 : 0, enabled ? 1 : 0};
    setBox(enabled, min, max, p, false, false);
  }
  // periodic per axis and plane walls at the z faces (axis 2 then closed)
  void setBox(bool enabled, const Vec3& min, const Vec3& max, const int periodic[3], bool floor_wall,
              bool ceil_wall) {
    box_.set = enabled;
    box_.lo[0] = min.x, box_.lo[1] = min.y, box_.lo[2] = min.z;
    box_.hi[0] = max.x, box_.hi[1] = max.y, box_.hi[2] = max.z;
    for (int k = 0; k < 3; ++k) box_.periodic[k] = periodic[k];
    box_.floor_wall = floor_wall ? 1 : 0;
    box_.ceil_wall = ceil_wall ? 1 : 0;
  }

  void detectAndBuildManifolds(const std::vector<RigidBody>& bodies) {
    rdk_ctx* c = ctx(bodies.size());
    detail::upload_bodies(c, scene_, bodies, box_, cfg_.mu, cfg_.delta);
    const rdk_detect_params dp = detect_params();
    int nc = 0, nm = 0;
    detail::check(rdk_detect(c, &dp, &nc), c, "rdk_detect");
    contacts_ = detail::download_contacts(c, nc);
    const rdk_nsc_params p = params();
    detail::check(rdk_nsc_build(c, &p, &nm), c, "rdk_nsc_build");
    download_manifolds(nm);
  }

  void solveVelocities(std::vector<RigidBody>& bodies, double dt) {
    rdk_ctx* c = ctx(bodies.size());
    set_dt(dt);
    const int n = (int)bodies.size();
    std::vector<double> x(3 * n), q(4 * n), v(3 * n), w(3 * n);
    pack(bodies, x, q, v, w);
    detail::check(rdk_upload_state(c, x.data(), q.data(), v.data(), w.data()), c, "rdk_upload_state");
    const rdk_nsc_params p = params();
    rdk_solve_stats st{};
    const int rc = rdk_nsc_solve(c, &p, &st);
    if (rc != RDK_OK && rc != RDK_ERR_NOCONV) detail::check(rc, c, "rdk_nsc_solve");
    lastResidual_ = cfg_.stop_rule == 0 ? st.residual_ref : st.residual;
    detail::check(rdk_download_state(c, nullptr, nullptr, v.data(), w.data(), nullptr, nullptr), c,
                  "rdk_download_state");
    for (int i = 0; i < n; ++i) {
      bodies[i].v = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
      bodies[i].w = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
    }
    download_manifolds((int)manifolds_.size());  // impulses after the solve
  }

  void projectPositions(std::vector<RigidBody>& bodies) {
    rdk_ctx* c = ctx(bodies.size());
    const int n = (int)bodies.size();
    std::vector<double> x(3 * n), q(4 * n), v(3 * n), w(3 * n);
    pack(bodies, x, q, v, w);
    detail::check(rdk_upload_state(c, x.data(), q.data(), v.data(), w.data()), c, "rdk_upload_state");
    const rdk_detect_params dp = detect_params();
    const rdk_nsc_params p = params();
    int passes = 0, first = 0;
    detail::check(rdk_nsc_project(c, &dp, &p, &passes, &first), c, "rdk_nsc_project");
    detail::check(rdk_download_state(c, x.data(), q.data(), nullptr, nullptr, nullptr, nullptr), c,
                  "rdk_download_state");
    for (int i = 0; i < n; ++i) {
      bodies[i].x = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
      bodies[i].q = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    }
  }

  size_t getNumContacts() const { return manifolds_.size(); }
  int countImpacts(double tol = 1e-6) const {
    int n = 0;
    for (const auto& m : manifolds_)
      if (m.v_n_pre < -tol) ++n;
    return n;
  }
  const std::vector<NscManifold>& getManifolds() const { return manifolds_; }
  const std::vector<ContactPrimitive>& getDetectedContacts() const { return contacts_; }
  double getLastResidual() const { return lastResidual_; }
  rdk_ctx* context() const { return ctx_ ? ctx_->get() : nullptr; }

 private:
  rdk_ctx* ctx(size_t) {
    if (!ctx_) {
      ctx_.reset(new detail::Context(cfg_.device));
      set_dt(dt_);
    }
    return ctx_->get();
  }
  void set_dt(double dt) {
    dt_ = dt;
    if (!ctx_) return;
    rdk_physics ph{};
    ph.dt = dt;
    ph.w_max = 1e30;
    detail::check(rdk_set_physics(ctx_->get(), &ph), ctx_->get(), "rdk_set_physics");
  }
  rdk_detect_params detect_params() const { return rdk_detect_params{cfg_.delta, cfg_.chunk_len, 1.0}; }
  rdk_nsc_params params() const {
    rdk_nsc_params p{};
    p.beta = cfg_.beta;
    p.cfm = cfg_.cfm;
    p.omega = cfg_.omega;
    p.tol = cfg_.tol;
    p.max_sweeps = cfg_.velocity_iters;
    p.check_every = cfg_.check_every;
    p.friction_proj = cfg_.friction_proj;
    p.position_stabilization = cfg_.position_stabilization ? 1 : 0;
    p.warm_mode = cfg_.enable_warm_start ? 1 : 0;
    p.stop_rule = cfg_.stop_rule;
    p.warm_walls = cfg_.warm_walls;
    p.stall_sweeps = cfg_.stall_sweeps;
    p.position_iters = cfg_.position_iters;
    p.position_psor_iters = cfg_.position_psor_iters;
    p.slop = cfg_.slop;
    p.stall_omega = cfg_.stall_omega;
    p.hot_set = cfg_.hot_set;
    return p;
  }
  static void pack(const std::vector<RigidBody>& B, std::vector<double>& x, std::vector<double>& q,
                   std::vector<double>& v, std::vector<double>& w) {
    for (size_t i = 0; i < B.size(); ++i) {
      x[3 * i] = B[i].x.x, x[3 * i + 1] = B[i].x.y, x[3 * i + 2] = B[i].x.z;
      v[3 * i] = B[i].v.x, v[3 * i + 1] = B[i].v.y, v[3 * i + 2] = B[i].v.z;
      w[3 * i] = B[i].w.x, w[3 * i + 1] = B[i].w.y, w[3 * i + 2] = B[i].w.z;
      q[4 * i] = B[i].q.w, q[4 * i + 1] = B[i].q.x, q[4 * i + 2] = B[i].q.y, q[4 * i + 3] = B[i].q.z;
    }
  }
  void download_manifolds(int nm) {
    rdk_ctx* c = ctx_->get();
    std::vector<double> rec((size_t)RDK_MF_FIELDS * nm), lam(3 * (size_t)nm);
    std::vector<int> ab(2 * (size_t)nm);
    if (nm > 0)
      detail::check(rdk_nsc_get_manifolds(c, nm, rec.data(), ab.data(), lam.data()), c, "rdk_nsc_get_manifolds");
    manifolds_.assign(nm, NscManifold{});
    for (int i = 0; i < nm; ++i) {
      const double* f = rec.data() + (size_t)RDK_MF_FIELDS * i;
      auto v3 = [&](int k) { return Vec3{f[k], f[k + 1], f[k + 2]}; };
      NscManifold& m = manifolds_[i];
      m.body_a = ab[2 * i];
      m.body_b = ab[2 * i + 1];
      m.isWall = m.body_b < 0;
      m.normal = v3(0), m.t1 = v3(3), m.t2 = v3(6), m.r_a = v3(9), m.r_b = v3(12);
      m.IinvA_rAxn = v3(15), m.IinvA_rAxt1 = v3(18), m.IinvA_rAxt2 = v3(21);
      m.IinvB_rBxn = v3(24), m.IinvB_rBxt1 = v3(27), m.IinvB_rBxt2 = v3(30);
      m.invMassA = f[33], m.invMassB = f[34];
      m.g_n = f[35], m.g_t1 = f[36], m.g_t2 = f[37];
      m.phi = f[38], m.v_n_pre = f[39], m.restitution = f[40];
      m.lambda_n = lam[3 * i], m.lambda_t1 = lam[3 * i + 1], m.lambda_t2 = lam[3 * i + 2];
    }
  }

  NscContactCfg cfg_;
  std::unique_ptr<detail::Context> ctx_;
  detail::Scene scene_;
  detail::Box box_;
  double dt_ = 1e-3;
  std::vector<ContactPrimitive> contacts_;
  std::vector<NscManifold> manifolds_;
  double lastResidual_ = 0.0;
};

}  // namespace rdk

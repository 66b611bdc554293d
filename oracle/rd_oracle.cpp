// rd_oracle.cpp -- FP64 CPU oracle (test infrastructure only; see rd_oracle.h).
//
// Every arithmetic expression keeps the operand order of the reference source
// it cites, and the file is compiled with -ffp-contract=off, so that a device
// implementation written with the same operation order can be compared
// bit for bit where the algorithm is deterministic (detection, integrators).
#include "rd_oracle.h"

#include <algorithm>
#include <limits>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
  double x, y, z;
};
static inline V3 mk(double a, double b, double c) { return V3{a, b, c}; }
static inline V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul(V3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
// glm::dot for vec3: (x*x + y*y) + z*z
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::cross
static inline V3 cross(V3 a, V3 b) {
  return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline double length(V3 a) { return std::sqrt(dot(a, a)); }
// glm::normalize(vec3) = v * inversesqrt(dot(v,v)), inversesqrt(x) = 1/sqrt(x)
static inline V3 normalize(V3 a) { return mul(a, 1.0 / std::sqrt(dot(a, a))); }

struct Q4 {
  double w, x, y, z;
};
// glm quaternion product p*q (glm/detail/type_quat.inl operator*=)
static inline Q4 qmul(Q4 p, Q4 q) {
  Q4 r;
  r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
  r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
  r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
  r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
  return r;
}
// glm::normalize(quat): len = sqrt(dot); multiply by 1/len
static inline Q4 qnormalize(Q4 q) {
  double len = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (len <= 0.0) return Q4{1, 0, 0, 0};
  double inv = 1.0 / len;
  return Q4{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}

// Column-major 3x3 like glm: m[c][r]
struct M3 {
  double m[3][3];
};
// glm::mat3_cast (glm/gtc/quaternion.inl)
static inline M3 mat3_cast(Q4 q) {
  double qxx = q.x * q.x, qyy = q.y * q.y, qzz = q.z * q.z;
  double qxz = q.x * q.z, qxy = q.x * q.y, qyz = q.y * q.z;
  double qwx = q.w * q.x, qwy = q.w * q.y, qwz = q.w * q.z;
  M3 R;
  R.m[0][0] = 1.0 - 2.0 * (qyy + qzz);
  R.m[0][1] = 2.0 * (qxy + qwz);
  R.m[0][2] = 2.0 * (qxz - qwy);
  R.m[1][0] = 2.0 * (qxy - qwz);
  R.m[1][1] = 1.0 - 2.0 * (qxx + qzz);
  R.m[1][2] = 2.0 * (qyz + qwx);
  R.m[2][0] = 2.0 * (qxz + qwy);
  R.m[2][1] = 2.0 * (qyz - qwx);
  R.m[2][2] = 1.0 - 2.0 * (qxx + qyy);
  return R;
}
static inline M3 mmul(const M3& A, const M3& B) {
  M3 C;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r)
      C.m[c][r] = A.m[0][r] * B.m[c][0] + A.m[1][r] * B.m[c][1] + A.m[2][r] * B.m[c][2];
  return C;
}
static inline M3 transpose(const M3& A) {
  M3 T;
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) T.m[c][r] = A.m[r][c];
  return T;
}
static inline V3 mv(const M3& M, V3 v) {
  return mk(M.m[0][0] * v.x + M.m[1][0] * v.y + M.m[2][0] * v.z,
            M.m[0][1] * v.x + M.m[1][1] * v.y + M.m[2][1] * v.z,
            M.m[0][2] * v.x + M.m[1][2] * v.y + M.m[2][2] * v.z);
}
static inline M3 diag(double a, double b, double c) {
  M3 D;
  std::memset(&D, 0, sizeof(D));
  D.m[0][0] = a;
  D.m[1][1] = b;
  D.m[2][2] = c;
  return D;
}

struct Body {
  V3 x, v, w, f, tau;
  Q4 q;
  double r, h, mass, invMass, rest;
  double Ixx, Iyy, Ixx_inv, Iyy_inv;  // solid cylinder about local +Y
};

// rigid_body.cpp:11-47: solid-cylinder inertia with L = 2h, axis local +Y
static void set_mass(Body& b, double mass) {
  b.mass = mass;
  if (mass > 0.0) {
    b.invMass = 1.0 / mass;
    double L = 2.0 * b.h;
    b.Ixx = mass * (3.0 * b.r * b.r + L * L) / 12.0;
    b.Iyy = mass * (b.r * b.r) / 2.0;
    b.Ixx_inv = b.Ixx > 0 ? 1.0 / b.Ixx : 0.0;
    b.Iyy_inv = b.Iyy > 0 ? 1.0 / b.Iyy : 0.0;
  } else {
    b.invMass = 0.0;
    b.Ixx = b.Iyy = b.Ixx_inv = b.Iyy_inv = 0.0;
  }
}

// rigid_body.cpp:116-121
static M3 iworld_inv(const Body& b) {
  if (b.invMass <= 0.0) return diag(0, 0, 0);
  M3 R = mat3_cast(b.q);
  return mmul(mmul(R, diag(b.Ixx_inv, b.Iyy_inv, b.Ixx_inv)), transpose(R));
}
static M3 iworld(const Body& b) {
  M3 R = mat3_cast(b.q);
  return mmul(mmul(R, diag(b.Ixx, b.Iyy, b.Ixx)), transpose(R));
}
// rigid_body.cpp:137-139 axisY = R()[1]
static V3 axis_of(Q4 q) {
  double qxx = q.x * q.x, qzz = q.z * q.z, qxy = q.x * q.y;
  double qyz = q.y * q.z, qwx = q.w * q.x, qwz = q.w * q.z;
  return mk(2.0 * (qxy - qwz), 1.0 - 2.0 * (qxx + qzz), 2.0 * (qyz + qwx));
}

struct Contact {
  int a, b, type;
  V3 pa, pb, n, shift;
  double dist, hsum, s, t;
};

struct Manifold {
  int a, b;  // b = -1 for wall
  V3 n, t1, t2, ra, rb;
  V3 IAn, IAt1, IAt2, IBn, IBt1, IBt2;
  double invMA, invMB, gn, gt1, gt2, phi, vnpre, e;
  double gtb;  // largest eigenvalue of the (t1,t2) Delassus block (incl. cfm)
  double ln, lt1, lt2;
};

struct HistEntry {
  V3 xi;
  uint64_t last_frame;
};

}  // namespace

struct orc_world {
  std::vector<Body> B;
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  int periodic[3] = {0, 0, 0};
  int floor_wall = 0, ceil_wall = 0;
  V3 g{0, 0, 0};
  double dt = 1e-3, lin_damp = 0, ang_damp = 0, w_max = 1e30;
  std::vector<Contact> contacts;
  std::vector<Manifold> mf;
  std::unordered_map<uint64_t, std::array<double, 3>> warm;
  std::unordered_map<uint64_t, std::array<double, 3>> warm_wall;  // key (rod << 2) | slot
  std::unordered_map<uint64_t, HistEntry> hm_hist;
  uint64_t hm_frame = 0;
  std::unordered_map<uint64_t, V3> cs_hist;  // Cundall-Strack, never pruned
};

namespace {

static inline uint64_t pair_key(int a, int b) {
  if (a > b) std::swap(a, b);
  return (uint64_t(uint32_t(a)) << 32) | uint64_t(uint32_t(b));
}

// contact_detector.cpp:978-1035 (Lumelsky, as in DisMech)
static void segseg(V3 a1, V3 a2, V3 b1, V3 b2, double& s, double& t) {
  const V3 e1 = sub(a2, a1);
  const V3 e2 = sub(b2, b1);
  const V3 e12 = sub(b1, a1);
  const double D1 = dot(e1, e1);
  const double D2 = dot(e2, e2);
  const double S1 = dot(e1, e12);
  const double S2 = dot(e2, e12);
  const double R = dot(e1, e2);
  const double den = D1 * D2 - R * R;
  auto fixBound = [](double& x) -> bool {
    if (x > 1.0) {
      x = 1.0;
      return true;
    } else if (x < 0.0) {
      x = 0.0;
      return true;
    }
    return false;
  };
  double uf;
  if (den == 0.0) {
    s = 0.0;
    t = -S2 / D2;
    uf = t;
    fixBound(uf);
    if (uf != t) {
      s = (uf * R + S1) / D1;
      fixBound(s);
      t = uf;
    }
  } else {
    s = (S1 * D2 - S2 * R) / den;
    fixBound(s);
    t = (s * R - S2) / D2;
    uf = t;
    fixBound(uf);
    if (uf != t) {
      s = (uf * R + S1) / D1;
      fixBound(s);
      t = uf;
    }
  }
}

// contact_detector.cpp:811-910 in FP64.  Returns true and fills c if active.
static bool capsule_capsule(const orc_world& W, int ia, int ib, double delta, Contact& c) {
  const Body& A = W.B[ia];
  const Body& Bd = W.B[ib];
  const V3 ua = axis_of(A.q), ub = axis_of(Bd.q);
  const V3 a1 = sub(A.x, mul(ua, A.h));
  const V3 a2 = add(A.x, mul(ua, A.h));
  V3 b1 = sub(Bd.x, mul(ub, Bd.h));
  V3 b2 = add(Bd.x, mul(ub, Bd.h));
  V3 shift{0, 0, 0};
  const double d3[3] = {Bd.x.x - A.x.x, Bd.x.y - A.x.y, Bd.x.z - A.x.z};
  double sh[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k) {
    double size = W.hi[k] - W.lo[k];
    if (W.periodic[k] && size > 0.0) {
      double nn = std::floor(d3[k] / size + 0.5);
      sh[k] = -nn * size;
    }
  }
  shift = mk(sh[0], sh[1], sh[2]);
  b1 = add(b1, shift);
  b2 = add(b2, shift);
  const double hs = A.r + Bd.r;
  double s, t;
  segseg(a1, a2, b1, b2, s, t);
  const V3 pa = add(a1, mul(sub(a2, a1), s));
  const V3 pb = add(b1, mul(sub(b2, b1), t));
  const V3 diff = sub(pb, pa);
  const double d2 = dot(diff, diff);
  const double act = hs + delta;
  if (!(d2 < act * act)) return false;
  const double dist = std::sqrt(d2);
  const double eps = 1e-6;
  int type;
  if (s < eps && t < eps) type = 0;
  else if (s > 1.0 - eps && t > 1.0 - eps) type = 0;
  else if (s < eps || s > 1.0 - eps) type = 1;
  else if (t < eps || t > 1.0 - eps) type = 1;
  else type = 2;
  c.a = ia;
  c.b = ib;
  c.type = type;
  c.pa = pa;
  c.pb = pb;
  c.dist = dist;
  c.hsum = hs;
  c.shift = shift;
  c.s = s;
  c.t = t;
  if (dist > 1e-10) c.n = mk(diff.x / dist, diff.y / dist, diff.z / dist);
  else c.n = mk(1, 0, 0);
  return true;
}

static void sort_unique(std::vector<Contact>& v) {
  std::sort(v.begin(), v.end(), [](const Contact& x, const Contact& y) {
    if (x.a != y.a) return x.a < y.a;
    return x.b < y.b;
  });
  v.erase(std::unique(v.begin(), v.end(),
                      [](const Contact& x, const Contact& y) { return x.a == y.a && x.b == y.b; }),
          v.end());
}

// contact_detector.cpp:213-281
static void detect_naive(orc_world& W, double delta) {
  const int n = (int)W.B.size();
  std::vector<std::vector<Contact>> per(n);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    for (int j = i + 1; j < n; ++j) {
      if (W.B[i].invMass <= 0.0 && W.B[j].invMass <= 0.0) continue;
      Contact c;
      if (capsule_capsule(W, i, j, delta, c)) per[i].push_back(c);
    }
  }
  W.contacts.clear();
  for (auto& p : per) W.contacts.insert(W.contacts.end(), p.begin(), p.end());
  sort_unique(W.contacts);
}

// contact_detector.cpp:366-399
static double adaptive_cell_size(const orc_world& W) {
  const size_t n = W.B.size();
  if (n == 0) return 1.0;
  double max_dim = 0, sum_dim = 0;
  int count = 0;
  size_t step = (n > 1000) ? (n / 100) : 1;
  for (size_t i = 0; i < n; i += step) {
    double dim = W.B[i].h * 2.0 + W.B[i].r * 2.0;
    max_dim = std::max(max_dim, dim);
    sum_dim += dim;
    count++;
  }
  double avg = count > 0 ? sum_dim / count : 0.0;
  return std::max(max_dim, avg * 1.5);
}

// contact_detector.cpp:473-480
static inline uint64_t hash_pos64(int x, int y, int z) {
  uint64_t ux = (uint64_t)(x + 1000000);
  uint64_t uy = (uint64_t)(y + 1000000);
  uint64_t uz = (uint64_t)(z + 1000000);
  return (ux << 42) | (uy << 21) | uz;
}

// contact_detector.cpp:487-809 (spatial hash; AABB test skipped under PBC)
static void detect_hash(orc_world& W, double delta, double cellSize) {
  const int n = (int)W.B.size();
  W.contacts.clear();
  if (n == 0) return;
  if (cellSize <= 0.0) cellSize = adaptive_cell_size(W);
  const bool pbc = W.periodic[0] || W.periodic[1] || W.periodic[2];
  int dims[3] = {1, 1, 1};
  double cell[3] = {cellSize, cellSize, cellSize};
  if (pbc) {
    for (int k = 0; k < 3; ++k) {
      double span = W.hi[k] - W.lo[k];
      dims[k] = std::max(1, (int)std::floor(span / cellSize));
      cell[k] = span / dims[k];
    }
  }
  std::vector<V3> amin(n), amax(n);
  std::vector<int> cnt(n);
  std::vector<std::array<int, 6>> rng(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const Body& b = W.B[i];
    V3 u = axis_of(b.q);
    V3 p1 = sub(b.x, mul(u, b.h)), p2 = add(b.x, mul(u, b.h));
    double rr = b.r + delta;
    V3 mn = mk(std::min(p1.x, p2.x) - rr, std::min(p1.y, p2.y) - rr, std::min(p1.z, p2.z) - rr);
    V3 mx = mk(std::max(p1.x, p2.x) + rr, std::max(p1.y, p2.y) + rr, std::max(p1.z, p2.z) + rr);
    amin[i] = mn;
    amax[i] = mx;
    int r6[6];
    double mnv[3] = {mn.x, mn.y, mn.z}, mxv[3] = {mx.x, mx.y, mx.z};
    for (int k = 0; k < 3; ++k) {
      if (pbc) {
        r6[k] = (int)std::floor((mnv[k] - W.lo[k]) / cell[k]);
        r6[3 + k] = (int)std::floor((mxv[k] - W.lo[k]) / cell[k]);
        r6[3 + k] = std::min(r6[3 + k], r6[k] + dims[k] - 1);
      } else {
        r6[k] = (int)std::floor(mnv[k] / cellSize);
        r6[3 + k] = (int)std::floor(mxv[k] / cellSize);
      }
    }
    for (int k = 0; k < 6; ++k) rng[i][k] = r6[k];
    cnt[i] = (r6[3] - r6[0] + 1) * (r6[4] - r6[1] + 1) * (r6[5] - r6[2] + 1);
  }
  std::vector<size_t> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
  std::vector<std::pair<uint64_t, int>> ent(off[n]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    size_t o = off[i];
    const auto& r = rng[i];
    for (int x = r[0]; x <= r[3]; ++x)
      for (int y = r[1]; y <= r[4]; ++y)
        for (int z = r[2]; z <= r[5]; ++z) {
          int kx = x, ky = y, kz = z;
          if (pbc) {
            kx = x % dims[0];
            if (kx < 0) kx += dims[0];
            ky = y % dims[1];
            if (ky < 0) ky += dims[1];
            kz = z % dims[2];
            if (kz < 0) kz += dims[2];
          }
          ent[o++] = {hash_pos64(kx, ky, kz), i};
        }
  }
  std::sort(ent.begin(), ent.end());
  std::vector<std::pair<size_t, size_t>> cells;
  if (!ent.empty()) {
    size_t start = 0;
    for (size_t i = 1; i < ent.size(); ++i) {
      if (ent[i].first != ent[start].first) {
        if (i - start > 1) cells.push_back({start, i});
        start = i;
      }
    }
    if (ent.size() - start > 1) cells.push_back({start, ent.size()});
  }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  std::vector<std::vector<Contact>> tc(nth);
#pragma omp parallel for schedule(dynamic)
  for (long c = 0; c < (long)cells.size(); ++c) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    for (size_t i = cells[c].first; i < cells[c].second; ++i)
      for (size_t j = i + 1; j < cells[c].second; ++j) {
        int ia = ent[i].second, ib = ent[j].second;
        if (ia == ib) continue;
        if (ia > ib) std::swap(ia, ib);
        if (W.B[ia].invMass <= 0.0 && W.B[ib].invMass <= 0.0) continue;
        if (!pbc) {
          const V3 &mna = amin[ia], &mxa = amax[ia], &mnb = amin[ib], &mxb = amax[ib];
          if (mxa.x < mnb.x || mna.x > mxb.x || mxa.y < mnb.y || mna.y > mxb.y ||
              mxa.z < mnb.z || mna.z > mxb.z)
            continue;
        }
        Contact ct;
        if (capsule_capsule(W, ia, ib, delta, ct)) tc[tid].push_back(ct);
      }
  }
  for (auto& v : tc) W.contacts.insert(W.contacts.end(), v.begin(), v.end());
  sort_unique(W.contacts);
}

static void detect(orc_world& W, int method, double delta, double cs) {
  if (method == 1) detect_hash(W, delta, cs);
  else detect_naive(W, delta);
}

// nsc_contact.cpp:78-85
static void tangent_basis(V3 n, V3& t1, V3& t2) {
  V3 ref = (std::fabs(n.x) < 0.9) ? mk(1, 0, 0) : mk(0, 1, 0);
  t1 = normalize(cross(n, ref));
  t2 = cross(n, t1);
}

// nsc_contact.cpp:106-117 / 180-188
static double diag_term(double ima, double imb, const M3& IA, const M3& IB, V3 ra, V3 rb, V3 d,
                        double cfm, bool wall) {
  V3 rAxd = cross(ra, d);
  double g;
  if (wall) {
    g = ima + dot(rAxd, mv(IA, rAxd)) + cfm;
  } else {
    V3 rBxd = cross(rb, d);
    g = ima + imb + dot(rAxd, mv(IA, rAxd)) + dot(rBxd, mv(IB, rBxd)) + cfm;
  }
  return std::max(g, 1e-12);
}

// Largest eigenvalue of the 2x2 tangential block [[g_t1, W12], [W12, g_t2]],
// W12 = (r_a x t1).I_a^-1(r_a x t2) + (r_b x t1).I_b^-1(r_b x t2).
static double tangent_block_max(const Manifold& m, bool wall) {
  double w12 = dot(cross(m.ra, m.t1), m.IAt2);
  if (!wall) w12 = w12 + dot(cross(m.rb, m.t1), m.IBt2);
  const double hm = 0.5 * (m.gt1 + m.gt2), hd = 0.5 * (m.gt1 - m.gt2);
  return hm + std::sqrt(hd * hd + w12 * w12);
}

// nsc_contact.cpp:239-321 (rod-rod) followed by plane-wall manifolds
// (nsc_contact.cpp:190-235 with capsule-plane geometry after collision.cpp:155-195,
// one contact per penetrating endpoint).
static void nsc_build(orc_world& W, const orc_nsc_params& P) {
  W.mf.clear();
  W.mf.reserve(W.contacts.size());
  for (const Contact& cp : W.contacts) {
    Manifold m;
    m.a = cp.a;
    m.b = cp.b;
    m.n = cp.n;
    m.phi = cp.dist - cp.hsum;
    tangent_basis(m.n, m.t1, m.t2);
    const Body& A = W.B[cp.a];
    const Body& Bb = W.B[cp.b];
    m.ra = sub(cp.pa, A.x);
    m.rb = sub(cp.pb, add(Bb.x, cp.shift));
    M3 IA = iworld_inv(A), IB = iworld_inv(Bb);
    m.gn = diag_term(A.invMass, Bb.invMass, IA, IB, m.ra, m.rb, m.n, P.cfm, false);
    m.gt1 = diag_term(A.invMass, Bb.invMass, IA, IB, m.ra, m.rb, m.t1, P.cfm, false);
    m.gt2 = diag_term(A.invMass, Bb.invMass, IA, IB, m.ra, m.rb, m.t2, P.cfm, false);
    m.invMA = A.invMass;
    m.invMB = Bb.invMass;
    m.IAn = mv(IA, cross(m.ra, m.n));
    m.IAt1 = mv(IA, cross(m.ra, m.t1));
    m.IAt2 = mv(IA, cross(m.ra, m.t2));
    m.IBn = mv(IB, cross(m.rb, m.n));
    m.IBt1 = mv(IB, cross(m.rb, m.t1));
    m.IBt2 = mv(IB, cross(m.rb, m.t2));
    V3 vrel = sub(add(Bb.v, cross(Bb.w, m.rb)), add(A.v, cross(A.w, m.ra)));
    m.vnpre = dot(m.n, vrel);
    m.e = std::min(A.rest, Bb.rest);
    m.gtb = tangent_block_max(m, false);
    m.ln = m.lt1 = m.lt2 = 0.0;
    if (P.warm_start) {
      auto it = W.warm.find(pair_key(cp.a, cp.b));
      if (it != W.warm.end()) {
        m.ln = std::max(0.0, it->second[0]);
        m.lt1 = it->second[1];
        m.lt2 = it->second[2];
      }
    }
    W.mf.push_back(m);
  }
  // plane walls (floor at lo.z, ceiling at hi.z), body order, endpoints a1 then a2
  if (W.floor_wall || W.ceil_wall) {
    for (int i = 0; i < (int)W.B.size(); ++i) {
      const Body& A = W.B[i];
      if (A.invMass <= 0.0) continue;
      V3 u = axis_of(A.q);
      V3 ends[2] = {sub(A.x, mul(u, A.h)), add(A.x, mul(u, A.h))};
      for (int plane = 0; plane < 2; ++plane) {
        if (plane == 0 && !W.floor_wall) continue;
        if (plane == 1 && !W.ceil_wall) continue;
        for (int e = 0; e < 2; ++e) {
          V3 p = ends[e];
          double sd;
          V3 nrm, pt;
          if (plane == 0) {
            sd = (p.z - W.lo[2]) - A.r;
            nrm = mk(0, 0, -1);
            pt = sub(p, mk(0, 0, A.r));
          } else {
            sd = (W.hi[2] - p.z) - A.r;
            nrm = mk(0, 0, 1);
            pt = add(p, mk(0, 0, A.r));
          }
          if (!(sd < 0.0)) continue;
          Manifold m;
          m.a = i;
          m.b = -1 - (2 * plane + e);
          m.n = nrm;
          m.phi = sd;  // = -penetration
          tangent_basis(m.n, m.t1, m.t2);
          m.ra = sub(pt, A.x);
          m.rb = mk(0, 0, 0);
          M3 IA = iworld_inv(A), Z = diag(0, 0, 0);
          m.gn = diag_term(A.invMass, 0, IA, Z, m.ra, m.rb, m.n, P.cfm, true);
          m.gt1 = diag_term(A.invMass, 0, IA, Z, m.ra, m.rb, m.t1, P.cfm, true);
          m.gt2 = diag_term(A.invMass, 0, IA, Z, m.ra, m.rb, m.t2, P.cfm, true);
          m.invMA = A.invMass;
          m.invMB = 0.0;
          m.IAn = mv(IA, cross(m.ra, m.n));
          m.IAt1 = mv(IA, cross(m.ra, m.t1));
          m.IAt2 = mv(IA, cross(m.ra, m.t2));
          m.IBn = m.IBt1 = m.IBt2 = mk(0, 0, 0);
          V3 vrel = neg(add(A.v, cross(A.w, m.ra)));
          m.vnpre = dot(m.n, vrel);
          m.e = A.rest;
          m.gtb = tangent_block_max(m, true);
          m.ln = m.lt1 = m.lt2 = 0.0;
          if (P.warm_start && P.warm_walls) {
            auto it = W.warm_wall.find(((uint64_t)(uint32_t)i << 2) | (uint64_t)(2 * plane + e));
            if (it != W.warm_wall.end()) {
              m.ln = std::max(0.0, it->second[0]);
              m.lt1 = it->second[1];
              m.lt2 = it->second[2];
            }
          }
          W.mf.push_back(m);
        }
      }
    }
  }
}

// std::clamp semantics (nsc_contact.cpp:508,529)
static inline double clampd(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

// nsc_contact.cpp:344-380
static void apply_warm(orc_world& W) {
  for (const Manifold& m : W.mf) {
    Body& A = W.B[m.a];
    const bool wall = m.b < 0;
    const V3* dirs[3] = {&m.n, &m.t1, &m.t2};
    const V3* IAs[3] = {&m.IAn, &m.IAt1, &m.IAt2};
    const V3* IBs[3] = {&m.IBn, &m.IBt1, &m.IBt2};
    const double lam[3] = {m.ln, m.lt1, m.lt2};
    for (int d = 0; d < 3; ++d) {
      if (lam[d] != 0.0) {
        A.v = sub(A.v, mul(*dirs[d], lam[d] * m.invMA));
        A.w = sub(A.w, mul(*IAs[d], lam[d]));
        if (!wall) {
          Body& Bb = W.B[m.b];
          Bb.v = add(Bb.v, mul(*dirs[d], lam[d] * m.invMB));
          Bb.w = add(Bb.w, mul(*IBs[d], lam[d]));
        }
      }
    }
  }
}

static inline double beta_eff(const orc_nsc_params& P) {
  return P.position_stabilization ? P.beta * 0.01 : P.beta;
}

static inline V3 rel_vel(const orc_world& W, const Manifold& m) {
  const Body& A = W.B[m.a];
  if (m.b < 0) return neg(add(A.v, cross(A.w, m.ra)));
  const Body& Bb = W.B[m.b];
  return sub(add(Bb.v, cross(Bb.w, m.rb)), add(A.v, cross(A.w, m.ra)));
}

// Projection of y onto the disc |x| <= r in the metric of the SPD tangential
// block G = [[a, b], [b, c]] (|y| > r > 0): x = (G + nu I)^-1 G y with the nu > 0
// at which |x| = r.  In G's eigenbasis (e1 >= e2 > 0, z = y rotated) the
// residual h(nu) = 1/r - 1/|x(nu)| is convex and decreasing, and the root lies
// in [e2 (|z|/r - 1), e1 (|z|/r - 1)]: Newton from the left end increases nu
// monotonically to the root, and stops once a step no longer increases it
// (at most 16 steps); x is then scaled onto the circle.
static void disc_project_metric(double a, double b, double c, double r, double& y1, double& y2) {
  const double hm = 0.5 * (a + c), hd = 0.5 * (a - c);
  const double s = std::sqrt(hd * hd + b * b);
  const double e1 = hm + s, e2 = hm - s;
  double cs = 1.0, sn = 0.0;
  if (s > 0.0) {
    const double c2 = hd / s;
    cs = std::sqrt(0.5 * (1.0 + c2));
    sn = std::sqrt(0.5 * (1.0 - c2));
    if (b < 0.0) sn = -sn;
  }
  const double z1 = cs * y1 + sn * y2, z2 = cs * y2 - sn * y1;
  double nu = e2 * (std::sqrt(z1 * z1 + z2 * z2) / r - 1.0);
  if (!(nu > 0.0)) nu = 0.0;
  double x1, x2, nx;
  for (int it = 0; it < 16; ++it) {
    const double i1 = 1.0 / (e1 + nu), i2 = 1.0 / (e2 + nu);
    x1 = e1 * z1 * i1;
    x2 = e2 * z2 * i2;
    nx = std::sqrt(x1 * x1 + x2 * x2);
    const double d = x1 * x1 * i1 + x2 * x2 * i2;
    const double nn = nu + (nx - r) * (nx * nx) / (r * d);
    if (!(nn > nu)) break;
    nu = nn;
  }
  x1 = e1 * z1 / (e1 + nu);
  x2 = e2 * z2 / (e2 + nu);
  nx = std::sqrt(x1 * x1 + x2 * x2);
  if (nx > r) {
    x1 = x1 * (r / nx);
    x2 = x2 * (r / nx);
  }
  y1 = cs * x1 - sn * x2;
  y2 = sn * x1 + cs * x2;
}

// One PSOR update of one manifold (nsc_contact.cpp:385-552).  Returns
// |dn|+|dt1|+|dt2|.
static double psor_one(orc_world& W, Manifold& m, const orc_nsc_params& P) {
  const double omega = P.omega, cfm = P.cfm, mu = P.mu, beta = beta_eff(P), dt = W.dt;
  Body& A = W.B[m.a];
  const bool wall = m.b < 0;
  V3 v_rel = rel_vel(W, m);
  double b_n = (m.phi < 0.0) ? beta * m.phi / dt : 0.0;
  double b_rest = (m.vnpre < 0.0) ? m.e * m.vnpre : 0.0;
  double w_n = dot(m.n, v_rel) + b_n + b_rest + cfm * m.ln;
  double delta_n = -(omega / m.gn) * w_n;
  double old_n = m.ln;
  m.ln = std::max(0.0, old_n + delta_n);
  double dn = m.ln - old_n;
  if (dn != 0.0) {
    V3 dvA = mul(m.n, dn * m.invMA);
    A.v = sub(A.v, dvA);
    A.w = sub(A.w, mul(m.IAn, dn));
    v_rel = add(v_rel, add(dvA, cross(mul(m.IAn, dn), m.ra)));
    if (!wall) {
      Body& Bb = W.B[m.b];
      V3 dvB = mul(m.n, dn * m.invMB);
      Bb.v = add(Bb.v, dvB);
      Bb.w = add(Bb.w, mul(m.IBn, dn));
      v_rel = add(v_rel, add(dvB, cross(mul(m.IBn, dn), m.rb)));
    }
  }
  double max_fric = mu * m.ln;
  double dt1, dt2;
  if (P.friction_proj == 0) {
    double limit_t1 = std::sqrt(std::max(0.0, max_fric * max_fric - m.lt2 * m.lt2));
    double w_t1 = dot(m.t1, v_rel) + cfm * m.lt1;
    double delta_t1 = -(omega / m.gt1) * w_t1;
    double old_t1 = m.lt1;
    m.lt1 = clampd(old_t1 + delta_t1, -limit_t1, limit_t1);
    dt1 = m.lt1 - old_t1;
    if (dt1 != 0.0) {
      V3 dvA = mul(m.t1, dt1 * m.invMA);
      A.v = sub(A.v, dvA);
      A.w = sub(A.w, mul(m.IAt1, dt1));
      v_rel = add(v_rel, add(dvA, cross(mul(m.IAt1, dt1), m.ra)));
      if (!wall) {
        Body& Bb = W.B[m.b];
        V3 dvB = mul(m.t1, dt1 * m.invMB);
        Bb.v = add(Bb.v, dvB);
        Bb.w = add(Bb.w, mul(m.IBt1, dt1));
        v_rel = add(v_rel, add(dvB, cross(mul(m.IBt1, dt1), m.rb)));
      }
    }
    double limit_t2 = std::sqrt(std::max(0.0, max_fric * max_fric - m.lt1 * m.lt1));
    double w_t2 = dot(m.t2, v_rel) + cfm * m.lt2;
    double delta_t2 = -(omega / m.gt2) * w_t2;
    double old_t2 = m.lt2;
    m.lt2 = clampd(old_t2 + delta_t2, -limit_t2, limit_t2);
    dt2 = m.lt2 - old_t2;
  } else if (P.friction_proj >= 2) {
    // block step lambda_t - omega G^-1 w_t with the 2x2 tangential block G
    // (G11 = g_t1, G22 = g_t2, G12 from the cached I^-1 (r x t2)): a sticking
    // contact's slip is solved in both directions at once.  Outside the disc,
    // 2 projects the step in G's metric and 3 takes the scalar step of 1 from
    // the current impulse instead; both keep 1's fixed points (maximal
    // dissipation), and a sliding solution's block step always leaves the disc.
    double g12 = dot(cross(m.ra, m.t1), m.IAt2);
    if (!wall) g12 = g12 + dot(cross(m.rb, m.t1), m.IBt2);
    // G^-1 from the manifold's fields alone (off the chain of velocity updates)
    const double det = m.gt1 * m.gt2 - g12 * g12;
    const double h11 = m.gt2 / det, h12 = g12 / det, h22 = m.gt1 / det;
    double w_t1 = dot(m.t1, v_rel) + cfm * m.lt1;
    double w_t2 = dot(m.t2, v_rel) + cfm * m.lt2;
    const double s1 = h11 * w_t1 - h12 * w_t2, s2 = h22 * w_t2 - h12 * w_t1;
    double y1 = m.lt1 - omega * s1;
    double y2 = m.lt2 - omega * s2;
    if (y1 * y1 + y2 * y2 > max_fric * max_fric) {
      if (P.friction_proj == 3) {
        const double st = omega / m.gtb;
        y1 = m.lt1 - st * w_t1;
        y2 = m.lt2 - st * w_t2;
        double ny = std::sqrt(y1 * y1 + y2 * y2);
        if (ny > max_fric) {
          double sc = ny > 0.0 ? max_fric / ny : 0.0;
          y1 = y1 * sc;
          y2 = y2 * sc;
        }
      } else if (max_fric > 0.0) {
        disc_project_metric(m.gt1, g12, m.gt2, max_fric, y1, y2);
      } else {
        y1 = y2 = 0.0;
      }
    }
    dt1 = y1 - m.lt1;
    dt2 = y2 - m.lt2;
    m.lt1 = y1;
    m.lt2 = y2;
    if (dt1 != 0.0) {
      A.v = sub(A.v, mul(m.t1, dt1 * m.invMA));
      A.w = sub(A.w, mul(m.IAt1, dt1));
      if (!wall) {
        Body& Bb = W.B[m.b];
        Bb.v = add(Bb.v, mul(m.t1, dt1 * m.invMB));
        Bb.w = add(Bb.w, mul(m.IBt1, dt1));
      }
    }
  } else {
    // Euclidean disc: projected step of the joint tangential impulse with one
    // scalar step omega/g_tb (g_tb bounds the 2x2 tangential block), whose
    // fixed point is the maximal-dissipation condition w_t = -alpha lambda_t.
    double w_t1 = dot(m.t1, v_rel) + cfm * m.lt1;
    double w_t2 = dot(m.t2, v_rel) + cfm * m.lt2;
    const double st = omega / m.gtb;
    double y1 = m.lt1 - st * w_t1;
    double y2 = m.lt2 - st * w_t2;
    double ny = std::sqrt(y1 * y1 + y2 * y2);
    if (ny > max_fric) {
      double sc = ny > 0.0 ? max_fric / ny : 0.0;
      y1 = y1 * sc;
      y2 = y2 * sc;
    }
    dt1 = y1 - m.lt1;
    dt2 = y2 - m.lt2;
    m.lt1 = y1;
    m.lt2 = y2;
    if (dt1 != 0.0) {
      A.v = sub(A.v, mul(m.t1, dt1 * m.invMA));
      A.w = sub(A.w, mul(m.IAt1, dt1));
      if (!wall) {
        Body& Bb = W.B[m.b];
        Bb.v = add(Bb.v, mul(m.t1, dt1 * m.invMB));
        Bb.w = add(Bb.w, mul(m.IBt1, dt1));
      }
    }
  }
  if (dt2 != 0.0) {
    A.v = sub(A.v, mul(m.t2, dt2 * m.invMA));
    A.w = sub(A.w, mul(m.IAt2, dt2));
    if (!wall) {
      Body& Bb = W.B[m.b];
      Bb.v = add(Bb.v, mul(m.t2, dt2 * m.invMB));
      Bb.w = add(Bb.w, mul(m.IBt2, dt2));
    }
  }
  return std::fabs(dn) + std::fabs(dt1) + std::fabs(dt2);
}

// Build-defined fixed-point residual of the PSOR map, in velocity units:
// r = max_c max(g_n |ln - P_n|, g_t1 |lt1 - P_t1|, g_t2 |lt2 - P_t2|)
// where P is one unrelaxed projection step from the current (lambda, v).
static double residual_one(const orc_world& W, const Manifold& m, const orc_nsc_params& P) {
  const double cfm = P.cfm, mu = P.mu, beta = beta_eff(P), dt = W.dt;
  V3 v_rel = rel_vel(W, m);
  double b_n = (m.phi < 0.0) ? beta * m.phi / dt : 0.0;
  double b_rest = (m.vnpre < 0.0) ? m.e * m.vnpre : 0.0;
  double w_n = dot(m.n, v_rel) + b_n + b_rest + cfm * m.ln;
  double pn = std::max(0.0, m.ln - w_n / m.gn);
  double r = m.gn * std::fabs(m.ln - pn);
  double mf = mu * m.ln;
  double w_t1 = dot(m.t1, v_rel) + cfm * m.lt1;
  double w_t2 = dot(m.t2, v_rel) + cfm * m.lt2;
  double p1, p2;
  if (P.friction_proj == 0) {
    double L1 = std::sqrt(std::max(0.0, mf * mf - m.lt2 * m.lt2));
    double L2 = std::sqrt(std::max(0.0, mf * mf - m.lt1 * m.lt1));
    p1 = clampd(m.lt1 - w_t1 / m.gt1, -L1, L1);
    p2 = clampd(m.lt2 - w_t2 / m.gt2, -L2, L2);
    r = std::max(r, m.gt1 * std::fabs(m.lt1 - p1));
    r = std::max(r, m.gt2 * std::fabs(m.lt2 - p2));
  } else {
    p1 = m.lt1 - w_t1 / m.gtb;
    p2 = m.lt2 - w_t2 / m.gtb;
    double ny = std::sqrt(p1 * p1 + p2 * p2);
    if (ny > mf) {
      double sc = ny > 0.0 ? mf / ny : 0.0;
      p1 = p1 * sc;
      p2 = p2 * sc;
    }
    const double d1 = m.lt1 - p1, d2 = m.lt2 - p2;
    r = std::max(r, m.gtb * std::sqrt(d1 * d1 + d2 * d2));
  }
  return r;
}

static double residual(const orc_world& W, const orc_nsc_params& P) {
  double r = 0.0;
  for (const Manifold& m : W.mf) r = std::max(r, residual_one(W, m, P));
  return r;
}

static double sweep(orc_world& W, const orc_nsc_params& P, const int* order) {
  double it = 0.0;
  const int M = (int)W.mf.size();
  for (int k = 0; k < M; ++k) {
    Manifold& m = W.mf[order ? order[k] : k];
    it = std::max(it, psor_one(W, m, P));
  }
  return it;
}

static int solve(orc_world& W, const orc_nsc_params& P, const int* order, double* res_out) {
  int sweeps = 0;
  if (W.mf.empty()) {
    if (res_out) *res_out = 0.0;
    return 0;
  }
  if (P.stop_rule == 0) {
    double last = 0.0;
    for (int it = 0; it < P.max_sweeps; ++it) {
      last = sweep(W, P, order);
      ++sweeps;
      if (last < 1e-6) break;
    }
    if (res_out) *res_out = residual(W, P);
    return sweeps;
  }
  const int k = std::max(1, P.check_every);
  double r = residual(W, P);
  while (r > P.tol && sweeps < P.max_sweeps) {
    for (int i = 0; i < k; ++i) sweep(W, P, order);
    sweeps += k;
    r = residual(W, P);
  }
  if (res_out) *res_out = r;
  return sweeps;
}

// nsc_contact.cpp:646-654
// Wall entries: the reference keys them (i,i) and never matches them; the
// build keeps them separately, keyed by (rod, slot), for the optional
// warm_walls mode.
static void store_warm(orc_world& W) {
  W.warm.clear();
  W.warm_wall.clear();
  for (const Manifold& m : W.mf) {
    if (m.b < 0) {
      W.warm_wall[((uint64_t)(uint32_t)m.a << 2) | (uint64_t)(-1 - m.b)] = {m.ln, m.lt1, m.lt2};
      continue;
    }
    W.warm[pair_key(m.a, m.b)] = {m.ln, m.lt1, m.lt2};
  }
}

// integrator.cpp:23-35
static void wrap(orc_world& W, Body& b) {
  double* p[3] = {&b.x.x, &b.x.y, &b.x.z};
  for (int k = 0; k < 3; ++k) {
    if (!W.periodic[k]) continue;
    double span = W.hi[k] - W.lo[k];
    if (span <= 0.0) continue;
    int guard = 0;
    while (*p[k] < W.lo[k] && guard++ < 8) *p[k] += span;
    guard = 0;
    while (*p[k] >= W.hi[k] && guard++ < 8) *p[k] -= span;
  }
}

static V3 gyro_rhs(const Body& b) {
  // main.cpp:4404-4407 / integrator.cpp:39-54: I_w^-1 (tau - w x (I_w w))
  M3 Iinv = iworld_inv(b);
  M3 Iw = iworld(b);
  V3 gyro = cross(b.w, mv(Iw, b.w));
  return mv(Iinv, sub(b.tau, gyro));
}

// angleAxis(angle, axis) = (cos(a/2), axis*sin(a/2))
static Q4 angle_axis(double angle, V3 ax) {
  double s = std::sin(angle * 0.5);
  return Q4{std::cos(angle * 0.5), ax.x * s, ax.y * s, ax.z * s};
}

static void integrate(orc_world& W, int stage) {
  const int n = (int)W.B.size();
  const double dt = W.dt;
#pragma omp parallel for schedule(static) if (n > 1000)
  for (int i = 0; i < n; ++i) {
    Body& b = W.B[i];
    if (stage == 4) {
      b.f = b.tau = mk(0, 0, 0);
      continue;
    }
    if (b.invMass <= 0.0) {
      if (stage == 1 || stage == 3) b.f = b.tau = mk(0, 0, 0);
      continue;
    }
    if (stage == 0) {
      // main.cpp:4397-4408
      b.v = add(b.v, mul(add(mul(b.f, b.invMass), W.g), dt));
      b.w = add(b.w, mul(gyro_rhs(b), dt));
    } else if (stage == 1) {
      // main.cpp:4469-4483, wrap :4508-4517, clear :4524-4527
      b.x = add(b.x, mul(b.v, dt));
      Q4 wq{0.0, b.w.x, b.w.y, b.w.z};
      double hdt = 0.5 * dt;
      Q4 sw{wq.w * hdt, wq.x * hdt, wq.y * hdt, wq.z * hdt};
      Q4 dq = qmul(sw, b.q);
      b.q = Q4{b.q.w + dq.w, b.q.x + dq.x, b.q.y + dq.y, b.q.z + dq.z};
      b.q = qnormalize(b.q);
      b.v = mul(b.v, 1.0 - W.lin_damp * dt);
      b.w = mul(b.w, 1.0 - W.ang_damp * dt);
      double wl = length(b.w);
      if (wl > W.w_max) b.w = mul(b.w, W.w_max / wl);
      wrap(W, b);
      b.f = b.tau = mk(0, 0, 0);
    } else if (stage == 2) {
      // integrator.cpp:113-135
      V3 al = add(mul(b.f, 1.0 / b.mass), W.g);
      V3 aa = gyro_rhs(b);
      b.v = add(b.v, mul(al, 0.5 * dt));
      b.w = add(b.w, mul(aa, 0.5 * dt));
      b.x = add(b.x, mul(b.v, dt));
      wrap(W, b);
      double wl = length(b.w);
      if (wl > 1e-8) {
        double angle = wl * dt;
        V3 ax = mul(b.w, 1.0 / wl);
        b.q = qnormalize(qmul(angle_axis(angle, ax), b.q));
      }
    } else if (stage == 3) {
      // integrator.cpp:138-159
      V3 al = add(mul(b.f, 1.0 / b.mass), W.g);
      V3 aa = gyro_rhs(b);
      b.v = add(b.v, mul(al, 0.5 * dt));
      b.w = add(b.w, mul(aa, 0.5 * dt));
      if (W.lin_damp > 0.0) b.v = mul(b.v, std::exp(-W.lin_damp * dt));
      if (W.ang_damp > 0.0) b.w = mul(b.w, std::exp(-W.ang_damp * dt));
      double wl = length(b.w);
      if (wl > W.w_max) b.w = mul(b.w, W.w_max / wl);
      b.f = b.tau = mk(0, 0, 0);
    }
  }
}

// soft_contact.cpp:215-237
static double soft_grad(double d, double h, double delta, double K1) {
  if (d > h - delta) {
    const double arg = K1 * (h - d);
    const double e = std::exp(arg);
    const double lt = std::log(1.0 + e);
    return -2.0 / (K1 * K1) * lt * e / (1.0 + e);
  }
  return -2.0 * (h - d);
}
static double soft_pe(double d, double h, double delta, double K1) {
  if (d > h - delta) {
    const double arg = K1 * (h - d);
    const double lt = std::log(1.0 + std::exp(arg));
    double base = lt / K1;
    return base * base;
  }
  double pen = h - d;
  return pen * pen;
}

// Accumulate per-contact body forces in sorted contact order: f[a] then f[b].
static void scatter(orc_world& W, const Contact& c, V3 Fa, V3 Fb, V3 Ta, V3 Tb) {
  Body& A = W.B[c.a];
  Body& Bb = W.B[c.b];
  A.f = add(A.f, Fa);
  Bb.f = add(Bb.f, Fb);
  A.tau = add(A.tau, Ta);
  Bb.tau = add(Bb.tau, Tb);
}

// soft_contact.cpp:80-213, 255-486
static double soft_forces(orc_world& W, const orc_soft_params& P) {
  const double K1 = 15.0 / P.delta, K2 = 15.0 / P.nu;
  double pe = 0.0;
  for (const Contact& c : W.contacts) {
    const Body& A = W.B[c.a];
    const Body& Bb = W.B[c.b];
    const double grad = soft_grad(c.dist, c.hsum, P.delta, K1);
    const double fm = -P.k_scaler * grad;
    V3 fa = mul(c.n, -fm), fb = mul(c.n, fm);
    const V3 ra = sub(c.pa, A.x);
    const V3 rb = sub(c.pb, add(Bb.x, c.shift));
    if (P.damping > 1e-9) {
      V3 va = add(A.v, cross(A.w, ra)), vb = add(Bb.v, cross(Bb.w, rb));
      V3 vr = sub(va, vb);
      double vn = dot(vr, c.n);
      V3 fd = mul(c.n, -P.damping * vn);
      fa = add(fa, fd);
      fb = sub(fb, fd);
    }
    V3 fra{0, 0, 0}, frb{0, 0, 0};
    if (P.enable_friction) {
      V3 va = add(A.v, cross(A.w, ra)), vb = add(Bb.v, cross(Bb.w, rb));
      V3 vr = sub(va, vb);
      double fn = length(fa);
      V3 vt = sub(vr, mul(c.n, dot(vr, c.n)));
      double vtm = length(vt);
      if (vtm >= 1e-10) {
        if (P.friction_model == 1) {
          V3& ft = W.cs_hist[pair_key(c.a, c.b)];
          ft = sub(ft, mul(vt, P.kt * W.dt));
          double nc = dot(ft, c.n);
          ft = sub(ft, mul(c.n, nc));
          double ftm = length(ft);
          double lim = P.mu * fn;
          if (ftm > lim) {
            if (ftm > 1e-8) ft = mul(ft, lim / ftm);
            else ft = mk(0, 0, 0);
          }
          fra = ft;
          frb = neg(ft);
        } else if (P.friction_model == 2) {
          if (vtm < P.vel_deadband) {
            double ma = A.invMass > 0 ? A.mass : 0.0, mb = Bb.invMass > 0 ? Bb.mass : 0.0;
            double me = 0.0;
            if (ma > 0 && mb > 0) me = (ma * mb) / (ma + mb);
            else if (ma > 0) me = ma;
            else if (mb > 0) me = mb;
            V3 Fstop = mul(vt, -(me / W.dt));
            V3 gt = sub(W.g, mul(c.n, dot(W.g, c.n)));
            V3 Fdrive = mul(gt, me);
            V3 Freq = add(neg(Fdrive), Fstop);
            double Freqm = length(Freq);
            double maxs = P.mu_static * fn;
            if (Freqm <= maxs) {
              fra = Freq;
              frb = neg(Freq);
            } else if (Freqm > 1e-12) {
              fra = mul(Freq, maxs / Freqm);
              frb = neg(fra);
            }
          } else {
            V3 dir = mul(vt, -1.0 / vtm);
            fra = mul(dir, P.mu * fn);
            frb = neg(fra);
          }
        } else {
          double gamma = (vtm > P.nu) ? 1.0 : 2.0 / (1.0 + std::exp(-K2 * vtm)) - 1.0;
          double mu_eff = P.mu;
          if (P.mu_static > P.mu) mu_eff = P.mu + (P.mu_static - P.mu) * std::exp(-vtm / P.nu);
          V3 dir = mul(vt, -1.0 / vtm);
          double fmag = (mu_eff * gamma) * fn;
          fra = mul(dir, fmag);
          frb = neg(fra);
        }
      }
    }
    V3 Ta = add(fa, fra), Tb = add(fb, frb);
    scatter(W, c, Ta, Tb, cross(ra, Ta), cross(rb, Tb));
    pe += P.k_scaler * soft_pe(c.dist, c.hsum, P.delta, K1);
  }
  return pe;
}

// hertz_mindlin.cpp:183-393 with capsule geometry (overlap = r_a+r_b-dist),
// torques r x F on both bodies (soft_contact.cpp:150-157 convention).
static double hm_forces(orc_world& W, const orc_hm_params& P) {
  double gn = 0.0;
  if (P.restitution < 1.0) {
    double le = std::log(P.restitution);
    gn = -le / std::sqrt(M_PI * M_PI + le * le);
  }
  const double gt = gn;
  const double Es = P.E / (2.0 * (1.0 - P.nu_poisson * P.nu_poisson));
  const double Gs = P.E / (2.0 * (1.0 + P.nu_poisson));
  double pe = 0.0;
  W.hm_frame++;
  for (const Contact& c : W.contacts) {
    const Body& A = W.B[c.a];
    const Body& Bb = W.B[c.b];
    HistEntry& st = W.hm_hist[pair_key(c.a, c.b)];
    const double delta = c.hsum - c.dist;
    const double Rs = (A.r * Bb.r) / (A.r + Bb.r);
    const double ms = (A.mass * Bb.mass) / (A.mass + Bb.mass);
    const V3 ra = sub(c.pa, A.x);
    const V3 rb = sub(c.pb, add(Bb.x, c.shift));
    // Hertz normal (hertz_mindlin.cpp:249-287)
    double sd = std::max(delta, 0.0);
    double sqd = std::sqrt(sd);
    double sqR = std::sqrt(std::max(Rs, 0.0));
    double kn = (4.0 / 3.0) * Es * sqR * sqd;
    double Fel = (P.law == 0) ? kn * sd * sqd : kn * sd;
    V3 va = add(A.v, cross(A.w, ra)), vb = add(Bb.v, cross(Bb.w, rb));
    V3 vr = sub(vb, va);
    double vn = dot(vr, c.n);
    double Cn = gn * 2.0 * std::sqrt(ms * kn);
    double Fn = std::max(0.0, Fel + (-Cn * vn));
    V3 Fnv = mul(c.n, Fn);
    V3 Ft{0, 0, 0};
    if (P.enable_tangential) {
      double kt = 8.0 * Gs * sqR * sqd;
      V3 vt = sub(vr, mul(c.n, vn));
      st.xi = add(st.xi, mul(vt, W.dt));
      V3 Fte = mul(st.xi, -kt);
      double Ct = gt * 2.0 * std::sqrt(ms * kt);
      V3 Ftd = mul(vt, -Ct);
      Ft = add(Fte, Ftd);
      double vtm = length(vt);
      double mue = P.mu_k;
      if (P.mu_s > P.mu_k && P.v_c > 0.0) mue += (P.mu_s - P.mu_k) * std::exp(-vtm / P.v_c);
      double Fnm = length(Fnv);
      double Ftm = length(Ft);
      double Ftmax = mue * Fnm;
      if (Ftm > Ftmax && Ftm > 1e-10) {
        Ft = mul(Ft, Ftmax / Ftm);
        if (kt > 1e-10) st.xi = mul(Ft, -1.0 / kt);
      }
    }
    V3 ta{0, 0, 0}, tb{0, 0, 0};
    if (P.enable_rolling) {
      V3 wr = sub(A.w, Bb.w);
      double wm = length(wr);
      if (wm >= 1e-10) {
        V3 wh = mul(wr, 1.0 / wm);
        double trm = P.mu_roll * Rs * length(Fnv);
        V3 tr = mul(wh, -trm);
        ta = mul(tr, 0.5);
        tb = mul(neg(tr), 0.5);
      }
    }
    V3 F = add(Fnv, Ft);
    scatter(W, c, neg(F), F, add(cross(ra, neg(F)), ta), add(cross(rb, F), tb));
    pe += (8.0 / 15.0) * Es * std::sqrt(Rs) * std::pow(delta, 2.5);
    st.last_frame = W.hm_frame;
  }
  // hertz_mindlin.cpp:385-393
  for (auto it = W.hm_hist.begin(); it != W.hm_hist.end();) {
    if (W.hm_frame - it->second.last_frame > 4 || it->second.last_frame == 0) it = W.hm_hist.erase(it);
    else ++it;
  }
  return pe;
}

}  // namespace

extern "C" {

orc_world* orc_create(int n, const double* x, const double* q, const double* v, const double* w,
                      const double* r, const double* h, const double* mass, const double* rest) {
  orc_world* W = new orc_world();
  W->B.resize(n);
  for (int i = 0; i < n; ++i) {
    Body& b = W->B[i];
    b.r = r[i];
    b.h = h[i];
    b.rest = rest ? rest[i] : 0.0;
    set_mass(b, mass[i]);
    b.f = b.tau = mk(0, 0, 0);
  }
  orc_set_state(W, x, q, v, w);
  return W;
}

void orc_destroy(orc_world* W) { delete W; }

void orc_set_box(orc_world* W, const double* lo, const double* hi, const int* periodic, int fw, int cw) {
  for (int k = 0; k < 3; ++k) {
    W->lo[k] = lo[k];
    W->hi[k] = hi[k];
    W->periodic[k] = periodic[k];
  }
  W->floor_wall = fw;
  W->ceil_wall = cw;
}

void orc_set_physics(orc_world* W, const double* g, double dt, double ld, double ad, double wmax) {
  W->g = mk(g[0], g[1], g[2]);
  W->dt = dt;
  W->lin_damp = ld;
  W->ang_damp = ad;
  W->w_max = wmax;
}

void orc_set_state(orc_world* W, const double* x, const double* q, const double* v, const double* w) {
  for (size_t i = 0; i < W->B.size(); ++i) {
    Body& b = W->B[i];
    if (x) b.x = mk(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
    if (q) b.q = Q4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    if (v) b.v = mk(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    if (w) b.w = mk(w[3 * i], w[3 * i + 1], w[3 * i + 2]);
  }
}

void orc_set_forces(orc_world* W, const double* f, const double* tau) {
  for (size_t i = 0; i < W->B.size(); ++i) {
    if (f) W->B[i].f = mk(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
    if (tau) W->B[i].tau = mk(tau[3 * i], tau[3 * i + 1], tau[3 * i + 2]);
  }
}

void orc_get_state(orc_world* W, double* x, double* q, double* v, double* w, double* f, double* tau) {
  for (size_t i = 0; i < W->B.size(); ++i) {
    const Body& b = W->B[i];
    const V3* src[5] = {&b.x, &b.v, &b.w, &b.f, &b.tau};
    double* dst[5] = {x, v, w, f, tau};
    for (int k = 0; k < 5; ++k)
      if (dst[k]) {
        dst[k][3 * i] = src[k]->x;
        dst[k][3 * i + 1] = src[k]->y;
        dst[k][3 * i + 2] = src[k]->z;
      }
    if (q) {
      q[4 * i] = b.q.w;
      q[4 * i + 1] = b.q.x;
      q[4 * i + 2] = b.q.y;
      q[4 * i + 3] = b.q.z;
    }
  }
}

void orc_axis(const double* q, double* u) {
  V3 a = axis_of(Q4{q[0], q[1], q[2], q[3]});
  u[0] = a.x;
  u[1] = a.y;
  u[2] = a.z;
}

void orc_segseg(const double* a1, const double* a2, const double* b1, const double* b2, double* s, double* t) {
  segseg(mk(a1[0], a1[1], a1[2]), mk(a2[0], a2[1], a2[2]), mk(b1[0], b1[1], b1[2]),
         mk(b2[0], b2[1], b2[2]), *s, *t);
}

int orc_detect(orc_world* W, int method, double delta, double cs) {
  detect(*W, method, delta, cs);
  return (int)W->contacts.size();
}

int orc_get_contacts(orc_world* W, int cap, int* a, int* b, double* pa, double* pb, double* nrm,
                     double* dist, double* shift, double* s, double* t, int* type) {
  int n = std::min(cap, (int)W->contacts.size());
  for (int i = 0; i < n; ++i) {
    const Contact& c = W->contacts[i];
    if (a) a[i] = c.a;
    if (b) b[i] = c.b;
    const V3* vs[4] = {&c.pa, &c.pb, &c.n, &c.shift};
    double* ds[4] = {pa, pb, nrm, shift};
    for (int k = 0; k < 4; ++k)
      if (ds[k]) {
        ds[k][3 * i] = vs[k]->x;
        ds[k][3 * i + 1] = vs[k]->y;
        ds[k][3 * i + 2] = vs[k]->z;
      }
    if (dist) dist[i] = c.dist;
    if (s) s[i] = c.s;
    if (t) t[i] = c.t;
    if (type) type[i] = c.type;
  }
  return (int)W->contacts.size();
}

int orc_nsc_build(orc_world* W, const orc_nsc_params* P) {
  nsc_build(*W, *P);
  return (int)W->mf.size();
}

int orc_nsc_num_manifolds(orc_world* W) { return (int)W->mf.size(); }

static void mf_to_rec(const Manifold& m, double* r) {
  const V3* vs[11] = {&m.n, &m.t1, &m.t2, &m.ra, &m.rb, &m.IAn, &m.IAt1, &m.IAt2, &m.IBn, &m.IBt1, &m.IBt2};
  for (int k = 0; k < 11; ++k) {
    r[3 * k] = vs[k]->x;
    r[3 * k + 1] = vs[k]->y;
    r[3 * k + 2] = vs[k]->z;
  }
  r[33] = m.invMA;
  r[34] = m.invMB;
  r[35] = m.gn;
  r[36] = m.gt1;
  r[37] = m.gt2;
  r[38] = m.phi;
  r[39] = m.vnpre;
  r[40] = m.e;
  r[41] = m.gtb;
}
static void rec_to_mf(const double* r, Manifold& m) {
  V3* vs[11] = {&m.n, &m.t1, &m.t2, &m.ra, &m.rb, &m.IAn, &m.IAt1, &m.IAt2, &m.IBn, &m.IBt1, &m.IBt2};
  for (int k = 0; k < 11; ++k) *vs[k] = mk(r[3 * k], r[3 * k + 1], r[3 * k + 2]);
  m.invMA = r[33];
  m.invMB = r[34];
  m.gn = r[35];
  m.gt1 = r[36];
  m.gt2 = r[37];
  m.phi = r[38];
  m.vnpre = r[39];
  m.e = r[40];
  m.gtb = r[41];
}

int orc_nsc_get_manifolds(orc_world* W, int cap, double* rec, int* ab, double* lambda) {
  int n = std::min(cap, (int)W->mf.size());
  for (int i = 0; i < n; ++i) {
    const Manifold& m = W->mf[i];
    if (rec) mf_to_rec(m, rec + (size_t)ORC_MF_FIELDS * i);
    if (ab) {
      ab[2 * i] = m.a;
      ab[2 * i + 1] = m.b;
    }
    if (lambda) {
      lambda[3 * i] = m.ln;
      lambda[3 * i + 1] = m.lt1;
      lambda[3 * i + 2] = m.lt2;
    }
  }
  return (int)W->mf.size();
}

int orc_nsc_set_manifolds(orc_world* W, int M, const double* rec, const int* ab, const double* lambda) {
  W->mf.resize(M);
  for (int i = 0; i < M; ++i) {
    Manifold& m = W->mf[i];
    rec_to_mf(rec + (size_t)ORC_MF_FIELDS * i, m);
    m.a = ab[2 * i];
    m.b = ab[2 * i + 1];
    m.ln = lambda ? lambda[3 * i] : 0.0;
    m.lt1 = lambda ? lambda[3 * i + 1] : 0.0;
    m.lt2 = lambda ? lambda[3 * i + 2] : 0.0;
  }
  return M;
}

// Island solve with the stall rule and the hot-set phase, in the order the
// library's island kernels run them (rdk_nsc.hip: StallTrack, k_island_lds).
// Hot set: once at least 64 sweeps have run and 0 < (manifolds above tol) <=
// max(16, m / 50), sweep only the manifolds sharing a body with one whose
// residual exceeds tol / 2 (in solve order; skipped when more than 512, the
// next attempt then waiting until the sweep count has doubled), in blocks of
// check_every sweeps until their residual is <= tol / 2, then check the
// island again.  hot == 2 adds selective rounds: when no hot set is taken and
// at least 32 sweeps have run, one block of check_every sweeps over that set
// (any size up to half the island), then the island check.
int orc_nsc_solve_island(orc_world* W, const orc_nsc_params* P, const int* ix, int m, int stall_sweeps,
                         double stall_omega, int hot, double* res_out, long long* updates_out,
                         int* hot_sweeps_out) {
  orc_nsc_params Q = *P;
  const double hot_thr = 0.5 * P->tol;
  const int kk = std::max(1, P->check_every);
  std::unordered_set<int> flagged;
  int above = 0;
  auto check = [&]() {
    double r = 0.0;
    above = 0;
    flagged.clear();
    for (int k = 0; k < m; ++k) {
      const Manifold& mf = W->mf[ix[k]];
      const double ri = residual_one(*W, mf, Q);
      r = std::max(r, ri);
      if (hot) {
        if (ri > P->tol) ++above;
        if (ri > hot_thr) {
          flagged.insert(mf.a);
          if (mf.b >= 0) flagged.insert(mf.b);
        }
      }
    }
    return r;
  };
  int sweeps = 0, hot_next = 64, hot_sw = 0;
  long long updates = 0;
  int next = stall_sweeps;  // stall milestones stall, 2 stall, 4 stall, ...
  double ref = -1.0;
  double res = check();
  std::vector<int> H;
  while (sweeps < P->max_sweeps && !(res <= P->tol)) {
    H.clear();
    bool sel = false;
    if (hot && sweeps >= hot_next && above > 0 && above <= std::max(16, m / 50)) {
      for (int k = 0; k < m; ++k) {
        const Manifold& mf = W->mf[ix[k]];
        if (flagged.count(mf.a) || (mf.b >= 0 && flagged.count(mf.b))) H.push_back(ix[k]);
      }
      if (H.size() > 512 || H.empty()) {
        H.clear();
        hot_next = 2 * sweeps;
      }
    }
    if (hot == 2 && H.empty() && sweeps >= 32) {
      // selective round: the same set, any size up to half the island, one
      // block of check_every sweeps, then the whole island is checked again
      for (int k = 0; k < m; ++k) {
        const Manifold& mf = W->mf[ix[k]];
        if (flagged.count(mf.a) || (mf.b >= 0 && flagged.count(mf.b))) H.push_back(ix[k]);
      }
      if (H.empty() || 2 * (int)H.size() > m) H.clear();
      else sel = true;
    }
    if (sel) {
      const int nh = (int)H.size();
      for (int it = 0; it < kk; ++it)
        for (int k = 0; k < nh; ++k) psor_one(*W, W->mf[H[k]], Q);
      sweeps += kk;
      updates += (long long)kk * nh;
    } else if (!H.empty()) {
      const int nh = (int)H.size();
      int ls = 0;
      while (true) {
        for (int it = 0; it < kk; ++it)
          for (int k = 0; k < nh; ++k) psor_one(*W, W->mf[H[k]], Q);
        ls += kk;
        double rh = 0.0;
        for (int k = 0; k < nh; ++k) rh = std::max(rh, residual_one(*W, W->mf[H[k]], Q));
        if (rh <= hot_thr || sweeps + ls >= P->max_sweeps) break;
      }
      sweeps += ls;
      updates += (long long)ls * nh;
      hot_sw += ls;
    } else {
      for (int it = 0; it < kk; ++it)
        for (int k = 0; k < m; ++k) psor_one(*W, W->mf[ix[k]], Q);
      sweeps += kk;
      updates += (long long)kk * m;
    }
    res = check();
    if (!(res <= P->tol) && next > 0 && sweeps >= next) {
      const bool st = ref >= 0.0 && !(res <= 0.9 * ref);
      ref = res;
      while (next <= sweeps) next *= 2;
      if (st) {
        if (stall_omega > 0.0 && Q.omega > stall_omega) Q.omega = stall_omega;
        else break;
      }
    }
  }
  if (res_out) *res_out = res;
  if (updates_out) *updates_out = updates;
  if (hot_sweeps_out) *hot_sweeps_out = hot_sw;
  return sweeps;
}

void orc_nsc_apply_warm(orc_world* W) { apply_warm(*W); }

int orc_nsc_sweeps(orc_world* W, const orc_nsc_params* P, const int* order, int nsweeps) {
  for (int i = 0; i < nsweeps; ++i) sweep(*W, *P, order);
  return nsweeps;
}

int orc_nsc_solve(orc_world* W, const orc_nsc_params* P, const int* order, double* res, int* sweeps) {
  int s = solve(*W, *P, order, res);
  if (sweeps) *sweeps = s;
  return 0;
}

double orc_nsc_residual(orc_world* W, const orc_nsc_params* P) { return residual(*W, *P); }

void orc_disc_project_metric(double a, double b, double c, double r, double* y) {
  disc_project_metric(a, b, c, r, y[0], y[1]);
}

void orc_nsc_store_warm(orc_world* W) { store_warm(*W); }

// main.cpp:4383-4527 (no random forces, no sleep, no cylinder; plane walls)
int orc_nsc_step(orc_world* W, const orc_nsc_params* P, int nsteps, int method, double cs,
                 double* sweeps_total, double* residual_max) {
  double st = 0.0, rmax = 0.0;
  for (int s = 0; s < nsteps; ++s) {
    integrate(*W, 0);
    detect(*W, method, 0.0, cs);
    nsc_build(*W, *P);
    if (P->warm_start) apply_warm(*W);
    double r = 0.0;
    int sw = solve(*W, *P, nullptr, &r);
    st += sw;
    rmax = std::max(rmax, r);
    if (P->warm_start) store_warm(*W);
    else W->warm.clear();
    integrate(*W, 1);
    if (P->position_stabilization && P->position_iters > 0) {
      detect(*W, method, 0.0, cs);
      orc_nsc_project(W, P, nullptr, nullptr);
    }
  }
  if (sweeps_total) *sweeps_total = st;
  if (residual_max) *residual_max = rmax;
  return 0;
}

// ---- position stabilisation (nsc_contact.cpp:659-880, multi-body branch) ----
namespace {
struct PosManifold {
  int a, b;
  V3 n, ra, rb, IAn, IBn;
  double phi, gn, lam;
};

// narrow phase of a cached pair at the current positions (nsc_contact.cpp:772-815)
static bool pos_manifold(const orc_world& W, int ia, int ib, double hs, double slop, double cfm, PosManifold& p) {
  Contact c;
  if (!capsule_capsule(W, ia, ib, std::numeric_limits<double>::infinity(), c)) return false;
  const double phi = c.dist - hs;
  if (!(phi < -slop)) return false;
  const Body& A = W.B[ia];
  const Body& Bd = W.B[ib];
  p.a = ia;
  p.b = ib;
  p.n = c.n;
  p.phi = phi + slop;
  p.ra = sub(c.pa, A.x);
  p.rb = sub(c.pb, add(Bd.x, c.shift));
  const M3 IA = iworld_inv(A), IB = iworld_inv(Bd);
  p.gn = diag_term(A.invMass, Bd.invMass, IA, IB, p.ra, p.rb, p.n, cfm, false);
  p.IAn = mv(IA, cross(p.ra, p.n));
  p.IBn = mv(IB, cross(p.rb, p.n));
  p.lam = 0.0;
  return true;
}
}  // namespace

int orc_nsc_project(orc_world* Wp, const orc_nsc_params* P, const int* order, int* npos) {
  orc_world& W = *Wp;
  const int n = (int)W.B.size();
  // cached pairs: every contact of the detection, sorted (a,b)
  std::vector<std::array<int, 2>> pairs;
  std::vector<double> hsum;
  for (const Contact& c : W.contacts) {
    pairs.push_back({c.a, c.b});
    hsum.push_back(c.hsum);
  }
  std::vector<V3> dx(n), dth(n);
  std::vector<PosManifold> pm;
  int passes = 0;
  if (npos) *npos = 0;
  for (int outer = 0; outer < P->position_iters; ++outer) {
    pm.clear();
    for (size_t k = 0; k < pairs.size(); ++k) {
      PosManifold p;
      if (pos_manifold(W, pairs[k][0], pairs[k][1], hsum[k], P->slop, P->cfm, p)) pm.push_back(p);
    }
    if (outer == 0 && npos) *npos = (int)pm.size();
    if (pm.empty()) break;
    ++passes;
    std::fill(dx.begin(), dx.end(), mk(0, 0, 0));
    std::fill(dth.begin(), dth.end(), mk(0, 0, 0));
    const int m = (int)pm.size();
    for (int inner = 0; inner < P->position_psor_iters; ++inner) {
      double pres = 0.0;
      for (int j = 0; j < m; ++j) {
        PosManifold& p = pm[(outer == 0 && order) ? order[j] : j];
        const Body& A = W.B[p.a];
        const Body& Bd = W.B[p.b];
        const double corr =
            dot(p.n, sub(sub(add(dx[p.b], cross(dth[p.b], p.rb)), dx[p.a]), cross(dth[p.a], p.ra)));
        const double res = corr + p.phi + P->cfm * p.lam;
        const double delta = -(P->omega / p.gn) * res;
        const double old = p.lam;
        p.lam = std::max(0.0, old + delta);
        const double dl = p.lam - old;
        if (dl != 0.0) {
          dx[p.a] = sub(dx[p.a], mul(p.n, dl * A.invMass));
          dth[p.a] = sub(dth[p.a], mul(p.IAn, dl));
          dx[p.b] = add(dx[p.b], mul(p.n, dl * Bd.invMass));
          dth[p.b] = add(dth[p.b], mul(p.IBn, dl));
        }
        pres = std::max(pres, std::fabs(dl));
      }
      if (pres < 1e-8) break;
    }
    for (int i = 0; i < n; ++i) {
      Body& b = W.B[i];
      if (b.invMass <= 0.0) continue;
      b.x = add(b.x, dx[i]);
      const double angle = length(dth[i]);
      if (angle > 1e-8) {
        const V3 ax = mk(dth[i].x / angle, dth[i].y / angle, dth[i].z / angle);
        b.q = qnormalize(qmul(angle_axis(angle, ax), b.q));
      }
      wrap(W, b);
    }
  }
  return passes;
}

void orc_integrate(orc_world* W, int stage) { integrate(*W, stage); }

int orc_soft_forces(orc_world* W, const orc_soft_params* P, double* pe) {
  double e = soft_forces(*W, *P);
  if (pe) *pe = e;
  return 0;
}

int orc_hm_forces(orc_world* W, const orc_hm_params* P, double* pe) {
  double e = hm_forces(*W, *P);
  if (pe) *pe = e;
  return 0;
}

int orc_hm_history_size(orc_world* W) { return (int)W->hm_hist.size(); }

// main.cpp:4529-4810: forces at t, half kick + drift, clear, forces at t+dt,
// second half kick.
int orc_vv_step(orc_world* W, int model, const void* params, int nsteps, int method, double cs) {
  for (int s = 0; s < nsteps; ++s) {
    for (int pass = 0; pass < 2; ++pass) {
      if (model == 0) {
        const orc_soft_params* P = (const orc_soft_params*)params;
        detect(*W, method, P->delta, cs);
        soft_forces(*W, *P);
      } else {
        const orc_hm_params* P = (const orc_hm_params*)params;
        detect(*W, method, 0.0, cs);
        hm_forces(*W, *P);
      }
      if (pass == 0) {
        integrate(*W, 2);
        integrate(*W, 4);
      } else {
        integrate(*W, 3);
      }
    }
  }
  return 0;
}

// main.cpp:3357-3368
double orc_kinetic_energy(orc_world* W) {
  double ke = 0.0;
  for (const Body& b : W->B) {
    double v2 = dot(b.v, b.v);
    ke += 0.5 * b.mass * v2;
    V3 Iw_w = mv(iworld(b), b.w);
    ke += 0.5 * dot(b.w, Iw_w);
  }
  return ke;
}

// ---- entanglement observer (main.cpp:4287-4331, entanglement_adapter.cpp:56-117)
// Gauss linking number of two straight segments a0a1, b0b1: the signed solid
// angle of the tetrahedron they span over 4 pi (Klenin & Langowski 2000):
// Omega = asin(n1.n2) + asin(n2.n3) + asin(n3.n4) + asin(n4.n1) with the face
// normals n1 = r13 x r14, n2 = r14 x r24, n3 = r24 x r23, n4 = r23 x r13
// (r_ij = p_j - p_i, p1..p4 = a0, a1, b0, b1), each normalised as
// n / (|n| + 1e-6), and the sign of (r34 x r12) . r13.  The 1e-6 is the
// convention of the reference's linking-number routine (its dependency
// entanglement-cpp is not vendored): with it this formula reproduces the
// published ebar0 of all 60 reference packings
// (assets/packings_metadata.csv) to 1e-6.
static double linking_number(V3 p1, V3 p2, V3 p3, V3 p4) {
  const V3 r13 = sub(p3, p1), r14 = sub(p4, p1), r23 = sub(p3, p2), r24 = sub(p4, p2);
  const V3 r12 = sub(p2, p1), r34 = sub(p4, p3);
  auto unit = [](V3 v) {
    const double n = std::sqrt(dot(v, v)) + 1e-6;
    return mk(v.x / n, v.y / n, v.z / n);
  };
  const V3 n1 = unit(cross(r13, r14)), n2 = unit(cross(r14, r24));
  const V3 n3 = unit(cross(r24, r23)), n4 = unit(cross(r23, r13));
  auto as = [](double x) { return std::asin(x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x)); };
  const double om = as(dot(n1, n2)) + as(dot(n2, n3)) + as(dot(n3, n4)) + as(dot(n4, n1));
  const double tp = dot(cross(r34, r12), r13);
  const double sg = tp > 0.0 ? 1.0 : (tp < 0.0 ? -1.0 : 0.0);
  return om * sg / (4.0 * M_PI);
}

// Rod axis endpoints x -/+ u h; on a periodic world each endpoint is wrapped
// into the box on its own, as the reference does before the pair sum
// (main.cpp:4294-4300).
static void rod_segment(const orc_world& W, const Body& b, V3& a, V3& e) {
  const V3 u = axis_of(b.q);
  a = sub(b.x, mul(u, b.h));
  e = add(b.x, mul(u, b.h));
  const bool pbc = W.periodic[0] || W.periodic[1] || W.periodic[2];
  if (!pbc) return;
  double* pa[3] = {&a.x, &a.y, &a.z};
  double* pe[3] = {&e.x, &e.y, &e.z};
  for (int k = 0; k < 3; ++k) {
    if (!W.periodic[k]) continue;
    const double span = W.hi[k] - W.lo[k];
    *pa[k] -= span * std::floor((*pa[k] - W.lo[k]) / span);
    *pe[k] -= span * std::floor((*pe[k] - W.lo[k]) / span);
  }
}

double orc_linking_number(const double* a0, const double* a1, const double* b0, const double* b1) {
  return linking_number(mk(a0[0], a0[1], a0[2]), mk(a1[0], a1[1], a1[2]), mk(b0[0], b0[1], b0[2]),
                        mk(b1[0], b1[1], b1[2]));
}

// Sum of |Lk| over the rod pairs (all of them when cutoff <= 0, else those
// whose axis segments are at most cutoff apart), and the pair count
// (entanglement_adapter.cpp:56-117).
double orc_entanglement(orc_world* W, double cutoff, long long* pairs) {
  const int n = (int)W->B.size();
  std::vector<V3> a(n), e(n);
  for (int i = 0; i < n; ++i) rod_segment(*W, W->B[i], a[i], e[i]);
  double sum = 0.0;
  long long np = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (cutoff > 0.0) {
        double s, t;
        segseg(a[i], e[i], a[j], e[j], s, t);
        const V3 d = sub(add(a[i], mul(sub(e[i], a[i]), s)), add(a[j], mul(sub(e[j], a[j]), t)));
        if (std::sqrt(dot(d, d)) > cutoff) continue;
      }
      sum += std::fabs(linking_number(a[i], e[i], a[j], e[j]));
      ++np;
    }
  if (pairs) *pairs = np;
  return sum;
}

double orc_packing_fraction(orc_world* W) {
  double v = 0.0;
  for (const Body& b : W->B) v += M_PI * b.r * b.r * (2.0 * b.h) + (4.0 / 3.0) * M_PI * b.r * b.r * b.r;
  double box = (W->hi[0] - W->lo[0]) * (W->hi[1] - W->lo[1]) * (W->hi[2] - W->lo[2]);
  return v / box;
}

}  // extern "C"

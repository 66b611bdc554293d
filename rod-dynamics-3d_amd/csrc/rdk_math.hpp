// rdk_math.hpp -- FP64 vector, quaternion and capsule-geometry helpers shared
// by the device kernels and the host-side scene generator.
//
// Operation order follows the reference's glm-based code (cited per
// function) and the library is compiled with -ffp-contract=off, so the
// narrow phase and the integrators round exactly like a sequential FP64
// evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define RDK_HD __host__ __device__ __forceinline__

namespace rdk {

struct d3 {
  double x, y, z;
};
RDK_HD d3 mk3(double a, double b, double c) { return d3{a, b, c}; }
RDK_HD d3 operator+(d3 a, d3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
RDK_HD d3 operator-(d3 a, d3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
RDK_HD d3 operator*(d3 a, double s) { return mk3(a.x * s, a.y * s, a.z * s); }
RDK_HD d3 operator-(d3 a) { return mk3(-a.x, -a.y, -a.z); }
// glm::dot (vec3): (x*x + y*y) + z*z
RDK_HD double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// glm::cross
RDK_HD d3 cross(d3 a, d3 b) {
  return mk3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
RDK_HD double len(d3 a) { return sqrt(dot(a, a)); }
// glm::normalize(vec3): v * (1/sqrt(dot(v,v)))
RDK_HD d3 normalize(d3 a) { return a * (1.0 / sqrt(dot(a, a))); }

// std::max / std::min / std::clamp value semantics
RDK_HD double dmax(double a, double b) { return (a < b) ? b : a; }
RDK_HD double dmin(double a, double b) { return (b < a) ? b : a; }
RDK_HD double dclamp(double v, double lo, double hi) { return (v < lo) ? lo : (hi < v) ? hi : v; }

struct q4 {
  double w, x, y, z;
};
// glm quaternion product p*q
RDK_HD q4 qmul(q4 p, q4 q) {
  q4 r;
  r.w = p.w * q.w - p.x * q.x - p.y * q.y - p.z * q.z;
  r.x = p.w * q.x + p.x * q.w + p.y * q.z - p.z * q.y;
  r.y = p.w * q.y + p.y * q.w + p.z * q.x - p.x * q.z;
  r.z = p.w * q.z + p.z * q.w + p.x * q.y - p.y * q.x;
  return r;
}
// glm::normalize(quat)
RDK_HD q4 qnormalize(q4 q) {
  double l = sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
  if (l <= 0.0) return q4{1, 0, 0, 0};
  double inv = 1.0 / l;
  return q4{q.w * inv, q.x * inv, q.y * inv, q.z * inv};
}

// Capsule axis = column 1 of glm::mat3_cast(q) (rigid_body.cpp:137-139):
// the body's local +Y in world coordinates.
RDK_HD d3 axis_of(q4 q) {
  double qxx = q.x * q.x, qzz = q.z * q.z, qxy = q.x * q.y;
  double qyz = q.y * q.z, qwx = q.w * q.x, qwz = q.w * q.z;
  return mk3(2.0 * (qxy - qwz), 1.0 - 2.0 * (qxx + qzz), 2.0 * (qyz + qwx));
}

// World inverse inertia of a solid cylinder with symmetry axis u:
// R diag(1/Ixx, 1/Iyy, 1/Ixx) R^T = (1/Ixx) Id + (1/Iyy - 1/Ixx) u u^T.
RDK_HD d3 inertia_apply(d3 u, double i_perp, double i_axial, d3 v) {
  double c = (i_axial - i_perp) * dot(u, v);
  return mk3(v.x * i_perp + u.x * c, v.y * i_perp + u.y * c, v.z * i_perp + u.z * c);
}

// Closest-point parameters of two segments, Lumelsky as in DisMech
// (contact_detector.cpp:978-1035).
RDK_HD void segseg(d3 a1, d3 a2, d3 b1, d3 b2, double& s, double& t) {
  const d3 e1 = a2 - a1;
  const d3 e2 = b2 - b1;
  const d3 e12 = b1 - a1;
  const double D1 = dot(e1, e1);
  const double D2 = dot(e2, e2);
  const double S1 = dot(e1, e12);
  const double S2 = dot(e2, e12);
  const double R = dot(e1, e2);
  const double den = D1 * D2 - R * R;
  double uf;
  if (den == 0.0) {
    s = 0.0;
    t = -S2 / D2;
  } else {
    s = (S1 * D2 - S2 * R) / den;
    if (s > 1.0) s = 1.0;
    else if (s < 0.0) s = 0.0;
    t = (s * R - S2) / D2;
  }
  uf = t;
  if (uf > 1.0) uf = 1.0;
  else if (uf < 0.0) uf = 0.0;
  if (uf != t) {
    s = (uf * R + S1) / D1;
    if (s > 1.0) s = 1.0;
    else if (s < 0.0) s = 0.0;
    t = uf;
  }
}

// Minimum-image shift for body B relative to body A from the centre
// difference (contact_detector.cpp:827-839): shift = -floor(d/L + 0.5) L.
RDK_HD double image_shift(double d, double span, int periodic) {
  if (!periodic || !(span > 0.0)) return 0.0;
  double n = floor(d / span + 0.5);
  return -n * span;
}

struct NarrowOut {
  d3 pa, pb, n, shift;
  double dist, s, t;
  int type;
};

// Capsule-capsule narrow phase (contact_detector.cpp:811-910) in FP64.
// Returns true when dist^2 < (r_a + r_b + delta)^2.  When `full` is false
// only s, t, shift and the activation test are produced.
RDK_HD bool capsule_narrow(d3 xa, d3 ua, double ha, double ra, d3 xb, d3 ub, double hb,
                           double rb, const double* span, const int* periodic, double delta,
                           bool full, NarrowOut& o) {
  const d3 a1 = xa - ua * ha;
  const d3 a2 = xa + ua * ha;
  d3 b1 = xb - ub * hb;
  d3 b2 = xb + ub * hb;
  const d3 shift = mk3(image_shift(xb.x - xa.x, span[0], periodic[0]),
                       image_shift(xb.y - xa.y, span[1], periodic[1]),
                       image_shift(xb.z - xa.z, span[2], periodic[2]));
  b1 = b1 + shift;
  b2 = b2 + shift;
  const double hs = ra + rb;
  double s, t;
  segseg(a1, a2, b1, b2, s, t);
  const d3 pa = a1 + (a2 - a1) * s;
  const d3 pb = b1 + (b2 - b1) * t;
  const d3 diff = pb - pa;
  const double d2 = dot(diff, diff);
  const double act = hs + delta;
  o.s = s;
  o.t = t;
  o.shift = shift;
  if (!(d2 < act * act)) return false;
  if (full) {
    const double dist = sqrt(d2);
    const double eps = 1e-6;
    int type;
    if (s < eps && t < eps) type = 0;
    else if (s > 1.0 - eps && t > 1.0 - eps) type = 0;
    else if (s < eps || s > 1.0 - eps) type = 1;
    else if (t < eps || t > 1.0 - eps) type = 1;
    else type = 2;
    o.type = type;
    o.pa = pa;
    o.pb = pb;
    o.dist = dist;
    if (dist > 1e-10) o.n = mk3(diff.x / dist, diff.y / dist, diff.z / dist);
    else o.n = mk3(1.0, 0.0, 0.0);
  }
  return true;
}

// Tangent basis (nsc_contact.cpp:78-85)
RDK_HD void tangent_basis(d3 n, d3& t1, d3& t2) {
  d3 ref = (fabs(n.x) < 0.9) ? mk3(1, 0, 0) : mk3(0, 1, 0);
  t1 = normalize(cross(n, ref));
  t2 = cross(n, t1);
}

}  // namespace rdk

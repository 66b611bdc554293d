// rdk_gen.cpp -- host-side synthetic packings: random sequential addition of
// non-overlapping capsules and thermal velocities (the reference's
// resetScene "nonoverlap" populate with its 27-image check,
// main.cpp:1857-2140, and the thermal randomInit, main.cpp:2243-2263).
// Deterministic for a given seed (xoshiro256** stream), independent of the
// standard library's distributions.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/rdk.h"
#include "rdk_math.hpp"

namespace {

using rdk::d3;
using rdk::mk3;

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; ++i) s[i] = splitmix(seed);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// uniformly random unit quaternion (Shoemake); its rotated +Y axis is uniform on S^2
static rdk::q4 random_quat(Rng& g) {
  const double u1 = g.uniform(), u2 = g.uniform(), u3 = g.uniform();
  const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
  return rdk::q4{b * std::cos(2.0 * M_PI * u3), a * std::sin(2.0 * M_PI * u2), a * std::cos(2.0 * M_PI * u2),
                 b * std::sin(2.0 * M_PI * u3)};
}

// Uniform grid of rod chunks: every rod is cut into K = ceil(2h / chunk)
// pieces and each piece's bounding-sphere centre is binned (cell >= 2 * max
// piece radius + gap), so a rod that touches the candidate has a piece in a
// cell adjacent to one of the candidate's pieces.  Cells hold linked lists.
struct ChunkGrid {
  int dims[3];
  double lo[3], cell[3];
  int periodic[3];
  std::vector<int> head, next, rod;
  int coord(int k, double c) const {
    int i = (int)std::floor((c - lo[k]) / cell[k]);
    if (periodic[k]) {
      i %= dims[k];
      if (i < 0) i += dims[k];
    } else {
      i = std::min(std::max(i, 0), dims[k] - 1);
    }
    return i;
  }
  int lin(int x, int y, int z) const { return x + dims[0] * (y + dims[1] * z); }
  void insert(int c, int i) {
    next.push_back(head[c]);
    rod.push_back(i);
    head[c] = (int)rod.size() - 1;
  }
};

int chunks_of(double h, double chunk) { return std::max(1, (int)std::ceil(2.0 * h / chunk)); }

}  // namespace

extern "C" int rdk_gen_rsa(int n, const double* r, const double* h, const rdk_system_desc* box, double gap,
                           uint64_t seed, int max_attempts, double* x, double* q, int* placed) {
  if (n < 0 || !r || !h || !box || !x || !q || !placed || gap < 0.0) return RDK_ERR_ARG;
  *placed = 0;
  if (n == 0) return RDK_OK;
  Rng g(seed);
  double span[3];
  double rmax = 0.0, hmax = 0.0;
  for (int i = 0; i < n; ++i) {
    rmax = std::max(rmax, r[i]);
    hmax = std::max(hmax, h[i]);
  }
  const double range = 2.0 * (hmax + rmax) + gap;  // reach of two whole rods (image selection)
  const double chunk = std::max(8.0 * rmax, 1e-6);
  double crad = 0.0;
  for (int i = 0; i < n; ++i) crad = std::max(crad, h[i] / chunks_of(h[i], chunk) + r[i]);
  const double cmin = (2.0 * crad + gap) * (1.0 + 1e-9) + 1e-12;
  ChunkGrid G;
  int64_t ncell = 1;
  for (int k = 0; k < 3; ++k) {
    span[k] = box->hi[k] - box->lo[k];
    if (!(span[k] > 0.0)) return RDK_ERR_ARG;
    int d = std::max(1, (int)std::floor(span[k] / cmin));
    d = std::min(d, 256);
    if (box->periodic[k] && d < 3) d = 1;
    G.dims[k] = d;
    G.lo[k] = box->lo[k];
    G.cell[k] = span[k] / d;
    G.periodic[k] = box->periodic[k];
    ncell *= d;
  }
  G.head.assign((size_t)ncell, -1);
  std::vector<d3> cen(n), ax(n);
  std::vector<int> stamp(n, -1);
  int tag = 0;
  const int attempts = max_attempts > 0 ? max_attempts : 1000;
  const bool wall = box->floor_wall || box->ceil_wall;
  for (int i = 0; i < n; ++i) {
    bool ok = false;
    const int K = chunks_of(h[i], chunk);
    for (int att = 0; att < attempts && !ok; ++att) {
      d3 c;
      double* cp[3] = {&c.x, &c.y, &c.z};
      for (int k = 0; k < 3; ++k) *cp[k] = box->lo[k] + span[k] * g.uniform();
      const rdk::q4 qq = random_quat(g);
      const d3 u = rdk::axis_of(qq);
      // walls on a non-periodic z: keep both caps inside with clearance gap
      if (wall) {
        const double zlo = std::min(c.z - u.z * h[i], c.z + u.z * h[i]) - r[i];
        const double zhi = std::max(c.z - u.z * h[i], c.z + u.z * h[i]) + r[i];
        if (box->floor_wall && zlo < box->lo[2] + gap) continue;
        if (box->ceil_wall && zhi > box->hi[2] - gap) continue;
      }
      const d3 a1 = c - u * h[i], a2 = c + u * h[i];
      ++tag;
      bool clash = false;
      for (int kc = 0; kc < K && !clash; ++kc) {
        const d3 pc = a1 + (a2 - a1) * (((double)kc + 0.5) / (double)K);
        const int cc[3] = {G.coord(0, pc.x), G.coord(1, pc.y), G.coord(2, pc.z)};
        for (int dz = -1; dz <= 1 && !clash; ++dz) {
          if (G.dims[2] == 1 && dz != 0) continue;
          int zz = cc[2] + dz;
          if (G.periodic[2]) zz = (zz + G.dims[2]) % G.dims[2];
          else if (zz < 0 || zz >= G.dims[2]) continue;
          for (int dy = -1; dy <= 1 && !clash; ++dy) {
            if (G.dims[1] == 1 && dy != 0) continue;
            int yy = cc[1] + dy;
            if (G.periodic[1]) yy = (yy + G.dims[1]) % G.dims[1];
            else if (yy < 0 || yy >= G.dims[1]) continue;
            for (int dx = -1; dx <= 1 && !clash; ++dx) {
              if (G.dims[0] == 1 && dx != 0) continue;
              int xx = cc[0] + dx;
              if (G.periodic[0]) xx = (xx + G.dims[0]) % G.dims[0];
              else if (xx < 0 || xx >= G.dims[0]) continue;
              for (int e = G.head[G.lin(xx, yy, zz)]; e >= 0 && !clash; e = G.next[e]) {
                const int j = G.rod[e];
                if (stamp[j] == tag) continue;  // rod already tested for this attempt
                stamp[j] = tag;
                const double lim = r[i] + r[j] + gap;
                // every periodic image within reach (27-image check)
                int lo_i[3], hi_i[3];
                double base[3];
                const double dd[3] = {cen[j].x - c.x, cen[j].y - c.y, cen[j].z - c.z};
                for (int k = 0; k < 3; ++k) {
                  base[k] = 0.0;
                  lo_i[k] = hi_i[k] = 0;
                  if (G.periodic[k]) {
                    base[k] = -std::floor(dd[k] / span[k] + 0.5) * span[k];
                    if (span[k] < 2.0 * range + 1e-9) {
                      lo_i[k] = -1;
                      hi_i[k] = 1;
                    }
                  }
                }
                for (int iz = lo_i[2]; iz <= hi_i[2] && !clash; ++iz)
                  for (int iy = lo_i[1]; iy <= hi_i[1] && !clash; ++iy)
                    for (int ix = lo_i[0]; ix <= hi_i[0] && !clash; ++ix) {
                      const d3 sh = mk3(base[0] + ix * span[0], base[1] + iy * span[1], base[2] + iz * span[2]);
                      const d3 cj = cen[j] + sh;
                      const d3 dc = cj - c;
                      const double reach = h[i] + h[j] + lim;
                      if (rdk::dot(dc, dc) >= reach * reach) continue;
                      const d3 b1 = cj - ax[j] * h[j], b2 = cj + ax[j] * h[j];
                      double s, t;
                      rdk::segseg(a1, a2, b1, b2, s, t);
                      const d3 d = (b1 + (b2 - b1) * t) - (a1 + (a2 - a1) * s);
                      if (rdk::dot(d, d) < lim * lim) clash = true;
                    }
              }
            }
          }
        }
      }
      if (clash) continue;
      cen[i] = c;
      ax[i] = u;
      x[3 * i] = c.x;
      x[3 * i + 1] = c.y;
      x[3 * i + 2] = c.z;
      q[4 * i] = qq.w;
      q[4 * i + 1] = qq.x;
      q[4 * i + 2] = qq.y;
      q[4 * i + 3] = qq.z;
      for (int kc = 0; kc < K; ++kc) {
        const d3 pc = a1 + (a2 - a1) * (((double)kc + 0.5) / (double)K);
        G.insert(G.lin(G.coord(0, pc.x), G.coord(1, pc.y), G.coord(2, pc.z)), i);
      }
      ok = true;
    }
    if (!ok) return RDK_OK;  // *placed holds the count placed so far
    *placed = i + 1;
  }
  return RDK_OK;
}

extern "C" int rdk_gen_thermal(int n, const double* q, const double* r, const double* h, const double* mass, double kT,
                               uint64_t seed, double* v, double* w) {
  if (n < 0 || !q || !r || !h || !mass || !v || !w || kT < 0.0) return RDK_ERR_ARG;
  Rng g(seed);
  for (int i = 0; i < n; ++i) {
    const double m = mass[i];
    const double L = 2.0 * h[i];
    const double Ip = m * (3.0 * r[i] * r[i] + L * L) / 12.0;
    const double sv = m > 0.0 ? std::sqrt(kT / m) : 0.0;
    const double sw = Ip > 0.0 ? std::sqrt(kT / Ip) : 0.0;
    for (int k = 0; k < 3; ++k) v[3 * i + k] = sv * g.normal();
    const d3 u = rdk::axis_of(rdk::q4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]});
    // a unit vector perpendicular to u, then e2 = u x e1
    const d3 ref = std::fabs(u.x) < 0.9 ? mk3(1, 0, 0) : mk3(0, 1, 0);
    const d3 e1 = rdk::normalize(rdk::cross(u, ref));
    const d3 e2 = rdk::cross(u, e1);
    const double w1 = sw * g.normal(), w2 = sw * g.normal();
    const d3 om = e1 * w1 + e2 * w2;
    w[3 * i] = om.x;
    w[3 * i + 1] = om.y;
    w[3 * i + 2] = om.z;
  }
  return RDK_OK;
}

// ---- the reference's randomInit stream (App::resetScene, main.cpp:2213-2294) ----
// One std::mt19937(seed) for all rods in order; a rod that is skipped draws
// nothing.  Per rod, FP32 throughout as the reference holds it:
//   gaussian: fresh normal_distribution<float>(0, vSigma) and (0, wSigma);
//             v = three draws, w = three draws (braced lists: left to right);
//   thermal:  sigma_v = sqrt(kBT_trans / m), sigma_w = sqrt(kBT_rot / I_perp)
//             (0 when m or I_perp is not positive); v = three draws;
//             w = b1 e1 + b2 e2 with e1 a unit vector perpendicular to the
//             axis u and e2 = u x e1; the reference's expression
//             `dw(gen) * e1 + dw(gen) * e2` is evaluated right operand first
//             by g++, so b2 is drawn before b1;
//   uniform:  uniform_real_distribution<float>(-vSigma, vSigma) for v, and
//             w = wSpeed * (s cos(phi), c, s sin(phi)) with c = 2 U - 1 and
//             phi = 2 pi U drawn in that order from U(0, 1);
// then, when project is set (gaussian, uniform), w -= (w . u) u.  The axis u
// is the rod's quaternion (w, x, y, z) applied to +Y with glm's quaternion
// rotation, v + 2 (q.w (qv x v) + qv x (qv x v)).
namespace {
struct F3 {
  float x, y, z;
};
F3 f3(float x, float y, float z) { return F3{x, y, z}; }
F3 fcross(const F3& a, const F3& b) {
  return f3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
float fdot(const F3& a, const F3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
F3 fscale(const F3& a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
F3 fadd(const F3& a, const F3& b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
F3 fsub(const F3& a, const F3& b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
F3 fnormalize(const F3& a) { return fscale(a, 1.0f / std::sqrt(fdot(a, a))); }
F3 frot_y(const float* q) {  // quaternion (w, x, y, z) applied to (0, 1, 0)
  const F3 qv = f3(q[1], q[2], q[3]), v = f3(0.0f, 1.0f, 0.0f);
  const F3 uv = fcross(qv, v), uuv = fcross(qv, uv);
  return fadd(v, fscale(fadd(fscale(uv, q[0]), uuv), 2.0f));
}
}  // namespace

extern "C" int rdk_random_init(int n, const float* q, const float* mass, const float* i_perp, int mode, float v_sigma,
                               float w_sigma, float w_speed, float kbt_trans, float kbt_rot, uint32_t seed,
                               int project, const unsigned char* skip, float* v, float* w) {
  if (n < 0 || !q || !v || !w || mode < 0 || mode > 2) return RDK_ERR_ARG;
  if (mode == RDK_RANDOM_INIT_THERMAL && (!mass || !i_perp)) return RDK_ERR_ARG;
  std::mt19937 gen(seed ? seed : std::random_device{}());
  std::uniform_real_distribution<float> uni(0.0f, 1.0f);
  for (int i = 0; i < n; ++i) {
    if (skip && skip[i]) continue;
    const float* qi = q + 4 * i;
    F3 vv, ww;
    if (mode == RDK_RANDOM_INIT_THERMAL) {
      const float sv = mass[i] > 0.0f ? std::sqrt(kbt_trans / mass[i]) : 0.0f;
      const float sw = i_perp[i] > 0.0f ? std::sqrt(kbt_rot / i_perp[i]) : 0.0f;
      std::normal_distribution<float> dv(0.0f, sv), dw(0.0f, sw);
      const float v0 = dv(gen), v1 = dv(gen), v2 = dv(gen);
      vv = f3(v0, v1, v2);
      const F3 u = frot_y(qi);
      const F3 e1 = fnormalize(fcross(u, std::fabs(u.x) < 0.9f ? f3(1, 0, 0) : f3(0, 1, 0)));
      const F3 e2 = fcross(u, e1);
      const float b2 = dw(gen);  // the right operand of the reference's sum, first under g++
      const float b1 = dw(gen);
      ww = fadd(fscale(e1, b1), fscale(e2, b2));
    } else if (mode == RDK_RANDOM_INIT_GAUSSIAN) {
      std::normal_distribution<float> dv(0.0f, v_sigma), dw(0.0f, w_sigma);
      const float v0 = dv(gen), v1 = dv(gen), v2 = dv(gen);
      const float w0 = dw(gen), w1 = dw(gen), w2 = dw(gen);
      vv = f3(v0, v1, v2);
      ww = f3(w0, w1, w2);
    } else {
      std::uniform_real_distribution<float> un(-v_sigma, v_sigma);
      const float v0 = un(gen), v1 = un(gen), v2 = un(gen);
      vv = f3(v0, v1, v2);
      const float c = 2.0f * uni(gen) - 1.0f;
      const float phi = 2.0f * float(M_PI) * uni(gen);
      const float s = std::sqrt(std::max(0.0f, 1.0f - c * c));
      ww = fscale(f3(s * std::cos(phi), c, s * std::sin(phi)), w_speed);
    }
    if (project && mode != RDK_RANDOM_INIT_THERMAL) {
      const F3 u = frot_y(qi);
      ww = fsub(ww, fscale(u, fdot(ww, u)));
    }
    v[3 * i] = vv.x;
    v[3 * i + 1] = vv.y;
    v[3 * i + 2] = vv.z;
    w[3 * i] = ww.x;
    w[3 * i + 1] = ww.y;
    w[3 * i + 2] = ww.z;
  }
  return RDK_OK;
}

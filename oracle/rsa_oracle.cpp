// rsa_oracle.cpp -- CPU oracle (test infrastructure only) of the reference's
// non-overlapping populate: App::resetScene, populate.mode "nonoverlap"
// (src/app/main.cpp:1857-2140).  Serial random sequential addition of equal
// capsules in FP32, restated step by step so that the GPU generator can be
// compared with it bit for bit:
//
//   * one std::mt19937 seeded with populate.seed and one
//     std::uniform_real_distribution<float>(0, 1) (main.cpp:1671-1675);
//   * each attempt draws 5 numbers: centre c = bmin + (r1, r2, r3) * box, then
//     the direction u = 2 r4 - 1 (y), phi = 2 pi r5, s = sqrt(max(0, 1 - u^2)),
//     dir = normalize(s cos phi, u, s sin phi) (main.cpp:1925-1930,2007-2011);
//   * placement grid of cell size cs over [bmin, bmax), n = max(1, floor(box /
//     cs)) per axis (main.cpp:1318-1322,1867-1872); each accepted rod is listed
//     in every cell its AABB (|dir| halfL + R) covers, wrapped on periodic
//     boxes, clamped otherwise (main.cpp:1910-1921,2118-2127);
//   * a candidate collides when a rod listed in a cell of its AABB has squared
//     segment distance < (2R)^2: the minimum over the 27 images on periodic
//     boxes (main.cpp:1942-1995, including its end-point rule that leaves s
//     unclamped above 1), the single-image variant otherwise
//     (main.cpp:2075-2101);
//   * a rod that fails max_attempts times stops the populate (main.cpp:2133).
//
// glm's operations are written out in its order: dot = (x x' + y y') + z z',
// normalize = v * (1 / sqrt(dot(v, v))), vector / scalar per component.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <vector>

namespace {

struct F3 {
  float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
inline float dot(F3 a, F3 b) {
  const float tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z;
  return (tx + ty) + tz;
}

// squared distance of segments p0p1 and q0q1 (one image), the periodic
// variant's end-point rule when pbc_rule is set
float seg_d2(F3 p0, F3 p1, F3 q0, F3 q1, bool pbc_rule) {
  const float eps = 1e-12f;
  const F3 u = p1 - p0, v = q1 - q0, w0 = p0 - q0;
  const float uu = dot(u, u), vv = dot(v, v), uv = dot(u, v);
  const float wu = dot(w0, u), wv = dot(w0, v);
  const float D = uu * vv - uv * uv;
  float s, t;
  if (std::fabs(D) < eps) {
    s = 0.0f;
    t = (vv >= eps) ? (-wv / vv) : 0.0f;
  } else {
    s = (uv * wv - vv * wu) / D;
    t = (uu * wv - uv * wu) / D;
  }
  s = s < 0.0f ? 0.0f : (s > 1.0f ? 1.0f : s);
  t = (s * uv + wv) / (vv >= eps ? vv : 1.0f);
  t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
  const float su = (-wu + t * uv) / (uu >= eps ? uu : 1.0f);
  if (!(t > 1e-6f && t < 1.0f - 1e-6f)) {
    if (pbc_rule)
      s = su < 0.0f ? 0.0f : su;
    else
      s = su < 0.0f ? 0.0f : (su > 1.0f ? 1.0f : su);
  }
  const F3 d = (w0 + s * u) - t * v;
  return dot(d, d);
}

float pair_d2(F3 p0, F3 p1, F3 q0, F3 q1, bool pbc, F3 L) {
  if (!pbc) return seg_d2(p0, p1, q0, q1, false);
  float m = std::numeric_limits<float>::max();
  for (int k = -1; k <= 1; ++k)
    for (int j = -1; j <= 1; ++j)
      for (int i = -1; i <= 1; ++i) {
        const F3 sh = f3((float)i * L.x, (float)j * L.y, (float)k * L.z);
        const float d2 = seg_d2(p0, p1, q0 + sh, q1 + sh, true);
        if (d2 < m) m = d2;
      }
  return m;
}

}  // namespace

extern "C" {

// Returns the number of rods placed; c and u receive n x 3 floats (centres and
// unit directions, in placement order); *attempts the number of candidates
// drawn.
int orc_populate_nonoverlap(int n, float length, float diameter, const float* bmin_, const float* bmax_,
                            int pbc, float cell_size, unsigned seed, int max_attempts, float* c_out,
                            float* u_out, long long* attempts) {
  const F3 bmin = f3(bmin_[0], bmin_[1], bmin_[2]), bmax = f3(bmax_[0], bmax_[1], bmax_[2]);
  const F3 box = bmax - bmin;
  const float halfL = 0.5f * length, R = 0.5f * diameter;
  const float diam2 = (2.0f * R) * (2.0f * R);
  const float cs = cell_size > 0.0f ? cell_size : std::max(0.25f * length, 2.5f * R);
  const int nx = std::max(1, (int)std::floor(box.x / cs));
  const int ny = std::max(1, (int)std::floor(box.y / cs));
  const int nz = std::max(1, (int)std::floor(box.z / cs));
  std::vector<std::vector<int>> cells((size_t)nx * ny * nz);
  auto cell_of = [&](int i, int dim) {
    if (pbc) {
      int r = i % dim;
      return r < 0 ? r + dim : r;
    }
    return std::min(std::max(i, 0), dim - 1);
  };
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> urand(0.0f, 1.0f);
  std::vector<F3> C, U;
  long long drawn = 0;
  int placed = 0;
  for (int i = 0; i < n; ++i) {
    bool ok = false;
    for (int att = 0; att < max_attempts && !ok; ++att) {
      ++drawn;
      const float r1 = urand(gen), r2 = urand(gen), r3 = urand(gen);
      const F3 c = f3(bmin.x + r1 * box.x, bmin.y + r2 * box.y, bmin.z + r3 * box.z);
      const float uy = 2.0f * urand(gen) - 1.0f;
      const float phi = 2.0f * float(M_PI) * urand(gen);
      const float s = std::sqrt(std::max(0.0f, 1.0f - uy * uy));
      const F3 raw = f3(s * std::cos(phi), uy, s * std::sin(phi));
      const F3 dir = raw * (1.0f / std::sqrt(dot(raw, raw)));
      const F3 p0 = c - dir * halfL, p1 = c + dir * halfL;
      const F3 ext = f3(std::fabs(dir.x) * halfL + R, std::fabs(dir.y) * halfL + R, std::fabs(dir.z) * halfL + R);
      const F3 mn = c - ext, mx = c + ext;
      const int i0[3] = {(int)std::floor((mn.x - bmin.x) / cs), (int)std::floor((mn.y - bmin.y) / cs),
                         (int)std::floor((mn.z - bmin.z) / cs)};
      const int i1[3] = {(int)std::floor((mx.x - bmin.x) / cs), (int)std::floor((mx.y - bmin.y) / cs),
                         (int)std::floor((mx.z - bmin.z) / cs)};
      bool hit = false;
      for (int iz = i0[2]; iz <= i1[2] && !hit; ++iz)
        for (int iy = i0[1]; iy <= i1[1] && !hit; ++iy)
          for (int ix = i0[0]; ix <= i1[0] && !hit; ++ix) {
            const int cell = cell_of(ix, nx) + nx * (cell_of(iy, ny) + ny * cell_of(iz, nz));
            for (int j : cells[cell]) {
              const F3 q0 = C[j] - U[j] * halfL, q1 = C[j] + U[j] * halfL;
              if (pair_d2(p0, p1, q0, q1, pbc != 0, box) < diam2) {
                hit = true;
                break;
              }
            }
          }
      if (hit) continue;
      C.push_back(c);
      U.push_back(dir);
      for (int iz = i0[2]; iz <= i1[2]; ++iz)
        for (int iy = i0[1]; iy <= i1[1]; ++iy)
          for (int ix = i0[0]; ix <= i1[0]; ++ix)
            cells[cell_of(ix, nx) + nx * (cell_of(iy, ny) + ny * cell_of(iz, nz))].push_back(i);
      ok = true;
    }
    if (!ok) break;
    c_out[3 * i] = C[i].x;
    c_out[3 * i + 1] = C[i].y;
    c_out[3 * i + 2] = C[i].z;
    u_out[3 * i] = U[i].x;
    u_out[3 * i + 1] = U[i].y;
    u_out[3 * i + 2] = U[i].z;
    ++placed;
  }
  if (attempts) *attempts = drawn;
  return placed;
}

// Squared segment distance of the populate test (one image); pbc_rule selects
// the periodic variant's end-point rule.
float orc_rsa_seg_d2(const float* p0, const float* p1, const float* q0, const float* q1, int pbc_rule) {
  return seg_d2(f3(p0[0], p0[1], p0[2]), f3(p1[0], p1[1], p1[2]), f3(q0[0], q0[1], q0[2]), f3(q1[0], q1[1], q1[2]),
                pbc_rule != 0);
}

// The first k outputs of std::uniform_real_distribution<float>(0, 1) on
// std::mt19937(seed): the candidate stream's raw numbers.
void orc_mt_uniform_float(unsigned seed, int k, float* out) {
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> urand(0.0f, 1.0f);
  for (int i = 0; i < k; ++i) out[i] = urand(gen);
}

// The k-th (1-based) raw output of std::mt19937(seed).
unsigned orc_mt_raw(unsigned seed, long long k) {
  std::mt19937 gen(seed);
  gen.discard((unsigned long long)(k - 1));
  return (unsigned)gen();
}

}  // extern "C"

// ---- randomInit (App::resetScene, main.cpp:2213-2294) --------------------------
// Restated in the reference's own expression forms on a minimal FP32 vector
// type whose operators are free functions as glm's are, so that g++ (the
// reference's compiler on Linux) decides the evaluation order of
// `dw(gen) * e1 + dw(gen) * e2` itself.  glm's quaternion rotation:
// v + ((qv x v) q.w + qv x (qv x v)) 2; normalize(v) = v * (1 / sqrt(dot(v, v))).
namespace orc_glm {
struct vec3 {
  float x, y, z;
  vec3(float a, float b, float c) : x(a), y(b), z(c) {}
};
inline vec3 operator*(float s, const vec3& v) { return vec3(s * v.x, s * v.y, s * v.z); }
inline vec3 operator*(const vec3& v, float s) { return vec3(v.x * s, v.y * s, v.z * s); }
inline vec3 operator+(const vec3& a, const vec3& b) { return vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline vec3& operator-=(vec3& a, const vec3& b) {
  a = vec3(a.x - b.x, a.y - b.y, a.z - b.z);
  return a;
}
inline float dot(const vec3& a, const vec3& b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline vec3 cross(const vec3& a, const vec3& b) {
  return vec3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
inline vec3 normalize(const vec3& v) { return v * (1.0f / std::sqrt(dot(v, v))); }
struct quat {
  float w, x, y, z;
};
inline vec3 operator*(const quat& q, const vec3& v) {
  const vec3 qv(q.x, q.y, q.z);
  const vec3 uv = cross(qv, v);
  const vec3 uuv = cross(qv, uv);
  return v + ((uv * q.w) + uuv) * 2.0f;
}
}  // namespace orc_glm

extern "C" {

// mode 0 gaussian, 1 thermal, 2 uniform; q: (w, x, y, z) per rod; skip may be NULL.
void orc_random_init(int n, const float* q, const float* mass, const float* i_perp, int mode, float v_sigma,
                     float w_sigma, float w_speed, float kbt_trans, float kbt_rot, unsigned seed, int project,
                     const unsigned char* skip, float* v, float* w) {
  using namespace orc_glm;
  std::mt19937 gen(seed);
  std::uniform_real_distribution<float> uni(0.0f, 1.0f);
  auto uniform_dir_s2 = [&](std::mt19937& g) {
    float u = 2.0f * uni(g) - 1.0f;
    float phi = 2.0f * float(M_PI) * uni(g);
    float s = std::sqrt(std::max(0.0f, 1.0f - u * u));
    return vec3(s * std::cos(phi), u, s * std::sin(phi));
  };
  auto arbitrary_perp = [](const vec3& u) -> vec3 {
    vec3 ref = (std::abs(u.x) < 0.9f) ? vec3(1, 0, 0) : vec3(0, 1, 0);
    return normalize(cross(u, ref));
  };
  for (int i = 0; i < n; ++i) {
    if (skip && skip[i]) continue;
    const quat qq{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
    vec3 rv(0, 0, 0), rw(0, 0, 0);
    if (mode == 1) {
      const float sigma_v = (mass[i] > 0.0f) ? std::sqrt(kbt_trans / mass[i]) : 0.0f;
      const float sigma_w = (i_perp[i] > 0.0f) ? std::sqrt(kbt_rot / i_perp[i]) : 0.0f;
      std::normal_distribution<float> dv(0.0f, sigma_v);
      std::normal_distribution<float> dw(0.0f, sigma_w);
      rv = {dv(gen), dv(gen), dv(gen)};
      vec3 u_ax = qq * vec3(0, 1, 0);
      vec3 e1 = arbitrary_perp(u_ax);
      vec3 e2 = cross(u_ax, e1);
      rw = dw(gen) * e1 + dw(gen) * e2;
    } else if (mode == 0) {
      std::normal_distribution<float> dv(0.0f, v_sigma);
      std::normal_distribution<float> dw(0.0f, w_sigma);
      rv = {dv(gen), dv(gen), dv(gen)};
      rw = {dw(gen), dw(gen), dw(gen)};
      if (project) {
        vec3 u_ax = qq * vec3(0, 1, 0);
        rw -= dot(rw, u_ax) * u_ax;
      }
    } else {
      std::uniform_real_distribution<float> uniform(-v_sigma, v_sigma);
      rv = {uniform(gen), uniform(gen), uniform(gen)};
      rw = w_speed * uniform_dir_s2(gen);
      if (project) {
        vec3 u_ax = qq * vec3(0, 1, 0);
        rw -= dot(rw, u_ax) * u_ax;
      }
    }
    v[3 * i] = rv.x;
    v[3 * i + 1] = rv.y;
    v[3 * i + 2] = rv.z;
    w[3 * i] = rw.x;
    w[3 * i + 1] = rw.y;
    w[3 * i + 2] = rw.z;
  }
}

// The first k outputs of std::normal_distribution<float>(0, 1) on std::mt19937(seed), one distribution object.
void orc_mt_normal_float(unsigned seed, int k, float* out) {
  std::mt19937 gen(seed);
  std::normal_distribution<float> d(0.0f, 1.0f);
  for (int i = 0; i < k; ++i) out[i] = d(gen);
}

}  // extern "C"

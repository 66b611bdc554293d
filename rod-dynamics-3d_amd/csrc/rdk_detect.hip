// rdk_detect.hip -- cell-list broad phase with periodic images and the FP64
// capsule narrow phase (replaces ContactDetector::detectContactsSpatialHash,
// contact_detector.cpp:487-809, and the O(N^2) kernels of
// cuda_broadphase.cu:133-581).
//
// Pipeline per call:
//   1. chunk_keys      every rod is cut into K = ceil(2h/chunk_len) sub-segments
//                      ("chunks"); each chunk's bounding-sphere centre is binned
//                      into a uniform per-system grid whose cell is at least
//                      2*max(chunk radius) + delta, so touching chunks always sit
//                      in the same or adjacent cells (27-cell stencil, periodic
//                      wrap on periodic axes); the FP32 chunk records below
//                      are written here too, in chunk order.
//   2. radix sort of (cell key, chunk id)
//   3. cell_start      dense start array: start[c] = first sorted entry with
//                      cell >= c, for every cell (and one past the last), so
//                      the three x-adjacent cells of a stencil row are one
//                      contiguous range [start[x-1], start[x+2])
//   4. chunk_gather    permutes the FP32 (centre, radius), FP32 (axis, half
//                      length), rod id and chunk index into sorted (cell)
//                      order so the stencil walk reads contiguous memory
//   5. broad_narrow    one thread per sorted chunk walks the 9 stencil rows
//                      (27 cells), several candidates in flight per lane;
//                      each unordered chunk pair is examined once (by the
//                      lower sorted position).  Workgroups are mapped so that
//                      each XCD (own L2) walks one contiguous run of the
//                      sorted chunks, i.e. one slab of the grid.  An FP32
//                      bounding-sphere test with the chunk-centre minimum
//                      image and a rounding margin rejects far pairs
//                      conservatively; survivors are queued per wavefront,
//                      pass a separating-axis lower bound on the chunk-segment
//                      distance (also FP32, also conservative), and what is
//                      left runs the full-rod FP64 Lumelsky narrow phase, 64
//                      pairs at a time.  A pair is emitted only by the chunk
//                      pair that contains the closest points (owner rule), so
//                      each contact appears exactly once whatever the
//                      chunking.  Emission is wavefront-aggregated (ballot +
//                      popcount, one atomic per wavefront).
//   6. radix sort of the 64-bit (a,b) keys -> canonical (a,b) order
//   7. contact_record  recomputes the narrow phase per sorted pair and writes
//                      the structure-of-arrays contact record.
#include <algorithm>
#include <cmath>

#include "rdk_internal.hpp"

namespace rdk {

__device__ __forceinline__ int cell_coord(double c, double lo, double inv_cell, int dims, int periodic) {
  int i = (int)floor((c - lo) * inv_cell);
  if (periodic) {
    i %= dims;
    if (i < 0) i += dims;
  } else {
    i = i < 0 ? 0 : (i >= dims ? dims - 1 : i);
  }
  return i;
}

// Cell order of a system's grid: x fastest, then rows (y, z) in bands of
// `T` y-rows; inside a band z runs over the whole grid before the next band.
// Every row stays contiguous in x (the stencil walk needs that) and the
// sorted chunks walk one band's z-planes in turn, so the three planes a
// stencil reads stay small enough for one XCD's L2 on large dense grids.
// T = d1 is the plain order row = y + d1 * z.
__host__ __device__ __forceinline__ int cell_row(int d1, int d2, int T, int y, int z) {
  const int y0 = (y / T) * T;
  const int tl = d1 - y0 < T ? d1 - y0 : T;
  return y0 * d2 + z * tl + (y - y0);
}
__host__ __device__ __forceinline__ void row_yz(int d1, int d2, int T, int row, int& y, int& z) {
  const int y0 = (row / (T * d2)) * T;
  const int tl = d1 - y0 < T ? d1 - y0 : T;
  const int rem = row - y0 * d2;
  z = rem / tl;
  y = y0 + rem % tl;
}

__device__ __forceinline__ q4 load_q(const BodyPtrs& B, int i) {
  return q4{B.q[0][i], B.q[1][i], B.q[2][i], B.q[3][i]};
}
__device__ __forceinline__ d3 load3(double* const* a, int i) { return mk3(a[0][i], a[1][i], a[2][i]); }

__device__ __forceinline__ void chunk_geom(const BodyPtrs& B, int i, int k, int K, d3& c, double& rad, d3& u) {
  const d3 x = load3(B.x, i);
  u = axis_of(load_q(B, i));
  const double h = B.h[i];
  const d3 a1 = x - u * h;
  const d3 a2 = x + u * h;
  const double sc = ((double)k + 0.5) / (double)K;
  c = a1 + (a2 - a1) * sc;
  rad = h / (double)K + B.r[i];
}

// Narrow-phase inputs of one rod, one 64-byte record (position, FP64 axis,
// half length, radius as four double2) plus (system, chunk count): the FP64
// narrow phase and the contact records read two lines per rod instead of one
// per field array.  Written by k_chunk_keys from the rod's chunk 0 with the
// same FP64 operations the narrow phase would apply to the body arrays.
struct RodRecs {
  double2* rec;  // 4 per rod
  int2* info;    // (system, chunk count)
};
struct RodIn {
  d3 x, u;
  double h, r;
};
__device__ __forceinline__ RodIn load_rod(const RodRecs& R, int i) {
  const double2 a = R.rec[4 * i], b = R.rec[4 * i + 1], c = R.rec[4 * i + 2], d = R.rec[4 * i + 3];
  return RodIn{mk3(a.x, a.y, b.x), mk3(b.y, c.x, c.y), d.x, d.y};
}
static RodRecs rod_recs(rdk_ctx* ctx) {
  RodRecs R;
  R.rec = (double2*)ctx->rodrec.p;
  R.info = (int2*)ctx->rodinfo.p;
  return R;
}

// Cell key of each chunk, and its FP32 prefilter records in chunk order
// (centre and radius; axis and half length of the chunk segment), which
// k_chunk_gather then permutes into sorted order: the rod's fields are read
// here, in chunk order (a rod's chunks are consecutive), instead of once per
// sorted chunk from scattered positions.
__global__ void k_chunk_keys(BodyPtrs B, const SysDev* __restrict__ sys, const int* __restrict__ chunk_rod,
                             const int* __restrict__ chunk_k, const int* __restrict__ rod_nchunk, int M,
                             uint32_t* __restrict__ key, int* __restrict__ id, float4* __restrict__ ucf,
                             float4* __restrict__ uca, RodRecs R) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int i = chunk_rod[m];
  const SysDev S = sys[B.sys[i]];
  d3 c, u;
  double rad;
  const int K = rod_nchunk[i];
  const int k = chunk_k[m];
  chunk_geom(B, i, k, K, c, rad, u);
  int cx = cell_coord(c.x, S.lo[0], S.inv_cell[0], S.dims[0], S.periodic[0]);
  int cy = cell_coord(c.y, S.lo[1], S.inv_cell[1], S.dims[1], S.periodic[1]);
  int cz = cell_coord(c.z, S.lo[2], S.inv_cell[2], S.dims[2], S.periodic[2]);
  key[m] = (uint32_t)(S.cell_off + cx + S.dims[0] * cell_row(S.dims[1], S.dims[2], S.ytile, cy, cz));
  id[m] = m;
  ucf[m] = make_float4((float)c.x, (float)c.y, (float)c.z, (float)rad);
  uca[m] = make_float4((float)u.x, (float)u.y, (float)u.z, (float)(B.h[i] / (double)K));
  if (k == 0) {
    const d3 x = load3(B.x, i);
    R.rec[4 * i] = make_double2(x.x, x.y);
    R.rec[4 * i + 1] = make_double2(x.z, u.x);
    R.rec[4 * i + 2] = make_double2(u.y, u.z);
    R.rec[4 * i + 3] = make_double2(B.h[i], B.r[i]);
    R.info[i] = make_int2(B.sys[i], K);
  }
}

__global__ void k_fill_int(int* __restrict__ a, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// start[c] = lower bound of cell c in the sorted keys, c = 0 .. ncell
__global__ void k_cell_start(const uint32_t* __restrict__ key, int M, int ncell, int* __restrict__ start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncell) return;
  int lo = 0, hi = M;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < (uint32_t)c) lo = mid + 1;
    else hi = mid;
  }
  start[c] = lo;
}

struct SortedChunks {
  const uint32_t* key;  // sorted cell keys
  int *rod, *k;
  float4* cf;  // FP32 (centre, radius) for the conservative prefilter
  float4* ca;  // FP32 (axis, half length of the chunk segment)
};

__global__ void k_chunk_gather(const int* __restrict__ sorted_id, const int* __restrict__ chunk_rod,
                               const int* __restrict__ chunk_k, const float4* __restrict__ ucf,
                               const float4* __restrict__ uca, int M, SortedChunks S) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  const int m = sorted_id[p];
  S.rod[p] = chunk_rod[m];
  S.k[p] = chunk_k[m];
  S.cf[p] = ucf[m];
  S.ca[p] = uca[m];
}

// Workgroups are dispatched round-robin over the XCDs (8 on MI355X, each with
// its own L2).  Renumber them so that XCD x runs one contiguous block range.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_block(int blk, int nblk) {
  const int q = nblk / NUM_XCD, r = nblk % NUM_XCD;
  const int x = blk % NUM_XCD, j = blk / NUM_XCD;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// Wavefront-aggregated append: one atomic per wavefront of active lanes.
__device__ __forceinline__ unsigned wave_append(unsigned* counter) {
  const unsigned long long act = __ballot(1);
  const int lane = __lane_id();
  const int leader = __ffsll((long long)act) - 1;
  const unsigned rank = (unsigned)__popcll(act & ((1ull << lane) - 1ull));
  unsigned base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned)__popcll(act));
  base = __shfl(base, leader);
  return base + rank;
}

__device__ __forceinline__ int owner_chunk(double s, int K) {
  int k = (int)floor(s * (double)K);
  return k < 0 ? 0 : (k >= K ? K - 1 : k);
}

// FP64 narrow phase of one surviving chunk pair (p, q), both in sorted
// order; emits the canonical (lower, higher) rod pair when this chunk pair
// owns the closest points.
__device__ __forceinline__ void narrow_pair(const RodRecs& R, const SysDev* __restrict__ sys, const SortedChunks& SC,
                                            int p, int q, double delta, uint64_t* __restrict__ out, unsigned cap,
                                            unsigned* __restrict__ counter) {
  const int a = SC.rod[p], b = SC.rod[q];
  const int A = a < b ? a : b, Bi = a < b ? b : a;
  const int kA = a < b ? SC.k[p] : SC.k[q], kB = a < b ? SC.k[q] : SC.k[p];
  const int2 ia = R.info[A], ib = R.info[Bi];
  const RodIn ra = load_rod(R, A), rb = load_rod(R, Bi);
  const SysDev& S = sys[ia.x];
  NarrowOut no;
  if (capsule_narrow(ra.x, ra.u, ra.h, ra.r, rb.x, rb.u, rb.h, rb.r, S.span, S.periodic, delta, false, no)) {
    if (owner_chunk(no.s, ia.y) == kA && owner_chunk(no.t, ib.y) == kB) {
      const unsigned pos = wave_append(counter);
      if (pos < cap) out[pos] = ((uint64_t)(uint32_t)A << 32) | (uint64_t)(uint32_t)Bi;
    }
  }
}

// Stencil ranges of one chunk: 9 rows (dz, dy), each as up to two contiguous
// x ranges of sorted positions (two after a periodic wrap).  Range r = 2 *
// row + part.  Returns false when no range is left; q is clipped to q > p.
struct StencilWalk {
  int p, cy, cz, xa0, xa1, xb;  // x ranges: [xa0, xa1] and, after a periodic wrap, the single cell xb
  int d0, d1, d2, off, ytile;  // the system's grid (kept in registers: on the path to the cstart loads)
  int kp;               // cell key of chunk p
  bool two, py, pz;     // periodic in y, z
  int r, q, q1;
  // Cell keys [c0, c1] of range r, false when the range does not exist (a
  // wrap part without a wrap, a row outside a closed box, a duplicate row of
  // a one-cell periodic axis) or holds only sorted positions below p: its
  // last cell key is below p's own cell (cstart is monotone and p >=
  // cstart[kp]), so it is empty after the q > p clip; about half of the
  // stencil is skipped this way without loading its bounds.
  __device__ __forceinline__ bool interval(int rr, int& c0, int& c1) const {
    const int row = rr >> 1, part = rr & 1;
    if (part == 1 && !two) return false;
    const int dz = row / 3 - 1, dy = row % 3 - 1;
    int z = cz + dz;
    if (pz) {
      if (d2 == 1 && dz != 0) return false;
      z += z < 0 ? d2 : (z >= d2 ? -d2 : 0);  // |dz| <= 1, d2 == 1 or >= 3
    } else if (z < 0 || z >= d2) {
      return false;
    }
    int y = cy + dy;
    if (py) {
      if (d1 == 1 && dy != 0) return false;
      y += y < 0 ? d1 : (y >= d1 ? -d1 : 0);
    } else if (y < 0 || y >= d1) {
      return false;
    }
    const int base = off + d0 * cell_row(d1, d2, ytile, y, z);
    c0 = base + (part ? xb : xa0);
    c1 = base + (part ? xb : xa1);
    return c1 >= kp;
  }
  // opens the next non-empty range below r_end (18: all of them)
  __device__ __forceinline__ bool open_next(const int* __restrict__ cstart, int r_end = 18) {
    while (r < r_end) {
      int c0, c1;
      const bool ok = interval(r, c0, c1);
      ++r;
      if (!ok) continue;
      q = cstart[c0];
      q1 = cstart[c1 + 1];
      if (q <= p) q = p + 1;
      if (q < q1) return true;
    }
    return false;
  }
};

constexpr int BN_QUEUE = 448;  // per wavefront: < 64 pending + one append round of <= 64 x BN_UNROLL
constexpr int BN_QUEUE2 = 128;  // separating-axis survivors: < 64 pending + one round of <= 64

// Separating-axis lower bound on the distance between the axis segments of
// chunks p and q, over the axes u_p, u_q and u_p x u_q (the last only when it
// is not short, so its normalisation stays accurate) and the closest-point
// direction described below: the distance between
// two point sets is at least the gap between their projections on any unit
// axis.  If the owner chunk pair of a touching rod pair is (p, q), its
// segments are closer than r_p + r_q + delta, so the bound is too; the pair is
// dropped only when the FP32 bound exceeds that by more than the rounding
// margin (centre rounding as in the sphere test, plus 1.6e-4 of the sphere
// threshold for the axes, half lengths and the normalisation, >> FP32 eps).
// Valid only when a single periodic image of q can lie within the sphere
// threshold (every periodic span > 2 x threshold): then the chunk-centre
// minimum image is the image the narrow phase uses; otherwise the pair is
// kept.
__device__ __forceinline__ bool sat_keep(const SortedChunks& SC, int p, int q, const float* spanf,
                                         const float* ispanf, int per, float extra) {
  const float4 cp = SC.cf[p], cq = SC.cf[q], ap = SC.ca[p], aq = SC.ca[q];
  float dx = cq.x - cp.x, dy = cq.y - cp.y, dz = cq.z - cp.z;
  if (per & 1) dx -= spanf[0] * rintf(dx * ispanf[0]);
  if (per & 2) dy -= spanf[1] * rintf(dy * ispanf[1]);
  if (per & 4) dz -= spanf[2] * rintf(dz * ispanf[2]);
  const float bound = cp.w + cq.w + extra;
  const float two = 2.01f * bound;
  if (((per & 1) && !(spanf[0] > two)) || ((per & 2) && !(spanf[1] > two)) || ((per & 4) && !(spanf[2] > two)))
    return true;
  const float hp = ap.w, hq = aq.w;
  const float lim = (cp.w - hp) + (cq.w - hq) + extra + 1.6e-4f * bound;
  const float c = fabsf(ap.x * aq.x + ap.y * aq.y + ap.z * aq.z);
  const float dp = fabsf(dx * ap.x + dy * ap.y + dz * ap.z);
  const float dq = fabsf(dx * aq.x + dy * aq.y + dz * aq.z);
  float lb = fmaxf(dp - hp - hq * c, dq - hq - hp * c);
  const float nx = ap.y * aq.z - ap.z * aq.y, ny = ap.z * aq.x - ap.x * aq.z, nz = ap.x * aq.y - ap.y * aq.x;
  const float nn = nx * nx + ny * ny + nz * nz;
  if (nn > 1.0f / 256.0f) lb = fmaxf(lb, fabsf(dx * nx + dy * ny + dz * nz) * rsqrtf(nn));
  if (lb > lim) return false;
  // Fourth axis: the direction between approximate closest points of the two
  // chunk axes (clamp s, solve t, clamp t, re-solve s).  Any unit axis gives
  // a valid bound, so the FP32 closest points need no accuracy of their own;
  // near the true closest points the bound is the segment distance itself,
  // which rejects the many pairs of long rods whose chunk axes pass each
  // other farther apart than the three fixed axes can show.  The margin
  // above covers the axis normalisation (a few ulps of rsqrt).
  const float b = ap.x * aq.x + ap.y * aq.y + ap.z * aq.z;
  const float f = dx * ap.x + dy * ap.y + dz * ap.z, g = dx * aq.x + dy * aq.y + dz * aq.z;
  const float den = 1.0f - b * b;
  float s = den > 1e-6f ? (f - b * g) * __builtin_amdgcn_rcpf(den) : 0.0f;
  s = fminf(fmaxf(s, -hp), hp);
  const float t = fminf(fmaxf(s * b - g, -hq), hq);
  s = fminf(fmaxf(f + t * b, -hp), hp);
  const float vx = dx + t * aq.x - s * ap.x, vy = dy + t * aq.y - s * ap.y, vz = dz + t * aq.z - s * ap.z;
  const float vv = vx * vx + vy * vy + vz * vz;
  if (!(vv > lim * lim)) return true;  // the closest points may touch: no axis to add
  const float iv = rsqrtf(vv);
  const float mx = vx * iv, my = vy * iv, mz = vz * iv;
  const float gap = fabsf(dx * mx + dy * my + dz * mz) - hp * fabsf(ap.x * mx + ap.y * my + ap.z * mz) -
                    hq * fabsf(aq.x * mx + aq.y * my + aq.z * mz);
  return gap <= lim;
}

// The system constants of one chunk lane, shared across the wavefront: every
// queued entry (p, q) was found by lane p - p0 of the same wavefront, so its
// box and margin are read from that lane.
struct LaneSys {
  float span[3], ispan[3], extra;
  int per;  // bit k: axis k periodic
};

// Wavefront-wide barrier around the LDS queue (a wavefront executes in
// lockstep; the fences order its LDS writes before other lanes' reads).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Runs the separating-axis test on n <= 64 queued sphere survivors (one per
// lane, entries src[0..n)) and appends the kept pairs to the narrow queue;
// whenever the narrow queue holds 64, all lanes run the FP64 narrow phase on
// one entry each.  Called by the whole wavefront (n and pend2 are uniform).
__device__ __forceinline__ void sat_stage(const int2* __restrict__ src, int n, int p0, const LaneSys& own,
                                          int2* __restrict__ Q2, int& pend2, const RodRecs& R,
                                          const SysDev* __restrict__ sys, const SortedChunks& SC, double delta,
                                          uint64_t* __restrict__ out, unsigned cap, unsigned* __restrict__ counter,
                                          unsigned long long& nnarrow) {
  const int lane = __lane_id();
  int2 e = make_int2(0, 0);
  if (lane < n) e = src[lane];
  const int from = lane < n ? e.x - p0 : lane;
  float span[3], ispan[3];
  for (int k = 0; k < 3; ++k) {
    span[k] = __shfl(own.span[k], from);
    ispan[k] = __shfl(own.ispan[k], from);
  }
  const float extra = __shfl(own.extra, from);
  const int per = __shfl(own.per, from);
  const bool keep = lane < n && sat_keep(SC, e.x, e.y, span, ispan, per, extra);
  const unsigned long long m = __ballot(keep);
  if (keep) {
    ++nnarrow;
    Q2[pend2 + __popcll(m & ((1ull << lane) - 1ull))] = e;
  }
  pend2 += __popcll(m);
  if (pend2 >= 64) {
    wave_sync();
    const int2 f = Q2[pend2 - 64 + lane];
    narrow_pair(R, sys, SC, f.x, f.y, delta, out, cap, counter);
    pend2 -= 64;
    wave_sync();
  }
}

constexpr int BN_UNROLL = 6;  // candidates loaded per lane per step (independent loads in flight)
static_assert(BN_QUEUE >= 64 + 64 * BN_UNROLL, "sphere-survivor queue: < 64 pending plus one append round");

// LDS staging of the stencil rows (STAGED kernels).  A workgroup's chunks are
// a contiguous run of sorted positions, i.e. consecutive cells of at most a
// few cell rows (row = one (system, y, z)); BN_GROUPS of those rows (groups)
// are staged.  The walk runs in three phases, one per z offset of the
// stencil (ranges 6 dz .. 6 dz + 5); before each, the workgroup copies the
// chunk records (FP32 centre/radius and rod) of the three rows (y - 1, y,
// y + 1) at that z offset, over the group's cells x0 - 1 .. x1 + 1, into LDS,
// and every lane whose range lies inside a staged entry reads its candidates
// from there.  A range outside every entry (a periodic wrap in x, a group
// beyond BN_GROUPS, an entry that did not fit in BN_CAP(NT) records) is read
// from global memory as in the plain kernel, so the candidate set and the
// contacts are the same either way.
constexpr int BN_GROUPS = 4;
template <int NT>
constexpr int bn_cap() {
  return 5 * NT;
}
// Each lane lists its non-empty-looking stencil ranges once, before the walk:
// their bounds are loaded together (one round trip instead of one per range
// opened during the walk) and the walk takes them from LDS.  A lane with more
// than BN_SLOTS ranges (periodic wraps) opens the rest as it goes.
constexpr int BN_SLOTS = 8;

// Broad + narrow phase, one thread per sorted chunk p.  Every unordered chunk
// pair (p, q) is examined once, by the thread with the lower sorted position.
// The stencil walk is latency-bound (a lane's candidates are loaded from L2
// one after another), so each step a lane loads the next BN_UNROLL candidates
// of its current stencil range at once and then runs the FP32 sphere
// prefilter on them (chunk-centre minimum image + rounding margin, never
// farther than the rod-centre image the narrow phase uses, so it never
// rejects a touching pair).  Survivors are queued per wavefront in LDS; each
// time 64 are queued the 64 lanes run the separating-axis test on one entry
// each and queue what passes for the narrow phase; each time that queue holds
// 64, all lanes run the FP64 narrow phase on one entry each.  The expensive
// stages thus run without divergence, and nothing is appended to global
// memory but the contacts themselves.
// NT = workgroup size (64 ... 512; rdk_set_tuning(RDK_TUNE_BROAD_BLOCK)): the
// queues are per wavefront, so NT only changes how many wavefronts share a
// workgroup slot and the granularity of the XCD-contiguous block order; the
// contact set does not depend on it, nor on STAGED
// (rdk_set_tuning(RDK_TUNE_BROAD_STAGE)).
template <int NT, bool STAGED>
__global__ void __launch_bounds__(NT, 4) k_broad_narrow(BodyPtrs B, const SysDev* __restrict__ sys,
                                                        SortedChunks SC, const int* __restrict__ cstart, RodRecs R,
                                                        int M, double delta,
                                                        uint64_t* __restrict__ out, unsigned cap,
                                                        unsigned* __restrict__ counter,
                                                        unsigned long long* __restrict__ cand_counter) {
  constexpr int CAP = STAGED ? bn_cap<NT>() : 1;
  constexpr int NE = 3 * BN_GROUPS;  // staged entries per phase: (group, dy)
  __shared__ int2 queue[NT / 64][BN_QUEUE];
  __shared__ int2 queue2[NT / 64][BN_QUEUE2];
  __shared__ float4 st_cf[CAP];
  __shared__ int st_rod[CAP];
  __shared__ int st_lo[NE], st_hi[NE], st_off[NE];
  __shared__ int g_x0[BN_GROUPS], g_x1[BN_GROUPS], g_cy[BN_GROUPS], g_cz[BN_GROUPS], g_rod[BN_GROUPS];
  __shared__ int s_rowkey[STAGED ? NT : 1];
  __shared__ int s_wcnt[NT / 64];
  __shared__ int s_ng, s_tot;
  __shared__ int rl_q[BN_SLOTS][NT], rl_q1[BN_SLOTS][NT];
  __shared__ unsigned char rl_r[BN_SLOTS][NT];
  int2* Q = queue[threadIdx.x >> 6];
  int2* Q2 = queue2[threadIdx.x >> 6];
  const int lane = __lane_id();
  const int tid = threadIdx.x;
  const int p = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  unsigned long long ncand = 0, nnarrow = 0;
  StencilWalk W;
  W.p = p;
  W.r = 0;
  W.q = W.q1 = 0;
  const bool valid = p < M;
  int a = 0, cx = 0;
  float4 me = make_float4(0.f, 0.f, 0.f, 0.f);
  bool a_static = false;
  LaneSys own;
  own.per = 0;
  own.extra = 0.f;
  for (int k = 0; k < 3; ++k) own.span[k] = own.ispan[k] = 0.f;
  if (valid) {
    a = SC.rod[p];
    me = SC.cf[p];
    const SysDev& S = sys[B.sys[a]];
    a_static = !(B.inv_mass[a] > 0.0);
    for (int k = 0; k < 3; ++k) {  // closed axes 0: the walk's image shift is then dx - 0 * rint(0) = dx
      own.span[k] = S.periodic[k] ? (float)S.span[k] : 0.0f;
      own.ispan[k] = S.periodic[k] ? (float)(1.0 / S.span[k]) : 0.0f;
      own.per |= S.periodic[k] ? (1 << k) : 0;
    }
    own.extra = (float)delta + S.fmargin;
    const int d0 = S.dims[0], d1 = S.dims[1];
    W.d0 = d0;
    W.d1 = d1;
    W.d2 = S.dims[2];
    W.off = S.cell_off;
    W.ytile = S.ytile;
    W.py = S.periodic[1] != 0;
    W.pz = S.periodic[2] != 0;
    const int cell = (int)SC.key[p] - S.cell_off;
    W.kp = (int)SC.key[p];
    cx = cell % d0;
    row_yz(d1, W.d2, S.ytile, cell / d0, W.cy, W.cz);
    W.two = false;
    W.xb = 0;
    if (d0 == 1) {
      W.xa0 = W.xa1 = 0;
    } else if (!S.periodic[0]) {
      W.xa0 = cx > 0 ? cx - 1 : 0;
      W.xa1 = cx < d0 - 1 ? cx + 1 : d0 - 1;
    } else if (cx == 0) {
      W.xa0 = 0;
      W.xa1 = 1;
      W.xb = d0 - 1;
      W.two = true;
    } else if (cx == d0 - 1) {
      W.xa0 = d0 - 2;
      W.xa1 = d0 - 1;
      W.xb = 0;
      W.two = true;
    } else {
      W.xa0 = cx - 1;
      W.xa1 = cx + 1;
    }
  }
  // ---- the lane's range list: cell intervals, then their bounds
  int nl = 0;
  W.r = 18;  // where open_next resumes after the list (18: nothing left)
  if (valid) {
    for (int rr = 0; rr < 18; ++rr) {
      int c0, c1;
      if (!W.interval(rr, c0, c1)) continue;
      if (nl == BN_SLOTS) {
        W.r = rr;
        break;
      }
      rl_q[nl][tid] = c0;
      rl_q1[nl][tid] = c1;
      rl_r[nl][tid] = (unsigned char)rr;
      ++nl;
    }
    int bq[BN_SLOTS], bq1[BN_SLOTS];
#pragma unroll
    for (int k = 0; k < BN_SLOTS; ++k) {
      if (k < nl) {
        bq[k] = cstart[rl_q[k][tid]];
        bq1[k] = cstart[rl_q1[k][tid] + 1];
      }
    }
#pragma unroll
    for (int k = 0; k < BN_SLOTS; ++k) {
      if (k < nl) {
        rl_q[k][tid] = bq[k] <= p ? p + 1 : bq[k];
        rl_q1[k][tid] = bq1[k];
      }
    }
  }
  // ---- row groups of the workgroup (STAGED): group = run of lanes with one
  // row key (key - x, unique per system row); lanes are in sorted order
  int grp = BN_GROUPS;
  if constexpr (STAGED) {
    const int rowkey = valid ? (int)SC.key[p] - cx : -1;
    s_rowkey[tid] = rowkey;
    __syncthreads();
    const bool start = valid && (tid == 0 || s_rowkey[tid - 1] != rowkey);
    const bool last = valid && (tid == NT - 1 || s_rowkey[tid + 1] != rowkey);
    const unsigned long long bal = __ballot(start);
    if (lane == 0) s_wcnt[tid >> 6] = __popcll(bal);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < NT / 64; ++w) {
      before += w < (tid >> 6) ? s_wcnt[w] : 0;
      total += s_wcnt[w];
    }
    const int g = before + __popcll(bal & ((2ull << lane) - 1ull)) - 1;  // starts at or before this lane, - 1
    if (valid && g < BN_GROUPS) {
      grp = g;
      if (start) {
        g_x0[g] = cx;
        g_cy[g] = W.cy;
        g_cz[g] = W.cz;
        g_rod[g] = a;
      }
      if (last) g_x1[g] = cx;
    }
    if (tid == 0) s_ng = total < BN_GROUPS ? total : BN_GROUPS;
  }
  const int p0 = p - lane;  // the wavefront's lanes hold chunks p0 .. p0 + 63
  const float me_ext = me.w + own.extra;  // the lane's sphere radius plus margin (sphere threshold less q's radius)
  int pending = 0, pend2 = 0;  // wave-uniform queue lengths
  int kl = 0;                  // next entry of the lane's range list
  constexpr int NPH = STAGED ? 3 : 1;
  for (int ph = 0; ph < NPH; ++ph) {
    const int r_end = STAGED ? 6 * (ph + 1) : 18;
    if constexpr (STAGED) {
      // ---- stage the rows of z offset dz = ph - 1 (entries (group, dy))
      __syncthreads();  // group data written; the previous phase's LDS reads are done
      const int ng = s_ng;
      if (tid < NE) {
        int lo = 0, hi = 0;
        const int g = tid / 3, dy = tid % 3 - 1, dz = ph - 1;
        if (g < ng) {
          const SysDev& S = sys[B.sys[g_rod[g]]];
          const int d0 = S.dims[0], d1 = S.dims[1], d2 = S.dims[2];
          int z = g_cz[g] + dz, y = g_cy[g] + dy;
          bool ok = true;
          if (S.periodic[2]) {
            if (d2 == 1 && dz != 0) ok = false;
            z += z < 0 ? d2 : (z >= d2 ? -d2 : 0);
          } else if (z < 0 || z >= d2) {
            ok = false;
          }
          if (S.periodic[1]) {
            if (d1 == 1 && dy != 0) ok = false;
            y += y < 0 ? d1 : (y >= d1 ? -d1 : 0);
          } else if (y < 0 || y >= d1) {
            ok = false;
          }
          if (ok) {
            const int xa = g_x0[g] > 0 ? g_x0[g] - 1 : 0;
            const int xb = g_x1[g] < d0 - 1 ? g_x1[g] + 1 : d0 - 1;
            const int base = S.cell_off + d0 * cell_row(d1, d2, S.ytile, y, z);
            // no lane of the group reads a row wholly below its first cell
            if (base + xb >= S.cell_off + g_x0[g] + d0 * cell_row(d1, d2, S.ytile, g_cy[g], g_cz[g])) {
              lo = cstart[base + xa];
              hi = cstart[base + xb + 1];
            }
          }
        }
        st_lo[tid] = lo;
        st_hi[tid] = hi;
      }
      __syncthreads();
      if (tid == 0) {
        int tot = 0;
        for (int t = 0; t < NE; ++t) {
          int n = st_hi[t] - st_lo[t];
          if (tot + n > CAP) {  // does not fit: read from global memory
            st_hi[t] = st_lo[t];
            n = 0;
          }
          st_off[t] = tot;
          tot += n;
        }
        s_tot = tot;
      }
      __syncthreads();
      const int tot = s_tot;
      for (int e = tid; e < tot; e += NT) {
        int t = 0;
        while (e >= st_off[t] + (st_hi[t] - st_lo[t])) ++t;
        const int q = st_lo[t] + (e - st_off[t]);
        st_cf[e] = SC.cf[q];
        st_rod[e] = SC.rod[q];
      }
      __syncthreads();
    }
    // the lane's current range [W.q, W.q1) and where its records are read:
    // LDS at lbase + q when a staged entry holds the whole range, else global
    int lbase = 0;
    bool lds = false;
    auto open = [&]() -> bool {
      int rr = -1;
      while (kl < nl) {
        const int r1 = rl_r[kl][tid];
        if (r1 >= r_end) break;  // a later phase's range
        W.q = rl_q[kl][tid];
        W.q1 = rl_q1[kl][tid];
        ++kl;
        if (W.q < W.q1) {
          rr = r1;
          break;
        }
      }
      if (rr < 0) {
        if (kl < nl || !W.open_next(cstart, r_end)) return false;
        rr = W.r - 1;
      }
      if constexpr (STAGED) {
        lds = false;
        if (grp < BN_GROUPS) {
          const int t = grp * 3 + (rr >> 1) % 3;  // entry (group, dy) of the range's row
          lds = W.q >= st_lo[t] && W.q1 <= st_hi[t];
          lbase = st_off[t] - st_lo[t];
        }
      }
      return true;
    };
    bool active = valid && open();
    while (__any(active)) {
      float4 o[BN_UNROLL];
      int rb[BN_UNROLL];
      const int qs = W.q;
      const int nh = active ? min(BN_UNROLL, W.q1 - W.q) : 0;
      // every lane loads BN_UNROLL records, no branch around the loads: past
      // the range's end it repeats its last candidate, an idle lane reads
      // record 0; the j < nh test below drops those
#pragma unroll
      for (int j = 0; j < BN_UNROLL; ++j) {
        const int q = nh > 0 ? qs + min(j, nh - 1) : 0;
        if (STAGED && lds && nh > 0) {
          o[j] = st_cf[lbase + q];
          rb[j] = st_rod[lbase + q];
        } else {
          o[j] = SC.cf[q];
          rb[j] = SC.rod[q];
        }
      }
      if (active) {
        W.q += nh;
        if (W.q >= W.q1) active = open();
      }
      bool hit[BN_UNROLL];
#pragma unroll
      for (int j = 0; j < BN_UNROLL; ++j) {
        // fused image shift and squared distance: one rounding each instead
        // of two, so the test stays inside the margin it was sized for
        float dx = o[j].x - me.x, dy = o[j].y - me.y, dz = o[j].z - me.z;
        dx = fmaf(-own.span[0], rintf(dx * own.ispan[0]), dx);
        dy = fmaf(-own.span[1], rintf(dy * own.ispan[1]), dy);
        dz = fmaf(-own.span[2], rintf(dz * own.ispan[2]), dz);
        const float thr = me_ext + o[j].w;
        hit[j] = j < nh && fmaf(dz, dz, fmaf(dy, dy, dx * dx)) < thr * thr && rb[j] != a;
      }
      if (a_static) {  // a fixed rod pairs only with movable ones (rare: the load only then)
#pragma unroll
        for (int j = 0; j < BN_UNROLL; ++j)
          if (hit[j]) hit[j] = B.inv_mass[rb[j]] > 0.0;
      }
      // one append round for the step's candidates
      const unsigned long long lt = (1ull << lane) - 1ull;
      int off = pending;
#pragma unroll
      for (int j = 0; j < BN_UNROLL; ++j) {
        const unsigned long long m = __ballot(hit[j]);
        if (hit[j]) {
          ++ncand;
          Q[off + __popcll(m & lt)] = make_int2(p, qs + j);
        }
        off += __popcll(m);
      }
      pending = off;
      while (pending >= 64) {
        wave_sync();
        sat_stage(Q + pending - 64, 64, p0, own, Q2, pend2, R, sys, SC, delta, out, cap, counter,
                  nnarrow);
        pending -= 64;
        wave_sync();
      }
    }
  }
  wave_sync();
  if (pending > 0)
    sat_stage(Q, pending, p0, own, Q2, pend2, R, sys, SC, delta, out, cap, counter, nnarrow);
  wave_sync();
  if (lane < pend2) {
    const int2 e = Q2[lane];
    narrow_pair(R, sys, SC, e.x, e.y, delta, out, cap, counter);
  }
  // wave-level reduction of the sphere-survivor and narrow-phase counts
  for (int off = 32; off > 0; off >>= 1) {
    ncand += __shfl_down(ncand, off);
    nnarrow += __shfl_down(nnarrow, off);
  }
  if (lane == 0 && ncand) atomicAdd(cand_counter, ncand);
  if (lane == 0 && nnarrow) atomicAdd(cand_counter + 1, nnarrow);
}

struct ContactPtrs {
  int *a, *b, *type;
  double *pa[3], *pb[3], *n[3], *sh[3], *dist, *hs, *s, *t;
};

__global__ void k_contact_record(RodRecs R, const SysDev* __restrict__ sys, const uint64_t* __restrict__ keys,
                                 Count ncc, double delta, ContactPtrs C) {
  const int nc = ncc.get();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const uint64_t key = keys[i];
  const int a = (int)(key >> 32), b = (int)(key & 0xffffffffu);
  const RodIn ra = load_rod(R, a), rb = load_rod(R, b);
  const SysDev S = sys[R.info[a].x];
  NarrowOut o;
  o.type = 0;
  o.dist = 0.0;
  o.pa = o.pb = o.n = mk3(0, 0, 0);
  capsule_narrow(ra.x, ra.u, ra.h, ra.r, rb.x, rb.u, rb.h, rb.r, S.span, S.periodic, delta, true, o);
  C.a[i] = a;
  C.b[i] = b;
  C.type[i] = o.type;
  C.pa[0][i] = o.pa.x;
  C.pa[1][i] = o.pa.y;
  C.pa[2][i] = o.pa.z;
  C.pb[0][i] = o.pb.x;
  C.pb[1][i] = o.pb.y;
  C.pb[2][i] = o.pb.z;
  C.n[0][i] = o.n.x;
  C.n[1][i] = o.n.y;
  C.n[2][i] = o.n.z;
  C.sh[0][i] = o.shift.x;
  C.sh[1][i] = o.shift.y;
  C.sh[2][i] = o.shift.z;
  C.dist[i] = o.dist;
  C.hs[i] = ra.r + rb.r;
  C.s[i] = o.s;
  C.t[i] = o.t;
}

// ---------------------------------------------------------------------------
// host orchestration

// Per-system cell grids for the current boxes: cell >= 2 * max chunk radius
// + delta (times cell_scale), periodic axes with fewer than 3 cells use one.
static int size_grids(rdk_ctx* ctx, double delta, double scale) {
  int cell_off = 0;
  for (int s = 0; s < ctx->nsys; ++s) {
    SysDev& S = ctx->sys_host[s];
    const double radmax = ctx->sys_radmax[s];
    const double cmin = (2.0 * radmax + delta) * scale * (1.0 + 1e-9) + 1e-9;
    int64_t total = 1;
    for (int k = 0; k < 3; ++k) {
      int d = (int)std::floor(S.span[k] / cmin);
      if (d < 1) d = 1;
      if (S.periodic[k] && d < 3) d = 1;
      if (d > 1024) d = 1024;
      S.dims[k] = d;
      total *= d;
    }
    // keep the grid modest for very dense or tiny-cell configurations
    while (total > (int64_t)(1 << 22)) {
      total = 1;
      for (int k = 0; k < 3; ++k) {
        int d = std::max(1, S.dims[k] / 2);
        if (S.periodic[k] && d < 3) d = 1;
        S.dims[k] = d;
        total *= d;
      }
    }
    double amax = 0.0;
    for (int k = 0; k < 3; ++k) {
      S.cell[k] = S.span[k] / S.dims[k];
      S.inv_cell[k] = 1.0 / S.cell[k];
      amax = std::max(amax, std::max(std::fabs(S.lo[k]), std::fabs(S.hi[k])) + 2.0 * S.span[k]);
    }
    // FP32 prefilter margin: rounding of centres, radii, the image shift and
    // the squared distance, bounded by a few ulps of the largest coordinate
    S.fmargin = (float)(64.0 * amax * 1.1920928955078125e-07 + 1e-6 * (2.0 * radmax + delta) + 1e-6);
    // rows per band of the cell order: three z-planes of a band (and the row
    // either side) within about 2 MB of chunk records (56 B each: centre and
    // radius, axis and half length, rod, chunk index), at least 4 rows
    const double nck = s < (int)ctx->sys_nchunk.size() ? (double)ctx->sys_nchunk[s] : 0.0;
    const double row_bytes = 56.0 * S.dims[0] * nck / (double)total;
    const double rows = row_bytes > 0.0 ? 2.0 * 1024.0 * 1024.0 / (3.0 * row_bytes) - 2.0 : 1e9;
    S.ytile = rows < (double)S.dims[1] ? std::min(S.dims[1], std::max(4, (int)rows)) : S.dims[1];
    if (ctx->cell_band > 0) S.ytile = std::min(S.dims[1], ctx->cell_band);
    S.cell_off = cell_off;
    cell_off += (int)total;
  }
  ctx->n_cells = cell_off;
  RDK_TRY(ensure(ctx, ctx->sys_dev, ctx->nsys));
  RDK_HIP(hipMemcpyAsync(ctx->sys_dev.p, ctx->sys_host.data(), ctx->nsys * sizeof(SysDev), hipMemcpyHostToDevice,
                         ctx->stream));
  RDK_TRY(ensure(ctx, ctx->cell_start, (size_t)ctx->n_cells + 1));
  ctx->grid_dirty = false;
  return RDK_OK;
}

int regrid(rdk_ctx* ctx) {
  if (ctx->n_chunks <= 0) return RDK_OK;  // no grid yet: the next detection sizes it
  return size_grids(ctx, ctx->built_delta, ctx->built_cell_scale);
}

static int build_grid(rdk_ctx* ctx, const rdk_detect_params* p) {
  const double cl = p->chunk_len;
  const double scale = p->cell_scale >= 1.0 ? p->cell_scale : 1.0;
  const bool chunks_ok = cl == ctx->built_chunk_len && ctx->n_chunks > 0;
  if (chunks_ok && scale == ctx->built_cell_scale && p->delta == ctx->built_delta && !ctx->grid_dirty)
    return RDK_OK;
  if (!chunks_ok) {
    const int n = ctx->n;
    std::vector<int> nchunk(n), crod, ck;
    crod.reserve(n);
    ck.reserve(n);
    ctx->sys_radmax.assign(ctx->nsys, 0.0);
    ctx->sys_nchunk.assign(ctx->nsys, 0);
    for (int i = 0; i < n; ++i) {
      int K = 1;
      if (cl > 0.0) K = std::max(1, (int)std::ceil(2.0 * ctx->h_h[i] / cl));
      nchunk[i] = K;
      for (int k = 0; k < K; ++k) {
        crod.push_back(i);
        ck.push_back(k);
      }
      const double rad = ctx->h_h[i] / K + ctx->h_r[i];
      const int s = ctx->h_sys_of[i];
      ctx->sys_radmax[s] = std::max(ctx->sys_radmax[s], rad);
      ctx->sys_nchunk[s] += K;
    }
    const int M = (int)crod.size();
    ctx->n_chunks = M;
    RDK_TRY(ensure(ctx, ctx->chunk_rod, M));
    RDK_TRY(ensure(ctx, ctx->chunk_k, M));
    RDK_TRY(ensure(ctx, ctx->rod_nchunk, n));
    RDK_TRY(ensure(ctx, ctx->rodrec, 8 * (size_t)n));
    RDK_TRY(ensure(ctx, ctx->rodinfo, 2 * (size_t)n));
    RDK_HIP(hipMemcpyAsync(ctx->chunk_rod.p, crod.data(), M * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    RDK_HIP(hipMemcpyAsync(ctx->chunk_k.p, ck.data(), M * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    RDK_HIP(hipMemcpyAsync(ctx->rod_nchunk.p, nchunk.data(), n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    RDK_TRY(ensure(ctx, ctx->ckey, M));
    RDK_TRY(ensure(ctx, ctx->ckey_alt, M));
    RDK_TRY(ensure(ctx, ctx->cid, M));
    RDK_TRY(ensure(ctx, ctx->cid_alt, M));
    RDK_TRY(ensure(ctx, ctx->scf, M));
    RDK_TRY(ensure(ctx, ctx->sca, M));
    RDK_TRY(ensure(ctx, ctx->ucf, M));
    RDK_TRY(ensure(ctx, ctx->uca, M));
    RDK_TRY(ensure(ctx, ctx->srod, M));
    RDK_TRY(ensure(ctx, ctx->sk, M));
  }
  RDK_TRY(size_grids(ctx, p->delta, scale));
  RDK_HIP(hipStreamSynchronize(ctx->stream));  // host chunk tables are released on return
  ctx->built_chunk_len = cl;
  ctx->built_cell_scale = scale;
  ctx->built_delta = p->delta;
  return RDK_OK;
}

static int ensure_contacts(rdk_ctx* ctx, int nc) {
  RDK_TRY(ensure(ctx, ctx->ca, nc));
  RDK_TRY(ensure(ctx, ctx->cb, nc));
  RDK_TRY(ensure(ctx, ctx->ctype, nc));
  for (int k = 0; k < 3; ++k) {
    RDK_TRY(ensure(ctx, ctx->cpa[k], nc));
    RDK_TRY(ensure(ctx, ctx->cpb[k], nc));
    RDK_TRY(ensure(ctx, ctx->cn[k], nc));
    RDK_TRY(ensure(ctx, ctx->csh[k], nc));
  }
  RDK_TRY(ensure(ctx, ctx->cdist, nc));
  RDK_TRY(ensure(ctx, ctx->chs, nc));
  RDK_TRY(ensure(ctx, ctx->cs, nc));
  RDK_TRY(ensure(ctx, ctx->ct, nc));
  return RDK_OK;
}

// The broad phase with the context's workgroup size.
static int launch_broad_narrow(rdk_ctx* ctx, const BodyPtrs& B, const SortedChunks& SC, int M, double delta,
                               unsigned cap) {
  const int nt = ctx->bn_threads;
  const bool stg = ctx->bn_stage != 0;
  auto go = [&](auto kern, int t) {
    hipLaunchKernelGGL(kern, dim3(grid_for(M, t)), dim3(t), 0, ctx->stream, B, ctx->sys_dev.p, SC, ctx->cell_start.p,
                       rod_recs(ctx), M, delta, ctx->pkeys.p, cap, ctx->counters.p,
                       (unsigned long long*)(ctx->counters.p + 2));
  };
  switch (nt) {
    case 64: stg ? go(k_broad_narrow<64, true>, 64) : go(k_broad_narrow<64, false>, 64); break;
    case 128: stg ? go(k_broad_narrow<128, true>, 128) : go(k_broad_narrow<128, false>, 128); break;
    case 256: stg ? go(k_broad_narrow<256, true>, 256) : go(k_broad_narrow<256, false>, 256); break;
    case 512: stg ? go(k_broad_narrow<512, true>, 512) : go(k_broad_narrow<512, false>, 512); break;
    default:
      ctx->err = "broad phase: workgroup size must be 64, 128, 256 or 512";
      return RDK_ERR_ARG;
  }
  return RDK_OK;
}

static ContactPtrs contact_ptrs(rdk_ctx* ctx) {
  ContactPtrs C;
  C.a = ctx->ca.p;
  C.b = ctx->cb.p;
  C.type = ctx->ctype.p;
  for (int k = 0; k < 3; ++k) {
    C.pa[k] = ctx->cpa[k].p;
    C.pb[k] = ctx->cpb[k].p;
    C.n[k] = ctx->cn[k].p;
    C.sh[k] = ctx->csh[k].p;
  }
  C.dist = ctx->cdist.p;
  C.hs = ctx->chs.p;
  C.s = ctx->cs.p;
  C.t = ctx->ct.p;
  return C;
}

// Reserve everything the fixed detection path touches (called before capture).
int detect_fixed_reserve(rdk_ctx* ctx, const rdk_detect_params* p, int pair_cap) {
  RDK_TRY(build_grid(ctx, p));
  RDK_TRY(ensure(ctx, ctx->counters, 8));
  if (!ctx->h_counters) RDK_HIP(hipHostMalloc((void**)&ctx->h_counters, 64, hipHostMallocDefault));
  RDK_TRY(ensure(ctx, ctx->pkeys, pair_cap));
  RDK_TRY(ensure(ctx, ctx->pkeys_alt, pair_cap));
  RDK_TRY(ensure_contacts(ctx, std::max(pair_cap, 1)));
  return RDK_OK;
}

// Detection of the fixed-capacity step: the launches of detect_run with the
// contact count left on the device (counters[0], capped at the reserved
// capacity).  The grid must be current (detect_fixed_reserve).
int detect_fixed_enqueue(rdk_ctx* ctx, const rdk_detect_params* p) {
  const int M = ctx->n_chunks, cap = ctx->fx_pair_cap;
  const int nb = 256;
  if (M <= 0 || cap <= 0) {
    ctx->err = "fixed step: detection not reserved";
    return RDK_ERR_STATE;
  }
  BodyPtrs B = body_ptrs(ctx);
  hipLaunchKernelGGL(k_chunk_keys, dim3(grid_for(M, nb)), dim3(nb), 0, ctx->stream, B, ctx->sys_dev.p,
                     ctx->chunk_rod.p, ctx->chunk_k.p, ctx->rod_nchunk.p, M, ctx->ckey.p, ctx->cid.p, ctx->ucf.p,
                     ctx->uca.p, rod_recs(ctx));
  bool alt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->ckey.p, ctx->ckey_alt.p, ctx->cid.p, ctx->cid_alt.p, Count::host(M),
                         bits_for((uint64_t)std::max(1, ctx->n_cells - 1)), alt, nullptr));
  const uint32_t* skey = alt ? ctx->ckey_alt.p : ctx->ckey.p;
  const int* sid = alt ? ctx->cid_alt.p : ctx->cid.p;
  hipLaunchKernelGGL(k_cell_start, dim3(grid_for((int64_t)ctx->n_cells + 1, nb)), dim3(nb), 0, ctx->stream, skey, M,
                     ctx->n_cells, ctx->cell_start.p);
  SortedChunks SC;
  SC.key = skey;
  SC.rod = ctx->srod.p;
  SC.k = ctx->sk.p;
  SC.cf = ctx->scf.p;
  SC.ca = ctx->sca.p;
  hipLaunchKernelGGL(k_chunk_gather, dim3(grid_for(M, nb)), dim3(nb), 0, ctx->stream, sid, ctx->chunk_rod.p,
                     ctx->chunk_k.p, ctx->ucf.p, ctx->uca.p, M, SC);
  // counters[] were cleared by fixed_clear_enqueue; more contacts than cap fail
  // the step in k_fx_counts
  RDK_TRY(launch_broad_narrow(ctx, B, SC, M, p->delta, (unsigned)cap));
  const Count nc = Count::dev((const int*)ctx->counters.p, cap);
  bool kalt = false;
  RDK_TRY(sort_pair_keys_u64(ctx, ctx->pkeys.p, ctx->pkeys_alt.p, nc, bits_for((uint64_t)ctx->n), kalt, nullptr));
  const uint64_t* keys = kalt ? ctx->pkeys_alt.p : ctx->pkeys.p;
  hipLaunchKernelGGL(k_contact_record, dim3(grid_for(cap, nb)), dim3(nb), 0, ctx->stream, rod_recs(ctx),
                     ctx->sys_dev.p, keys, nc, p->delta, contact_ptrs(ctx));
  if (kalt) {  // the sorted keys are the canonical ones (warm cache, history)
    std::swap(ctx->pkeys.p, ctx->pkeys_alt.p);
    std::swap(ctx->pkeys.cap, ctx->pkeys_alt.cap);
  }
  ctx->last_delta = p->delta;
  return RDK_OK;
}

int detect_run(rdk_ctx* ctx, const rdk_detect_params* p) {
  if (ctx->n <= 0 || ctx->nsys <= 0) {
    ctx->err = "rdk_detect: bodies and systems must be set first";
    return RDK_ERR_STATE;
  }
  RDK_TRY(build_grid(ctx, p));
  const int M = ctx->n_chunks;
  const int nb = 256;
  BodyPtrs B = body_ptrs(ctx);
  {
    ProfScope ps(ctx, K_CHUNK_KEYS);
    hipLaunchKernelGGL(k_chunk_keys, dim3(grid_for(M, nb)), dim3(nb), 0, ctx->stream, B, ctx->sys_dev.p,
                       ctx->chunk_rod.p, ctx->chunk_k.p, ctx->rod_nchunk.p, M, ctx->ckey.p, ctx->cid.p, ctx->ucf.p,
                       ctx->uca.p, rod_recs(ctx));
  }
  bool alt = false;
  RDK_TRY(sort_pairs_u32(ctx, ctx->ckey.p, ctx->ckey_alt.p, ctx->cid.p, ctx->cid_alt.p, M,
                         bits_for((uint64_t)std::max(1, ctx->n_cells - 1)), alt));
  const uint32_t* skey = alt ? ctx->ckey_alt.p : ctx->ckey.p;
  const int* sid = alt ? ctx->cid_alt.p : ctx->cid.p;
  {
    ProfScope ps(ctx, K_CELL_RANGES);
    hipLaunchKernelGGL(k_cell_start, dim3(grid_for((int64_t)ctx->n_cells + 1, nb)), dim3(nb), 0, ctx->stream, skey,
                       M, ctx->n_cells, ctx->cell_start.p);
  }
  SortedChunks SC;
  SC.key = skey;
  SC.rod = ctx->srod.p;
  SC.k = ctx->sk.p;
  SC.cf = ctx->scf.p;
  SC.ca = ctx->sca.p;
  {
    ProfScope ps(ctx, K_CHUNK_GATHER);
    hipLaunchKernelGGL(k_chunk_gather, dim3(grid_for(M, nb)), dim3(nb), 0, ctx->stream, sid, ctx->chunk_rod.p,
                       ctx->chunk_k.p, ctx->ucf.p, ctx->uca.p, M, SC);
  }
  RDK_TRY(ensure(ctx, ctx->counters, 8));
  if (!ctx->h_counters) RDK_HIP(hipHostMalloc((void**)&ctx->h_counters, 64, hipHostMallocDefault));
  unsigned cap;
  if (ctx->pair_cap_forced > 0) cap = (unsigned)ctx->pair_cap_forced;
  else cap = (unsigned)std::max<size_t>(ctx->pkeys.cap, (size_t)std::max(1024, 4 * ctx->n));
  unsigned total = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    RDK_TRY(ensure(ctx, ctx->pkeys, cap));
    RDK_TRY(ensure(ctx, ctx->pkeys_alt, cap));
    RDK_HIP(hipMemsetAsync(ctx->counters.p, 0, 8 * sizeof(unsigned), ctx->stream));
    {
      ProfScope ps(ctx, K_BROAD_NARROW);
      RDK_TRY(launch_broad_narrow(ctx, B, SC, M, p->delta, cap));
    }
    RDK_HIP(hipMemcpyAsync(ctx->h_counters, ctx->counters.p, 6 * sizeof(unsigned), hipMemcpyDeviceToHost,
                           ctx->stream));
    RDK_HIP(hipStreamSynchronize(ctx->stream));
    total = ctx->h_counters[0];
    ctx->last_candidates = (int64_t)(((uint64_t)ctx->h_counters[3] << 32) | ctx->h_counters[2]);
    ctx->last_narrow = (int64_t)(((uint64_t)ctx->h_counters[5] << 32) | ctx->h_counters[4]);
    if (total <= cap) break;
    // capacity exceeded: the kernel counted every contact; regrow and rerun
    ctx->regrows++;
    ctx->pair_cap_forced = 0;
    cap = total + total / 4 + 64;
    if (attempt == 2) {
      ctx->err = "rdk_detect: pair capacity still exceeded after regrow";
      return RDK_ERR_CAPACITY;
    }
  }
  const int nc = (int)total;
  ctx->nc = nc;
  ctx->last_delta = p->delta;
  ctx->work[W_BODY_PASSES] += 1;
  ctx->work[W_CHUNKS] += M;
  ctx->work[W_CANDIDATES] += ctx->last_candidates;
  ctx->work[W_NARROW] += ctx->last_narrow;
  ctx->work[W_CONTACTS] += nc;
  RDK_TRY(ensure_contacts(ctx, std::max(nc, 1)));
  if (nc == 0) return RDK_OK;
  bool kalt = false;
  RDK_TRY(sort_pair_keys_u64(ctx, ctx->pkeys.p, ctx->pkeys_alt.p, Count::host(nc), bits_for((uint64_t)ctx->n), kalt,
                             nullptr));
  const uint64_t* keys = kalt ? ctx->pkeys_alt.p : ctx->pkeys.p;
  ContactPtrs C;
  C.a = ctx->ca.p;
  C.b = ctx->cb.p;
  C.type = ctx->ctype.p;
  for (int k = 0; k < 3; ++k) {
    C.pa[k] = ctx->cpa[k].p;
    C.pb[k] = ctx->cpb[k].p;
    C.n[k] = ctx->cn[k].p;
    C.sh[k] = ctx->csh[k].p;
  }
  C.dist = ctx->cdist.p;
  C.hs = ctx->chs.p;
  C.s = ctx->cs.p;
  C.t = ctx->ct.p;
  {
    ProfScope ps(ctx, K_CONTACT_RECORD);
    hipLaunchKernelGGL(k_contact_record, dim3(grid_for(nc, nb)), dim3(nb), 0, ctx->stream, rod_recs(ctx),
                       ctx->sys_dev.p, keys, Count::host(nc), p->delta, C);
  }
  // keep the canonical keys for history matching
  if (kalt) {
    uint64_t* tmp = ctx->pkeys.p;
    size_t tc = ctx->pkeys.cap;
    ctx->pkeys.p = ctx->pkeys_alt.p;
    ctx->pkeys.cap = ctx->pkeys_alt.cap;
    ctx->pkeys_alt.p = tmp;
    ctx->pkeys_alt.cap = tc;
  }
  return RDK_OK;
}

}  // namespace rdk

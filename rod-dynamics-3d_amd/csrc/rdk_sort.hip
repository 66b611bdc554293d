// rdk_sort.hip -- stable LSD radix sorts and the exclusive scan used between
// the hand-written stages (cell binning, pair canonicalisation, colour
// grouping, body CSR, island grouping, history merge).
//
// Launches only (no memset, no host read-back, no host-side decision that
// depends on the data), so a whole step can be captured into a graph.  The element count can be a device value (Count):
// launches are sized for the capacity and read the count when they run.
//
// One pass over 8 key bits, three launches:
//   k_rs_hist     one workgroup per tile of RS_TILE keys: a 256-bin histogram
//                 of the tile's digits (LDS atomics), stored digit-major;
//   k_rs_scan     one workgroup per digit: exclusive scan of that digit's tile
//                 counts, and the digit's total;
//   k_rs_scatter  one workgroup per tile: digit bases (scan of the 256 totals)
//                 plus the tile's offset, then a stable rank inside the tile.
//                 Keys are taken in rounds of 256 consecutive keys; inside a
//                 wavefront the lanes holding the same digit are found with
//                 eight ballots (one per digit bit), the lowest of them
//                 publishes the group size, and a key's rank is the running
//                 count of its digit plus the groups of lower wavefronts plus
//                 its position among the lower lanes of its group.  Rank order
//                 is therefore index order: the sort is stable.
// When the tiles are very few (RS_FUSE_TILES) the scatter kernel sums the
// digit counts of the earlier tiles itself and k_rs_scan is skipped (each of
// its threads then loads one count per tile, one after another, so beyond a
// few tiles the extra scan launch is much cheaper); at most
// RS_SMALL keys are sorted by one workgroup in LDS in a single launch (every
// pass: histogram, digit bases, the same stable ranking).
// 64-bit pair keys (lo << 32 | hi, both below 2^b) are ranked on the packed
// 2b-bit value (hi bits first), which orders them like the 64-bit key with
// fewer passes.
#include <cstring>

#include "rdk_internal.hpp"

namespace rdk {

// Tiles of 1024 keys: a pass is a chain of RS_ITEMS dependent ranking rounds
// per workgroup, so small tiles (more workgroups, fewer rounds) finish sooner;
// measured 21 us per scatter launch at 4096-key tiles on 3e5 keys (73
// workgroups for 256 CUs).
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 4;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;
// measured on the 2e5 chunk keys of the settling bed (196 tiles): fused
// scatter 32 us per pass, scan + scatter 10 us
constexpr int RS_FUSE_TILES = 16;
constexpr int RS_SMALL_THREADS = 1024;
constexpr int RS_SMALL = 8192;
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__device__ __forceinline__ unsigned rs_digit(uint32_t k, int shift, int) { return (k >> shift) & 255u; }
__device__ __forceinline__ unsigned rs_digit(uint64_t k, int shift, int lo_bits) {
  const uint64_t c = lo_bits > 0 ? (((k >> 32) << lo_bits) | (k & ((uint64_t(1) << lo_bits) - 1))) : k;
  return (unsigned)(c >> shift) & 255u;
}

// Inclusive scan across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan across a workgroup of NT threads (NT / 64 <= 16); *total gets
// the sum.  ws holds NT / 64 ints.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int* ws, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) ws[wave] = inc;
  __syncthreads();
  int before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    const int s = ws[w];
    if (w < wave) before += s;
    all += s;
  }
  __syncthreads();  // ws may be reused by the caller
  *total = all;
  return before + inc - v;
}

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const K* __restrict__ keys, Count cnt, int shift, int lo_bits,
                                                        int ntiles_cap, int* __restrict__ hist) {
  __shared__ int h[256];
  const int n = cnt.get();
  const int tile = blockIdx.x;
  if ((int64_t)tile * RS_TILE >= n) return;  // columns past the used tiles are never read
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = tile * RS_TILE;
#pragma unroll 4
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int i = base + r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[rs_digit(keys[i], shift, lo_bits)], 1);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles_cap + tile] = h[threadIdx.x];
}

// blockIdx.x = digit: exclusive scan of its tile counts in place.
__global__ void __launch_bounds__(RS_THREADS) k_rs_scan(int* __restrict__ hist, Count cnt, int ntiles_cap,
                                                        int* __restrict__ total) {
  __shared__ int ws[RS_WAVES];
  const int n = cnt.get();
  const int nt = (int)(((int64_t)n + RS_TILE - 1) / RS_TILE);
  int* row = hist + (size_t)blockIdx.x * ntiles_cap;
  int carry = 0;
  for (int c0 = 0; c0 < nt; c0 += RS_THREADS) {
    const int i = c0 + threadIdx.x;
    const int v = i < nt ? row[i] : 0;
    int sum;
    const int ex = block_excl_scan<RS_THREADS>(v, ws, &sum);
    if (i < nt) row[i] = carry + ex;
    carry += sum;
  }
  if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

template <class K, bool VAL>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const K* __restrict__ kin, K* __restrict__ kout,
                                                           const int* __restrict__ vin, int* __restrict__ vout,
                                                           Count cnt, int shift, int lo_bits, int ntiles_cap,
                                                           const int* __restrict__ hist,
                                                           const int* __restrict__ total) {
  __shared__ int base[256];
  __shared__ int run[256];
  __shared__ int wc[RS_WAVES][256];
  __shared__ int ws[RS_WAVES];
  const int n = cnt.get();
  const int tile = blockIdx.x;
  if ((int64_t)tile * RS_TILE >= n) return;
  const int t = threadIdx.x;
  int all, before, tot;
  if (total) {  // k_rs_scan ran: hist holds each tile's offset within its digit
    tot = total[t];
    before = hist[t * ntiles_cap + tile];
  } else {      // few tiles: sum the earlier tiles' counts of digit t here
    const int nt = (int)(((int64_t)n + RS_TILE - 1) / RS_TILE);
    const int* row = hist + (size_t)t * ntiles_cap;
    before = 0;
    tot = 0;
    for (int k = 0; k < nt; ++k) {
      const int c = row[k];
      before += k < tile ? c : 0;
      tot += c;
    }
  }
  const int ex = block_excl_scan<RS_THREADS>(tot, ws, &all);
  base[t] = ex + before;
  run[t] = 0;
#pragma unroll
  for (int w = 0; w < RS_WAVES; ++w) wc[w][t] = 0;
  __syncthreads();
  const int lane = t & 63, wave = t >> 6;
  const uint64_t lt = (uint64_t(1) << lane) - 1;
  const int t0 = tile * RS_TILE;
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int i = t0 + r * RS_THREADS + t;
    const bool valid = i < n;
    const K k = valid ? kin[i] : K(0);
    const int v = (VAL && valid) ? vin[i] : 0;
    const unsigned d = rs_digit(k, shift, lo_bits);
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      m &= bit ? bb : ~bb;
    }
    const int rank = __popcll(m & lt);
    if (valid && rank == 0) wc[wave][d] = __popcll(m);
    __syncthreads();
    if (valid) {
      int pos = run[d] + rank;
      for (int w2 = 0; w2 < wave; ++w2) pos += wc[w2][d];
      const int o = base[d] + pos;
      kout[o] = k;
      if (VAL) vout[o] = v;
    }
    __syncthreads();
    int s = 0;
#pragma unroll
    for (int w2 = 0; w2 < RS_WAVES; ++w2) {
      s += wc[w2][t];
      wc[w2][t] = 0;
    }
    run[t] += s;
    __syncthreads();
  }
}

// One workgroup sorts at most RS_SMALL keys in LDS: per pass a histogram, the
// digit bases, then rounds of RS_SMALL_THREADS consecutive keys ranked as in
// k_rs_scatter (ballot groups inside a wavefront, lower wavefronts' groups,
// running digit counts).  The result is written to the buffer the multi-tile
// sort would leave it in (alt after an odd number of passes).
template <class K, bool VAL>
__global__ void __launch_bounds__(RS_SMALL_THREADS) k_rs_small(K* __restrict__ keys, K* __restrict__ keys_alt,
                                                               int* __restrict__ vals, int* __restrict__ vals_alt,
                                                               Count cnt, int passes, int lo_bits) {
  constexpr int NW = RS_SMALL_THREADS / 64;
  __shared__ K kb[2][RS_SMALL];
  __shared__ int vb[VAL ? 2 : 1][VAL ? RS_SMALL : 1];
  __shared__ int h[256], base[256];
  __shared__ int wc[NW][256];
  __shared__ int ws[NW];
  const int n = cnt.get();
  if (n > RS_SMALL || n <= 0) return;  // above RS_SMALL: the multi-tile sort of a captured step handles it
  const int rounds = (n + RS_SMALL_THREADS - 1) / RS_SMALL_THREADS;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint64_t lt = (uint64_t(1) << lane) - 1;
  for (int i = t; i < n; i += RS_SMALL_THREADS) {
    kb[0][i] = keys[i];
    if (VAL) vb[0][i] = vals[i];
  }
  for (int w = 0; w < NW; ++w)
    if (t < 256) wc[w][t] = 0;
  for (int p = 0; p < passes; ++p) {
    const int src = p & 1, dst = src ^ 1, shift = 8 * p;
    if (t < 256) h[t] = 0;
    __syncthreads();
    for (int i = t; i < n; i += RS_SMALL_THREADS) atomicAdd(&h[rs_digit(kb[src][i], shift, lo_bits)], 1);
    __syncthreads();
    {
      const int v = t < 256 ? h[t] : 0;
      int all;
      const int ex = block_excl_scan<RS_SMALL_THREADS>(v, ws, &all);
      if (t < 256) base[t] = ex;  // running start of each digit
    }
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
      const int i = r * RS_SMALL_THREADS + t;
      const bool valid = i < n;
      const K k = valid ? kb[src][i] : K(0);
      const int v = (VAL && valid) ? vb[src][i] : 0;
      const unsigned d = rs_digit(k, shift, lo_bits);
      uint64_t m = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
      }
      const int rank = __popcll(m & lt);
      if (valid && rank == 0) wc[wave][d] = __popcll(m);
      __syncthreads();
      if (valid) {
        int pos = base[d] + rank;
        for (int w2 = 0; w2 < wave; ++w2) pos += wc[w2][d];
        kb[dst][pos] = k;
        if (VAL) vb[dst][pos] = v;
      }
      __syncthreads();
      if (t < 256) {
        int s = 0;
        for (int w2 = 0; w2 < NW; ++w2) {
          s += wc[w2][t];
          wc[w2][t] = 0;
        }
        base[t] += s;
      }
      __syncthreads();
    }
  }
  const int fin = passes & 1;
  K* ko = fin ? keys_alt : keys;
  int* vo = fin ? vals_alt : vals;
  for (int i = t; i < n; i += RS_SMALL_THREADS) {
    ko[i] = kb[fin][i];
    if (VAL) vo[i] = vb[fin][i];
  }
}

static int hist_bytes(int cap) {
  const int ntiles = std::max(1, grid_for(cap, RS_TILE));
  return (256 * ntiles + 256) * (int)sizeof(int);
}

// Scratch for sorts and scans of up to max_count entries (reserved before a
// step is captured: a captured step allocates nothing).
int sort_reserve(rdk_ctx* ctx, int max_count) {
  RDK_TRY(ensure(ctx, ctx->temp, (size_t)hist_bytes(std::max(max_count, 1)) + 256));
  RDK_TRY(ensure(ctx, ctx->scan_tmp, (size_t)grid_for(std::max(max_count, 1), SC_TILE) + 16));
  return RDK_OK;
}

template <class K, bool VAL>
static int radix_sort(rdk_ctx* ctx, K* keys, K* keys_alt, int* vals, int* vals_alt, Count cnt, int passes,
                      int lo_bits, bool& in_alt, hipStream_t st) {
  in_alt = false;
  const int cap = cnt.cap;
  if (passes <= 0 || cap <= 1) return RDK_OK;
  in_alt = (passes & 1) != 0;
  if (cap <= RS_SMALL) {
    hipLaunchKernelGGL((k_rs_small<K, VAL>), dim3(1), dim3(RS_SMALL_THREADS), 0, st, keys, keys_alt, vals, vals_alt,
                       cnt, passes, lo_bits);
    return RDK_OK;
  }
  const int ntiles = grid_for(cap, RS_TILE);
  RDK_TRY(ensure(ctx, ctx->temp, (size_t)hist_bytes(cap) + 256));
  int* hist = (int*)ctx->temp.p;
  int* total = hist + 256 * ntiles;
  const bool fuse = ntiles <= RS_FUSE_TILES;
  K* kb[2] = {keys, keys_alt};
  int* vb[2] = {vals, vals_alt};
  // a device count that turns out small is sorted by one workgroup; the
  // multi-tile launches then see the same count and return at once
  if (cnt.p)
    hipLaunchKernelGGL((k_rs_small<K, VAL>), dim3(1), dim3(RS_SMALL_THREADS), 0, st, keys, keys_alt, vals, vals_alt,
                       cnt, passes, lo_bits);
  Count big = cnt;
  big.skip_below = cnt.p ? RS_SMALL + 1 : 0;
  for (int p = 0; p < passes; ++p) {
    const int src = p & 1, dst = src ^ 1, shift = 8 * p;
    hipLaunchKernelGGL((k_rs_hist<K>), dim3(ntiles), dim3(RS_THREADS), 0, st, kb[src], big, shift, lo_bits, ntiles,
                       hist);
    if (!fuse) hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(RS_THREADS), 0, st, hist, big, ntiles, total);
    hipLaunchKernelGGL((k_rs_scatter<K, VAL>), dim3(ntiles), dim3(RS_THREADS), 0, st, kb[src], kb[dst], vb[src],
                       vb[dst], big, shift, lo_bits, ntiles, hist, fuse ? (const int*)nullptr : total);
  }
  return RDK_OK;
}

int sort_pairs_u32(rdk_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, int* vals, int* vals_alt, Count n,
                   int end_bit, bool& in_alt, hipStream_t st) {
  in_alt = false;
  if (!n.p && n.cap <= 1) return RDK_OK;
  ProfScope ps(ctx, K_SORT, st);
  return radix_sort<uint32_t, true>(ctx, keys, keys_alt, vals, vals_alt, n, (end_bit + 7) / 8, 0, in_alt,
                                    st ? st : ctx->stream);
}

int sort_pairs_u32(rdk_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, int* vals, int* vals_alt, int n, int end_bit,
                   bool& in_alt) {
  return sort_pairs_u32(ctx, keys, keys_alt, vals, vals_alt, Count::host(n), end_bit, in_alt, nullptr);
}

int sort_pair_keys_u64(rdk_ctx* ctx, uint64_t* keys, uint64_t* keys_alt, Count n, int index_bits, bool& in_alt,
                       hipStream_t st) {
  in_alt = false;
  if (!n.p && n.cap <= 1) return RDK_OK;
  if (index_bits < 1 || index_bits > 32) {
    ctx->err = "sort_pair_keys_u64: index bits out of range";
    return RDK_ERR_ARG;
  }
  ProfScope ps(ctx, K_SORT, st);
  return radix_sort<uint64_t, false>(ctx, keys, keys_alt, nullptr, nullptr, n, (2 * index_bits + 7) / 8, index_bits,
                                     in_alt, st ? st : ctx->stream);
}

// ---- exclusive scan of ints ----------------------------------------------------
// Three launches: per-tile sums, a one-workgroup scan of the tile sums, and the
// per-tile scan plus its offset.  Thread t of a tile owns SC_ITEMS consecutive
// entries.

__global__ void __launch_bounds__(SC_THREADS) k_scan_tiles(const int* __restrict__ in, Count cnt,
                                                           int* __restrict__ tsum) {
  __shared__ int ws[SC_THREADS / 64];
  const int n = cnt.get();
  const int t0 = blockIdx.x * SC_TILE;
  if (t0 >= n) return;
  int s = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    const int i = t0 + threadIdx.x * SC_ITEMS + k;
    if (i < n) s += in[i];
  }
  int all;
  (void)block_excl_scan<SC_THREADS>(s, ws, &all);
  if (threadIdx.x == 0) tsum[blockIdx.x] = all;
}

__global__ void __launch_bounds__(1024) k_scan_tile_sums(int* __restrict__ tsum, Count cnt) {
  __shared__ int ws[16];
  const int n = cnt.get();
  const int nt = (n + SC_TILE - 1) / SC_TILE;
  int carry = 0;
  for (int c0 = 0; c0 < nt; c0 += 1024) {
    const int i = c0 + threadIdx.x;
    const int v = i < nt ? tsum[i] : 0;
    int all;
    const int ex = block_excl_scan<1024>(v, ws, &all);
    if (i < nt) tsum[i] = carry + ex;
    carry += all;
  }
}

__global__ void __launch_bounds__(SC_THREADS) k_scan_apply(const int* __restrict__ in, int* __restrict__ out, Count cnt,
                                                           const int* __restrict__ tsum) {
  __shared__ int ws[SC_THREADS / 64];
  const int n = cnt.get();
  const int t0 = blockIdx.x * SC_TILE;
  if (t0 >= n) return;
  int v[SC_ITEMS];
  int s = 0;
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    const int i = t0 + threadIdx.x * SC_ITEMS + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int all;
  int run = tsum[blockIdx.x] + block_excl_scan<SC_THREADS>(s, ws, &all);
#pragma unroll
  for (int k = 0; k < SC_ITEMS; ++k) {
    const int i = t0 + threadIdx.x * SC_ITEMS + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

// One workgroup for at most SC_SMALL entries: thread t owns SC_SMALL / 1024
// consecutive entries.
constexpr int SC_SMALL = 8192;
__global__ void __launch_bounds__(1024) k_scan_small(const int* __restrict__ in, int* __restrict__ out, Count cnt) {
  constexpr int PER = SC_SMALL / 1024;
  __shared__ int ws[16];
  const int n = cnt.get();
  if (n > SC_SMALL) return;  // the three-launch scan of a captured step handles it
  int v[PER];
  int s = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x * PER + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  int all;
  int run = block_excl_scan<1024>(s, ws, &all);
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x * PER + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

int exclusive_scan_int(rdk_ctx* ctx, const int* in, int* out, Count n, hipStream_t st) {
  if (n.cap <= 0) return RDK_OK;
  st = st ? st : ctx->stream;
  ProfScope ps(ctx, K_SORT, st);
  if (n.cap <= SC_SMALL || n.p) hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, st, in, out, n);
  if (n.cap <= SC_SMALL) return RDK_OK;
  Count big = n;
  big.skip_below = n.p ? SC_SMALL + 1 : 0;
  const int ntiles = grid_for(n.cap, SC_TILE);
  RDK_TRY(ensure(ctx, ctx->scan_tmp, (size_t)ntiles + 16));
  hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(SC_THREADS), 0, st, in, big, ctx->scan_tmp.p);
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, st, ctx->scan_tmp.p, big);
  hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(SC_THREADS), 0, st, in, out, big, ctx->scan_tmp.p);
  return RDK_OK;
}

int exclusive_scan_int(rdk_ctx* ctx, const int* in, int* out, int n) {
  return exclusive_scan_int(ctx, in, out, Count::host(n), nullptr);
}

}  // namespace rdk

// rdk_sort.hip -- radix sorts and scans used between the hand-written
// stages (cell binning, pair canonicalisation, colour grouping, body CSR).
// These are plain library primitives (rocPRIM, header-only), sized per call
// with a growing scratch buffer owned by the context.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "rdk_internal.hpp"

namespace rdk {

static int ensure_temp(rdk_ctx* ctx, size_t bytes) { return ensure(ctx, ctx->temp, bytes + 256); }

int sort_pairs_u32(rdk_ctx* ctx, uint32_t* keys, uint32_t* keys_alt, int* vals, int* vals_alt, int n,
                   int end_bit, bool& in_alt) {
  in_alt = false;
  if (n <= 1) return RDK_OK;
  ProfScope ps(ctx, K_SORT);
  size_t bytes = 0;
  RDK_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys, keys_alt, vals, vals_alt, (size_t)n, 0,
                                    (unsigned)end_bit, ctx->stream));
  RDK_TRY(ensure_temp(ctx, bytes));
  bytes = ctx->temp.cap;
  RDK_HIP(rocprim::radix_sort_pairs((void*)ctx->temp.p, bytes, keys, keys_alt, vals, vals_alt, (size_t)n, 0,
                                    (unsigned)end_bit, ctx->stream));
  in_alt = true;
  return RDK_OK;
}

int sort_keys_u64(rdk_ctx* ctx, uint64_t* keys, uint64_t* keys_alt, int n, int end_bit, bool& in_alt) {
  in_alt = false;
  if (n <= 1) return RDK_OK;
  ProfScope ps(ctx, K_SORT);
  size_t bytes = 0;
  RDK_HIP(rocprim::radix_sort_keys(nullptr, bytes, keys, keys_alt, (size_t)n, 0, (unsigned)end_bit,
                                   ctx->stream));
  RDK_TRY(ensure_temp(ctx, bytes));
  bytes = ctx->temp.cap;
  RDK_HIP(rocprim::radix_sort_keys((void*)ctx->temp.p, bytes, keys, keys_alt, (size_t)n, 0,
                                   (unsigned)end_bit, ctx->stream));
  in_alt = true;
  return RDK_OK;
}

int exclusive_scan_int(rdk_ctx* ctx, const int* in, int* out, int n) {
  if (n <= 0) return RDK_OK;
  ProfScope ps(ctx, K_SORT);
  size_t bytes = 0;
  RDK_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), ctx->stream));
  RDK_TRY(ensure_temp(ctx, bytes));
  bytes = ctx->temp.cap;
  RDK_HIP(rocprim::exclusive_scan((void*)ctx->temp.p, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(),
                                  ctx->stream));
  return RDK_OK;
}

}  // namespace rdk

// noalloc.hip — segtaco_spmm_noalloc: segtaco_spmm for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). The element-balanced schedules need a
// workspace (tile rows, fix-up records, carry partials; eb_cs also its sorted
// copy of A); here it is a fixed array of this library in device memory,
// `scratch`, instead of a caller buffer. Each GPU the library runs on holds
// its own copy of the array until the process ends.
//
// `scratch` names a host-side stand-in in host code: its device address is
// taken once per device with hipGetSymbolAddress (with that device current,
// on the first call there) and only that address is ever handed to kernels.
#include <mutex>

#include "engine.h"
#include "segtaco/spmm.h"

namespace segtaco {
namespace {

constexpr size_t kScratchBytes = SEGTACO_NOALLOC_WORKSPACE_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up once per
// device, under the lock).
int32_t scratch_address(void **out) {
  int dev = -1;
  SEGTACO_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(SEGTACO_ERR_UNSUPPORTED, "segtaco_spmm_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    SEGTACO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(SEGTACO_ERR_HIP, "segtaco_spmm_noalloc: no device address");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return SEGTACO_OK;
}

}  // namespace
}  // namespace segtaco

using namespace segtaco;

extern "C" int32_t segtaco_spmm_noalloc(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                                        int64_t ldb, int32_t N, float *C, int64_t ldc,
                                        const segtaco_plan_t *plan, void *stream_v) {
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  if (!A || !plan) return set_error(SEGTACO_ERR_INVALID, "segtaco_spmm_noalloc: null argument");
  // products with no entries, rows or columns need no workspace
  if (A->M <= 0 || N <= 0 || A->nnz <= 0)
    return spmm_impl(A, B, b_dtype, ldb, N, C, ldc, plan, nullptr, 0, stream, nullptr);
  // the plan segtaco_spmm would run for these operands, and its workspace:
  // refused on the host, before any device work, when it does not fit
  segtaco_plan_t p = *plan;
  SEGTACO_RET(plan_resolve_ptrs_impl(A, N, b_dtype, ldb, ldc, B, C, &p));
  const size_t need = spmm_workspace_bytes_impl(A, N, &p);
  if (need > kScratchBytes)
    return set_error(SEGTACO_ERR_CAPACITY,
                     "segtaco_spmm_noalloc: this product needs %zu bytes of workspace; the "
                     "library's fixed device array holds %zu (use segtaco_spmm with a caller "
                     "workspace)",
                     need, kScratchBytes);
  void *ws = nullptr;
  if (need > 0) SEGTACO_RET(scratch_address(&ws));
  return spmm_impl(A, B, b_dtype, ldb, N, C, ldc, &p, ws, need > 0 ? kScratchBytes : 0, stream,
                   nullptr);
}

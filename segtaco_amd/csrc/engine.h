// engine.h — internal host-side helpers shared by the libsegtaco sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <string>

#include "segtaco/spmm.h"

namespace segtaco {

// Records (status, message) as this thread's last error and returns status.
int32_t set_error(int32_t status, const char *fmt, ...);
int32_t hip_error(hipError_t e, const char *what);

#define SEGTACO_HIP(call)                                     \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return ::segtaco::hip_error(e_, #call); \
  } while (0)

#define SEGTACO_RET(call)                 \
  do {                                    \
    int32_t s_ = (call);                  \
    if (s_ != SEGTACO_OK) return s_;      \
  } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }
inline int32_t ilog2(int64_t x) {
  int32_t l = 0;
  while ((int64_t(1) << (l + 1)) <= x) ++l;
  return l;
}
inline int64_t pow2_ceil(int64_t x) {
  int64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int32_t elem_bytes(int32_t dt) { return dt == SEGTACO_DT_F32 ? 4 : 2; }

// Workspace layout of an element-balanced plan: ntiles + 1 ints at part_off
// (a batch keeps the K21 partition there, a standalone call the fix-up's
// per-tile records), ntiles + 1 tile rows, then the two carry arrays.
struct EbLayout {
  int32_t ntiles;
  size_t part_off, rows_off, head_off, tail_off, total;
};
EbLayout eb_layout(int32_t nnz, int32_t N, int32_t tile);

// Largest nnz the engine accepts. Entry positions are 32-bit, and kernels
// form positions up to one tile (at most 2^22 entries) plus a staging chunk
// past a tile's start; below this bound none of them can wrap.
constexpr int32_t kMaxNnz = 0x7fffffff - (1 << 23);

// Compute units of the current device (256 on MI355X; queried once).
int32_t device_cus();

// Phases of one product: the steps that depend on A alone (kPhasePrepare:
// row partition and tile rows; eb_cs's slot plan and column-sorted records)
// and the product itself (kPhaseRun).
constexpr int32_t kPhasePrepare = 1;
constexpr int32_t kPhaseRun = 2;
constexpr int32_t kPhaseAll = kPhasePrepare | kPhaseRun;

// The device API's implementation (engine.cpp). ev: 4 optional events
// recorded around partition, main kernel and fix-up. phases: kPhaseAll for
// segtaco_spmm; kPhasePrepare alone (B and C unused, may be null) for
// segtaco_spmm_prepare; kPhaseRun alone for segtaco_spmm_prepared, which
// requires the workspace's last preparation to be of this product.
int32_t spmm_impl(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                  int64_t ldb, int32_t N, float *C, int64_t ldc,
                  const segtaco_plan_t *plan, void *ws, size_t ws_bytes,
                  hipStream_t stream, hipEvent_t *ev, int32_t phases = kPhaseAll);
size_t spmm_workspace_bytes_impl(const segtaco_csr_t *A, int32_t N,
                                 const segtaco_plan_t *plan);
// Resolves open plan fields without operand pointers (ld's as given).
int32_t plan_resolve_impl(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                          int64_t ldb, int64_t ldc, segtaco_plan_t *plan);
// The same with the operand pointers, whose alignment bounds the vector
// width: the plan segtaco_spmm runs for these operands.
int32_t plan_resolve_ptrs_impl(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                               int64_t ldb, int64_t ldc, const void *B, const void *C,
                               segtaco_plan_t *plan);

// Default element-balanced tile of a product with nnz entries: `tile`,
// halved (not below 8) while the product would have fewer than two
// workgroups of `tiles_per_block` tiles per compute unit. Each lane group
// walks its tile serially, so on a small product a large tile leaves most of
// the device idle and the call's time is one long serial walk.
int32_t latency_tile(int64_t nnz, int32_t tile, int32_t tiles_per_block);

}  // namespace segtaco

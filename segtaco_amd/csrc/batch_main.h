// batch_main.h — the persistent batched kernel over a list of heterogeneous
// SpMMs, as a template over the storage type (DT) and whether the N = 1 form
// of eb_pr is compiled in (kSv). Each (DT, kSv) pair is instantiated in its
// own translation unit (k_batch_inst.hip, compiled once per pair) so the
// pairs compile in parallel; k_batch.hip dispatches between them.
//
// The batch is a list of work units: product i owns units
// [prefix[i], prefix[i+1]), one per (row or tile block, column panel) of its
// own plan. A grid sized to what the device holds at once loops: thread 0 of
// each workgroup takes the next unit from a shared counter (an integer
// atomic; no floating-point atomics anywhere), finds its product by bisection
// over the prefix array, and the workgroup runs that product's schedule body
// (sched_rb.h / sched_eb.h) on the unit. Units of different products, sizes,
// column counts and schedules interleave freely, so a large product does not
// leave the device idle while small ones finish. Every workgroup leaves when
// the counter passes the last unit, so the grid always drains.
#pragma once

#include "launch.h"

namespace segtaco {

// Last product i with prefix[i] <= u (products with no units are skipped).
__device__ __forceinline__ int32_t find_product(const int64_t *prefix,
                                                int32_t n, int64_t u) {
  int32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= u)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Occupancy-sized grid and launch of k_batch_main<DT, kSv> (defined in
// k_batch_inst.hip, one instantiation per object file).
template <int DT, bool kSv>
hipError_t batch_main_grid_t(size_t lds, int32_t *grid);
template <int DT, bool kSv>
hipError_t launch_batch_main_t(const BatchDesc *d, int32_t n,
                               const int64_t *prefix, int64_t total,
                               unsigned long long *counter, int32_t grid,
                               size_t lds, hipStream_t s);

}  // namespace segtaco

// k_rb.hip — standalone launches of the row-balanced schedules (bodies in
// sched_rb.h): one workgroup per (row block, column panel).
#include "sched_rb.h"

namespace segtaco {

template <int DT, int V>
__global__ __launch_bounds__(kBlockThreads) void k_rb_sr(const KArgs a) {
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  rb_sr_body<DT, V>(a, blk, panel);
}

template <int DT, int V>
__global__ __launch_bounds__(kBlockThreads) void k_rb_pr(const KArgs a) {
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  rb_pr_body<DT, V>(a, blk, panel);
}

#define SEGTACO_DISPATCH(KERNEL, dt, vec, grid, args, stream)                  \
  do {                                                                        \
    const dim3 g_((unsigned)(grid)), b_(kBlockThreads);                       \
    switch ((dt) * 16 + (vec)) {                                              \
      case 0 * 16 + 1: hipLaunchKernelGGL((KERNEL<0, 1>), g_, b_, 0, stream, args); break; \
      case 0 * 16 + 2: hipLaunchKernelGGL((KERNEL<0, 2>), g_, b_, 0, stream, args); break; \
      case 0 * 16 + 4: hipLaunchKernelGGL((KERNEL<0, 4>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 1: hipLaunchKernelGGL((KERNEL<1, 1>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 2: hipLaunchKernelGGL((KERNEL<1, 2>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 4: hipLaunchKernelGGL((KERNEL<1, 4>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 8: hipLaunchKernelGGL((KERNEL<1, 8>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 1: hipLaunchKernelGGL((KERNEL<2, 1>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 2: hipLaunchKernelGGL((KERNEL<2, 2>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 4: hipLaunchKernelGGL((KERNEL<2, 4>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 8: hipLaunchKernelGGL((KERNEL<2, 8>), g_, b_, 0, stream, args); break; \
      default: return hipErrorInvalidValue;                                   \
    }                                                                         \
  } while (0)

hipError_t launch_rb_sr(int dt, int vec, const KArgs &a, hipStream_t s) {
  const int64_t grid = grid_for(a.panels, a.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  SEGTACO_DISPATCH(k_rb_sr, dt, vec, grid, a, s);
  return hipGetLastError();
}

hipError_t launch_rb_pr(int dt, int vec, const KArgs &a, hipStream_t s) {
  const int64_t grid = grid_for(a.panels, a.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  SEGTACO_DISPATCH(k_rb_pr, dt, vec, grid, a, s);
  return hipGetLastError();
}

}  // namespace segtaco

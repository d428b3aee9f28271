// k_eb_inst.hip — the standalone eb_sr / eb_pr kernels for one storage type,
// chosen by SEGTACO_EB_DT (0 FP32, 1 BF16, 2 FP16) at compile time; the
// build compiles this file once per type. Bodies in sched_eb.h.
#include "sched_eb.h"

#ifndef SEGTACO_EB_DT
#error "SEGTACO_EB_DT must be defined (0, 1 or 2)"
#endif

namespace segtaco {

template <int DT, int V>
__global__ __launch_bounds__(kBlockThreads) void k_eb_sr(const KArgs a) {
  extern __shared__ int32_t smem[];
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  eb_sr_body<DT, V>(a, blk, panel, smem);
}

template <int DT, int V>
__global__ __launch_bounds__(kBlockThreads) void k_eb_pr(const KArgs a) {
  extern __shared__ int32_t smem[];
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  eb_pr_body<DT, V>(a, blk, panel, smem);
}

// Launches KERNEL<DT, vec> (vec 1, 2, 4, and 8 for 16-bit storage).
#define SEGTACO_LAUNCH_VEC(KERNEL, DT_, vec, grid, lds, args, stream)                 \
  do {                                                                                \
    const dim3 g_((unsigned)(grid)), b_(kBlockThreads);                               \
    switch (vec) {                                                                    \
      case 1: hipLaunchKernelGGL((KERNEL<DT_, 1>), g_, b_, lds, stream, args); break; \
      case 2: hipLaunchKernelGGL((KERNEL<DT_, 2>), g_, b_, lds, stream, args); break; \
      case 4: hipLaunchKernelGGL((KERNEL<DT_, 4>), g_, b_, lds, stream, args); break; \
      case 8:                                                                         \
        if constexpr (DT_ != 0) {                                                     \
          hipLaunchKernelGGL((KERNEL<DT_, 8>), g_, b_, lds, stream, args);            \
          break;                                                                      \
        }                                                                             \
        return hipErrorInvalidValue;                                                  \
      default: return hipErrorInvalidValue;                                           \
    }                                                                                 \
  } while (0)

template <int DT>
hipError_t launch_eb_sr_t(int vec, const KArgs &a, hipStream_t s) {
  const int64_t grid = grid_for(a.panels, a.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  const size_t lds = eb_sr_lds_bytes(a);
  SEGTACO_LAUNCH_VEC(k_eb_sr, DT, vec, grid, lds, a, s);
  return hipGetLastError();
}

template <int DT>
hipError_t launch_eb_pr_t(int vec, const KArgs &a, hipStream_t s) {
  const int64_t grid = grid_for(a.panels, a.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  const size_t lds = eb_pr_lds_bytes(a);
  SEGTACO_LAUNCH_VEC(k_eb_pr, DT, vec, grid, lds, a, s);
  return hipGetLastError();
}

template hipError_t launch_eb_sr_t<SEGTACO_EB_DT>(int, const KArgs &, hipStream_t);
template hipError_t launch_eb_pr_t<SEGTACO_EB_DT>(int, const KArgs &, hipStream_t);

}  // namespace segtaco

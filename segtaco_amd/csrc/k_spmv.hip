// k_spmv.hip — the standalone launch of the N = 1 form of eb_pr (SpMV:
// y = A x, reference K5/K6, spmm_eb_pr.h:11-75, 281-946); the per-tile body
// is in spmv_body.h. One wave per tile of 64 * E entries, 4 waves per block.
#include "spmv_body.h"

namespace segtaco {
namespace {

template <int DT, int E, bool kOff32>
__global__ __launch_bounds__(256) void k_spmv(const KArgs a) {
  __shared__ int32_t smem[4 * kSvWindow];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // scalar tile index
  const int32_t t = blockIdx.x * 4 + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  spmv_tile_any<DT, E, kOff32>(a, t, smem + wave * kSvWindow);
}

template <int DT, bool kOff32>
hipError_t launch_spmv_e(const KArgs &a, dim3 g, dim3 b, hipStream_t s) {
  switch (a.tile >> 6) {
    case 1: hipLaunchKernelGGL((k_spmv<DT, 1, kOff32>), g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_spmv<DT, 2, kOff32>), g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL((k_spmv<DT, 4, kOff32>), g, b, 0, s, a); break;
    case 8: hipLaunchKernelGGL((k_spmv<DT, 8, kOff32>), g, b, 0, s, a); break;
    case 16: hipLaunchKernelGGL((k_spmv<DT, 16, kOff32>), g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_spmv(int dt, const KArgs &a, hipStream_t s) {
  const int64_t grid = ((int64_t)a.ntiles + 3) / 4;
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)grid), b(256);
  switch (dt * 2 + (a.b_off32 ? 1 : 0)) {
    case 0: return launch_spmv_e<0, false>(a, g, b, s);
    case 1: return launch_spmv_e<0, true>(a, g, b, s);
    case 2: return launch_spmv_e<1, false>(a, g, b, s);
    case 3: return launch_spmv_e<1, true>(a, g, b, s);
    case 4: return launch_spmv_e<2, false>(a, g, b, s);
    case 5: return launch_spmv_e<2, true>(a, g, b, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace segtaco

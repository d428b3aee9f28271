// k_batch.hip — host dispatch of the persistent batched launch
// (k_batch_main, batch_main.h / k_batch_inst.hip), its counter reset and the
// batched carry fix-up.
#include "batch_main.h"
#include "sched_eb.h"

namespace segtaco {

__global__ void k_batch_reset(unsigned long long *counter) { *counter = 0ull; }

__global__ __launch_bounds__(kBlockThreads) void k_batch_fixup(
    const BatchDesc *__restrict__ d, int32_t n,
    const int64_t *__restrict__ prefix, int64_t total) {
  for (int64_t u = blockIdx.x; u < total; u += gridDim.x) {
    const int32_t m = find_product(prefix, n, u);
    const BatchDesc &D = d[m];
    KArgs a = D.a;
    const int64_t local = u - D.fix_begin;
    const int32_t panel = (int32_t)(local % a.panels);
    const int32_t blk = (int32_t)(local / a.panels);
    switch (D.vec) {
      case 1: fixup_body<1>(a, blk, panel); break;
      case 2: fixup_body<2>(a, blk, panel); break;
      case 4: fixup_body<4>(a, blk, panel); break;
      default: fixup_body<8>(a, blk, panel); break;
    }
  }
}

hipError_t batch_main_grid(int dt, size_t lds, int32_t *grid, bool sv) {
  switch (dt * 2 + (sv ? 1 : 0)) {
    case 0: return batch_main_grid_t<0, false>(lds, grid);
    case 1: return batch_main_grid_t<0, true>(lds, grid);
    case 2: return batch_main_grid_t<1, false>(lds, grid);
    case 3: return batch_main_grid_t<1, true>(lds, grid);
    case 4: return batch_main_grid_t<2, false>(lds, grid);
    case 5: return batch_main_grid_t<2, true>(lds, grid);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batch_main(int dt, const BatchDesc *d, int32_t n,
                             const int64_t *prefix, int64_t total,
                             unsigned long long *counter, int32_t grid,
                             size_t lds, hipStream_t s, bool reset_counter, bool sv) {
  if (reset_counter) hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || total <= 0) return e;
  switch (dt * 2 + (sv ? 1 : 0)) {
    case 0: return launch_batch_main_t<0, false>(d, n, prefix, total, counter, grid, lds, s);
    case 1: return launch_batch_main_t<0, true>(d, n, prefix, total, counter, grid, lds, s);
    case 2: return launch_batch_main_t<1, false>(d, n, prefix, total, counter, grid, lds, s);
    case 3: return launch_batch_main_t<1, true>(d, n, prefix, total, counter, grid, lds, s);
    case 4: return launch_batch_main_t<2, false>(d, n, prefix, total, counter, grid, lds, s);
    case 5: return launch_batch_main_t<2, true>(d, n, prefix, total, counter, grid, lds, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batch_fixup(const BatchDesc *d, int32_t n,
                              const int64_t *fix_prefix, int64_t total,
                              hipStream_t s) {
  if (total <= 0) return hipSuccess;
  const int64_t grid = total < (1 << 20) ? total : (1 << 20);
  hipLaunchKernelGGL(k_batch_fixup, dim3((unsigned)grid), dim3(kBlockThreads),
                     0, s, d, n, fix_prefix, total);
  return hipGetLastError();
}

}  // namespace segtaco

// k_batch.hip — one persistent launch over a batch of heterogeneous SpMMs.
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
#include "sched_eb.h"
#include "sched_rb.h"
#include "segtaco/spmm.h"
#include "spmv_body.h"

namespace segtaco {

// kSv: the batch holds a product in the N = 1 form of eb_pr (spmv_form);
// only then is that body compiled in, since it costs the 16-bit kernel
// registers and 3 % of served throughput.
template <int DT, int V, bool kSv>
__device__ __forceinline__ void run_body(int32_t sched, const KArgs &a,
                                         int32_t blk, int32_t panel,
                                         int32_t *smem) {
  switch (sched) {
    case SEGTACO_RB_SR: rb_sr_body<DT, V>(a, blk, panel); break;
    case SEGTACO_RB_PR: rb_pr_body<DT, V>(a, blk, panel); break;
    case SEGTACO_EB_SR: eb_sr_body<DT, V>(a, blk, panel, smem); break;
    default:
      if constexpr (kSv && V == 1) {
        // eb_pr at N = 1: the form a standalone call takes (one wave per
        // tile, as in k_spmv), so the product's bits match a single call
        if (spmv_form(a, 1)) {
          const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
          const int32_t t = blk * kWavesPerBlock + wave;
          if (t < a.ntiles) spmv_tile_rt<DT>(a, t, smem + wave * kSvWindow);
          break;
        }
      }
      eb_pr_body<DT, V>(a, blk, panel, smem);
      break;
  }
}

template <int DT, bool kSv>
__device__ __forceinline__ void dispatch(int32_t sched, int32_t vec,
                                         const KArgs &a, int32_t blk,
                                         int32_t panel, int32_t *smem) {
  if (vec == 4) {
    run_body<DT, 4, kSv>(sched, a, blk, panel, smem);
  } else if (vec == 2) {
    run_body<DT, 2, kSv>(sched, a, blk, panel, smem);
  } else if (vec == 1) {
    run_body<DT, 1, kSv>(sched, a, blk, panel, smem);
  } else if constexpr (DT != 0) {
    if (vec == 8) run_body<DT, 8, kSv>(sched, a, blk, panel, smem);
  }
}

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

__global__ void k_batch_reset(unsigned long long *counter) { *counter = 0ull; }

// FP32: held at 4 waves per SIMD (128 VGPRs), which every FP32 body fits;
// the 16-bit forms need more registers and run at 3.
template <int DT, bool kSv>
__global__ __launch_bounds__(kBlockThreads)
__attribute__((amdgpu_waves_per_eu(DT == 0 ? 4 : 1))) void k_batch_main(
    const BatchDesc *__restrict__ d, int32_t n,
    const int64_t *__restrict__ prefix, int64_t total,
    unsigned long long *counter) {
  extern __shared__ int32_t smem[];
  __shared__ int64_t s_unit;
  __shared__ int32_t s_prod;
  for (;;) {
    if (threadIdx.x == 0) {
      const int64_t u = (int64_t)atomicAdd(counter, 1ull);
      s_unit = u;
      s_prod = u < total ? find_product(prefix, n, u) : 0;
    }
    __syncthreads();
    const int64_t u = s_unit;
    const int32_t m = __builtin_amdgcn_readfirstlane(s_prod);
    __syncthreads();  // every wave has read the unit before it is replaced
    if (u >= total) break;
    const BatchDesc &D = d[m];
    const KArgs a = D.a;
    const int64_t local = u - D.unit_begin;
    const int32_t panel = (int32_t)(local % a.panels);
    const int32_t blk = (int32_t)(local / a.panels);
    dispatch<DT, kSv>(D.schedule, D.vec, a, blk, panel, smem);
  }
}

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
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  const void *fn =
      sv ? (dt == 0 ? (const void *)k_batch_main<0, true>
            : dt == 1 ? (const void *)k_batch_main<1, true>
                      : (const void *)k_batch_main<2, true>)
         : (dt == 0 ? (const void *)k_batch_main<0, false>
            : dt == 1 ? (const void *)k_batch_main<1, false>
                      : (const void *)k_batch_main<2, false>);
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlockThreads,
                                                   lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  *grid = cus * per_cu;
  return hipSuccess;
}

hipError_t launch_batch_main(int dt, const BatchDesc *d, int32_t n,
                             const int64_t *prefix, int64_t total,
                             unsigned long long *counter, int32_t grid,
                             size_t lds, hipStream_t s, bool reset_counter, bool sv) {
  if (reset_counter) hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(1), 0, s, counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || total <= 0) return e;
  const dim3 g((unsigned)grid), b(kBlockThreads);
  switch (dt * 2 + (sv ? 1 : 0)) {
    case 0: hipLaunchKernelGGL((k_batch_main<0, false>), g, b, lds, s, d, n, prefix, total, counter); break;
    case 1: hipLaunchKernelGGL((k_batch_main<0, true>), g, b, lds, s, d, n, prefix, total, counter); break;
    case 2: hipLaunchKernelGGL((k_batch_main<1, false>), g, b, lds, s, d, n, prefix, total, counter); break;
    case 3: hipLaunchKernelGGL((k_batch_main<1, true>), g, b, lds, s, d, n, prefix, total, counter); break;
    case 4: hipLaunchKernelGGL((k_batch_main<2, false>), g, b, lds, s, d, n, prefix, total, counter); break;
    case 5: hipLaunchKernelGGL((k_batch_main<2, true>), g, b, lds, s, d, n, prefix, total, counter); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
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

// k_batch_inst.hip — k_batch_main for one (storage type, N = 1 form) pair,
// chosen by SEGTACO_BATCH_DT (0 FP32, 1 BF16, 2 FP16) and SEGTACO_BATCH_SV
// (0 or 1) at compile time; the build compiles this file once per pair.
// The persistent loop itself is described in batch_main.h.
#include "batch_main.h"
#include "sched_eb.h"
#include "sched_rb.h"
#include "segtaco/spmm.h"
#include "spmv_body.h"

#ifndef SEGTACO_BATCH_DT
#error "SEGTACO_BATCH_DT must be defined (0, 1 or 2)"
#endif
#ifndef SEGTACO_BATCH_SV
#error "SEGTACO_BATCH_SV must be defined (0 or 1)"
#endif

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

// Held at 4 waves per SIMD (128 VGPRs). Every FP32 body fits; the 16-bit
// forms would take 145 VGPRs (3 waves) and instead spill 48-60 bytes per
// lane. Measured in BF16 on the 512-matrix power-law pool, 3 waves -> 4:
// server p50 275 -> 252 us, p99 908 -> 866 us, saturated 4668 -> 4828
// GFLOP/s, power-law N = 128 burst 5701 -> 5964 GFLOP/s; the 512-product
// launch 57.3 -> 58.8 ms.
template <int DT, bool kSv>
__global__ __launch_bounds__(kBlockThreads)
__attribute__((amdgpu_waves_per_eu(4))) void k_batch_main(
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

template <int DT, bool kSv>
hipError_t batch_main_grid_t(size_t lds, int32_t *grid) {
  int dev = 0, cus = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void *)k_batch_main<DT, kSv>, kBlockThreads, lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  *grid = cus * per_cu;
  return hipSuccess;
}

template <int DT, bool kSv>
hipError_t launch_batch_main_t(const BatchDesc *d, int32_t n,
                               const int64_t *prefix, int64_t total,
                               unsigned long long *counter, int32_t grid,
                               size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((k_batch_main<DT, kSv>), dim3((unsigned)grid), dim3(kBlockThreads),
                     lds, s, d, n, prefix, total, counter);
  return hipGetLastError();
}

template hipError_t batch_main_grid_t<SEGTACO_BATCH_DT, (SEGTACO_BATCH_SV != 0)>(size_t,
                                                                               int32_t *);
template hipError_t launch_batch_main_t<SEGTACO_BATCH_DT, (SEGTACO_BATCH_SV != 0)>(
    const BatchDesc *, int32_t, const int64_t *, int64_t, unsigned long long *, int32_t,
    size_t, hipStream_t);

}  // namespace segtaco

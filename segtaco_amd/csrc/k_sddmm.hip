// k_sddmm.hip — sampled dense-dense product on a CSR pattern (SDDMM):
//
//   out[p] = s_p * sum_j X[i, j] * Y[k, j]   for every stored entry p = (i, k)
//
// (reference: sddmm_csr_gpu_taco.h:70-148, TACO kernel over 2048-entry blocks
// with the K21 partition, FP64 values and one warp atomic per entry; :5-44,
// a warp per row; both are disabled in the reference's build.)
//
// Here the entries are cut into tiles of kSdTile consecutive entries, one
// wave per tile, and the tile's first row comes from the same partition pass
// the SpMM schedules use (k_tile_rows: the repaired K21 row). Each chunk of the
// tile stages its column indices, values and a window of the row pointer in
// LDS and resolves every entry's row by a binary search there. The wave is
// P = 64/Q slots of Q lanes; a slot takes one entry, each lane multiplies V
// consecutive elements of the X row and the gathered Y row, and the slot's Q
// partial sums are combined by a fixed butterfly. Every out[p] has exactly
// one writer and a fixed summation order: no atomics, no clearing, and
// results are identical run to run. U entries per slot are in flight.
#include <climits>

#include "device_common.h"
#include "sddmm.h"

namespace segtaco {

namespace {

constexpr int32_t kSdChunk = 128;
constexpr int32_t kSdPosSlack = 8;
constexpr int kSdU = 4;  // entries per slot with loads in flight

template <int DT, int V>
__global__ __launch_bounds__(256) void k_sddmm(const SddmmArgs a) {
  __shared__ int32_t smem[4 * (4 * kSdChunk + kSdPosSlack)];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t t = blockIdx.x * 4 + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  const int lq = a.lq;
  const int32_t Q = 1 << lq;
  const int32_t P = 64 >> lq;
  const int32_t q = lane & (Q - 1);
  const int32_t sl = lane >> lq;
  constexpr int32_t CH = kSdChunk;
  constexpr int32_t PW = kSdChunk + kSdPosSlack;
  int32_t *s_crd = smem + wave * (4 * CH + kSdPosSlack);
  float *s_val = reinterpret_cast<float *>(s_crd + CH);
  int32_t *s_row = s_crd + 2 * CH;
  int32_t *s_pos = s_crd + 3 * CH;

  const int32_t *__restrict__ pos = a.pos;
  const int32_t *__restrict__ crd = a.crd;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  int32_t r = a.tile_row[t];
  using R = typename RawT<DT, V>::type;

  for (int32_t cs = s; cs < e; cs += CH) {
    const int32_t cn = min(CH, e - cs);
    for (int32_t i = lane; i < cn; i += 64) {
      s_crd[i] = crd[cs + i];
      s_val[i] = load_val<DT>(a.vals, cs + i);
    }
    const int32_t rbase = r;
    const int32_t pw = min(PW, a.M + 1 - rbase);
    for (int32_t i = lane; i < pw; i += 64) s_pos[i] = pos[rbase + i];
    wave_sync();
    // row of every entry of the chunk: last j with pos[rbase + j] <= p
    for (int32_t i = lane; i < cn; i += 64) {
      const int32_t p = cs + i;
      int32_t lo = 0, hi = pw - 1;
      while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (s_pos[mid] <= p)
          lo = mid;
        else
          hi = mid - 1;
      }
      s_row[i] = (lo == pw - 1) ? row_of(pos, rbase + lo, r_hi, p) : rbase + lo;
    }
    wave_sync();
    for (int32_t pb = 0; pb < cn; pb += P * kSdU) {
      int32_t rows[kSdU], cols[kSdU];
      bool ok[kSdU];
      float part[kSdU];
#pragma unroll
      for (int u = 0; u < kSdU; ++u) {
        const int32_t i = pb + u * P + sl;
        ok[u] = i < cn;
        rows[u] = ok[u] ? s_row[i] : 0;
        cols[u] = ok[u] ? s_crd[i] : 0;
        part[u] = 0.0f;
      }
      for (int32_t jc = 0; jc < a.J; jc += Q * V) {
        const int32_t c = jc + q * V;
        const bool cok = c < a.J;
        R xr[kSdU], yr[kSdU];
#pragma unroll
        for (int u = 0; u < kSdU; ++u) {
          if (ok[u] && cok) {
            yr[u] = load_raw<DT, V>(a.Y, (int64_t)cols[u] * a.ldy + c);
            xr[u] = load_raw<DT, V>(a.X, (int64_t)rows[u] * a.ldx + c);
          }
        }
#pragma unroll
        for (int u = 0; u < kSdU; ++u) {
          if (ok[u] && cok) {
            float x[V], y[V];
            unpack<DT, V>(xr[u], x);
            unpack<DT, V>(yr[u], y);
#pragma unroll
            for (int j = 0; j < V; ++j) part[u] = fmaf(x[j], y[j], part[u]);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kSdU; ++u) {
        for (int32_t d = 1; d < Q; d <<= 1) part[u] += __shfl_xor(part[u], d, 64);
        const int32_t i = pb + u * P + sl;
        if (ok[u] && q == 0) a.out[cs + i] = s_val[i] * part[u];
      }
    }
    r = s_row[cn - 1];
    wave_sync();
  }
}

}  // namespace

int32_t sddmm_tile() { return kSdTile; }

hipError_t launch_sddmm(int dt, int vec, const SddmmArgs &a, hipStream_t s) {
  const int64_t grid = (a.ntiles + 3) / 4;
  if (grid <= 0) return hipSuccess;
  const dim3 g((unsigned)grid), b(256);
  switch (dt * 16 + vec) {
    case 0 * 16 + 1: hipLaunchKernelGGL((k_sddmm<0, 1>), g, b, 0, s, a); break;
    case 0 * 16 + 2: hipLaunchKernelGGL((k_sddmm<0, 2>), g, b, 0, s, a); break;
    case 0 * 16 + 4: hipLaunchKernelGGL((k_sddmm<0, 4>), g, b, 0, s, a); break;
    case 1 * 16 + 1: hipLaunchKernelGGL((k_sddmm<1, 1>), g, b, 0, s, a); break;
    case 1 * 16 + 2: hipLaunchKernelGGL((k_sddmm<1, 2>), g, b, 0, s, a); break;
    case 1 * 16 + 4: hipLaunchKernelGGL((k_sddmm<1, 4>), g, b, 0, s, a); break;
    case 1 * 16 + 8: hipLaunchKernelGGL((k_sddmm<1, 8>), g, b, 0, s, a); break;
    case 2 * 16 + 1: hipLaunchKernelGGL((k_sddmm<2, 1>), g, b, 0, s, a); break;
    case 2 * 16 + 2: hipLaunchKernelGGL((k_sddmm<2, 2>), g, b, 0, s, a); break;
    case 2 * 16 + 4: hipLaunchKernelGGL((k_sddmm<2, 4>), g, b, 0, s, a); break;
    case 2 * 16 + 8: hipLaunchKernelGGL((k_sddmm<2, 8>), g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace segtaco

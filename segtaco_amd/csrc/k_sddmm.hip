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
// P = 64/Q slots of Q lanes; per batch a slot takes U consecutive entries,
// each lane multiplies V consecutive elements of the X row and the gathered
// Y row of each, and the slot's U x Q partial sums are reduced by a fixed
// tree (transposed across the U entries, then a butterfly). Every out[p]
// has exactly one writer and a fixed summation order: no atomics, no
// clearing, and results are identical run to run.
#include <climits>

#include "device_common.h"
#include "sddmm.h"

namespace segtaco {

namespace {

constexpr int32_t kSdChunk = 128;
constexpr int32_t kSdPosSlack = 8;
constexpr int kSdU = 4;  // consecutive entries per slot and batch

template <int DT, int V, bool kOff32>
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
  // With Q >= kSdU lanes per slot the kSdU partial sums of a batch are
  // reduced together (transposed); with fewer, one butterfly per entry.
  const bool transposed = Q >= kSdU;
  const int32_t lq_rest = lq - 2;  // log2(Q / kSdU) when transposed

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
    // A batch gives slot sl the kSdU consecutive entries pb + sl*kSdU + u.
    // Every lane issues all 2*kSdU loads of a batch with no branch around
    // them (entries past the chunk re-load its last entry, columns past J
    // column 0; neither is used), so the load count on every path is the
    // same and no wait covers more than the batch.
    for (int32_t pb = 0; pb < cn; pb += P * kSdU) {
      int32_t rows[kSdU], cols[kSdU];
      float part[kSdU];
#pragma unroll
      for (int u = 0; u < kSdU; ++u) {
        const int32_t i = min(pb + sl * kSdU + u, cn - 1);
        rows[u] = s_row[i];
        cols[u] = s_crd[i];
        part[u] = 0.0f;
      }
      for (int32_t jc = 0; jc < a.J; jc += Q * V) {
        const int32_t c = jc + q * V;
        const bool cok = c < a.J;
        const int32_t c_ld = cok ? c : 0;
        R xr[kSdU], yr[kSdU];
#pragma unroll
        for (int u = 0; u < kSdU; ++u) {
          yr[u] = load_row_raw<DT, V, kOff32>(a.Y, cols[u], a.ldy, c_ld);
          xr[u] = load_row_raw<DT, V, kOff32>(a.X, rows[u], a.ldx, c_ld);
        }
        if (cok) {
#pragma unroll
          for (int u = 0; u < kSdU; ++u) {
            float x[V], y[V];
            unpack<DT, V>(xr[u], x);
            unpack<DT, V>(yr[u], y);
#pragma unroll
            for (int j = 0; j < V; ++j) part[u] = fmaf(x[j], y[j], part[u]);
          }
        }
      }
      const int32_t i0 = pb + sl * kSdU;
      if (transposed) {
        // Step d = Q/2: the lower half of the slot keeps entries {0,1}, the
        // upper half {2,3}, each adding the partner's partials of the kept
        // entries; step d = Q/4 splits again. Every lane then holds one
        // entry, (q >> lq_rest) & 3, summed over 4 lanes; a butterfly over
        // the remaining Q/4 lanes completes it.
        const int32_t d1 = Q >> 1, d2 = Q >> 2;
        const bool hi1 = (q & d1) != 0;
        float h0 = hi1 ? part[0] : part[2];
        float h1 = hi1 ? part[1] : part[3];
        h0 = (hi1 ? part[2] : part[0]) + __shfl_xor(h0, d1, 64);
        h1 = (hi1 ? part[3] : part[1]) + __shfl_xor(h1, d1, 64);
        const bool hi2 = (q & d2) != 0;
        const float send = hi2 ? h0 : h1;
        float v = (hi2 ? h1 : h0) + __shfl_xor(send, d2, 64);
        for (int32_t d = 1; d < d2; d <<= 1) v += __shfl_xor(v, d, 64);
        const int32_t u = (q >> lq_rest) & 3;
        const int32_t i = i0 + u;
        if ((q & (d2 - 1)) == 0 && i < cn) a.out[cs + i] = s_val[i] * v;
      } else {
#pragma unroll
        for (int u = 0; u < kSdU; ++u) {
          for (int32_t d = 1; d < Q; d <<= 1) part[u] += __shfl_xor(part[u], d, 64);
          const int32_t i = i0 + u;
          if (q == 0 && i < cn) a.out[cs + i] = s_val[i] * part[u];
        }
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
#define SEGTACO_SD_CASE(DT, V)                                                  \
  case DT * 16 + V:                                                             \
    if (a.off32)                                                                \
      hipLaunchKernelGGL((k_sddmm<DT, V, true>), g, b, 0, s, a);                \
    else                                                                        \
      hipLaunchKernelGGL((k_sddmm<DT, V, false>), g, b, 0, s, a);               \
    break;
  switch (dt * 16 + vec) {
    SEGTACO_SD_CASE(0, 1)
    SEGTACO_SD_CASE(0, 2)
    SEGTACO_SD_CASE(0, 4)
    SEGTACO_SD_CASE(1, 1)
    SEGTACO_SD_CASE(1, 2)
    SEGTACO_SD_CASE(1, 4)
    SEGTACO_SD_CASE(1, 8)
    SEGTACO_SD_CASE(2, 1)
    SEGTACO_SD_CASE(2, 2)
    SEGTACO_SD_CASE(2, 4)
    SEGTACO_SD_CASE(2, 8)
    default: return hipErrorInvalidValue;
  }
#undef SEGTACO_SD_CASE
  return hipGetLastError();
}

}  // namespace segtaco

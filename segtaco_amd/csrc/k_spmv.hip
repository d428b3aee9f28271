// k_spmv.hip — the element-balanced parallel-reduction schedule (eb_pr) for a
// single column of B (SpMV: y = A x), reference K5/K6 (spmm_eb_pr.h:11-75,
// 281-946) at N = 1.
//
// With one column there is one lane per entry slot, and eb_pr's general body
// pays a six-step segmented scan for every 64 entries. Here each lane of a
// wave owns E = tile/64 consecutive entries of the wave's tile and sums them
// sequentially (a row that starts and ends inside the lane is written at
// once); one segmented scan over the 64 lanes then joins the rows that cross
// lane boundaries, and each such row is written by the lane that holds its
// last entry. Rows that cross the tile's boundaries go to the same carry
// arrays as every element-balanced schedule (carry_head[t] for a row that
// started in an earlier tile, carry_tail[t] for one that continues past it)
// and k_fixup combines them. Each y element has one writer and a fixed
// summation order (sequential inside a lane, then a fixed tree over lanes):
// results are identical run to run.
#include "sched_eb.h"
#include "spmv.h"

namespace segtaco {
namespace {

// Row-pointer window per wave: the rows of a tile of 64 * kSvMaxE entries
// plus room for empty rows between them (a tile spanning more rows reads
// pos from memory past the window).
constexpr int32_t kSvWindow = 64 * kSvMaxE + 64;

template <int DT, int E, bool kWide, bool kOff32>
__device__ __forceinline__ void spmv_tile(const KArgs &a, int32_t t, int32_t *s_pos) {
  const int lane = threadIdx.x & 63;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  const int32_t r0 = a.tile_row[t];
  // the row holding entry e, M when e == nnz: no entry of the tile is later
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  const int32_t p0 = s + lane * E;
  const int32_t p1 = min(e, p0 + E);
  const bool active = p0 < p1;
  const int32_t n = active ? p1 - p0 : 0;
  // The lane's entry loads first (independent of the row pointer), all of
  // them unconditional — past the tile they re-read the last entry and are
  // not used — so that none of them sits behind a branch and they are in
  // flight together.
  int32_t cc[E];
  float vv[E];
#pragma unroll
  for (int u = 0; u < E; ++u) {
    const int32_t q = min(p0 + u, a.nnz - 1);
    cc[u] = a.crd[q];
    vv[u] = load_val<DT>(a.vals, q);
  }
  // stage pos[r0 .. r_hi + 1], the rows the tile can touch
  const int32_t pw = min(kSvWindow, r_hi + 2 - r0);
  for (int32_t i = lane; i < pw; i += 64) s_pos[i] = a.pos[r0 + i];
  wave_sync();
  // pos[row] for rows r0 .. r_hi + 1 (from memory past the window when kWide)
  auto P = [&](int32_t row) -> int32_t {
    const int32_t i = row - r0;
    if constexpr (kWide) return i < pw ? s_pos[i] : a.pos[row];
    return s_pos[i];
  };
  auto dst = [&](int32_t row) -> float * {
    if (P(row) < s) return a.carry_head + t;
    if (P(row + 1) > e) return a.carry_tail + t;
    return a.C + (int64_t)row * a.ldc;
  };

  // head row: the row holding p0 (last row r with pos[r] <= p0)
  int32_t hr = -1 - lane, tr = -1 - lane;  // distinct per lane when inactive
  float hp = 0.0f, tp = 0.0f;
  if (active) {
    // the gathers of x, in flight behind the row search below
    float xx[E];
#pragma unroll
    for (int u = 0; u < E; ++u)
      xx[u] = load_val<DT>(a.B, kOff32 ? (int64_t)__umul24((uint32_t)cc[u], (uint32_t)a.ldb)
                                       : (int64_t)cc[u] * a.ldb);
    int32_t lo = r0, hi = r_hi;
    while (lo < hi) {
      const int32_t mid = (lo + hi + 1) >> 1;
      if (P(mid) <= p0)
        lo = mid;
      else
        hi = mid - 1;
    }
    hr = lo;
    int32_t cur = hr, cur_end = P(hr + 1);
    float acc = 0.0f;
    bool first = true;
#pragma unroll
    for (int u = 0; u < E; ++u) {
      if (u < n) {
        const int32_t p = p0 + u;
        if (p == cur_end) {
          // row `cur` ended before entry p
          if (first) {
            hp = acc;
            first = false;
          } else {
            *dst(cur) = acc;  // started and ended inside this lane
          }
          ++cur;
          while (P(cur + 1) <= p) ++cur;  // skip empty rows
          cur_end = P(cur + 1);
          acc = 0.0f;
        }
        acc = fmaf(vv[u], xx[u], acc);
      }
    }
    tr = cur;
    tp = acc;
    if (first) hp = acc;  // one row in this lane: head == tail
  }
  const bool multi = hr != tr;
  // Carry chain over lanes for each lane's tail row: a segment starts at a
  // lane whose tail row began in it (multi) or whose row differs from the
  // previous lane's tail row; inclusive segmented scan of tp.
  const int32_t tr_prev = __shfl_up(tr, 1, 64);
  const bool cont = lane > 0 && hr == tr_prev;  // head row continues from lane-1
  int32_t f = (multi || !cont) ? 1 : 0;
  float cv = tp;
  for (int d = 1; d < 64; d <<= 1) {
    const float v_up = __shfl_up(cv, d, 64);
    const int32_t f_up = __shfl_up(f, d, 64);
    if (lane >= d && !f) {
      cv = v_up + cv;
      f = f_up;
    }
  }
  const float c_prev = __shfl_up(cv, 1, 64);
  const int32_t hr_next = __shfl_down(hr, 1, 64);
  if (active) {
    // head row of a multi-row lane ends here
    if (multi) *dst(hr) = cont ? c_prev + hp : hp;
    // tail row ends here unless the next lane continues it
    const bool next_cont = lane < 63 && hr_next == tr;
    if (!next_cont) *dst(tr) = cv;
  }
}

template <int DT, int E, bool kOff32>
__global__ __launch_bounds__(256) void k_spmv(const KArgs a) {
  __shared__ int32_t smem[4 * kSvWindow];
  const int wave = threadIdx.x >> 6;
  const int32_t t = blockIdx.x * 4 + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  int32_t *s_pos = smem + wave * kSvWindow;
  const bool wide = a.tile_row[t + 1] - a.tile_row[t] > kSvWindow - 2;
  if (wide)
    spmv_tile<DT, E, true, kOff32>(a, t, s_pos);
  else
    spmv_tile<DT, E, false, kOff32>(a, t, s_pos);
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

bool spmv_applies(const KArgs &a, int vec) {
  const int32_t E = a.tile / 64;
  return a.N == 1 && vec == 1 && a.lq == 0 && a.tile % 64 == 0 && E <= kSvMaxE &&
         E > 0 && (E & (E - 1)) == 0;
}

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

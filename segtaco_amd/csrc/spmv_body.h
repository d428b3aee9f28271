// spmv_body.h — the N = 1 form of the eb_pr schedule (reference K5/K6,
// spmm_eb_pr.h:11-75, 281-946), shared by the standalone kernel (k_spmv.hip)
// and the batched kernel (k_batch.hip), so that a product gives the same bits
// in a batch as alone.
//
// With one column there is one lane per entry slot, and eb_pr's general body
// pays a six-step segmented scan for every 64 entries. Here each lane of a
// wave owns E = tile/64 consecutive entries of the wave's tile and sums them
// sequentially (a row that starts and ends inside the lane is written at
// once); one segmented scan over the 64 lanes then joins the rows that cross
// lane boundaries, and each such row is written by the lane that holds its
// last entry. Rows that cross the tile's boundaries go to the same carry
// arrays as every element-balanced schedule (carry_head[t] for a row that
// started in an earlier tile, carry_tail[t] for one that continues past it),
// which the fix-up combines. Each y element has one writer and a fixed
// summation order (sequential inside a lane, then a fixed tree over lanes):
// results are identical run to run.
//
// On a full tile whose crd and vals are aligned to one lane's run, each lane
// reads its run with 64/128-bit loads, so the wave reads the tile's crd and
// vals as contiguous blocks; otherwise one load per entry. (A persistent
// form that loaded the next tile's entries during the current tile's
// gathers measured slower: 28.9 vs 25.9 us on the power-law 262144^2 matrix
// at tile 256.)
#pragma once

#include "sched_eb.h"
#include "spmv.h"

namespace segtaco {

// Entries per 64/128-bit load of the lane's run: crd (int32) and vals.
template <int E>
inline constexpr int kSvCrdRun = E < 4 ? E : 4;
template <int DT, int E>
inline constexpr int kSvValRun = E < (DT == 0 ? 4 : 8) ? E : (DT == 0 ? 4 : 8);

// One stored element of a 4- or 2-byte array (no null check).
template <int DT>
__device__ __forceinline__ float spmv_elem(const void *base, int64_t i) {
  if constexpr (DT == 0)
    return static_cast<const float *>(base)[i];
  else
    return half_to_f32<DT>(static_cast<const uint16_t *>(base)[i]);
}

// One lane's run of E entries from p0. kVec: 64/128-bit loads (p0 + E <=
// nnz, p0 a multiple of E, crd and vals aligned to one run). Otherwise one
// load per entry at min(p, nnz - 1); entries past the tile re-read the last
// entry and are not used. Every load is unconditional, so that none sits
// behind a branch (which makes the compiler wait for all loads in flight):
// without values, the value loads read crd (at most 4 bytes per entry, so
// in bounds and aligned like crd) and the results are replaced by 1.0.
template <int DT, int E, bool kVec>
__device__ __forceinline__ void spmv_load_run(const KArgs &a, int32_t p0, int32_t (&cc)[E],
                                              float (&vv)[E]) {
  const bool has_vals = a.vals != nullptr;
  const void *vsrc = has_vals ? a.vals : static_cast<const void *>(a.crd);
  if constexpr (kVec) {
    constexpr int CI = kSvCrdRun<E>, CV = kSvValRun<DT, E>;
#pragma unroll
    for (int j = 0; j < E; j += CI) {
      if constexpr (CI == 4) {
        const int4 w = *reinterpret_cast<const int4 *>(a.crd + p0 + j);
        cc[j] = w.x; cc[j + 1] = w.y; cc[j + 2] = w.z; cc[j + 3] = w.w;
      } else if constexpr (CI == 2) {
        const int2 w = *reinterpret_cast<const int2 *>(a.crd + p0 + j);
        cc[j] = w.x; cc[j + 1] = w.y;
      } else {
        cc[j] = a.crd[p0 + j];
      }
    }
#pragma unroll
    for (int j = 0; j < E; j += CV) {
      float w[CV];
      load_vec<DT, CV>(vsrc, p0 + j, w);
#pragma unroll
      for (int u = 0; u < CV; ++u) vv[j + u] = w[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < E; ++u) {
      const int32_t q = min(p0 + u, a.nnz - 1);
      cc[u] = a.crd[q];
      vv[u] = spmv_elem<DT>(vsrc, q);
    }
  }
#pragma unroll
  for (int u = 0; u < E; ++u) vv[u] = has_vals ? vv[u] : 1.0f;
}

// x[crd] for every slot of the run, unconditionally (every slot holds the
// column of a stored entry).
template <int DT, int E, bool kOff32>
__device__ __forceinline__ void spmv_gather(const KArgs &a, const int32_t (&cc)[E],
                                            float (&xx)[E]) {
#pragma unroll
  for (int u = 0; u < E; ++u)
    xx[u] = spmv_elem<DT>(a.B, kOff32 ? (int64_t)__umul24((uint32_t)cc[u], (uint32_t)a.ldb)
                                      : (int64_t)cc[u] * a.ldb);
}

// Rows, sums and writes of tile t, once pos[r0 .. r0 + pw) is staged in
// s_pos (every row the tile touches when !kWide; past the window a row
// pointer is read from memory) and the lane's entries and x values are in
// registers.
template <int E, bool kWide>
__device__ __forceinline__ void spmv_finish(const KArgs &a, int32_t t, int32_t r0, int32_t pw,
                                            const int32_t *s_pos, const float (&vv)[E],
                                            const float (&xx)[E]) {
  const int lane = threadIdx.x & 63;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  // the row holding entry e, M when e == nnz: no entry of the tile is later
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  const int32_t p0 = s + lane * E;
  const int32_t p1 = min(e, p0 + E);
  const bool active = p0 < p1;
  const int32_t n = active ? p1 - p0 : 0;
  // pos[row] for rows r0 .. r_hi + 1
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

// One tile: the lane's run (64/128-bit loads when kVec), pos[r0 .. r_hi + 1]
// staged (up to the window), then the x gathers.
template <int DT, int E, bool kWide, bool kOff32, bool kVec>
__device__ __forceinline__ void spmv_tile(const KArgs &a, int32_t t, int32_t *s_pos) {
  const int lane = threadIdx.x & 63;
  int32_t cc[E];
  float vv[E];
  spmv_load_run<DT, E, kVec>(a, t * a.tile + lane * E, cc, vv);
  const int32_t r0 = a.tile_row[t];
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  const int32_t pw = min(kSvWindow, r_hi + 2 - r0);
  for (int32_t i = lane; i < pw; i += 64) s_pos[i] = a.pos[r0 + i];
  wave_sync();
  float xx[E];
  spmv_gather<DT, E, kOff32>(a, cc, xx);
  spmv_finish<E, kWide>(a, t, r0, pw, s_pos, vv, xx);
}

// Whether tile t can read its runs with vector loads: the tile is full (no
// lane's run passes nnz) and crd and vals are aligned to one run (p0 = s +
// lane * E is a multiple of E). Without values the value loads read crd,
// whose alignment covers theirs.
template <int DT, int E>
__device__ __forceinline__ bool spmv_vec_ok(const KArgs &a, int32_t t) {
  const uintptr_t elem = DT == 0 ? 4 : 2;
  return (int64_t)(t + 1) * a.tile <= a.nnz &&
         ((uintptr_t)a.crd & (4 * kSvCrdRun<E> - 1)) == 0 &&
         ((uintptr_t)a.vals & (elem * kSvValRun<DT, E> - 1)) == 0;
}


// Tile t of an SpMV-form product, E entries per lane (E = tile / 64): picks
// the row-pointer window form and the load form for this tile.
template <int DT, int E, bool kOff32>
__device__ __forceinline__ void spmv_tile_any(const KArgs &a, int32_t t, int32_t *s_pos) {
  const bool wide = a.tile_row[t + 1] - a.tile_row[t] > kSvWindow - 2;
  if (spmv_vec_ok<DT, E>(a, t)) {
    if (wide)
      spmv_tile<DT, E, true, kOff32, true>(a, t, s_pos);
    else
      spmv_tile<DT, E, false, kOff32, true>(a, t, s_pos);
  } else {
    if (wide)
      spmv_tile<DT, E, true, kOff32, false>(a, t, s_pos);
    else
      spmv_tile<DT, E, false, kOff32, false>(a, t, s_pos);
  }
}

// The same with E and the offset form read from the arguments (the batched
// kernel's per-product dispatch); `a` must satisfy spmv_form.
template <int DT>
__device__ __forceinline__ void spmv_tile_rt(const KArgs &a, int32_t t, int32_t *s_pos) {
  if (a.b_off32) {
    switch (a.tile >> 6) {
      case 1: spmv_tile_any<DT, 1, true>(a, t, s_pos); break;
      case 2: spmv_tile_any<DT, 2, true>(a, t, s_pos); break;
      case 4: spmv_tile_any<DT, 4, true>(a, t, s_pos); break;
      case 8: spmv_tile_any<DT, 8, true>(a, t, s_pos); break;
      default: spmv_tile_any<DT, 16, true>(a, t, s_pos); break;
    }
  } else {
    switch (a.tile >> 6) {
      case 1: spmv_tile_any<DT, 1, false>(a, t, s_pos); break;
      case 2: spmv_tile_any<DT, 2, false>(a, t, s_pos); break;
      case 4: spmv_tile_any<DT, 4, false>(a, t, s_pos); break;
      case 8: spmv_tile_any<DT, 8, false>(a, t, s_pos); break;
      default: spmv_tile_any<DT, 16, false>(a, t, s_pos); break;
    }
  }
}

}  // namespace segtaco

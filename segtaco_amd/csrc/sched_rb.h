// sched_rb.h — row-balanced schedule bodies for gfx950 (wave64), shared by
// the standalone kernels (k_rb.hip) and the batched kernel (k_batch.hip).
//
// rb_sr (reference: spmm_rb_sr.h:187-243, K1): each lane group of Q lanes
// owns one output row and one column tile of Q*V columns; it walks the row's
// stored entries in CSR order, gathering V-wide slices of B rows (one
// coalesced Q*V-element request per entry) and accumulating sequentially, so
// every C element is a sequential sum in entry order (Eigen's order, with
// FMA). Eight entries are in flight per lane group to cover HBM/MALL latency.
//
// rb_pr (reference: spmm_rb_pr.h:223-375, K2): P = group slots of Q lanes
// share one row; slot s takes entries s, s+P, s+2P, ... and the P partial
// sums are combined with a butterfly over lane distances Q..Q*P/2, a fixed
// tree order, so results are deterministic. The reference's 32-lane,
// 4-column warps become 64-lane waves with Q*P*R = 64.
//
// Neither schedule needs C cleared first: every row (empty ones included) is
// written exactly once.
#pragma once

#include "device_common.h"

namespace segtaco {

// One output row of rb_sr for this lane's V columns.
template <int DT, int V, bool kOff32>
__device__ __forceinline__ void rb_sr_row(const KArgs &a, int32_t row,
                                          int32_t col) {
  const int32_t *__restrict__ crd = a.crd;
  const void *__restrict__ vals = a.vals;
  const void *__restrict__ B = a.B;
  const int64_t ldb = a.ldb;
  int32_t p = a.pos[row];
  const int32_t e = a.pos[row + 1];
  float acc[V];
  zero_vec(acc);

  if (e - p > a.comp_threshold) {
    float cmp[V];
    zero_vec(cmp);
    for (; p < e; ++p) {
      float b[V];
      const float v = load_val<DT>(vals, p);
      load_row_vec<DT, V, kOff32>(B, crd[p], ldb, col, b);
#pragma unroll
      for (int j = 0; j < V; ++j) neumaier_add(acc[j], cmp[j], v, b[j]);
    }
#pragma unroll
    for (int j = 0; j < V; ++j) acc[j] += cmp[j];
  } else {
    constexpr int U = (V >= 8) ? 4 : 8;
    for (; p + U <= e; p += U) {
      int32_t c[U];
      float v[U];
      float b[U][V];
#pragma unroll
      for (int u = 0; u < U; ++u) c[u] = crd[p + u];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = load_val<DT>(vals, p + u);
#pragma unroll
      for (int u = 0; u < U; ++u)
        load_row_vec<DT, V, kOff32>(B, c[u], ldb, col, b[u]);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] = fmaf(v[u], b[u][j], acc[j]);
    }
    for (; p < e; ++p) {
      float b[V];
      const float v = load_val<DT>(vals, p);
      load_row_vec<DT, V, kOff32>(B, crd[p], ldb, col, b);
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] = fmaf(v, b[j], acc[j]);
    }
  }
  store_vec<V>(a.C + (int64_t)row * a.ldc + col, acc);
}

// rb_sr over row block `blk`, column panel `panel`: lane group gi owns rows
// [gi*R, gi*R + R), R = 2^lg, one after another.
template <int DT, int V, bool kOff32>
__device__ __forceinline__ void rb_sr_walk(const KArgs &a, int32_t blk,
                                           int32_t panel) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t gi =
      ((blk * kWavesPerBlock + wave) << (6 - lq)) + (lane >> lq);
  const int32_t col = ((panel << lq) + (lane & ((1 << lq) - 1))) * V;
  if (col >= a.N) return;
  const int64_t first = (int64_t)gi << a.lg;
  const int32_t R = 1 << a.lg;
  for (int32_t i = 0; i < R; ++i) {
    const int64_t row = first + i;
    if (row >= a.M) return;
    rb_sr_row<DT, V, kOff32>(a, (int32_t)row, col);
  }
}

template <int DT, int V>
__device__ __forceinline__ void rb_sr_body(const KArgs &a, int32_t blk,
                                           int32_t panel) {
  if (a.b_off32)
    rb_sr_walk<DT, V, true>(a, blk, panel);
  else
    rb_sr_walk<DT, V, false>(a, blk, panel);
}

// rb_pr over row block `blk`, column panel `panel`.
template <int DT, int V, bool kOff32>
__device__ __forceinline__ void rb_pr_walk(const KArgs &a, int32_t blk,
                                           int32_t panel) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq, lg = a.lg;
  const int lr = 6 - lq - lg;  // log2(rows per wave)
  const int32_t row0 = (blk * kWavesPerBlock + wave) << lr;
  if (row0 >= a.M) return;  // uniform over the wave
  const int32_t row = row0 + (lane >> (lq + lg));
  const int32_t slot = (lane >> lq) & ((1 << lg) - 1);
  const int32_t P = 1 << lg;
  const int32_t col = ((panel << lq) + (lane & ((1 << lq) - 1))) * V;
  const bool live = row < a.M && col < a.N;

  float acc[V];
  zero_vec(acc);
  if (live) {
    const int32_t *__restrict__ crd = a.crd;
    const void *__restrict__ vals = a.vals;
    const void *__restrict__ B = a.B;
    const int64_t ldb = a.ldb;
    const int32_t s = a.pos[row];
    const int32_t e = a.pos[row + 1];
    int32_t p = s + slot;
    if (e - s > a.comp_threshold) {
      float cmp[V];
      zero_vec(cmp);
      for (; p < e; p += P) {
        float b[V];
        const float v = load_val<DT>(vals, p);
        load_row_vec<DT, V, kOff32>(B, crd[p], ldb, col, b);
#pragma unroll
        for (int j = 0; j < V; ++j) neumaier_add(acc[j], cmp[j], v, b[j]);
      }
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] += cmp[j];
    } else {
      constexpr int U = (V >= 8) ? 2 : 4;
      for (; p + (U - 1) * P < e; p += U * P) {
        int32_t c[U];
        float v[U];
        float b[U][V];
#pragma unroll
        for (int u = 0; u < U; ++u) c[u] = crd[p + u * P];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = load_val<DT>(vals, p + u * P);
#pragma unroll
        for (int u = 0; u < U; ++u)
          load_row_vec<DT, V, kOff32>(B, c[u], ldb, col, b[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < V; ++j) acc[j] = fmaf(v[u], b[u][j], acc[j]);
      }
      for (; p < e; p += P) {
        float b[V];
        const float v = load_val<DT>(vals, p);
        load_row_vec<DT, V, kOff32>(B, crd[p], ldb, col, b);
#pragma unroll
        for (int j = 0; j < V; ++j) acc[j] = fmaf(v, b[j], acc[j]);
      }
    }
  }
  // Butterfly over the P slots of each row (lane distances Q .. Q*P/2).
  for (int d = P >> 1; d >= 1; d >>= 1) {
#pragma unroll
    for (int j = 0; j < V; ++j) acc[j] += __shfl_xor(acc[j], d << lq, 64);
  }
  if (live && slot == 0) store_vec<V>(a.C + (int64_t)row * a.ldc + col, acc);
}

template <int DT, int V>
__device__ __forceinline__ void rb_pr_body(const KArgs &a, int32_t blk,
                                           int32_t panel) {
  if (a.b_off32)
    rb_pr_walk<DT, V, true>(a, blk, panel);
  else
    rb_pr_walk<DT, V, false>(a, blk, panel);
}

}  // namespace segtaco

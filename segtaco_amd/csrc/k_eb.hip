// k_eb.hip — element-balanced schedules for gfx950 (wave64), the K21 row
// partition, and the carry fix-up that replaces the reference's atomics.
//
// Stored entries are cut into tiles of T consecutive entries (tile t covers
// [t*T, min((t+1)*T, nnz))); the first row of each tile comes from the K21
// partition (segtaco_block_starts, gpu_library.h:206-215) followed by the
// reference's repair step that skips empty rows (e.g. spmm_eb_sr.h:39-42).
//
// Carry-out without atomics (the reference uses atomicAdd for every row that
// crosses a warp or block, spmm_eb_sr.h:232-256, spmm_eb_pr.h:340-371):
//   * a row that starts and ends inside tile t is written to C by tile t;
//   * a row that started before tile t contributes its partial to
//     carry_head[t]; a row that starts in tile t and continues past it
//     contributes to carry_tail[t];
//   * k_fixup then writes every row that spans tiles t0..t1 as
//     carry_tail[t0] + carry_head[t0+1] + ... + carry_head[t1], in tile order.
// Every C element therefore has exactly one writer and a fixed summation
// order: results are bit-identical run to run.
//
// Empty rows are written as zeros by the tile whose row slice
// [t*rows_per_slice, (t+1)*rows_per_slice) holds them, so C needs no clear.
//
// eb_sr (reference K3/K14/K18: spmm_eb_sr.h:153-423, spmm_csr_gpu_taco.h:
// 114-163): a lane group of Q lanes walks its tile in entry order, eight
// entries in flight, flushing each row as it ends.
// eb_pr (reference K5/K16: spmm_eb_pr.h:11-75, 281-480): one wave per tile;
// P = 64/Q slots take P consecutive entries per pass; a segmented suffix scan
// over the slots (lane distances Q, 2Q, .. 32) leaves each row segment's sum
// in its first slot; a segment that runs into the next pass is carried in
// registers.
#include <climits>

#include "device_common.h"

namespace segtaco {

// Zero the empty rows of [r0, r1) for column panel `panel` (all Q*V columns
// of the panel), each checked by one of `nl` cooperating lanes.
template <int V>
__device__ __forceinline__ void zero_empty_rows(const KArgs &a, int32_t r0,
                                                int32_t r1, int32_t lane_in,
                                                int32_t nl, int32_t panel) {
  const int32_t *__restrict__ pos = a.pos;
  const int32_t Q = 1 << a.lq;
  float z[V];
  zero_vec(z);
  for (int32_t r = r0 + lane_in; r < r1; r += nl) {
    if (pos[r] == pos[r + 1]) {
      float *crow = a.C + (int64_t)r * a.ldc;
      for (int32_t qq = 0; qq < Q; ++qq) {
        const int32_t cc = ((panel << a.lq) + qq) * V;
        if (cc < a.N) store_vec<V>(crow + cc, z);
      }
    }
  }
}

template <int DT, int V>
__global__ __launch_bounds__(kBlockThreads) void k_eb_sr(const KArgs a) {
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t q = lane & ((1 << lq) - 1);
  const int32_t t = ((blk * kWavesPerBlock + wave) << (6 - lq)) + (lane >> lq);
  if (t >= a.ntiles) return;
  const int32_t *__restrict__ pos = a.pos;

  {
    const int64_t r0 = (int64_t)t * a.rows_per_slice;
    const int64_t r1 = r0 + a.rows_per_slice < a.M ? r0 + a.rows_per_slice : a.M;
    if (r0 < r1) zero_empty_rows<V>(a, (int32_t)r0, (int32_t)r1, q, 1 << lq, panel);
  }

  const int32_t col = ((panel << lq) + q) * V;
  if (col >= a.N) return;

  const int32_t *__restrict__ crd = a.crd;
  const void *__restrict__ vals = a.vals;
  const void *__restrict__ B = a.B;
  const int64_t ldb = a.ldb;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);

  int32_t r = a.part[t];
  while (pos[r + 1] <= s) ++r;
  int32_t rs = pos[r], re = pos[r + 1];
  bool comp = (re - rs) > a.comp_threshold;
  float acc[V], cmp[V];
  zero_vec(acc);
  zero_vec(cmp);

  auto flush = [&]() {
    const bool before = rs < s;
    const bool after = re > e;
    if (comp) {
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] += cmp[j];
    }
    float *dst = before  ? a.carry_head + (int64_t)t * a.N + col
                 : after ? a.carry_tail + (int64_t)t * a.N + col
                         : a.C + (int64_t)r * a.ldc + col;
    store_vec<V>(dst, acc);
  };
  auto next_row = [&](int32_t pp) {
    flush();
    ++r;
    while (pos[r + 1] <= pp) ++r;  // skip empty rows (zeroed by their slice)
    rs = pos[r];
    re = pos[r + 1];
    comp = (re - rs) > a.comp_threshold;
    zero_vec(acc);
    zero_vec(cmp);
  };
  auto accumulate = [&](float v, const float (&b)[V]) {
    if (comp) {
#pragma unroll
      for (int j = 0; j < V; ++j) neumaier_add(acc[j], cmp[j], v, b[j]);
    } else {
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] = fmaf(v, b[j], acc[j]);
    }
  };

  constexpr int U = (V >= 8) ? 4 : 8;
  int32_t p = s;
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
      load_vec<DT, V>(B, (int64_t)c[u] * ldb + col, b[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (p + u == re) next_row(p + u);
      accumulate(v[u], b[u]);
    }
  }
  for (; p < e; ++p) {
    float b[V];
    const float v = load_val<DT>(vals, p);
    load_vec<DT, V>(B, (int64_t)crd[p] * ldb + col, b);
    if (p == re) next_row(p);
    accumulate(v, b);
  }
  flush();
}

template <int DT, int V>
__global__ __launch_bounds__(kBlockThreads) void k_eb_pr(const KArgs a) {
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t P = 64 >> lq;  // entry slots per pass
  const int32_t t = blk * kWavesPerBlock + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  const int32_t *__restrict__ pos = a.pos;
  const int32_t q = lane & ((1 << lq) - 1);
  const int32_t sl = lane >> lq;

  {
    const int64_t r0 = (int64_t)t * a.rows_per_slice;
    const int64_t r1 = r0 + a.rows_per_slice < a.M ? r0 + a.rows_per_slice : a.M;
    if (r0 < r1) zero_empty_rows<V>(a, (int32_t)r0, (int32_t)r1, lane, 64, panel);
  }

  const int32_t col = ((panel << lq) + q) * V;
  const bool colok = col < a.N;
  const int32_t *__restrict__ crd = a.crd;
  const void *__restrict__ vals = a.vals;
  const void *__restrict__ B = a.B;
  const int64_t ldb = a.ldb;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);

  int32_t r_lo = a.part[t];
  while (pos[r_lo + 1] <= s) ++r_lo;
  const int32_t r_hi = min(a.M - 1, a.part[t + 1]);

  float open[V];
  zero_vec(open);
  int32_t open_row = -1;

  for (int32_t base = s; base < e; base += P) {
    const int32_t pidx = base + sl;
    const bool valid = pidx < e;
    int32_t rid = INT_MAX;
    float prod[V];
    zero_vec(prod);
    if (valid) {
      rid = row_of(pos, r_lo, r_hi, pidx);
      if (colok) {
        float b[V];
        const float v = load_val<DT>(vals, pidx);
        load_vec<DT, V>(B, (int64_t)crd[pidx] * ldb + col, b);
#pragma unroll
        for (int j = 0; j < V; ++j) prod[j] = v * b[j];
      }
    }
    // Segmented suffix scan: slot sl accumulates slot sl+d of the same row.
    for (int32_t d = 1; d < P; d <<= 1) {
      const int32_t tr = __shfl_down(rid, d << lq, 64);
      float tv[V];
#pragma unroll
      for (int j = 0; j < V; ++j) tv[j] = __shfl_down(prod[j], d << lq, 64);
      if (sl + d < P && tr == rid) {
#pragma unroll
        for (int j = 0; j < V; ++j) prod[j] += tv[j];
      }
    }
    const int32_t prev = __shfl_up(rid, 1 << lq, 64);
    const bool head = valid && (sl == 0 || prev != rid);
    if (sl == 0 && rid == open_row) {
#pragma unroll
      for (int j = 0; j < V; ++j) prod[j] = open[j] + prod[j];
    }
    const uint64_t hm = __ballot(head);
    const int32_t last_head = (63 - __clzll((long long)hm)) >> lq;
    const int32_t last_row = __shfl(rid, last_head << lq, 64);
    const int32_t pass_end = min(base + P, e);
    const bool last_open = pass_end < e && pos[last_row + 1] > pass_end;
    if (head && colok && !(last_open && sl == last_head)) {
      const int32_t rs = pos[rid], re = pos[rid + 1];
      const bool before = rs < s;
      const bool after = re > e;
      float *dst = before  ? a.carry_head + (int64_t)t * a.N + col
                   : after ? a.carry_tail + (int64_t)t * a.N + col
                           : a.C + (int64_t)rid * a.ldc + col;
      store_vec<V>(dst, prod);
    }
    if (last_open) {
#pragma unroll
      for (int j = 0; j < V; ++j)
        open[j] = __shfl(prod[j], (last_head << lq) + q, 64);
      open_row = last_row;
    } else {
      open_row = -1;
    }
  }
}

// Rows spanning tiles t0..t (t is the tile where the row ends):
// C[r] = carry_tail[t0] + carry_head[t0+1] + ... + carry_head[t].
template <int V>
__global__ __launch_bounds__(kBlockThreads) void k_fixup(const KArgs a) {
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t t = ((blk * kWavesPerBlock + wave) << (6 - lq)) + (lane >> lq);
  const int32_t col = ((panel << lq) + (lane & ((1 << lq) - 1))) * V;
  if (t >= a.ntiles || t == 0 || col >= a.N) return;
  const int32_t *__restrict__ pos = a.pos;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  int32_t r = a.part[t];
  while (pos[r + 1] <= s) ++r;
  const int32_t rs = pos[r], re = pos[r + 1];
  if (!(rs < s && re <= e)) return;
  const int32_t t0 = rs / a.tile;
  float acc[V];
  load_f32vec<V>(a.carry_tail + (int64_t)t0 * a.N + col, acc);
  for (int32_t tt = t0 + 1; tt <= t; ++tt) {
    float h[V];
    load_f32vec<V>(a.carry_head + (int64_t)tt * a.N + col, h);
#pragma unroll
    for (int j = 0; j < V; ++j) acc[j] += h[j];
  }
  store_vec<V>(a.C + (int64_t)r * a.ldc + col, acc);
}

__global__ __launch_bounds__(kBlockThreads) void k_block_starts(
    const int32_t *__restrict__ pos, int32_t M, int32_t T, int32_t nb,
    int32_t *__restrict__ out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  out[b] = search_before(pos, 0, M, b * (int64_t)T);
}

__global__ __launch_bounds__(kBlockThreads) void k_fill_zero(float *C,
                                                             int32_t M,
                                                             int32_t N,
                                                             int64_t ldc) {
  const int64_t total = (int64_t)M * N;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / N, c = i - r * N;
    C[r * ldc + c] = 0.0f;
  }
}

#define SEGTACO_DISPATCH(KERNEL, dt, vec, grid, args, stream)                  \
  do {                                                                        \
    const dim3 g_((unsigned)(grid)), b_(kBlockThreads);                       \
    switch ((dt) * 16 + (vec)) {                                              \
      case 0 * 16 + 1: hipLaunchKernelGGL((KERNEL<0, 1>), g_, b_, 0, stream, args); break; \
      case 0 * 16 + 2: hipLaunchKernelGGL((KERNEL<0, 2>), g_, b_, 0, stream, args); break; \
      case 0 * 16 + 4: hipLaunchKernelGGL((KERNEL<0, 4>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 1: hipLaunchKernelGGL((KERNEL<1, 1>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 2: hipLaunchKernelGGL((KERNEL<1, 2>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 4: hipLaunchKernelGGL((KERNEL<1, 4>), g_, b_, 0, stream, args); break; \
      case 1 * 16 + 8: hipLaunchKernelGGL((KERNEL<1, 8>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 1: hipLaunchKernelGGL((KERNEL<2, 1>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 2: hipLaunchKernelGGL((KERNEL<2, 2>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 4: hipLaunchKernelGGL((KERNEL<2, 4>), g_, b_, 0, stream, args); break; \
      case 2 * 16 + 8: hipLaunchKernelGGL((KERNEL<2, 8>), g_, b_, 0, stream, args); break; \
      default: return hipErrorInvalidValue;                                   \
    }                                                                         \
  } while (0)

hipError_t launch_eb_sr(int dt, int vec, const KArgs &a, hipStream_t s) {
  const int64_t grid = grid_for(a.panels, a.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  SEGTACO_DISPATCH(k_eb_sr, dt, vec, grid, a, s);
  return hipGetLastError();
}

hipError_t launch_eb_pr(int dt, int vec, const KArgs &a, hipStream_t s) {
  const int64_t grid = grid_for(a.panels, a.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  SEGTACO_DISPATCH(k_eb_pr, dt, vec, grid, a, s);
  return hipGetLastError();
}

hipError_t launch_fixup(int vec, const KArgs &a, hipStream_t s) {
  // One lane group per tile, as k_eb_sr.
  KArgs f = a;
  const int64_t groups_per_block = (int64_t)kWavesPerBlock << (6 - a.lq);
  f.nblocks = (int32_t)((a.ntiles + groups_per_block - 1) / groups_per_block);
  const int64_t grid = grid_for(f.panels, f.nblocks);
  if (grid <= 0) return hipSuccess;
  const dim3 g((unsigned)grid), b(kBlockThreads);
  switch (vec) {
    case 1: hipLaunchKernelGGL((k_fixup<1>), g, b, 0, s, f); break;
    case 2: hipLaunchKernelGGL((k_fixup<2>), g, b, 0, s, f); break;
    case 4: hipLaunchKernelGGL((k_fixup<4>), g, b, 0, s, f); break;
    case 8: hipLaunchKernelGGL((k_fixup<8>), g, b, 0, s, f); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_block_starts(const int32_t *pos, int32_t M, int32_t T,
                               int32_t num_blocks, int32_t *out,
                               hipStream_t s) {
  const int64_t n = (int64_t)num_blocks + 1;
  const int64_t grid = (n + kBlockThreads - 1) / kBlockThreads;
  hipLaunchKernelGGL(k_block_starts, dim3((unsigned)grid), dim3(kBlockThreads),
                     0, s, pos, M, T, num_blocks, out);
  return hipGetLastError();
}

hipError_t launch_fill_zero(float *C, int32_t M, int32_t N, int64_t ldc,
                            hipStream_t s) {
  const int64_t total = (int64_t)M * N;
  if (total <= 0) return hipSuccess;
  int64_t grid = (total + kBlockThreads - 1) / kBlockThreads;
  if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)grid), dim3(kBlockThreads), 0,
                     s, C, M, N, ldc);
  return hipGetLastError();
}

}  // namespace segtaco

// sched_eb.h — element-balanced schedule bodies for gfx950 (wave64) and the
// carry fix-up that replaces the reference's atomics; shared by the
// standalone kernels (k_eb.hip) and the batched kernel (k_batch.hip).
//
// Stored entries are cut into tiles of T consecutive entries (tile t covers
// [t*T, min((t+1)*T, nnz))). The tiles need each tile's first row after the
// reference's repair step that skips empty rows (e.g. spmm_eb_sr.h:39-42)
// applied to its K21 search (gpu_library.h:206-215): k_tile_rows computes
// exactly those rows in one pass over pos; k_block_starts computes the K21
// array itself bit-exactly (segtaco_block_starts) and the same rows from the
// search.
//
// Carry-out without atomics (the reference uses atomicAdd for every row that
// crosses a warp or block, spmm_eb_sr.h:232-256, spmm_eb_pr.h:340-371):
//   * a row that starts and ends inside tile t is written to C by tile t;
//   * a row that started before tile t contributes its partial to
//     carry_head[t]; a row that starts in tile t and continues past it
//     contributes to carry_tail[t];
//   * k_fixup writes every row that spans tiles t0..t1 as
//     carry_tail[t0] + carry_head[t0+1] + ... + carry_head[t1], in tile order;
//   * the empty rows are written as zeros: by k_tile_rows_fix before the main
//     kernel in a standalone call, by the fix-up (tile t zeroes the empty rows
//     of its slice [t*rows_per_slice, (t+1)*rows_per_slice)) in a batch.
// Every C element has exactly one writer and a fixed summation order: results
// are bit-identical run to run, and C needs no clearing.
//
// Both schedules stage each chunk of their tile in LDS with coalesced loads —
// column indices, values and a window of the row pointer starting at the
// chunk's first row — so that the walk over the chunk issues only B-row
// gathers to memory (the reference re-reads pos per row and, in K3/K5,
// bisects all of pos per entry, spmm_eb_sr.h:215, spmm_eb_pr.h:320).
//
// eb_sr (reference K3/K4/K14/K18: spmm_eb_sr.h:5-78, 153-712,
// spmm_csr_gpu_taco.h:114-163): a lane group of Q lanes walks its tile in
// entry order with two batches of gathers in flight, flushing each row as
// it ends.
// eb_pr (reference K5/K6/K16/K17: spmm_eb_pr.h:11-75, 281-946): one wave per
// tile; P = 64/Q slots take P consecutive entries per pass; a segmented
// suffix scan over the slots (lane distances Q, 2Q, .. 32) leaves each row
// segment's sum in its first slot; a segment that runs into the next pass is
// carried in registers.
#pragma once

#include <climits>

#include "cs.h"
#include "device_common.h"

namespace segtaco {

constexpr int32_t kPosSlack = 8;   // extra pos entries staged past a chunk
constexpr int32_t kSrChunkMax = 128;
#ifndef SEGTACO_SR_BATCH
#define SEGTACO_SR_BATCH 8
#endif
// eb_sr gathers per batch (two batches in flight per lane group)
constexpr int kSrBatch = SEGTACO_SR_BATCH;
constexpr int32_t kPrChunk = 128;
constexpr int kPrPasses = 8;       // eb_pr passes whose gathers are in flight

// eb_sr chunk: the LDS per block stays under 64 KiB for every lane count.
__host__ __device__ inline int32_t eb_sr_chunk(int32_t lq) {
  const int32_t c = 16 << lq;
  return c < kSrChunkMax ? c : kSrChunkMax;
}

__host__ __device__ inline size_t eb_sr_lds_bytes_lq(int32_t lq) {
  const int32_t ch = eb_sr_chunk(lq);
  const size_t groups = (size_t)kWavesPerBlock << (6 - lq);
  return groups * (size_t)(3 * ch + kPosSlack) * sizeof(int32_t);
}

__host__ __device__ inline size_t eb_pr_lds_bytes_fixed() {
  return (size_t)kWavesPerBlock * (4 * kPrChunk + kPosSlack) * sizeof(int32_t);
}

// Zero the empty rows of [r0, r1) for column panel `panel` (all Q*V columns
// of the panel), each row checked by one of `nl` cooperating lanes.
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

// eb_sr over tile block `blk` (kWavesPerBlock * 64/Q tiles, one per lane
// group), column panel `panel`; smem holds eb_sr_lds_bytes_lq(lq) bytes.
// kWide: some tile of the wave spans more rows than the chunk's LDS window of
// the row pointer holds, so row lookups past the window read pos from memory.
// Otherwise every lookup is an LDS read: a lookup that may go to either
// address space compiles to a flat load, whose wait drains every gather the
// lane group has in flight.
template <int DT, int V, bool kWide, bool kOff32>
__device__ __forceinline__ void eb_sr_walk(const KArgs &a, int32_t blk,
                                           int32_t panel, int32_t *smem) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t Q = 1 << lq;
  const int32_t q = lane & (Q - 1);
  const int32_t gib = (wave << (6 - lq)) + (lane >> lq);  // group in block
  const int32_t t = blk * (kWavesPerBlock << (6 - lq)) + gib;
  const int32_t CH = eb_sr_chunk(lq);
  const int32_t PW = CH + kPosSlack;
  int32_t *s_crd = smem + gib * (3 * CH + kPosSlack);
  float *s_val = reinterpret_cast<float *>(s_crd + CH);
  int32_t *s_pos = s_crd + 2 * CH;

  const int32_t *__restrict__ pos = a.pos;
  const int32_t *__restrict__ crd = a.crd;
  const void *__restrict__ vals = a.vals;
  const void *__restrict__ B = a.B;
  const int64_t ldb = a.ldb;
  const int32_t col = ((panel << lq) + q) * V;
  const bool colok = col < a.N;
  const bool tok = t < a.ntiles;
  const int32_t s = tok ? t * a.tile : 0;
  const int32_t e =
      tok ? (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile) : 0;

  int32_t r = tok ? a.tile_row[t] : 0;
  int32_t rs = 0, re = 0;
  bool comp = false;
  float acc[V], cmp[V];
  zero_vec(acc);
  zero_vec(cmp);
  int32_t rbase = 0, pw = 0;

  // pos[row]: from the chunk's LDS window (rows rbase .. rbase+pw-1)
  auto P = [&](int32_t row) -> int32_t {
    const int32_t i = row - rbase;
    if constexpr (kWide) return i < pw ? s_pos[i] : pos[row];
    return s_pos[i];
  };
  auto flush = [&]() {
    if (!colok) return;
    const bool before = rs < s;
    const bool after = re > e;
    if (comp) {
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] += cmp[j];
    }
    if (before || after) {
      // carries: read back by the fix-up
      store_vec<V>((before ? a.carry_head : a.carry_tail) + (int64_t)t * a.N + col, acc);
    } else {
      // C rows are never read back here: non-temporal, to leave L2 to B
      store_vec_nt<V>(a.C + (int64_t)r * a.ldc + col, acc);
    }
  };
  auto next_row = [&](int32_t pp) {
    flush();
    ++r;
    while (P(r + 1) <= pp) ++r;  // skip empty rows (zeroed elsewhere)
    rs = P(r);
    re = P(r + 1);
    comp = (re - rs) > a.comp_threshold;
    zero_vec(acc);
    zero_vec(cmp);
  };

  using R = typename RawT<DT, V>::type;
  const int32_t nchunks = (a.tile + CH - 1) / CH;  // uniform over the block
  for (int32_t c = 0; c < nchunks; ++c) {
    const int32_t cs = s + c * CH;
    const int32_t cn = max(0, min(e - cs, CH));
    if (cn > 0) {
      // the tile's entries are read once: non-temporal (power-law 262144^2,
      // N = 128, FP32, with the C stores above: 343 vs 352 us per step on one
      // host, the same time within 1 % on others)
      for (int32_t i = q; i < cn; i += Q) {
        s_crd[i] = __builtin_nontemporal_load(crd + cs + i);
        s_val[i] = load_val_stream<DT>(vals, cs + i);
      }
      rbase = r;
      pw = min(PW, a.M + 1 - rbase);
      for (int32_t i = q; i < pw; i += Q) s_pos[i] = pos[rbase + i];
    }
    wave_sync();
    if (cn > 0) {
      if (c == 0) {
        rs = P(r);
        re = P(r + 1);
        comp = (re - rs) > a.comp_threshold;
      }
      // Double-buffered: batch k+1's gathers are in flight while batch k is
      // consumed, so the memory pipe never drains between batches. Every
      // batch issues exactly kSrBatch loads, with no branch around them
      // (entries past the chunk and columns past N re-load the chunk's last
      // entry / column 0, and are not consumed): the number of loads in flight
      // is then the same on every path, so the wait before consuming a batch
      // only covers that batch and leaves the next one in flight.
      R ra[kSrBatch], rb[kSrBatch];
      const int32_t col_ld = colok ? col : 0;
      auto issue = [&](R (&buf)[kSrBatch], int32_t b0) {
#pragma unroll
        for (int u = 0; u < kSrBatch; ++u) {
          const int32_t i = min(b0 + u, cn - 1);
          buf[u] = load_row_raw<DT, V, kOff32>(B, s_crd[i], ldb, col_ld);
        }
      };
      auto consume = [&](const R (&buf)[kSrBatch], int32_t b0) {
        const int32_t nb = min(kSrBatch, cn - b0);
#pragma unroll
        for (int u = 0; u < kSrBatch; ++u) {
          if (u < nb) {
            const int32_t pp = cs + b0 + u;
            if (pp == re) next_row(pp);
            if (colok) {
              const float v = s_val[b0 + u];
              float b[V];
              unpack<DT, V>(buf[u], b);
              if (comp) {
#pragma unroll
                for (int j = 0; j < V; ++j) neumaier_add(acc[j], cmp[j], v, b[j]);
              } else {
#pragma unroll
                for (int j = 0; j < V; ++j) acc[j] = fmaf(v, b[j], acc[j]);
              }
            }
          }
        }
      };
      // the final pair of batches issues nothing after it
      const int32_t last = (cn - 1) / (2 * kSrBatch) * (2 * kSrBatch);
      issue(ra, 0);
      for (int32_t b0 = 0; b0 < last; b0 += 2 * kSrBatch) {
        issue(rb, b0 + kSrBatch);
        consume(ra, b0);
        issue(ra, b0 + 2 * kSrBatch);
        consume(rb, b0 + kSrBatch);
      }
      issue(rb, last + kSrBatch);
      consume(ra, last);
      consume(rb, last + kSrBatch);
    }
    wave_sync();
  }
  if (tok) flush();
}

// A lane group's lookups stay inside its LDS window when its tile's rows
// r0 = tile_row[t] .. r1 = tile_row[t+1] satisfy r1 - r0 <= PW - 2: each chunk's
// window starts at a row >= r0, and the walk reads pos at rows up to
// min(r1, M-1) + 1 (the row after the one holding the tile's last entry).
template <int DT, int V>
__device__ __forceinline__ void eb_sr_body(const KArgs &a, int32_t blk,
                                           int32_t panel, int32_t *smem) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t t = blk * (kWavesPerBlock << (6 - lq)) + (wave << (6 - lq)) + (lane >> lq);
  const int32_t PW = eb_sr_chunk(lq) + kPosSlack;
  const bool wide = t < a.ntiles && a.tile_row[t + 1] - a.tile_row[t] > PW - 2;
  const bool any_wide = __ballot(wide) != 0;
  if (a.b_off32) {
    if (any_wide)
      eb_sr_walk<DT, V, true, true>(a, blk, panel, smem);
    else
      eb_sr_walk<DT, V, false, true>(a, blk, panel, smem);
  } else {
    if (any_wide)
      eb_sr_walk<DT, V, true, false>(a, blk, panel, smem);
    else
      eb_sr_walk<DT, V, false, false>(a, blk, panel, smem);
  }
}

// Row of every entry p = cs + i (i < cn) of a staged eb_pr chunk, into
// s_row[i]: the last row j with pos[j] <= p (the non-empty row holding p),
// by bisection in the LDS window s_pos = pos[rbase .. rbase + pw); an entry
// beyond the window's last row continues with a bisection of pos in memory
// up to r_hi, the last row the tile can touch. One wave; lane = lane id.
__device__ __forceinline__ void chunk_entry_rows(const int32_t *__restrict__ pos,
                                                 const int32_t *s_pos, int32_t pw,
                                                 int32_t rbase, int32_t r_hi,
                                                 int32_t cs, int32_t cn, int lane,
                                                 int32_t *s_row) {
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
}

// eb_pr over tile block `blk` (one tile per wave), column panel `panel`;
// smem holds eb_pr_lds_bytes_fixed() bytes.
template <int DT, int V, bool kWide, bool kOff32>
__device__ __forceinline__ void eb_pr_walk(const KArgs &a, int32_t blk,
                                           int32_t panel, int32_t *smem) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t P = 64 >> lq;  // entry slots per pass
  const int32_t t = blk * kWavesPerBlock + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  const int32_t q = lane & ((1 << lq) - 1);
  const int32_t sl = lane >> lq;
  constexpr int32_t CH = kPrChunk;
  constexpr int32_t PW = kPrChunk + kPosSlack;
  int32_t *s_crd = smem + wave * (4 * CH + kPosSlack);
  float *s_val = reinterpret_cast<float *>(s_crd + CH);
  int32_t *s_row = s_crd + 2 * CH;
  int32_t *s_pos = s_crd + 3 * CH;

  const int32_t *__restrict__ pos = a.pos;
  const int32_t *__restrict__ crd = a.crd;
  const void *__restrict__ vals = a.vals;
  const void *__restrict__ B = a.B;
  const int64_t ldb = a.ldb;
  const int32_t col = ((panel << lq) + q) * V;
  const bool colok = col < a.N;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  // Upper bound on the rows of this tile: the row holding entry e (M when
  // e == nnz) is at or after the row holding entry e - 1.
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  int32_t r = a.tile_row[t];
  int32_t rbase = r, pw = 0;
  // pos[row] from the chunk's LDS window (memory past it only when kWide)
  auto Pos = [&](int32_t row) -> int32_t {
    const int32_t i = row - rbase;
    if constexpr (kWide) return i < pw ? s_pos[i] : pos[row];
    return s_pos[i];
  };

  float open[V];
  zero_vec(open);
  int32_t open_row = -1;
  using R = typename RawT<DT, V>::type;

  for (int32_t cs = s; cs < e; cs += CH) {
    const int32_t cn = min(CH, e - cs);
    for (int32_t i = lane; i < cn; i += 64) {
      s_crd[i] = crd[cs + i];
      s_val[i] = load_val<DT>(vals, cs + i);
    }
    rbase = r;
    pw = min(PW, a.M + 1 - rbase);
    for (int32_t i = lane; i < pw; i += 64) s_pos[i] = pos[rbase + i];
    wave_sync();
    chunk_entry_rows(pos, s_pos, pw, rbase, r_hi, cs, cn, lane, s_row);
    wave_sync();
    for (int32_t pb = 0; pb < cn; pb += P * kPrPasses) {
      int32_t rid[kPrPasses];
      float vv[kPrPasses];
      R raw[kPrPasses];
#pragma unroll
      for (int u = 0; u < kPrPasses; ++u) {
        // every pass issues its load, with no branch around it (slots past
        // the chunk re-load its last entry, columns past N column 0), so the
        // wait before a pass covers only that pass's load
        const int32_t i = pb + u * P + sl;
        const bool ok = i < cn;
        rid[u] = ok ? s_row[i] : INT_MAX;
        vv[u] = ok ? s_val[i] : 0.0f;
        raw[u] = load_row_raw<DT, V, kOff32>(B, s_crd[min(i, cn - 1)], ldb,
                                             colok ? col : 0);
      }
#pragma unroll
      for (int u = 0; u < kPrPasses; ++u) {
        const int32_t i0 = pb + u * P;
        if (i0 >= cn) break;  // uniform over the wave
        float prod[V];
        zero_vec(prod);
        if (rid[u] != INT_MAX && colok) {
          float b[V];
          unpack<DT, V>(raw[u], b);
#pragma unroll
          for (int j = 0; j < V; ++j) prod[j] = vv[u] * b[j];
        }
        const int32_t my = rid[u];
        // Segmented suffix scan: slot sl accumulates slot sl+d of its row.
        for (int32_t d = 1; d < P; d <<= 1) {
          const int32_t tr = __shfl_down(my, d << lq, 64);
          float tv[V];
#pragma unroll
          for (int j = 0; j < V; ++j) tv[j] = __shfl_down(prod[j], d << lq, 64);
          if (sl + d < P && tr == my) {
#pragma unroll
            for (int j = 0; j < V; ++j) prod[j] += tv[j];
          }
        }
        const int32_t prev = __shfl_up(my, 1 << lq, 64);
        const bool head = my != INT_MAX && (sl == 0 || prev != my);
        if (sl == 0 && my == open_row) {
#pragma unroll
          for (int j = 0; j < V; ++j) prod[j] = open[j] + prod[j];
        }
        const uint64_t hm = __ballot(head);
        const int32_t last_head = (63 - __clzll((long long)hm)) >> lq;
        const int32_t last_row = __shfl(my, last_head << lq, 64);
        const int32_t pass_end = cs + min(i0 + P, cn);
        const bool last_open = pass_end < e && Pos(last_row + 1) > pass_end;
        if (head && colok && !(last_open && sl == last_head)) {
          const bool before = Pos(my) < s;
          const bool after = Pos(my + 1) > e;
          float *dst = before  ? a.carry_head + (int64_t)t * a.N + col
                       : after ? a.carry_tail + (int64_t)t * a.N + col
                               : a.C + (int64_t)my * a.ldc + col;
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
    r = s_row[cn - 1];
    wave_sync();
  }
}

// The tile's row lookups stay in the LDS window when its rows r0 = tile_row[t]
// .. r1 = tile_row[t+1] satisfy r1 - r0 <= PW - 2 (as for eb_sr).
template <int DT, int V>
__device__ __forceinline__ void eb_pr_body(const KArgs &a, int32_t blk,
                                           int32_t panel, int32_t *smem) {
  const int32_t t = blk * kWavesPerBlock + (threadIdx.x >> 6);
  if (t >= a.ntiles) return;  // uniform over the wave
  const bool wide = a.tile_row[t + 1] - a.tile_row[t] > kPrChunk + kPosSlack - 2;
  if (a.b_off32) {
    if (wide)
      eb_pr_walk<DT, V, true, true>(a, blk, panel, smem);
    else
      eb_pr_walk<DT, V, false, true>(a, blk, panel, smem);
  } else {
    if (wide)
      eb_pr_walk<DT, V, true, false>(a, blk, panel, smem);
    else
      eb_pr_walk<DT, V, false, false>(a, blk, panel, smem);
  }
}

// C[r][col..col+V) = carry_tail[t0] + carry_head[t0+1] + ... + carry_head[t],
// added in tile order. A long row spans many tiles: its heads are loaded
// eight at a time (all in flight together), then added one by one.
template <int V>
__device__ __forceinline__ void combine_carries(const KArgs &a, int32_t t0,
                                                int32_t t, int32_t r,
                                                int32_t col) {
  float acc[V];
  load_f32vec<V>(a.carry_tail + (int64_t)t0 * a.N + col, acc);
  constexpr int kH = 8;
  int32_t tt = t0 + 1;
  for (; tt + kH - 1 <= t; tt += kH) {
    float h[kH][V];
#pragma unroll
    for (int k = 0; k < kH; ++k)
      load_f32vec<V>(a.carry_head + (int64_t)(tt + k) * a.N + col, h[k]);
#pragma unroll
    for (int k = 0; k < kH; ++k)
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] += h[k][j];
  }
  for (; tt <= t; ++tt) {
    float h[V];
    load_f32vec<V>(a.carry_head + (int64_t)tt * a.N + col, h);
#pragma unroll
    for (int j = 0; j < V; ++j) acc[j] += h[j];
  }
  store_vec<V>(a.C + (int64_t)r * a.ldc + col, acc);
}

// Per tile t (one lane group). With a.fix_t0 (standalone calls, where
// k_tile_rows_fix already zeroed the empty rows): if a row that started in
// tile t0 < t ends in t, write it from its carries. Without (batched calls,
// whose partition is made once and reused): zero the empty rows of t's row
// slice, then find such a row from tile_row and pos and write it the same way.
template <int V>
__device__ __forceinline__ void fixup_body(const KArgs &a, int32_t blk,
                                           int32_t panel) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int lq = a.lq;
  const int32_t Q = 1 << lq;
  const int32_t q = lane & (Q - 1);
  const int32_t t = ((blk * kWavesPerBlock + wave) << (6 - lq)) + (lane >> lq);
  if (t >= a.ntiles) return;
  const int32_t col = ((panel << lq) + q) * V;
  if (a.fix_t0) {
    const int32_t t0 = a.fix_t0[t];
    if (t0 < 0 || col >= a.N) return;
    // eb_cs has already written the rows it could sum from its LDS slots
    if (a.cs_flags && cs_row_folded(a.cs_flags, a.cs_group, t0, t)) return;
    combine_carries<V>(a, t0, t, a.tile_row[t], col);
    return;
  }
  {
    const int64_t r0 = (int64_t)t * a.rows_per_slice;
    const int64_t r1 = r0 + a.rows_per_slice < a.M ? r0 + a.rows_per_slice : a.M;
    if (r0 < r1) zero_empty_rows<V>(a, (int32_t)r0, (int32_t)r1, q, Q, panel);
  }
  if (t == 0 || col >= a.N) return;
  const int32_t *__restrict__ pos = a.pos;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  const int32_t r = a.tile_row[t];
  const int32_t rs = pos[r], re = pos[r + 1];
  if (!(rs < s && re <= e)) return;
  combine_carries<V>(a, rs / a.tile, t, r, col);
}

}  // namespace segtaco

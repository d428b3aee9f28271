// k_eb.hip — host dispatch of the standalone element-balanced kernels (one
// object per storage type, k_eb_inst.hip), their carry fix-up (bodies in
// sched_eb.h), the K21 row partition, the tile-row scatter and a zero fill.
#include "sched_eb.h"

namespace segtaco {

size_t eb_sr_lds_bytes(const KArgs &a) { return eb_sr_lds_bytes_lq(a.lq); }
size_t eb_pr_lds_bytes(const KArgs &) { return eb_pr_lds_bytes_fixed(); }

template <int V>
__global__ __launch_bounds__(kBlockThreads) void k_fixup(const KArgs a) {
  int32_t blk, panel;
  if (!decode_block(a.panels, a.nblocks, blk, panel)) return;
  fixup_body<V>(a, blk, panel);
}

__global__ __launch_bounds__(kBlockThreads) void k_block_starts(
    const int32_t *__restrict__ pos, int32_t M, int32_t T, int32_t nb,
    int32_t *__restrict__ out, int32_t *__restrict__ rows) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nb) return;
  const int64_t target = b * (int64_t)T;
  int32_t r = search_before(pos, 0, M, target);
  if (out) out[b] = r;
  if (rows) {
    if (r < M)
      while ((int64_t)pos[r + 1] <= target) ++r;
    rows[b] = r;
  }
}

// rows[b] = the row holding entry b*T (the repaired K21 row), by scatter:
// non-empty row r holds entries [pos[r], pos[r+1]), i.e. the tile starts
// b*T for b in [ceil(pos[r]/T), floor((pos[r+1]-1)/T)]; rows[ntiles] = M.
// One coalesced pass over pos instead of a bisection per tile.
__global__ __launch_bounds__(kBlockThreads) void k_tile_rows(
    const int32_t *__restrict__ pos, int32_t M, int32_t T, int32_t ntiles,
    int32_t *__restrict__ rows) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < M) {
    const int32_t p0 = pos[r], p1 = pos[r + 1];
    if (p0 < p1) {
      const int32_t b1 = (p1 - 1) / T;
      for (int32_t b = (int32_t)(((int64_t)p0 + T - 1) / T); b <= b1; ++b)
        rows[b] = (int32_t)r;
    }
  } else if (r == M) {
    rows[ntiles] = M;
  }
}

// k_tile_rows plus the fix-up's bookkeeping, for a standalone call:
// fix_t0[b] = the tile where the row ending in tile b started, if it started
// before tile b (the fix-up then sums that row's carries), else -1; and every
// empty row of C is written as zeros here, since no later kernel of the call
// writes it. Each of tiles ceil(p0/T)..floor((p1-1)/T) starts inside row r,
// so every tile's entries are written exactly once. One thread per row; a
// wave zeroes its empty rows together, 64 lanes x V floats across a row.
template <int V>
__global__ __launch_bounds__(kBlockThreads) void k_tile_rows_fix(
    const int32_t *__restrict__ pos, int32_t M, int32_t T, int32_t ntiles,
    int32_t *__restrict__ rows, int32_t *__restrict__ fix_t0,
    float *__restrict__ C, int32_t N, int64_t ldc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool empty = false;
  if (r < M) {
    const int32_t p0 = pos[r], p1 = pos[r + 1];
    if (p0 < p1) {
      const int32_t b1 = (p1 - 1) / T;
      for (int32_t b = (int32_t)(((int64_t)p0 + T - 1) / T); b <= b1; ++b) {
        rows[b] = (int32_t)r;
        fix_t0[b] = (b == b1 && (int64_t)p0 < (int64_t)b * T) ? p0 / T : -1;
      }
    } else {
      empty = true;
    }
  } else if (r == M) {
    rows[ntiles] = M;
  }
  if (C == nullptr) return;  // tile rows alone (uniform over the launch)
  unsigned long long m = __ballot(empty);
  const int lane = threadIdx.x & 63;
  const int64_t wave_row0 = r - lane;
  float z[V];
  zero_vec(z);
  while (m) {
    const int i = __ffsll(m) - 1;
    m &= m - 1;
    float *row = C + (wave_row0 + i) * ldc;
    for (int32_t j = lane * V; j < N; j += 64 * V) store_vec<V>(row + j, z);
  }
}

// The row of every stored entry as eb_pr's walk derives it (one wave per
// tile, the same chunking, row-pointer window and bisection as
// eb_pr_walk), written to rows[p]: exposes those per-entry row ids so they
// can be compared with the row each entry belongs to.
__global__ __launch_bounds__(kBlockThreads) void k_entry_rows(
    const int32_t *__restrict__ pos, int32_t M, int32_t nnz, int32_t T, int32_t ntiles,
    const int32_t *__restrict__ tile_row, int32_t *__restrict__ rows) {
  __shared__ int32_t sm[kWavesPerBlock][2 * kPrChunk + kPosSlack];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t t = (int32_t)blockIdx.x * kWavesPerBlock + wave;
  if (t >= ntiles) return;  // uniform over the wave
  int32_t *s_row = sm[wave];
  int32_t *s_pos = sm[wave] + kPrChunk;
  constexpr int32_t PW = kPrChunk + kPosSlack;
  const int32_t s = t * T;
  const int32_t e = (int32_t)min((int64_t)nnz, (int64_t)s + T);
  const int32_t r_hi = min(M - 1, tile_row[t + 1]);
  int32_t r = tile_row[t];
  for (int32_t cs = s; cs < e; cs += kPrChunk) {
    const int32_t cn = min(kPrChunk, e - cs);
    const int32_t rbase = r;
    const int32_t pw = min(PW, M + 1 - rbase);
    for (int32_t i = lane; i < pw; i += 64) s_pos[i] = pos[rbase + i];
    wave_sync();
    chunk_entry_rows(pos, s_pos, pw, rbase, r_hi, cs, cn, lane, s_row);
    wave_sync();
    for (int32_t i = lane; i < cn; i += 64) rows[cs + i] = s_row[i];
    r = s_row[cn - 1];
    wave_sync();
  }
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

hipError_t launch_eb_sr(int dt, int vec, const KArgs &a, hipStream_t s) {
  switch (dt) {
    case 0: return launch_eb_sr_t<0>(vec, a, s);
    case 1: return launch_eb_sr_t<1>(vec, a, s);
    case 2: return launch_eb_sr_t<2>(vec, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_eb_pr(int dt, int vec, const KArgs &a, hipStream_t s) {
  switch (dt) {
    case 0: return launch_eb_pr_t<0>(vec, a, s);
    case 1: return launch_eb_pr_t<1>(vec, a, s);
    case 2: return launch_eb_pr_t<2>(vec, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fixup(int vec, const KArgs &a, hipStream_t s) {
  // One lane group of Q lanes per tile, whatever the main schedule.
  KArgs f = a;
  const int64_t groups_per_block = (int64_t)kWavesPerBlock << (6 - a.lq);
  f.nblocks = (int32_t)((a.ntiles + groups_per_block - 1) / groups_per_block);
  const int64_t grid = grid_for(f.panels, f.nblocks);
  if (grid <= 0) return hipSuccess;
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
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
                               int32_t num_blocks, int32_t *out, int32_t *rows,
                               hipStream_t s) {
  const int64_t n = (int64_t)num_blocks + 1;
  const int64_t grid = (n + kBlockThreads - 1) / kBlockThreads;
  hipLaunchKernelGGL(k_block_starts, dim3((unsigned)grid), dim3(kBlockThreads),
                     0, s, pos, M, T, num_blocks, out, rows);
  return hipGetLastError();
}

hipError_t launch_tile_rows(const int32_t *pos, int32_t M, int32_t T,
                            int32_t ntiles, int32_t *rows, hipStream_t s) {
  const int64_t n = (int64_t)M + 1;
  const int64_t grid = (n + kBlockThreads - 1) / kBlockThreads;
  hipLaunchKernelGGL(k_tile_rows, dim3((unsigned)grid), dim3(kBlockThreads), 0,
                     s, pos, M, T, ntiles, rows);
  return hipGetLastError();
}

hipError_t launch_tile_rows_fix(int vec, const KArgs &a, hipStream_t s) {
  const int64_t n = (int64_t)a.M + 1;
  const dim3 g((unsigned)((n + kBlockThreads - 1) / kBlockThreads)), b(kBlockThreads);
  int32_t *rows = const_cast<int32_t *>(a.tile_row);
  int32_t *fix = const_cast<int32_t *>(a.fix_t0);
  switch (vec) {
    case 1: hipLaunchKernelGGL((k_tile_rows_fix<1>), g, b, 0, s, a.pos, a.M, a.tile, a.ntiles, rows, fix, a.C, a.N, a.ldc); break;
    case 2: hipLaunchKernelGGL((k_tile_rows_fix<2>), g, b, 0, s, a.pos, a.M, a.tile, a.ntiles, rows, fix, a.C, a.N, a.ldc); break;
    case 4: hipLaunchKernelGGL((k_tile_rows_fix<4>), g, b, 0, s, a.pos, a.M, a.tile, a.ntiles, rows, fix, a.C, a.N, a.ldc); break;
    case 8: hipLaunchKernelGGL((k_tile_rows_fix<8>), g, b, 0, s, a.pos, a.M, a.tile, a.ntiles, rows, fix, a.C, a.N, a.ldc); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_entry_rows(const int32_t *pos, int32_t M, int32_t nnz, int32_t T,
                             const int32_t *tile_row, int32_t *rows, hipStream_t s) {
  const int64_t ntiles = ((int64_t)nnz + T - 1) / T;
  if (ntiles <= 0) return hipSuccess;
  const int64_t grid = (ntiles + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL(k_entry_rows, dim3((unsigned)grid), dim3(kBlockThreads), 0, s, pos, M,
                     nnz, T, (int32_t)ntiles, tile_row, rows);
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

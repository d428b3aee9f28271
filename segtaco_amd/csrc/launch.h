// launch.h — internal kernel argument block and launchers (not public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace segtaco {

constexpr int kBlockThreads = 256;  // 4 waves of 64 lanes
constexpr int kWavesPerBlock = kBlockThreads / 64;

// Everything a schedule kernel needs, passed by value (kernarg segment).
struct KArgs {
  int32_t M, K, N, nnz;
  const int32_t *pos;
  const int32_t *crd;
  const void *vals;  // nullptr: every stored value is 1.0
  const void *B;
  int64_t ldb;
  float *C;
  int64_t ldc;
  int32_t lq;              // log2(col_lanes): lanes per B-row segment
  int32_t lg;              // log2(group): *_pr nnz slots per row / pass
  int32_t tile;            // eb_*: nnz per tile
  int32_t ntiles;          // eb_*: ceil(nnz / tile)
  int32_t comp_threshold;  // rows longer than this are compensated
  int32_t panels;          // column panels: ceil(N / (col_lanes * vec))
  int32_t nblocks;         // row / tile blocks before panel expansion
  int32_t rows_per_slice;  // eb_*: rows per tile for empty-row zero fill
  const int32_t *part;     // eb_*: K21 partition, ntiles + 1 entries
  const int32_t *tile_row; // eb_*: row holding each tile's first entry
  float *carry_head;       // eb_*: [ntiles][N] partial of a continued row
  float *carry_tail;       // eb_*: [ntiles][N] partial of a row left open
};

hipError_t launch_rb_sr(int dt, int vec, const KArgs &a, hipStream_t s);
hipError_t launch_rb_pr(int dt, int vec, const KArgs &a, hipStream_t s);
hipError_t launch_eb_sr(int dt, int vec, const KArgs &a, hipStream_t s);
hipError_t launch_eb_pr(int dt, int vec, const KArgs &a, hipStream_t s);
// Combines carry partials of rows that span eb tiles (vec in floats).
hipError_t launch_fixup(int vec, const KArgs &a, hipStream_t s);
// out[b] = K21 search for b*T (bit-exact); rows (optional) gets the row that
// holds entry b*T after the reference's repair step (M when b*T >= nnz).
hipError_t launch_block_starts(const int32_t *pos, int32_t M, int32_t T,
                               int32_t num_blocks, int32_t *out, int32_t *rows,
                               hipStream_t s);
// Dynamic LDS bytes a launch of k_eb_sr / k_eb_pr uses.
size_t eb_sr_lds_bytes(const KArgs &a);
size_t eb_pr_lds_bytes(const KArgs &a);
hipError_t launch_fill_zero(float *C, int32_t M, int32_t N, int64_t ldc,
                            hipStream_t s);

}  // namespace segtaco

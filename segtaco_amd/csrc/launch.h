// launch.h — internal kernel argument block and launchers (not public ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "segtaco/spmm.h"

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
  // eb_*: per tile, the tile where the row that ends in it started, if that
  // row began in an earlier tile, else -1 (written by k_tile_rows_fix, which
  // also zeroes the empty rows of C); null: the fix-up derives both from pos
  const int32_t *fix_t0;
  // eb_cs: per tile bit 2 = the tile walks (k_cs_sort), and tiles per unit.
  // The rows that cs_row_folded names were written by k_eb_cs itself, so the
  // fix-up skips them. Null / 0 for the other schedules.
  const uint8_t *cs_flags;
  int32_t cs_group;
  float *carry_head;       // eb_*: [ntiles][N] partial of a continued row
  float *carry_tail;       // eb_*: [ntiles][N] partial of a row left open
  // K, ldb < 2^24 and every byte offset into B fits 32 bits: gathers form
  // their offsets in 32 bits (load_row_raw<.., true>)
  int32_t b_off32;
};

// Whether a K x ldb matrix of elem-byte elements qualifies for 32-bit
// gather offsets (see KArgs::b_off32).
inline int32_t b_off32_ok(int32_t K, int64_t ldb, int32_t elem) {
  const int64_t lim24 = (int64_t)1 << 24;
  return K < lim24 && ldb < lim24 && ((int64_t)K + 1) * ldb * elem < ((int64_t)1 << 32);
}

hipError_t launch_rb_sr(int dt, int vec, const KArgs &a, hipStream_t s);
hipError_t launch_rb_pr(int dt, int vec, const KArgs &a, hipStream_t s);
hipError_t launch_eb_sr(int dt, int vec, const KArgs &a, hipStream_t s);
hipError_t launch_eb_pr(int dt, int vec, const KArgs &a, hipStream_t s);
// The same for one storage type DT (k_eb_inst.hip, one object per DT).
template <int DT>
hipError_t launch_eb_sr_t(int vec, const KArgs &a, hipStream_t s);
template <int DT>
hipError_t launch_eb_pr_t(int vec, const KArgs &a, hipStream_t s);
// Combines carry partials of rows that span eb tiles (vec in floats).
hipError_t launch_fixup(int vec, const KArgs &a, hipStream_t s);
// out[b] = K21 search for b*T (bit-exact); rows (optional) gets the row that
// holds entry b*T after the reference's repair step (M when b*T >= nnz).
hipError_t launch_block_starts(const int32_t *pos, int32_t M, int32_t T,
                               int32_t num_blocks, int32_t *out, int32_t *rows,
                               hipStream_t s);
// rows[b] = row holding entry b*T for b < ntiles, rows[ntiles] = M: the same
// values as launch_block_starts' `rows`, from one pass over pos.
hipError_t launch_tile_rows(const int32_t *pos, int32_t M, int32_t T,
                            int32_t ntiles, int32_t *rows, hipStream_t s);
// launch_tile_rows for a standalone SpMM call: also writes a.fix_t0 and zeroes
// the empty rows of a.C (vec floats per store), so that the fix-up only sums
// the carries of rows that span tiles.
hipError_t launch_tile_rows_fix(int vec, const KArgs &a, hipStream_t s);
// rows[p] for every stored entry p: its row as eb_pr's walk over tiles of T
// entries derives it (tile_row as from launch_tile_rows).
hipError_t launch_entry_rows(const int32_t *pos, int32_t M, int32_t nnz, int32_t T,
                             const int32_t *tile_row, int32_t *rows, hipStream_t s);
// Dynamic LDS bytes a launch of k_eb_sr / k_eb_pr uses.
size_t eb_sr_lds_bytes(const KArgs &a);
size_t eb_pr_lds_bytes(const KArgs &a);
hipError_t launch_fill_zero(float *C, int32_t M, int32_t N, int64_t ldc,
                            hipStream_t s);

// One product of a batch, as the batched kernels see it.
struct BatchDesc {
  KArgs a;
  int32_t schedule;    // SEGTACO_RB_SR .. SEGTACO_EB_PR
  int32_t vec;         // elements per lane per gather
  int64_t unit_begin;  // first unit of this product in the persistent launch
  int64_t fix_begin;   // first unit of this product in the batched fix-up
};

// Persistent launch over `total` units (prefix[i] = first unit of product i,
// non-decreasing, n entries); workgroups take units from *counter, which the
// launcher resets first with a one-thread kernel unless reset_counter is
// false (the caller has already put a zero there on the same stream).
// grid = persistent workgroups (batch_main_grid), lds = dynamic LDS; sv: some
// product takes the N = 1 form of eb_pr (batch_uses_spmv), which selects the
// kernel that carries that body.
hipError_t launch_batch_main(int dt, const BatchDesc *d, int32_t n,
                             const int64_t *prefix, int64_t total,
                             unsigned long long *counter, int32_t grid,
                             size_t lds, hipStream_t s, bool reset_counter = true,
                             bool sv = false);
hipError_t launch_batch_fixup(const BatchDesc *d, int32_t n,
                              const int64_t *fix_prefix, int64_t total,
                              hipStream_t s);
// Fills D (and the product's unit counts, dynamic LDS need and, for
// element-balanced plans, its workspace layout) for a resolved plan. The
// element-balanced pointers (part, tile_row, carries) are left null.
struct EbLayout;
void fill_batch_desc(const segtaco_csr_t &A, const void *B, int64_t ldb,
                     float *C, int64_t ldc, int32_t N,
                     const segtaco_plan_t &plan, BatchDesc *D, int64_t *units,
                     int64_t *fix_units, size_t *lds, EbLayout *lay);
// Workgroups of the persistent launch that fit the device at once.
hipError_t batch_main_grid(int dt, size_t lds, int32_t *grid, bool sv = false);

}  // namespace segtaco

// cs.h — internal argument block, constants and launchers of the eb_cs
// schedule (k_cs.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace segtaco {

constexpr int kCsThreads = 512;           // main kernel: 8 waves
constexpr int kCsWaves = kCsThreads / 64;
constexpr int kCsPrepThreads = 256;       // prep kernel: 4 waves
constexpr int32_t kCsLanes = 8;           // lanes per entry (Q)
constexpr int32_t kCsMaxGroup = kCsWaves * (64 / kCsLanes);  // tiles per unit
constexpr int32_t kCsMaxTile = 512;       // entries per tile
constexpr uint16_t kCsHead = 0xFFFF;      // slot code: row began before tile
constexpr uint16_t kCsTail = 0xFFFE;      // slot code: row continues after
constexpr size_t kCsLdsBudget = 152 * 1024;  // accumulator bytes per unit

struct CsArgs {
  int32_t M, K, N, nnz;
  const int32_t *pos;
  const int32_t *crd;
  const void *vals;     // nullptr: values are 1.0
  const void *B;
  int64_t ldb;
  float *C;
  int64_t ldc;
  int32_t tile, ntiles;
  int32_t group;        // tiles per unit
  int32_t units, rounds;  // units = ceil(ntiles / group); units per block
  int32_t cap;          // LDS accumulator slots per unit
  int32_t elem;         // bytes per stored value
  const int32_t *tile_row;
  float *carry_head, *carry_tail;
  // written by k_cs_prep
  int32_t *s_col;       // nnz: columns, each tile sorted
  void *s_val;          // nnz: values in the same order (elem bytes)
  uint16_t *s_slot;     // nnz: LDS slot, kCsHead or kCsTail
  int32_t *slot_rows;   // units x cap: row of each interior slot
  int32_t *n_slots;     // units
  int32_t *tile_count;  // ntiles: interior rows of each tile
  uint8_t *tile_flags;  // ntiles: bit 0 head row, bit 1 tail row
};

hipError_t launch_cs_prep(const CsArgs &a, hipStream_t s);
hipError_t launch_eb_cs(int dt, int vec, const CsArgs &a, hipStream_t s);
size_t cs_lds_bytes(int32_t vec, int32_t cap);

}  // namespace segtaco

// cs.h — internal argument block, constants and launchers of the eb_cs
// schedule (k_cs.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace segtaco {

constexpr int kCsThreads = 512;           // main kernel: 8 waves
constexpr int kCsWaves = kCsThreads / 64;
constexpr int kCsPrepThreads = 256;       // prep kernels: one tile per wave
constexpr int32_t kCsLanes = 8;           // lanes per entry (Q)
constexpr int32_t kCsMaxGroup = kCsWaves * (64 / kCsLanes);  // tiles per unit
constexpr int32_t kCsMaxTile = 512;       // entries per tile
// Accumulator bytes per unit: a whole CU's LDS on gfx950. The slots' row ids
// (4 bytes each) share it, so the slot count is also capped at the device's
// LDS / (slot bytes + 4): 1241 slots of 128 bytes.
constexpr size_t kCsLdsBudget = 160 * 1024;
// One record per sorted entry, read with one load: {(column << kCsCodeBits)
// | slot, value bits} in 64 bits when every column is below
// 2^(32 - kCsCodeBits) (slots are below 2^kCsCodeBits), otherwise {column,
// slot, value bits} in 96 bits. Value bits: the FP32 word, or the
// 16-bit storage word zero-extended; 1.0 when A has no values.
constexpr int kCsCodeBits = 12;

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
  int32_t packed;       // 64-bit records (else 96)
  uint32_t one_bits;    // 1.0 in the storage type (FP32, BF16 or FP16 bits)
  const int32_t *tile_row;
  float *carry_head, *carry_tail;
  // written by the preparation kernels
  void *s_rec;          // nnz records (column, code, value). Sweeping tiles:
                        // sorted by (column, position), code = LDS slot
                        // (2g / 2g + 1 = head / tail row of the unit's tile
                        // g, from 2G on the unit's interior rows). Walking
                        // tiles: CSR order, code 0 head, 1 tail, 2 + k the
                        // tile's k-th interior row
  int32_t *slot_rows;   // units x cap: row of each interior slot
  int32_t *int_rows;    // ntiles x tile: row of each tile's k-th interior
                        // row (one that starts and ends inside the tile)
  uint16_t *codes;      // nnz: row of each entry by its place in the tile
                        // (0 head, 1 tail, 2 + k the k-th interior row)
  int32_t *n_slots;     // units: slots used (2G + swept interior rows)
  int32_t *tile_count;  // ntiles: interior rows of each tile
  uint8_t *tile_flags;  // ntiles: bit 0 head row, bit 1 tail row, bit 2
                        // the tile walks
};

// Whether k_eb_cs writes the row that spans tiles t0 < t1 itself, summing
// its LDS tail / head slots in the fix-up's order (tail of t0, then the head
// of t1), instead of leaving it to the carries and k_fixup: a row over two
// neighbouring tiles of one unit of `group` tiles, neither walking (flags:
// bit 2 per walking tile). Rows over more tiles keep the carries: summing
// them in the kernel is a serial chain of slot reads that holds the unit's
// other waves at the panel barrier (64-entry tiles on a 65536^2 power-law
// product: 104 us per call against 90 us with every spanning row folded).
__device__ inline bool cs_row_folded(const uint8_t *flags, int32_t group, int32_t t0,
                                     int32_t t1) {
  return t1 == t0 + 1 && t0 / group == t1 / group && !((flags[t0] | flags[t1]) & 4);
}

hipError_t launch_cs_prep(const CsArgs &a, hipStream_t s);
hipError_t launch_eb_cs(int dt, int vec, const CsArgs &a, hipStream_t s);
size_t cs_lds_bytes(int32_t vec, int32_t cap);

}  // namespace segtaco

// sddmm.h — internal argument block and launcher of the SDDMM kernel
// (k_sddmm.hip); the public C-ABI is include/segtaco/sddmm.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace segtaco {

constexpr int32_t kSdTile = 512;  // entries per wave (one tile)

struct SddmmArgs {
  int32_t M, K, J, nnz;
  const int32_t *pos;
  const int32_t *crd;
  const void *vals;  // nullptr: every sampled value is 1.0
  const void *X;     // M x J (ldx), row-major
  int64_t ldx;
  const void *Y;     // K x J (ldy), row-major
  int64_t ldy;
  float *out;        // nnz
  int32_t lq;        // log2(lanes per entry)
  int32_t tile;      // entries per tile (kSdTile)
  int32_t ntiles;
  const int32_t *tile_row;  // row holding each tile's first entry (ntiles+1)
  // row * ld + column (elements) of X and Y fits 32 bits after scaling to
  // bytes, and every row index and leading dimension is below 2^24
  int32_t off32;
};

int32_t sddmm_tile();
hipError_t launch_sddmm(int dt, int vec, const SddmmArgs &a, hipStream_t s);

}  // namespace segtaco

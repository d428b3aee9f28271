// spmv.h — the N = 1 form of the eb_pr schedule: which launches take it, its
// LDS need, and the standalone launcher (k_spmv.hip; the per-tile body is
// spmv_body.h, also run by the batched kernel).
#pragma once

#include "launch.h"

namespace segtaco {

constexpr int kSvMaxE = 16;  // entries per lane at most (tile <= 1024)

// Row-pointer window per wave: the rows of a tile of 64 * kSvMaxE entries
// plus room for empty rows between them (a tile spanning more rows reads
// pos from memory past the window).
constexpr int32_t kSvWindow = 64 * kSvMaxE + 64;
// LDS of a 4-wave block running the N = 1 form.
constexpr size_t kSvLdsBytes = 4 * (size_t)kSvWindow * sizeof(int32_t);

// Whether a resolved eb_pr launch (vec elements per lane) takes the N = 1
// form: one column, one lane per slot, and a tile of 64 * E entries, E a
// power of two up to kSvMaxE. Standalone calls and batches decide alike, so
// a product gives the same bits in both.
__host__ __device__ inline bool spmv_form(const KArgs &a, int vec) {
  const int32_t E = a.tile / 64;
  return a.N == 1 && vec == 1 && a.lq == 0 && a.tile % 64 == 0 && E <= kSvMaxE && E > 0 &&
         (E & (E - 1)) == 0;
}
inline bool spmv_applies(const KArgs &a, int vec) { return spmv_form(a, vec); }
// Whether a batched product runs the N = 1 form.
inline bool batch_uses_spmv(const BatchDesc &D) {
  return D.schedule == SEGTACO_EB_PR && spmv_form(D.a, D.vec);
}

hipError_t launch_spmv(int dt, const KArgs &a, hipStream_t s);

}  // namespace segtaco

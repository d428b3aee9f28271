// spmv.h — the N = 1 form of the eb_pr schedule (k_spmv.hip).
#pragma once

#include "launch.h"

namespace segtaco {

constexpr int kSvMaxE = 16;  // entries per lane at most (tile <= 1024)

// Whether a resolved eb_pr launch (vec floats per lane) can use k_spmv: one
// column, one lane per slot, and a tile of 64 * E entries, E a power of two
// up to kSvMaxE.
bool spmv_applies(const KArgs &a, int vec);
hipError_t launch_spmv(int dt, const KArgs &a, hipStream_t s);

}  // namespace segtaco

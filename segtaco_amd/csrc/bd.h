// bd.h — the block-dense schedule (k_bd.hip): 32-row blocks of A whose
// entries span at most kBdMaxW columns are multiplied on the matrix cores.
#pragma once

#include "launch.h"

namespace segtaco {

constexpr int32_t kBdRows = 32;   // rows per block (the MFMA tile height)
constexpr int32_t kBdMaxW = 256;  // widest column window of the MFMA path

// Column tiles of 32 per wave: 2 (a 256-column panel) when N > 128, else 1.
int32_t bd_col_tiles(int32_t N);
// a.nblocks = row blocks, a.panels = column panels of 128 * bd_col_tiles(N).
hipError_t launch_bd(int dt, const KArgs &a, hipStream_t s);

}  // namespace segtaco

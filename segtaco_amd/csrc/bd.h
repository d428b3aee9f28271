// bd.h — the block-dense schedule (k_bd.hip): blocks of 32 or 64 rows of A
// whose entries span a narrow column window are multiplied on the matrix
// cores.
#pragma once

#include "launch.h"

namespace segtaco {

constexpr int32_t kBdRows = 32;    // rows of one MFMA tile
constexpr int32_t kBdMaxW = 256;   // widest column window of a 32-row block
constexpr int32_t kBdMaxW2 = 192;  // widest column window of a 64-row block

// Column tiles of 32 per wave: 2 (a 256-column panel) when N > 128, else 1.
int32_t bd_col_tiles(int32_t N);
// a.tile = rows per block (32 or 64), a.nblocks = row blocks, a.panels =
// column panels of 128 * bd_col_tiles(N).
hipError_t launch_bd(int dt, const KArgs &a, hipStream_t s);

}  // namespace segtaco

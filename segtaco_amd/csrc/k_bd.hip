// k_bd.hip — block-dense schedule (bd) for 16-bit storage: C = A * B with
// the products of dense-enough row blocks on the matrix cores.
//
// The reference's schedules are all gather loops on vector ALUs (K1..K18);
// for a banded or block-structured A that leaves the VALU as the bound (per
// element one 16-bit unpack and one FMA). Here the rows are cut into blocks
// of 32 * RB (RB = 1 or 2 row tiles of 32). For a block whose entries span a
// column window cmin .. cmax narrow enough for the LDS tile, the block is
// scattered into a dense (32 RB) x W tile in LDS (zeros elsewhere) and
// multiplied by rows cmin .. cmin+W of B with v_mfma_f32_32x32x16_{bf16,f16}:
// B is staged in LDS kBdKC rows at a time as it lies in memory (each lane
// gathers its eight K-consecutive operands with 2-byte reads), the next
// chunk's loads in flight during this chunk's MFMAs; the RB row tiles share
// every staged chunk, so B is read once per 32 RB rows. A 64-row block whose
// window is too wide runs as two 32-row blocks.
// Each of the 4 waves owns CT tiles of 32 columns per row tile; accumulation
// is in FP32 and every 16-bit product is exact in FP32. A block whose window
// is wider still falls back to a per-row walk on the vector ALUs (Neumaier
// compensated for rows longer than comp_threshold). Every C element of a
// block is written once (rows without entries as zeros): no clearing, no
// atomics, deterministic.
#include <climits>

#include "bd.h"
#include "device_common.h"

namespace segtaco {
namespace {

// LDS A-tile row strides (elements), multiples of 8 so that each lane's
// eight K-consecutive A operands are one aligned 16-byte read: 32-row tiles
// up to kBdMaxW columns wide (264: rows 4 banks apart, so the 16 rows of one
// read pass start on distinct 16-byte bank groups), 64-row tiles up to
// kBdMaxW2 wide (200: rows 36 banks apart, likewise distinct). One buffer
// holds either.
constexpr int32_t kBdStride1 = kBdMaxW + 8;
constexpr int32_t kBdStride2 = kBdMaxW2 + 8;
template <int RB>
constexpr int32_t bd_a_elems() {
  return RB == 1 ? kBdRows * kBdStride1
                 : (kBdRows * kBdStride1 > 2 * kBdRows * kBdStride2 ? kBdRows * kBdStride1
                                                                   : 2 * kBdRows * kBdStride2);
}
// B rows staged per chunk: one MFMA K step (banded 1M^2, N = 256, BF16 with
// 32-row blocks: 16 rows 530 us per call, 32 rows 587 us, 64 rows 667 us: a
// smaller chunk fits more workgroups)
constexpr int kBdKC = 16;
// Scatter: entries per thread per round, and per round. The scatter's
// per-entry registers set the kernel's peak register count: six entries
// instead of eight bring it from 53 to 47 VGPRs (plus 32 accumulators),
// which lets 6 rather than 5 workgroups share a CU (banded 1M^2 BF16,
// N = 256: 524-537 vs 561-565 us per call in one side-by-side run; four
// entries: 528-543 us)
constexpr int kBdScat = 6;
constexpr int32_t kBdRound = kBlockThreads * kBdScat;

typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// C[32x32] += A[32x16] B[16x32]: lane l holds A[l % 32][8 (l / 32) .. + 7]
// and B[8 (l / 32) .. + 7][l % 32]; C[i][j] comes back in lane j + 32 ((i / 4)
// % 2), register (i % 4) + 4 (i / 8).
template <int DT>
__device__ __forceinline__ v16f mfma32(v8s a, v8s b, v16f c) {
  if constexpr (DT == 1)
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a),
                                                   __builtin_bit_cast(v8bf, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, a),
                                                  __builtin_bit_cast(v8h, b), c, 0, 0, 0);
}

// One stored value as its 16-bit encoding (1.0 when A has no values).
template <int DT>
__device__ __forceinline__ uint16_t raw_val(const void *vals, int32_t p) {
  if (vals == nullptr) return DT == 1 ? 0x3F80 : 0x3C00;
  return static_cast<const uint16_t *>(vals)[p];
}

// Rows r0 .. r0+rows-1, columns col0 .. col0+cw-1 (clipped to N), on the
// vector ALUs: wave w takes rows w, w+4, ...; lanes take columns.
template <int DT>
__device__ void bd_rows_valu(const KArgs &a, int32_t r0, int32_t rows, int32_t col0,
                             int32_t cw) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint16_t *B = static_cast<const uint16_t *>(a.B);
  const int32_t cend = min(a.N, col0 + cw);
  for (int32_t i = wave; i < rows; i += kWavesPerBlock) {
    const int32_t row = r0 + i;
    const int32_t s = a.pos[row], e = a.pos[row + 1];
    const bool comp = e - s > a.comp_threshold;
    for (int32_t c = col0 + lane; c < cend; c += 64) {
      float acc = 0.0f, cmp = 0.0f;
      for (int32_t p = s; p < e; ++p) {
        const float v = load_val<DT>(a.vals, p);
        const float b = half_to_f32<DT>(B[(int64_t)a.crd[p] * a.ldb + c]);
        if (comp)
          neumaier_add(acc, cmp, v, b);
        else
          acc = fmaf(v, b, acc);
      }
      a.C[(int64_t)row * a.ldc + c] = comp ? acc + cmp : acc;
    }
  }
}

// B rows cmin+k0 .. cmin+k0+kBdKC-1, the panel's 8-column segment of this
// thread in each of its kBdKC / kRowsPerPass rows (zeros past the window, K
// or N): the loads of one chunk, held in registers until written to LDS.
template <int CT>
struct BChunk {
  static constexpr int32_t kCols = kWavesPerBlock * CT * 32;
  static constexpr int32_t kSegs = kCols / 8;
  static constexpr int32_t kRowsPerPass = kBlockThreads / kSegs;
  static constexpr int32_t kPasses = kBdKC / kRowsPerPass;
  uint4 q[kPasses];
};

// full8: B 16-byte aligned with ldb and N multiples of 8 (uniform over the
// launch), so each thread's 8-column segment lies wholly inside or wholly
// past N. Then every segment is one 16-byte load with no branch around it —
// rows past the window or K read row cmin, segments past N column 0, and
// both are replaced by zeros — so the number of loads in flight is the same
// on every path and a wait for one chunk leaves the later chunks in flight.
template <int CT>
__device__ __forceinline__ void bd_load_chunk(const KArgs &a, int32_t cmin, int32_t W,
                                              int32_t k0, int32_t col0, bool full8,
                                              BChunk<CT> &ch) {
  using C = BChunk<CT>;
  const uint16_t *B = static_cast<const uint16_t *>(a.B);
  const int tid = threadIdx.x;
  const int32_t cs = (tid % C::kSegs) * 8, col = col0 + cs;
#pragma unroll
  for (int i = 0; i < C::kPasses; ++i) {
    const int32_t k = k0 + tid / C::kSegs + i * C::kRowsPerPass;
    const int32_t grow = cmin + k;
    const bool row_ok = k < W && grow < a.K;
    if (full8) {
      const bool ok = row_ok && col < a.N;
      const uint16_t *src = B + (int64_t)(row_ok ? grow : cmin) * a.ldb + (col < a.N ? col : 0);
      const uint4 q = *reinterpret_cast<const uint4 *>(src);
      ch.q[i] = ok ? q : make_uint4(0u, 0u, 0u, 0u);
    } else {
      uint4 q = make_uint4(0u, 0u, 0u, 0u);
      if (row_ok) {
        const uint16_t *src = B + (int64_t)grow * a.ldb + col;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (col + u < a.N) w[u >> 1] |= (uint32_t)src[u] << (16 * (u & 1));
        q = make_uint4(w[0], w[1], w[2], w[3]);
      }
      ch.q[i] = q;
    }
  }
}

// Writes a loaded chunk to LDS as it lies in B, sB[k][column]: each thread
// one 16-byte store, consecutive lanes at consecutive addresses (a
// transposed layout would put all lanes of a store on one bank).
template <int CT>
__device__ __forceinline__ void bd_store_chunk(uint16_t *sB, const BChunk<CT> &ch) {
  using C = BChunk<CT>;
  const int tid = threadIdx.x;
  const int32_t cs = (tid % C::kSegs) * 8;
#pragma unroll
  for (int i = 0; i < C::kPasses; ++i) {
    const int32_t kr = tid / C::kSegs + i * C::kRowsPerPass;
    *reinterpret_cast<uint4 *>(&sB[kr * C::kCols + cs]) = ch.q[i];
  }
}

// The MFMA B operand of one lane: B[k .. k+7][jc] from the row-major chunk.
__device__ __forceinline__ v8s bd_b_operand(const uint16_t *sB, int32_t cols, int32_t k,
                                            int32_t jc) {
  v8s b;
#pragma unroll
  for (int e = 0; e < 8; ++e) b[e] = (short)sB[(k + e) * cols + jc];
  return b;
}

// LDS of one workgroup.
template <int CT, int RB>
struct BdShared {
  __align__(16) uint16_t A[bd_a_elems<RB>()];
  __align__(16) uint16_t B[kBdKC * kWavesPerBlock * CT * 32];
  int32_t pos[kBdRows * RB + 1];
  int32_t win[2 * RB];                      // per 32-row half: first, last column
  int32_t edge[kBdScat][kWavesPerBlock];    // each wave's last column per round
};

// The MFMA path for the RT row tiles r0 .. r0 + rows - 1 (rows <= 32 RT;
// pos = their row pointer in LDS, rows + 1 entries) over the column window
// [cmin, cmax] (at most stride - 8 wide), panel columns col0 ..; writes C
// and returns true, or returns false without writing when the scatter finds
// an entry outside the window or a row whose columns do not increase.
// Called by the whole workgroup (uniform arguments).
template <int DT, int CT, int RB, int RT>
__device__ bool bd_mfma(const KArgs &a, BdShared<CT, RB> &sh, const int32_t *pos, int32_t r0,
                        int32_t rows, int32_t col0, int32_t cmin, int32_t cmax,
                        int32_t stride) {
  constexpr int32_t kCols = BChunk<CT>::kCols;  // panel width
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t W = cmax - cmin + 1;
  const int32_t Wk = (W + kBdKC - 1) & ~(kBdKC - 1);  // MFMA K steps of 16
  // branch-free 16-byte B loads (bd_load_chunk)
  const bool full8 = (reinterpret_cast<uintptr_t>(a.B) & 15) == 0 && (a.ldb & 7) == 0 &&
                     (a.N & 7) == 0;
  uint16_t *sA = sh.A;
  __syncthreads();  // a previous pass has finished with the LDS tiles

  // the first B chunk and the block's entries in flight together, while the
  // A tile's window is zeroed
  BChunk<CT> ch;
  bd_load_chunk<CT>(a, cmin, W, 0, col0, full8, ch);
  const int32_t pb = pos[0], n = pos[rows] - pb;
  for (int32_t i = wave; i < kBdRows * RT; i += kWavesPerBlock)
    for (int32_t w = lane; w < Wk / 2; w += 64)
      reinterpret_cast<uint32_t *>(&sA[i * stride])[w] = 0u;
  // Scatter, checking the column order it relies on as it goes: the window
  // and one tile cell per entry need every row's columns strictly
  // increasing. An entry outside the window is not written, and a block with
  // such an entry, or with a column that does not exceed its predecessor in
  // the same row (unsorted or duplicate columns), takes the per-row walk,
  // which sums every entry in CSR order. The predecessor's column comes from
  // the neighbouring lane (entry j - 1 is thread tid - 1's), across waves and
  // rounds from a few staged edge values, so it costs no second load.
  int bad = 0;
  int32_t prev_last = INT_MIN;  // column of the entry before this round
  for (int32_t j0 = 0; j0 < n; j0 += kBdRound) {
    int32_t cc[kBdScat];
    uint16_t vv[kBdScat];
#pragma unroll
    for (int u = 0; u < kBdScat; ++u) {
      const int32_t q = pb + min(j0 + tid + u * kBlockThreads, n - 1);
      cc[u] = a.crd[q];
      vv[u] = raw_val<DT>(a.vals, q);
    }
    __syncthreads();  // tile zeroed (first round) / previous round's edges read
    if (lane == 63) {
#pragma unroll
      for (int u = 0; u < kBdScat; ++u) sh.edge[u][wave] = cc[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kBdScat; ++u) {
      const int32_t j = j0 + tid + u * kBlockThreads;
      const int32_t up = __shfl_up(cc[u], 1, 64);
      if (j < n) {
        const int32_t p = pb + j;
        int32_t lo = 0, hi = rows - 1;  // row: last i with pos[i] <= p
        while (lo < hi) {
          const int32_t mid = (lo + hi + 1) >> 1;
          if (pos[mid] <= p)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int32_t c = cc[u];
        const bool in = c >= cmin && c <= cmax;
        const int32_t prev = lane > 0  ? up
                             : wave > 0 ? sh.edge[u][wave - 1]
                             : u > 0    ? sh.edge[u - 1][kWavesPerBlock - 1]
                                        : prev_last;
        bad |= (!in || (p != pos[lo] && prev >= c)) ? 1 : 0;
        if (in) sA[lo * stride + (c - cmin)] = vv[u];
      }
    }
    // a round that is not the last is full: its last entry is thread
    // kBlockThreads - 1's last
    prev_last = sh.edge[kBdScat - 1][kWavesPerBlock - 1];
  }
  if (__syncthreads_or(bad)) return false;

  // K loop over the window, kBdKC rows of B at a time: a chunk's registers go
  // to LDS and are refilled at once with the next chunk, whose loads are in
  // flight during this chunk's MFMAs. Measured on the banded 1M^2 BF16
  // workload at N = 256 with 32-row blocks, 500 us per call; three chunks in
  // flight, each refilled with the chunk three steps ahead: 564 us; two LDS
  // buffers with one barrier per chunk instead of two: 524 us (more LDS per
  // workgroup, fewer workgroups per CU).
  v16f acc[RT][CT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[rt][ct][r] = 0.0f;
  const int32_t kl = 8 * (lane >> 5);  // this lane's K offset in a step
  for (int32_t k0 = 0; k0 < Wk; k0 += kBdKC) {
    __syncthreads();  // A scattered / previous chunk consumed
    bd_store_chunk<CT>(sh.B, ch);
    if (k0 + kBdKC < Wk) bd_load_chunk<CT>(a, cmin, W, k0 + kBdKC, col0, full8, ch);
    __syncthreads();
    v8s bv[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
      bv[ct] = bd_b_operand(sh.B, kCols, kl, (wave * CT + ct) * 32 + (lane & 31));
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      const v8s av =
          *reinterpret_cast<const v8s *>(&sA[(rt * 32 + (lane & 31)) * stride + k0 + kl]);
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = mfma32<DT>(av, bv[ct], acc[rt][ct]);
    }
  }

  // C tiles: lane holds rows (r%4) + 4*(lane/32) + 8*(r/4), column lane%32.
  // Non-temporal stores: C (1.07 GB on the banded 1M^2 workload at N = 256)
  // is never read back here, and allocating it in L2 evicts the B-window
  // lines neighbouring blocks reuse (5 % less FETCH, 1-3 % less time).
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int32_t col = col0 + (wave * CT + ct) * 32 + (lane & 31);
    if (col >= a.N) continue;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int32_t i = rt * 32 + (r & 3) + 4 * (lane >> 5) + 8 * (r >> 2);
        if (i < rows) __builtin_nontemporal_store(acc[rt][ct][r], &a.C[(int64_t)(r0 + i) * a.ldc + col]);
      }
  }
  return true;
}

template <int DT, int CT, int RB>
__global__ __launch_bounds__(kBlockThreads) void k_bd(const KArgs a) {
  __shared__ BdShared<CT, RB> sh;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int32_t kCols = BChunk<CT>::kCols;  // panel width
  constexpr int32_t kRowsB = kBdRows * RB;      // rows per block
  // XCD-contiguous order: workgroup b runs on XCD b % 8, and XCD x takes the
  // units x*per .. x*per + per - 1 in ascending order, so the blocks resident
  // on one XCD at a time are neighbours whose B windows overlap (a banded
  // block's window shares 64 rows with each neighbour) and its L2 serves the
  // overlap. Round-robin order put neighbours on different XCDs, and every
  // block fetched its whole window from memory.
  const int32_t units = a.nblocks * a.panels;
  const int32_t per = (units + 7) / 8;
  const int32_t u = (int32_t)(blockIdx.x & 7) * per + (int32_t)(blockIdx.x >> 3);
  if (u >= units) return;  // padding of the last round (whole workgroup)
  const int32_t rb = u / a.panels, panel = u % a.panels;
  const int32_t r0 = rb * kRowsB;
  const int32_t rows = min(kRowsB, a.M - r0);
  const int32_t col0 = panel * kCols;

  // 1. the block's row pointer, and each 32-row half's column window (rows
  // are sorted: their first and last entries bound it)
  if (wave == 0) {
    int32_t lo = INT_MAX, hi = -1;
    if (lane < rows) {
      const int32_t s = a.pos[r0 + lane], e = a.pos[r0 + lane + 1];
      sh.pos[lane] = s;
      if (lane == rows - 1) sh.pos[rows] = e;
      if (s < e) {
        lo = a.crd[s];
        hi = a.crd[e - 1];
      }
    }
    for (int d = 16; d >= 1; d >>= 1) {
      lo = min(lo, __shfl_xor(lo, d, 64));
      hi = max(hi, __shfl_xor(hi, d, 64));
    }
    if ((lane & 31) == 0 && (lane >> 5) < RB) {
      sh.win[2 * (lane >> 5)] = lo;
      sh.win[2 * (lane >> 5) + 1] = hi;
    }
  }
  __syncthreads();
  if constexpr (RB == 2) {
    const int32_t lo = min(sh.win[0], sh.win[2]), hi = max(sh.win[1], sh.win[3]);
    if (hi >= 0 && hi - lo + 1 <= kBdMaxW2) {
      if (!bd_mfma<DT, CT, RB, 2>(a, sh, sh.pos, r0, rows, col0, lo, hi, kBdStride2))
        bd_rows_valu<DT>(a, r0, rows, col0, kCols);
      return;
    }
  }
  // 32-row halves, each on its own window
  for (int32_t h = 0; h < RB && 32 * h < rows; ++h) {
    const int32_t hr0 = r0 + 32 * h, hrows = min(32, rows - 32 * h);
    const int32_t lo = sh.win[2 * h], hi = sh.win[2 * h + 1];
    // no entries (every row written as zeros) or too wide for the tile
    if (hi < 0 || hi - lo + 1 > kBdMaxW ||
        !bd_mfma<DT, CT, RB, 1>(a, sh, sh.pos + 32 * h, hr0, hrows, col0, lo, hi, kBdStride1))
      bd_rows_valu<DT>(a, hr0, hrows, col0, kCols);
  }
}

}  // namespace

int32_t bd_col_tiles(int32_t N) { return N > kWavesPerBlock * 32 ? 2 : 1; }

hipError_t launch_bd(int dt, const KArgs &a, hipStream_t s) {
  const int64_t units = (int64_t)a.nblocks * a.panels;
  if (units <= 0) return hipSuccess;
  if (units > 0x7fffffff - 8) return hipErrorInvalidConfiguration;
  const int64_t grid = 8 * ((units + 7) / 8);  // the kernel's XCD-contiguous decode
  if (grid > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 g((unsigned)grid), b(kBlockThreads);
  const int ct = bd_col_tiles(a.N);
  const int rb = a.tile / kBdRows;
  switch (dt * 16 + ct * 4 + rb) {
    case 1 * 16 + 1 * 4 + 1: hipLaunchKernelGGL((k_bd<1, 1, 1>), g, b, 0, s, a); break;
    case 1 * 16 + 2 * 4 + 1: hipLaunchKernelGGL((k_bd<1, 2, 1>), g, b, 0, s, a); break;
    case 2 * 16 + 1 * 4 + 1: hipLaunchKernelGGL((k_bd<2, 1, 1>), g, b, 0, s, a); break;
    case 2 * 16 + 2 * 4 + 1: hipLaunchKernelGGL((k_bd<2, 2, 1>), g, b, 0, s, a); break;
    case 1 * 16 + 1 * 4 + 2: hipLaunchKernelGGL((k_bd<1, 1, 2>), g, b, 0, s, a); break;
    case 1 * 16 + 2 * 4 + 2: hipLaunchKernelGGL((k_bd<1, 2, 2>), g, b, 0, s, a); break;
    case 2 * 16 + 1 * 4 + 2: hipLaunchKernelGGL((k_bd<2, 1, 2>), g, b, 0, s, a); break;
    case 2 * 16 + 2 * 4 + 2: hipLaunchKernelGGL((k_bd<2, 2, 2>), g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace segtaco

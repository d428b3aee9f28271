// k_bd.hip — block-dense schedule (bd) for 16-bit storage: C = A * B with
// the products of dense-enough row blocks on the matrix cores.
//
// The reference's schedules are all gather loops on vector ALUs (K1..K18);
// for a banded or block-structured A that leaves the VALU as the bound (per
// element one 16-bit unpack and one FMA). Here the rows are cut into blocks
// of 32. For a block whose entries span a column window of at most 256
// (cmin .. cmax), the block is scattered into a dense 32 x W tile in LDS
// (zeros elsewhere) and multiplied by rows cmin .. cmin+W of B with
// v_mfma_f32_32x32x8{bf16_1k,f16}: B is staged in LDS kBdKC rows at a time
// as it lies in memory (each lane gathers its four K-consecutive operands
// with 2-byte reads), the next chunk's loads in flight during this chunk's
// MFMAs.
// Each of the 4 waves owns CT tiles of 32 columns; accumulation is in
// FP32 and every 16-bit product is exact in FP32. A block whose window is
// wider falls back to a per-row walk on the vector ALUs (Neumaier
// compensated for rows longer than comp_threshold). Every C element of a
// block is written once (rows without entries as zeros): no clearing, no
// atomics, deterministic.
#include <climits>

#include "bd.h"
#include "device_common.h"

namespace segtaco {
namespace {

// LDS A-tile row stride (elements): 130 words, so the 32 rows that one MFMA
// operand read touches start in 16 different bank pairs
constexpr int kBdAStride = kBdMaxW + 4;
// B rows staged per chunk (banded 1M^2, N = 256, BF16: 16 rows 530 us per
// call, 32 rows 587 us, 64 rows 667 us: a smaller chunk fits more workgroups)
constexpr int kBdKC = 16;
// Scatter: entries per thread per round, and per round
constexpr int kBdScat = 8;
constexpr int32_t kBdRound = kBlockThreads * kBdScat;

typedef short v4s __attribute__((ext_vector_type(4)));
typedef _Float16 v4h __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

template <int DT>
__device__ __forceinline__ v16f mfma32(v4s a, v4s b, v16f c) {
  if constexpr (DT == 1)
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x8f16(__builtin_bit_cast(v4h, a),
                                                __builtin_bit_cast(v4h, b), c, 0, 0, 0);
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

// The MFMA B operand of one lane: B[k .. k+3][jc] from the row-major chunk.
__device__ __forceinline__ v4s bd_b_operand(const uint16_t *sB, int32_t cols, int32_t k,
                                            int32_t jc) {
  v4s b;
#pragma unroll
  for (int e = 0; e < 4; ++e) b[e] = (short)sB[(k + e) * cols + jc];
  return b;
}

template <int DT, int CT>
__global__ __launch_bounds__(kBlockThreads) void k_bd(const KArgs a) {
  __shared__ __align__(16) uint16_t sA[kBdRows * kBdAStride];
  __shared__ __align__(16) uint16_t sB[kBdKC * kWavesPerBlock * CT * 32];
  __shared__ int32_t s_pos[kBdRows + 1];
  __shared__ int32_t s_win[2];
  __shared__ int32_t s_edge[kBdScat][kWavesPerBlock];  // each wave's last column
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int32_t kCols = BChunk<CT>::kCols;  // panel width
  // XCD-contiguous order: workgroup b runs on XCD b % 8, and XCD x takes the
  // units x*per .. x*per + per - 1 in ascending order, so the blocks resident
  // on one XCD at a time are neighbours whose B windows overlap (a banded
  // block's 96-row window shares 64 rows with each neighbour) and its L2
  // serves the overlap. Round-robin order put neighbours on different XCDs,
  // and every block fetched its whole window from memory.
  const int32_t units = a.nblocks * a.panels;
  const int32_t per = (units + 7) / 8;
  const int32_t u = (int32_t)(blockIdx.x & 7) * per + (int32_t)(blockIdx.x >> 3);
  if (u >= units) return;  // padding of the last round (whole workgroup)
  const int32_t rb = u / a.panels, panel = u % a.panels;
  const int32_t r0 = rb * kBdRows;
  const int32_t rows = min(kBdRows, a.M - r0);
  const int32_t col0 = panel * kCols;

  // 1. the block's row pointer, and its column window (rows are sorted:
  // their first and last entries bound it)
  if (wave == 0) {
    if (lane <= rows) s_pos[lane] = a.pos[r0 + lane];
    int32_t lo = INT_MAX, hi = -1;
    if (lane < rows) {
      const int32_t s = a.pos[r0 + lane], e = a.pos[r0 + lane + 1];
      if (s < e) {
        lo = a.crd[s];
        hi = a.crd[e - 1];
      }
    }
    for (int d = 32; d >= 1; d >>= 1) {
      lo = min(lo, __shfl_xor(lo, d, 64));
      hi = max(hi, __shfl_xor(hi, d, 64));
    }
    if (lane == 0) {
      s_win[0] = lo;
      s_win[1] = hi;
    }
  }
  __syncthreads();
  const int32_t cmin = s_win[0], cmax = s_win[1];
  if (cmax < 0 || cmax - cmin + 1 > kBdMaxW) {
    // no entries (every row written as zeros) or too wide for the tile
    bd_rows_valu<DT>(a, r0, rows, col0, kCols);
    return;
  }
  const int32_t W = cmax - cmin + 1;
  const int32_t Wk = (W + 7) & ~7;  // MFMA K steps of 8
  // branch-free 16-byte B loads (bd_load_chunk)
  const bool full8 = (reinterpret_cast<uintptr_t>(a.B) & 15) == 0 && (a.ldb & 7) == 0 &&
                     (a.N & 7) == 0;

  // 2. the first B chunk and the block's entries in flight together, while
  // the A tile's window is zeroed
  BChunk<CT> ch;
  bd_load_chunk<CT>(a, cmin, W, 0, col0, full8, ch);
  const int32_t pb = s_pos[0], n = s_pos[rows] - pb;
  for (int32_t i = wave; i < kBdRows; i += kWavesPerBlock)
    for (int32_t w = lane; w < Wk / 2; w += 64)
      reinterpret_cast<uint32_t *>(&sA[i * kBdAStride])[w] = 0u;
  // Scatter, checking the column order it relies on as it goes: the window
  // above and one tile cell per entry need every row's columns strictly
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
      for (int u = 0; u < kBdScat; ++u) s_edge[u][wave] = cc[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kBdScat; ++u) {
      const int32_t j = j0 + tid + u * kBlockThreads;
      const int32_t up = __shfl_up(cc[u], 1, 64);
      if (j < n) {
        const int32_t p = pb + j;
        int32_t lo = 0, hi = rows - 1;  // row: last i with s_pos[i] <= p
        while (lo < hi) {
          const int32_t mid = (lo + hi + 1) >> 1;
          if (s_pos[mid] <= p)
            lo = mid;
          else
            hi = mid - 1;
        }
        const int32_t c = cc[u];
        const bool in = c >= cmin && c <= cmax;
        const int32_t prev = lane > 0  ? up
                             : wave > 0 ? s_edge[u][wave - 1]
                             : u > 0    ? s_edge[u - 1][kWavesPerBlock - 1]
                                        : prev_last;
        bad |= (!in || (p != s_pos[lo] && prev >= c)) ? 1 : 0;
        if (in) sA[lo * kBdAStride + (c - cmin)] = vv[u];
      }
    }
    // a round that is not the last is full: its last entry is thread
    // kBlockThreads - 1's last
    prev_last = s_edge[kBdScat - 1][kWavesPerBlock - 1];
  }
  if (__syncthreads_or(bad)) {
    bd_rows_valu<DT>(a, r0, rows, col0, kCols);
    return;
  }

  // 3. K loop over the window, kBdKC rows of B at a time: a chunk's
  // registers go to LDS and are refilled at once with the next chunk, whose
  // loads are in flight during this chunk's MFMAs. (Three chunks in flight,
  // each refilled with the chunk three steps ahead, measured slower on the
  // banded 1M^2 BF16 workload at N = 256: 564 us vs 505 us per call.)
  v16f acc[CT];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[ct][r] = 0.0f;
  for (int32_t k0 = 0; k0 < Wk; k0 += kBdKC) {
    __syncthreads();  // A scattered / previous chunk consumed
    bd_store_chunk<CT>(sB, ch);
    if (k0 + kBdKC < Wk) bd_load_chunk<CT>(a, cmin, W, k0 + kBdKC, col0, full8, ch);
    __syncthreads();
    const int32_t ksteps = min(kBdKC, Wk - k0) / 8;
    for (int32_t ks = 0; ks < ksteps; ++ks) {
      const int32_t kk = ks * 8 + 4 * (lane >> 5);
      const v4s av =
          *reinterpret_cast<const v4s *>(&sA[(lane & 31) * kBdAStride + k0 + kk]);
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int32_t jc = (wave * CT + ct) * 32 + (lane & 31);
        const v4s bv = bd_b_operand(sB, kCols, kk, jc);
        acc[ct] = mfma32<DT>(av, bv, acc[ct]);
      }
    }
  }

  // 4. C tile: lane holds rows (r%4) + 4*(lane/32) + 8*(r/4), column lane%32
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int32_t col = col0 + (wave * CT + ct) * 32 + (lane & 31);
    if (col >= a.N) continue;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int32_t i = (r & 3) + 4 * (lane >> 5) + 8 * (r >> 2);
      if (i < rows) a.C[(int64_t)(r0 + i) * a.ldc + col] = acc[ct][r];
    }
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
  switch (dt * 4 + ct) {
    case 1 * 4 + 1: hipLaunchKernelGGL((k_bd<1, 1>), g, b, 0, s, a); break;
    case 1 * 4 + 2: hipLaunchKernelGGL((k_bd<1, 2>), g, b, 0, s, a); break;
    case 2 * 4 + 1: hipLaunchKernelGGL((k_bd<2, 1>), g, b, 0, s, a); break;
    case 2 * 4 + 2: hipLaunchKernelGGL((k_bd<2, 2>), g, b, 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace segtaco

// k_cs.hip — eb_cs: element-balanced tiles swept in column order, with the
// rows of a CU-sized unit of tiles accumulated in LDS.
//
// Why: on the power-law workload every stored entry gathers a B row and
// columns are uniform over K, so in CSR order the gathers of all waves land
// all over B and nearly every one misses the XCD's 4 MB L2 (measured traffic
// equals the no-reuse gather model). If instead every lane group walks its
// tile's entries in ascending column order, all groups of the chip sweep B
// from row 0 to row K-1 together and the rows in flight at any moment form a
// narrow band that L2 serves to every CU of the XCD.
//
// How, keeping the reference's arithmetic (spmm_eb_sr.h:5-78 semantics,
// Eigen's per-row sequential order):
//   * tiles are the eb_sr tiles (T consecutive entries and the repaired
//     first row of each, k_tile_rows_fix); a unit is G consecutive tiles,
//     one per lane group of one 512-thread workgroup, eight per wave;
//   * a preparation pass, k_cs_count then k_cs_sort (one wave per tile),
//     names each entry's row by its place in the tile (head, tail, k-th
//     interior row, i.e. a row that starts and ends inside the tile, whose
//     row id goes to a side array) and decides per wave of the unit whether
//     its tiles sweep: they do when their interior rows fit the LDS slots
//     beside those of the unit's earlier sweeping waves. A sweeping tile's
//     rows get LDS slots — slot 2g for the row tile g continues from an
//     earlier tile (head), 2g + 1 for the row it leaves open for a later one
//     (tail), from 2G on the unit's interior rows — and its entries are
//     written sorted by (column, position) as one record each (column, slot,
//     value). A walking tile's entries stay in CSR order with the in-tile
//     names;
//   * k_eb_cs, per unit: the rows of its interior slots into LDS and its
//     slots zeroed, then per column panel of W = 8*V columns: every sweeping
//     lane group runs its sorted tile — per entry one 16-byte gather per lane
//     and one read-modify-write of its row's slot, with no branch on the kind
//     of row — and every walking group runs its tile in CSR order with
//     register accumulators, written out at each change of row; then interior
//     slots go to C, a row over two neighbouring sweeping tiles is summed from
//     its head / tail slots by the tile where it ends (in k_fixup's order),
//     and the other head / tail slots go to the carry arrays; each slot is
//     zeroed as it is read, for the next panel. Block barriers around each
//     panel's sweep keep the unit's waves on the same band of B rows.
//     k_tile_rows_fix zeroes the empty rows and k_fixup combines the
//     remaining carries, as for eb_sr.
// Each row's entries meet its accumulator in ascending column order (sweep)
// or CSR order (walk), the same order for canonical CSR, with one fmaf per
// entry starting from zero: C is bit-identical to eb_sr with the same tile
// size. Walking is decided per wave so that no wave runs both paths.
#include <climits>

#include "cs.h"
#include "device_common.h"

namespace segtaco {

namespace {

constexpr int kCsPrepWaves = kCsPrepThreads / 64;
// Bits of an entry's place in its tile inside a sort key (kCsMaxTile = 2^9).
constexpr int kCsIdxBits = 9;
static_assert((1 << kCsIdxBits) == kCsMaxTile, "sort key index field");
// The sort's column buckets (one per lane) and the most keys a lane sorts in
// registers; a tile with a fuller bucket (its columns clustered on a few
// values) is sorted by the whole wave instead.
constexpr int kCsBuckets = 64;
constexpr int kCsBucketMax = 16;
// Mark of the tail row's first entry: above every interior mark (k + 1 for
// interior row k < kCsMaxTile).
constexpr int32_t kCsTailMark = 1 << 20;

// Record of one entry when columns do not fit the packed form (one 12-byte
// load).
struct CsWide {
  uint32_t col, code, vbits;
};

// x of the lane a DPP row shift (CTRL) names; 0 where that lane lies outside
// the row of 16 lanes.
template <int CTRL>
__device__ __forceinline__ int32_t dpp_or_zero(int32_t x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}

// Inclusive scan over the wave's 64 lanes, sum (kMax false) or maximum
// (kMax true), of values >= 0 (0 is the identity of both): shifts by 1, 2,
// 4 and 8 lanes inside each row of 16 (DPP, no LDS), then each row adds the
// totals of the rows below it (lanes 15, 31 and 47).
template <bool kMax>
__device__ __forceinline__ int32_t wave_scan(int32_t x) {
  auto op = [](int32_t u, int32_t v) { return kMax ? max(u, v) : u + v; };
  x = op(x, dpp_or_zero<0x111>(x));  // row_shr:1
  x = op(x, dpp_or_zero<0x112>(x));  // row_shr:2
  x = op(x, dpp_or_zero<0x114>(x));  // row_shr:4
  x = op(x, dpp_or_zero<0x118>(x));  // row_shr:8
  const int32_t b1 = __builtin_amdgcn_readlane(x, 15);
  const int32_t b2 = op(b1, __builtin_amdgcn_readlane(x, 31));
  const int32_t b3 = op(b2, __builtin_amdgcn_readlane(x, 47));
  const int row = (threadIdx.x & 63) >> 4;
  return row == 0 ? x : op(x, row == 1 ? b1 : row == 2 ? b2 : b3);
}

// Slot plan of tile t's unit. The main kernel's wave w runs tiles 8w ..
// 8w + 7 of the unit; in wave order, a wave sweeps when the interior rows of
// its tiles fit the slots left after the 2G head / tail slots and the
// earlier sweeping waves' rows, and walks otherwise (a later, smaller wave
// may still sweep). Returns tile t's first interior slot, or -1 when its
// wave walks; n_slots = slots the unit uses, count = tile t's interior rows.
__device__ int32_t cs_slot_base(const CsArgs &a, int32_t t, int32_t &n_slots, int32_t &count) {
  const int lane = threadIdx.x & 63;
  const int32_t t0 = (t / a.group) * a.group;
  const int32_t g = t - t0;
  const int32_t cnt = (lane < a.group && t0 + lane < a.ntiles) ? a.tile_count[t0 + lane] : 0;
  // interior rows of the unit's tiles 0 .. j, and of none for j = -1
  const int32_t inc = wave_scan<false>(cnt);
  auto upto = [&](int32_t j) { return j < 0 ? 0 : __builtin_amdgcn_readlane(inc, j); };
  const int32_t w_g = g >> 3;
  const int32_t before = upto(g - 1) - upto(8 * w_g - 1);  // tiles before g in its wave
  const int32_t room = a.cap - 2 * a.group;
  int32_t used = 0, base = -1;
#pragma unroll
  for (int w = 0; w < 8; ++w) {
    const int32_t sw = upto(8 * w + 7) - upto(8 * w - 1);
    if (used + sw <= room) {
      if (w == w_g) base = 2 * a.group + used + before;
      used += sw;
    }
  }
  n_slots = 2 * a.group + used;
  count = __builtin_amdgcn_readlane(cnt, g);
  return base;
}

// Ascending bitonic sort of the 64 * R keys held as key[r] of lane l =
// element 64r + l. Distances of 64 and more pair registers of one lane;
// shorter ones pair lanes through a shuffle.
template <int R>
__device__ __forceinline__ void wave_bitonic(uint32_t (&key)[R]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 2; k <= 64 * R; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= 64) {
        const int jr = j / 64;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if ((r & jr) != 0) continue;
          const int rp = r | jr;
          // element 64r + l is below its partner; the pair is ascending
          // when bit k of the element index is clear (k >= 128 here)
          const bool asc = ((64 * r) & k) == 0;
          const uint32_t lo = min(key[r], key[rp]), hi = max(key[r], key[rp]);
          key[r] = asc ? lo : hi;
          key[rp] = asc ? hi : lo;
        }
      } else {
        const bool lower = (lane & j) == 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint32_t p = (uint32_t)__shfl_xor((int32_t)key[r], j, 64);
          const bool asc = (((64 * r) | lane) & k) == 0;
          key[r] = (asc == lower) ? min(key[r], p) : max(key[r], p);
        }
      }
    }
  }
}

// Ascending sort of W <= 16 keys in one lane's registers: Batcher's odd-even
// merge network for the next power of two without the compare-exchanges
// that reach past W (they would meet +infinity padding, which stays on top):
// 19, 32, 42 and 63 compare-exchanges for W = 8, 10, 12 and 16.
template <int W>
__device__ __forceinline__ void lane_sort(uint32_t (&v)[W]) {
#pragma unroll
  for (int p = 1; p < W; p <<= 1) {
#pragma unroll
    for (int k = p; k >= 1; k >>= 1) {
#pragma unroll
      for (int j = k % p; j + k < W; j += 2 * k) {
#pragma unroll
        for (int i = 0; i < k; ++i) {
          if (i + j + k < W && (i + j) / (2 * p) == (i + j + k) / (2 * p)) {
            const uint32_t x = v[i + j], y = v[i + j + k];
            v[i + j] = min(x, y);
            v[i + j + k] = max(x, y);
          }
        }
      }
    }
  }
}

// Lane b sorts bucket b in place: sk[st .. st + n), n <= W.
template <int W>
__device__ __forceinline__ void cs_sort_bucket(uint32_t *sk, int32_t st, int32_t n) {
  uint32_t v[W];
#pragma unroll
  for (int m = 0; m < W; ++m) v[m] = m < n ? sk[st + m] : 0xffffffffu;
  lane_sort<W>(v);
#pragma unroll
  for (int m = 0; m < W; ++m)
    if (m < n) sk[st + m] = v[m];
}

// Prep, one wave per tile.
//
// k_cs_count: the tile's rows, 64 per step from its first row (the repaired
// tile row) until the rows begin past the tile. Per row: head (began before
// the tile), tail (begins inside, ends past it) or interior (begins and ends
// inside, not empty); interior rows are numbered in row order and their ids
// go to int_rows. Each row that begins inside marks its first entry in LDS
// (k + 1 for interior row k, kCsTailMark for the tail), and a running
// maximum over the tile's entries names every entry's row: code 0 head, 1
// tail, 2 + k interior row k. Then the tile's count of interior rows and its
// head / tail flags.
template <int E>
__global__ __launch_bounds__(kCsPrepThreads) void k_cs_count(const CsArgs a) {
  __shared__ int32_t s_mark[kCsPrepWaves][64 * E];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar tile math
  const int32_t t = blockIdx.x * kCsPrepWaves + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  int32_t *mk = s_mark[wave];
  const int32_t s = t * a.tile;
  const int32_t cn = min(a.nnz - s, a.tile);
  const int32_t e = s + cn;
#pragma unroll
  for (int k = 0; k < E; ++k)
    if (lane + 64 * k < cn) mk[lane + 64 * k] = 0;
  wave_sync();
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  int32_t count = 0;
  bool head = false, tail = false;
  // 64-bit row steps: rows near INT32_MAX must not wrap the loop
  for (int64_t rb = a.tile_row[t]; rb <= r_hi; rb += 64) {
    const int32_t r = (int32_t)min(rb + lane, (int64_t)r_hi + 1);
    int32_t p0 = e, p1 = e;  // past r_hi: as an empty row at the tile's end
    if (r <= r_hi) {
      p0 = a.pos[r];
      p1 = a.pos[r + 1];
    }
    const bool inside = p0 >= s && p0 < e;
    const bool intr = inside && p1 <= e && p1 > p0;
    const bool tl = inside && p1 > e;
    head = head || __ballot(p0 < s && p1 > s) != 0;
    tail = tail || __ballot(tl) != 0;
    const uint64_t im = __ballot(intr);
    const int32_t k = count + (int32_t)__builtin_amdgcn_mbcnt_hi(
                                  (uint32_t)(im >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)im, 0u));
    if (intr) {
      a.int_rows[(int64_t)t * a.tile + k] = r;
      mk[p0 - s] = k + 1;
    }
    if (tl) mk[p0 - s] = kCsTailMark;
    count += __popcll(im);
    if (__builtin_amdgcn_readlane(p1, 63) >= e) break;  // later rows begin past the tile
  }
  wave_sync();
  int32_t carry = 0;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int32_t i = lane + 64 * k;
    const int32_t m = max(wave_scan<true>(i < cn ? mk[i] : 0), carry);
    carry = __builtin_amdgcn_readlane(m, 63);
    if (i < cn)
      a.codes[s + i] = (uint16_t)(m == 0 ? 0 : m == kCsTailMark ? 1 : min(m, count) + 1);
  }
  if (lane == 0) {
    a.tile_count[t] = count;
    a.tile_flags[t] = (uint8_t)((head ? 1 : 0) | (tail ? 2 : 0));
  }
}

// k_cs_sort: the unit's slot plan (cs_slot_base), then the records. A
// walking tile keeps CSR order and the tile-local codes. A sweeping tile maps
// its codes to LDS slots, writes the rows of its interior slots, and sorts
// its entries by (column, position): each entry goes to one of 64 buckets of
// the tile's column range (a monotone map of the column), a count and a
// prefix sum place the buckets, and lane b sorts bucket b in registers (a
// bucket above kCsBucketMax keys sends the tile to a whole-wave bitonic
// sort). The keys (column << 9 | place) are unique, so the order does not
// depend on the order the buckets were filled in. Every load is issued
// before anything waits.
template <int E>
__global__ __launch_bounds__(kCsPrepThreads, E <= 6 ? 8 : 5) void k_cs_sort(const CsArgs a) {
  constexpr int R = E <= 1 ? 1 : E <= 2 ? 2 : E <= 4 ? 4 : 8;  // whole-wave sort registers
  __shared__ uint32_t s_key[kCsPrepWaves][64 * E];
  __shared__ uint2 s_rec[kCsPrepWaves][64 * E];
  __shared__ int32_t s_bkt[kCsPrepWaves][kCsBuckets];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar tile math
  const int32_t t = blockIdx.x * kCsPrepWaves + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  const int32_t unit = t / a.group;
  const int32_t g = t - unit * a.group;
  const int32_t s = t * a.tile;
  const int32_t cn = min(a.nnz - s, a.tile);
  const uint16_t *codes = a.codes + s;
  const int32_t *crd = a.crd + s;
  const int32_t *irows = a.int_rows + (int64_t)t * a.tile;
  int32_t cd[E], col[E], irow[E];
  uint32_t vb[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int32_t i = min(lane + 64 * k, cn - 1);
    cd[k] = codes[i];
    col[k] = crd[i];
    irow[k] = irows[i];
  }
  // value bits: the storage type is uniform, so one branch around all loads
  if (a.vals == nullptr) {
#pragma unroll
    for (int k = 0; k < E; ++k) vb[k] = a.one_bits;  // A without values reads 1.0
  } else if (a.elem == 4) {
    const uint32_t *v = static_cast<const uint32_t *>(a.vals) + s;
#pragma unroll
    for (int k = 0; k < E; ++k) vb[k] = v[min(lane + 64 * k, cn - 1)];
  } else {
    const uint16_t *v = static_cast<const uint16_t *>(a.vals) + s;
#pragma unroll
    for (int k = 0; k < E; ++k) vb[k] = v[min(lane + 64 * k, cn - 1)];
  }
  int32_t n_slots = 0, count = 0;
  const int32_t base = cs_slot_base(a, t, n_slots, count);
  const bool walk = base < 0;
  if (g == 0 && lane == 0) a.n_slots[unit] = n_slots;
  auto put = [&](int32_t d, int32_t c, int32_t code, uint32_t v) {
    const int64_t dst = (int64_t)s + d;
    if (a.packed)
      static_cast<uint2 *>(a.s_rec)[dst] =
          make_uint2(((uint32_t)c << kCsCodeBits) | (uint32_t)code, v);
    else
      static_cast<CsWide *>(a.s_rec)[dst] = CsWide{(uint32_t)c, (uint32_t)code, v};
  };
  if (walk) {
    if (lane == 0) a.tile_flags[t] |= 4;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int32_t i = lane + 64 * k;
      if (i < cn) put(i, col[k], cd[k], vb[k]);
    }
    return;
  }
  // sweeping: tile-local codes to LDS slots, and the rows of its interior
  // slots (interior row i of the tile holds slot base + i)
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int32_t i = lane + 64 * k;
    if (i < count) a.slot_rows[(int64_t)unit * a.cap + base + i] = irow[k];
    cd[k] = cd[k] == 0 ? 2 * g : cd[k] == 1 ? 2 * g + 1 : base + cd[k] - 2;
  }
  if (a.packed) {
    // columns below 2^(32 - kCsCodeBits) <= 2^20: the key (column << 9 |
    // place) is unique and fits 29 bits; the records wait in LDS by place
    uint32_t *sk = s_key[wave];
    uint2 *sr = s_rec[wave];
    int32_t *nb = s_bkt[wave];
    int32_t hi = 0, flip = 0;  // largest column; largest INT_MAX - column
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int32_t i = lane + 64 * k;
      if (i < cn) {
        hi = max(hi, col[k]);
        flip = max(flip, INT_MAX - col[k]);
        sr[i] = make_uint2(((uint32_t)col[k] << kCsCodeBits) | (uint32_t)cd[k], vb[k]);
      }
    }
    hi = __builtin_amdgcn_readlane(wave_scan<true>(hi), 63);
    const int32_t lo = INT_MAX - __builtin_amdgcn_readlane(wave_scan<true>(flip), 63);
    // bucket of a column: (column - lo) * 64 / span, rounded down in FP32 —
    // exact inputs (span <= 2^20) and a positive factor keep it monotone
    const float scale = (float)kCsBuckets / (float)(hi - lo + 1);
    int32_t bk[E];
#pragma unroll
    for (int k = 0; k < E; ++k)
      bk[k] = min(kCsBuckets - 1, (int32_t)((float)(col[k] - lo) * scale));
    nb[lane] = 0;
    wave_sync();
#pragma unroll
    for (int k = 0; k < E; ++k)
      if (lane + 64 * k < cn) atomicAdd(&nb[bk[k]], 1);
    wave_sync();
    const int32_t n = nb[lane];                   // keys in bucket `lane`
    const int32_t st = wave_scan<false>(n) - n;   // its first place
    const int32_t n_max = __builtin_amdgcn_readlane(wave_scan<true>(n), 63);
    nb[lane] = st;
    wave_sync();
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int32_t i = lane + 64 * k;
      if (i < cn) sk[atomicAdd(&nb[bk[k]], 1)] = ((uint32_t)col[k] << kCsIdxBits) | (uint32_t)i;
    }
    wave_sync();
    // the network for the fullest bucket (19, 32, 42 or 63 compare-exchanges)
    if (n_max <= 8) {
      cs_sort_bucket<8>(sk, st, n);
    } else if (n_max <= 10) {
      cs_sort_bucket<10>(sk, st, n);
    } else if (n_max <= 12) {
      cs_sort_bucket<12>(sk, st, n);
    } else if (n_max <= kCsBucketMax) {
      cs_sort_bucket<kCsBucketMax>(sk, st, n);
    } else {
      uint32_t key[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int32_t i = lane + 64 * r;
        key[r] = 0xffffffffu;
        if (r < E && i < cn) key[r] = ((uint32_t)col[r] << kCsIdxBits) | (uint32_t)i;
      }
      wave_bitonic<R>(key);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int32_t d = lane + 64 * r;
        if (d < cn) sk[d] = key[r];
      }
    }
    wave_sync();
    uint2 *out = static_cast<uint2 *>(a.s_rec) + s;
#pragma unroll
    for (int k = 0; k < E; ++k) {
      const int32_t d = lane + 64 * k;
      if (d < cn) out[d] = sr[sk[d] & (kCsMaxTile - 1)];
    }
    return;
  }
  // wide columns: rank of each entry in (column, position) order
  int32_t rank[E];
#pragma unroll
  for (int k = 0; k < E; ++k) rank[k] = 0;
#pragma unroll
  for (int k2 = 0; k2 < E; ++k2) {
    if (64 * k2 >= cn) break;  // uniform
    const int32_t lim = min(64, cn - 64 * k2);
    for (int32_t l = 0; l < lim; ++l) {
      const int32_t cj = __builtin_amdgcn_readlane(col[k2], l);
      const int32_t j = 64 * k2 + l;
#pragma unroll
      for (int k = 0; k < E; ++k)
        rank[k] += (cj < col[k] || (cj == col[k] && j < lane + 64 * k)) ? 1 : 0;
    }
  }
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int32_t i = lane + 64 * k;
    if (i < cn) put(rank[k], col[k], cd[k], vb[k]);
  }
}

// Record of sorted entry i: column, slot and value bits, one load.
template <bool kPacked>
__device__ __forceinline__ void cs_rec(const void *rec, int32_t i, int32_t &col,
                                       int32_t &code, uint32_t &vbits) {
  if constexpr (kPacked) {
    const uint2 w = static_cast<const uint2 *>(rec)[i];
    col = (int32_t)(w.x >> kCsCodeBits);
    code = (int32_t)(w.x & ((1u << kCsCodeBits) - 1));
    vbits = w.y;
  } else {
    const CsWide w = static_cast<const CsWide *>(rec)[i];
    col = (int32_t)w.col;
    code = (int32_t)w.code;
    vbits = w.vbits;
  }
}

// Value bits (FP32, or 16-bit storage zero-extended) widened to FP32.
template <int DT>
__device__ __forceinline__ float cs_widen(uint32_t r) {
  if constexpr (DT == 0)
    return __uint_as_float(r);
  else
    return half_to_f32<DT>(r);
}

// Entries per batch: metadata of batch b + 2, gathers of batch b + 1 and
// the accumulation of batch b are in flight together. Measured on the
// power-law 262144^2 product at N = 128 (prepared run, us per call): 1: 408,
// 2: 274, 3: 260, 4: 272, 5: 292, 6: 298, 7: 303, 8: 303 — small batches
// keep each wave's gathers and LDS updates interleaved more finely.
template <bool kPacked>
constexpr int cs_batch() {
  return 3;
}

template <int NB>
struct CsBatch {
  static constexpr int kN = NB;
  int32_t c[NB], s[NB];
  uint32_t v[NB];
};

// The pipelined loop shared by the sweep and the walk over entries [s, e):
// Every load is unconditional (entries past the tile re-read its last
// entry; columns past N read column 0; neither is used), so the loads in
// flight are the same on every path and each wait covers only its batch:
// loads complete in issue order, and per iteration the metadata of batch
// b + 2 goes out before the gathers of batch b + 1, so waiting for the
// metadata those gathers need (issued an iteration earlier) leaves the
// gathers of batch b in flight. Three metadata buffers and two gather
// buffers rotate by name through an unrolled cycle of six steps: a copy of a
// register whose load is still in flight would make the wave wait for that
// load, so no buffer is copied. A step consumes batch b (metadata X, gathers
// GX), gathers batch b + 1 (metadata Y into GY) and loads the metadata of
// batch b + 2 into Z. G is the gather buffer type, filled by gather(m, g)
// and read by consume(m, g, i0).
template <bool kPacked, typename M, typename G, typename Gather, typename Consume>
__device__ __forceinline__ void cs_pipeline(const CsArgs &a, int32_t s, int32_t e,
                                            Gather &&gather, Consume &&consume) {
  if (s >= e) return;
  const int32_t last = e - 1;
  constexpr int kB = M::kN;
  auto load_meta = [&](M &m, int32_t i0) {
#pragma unroll
    for (int u = 0; u < kB; ++u)
      cs_rec<kPacked>(a.s_rec, min(i0 + u, last), m.c[u], m.s[u], m.v[u]);
  };
  M mA, mB, mC;
  G gA, gB;
  load_meta(mA, s);
  load_meta(mB, s + kB);  // past the tile: clamped, unused
  gather(mA, gA);
  int32_t i0 = s;
  auto step = [&](M &X, G &GX, M &Y, G &GY, M &Z) {
    load_meta(Z, i0 + 2 * kB);
    gather(Y, GY);
    consume(X, GX, i0);
    i0 += kB;
    return i0 < e;
  };
  for (;;) {
    if (!step(mA, gA, mB, gB, mC)) break;
    if (!step(mB, gB, mC, gA, mA)) break;
    if (!step(mC, gA, mA, gB, mB)) break;
    if (!step(mA, gB, mB, gA, mC)) break;
    if (!step(mB, gA, mC, gB, mA)) break;
    if (!step(mC, gB, mA, gA, mB)) break;
  }
}

// Sweep of a sorted tile: each entry's product added to its row's LDS slot.
template <int DT, int V, bool kPacked>
__device__ __forceinline__ void cs_sweep(const CsArgs &a, float *acc_lds, int32_t s,
                                         int32_t e, int32_t col, bool colok) {
  constexpr int32_t Q = kCsLanes;
  constexpr int32_t W = Q * V;
  const int32_t q = (threadIdx.x & 63) & (Q - 1);
  constexpr int kB = cs_batch<kPacked>();
  using M = CsBatch<kB>;
  using R = typename RawT<DT, V>::type;
  struct G {
    R b[kB];
  };
  const int32_t col_ld = colok ? col : 0;
  auto gather = [&](const M &m, G &g) {
#pragma unroll
    for (int u = 0; u < kB; ++u)
      g.b[u] = load_raw<DT, V>(a.B, (int64_t)m.c[u] * a.ldb + col_ld);
  };
  auto consume = [&](const M &m, const G &g, int32_t i0) {
    const int32_t nb = min(kB, e - i0);
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      if (u < nb && colok) {
        const float v = cs_widen<DT>(m.v[u]);
        float b[V];
        unpack<DT, V>(g.b[u], b);
        float *p = acc_lds + m.s[u] * W + q * V;
        float x[V];
        load_f32vec<V>(p, x);
#pragma unroll
        for (int k = 0; k < V; ++k) x[k] = fmaf(v, b[k], x[k]);
        store_vec<V>(p, x);
      }
    }
  };
  cs_pipeline<kPacked, M, G>(a, s, e, gather, consume);
}

// Walk of a tile in CSR order (its wave did not fit the LDS slots): a
// register accumulator per row, stored when the row code changes — code 0
// to the head carry, 1 to the tail carry, 2 + k to C's row int_rows[k] of
// the tile (loaded with the gathers, so it is in registers by then).
template <int DT, int V, bool kPacked>
__device__ __forceinline__ void cs_walk(const CsArgs &a, int32_t t, int32_t s, int32_t e,
                                        int32_t col, bool colok) {
  constexpr int kB = cs_batch<kPacked>();
  using M = CsBatch<kB>;
  using R = typename RawT<DT, V>::type;
  struct G {
    R b[kB];
    int32_t row[kB];
  };
  const int32_t col_ld = colok ? col : 0;
  const int32_t *rows = a.int_rows + (int64_t)t * a.tile;
  auto gather = [&](const M &m, G &g) {
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      g.b[u] = load_raw<DT, V>(a.B, (int64_t)m.c[u] * a.ldb + col_ld);
      g.row[u] = rows[max(m.s[u] - 2, 0)];
    }
  };
  float acc[V];
  zero_vec(acc);
  int32_t cur = -1, row = 0;
  auto flush = [&]() {
    if (cur < 0 || !colok) return;
    float *dst = cur == 0   ? a.carry_head + (int64_t)t * a.N + col
                 : cur == 1 ? a.carry_tail + (int64_t)t * a.N + col
                            : a.C + (int64_t)row * a.ldc + col;
    store_vec<V>(dst, acc);
  };
  auto consume = [&](const M &m, const G &g, int32_t i0) {
    const int32_t nb = min(kB, e - i0);
#pragma unroll
    for (int u = 0; u < kB; ++u) {
      if (u < nb) {
        if (m.s[u] != cur) {
          flush();
          cur = m.s[u];
          row = g.row[u];
          zero_vec(acc);
        }
        const float v = cs_widen<DT>(m.v[u]);
        float b[V];
        unpack<DT, V>(g.b[u], b);
#pragma unroll
        for (int k = 0; k < V; ++k) acc[k] = fmaf(v, b[k], acc[k]);
      }
    }
  };
  cs_pipeline<kPacked, M, G>(a, s, e, gather, consume);
  flush();
}

template <int DT, int V, bool kPacked>
__global__ __launch_bounds__(kCsThreads) void k_eb_cs(const CsArgs a) {
  extern __shared__ float acc_lds[];  // [cap][W] accumulators, then [cap] slot rows
  constexpr int32_t Q = kCsLanes;
  constexpr int32_t W = Q * V;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t q = lane & (Q - 1);
  const int32_t g = wave * (64 / Q) + (lane / Q);  // group in workgroup
  const int32_t panels = (a.N + W - 1) / W;

  for (int32_t round = 0; round < a.rounds; ++round) {
    const int32_t unit = blockIdx.x * a.rounds + round;
    if (unit >= a.units) break;  // uniform over the workgroup
    const int32_t t = unit * a.group + g;
    const bool tok = g < a.group && t < a.ntiles;
    const int32_t n_slots = a.n_slots[unit];
    const int32_t s = tok ? t * a.tile : 0;
    const int32_t e = tok ? (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile) : 0;
    const int32_t flags = tok ? a.tile_flags[t] : 0;
    const bool walk = (flags & 4) != 0;  // the same for the wave's tiles
    // A sweeping tile's head and tail rows go to the carries unless the
    // unit writes them itself (cs_row_folded): then the tile where such a
    // row ends sums its slots after each panel, in the fix-up's order.
    bool head_carry = false, tail_carry = false;
    int32_t fold_t0 = -1, fold_row = 0;
    if (tok && !walk) {
      if (flags & 1) {  // row holding entry t * T, begun in an earlier tile
        const int32_t r = a.tile_row[t];
        const int32_t t0 = a.pos[r] / a.tile, t1 = (a.pos[r + 1] - 1) / a.tile;
        if (!cs_row_folded(a.tile_flags, a.group, t0, t1)) {
          head_carry = true;
        } else if (t1 == t) {
          fold_t0 = t0;
          fold_row = r;
        }
      }
      if (flags & 2) {  // row holding entry (t + 1) * T, begun in this tile
        const int32_t r = a.tile_row[t + 1];
        tail_carry = !cs_row_folded(a.tile_flags, a.group, t, (a.pos[r + 1] - 1) / a.tile);
      }
    }
    // the rows of the interior slots, read once for all panels
    int32_t *s_rows = reinterpret_cast<int32_t *>(acc_lds + (size_t)a.cap * W);
    for (int32_t i = threadIdx.x; i < n_slots - 2 * a.group; i += kCsThreads)
      s_rows[i] = a.slot_rows[(int64_t)unit * a.cap + 2 * a.group + i];
    // slots start at zero; each panel's write-out zeroes what it read
    for (int32_t i = threadIdx.x; i < n_slots * W / 4; i += kCsThreads)
      reinterpret_cast<float4 *>(acc_lds)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    float z[V];
    zero_vec(z);
    for (int32_t panel = 0; panel < panels; ++panel) {
      const int32_t col = panel * W + q * V;
      const bool colok = col < a.N;
      if (walk)
        cs_walk<DT, V, kPacked>(a, t, s, e, col, colok);
      else
        cs_sweep<DT, V, kPacked>(a, acc_lds, s, e, col, colok);
      __syncthreads();
      // head / tail slots of sweeping tiles to the carries, or rows that
      // end in this tile summed from the unit's slots (each slot is read
      // and zeroed by one lane group: its own, or the group that sums it)
      if (tok && colok && !walk) {
        float x[V];
        if (head_carry) {
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[(2 * g) * W + q * V + k];
          store_vec<V>(a.carry_head + (int64_t)t * a.N + col, x);
          store_vec<V>(acc_lds + (2 * g) * W + q * V, z);
        }
        if (tail_carry) {
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[(2 * g + 1) * W + q * V + k];
          store_vec<V>(a.carry_tail + (int64_t)t * a.N + col, x);
          store_vec<V>(acc_lds + (2 * g + 1) * W + q * V, z);
        }
        if (fold_t0 >= 0) {
          const int32_t u0 = unit * a.group;
          float *sl = acc_lds + (2 * (fold_t0 - u0) + 1) * W + q * V;
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = sl[k];
          store_vec<V>(sl, z);
          for (int32_t tt = fold_t0 + 1; tt <= t; ++tt) {
            sl = acc_lds + (2 * (tt - u0)) * W + q * V;
#pragma unroll
            for (int k = 0; k < V; ++k) x[k] += sl[k];
            store_vec<V>(sl, z);
          }
          store_vec<V>(a.C + (int64_t)fold_row * a.ldc + col, x);
        }
      }
      // interior slots to C
      constexpr int32_t PER = W / V;  // vectors per slot
      const int32_t first = 2 * a.group;
      for (int32_t i = threadIdx.x; i < (n_slots - first) * PER; i += kCsThreads) {
        const int32_t sl = first + i / PER;
        const int32_t seg = i % PER;
        const int32_t c = panel * W + seg * V;
        if (c < a.N) {
          const int32_t row = s_rows[sl - first];
          float x[V];
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[sl * W + seg * V + k];
          store_vec<V>(a.C + (int64_t)row * a.ldc + c, x);
          store_vec<V>(acc_lds + sl * W + seg * V, z);
        }
      }
      __syncthreads();
    }
  }
}

}  // namespace

hipError_t launch_cs_prep(const CsArgs &a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  const dim3 g((unsigned)((a.ntiles + kCsPrepWaves - 1) / kCsPrepWaves)), b(kCsPrepThreads);
#define SEGTACO_CS_PREP(E_)                                 \
  case E_:                                                  \
    hipLaunchKernelGGL(k_cs_count<E_>, g, b, 0, s, a);      \
    hipLaunchKernelGGL(k_cs_sort<E_>, g, b, 0, s, a);       \
    break;
  switch ((a.tile + 63) / 64) {
    SEGTACO_CS_PREP(1)
    SEGTACO_CS_PREP(2)
    SEGTACO_CS_PREP(3)
    SEGTACO_CS_PREP(4)
    SEGTACO_CS_PREP(5)
    SEGTACO_CS_PREP(6)
    SEGTACO_CS_PREP(7)
    SEGTACO_CS_PREP(8)
    default: return hipErrorInvalidValue;
  }
#undef SEGTACO_CS_PREP
  return hipGetLastError();
}

size_t cs_lds_bytes(int32_t vec, int32_t cap) {
  return (size_t)cap * (kCsLanes * vec * sizeof(float) + sizeof(int32_t));
}

template <int DT, int V, bool kPacked>
hipError_t launch_eb_cs_t(const CsArgs &a, size_t lds, hipStream_t s) {
  // dynamic LDS above the default per-kernel limit has to be declared; the
  // capacity is fixed per device, so once per instantiation is enough
  static const hipError_t declared = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&k_eb_cs<DT, V, kPacked>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (declared != hipSuccess) return declared;
  const unsigned grid = (unsigned)((a.units + a.rounds - 1) / a.rounds);
  hipLaunchKernelGGL((k_eb_cs<DT, V, kPacked>), dim3(grid), dim3(kCsThreads), lds, s, a);
  return hipGetLastError();
}

template <int DT, int V>
hipError_t launch_eb_cs_p(const CsArgs &a, size_t lds, hipStream_t s) {
  return a.packed ? launch_eb_cs_t<DT, V, true>(a, lds, s)
                  : launch_eb_cs_t<DT, V, false>(a, lds, s);
}

hipError_t launch_eb_cs(int dt, int vec, const CsArgs &a, hipStream_t s) {
  if (a.units <= 0) return hipSuccess;
  const size_t lds = cs_lds_bytes(vec, a.cap);
  switch (dt * 16 + vec) {
    case 0 * 16 + 1: return launch_eb_cs_p<0, 1>(a, lds, s);
    case 0 * 16 + 2: return launch_eb_cs_p<0, 2>(a, lds, s);
    case 0 * 16 + 4: return launch_eb_cs_p<0, 4>(a, lds, s);
    case 1 * 16 + 1: return launch_eb_cs_p<1, 1>(a, lds, s);
    case 1 * 16 + 2: return launch_eb_cs_p<1, 2>(a, lds, s);
    case 1 * 16 + 4: return launch_eb_cs_p<1, 4>(a, lds, s);
    case 2 * 16 + 1: return launch_eb_cs_p<2, 1>(a, lds, s);
    case 2 * 16 + 2: return launch_eb_cs_p<2, 2>(a, lds, s);
    case 2 * 16 + 4: return launch_eb_cs_p<2, 4>(a, lds, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace segtaco

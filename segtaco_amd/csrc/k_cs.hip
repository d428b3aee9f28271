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
//   * k_eb_cs, per unit and per column panel of W = 8*V columns: zero the
//     unit's LDS slots, let every sweeping lane group run its sorted tile —
//     per entry one 16-byte gather per lane and one read-modify-write of its
//     row's slot, with no branch on the kind of row — and every walking group
//     run its tile in CSR order with register accumulators, written out at
//     each change of row; then write interior slots to C and head / tail
//     slots to the carry arrays. k_tile_rows_fix zeroes the empty rows and
//     k_fixup combines the carries, as for eb_sr.
// Each row's entries meet its accumulator in ascending column order (sweep)
// or CSR order (walk), the same order for canonical CSR, with one fmaf per
// entry starting from zero: C is bit-identical to eb_sr with the same tile
// size. Walking is decided per wave so that no wave runs both paths.
#include <climits>

#include "cs.h"
#include "device_common.h"

namespace segtaco {

namespace {

constexpr int32_t kCsPosSlack = 8;
constexpr int kCsPrepWaves = kCsPrepThreads / 64;
// Bits of an entry's place in its tile inside a sort key (kCsMaxTile = 2^9).
constexpr int kCsIdxBits = 9;
static_assert((1 << kCsIdxBits) == kCsMaxTile, "sort key index field");

// Record of one entry when columns do not fit the packed form (one 12-byte
// load).
struct CsWide {
  uint32_t col, code, vbits;
};

// Row of entry p given the staged window s_pos[0..pw) = pos[rbase..] (falls
// back to a bisection of pos in [rbase + pw - 1, r_hi] past the window).
__device__ __forceinline__ int32_t cs_row(const int32_t *s_pos, int32_t pw,
                                          int32_t rbase, const int32_t *pos,
                                          int32_t r_hi, int32_t p) {
  int32_t lo = 0, hi = pw - 1;
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (s_pos[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return (lo == pw - 1) ? row_of(pos, rbase + lo, r_hi, p) : rbase + lo;
}

__device__ __forceinline__ int32_t cs_P(const int32_t *s_pos, int32_t pw,
                                        int32_t rbase, const int32_t *pos,
                                        int32_t row) {
  const int32_t i = row - rbase;
  return i < pw ? s_pos[i] : pos[row];
}

// Wave-wide inclusive prefix sum of x (64 lanes).
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// Wave-wide sum of x.
__device__ __forceinline__ int32_t wave_sum(int32_t x) {
  for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
  return x;
}

// One tile as the prep kernel sees it: each lane holds entries lane + 64k
// (k < E), their rows, and whether the tile has a head / tail row.
template <int E>
struct PrepTile {
  int32_t s, e, cn;
  int32_t pw;  // rows staged in the wave's window from r_first
  int32_t row[E];
  bool head, tail;
  int32_t r_first, r_last;
};

// Rows of tile t's entries and its head / tail rows. s_pos: this wave's
// LDS window of the row pointer (kCsMaxTile + kCsPosSlack entries).
template <int E>
__device__ void cs_tile_rows(const CsArgs &a, int32_t t, int32_t *s_pos, PrepTile<E> &P) {
  const int lane = threadIdx.x & 63;
  const int32_t *__restrict__ pos = a.pos;
  P.s = t * a.tile;
  P.e = (int32_t)min((int64_t)a.nnz, (int64_t)P.s + a.tile);
  P.cn = P.e - P.s;
  const int32_t rbase = a.tile_row[t];
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  const int32_t pw = min(a.tile + kCsPosSlack, a.M + 1 - rbase);
  P.pw = pw;
  for (int32_t i = lane; i < pw; i += 64) s_pos[i] = pos[rbase + i];
  wave_sync();
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int32_t i = lane + 64 * k;
    P.row[k] = i < P.cn ? cs_row(s_pos, pw, rbase, pos, r_hi, P.s + i) : INT_MAX;
  }
  P.r_first = rbase;  // the repaired row of the tile's first entry
  // the row of the last entry: the lane / register holding entry cn - 1
  const int32_t last = P.cn - 1;
  int32_t rl = 0;
#pragma unroll
  for (int k = 0; k < E; ++k)
    if (last / 64 == k) rl = __shfl(P.row[k], last & 63, 64);
  P.r_last = rl;
  P.head = cs_P(s_pos, pw, rbase, pos, P.r_first) < P.s;
  P.tail = !(P.head && P.r_last == P.r_first) && cs_P(s_pos, pw, rbase, pos, P.r_last + 1) > P.e;
  // callers read s_pos again: it stays valid until the next tile
}

// Interior rows of the tile: those that start and end inside it. Entry i
// starts one when its row begins at i and is neither the head nor the tail.
template <int E>
__device__ __forceinline__ bool cs_starts_interior(const CsArgs &a, const PrepTile<E> &P,
                                                   const int32_t *s_pos, int k) {
  const int lane = threadIdx.x & 63;
  const int32_t i = lane + 64 * k;
  if (i >= P.cn) return false;
  const int32_t r = P.row[k];
  if (cs_P(s_pos, P.pw, P.r_first, a.pos, r) != P.s + i) return false;
  if (P.head && r == P.r_first) return false;
  if (P.tail && r == P.r_last) return false;
  return true;
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
  // interior rows of the tiles before g in g's wave, and of each wave
  const int32_t w_g = g >> 3;
  const int32_t before = wave_sum((lane >> 3) == w_g && lane < g ? cnt : 0);
  int32_t seg = cnt;
  for (int d = 4; d > 0; d >>= 1) seg += __shfl_xor(seg, d, 64);
  const int32_t room = a.cap - 2 * a.group;
  int32_t used = 0, base = -1;
  for (int w = 0; w < 8; ++w) {
    const int32_t sw = __shfl(seg, 8 * w, 64);
    if (used + sw <= room) {
      if (w == w_g) base = 2 * a.group + used + before;
      used += sw;
    }
  }
  n_slots = 2 * a.group + used;
  count = __shfl(cnt, g, 64);
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

// Prep, one wave per tile (many waves per CU hide the latency of the
// per-entry row searches). k_cs_count: per tile its head / tail flags, its
// interior rows (count, and the row of the k-th in int_rows) and per entry
// the tile-local code of its row (0 head, 1 tail, 2 + k the k-th interior
// row). k_cs_sort: the unit's slot plan (cs_slot_base), then the records:
// walking tiles in CSR order with the tile-local codes, sweeping tiles
// sorted by (column, position) with the codes mapped to LDS slots, and the
// rows of the interior slots. Every load k_cs_sort makes is issued before
// anything waits, so the kernel costs one memory round trip.
template <int E>
__global__ __launch_bounds__(kCsPrepThreads) void k_cs_count(const CsArgs a) {
  __shared__ int32_t s_pos[kCsPrepWaves][kCsMaxTile + kCsPosSlack];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t t = blockIdx.x * kCsPrepWaves + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  int32_t *sp = s_pos[wave];
  PrepTile<E> P;
  cs_tile_rows<E>(a, t, sp, P);
  int32_t code[E];
  int32_t running = 0;  // interior rows started by earlier chunks
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const bool st = cs_starts_interior<E>(a, P, sp, k);
    const int32_t inc = wave_incl_scan(st ? 1 : 0, lane) + running;
    running = __shfl(inc, 63, 64);
    // the entry that starts an interior row names the row's place k in the
    // tile; the row's other entries take it from the running maximum below
    code[k] = st ? inc - 1 : -1;
    if (st) a.int_rows[(int64_t)t * a.tile + inc - 1] = P.row[k];
  }
  int32_t carry = -1;
#pragma unroll
  for (int k = 0; k < E; ++k) {
    int32_t m = code[k];
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t y = __shfl_up(m, d, 64);
      if (lane >= d) m = max(m, y);
    }
    m = max(m, carry);
    carry = __shfl(m, 63, 64);
    const int32_t i = lane + 64 * k;
    const int32_t r = P.row[k];
    const bool hd = P.head && r == P.r_first;
    const bool tl = !hd && P.tail && r == P.r_last;
    if (i < P.cn) a.codes[P.s + i] = (uint16_t)(hd ? 0 : tl ? 1 : 2 + m);
  }
  if (lane == 0) {
    a.tile_count[t] = running;
    a.tile_flags[t] = (uint8_t)((P.head ? 1 : 0) | (P.tail ? 2 : 0));
  }
}

template <int E>
__global__ __launch_bounds__(kCsPrepThreads) void k_cs_sort(const CsArgs a) {
  constexpr int R = E <= 1 ? 1 : E <= 2 ? 2 : E <= 4 ? 4 : 8;  // sort registers
  __shared__ int32_t s_code[kCsPrepWaves][kCsMaxTile];
  __shared__ uint32_t s_val[kCsPrepWaves][kCsMaxTile];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t t = blockIdx.x * kCsPrepWaves + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  const int32_t unit = t / a.group;
  const int32_t g = t - unit * a.group;
  const int32_t s = t * a.tile;
  const int32_t cn = (int32_t)min((int64_t)a.nnz - s, (int64_t)a.tile);
  // every load of the tile at once: codes, columns, values, interior rows
  int32_t cd[E], col[E], irow[E];
  uint32_t vb[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int32_t i = min(lane + 64 * k, cn - 1);
    cd[k] = a.codes[s + i];
    col[k] = a.crd[s + i];
    vb[k] = a.one_bits;  // A without values reads 1.0
    if (a.vals)
      vb[k] = a.elem == 4 ? static_cast<const uint32_t *>(a.vals)[s + i]
                          : (uint32_t) static_cast<const uint16_t *>(a.vals)[s + i];
    irow[k] = a.int_rows[(int64_t)t * a.tile + i];
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
    // place) is unique, fits 29 bits and orders by (column, position);
    // places past the tile sort last
    int32_t *sc = s_code[wave];
    uint32_t *sv = s_val[wave];
    uint32_t key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t i = lane + 64 * r;
      key[r] = 0xffffffffu;
      if (r < E && i < cn) {
        key[r] = ((uint32_t)col[r] << kCsIdxBits) | (uint32_t)i;
        sc[i] = cd[r];
        sv[i] = vb[r];
      }
    }
    wave_sync();
    wave_bitonic<R>(key);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int32_t d = lane + 64 * r;
      if (d < cn) {
        const int32_t i = (int32_t)(key[r] & (kCsMaxTile - 1));
        put(d, (int32_t)(key[r] >> kCsIdxBits), sc[i], sv[i]);
      }
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
// the accumulation of batch b are in flight together. Wide records hold
// three words, so their batches are smaller (all buffers stay in registers).
template <bool kPacked>
constexpr int cs_batch() {
  return kPacked ? 8 : 6;
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
  extern __shared__ float acc_lds[];  // [cap][W]
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
    for (int32_t panel = 0; panel < panels; ++panel) {
      const int32_t col = panel * W + q * V;
      const bool colok = col < a.N;
      for (int32_t i = threadIdx.x; i < n_slots * W / 4; i += kCsThreads)
        reinterpret_cast<float4 *>(acc_lds)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      __syncthreads();
      if (walk)
        cs_walk<DT, V, kPacked>(a, t, s, e, col, colok);
      else
        cs_sweep<DT, V, kPacked>(a, acc_lds, s, e, col, colok);
      __syncthreads();
      // head / tail slots of sweeping tiles to the carries
      if (tok && colok && !walk) {
        float x[V];
        if (flags & 1) {
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[(2 * g) * W + q * V + k];
          store_vec<V>(a.carry_head + (int64_t)t * a.N + col, x);
        }
        if (flags & 2) {
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[(2 * g + 1) * W + q * V + k];
          store_vec<V>(a.carry_tail + (int64_t)t * a.N + col, x);
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
          const int32_t row = a.slot_rows[(int64_t)unit * a.cap + sl];
          float x[V];
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[sl * W + seg * V + k];
          store_vec<V>(a.C + (int64_t)row * a.ldc + c, x);
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
  return (size_t)cap * kCsLanes * vec * sizeof(float);
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

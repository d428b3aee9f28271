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
//   * tiles are the eb_sr tiles (T consecutive entries, the K21 partition
//     and its repaired first row, k_block_starts); a unit is G consecutive
//     tiles, one per lane group of one 512-thread workgroup;
//   * once per call, k_cs_count and k_cs_sort (one wave per tile) classify
//     each entry's row as head (began in an earlier tile), tail (continues
//     into a later tile) or interior, number the unit's interior rows (LDS
//     slots), and write each tile's entries sorted by (column, position)
//     with their slot codes;
//   * k_eb_cs, per unit and per column panel of W = 8*V columns: zero the
//     unit's LDS accumulators, let every lane group sweep its sorted tile —
//     interior rows accumulate in LDS, head and tail rows in registers —
//     then write the unit's interior rows to C and the head/tail partials to
//     the carry arrays; as for eb_sr, k_tile_rows_fix (k_eb.hip) zeroes the
//     empty rows and k_fixup (sched_eb.h) combines the carries.
// Each row's entries still meet its accumulator in ascending column order,
// which is CSR order for canonical CSR, with one fmaf per entry starting from
// zero: C is bit-identical to eb_sr with the same tile size. A unit whose
// interior rows exceed the LDS capacity runs its tiles in CSR order with
// register accumulators instead (same arithmetic).
#include <climits>

#include "cs.h"
#include "device_common.h"

namespace segtaco {

namespace {

constexpr int32_t kCsPosSlack = 8;

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

// Per-wave LDS of k_cs_prep (tile <= kCsMaxTile entries).
struct PrepLds {
  uint64_t key[kCsMaxTile];
  int32_t pos[kCsMaxTile + kCsPosSlack];
  int32_t row[kCsMaxTile];
  int32_t slot[kCsMaxTile];
};

// Wave-wide inclusive prefix sum of x (64 lanes).
__device__ __forceinline__ int32_t wave_incl_scan(int32_t x, int lane) {
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// Tile t's rows (phase A) or rows + slot codes + sorted output (phase B).
// Returns the tile's interior row count and sets *flags (bit 0 head, bit 1
// tail). Uniform over the wave.
template <bool kWrite>
__device__ int32_t cs_prep_tile(const CsArgs &a, int32_t t, PrepLds &L,
                                int32_t base, int32_t unit, int32_t *flags) {
  const int lane = threadIdx.x & 63;
  const int32_t *__restrict__ pos = a.pos;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  const int32_t cn = e - s;
  const int32_t rbase = a.tile_row[t];
  const int32_t r_hi = min(a.M - 1, a.tile_row[t + 1]);
  const int32_t pw = min(a.tile + kCsPosSlack, a.M + 1 - rbase);
  for (int32_t i = lane; i < pw; i += 64) L.pos[i] = pos[rbase + i];
  wave_sync();
  for (int32_t i = lane; i < cn; i += 64)
    L.row[i] = cs_row(L.pos, pw, rbase, pos, r_hi, s + i);
  wave_sync();
  const int32_t r_first = L.row[0], r_last = L.row[cn - 1];
  const bool head = cs_P(L.pos, pw, rbase, pos, r_first) < s;
  const bool tail = !(head && r_last == r_first) &&
                    cs_P(L.pos, pw, rbase, pos, r_last + 1) > e;
  // interior rank: starts in [s, p] minus one, entries processed in order
  // in chunks of 64 (one per lane) with a running total
  int32_t total = 0;
  for (int32_t c0 = 0; c0 < cn; c0 += 64) {
    const int32_t i = c0 + lane;
    const bool ok = i < cn;
    const int32_t r = ok ? L.row[i] : -1;
    const int32_t st = ok && cs_P(L.pos, pw, rbase, pos, r) == s + i ? 1 : 0;
    const int32_t inc = wave_incl_scan(st, lane) + total;
    if (kWrite && ok) {
      int32_t code;
      if (head && r == r_first)
        code = kCsHead;
      else if (tail && r == r_last)
        code = kCsTail;
      else
        code = base + inc - 1;
      L.slot[i] = code;
      if (st && code < kCsTail && code < a.cap)
        a.slot_rows[(int64_t)unit * a.cap + code] = r;
    }
    total = __shfl(inc, 63, 64);
  }
  *flags = (head ? 1 : 0) | (tail ? 2 : 0);
  if (!kWrite) {
    wave_sync();
    return total - (tail ? 1 : 0);
  }
  // sort by (column, position): bitonic over the next power of two
  int32_t n2 = 1;
  while (n2 < cn) n2 <<= 1;
  const int32_t *__restrict__ crd = a.crd;
  for (int32_t i = lane; i < n2; i += 64)
    L.key[i] = i < cn ? ((uint64_t)(uint32_t)crd[s + i] << 32) | (uint32_t)i
                      : ~(uint64_t)0;
  wave_sync();
  for (int32_t k = 2; k <= n2; k <<= 1) {
    for (int32_t j = k >> 1; j > 0; j >>= 1) {
      for (int32_t i = lane; i < n2; i += 64) {
        const int32_t l = i ^ j;
        if (l > i) {
          const uint64_t x = L.key[i], y = L.key[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            L.key[i] = y;
            L.key[l] = x;
          }
        }
      }
      wave_sync();
    }
  }
  for (int32_t j = lane; j < cn; j += 64) {
    const uint64_t key = L.key[j];
    const int32_t i = (int32_t)(key & 0xffffffffu);
    a.s_col[s + j] = (int32_t)(key >> 32);
    a.s_slot[s + j] = (uint16_t)L.slot[i];
    if (a.vals) {
      if (a.elem == 4)
        static_cast<uint32_t *>(a.s_val)[s + j] =
            static_cast<const uint32_t *>(a.vals)[s + i];
      else
        static_cast<uint16_t *>(a.s_val)[s + j] =
            static_cast<const uint16_t *>(a.vals)[s + i];
    }
  }
  wave_sync();
  return total - (tail ? 1 : 0);
}

// Phase A, one wave per tile: interior row count and head/tail flags.
__global__ __launch_bounds__(kCsPrepThreads) void k_cs_count(const CsArgs a) {
  __shared__ PrepLds lds[kCsPrepThreads / 64];
  const int wave = threadIdx.x >> 6;
  const int32_t t = blockIdx.x * (kCsPrepThreads / 64) + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  int32_t fl;
  const int32_t n = cs_prep_tile<false>(a, t, lds[wave], 0, 0, &fl);
  if ((threadIdx.x & 63) == 0) {
    a.tile_count[t] = n;
    a.tile_flags[t] = (uint8_t)fl;
  }
}

// Phase B, one wave per tile: the tile's first slot is the sum of the counts
// of the unit's earlier tiles; slot codes, slot rows and the sorted copy.
__global__ __launch_bounds__(kCsPrepThreads) void k_cs_sort(const CsArgs a) {
  __shared__ PrepLds lds[kCsPrepThreads / 64];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int32_t t = blockIdx.x * (kCsPrepThreads / 64) + wave;
  if (t >= a.ntiles) return;  // uniform over the wave
  const int32_t unit = t / a.group;
  const int32_t t0 = unit * a.group;
  int32_t part = 0;
  for (int32_t i = t0 + lane; i < t; i += 64) part += a.tile_count[i];
  for (int d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
  int32_t fl;
  cs_prep_tile<true>(a, t, lds[wave], part, unit, &fl);
  const bool last = t == a.ntiles - 1 || t == t0 + a.group - 1;
  if (last && lane == 0) a.n_slots[unit] = part + a.tile_count[t];
}

// CSR-order walk of tile t, panel columns [col, col + V): the fallback for a
// unit whose interior rows do not fit the LDS accumulators.
template <int DT, int V>
__device__ void cs_fallback_tile(const CsArgs &a, int32_t t, int32_t col) {
  const int32_t *__restrict__ pos = a.pos;
  const int32_t s = t * a.tile;
  const int32_t e = (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile);
  int32_t r = a.tile_row[t];
  int32_t rs = pos[r], re = pos[r + 1];
  float acc[V];
  zero_vec(acc);
  const bool colok = col < a.N;
  auto flush = [&]() {
    if (!colok) return;
    float *dst = rs < s   ? a.carry_head + (int64_t)t * a.N + col
                 : re > e ? a.carry_tail + (int64_t)t * a.N + col
                          : a.C + (int64_t)r * a.ldc + col;
    store_vec<V>(dst, acc);
  };
  for (int32_t p = s; p < e; ++p) {
    if (p == re) {
      flush();
      ++r;
      while (pos[r + 1] <= p) ++r;
      rs = pos[r];
      re = pos[r + 1];
      zero_vec(acc);
    }
    if (colok) {
      const float v = load_val<DT>(a.vals, p);
      float b[V];
      unpack<DT, V>(load_raw<DT, V>(a.B, (int64_t)a.crd[p] * a.ldb + col), b);
#pragma unroll
      for (int j = 0; j < V; ++j) acc[j] = fmaf(v, b[j], acc[j]);
    }
  }
  flush();
}

template <int DT, int V>
__global__ __launch_bounds__(kCsThreads) void k_eb_cs(const CsArgs a) {
  extern __shared__ float acc_lds[];  // [cap][W]
  constexpr int32_t Q = kCsLanes;
  constexpr int32_t W = Q * V;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int32_t q = lane & (Q - 1);
  const int32_t gbase = lane & ~(Q - 1);        // first lane of my group
  const int32_t g = wave * (64 / Q) + (lane / Q);  // group in workgroup
  using R = typename RawT<DT, V>::type;
  const int32_t panels = (a.N + W - 1) / W;

  for (int32_t round = 0; round < a.rounds; ++round) {
    const int32_t unit = blockIdx.x * a.rounds + round;
    if (unit >= a.units) break;  // uniform over the workgroup
    const int32_t t = unit * a.group + g;
    const bool tok = g < a.group && t < a.ntiles;
    const int32_t n_slots = a.n_slots[unit];
    const bool fits = n_slots <= a.cap;
    const int32_t s = tok ? t * a.tile : 0;
    const int32_t e = tok ? (int32_t)min((int64_t)a.nnz, (int64_t)s + a.tile) : 0;
    const int32_t flags = tok ? a.tile_flags[t] : 0;
    for (int32_t panel = 0; panel < panels; ++panel) {
      const int32_t col = panel * W + q * V;
      const bool colok = col < a.N;
      if (!fits) {
        if (tok) cs_fallback_tile<DT, V>(a, t, col);
        continue;
      }
      for (int32_t i = threadIdx.x; i < n_slots * W; i += kCsThreads) acc_lds[i] = 0.0f;
      __syncthreads();
      float head[V], tail[V];
      zero_vec(head);
      zero_vec(tail);
      // Software pipeline over chunks of Q*Q entries: lane q holds the
      // metadata of entries c0 + Q*k + q (k < Q); batch k of a chunk is
      // entries c0 + Q*k .. c0 + Q*k + Q - 1, broadcast from lane u of the
      // group. The next chunk's metadata loads when a chunk starts, and
      // batch k+1's gathers are in flight while batch k is consumed.
      constexpr int32_t CK = Q * Q;
      int32_t mc[Q], ms[Q], nc[Q], ns[Q];
      float mv[Q], nv[Q];
      // Every load below is unconditional (entries past the tile re-read its
      // last entry, columns past N column 0, neither used), so the number of
      // loads in flight is the same on every path and each wait covers only
      // the batch it needs. Called only while s < e.
      const int32_t col_ld = colok ? col : 0;
      auto load_meta = [&](int32_t c0, int32_t (&c)[Q], float (&v)[Q], int32_t (&sl)[Q]) {
#pragma unroll
        for (int32_t k = 0; k < Q; ++k) {
          const int32_t jj = c0 + Q * k + q;
          const int32_t jc = min(jj, e - 1);
          const bool ok = jj < e;
          c[k] = a.s_col[jc];
          const float vk = load_val<DT>(a.s_val, jc);
          const int32_t sk = (int32_t)a.s_slot[jc];
          v[k] = ok ? vk : 0.0f;
          sl[k] = ok ? sk : -1;
        }
      };
      auto issue = [&](R (&buf)[Q], int32_t cq, int32_t) {
#pragma unroll
        for (int32_t u = 0; u < Q; ++u) {
          const int32_t cu = __shfl(cq, gbase + u, 64);
          buf[u] = load_raw<DT, V>(a.B, (int64_t)cu * a.ldb + col_ld);
        }
      };
      auto consume = [&](const R (&buf)[Q], float vq, int32_t sq, int32_t j0) {
#pragma unroll
        for (int32_t u = 0; u < Q; ++u) {
          const float v = __shfl(vq, gbase + u, 64);
          const int32_t sl = __shfl(sq, gbase + u, 64);
          if (j0 + u < e && colok) {
            float b[V];
            unpack<DT, V>(buf[u], b);
            if (sl < kCsTail) {
              float *p = acc_lds + sl * W + q * V;
              float x[V];
#pragma unroll
              for (int k = 0; k < V; ++k) x[k] = p[k];
#pragma unroll
              for (int k = 0; k < V; ++k) x[k] = fmaf(v, b[k], x[k]);
#pragma unroll
              for (int k = 0; k < V; ++k) p[k] = x[k];
            } else if (sl == kCsHead) {
#pragma unroll
              for (int k = 0; k < V; ++k) head[k] = fmaf(v, b[k], head[k]);
            } else {
#pragma unroll
              for (int k = 0; k < V; ++k) tail[k] = fmaf(v, b[k], tail[k]);
            }
          }
        }
      };
      R buf0[Q], buf1[Q];
      if (s < e) {
        load_meta(s, mc, mv, ms);
        issue(buf0, mc[0], s);
      }
      for (int32_t c0 = s; c0 < e; c0 += CK) {
        const bool more = c0 + CK < e;
        load_meta(c0 + CK, nc, nv, ns);  // past the tile: clamped, unused
#pragma unroll
        for (int32_t k = 0; k < Q; ++k) {
          const int32_t j0 = c0 + Q * k;
          // batch k+1 (the next chunk's batch 0 after the last) goes out first
          if (k + 1 < Q) {
            if (k & 1)
              issue(buf0, mc[k + 1], j0 + Q);
            else
              issue(buf1, mc[k + 1], j0 + Q);
          } else {
            issue(buf0, nc[0], c0 + CK);  // Q is even: batch 0 uses buf0
          }
          if (k & 1)
            consume(buf1, mv[k], ms[k], j0);
          else
            consume(buf0, mv[k], ms[k], j0);
        }
        if (more) {
#pragma unroll
          for (int32_t k = 0; k < Q; ++k) {
            mc[k] = nc[k];
            mv[k] = nv[k];
            ms[k] = ns[k];
          }
        }
      }
      if (tok && colok) {
        if (flags & 1) store_vec<V>(a.carry_head + (int64_t)t * a.N + col, head);
        if (flags & 2) store_vec<V>(a.carry_tail + (int64_t)t * a.N + col, tail);
      }
      __syncthreads();
      constexpr int32_t PER = W / V;  // vectors per slot
      for (int32_t i = threadIdx.x; i < n_slots * PER; i += kCsThreads) {
        const int32_t sl = i / PER;
        const int32_t c = panel * W + (i - sl * PER) * V;
        if (c < a.N) {
          const int32_t row = a.slot_rows[(int64_t)unit * a.cap + sl];
          float x[V];
#pragma unroll
          for (int k = 0; k < V; ++k) x[k] = acc_lds[sl * W + (i - sl * PER) * V + k];
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
  constexpr int kWaves = kCsPrepThreads / 64;
  const dim3 g((unsigned)((a.ntiles + kWaves - 1) / kWaves)), b(kCsPrepThreads);
  hipLaunchKernelGGL(k_cs_count, g, b, 0, s, a);
  hipLaunchKernelGGL(k_cs_sort, g, b, 0, s, a);
  return hipGetLastError();
}

size_t cs_lds_bytes(int32_t vec, int32_t cap) {
  return (size_t)cap * kCsLanes * vec * sizeof(float);
}

template <int DT, int V>
hipError_t launch_eb_cs_t(const CsArgs &a, size_t lds, hipStream_t s) {
  // dynamic LDS above the default per-kernel limit has to be declared; the
  // capacity is fixed per device, so once per instantiation is enough
  static const hipError_t declared = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&k_eb_cs<DT, V>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (declared != hipSuccess) return declared;
  const unsigned grid = (unsigned)((a.units + a.rounds - 1) / a.rounds);
  hipLaunchKernelGGL((k_eb_cs<DT, V>), dim3(grid), dim3(kCsThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_eb_cs(int dt, int vec, const CsArgs &a, hipStream_t s) {
  if (a.units <= 0) return hipSuccess;
  const size_t lds = cs_lds_bytes(vec, a.cap);
  switch (dt * 16 + vec) {
    case 0 * 16 + 1: return launch_eb_cs_t<0, 1>(a, lds, s);
    case 0 * 16 + 2: return launch_eb_cs_t<0, 2>(a, lds, s);
    case 0 * 16 + 4: return launch_eb_cs_t<0, 4>(a, lds, s);
    case 1 * 16 + 1: return launch_eb_cs_t<1, 1>(a, lds, s);
    case 1 * 16 + 2: return launch_eb_cs_t<1, 2>(a, lds, s);
    case 1 * 16 + 4: return launch_eb_cs_t<1, 4>(a, lds, s);
    case 2 * 16 + 1: return launch_eb_cs_t<2, 1>(a, lds, s);
    case 2 * 16 + 2: return launch_eb_cs_t<2, 2>(a, lds, s);
    case 2 * 16 + 4: return launch_eb_cs_t<2, 4>(a, lds, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace segtaco

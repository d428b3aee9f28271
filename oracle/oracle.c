/*
 * oracle.c — CPU restatement of the reference's CSR SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg. The product never links or calls it.
 *
 * Build with -ffp-contract=off: the Eigen-order product must round the
 * multiply and the add separately, as the reference's g++ build of Eigen does
 * on x86-64 without -mfma.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* Restates taco_binarySearchBefore (gpu_library.h:185-205): bisection over
 * array[start..end] for the last index whose value is <= target, except that
 * an exact hit returns the probed index at once. Inside a run of equal values
 * (empty CSR rows) that is whichever member of the run the bisection probes
 * first, not necessarily the last one. */
int32_t or_search_before(const int32_t *array, int32_t start, int32_t end,
                         int32_t target) {
  if (array[end] <= target) return end;
  int32_t lo = start; /* array[lo] <= target */
  int32_t hi = end;   /* array[hi] > target */
  while (hi - lo > 1) {
    int32_t mid = (hi + lo) / 2;
    int32_t v = array[mid];
    if (v < target) {
      lo = mid;
    } else if (v > target) {
      hi = mid;
    } else {
      return mid;
    }
  }
  return lo;
}

/* Restates taco_binarySearchBeforeBlock and its launcher
 * (gpu_library.h:206-215, 236-240): out[b] = search_before(pos, 0, M, b*T)
 * for b = 0 .. num_blocks inclusive (num_blocks + 1 entries). */
void or_block_starts(const int32_t *pos, int32_t M, int32_t values_per_block,
                     int32_t num_blocks, int32_t *out) {
  for (int32_t b = 0; b <= num_blocks; ++b)
    out[b] = or_search_before(pos, 0, M,
                              (int32_t)((int64_t)b * values_per_block));
}

/* Restates binary_search_segment_number (gpu_library.h:138-152): returns
 * hi - 1 where hi is the first index in [1, n_seg] with seg_offsets[hi] >
 * elem_id (or n_seg), i.e. the last row r with pos[r] <= elem_id. */
int32_t or_segment_number(const int32_t *seg_offsets, int32_t n_seg,
                          int32_t elem_id) {
  int32_t lo = 1, hi = n_seg;
  while (lo < hi) {
    int32_t mid = (lo + hi) >> 1;
    if (seg_offsets[mid] <= elem_id)
      lo = mid + 1;
    else
      hi = mid;
  }
  return hi - 1;
}

/* Row id of every stored entry (the COO row array), by a linear walk. */
void or_row_of_nnz(const int32_t *pos, int32_t M, int32_t *rows) {
  for (int32_t r = 0; r < M; ++r)
    for (int32_t p = pos[r]; p < pos[r + 1]; ++p) rows[p] = r;
}

/* First row of every element-balanced tile, as the reference's tiled
 * kernels reach it: the K21 search (taco_binarySearchBefore for b*T,
 * gpu_library.h:185-205) followed by the repair loop that steps over the
 * empty rows sharing that row pointer (spmm_eb_sr.h:39-42:
 * while (fposA == A2_pos[i_pos + 1]) i_pos++). rows[b] for b < ntiles is the
 * row holding entry b*T; rows[ntiles] = M (the search's early return for a
 * target at or past nnz). ntiles = ceil(nnz / T). */
void or_tile_rows(const int32_t *pos, int32_t M, int32_t T, int32_t ntiles,
                  int32_t *rows) {
  for (int32_t b = 0; b <= ntiles; ++b) {
    const int64_t target64 = (int64_t)b * T;
    const int32_t target = target64 > pos[M] ? pos[M] : (int32_t)target64;
    int32_t r = or_search_before(pos, 0, M, target);
    if (b < ntiles)
      while (pos[r + 1] == target) ++r;
    rows[b] = r;
  }
}

/* Carry bookkeeping of a standalone element-balanced call: fix[b] = the tile
 * in which the row holding tile b's first entry started, when that row began
 * before tile b and ends inside tile b (its carries are then combined by the
 * fix-up of tile b); -1 otherwise. Walks the rows: a non-empty row with
 * entries [p0, p1) spans tiles p0/T .. (p1-1)/T. */
void or_fix_t0(const int32_t *pos, int32_t M, int32_t T, int32_t ntiles,
               int32_t *fix) {
  for (int32_t b = 0; b < ntiles; ++b) fix[b] = -1;
  for (int32_t r = 0; r < M; ++r) {
    const int32_t p0 = pos[r], p1 = pos[r + 1];
    if (p0 == p1) continue;
    const int32_t t0 = p0 / T, t1 = (p1 - 1) / T;
    if (t0 < t1) fix[t1] = t0;
  }
}

/* Restates Eigen 3.4.0 sparse_time_dense_product_impl<..., RowMajor>::
 * processRow (SparseDenseProduct.h:141-146) under C.noalias() = A * B:
 * C is zeroed, then for each row, for each stored entry in CSR order,
 * C[i,:] += (1 * v) * B[crd,:], each multiply and add rounded to float.
 * vals == NULL reads every value as 1.0 (gpu_library.h:154-161). */
void or_spmm_f32_seq(int32_t M, int32_t N, const int32_t *pos,
                     const int32_t *crd, const float *vals, const float *B,
                     int64_t ldb, float *C, int64_t ldc) {
  for (int32_t i = 0; i < M; ++i) {
    float *c = C + (int64_t)i * ldc;
    for (int32_t k = 0; k < N; ++k) c[k] = 0.0f;
    for (int32_t p = pos[i]; p < pos[i + 1]; ++p) {
      const float v = vals ? vals[p] : 1.0f;
      const float *b = B + (int64_t)crd[p] * ldb;
      for (int32_t k = 0; k < N; ++k) {
        float prod = v * b[k];
        c[k] = c[k] + prod;
      }
    }
  }
}

/* Closed form C[i,k] = sum_p v_p * B[crd_p, k], accumulated in double from
 * the float inputs (every float product is exact in double). */
void or_spmm_f64(int32_t M, int32_t N, const int32_t *pos, const int32_t *crd,
                 const float *vals, const float *B, int64_t ldb, double *C,
                 int64_t ldc) {
  for (int32_t i = 0; i < M; ++i) {
    double *c = C + (int64_t)i * ldc;
    for (int32_t k = 0; k < N; ++k) c[k] = 0.0;
    for (int32_t p = pos[i]; p < pos[i + 1]; ++p) {
      const double v = vals ? (double)vals[p] : 1.0;
      const float *b = B + (int64_t)crd[p] * ldb;
      for (int32_t k = 0; k < N; ++k) c[k] += v * (double)b[k];
    }
  }
}

/* S[i,k] = sum_p |v_p| * |B[crd_p, k]|: the scale that rounding error in any
 * summation order is proportional to. */
void or_spmm_absbound(int32_t M, int32_t N, const int32_t *pos,
                      const int32_t *crd, const float *vals, const float *B,
                      int64_t ldb, double *S, int64_t lds) {
  for (int32_t i = 0; i < M; ++i) {
    double *s = S + (int64_t)i * lds;
    for (int32_t k = 0; k < N; ++k) s[k] = 0.0;
    for (int32_t p = pos[i]; p < pos[i + 1]; ++p) {
      const double v = vals ? fabs((double)vals[p]) : 1.0;
      const float *b = B + (int64_t)crd[p] * ldb;
      for (int32_t k = 0; k < N; ++k) s[k] += v * fabs((double)b[k]);
    }
  }
}

/* SDDMM in double (the reference's sddmm_csr_gpu_taco.h:70-148 computes
 * A_vals[p] = sum_j B_vals[p] * C[i, j] * D[k, j] in FP64): out[p] =
 * s_p * sum_j X[i, j] * Y[k, j] for stored entry p = (i, k); with absval set,
 * |s_p| * sum_j |X[i, j]| * |Y[k, j]| (the rounding-error scale) instead. */
void or_sddmm_f64(int32_t M, int32_t J, const int32_t *pos, const int32_t *crd,
                  const float *vals, const float *X, int64_t ldx,
                  const float *Y, int64_t ldy, int32_t absval, double *out) {
  for (int32_t i = 0; i < M; ++i) {
    const float *x = X + (int64_t)i * ldx;
    for (int32_t p = pos[i]; p < pos[i + 1]; ++p) {
      const float *y = Y + (int64_t)crd[p] * ldy;
      double d = 0.0;
      for (int32_t j = 0; j < J; ++j)
        d += absval ? fabs((double)x[j]) * fabs((double)y[j])
                    : (double)x[j] * (double)y[j];
      const double v = vals ? (double)vals[p] : 1.0;
      out[p] = (absval ? fabs(v) : v) * d;
    }
  }
}

/* max_i |C_i - ref_i| / (scale_i + 1e-30); NaN anywhere returns +inf. */
double or_max_scaled_err(int64_t n, const float *C, const double *ref,
                         const double *scale) {
  double worst = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double d = fabs((double)C[i] - ref[i]);
    if (d != d) return INFINITY;
    double e = d / (scale[i] + 1e-30);
    if (e > worst) worst = e;
  }
  return worst;
}

static uint32_t f32_bits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}
static float bits_f32(uint32_t u) {
  float x;
  memcpy(&x, &u, 4);
  return x;
}

/* float -> bfloat16, round to nearest, ties to even; NaN stays NaN. */
uint16_t or_bf16_round(float x) {
  uint32_t u = f32_bits(x);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
    return (uint16_t)((u >> 16) | 0x40u);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

float or_bf16_to_f32(uint16_t h) { return bits_f32((uint32_t)h << 16); }

/* float -> IEEE binary16, round to nearest, ties to even, with subnormals
 * and overflow to infinity. */
uint16_t or_f16_round(float x) {
  uint32_t u = f32_bits(x);
  uint32_t sign = (u >> 16) & 0x8000u;
  uint32_t ex = (u >> 23) & 0xffu;
  uint32_t man = u & 0x007fffffu;
  if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
  int32_t e = (int32_t)ex - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    /* subnormal half (or zero): value = man24 * 2^(ex-150), unit 2^-24 */
    if (e < -10) return (uint16_t)sign;
    uint32_t m = man | 0x00800000u;
    uint32_t shift = (uint32_t)(14 - e); /* 24-bit mantissa -> units of 2^-24 */
    uint32_t q = m >> shift;
    uint32_t rem = m & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t q = ((uint32_t)e << 10) | (man >> 13);
  uint32_t rem = man & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q); /* carry into exponent is correct rounding */
}

float or_f16_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t ex = ((uint32_t)h >> 10) & 0x1fu;
  uint32_t man = (uint32_t)h & 0x3ffu;
  if (ex == 0x1fu) return bits_f32(sign | 0x7f800000u | (man << 13));
  if (ex == 0) {
    float v = (float)man * (1.0f / 16777216.0f); /* man * 2^-24, exact */
    return sign ? -v : v;
  }
  return bits_f32(sign | ((ex - 15 + 127) << 23) | (man << 13));
}

/* dtype 1 = bf16, 2 = fp16: out[i] = round(in[i]); back[i] = widen(out[i]). */
void or_round_array(int32_t dtype, int64_t n, const float *in, uint16_t *out,
                    float *back) {
  for (int64_t i = 0; i < n; ++i) {
    uint16_t h = dtype == 1 ? or_bf16_round(in[i]) : or_f16_round(in[i]);
    if (out) out[i] = h;
    if (back) back[i] = dtype == 1 ? or_bf16_to_f32(h) : or_f16_to_f32(h);
  }
}

/*
 * oracle.h — CPU restatement of the reference's CSR SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (segtaco_amd/) links,
 * loads or calls this code; it is used by tests/, by __graft_entry__.smoke()
 * as the checker, and by bench.py's cpu_baseline leg.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the segTACO tree). How each function is pinned:
 *   or_search_before      restates gpu_library.h:185-205; pinned by brute-force
 *                         invariants + exhaustive tie-break enumeration in tests.
 *   or_block_starts       restates gpu_library.h:206-215,236-240; pinned as above.
 *   or_segment_number     restates gpu_library.h:138-152; pinned by brute force.
 *   or_row_of_nnz         per-entry row ids (the COO row array); pinned against
 *                         numpy's repeat(arange(M), diff(pos)) and against
 *                         or_segment_number at every entry.
 *   or_tile_rows          K21 search + the repair loop of spmm_eb_sr.h:39-42
 *                         (the row each element-balanced tile starts at);
 *                         pinned against the row holding entry b*T (brute
 *                         force) and by a planted repair-less variant.
 *   or_fix_t0             the fix-up's per-tile record (this build's carry
 *                         bookkeeping); pinned by a tile-wise brute force.
 *   or_spmm_f32_seq       restates Eigen 3.4.0 SparseDenseProduct.h:109-147
 *                         (row-major CSR x row-major dense); pinned bit-exactly
 *                         against the real Eigen product (oracle/_ref) and the
 *                         committed golden vectors made from it.
 *   or_spmm_f64           closed form sum_p v_p * B[crd_p, k] in double; pinned
 *                         by special cases (A = I, B = I, B = ones, N = 1) and a
 *                         dense brute-force product.
 *   or_spmm_absbound      sum_p |v_p| * |B[crd_p, k]| in double (error scale);
 *                         pinned against the dense |A| |B| product in float64
 *                         (empty rows, no values), planted x2 / signed variants
 *                         rejected.
 *   or_sddmm_f64          SDDMM closed form in double (sddmm_csr_gpu_taco.h:
 *                         70-148 computes it in FP64); pinned by a dense
 *                         brute-force (X Y^T masked by the pattern) and special
 *                         cases (J = 1, X = Y rows -> squared norms).
 *   or_bf16_round / or_f16_round  IEEE round-to-nearest-even narrowing; pinned
 *                         against numpy / torch conversions in tests.
 */
#ifndef SEGTACO_ORACLE_H
#define SEGTACO_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int32_t or_search_before(const int32_t *array, int32_t start, int32_t end,
                         int32_t target);
void or_block_starts(const int32_t *pos, int32_t M, int32_t values_per_block,
                     int32_t num_blocks, int32_t *out);
int32_t or_segment_number(const int32_t *seg_offsets, int32_t n_seg,
                          int32_t elem_id);
void or_row_of_nnz(const int32_t *pos, int32_t M, int32_t *rows);
void or_tile_rows(const int32_t *pos, int32_t M, int32_t T, int32_t ntiles,
                  int32_t *rows);
void or_fix_t0(const int32_t *pos, int32_t M, int32_t T, int32_t ntiles,
               int32_t *fix);

void or_spmm_f32_seq(int32_t M, int32_t N, const int32_t *pos,
                     const int32_t *crd, const float *vals, const float *B,
                     int64_t ldb, float *C, int64_t ldc);
void or_spmm_f64(int32_t M, int32_t N, const int32_t *pos, const int32_t *crd,
                 const float *vals, const float *B, int64_t ldb, double *C,
                 int64_t ldc);
void or_spmm_absbound(int32_t M, int32_t N, const int32_t *pos,
                      const int32_t *crd, const float *vals, const float *B,
                      int64_t ldb, double *S, int64_t lds);
void or_sddmm_f64(int32_t M, int32_t J, const int32_t *pos, const int32_t *crd,
                  const float *vals, const float *X, int64_t ldx,
                  const float *Y, int64_t ldy, int32_t absval, double *out);
double or_max_scaled_err(int64_t n, const float *C, const double *ref,
                         const double *scale);

uint16_t or_bf16_round(float x);
uint16_t or_f16_round(float x);
float or_bf16_to_f32(uint16_t h);
float or_f16_to_f32(uint16_t h);
void or_round_array(int32_t dtype, int64_t n, const float *in, uint16_t *out,
                    float *back);

#ifdef __cplusplus
}
#endif
#endif

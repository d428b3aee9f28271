/*
 * io.h — host-side operand construction for drivers: MatrixMarket reading,
 * COO -> CSR canonicalisation, the reference's value fill, and the
 * reference's result comparators. Host code only (no device memory).
 *
 * Arrays returned through pointer-to-pointer arguments are allocated by the
 * library and released with segtaco_host_free.
 */
#ifndef SEGTACO_IO_H
#define SEGTACO_IO_H
#include <stddef.h>
#include <stdint.h>

#include "segtaco/spmm.h"
#include "segtaco/taco_tensor_t.h"

#ifdef __cplusplus
extern "C" {
#endif

/* segtaco_mtx_read flags */
/* Keep only the first `entries` (header count) coordinates after symmetric
 * expansion, as the reference's read_mtx_file does (ds.h:254-265 loops
 * `curr_pos < nnz` with nnz still the header count), instead of the whole
 * expanded matrix. Only meaningful for symmetric files. */
#define SEGTACO_MTX_REFERENCE_TRUNCATE 1

/* Banner of the last file read (segtaco_mtx_read's *info). */
typedef struct segtaco_mtx_info {
  int32_t coordinate; /* 1: coordinate, 0: array (rejected) */
  int32_t field;      /* 0 real, 1 complex, 2 pattern, 3 integer */
  int32_t symmetry;   /* 0 general, 1 symmetric, 2 hermitian, 3 skew */
  int32_t entries;    /* entry count stated in the size line */
} segtaco_mtx_info_t;

/* Reads the sparsity pattern of a MatrixMarket coordinate file
 * (replaces read_mtx_file, ds.h:183-268, over mm_read_banner /
 * mm_read_mtx_crd_size, mmio.h:203-310): 1-based coordinates -> 0-based,
 * values ignored, symmetric files expanded (i,j) -> (i,j),(j,i) and
 * de-duplicated, sorted by (row, column). General files keep duplicate
 * coordinates (summed later by segtaco_coo_to_csr, as the reference's
 * to_eigen_csr does). Hermitian and skew files are read as stored, like the
 * reference. Errors (missing file, bad banner, short file, a coordinate
 * outside the stated shape, array format) return SEGTACO_ERR_INVALID with
 * segtaco_last_error_string() describing it; the reference exits instead.
 * *rows / *cols (length *nnz) are sorted coordinates. info may be NULL. */
int32_t segtaco_mtx_read(const char *path, int32_t flags, int32_t *nrow,
                         int32_t *ncol, int32_t *nnz, int32_t **rows,
                         int32_t **cols, segtaco_mtx_info_t *info);

/* Value fill of the reference's new_get_matrix (fill_random, ds.h:270-274:
 * (rand() % 3) / 10, i.e. values in {0, 0.1f, 0.2f}) drawn from a seeded
 * splitmix64 stream instead of the time-seeded rand(), so inputs are
 * reproducible. */
void segtaco_fill_reference_vals(uint64_t seed, int64_t n, float *vals);

/* COO triplets -> compressed CSR (to_eigen_csr, ds.h:380-397, i.e. Eigen's
 * setFromTriplets + makeCompressed): entries sorted by (row, column),
 * duplicate coordinates summed in input order (((v1 + v2) + v3) ...),
 * explicit zeros kept. Input order is arbitrary. pos has m + 1 entries; crd
 * and out_vals need room for nnz entries (the result has *out_nnz <= nnz).
 * vals / out_vals may be NULL (pattern only). */
int32_t segtaco_coo_to_csr(int32_t m, int32_t n, int32_t nnz,
                           const int32_t *rows, const int32_t *cols,
                           const float *vals, int32_t *pos, int32_t *crd,
                           float *out_vals, int32_t *out_nnz);

/* new_get_matrix + to_eigen_csr in one call (ds.h:288-313, 380-397): reads
 * the pattern, fills values with segtaco_fill_reference_vals(seed) (or all
 * ones when seed == UINT64_MAX, ds.h:276-280 fill_one), and canonicalises.
 * *pos (M + 1), *crd, *vals (nnz) are library-allocated. */
int32_t segtaco_mtx_read_csr(const char *path, int32_t flags, uint64_t seed,
                             int32_t *M, int32_t *K, int32_t *nnz,
                             int32_t **pos, int32_t **crd, float **vals);

void segtaco_host_free(void *p);

/* compare_array (ds.h:463-477): max |y[i] / x[i] - 1| over x[i] != 0;
 * +infinity as soon as some x[i] == 0 has y[i] != 0. */
float segtaco_compare_array(const float *x, const float *y, size_t n);
/* compare_array_my (ds.h:478-492): number of i with |x[i] - y[i]| > 1e-3,
 * counting stops once it exceeds 10000 (so at most 10001). The reference
 * prints each mismatch; this does not. */
float segtaco_compare_array_my(const float *x, const float *y, size_t n);
/* compare_matrices (ds.h:493-499): compare_array_my over the dense
 * a.dimensions[0] x a.dimensions[1] values. Mismatched shapes return
 * +infinity (the reference asserts). */
float segtaco_compare_matrices(const taco_tensor_t *a, const taco_tensor_t *b);

#ifdef __cplusplus
}
#endif
#endif

// ref_shim.cpp — C entry points over the reference's own CPU code.
//
// TEST INFRASTRUCTURE ONLY. Compiled (by oracle/Makefile) against the
// reference tree in place: its ds.h (operand construction, MatrixMarket
// reading, COO -> Eigen CSR) and its vendored Eigen 3.4.0, whose
// sparse x dense product is the reference's CPU SpMM path. Output goes to
// oracle/_ref/ only. Nothing here is re-implemented: each function forwards
// to the reference code named in its comment.
#include "ds.h"  // reference: ds.h (includes mmio.h and <Eigen/Sparse>)

#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// C = A * B through the reference's Eigen types (ds.h:32-33:
// EigenCSR = SparseMatrix<float, RowMajor, int>, EigenRowMajor), i.e.
// SparseDenseProduct.h:109-147. threads > 1 enables Eigen's OpenMP row loop.
void ref_eigen_spmm(int32_t M, int32_t K, int32_t N, int32_t nnz,
                    const int32_t *pos, const int32_t *crd, const float *vals,
                    const float *B, float *C, int32_t threads) {
#ifdef _OPENMP
  Eigen::setNbThreads(threads > 0 ? threads : 1);
#endif
  Eigen::Map<const EigenCSR> A(M, K, nnz, pos, crd, vals);
  Eigen::Map<const EigenRowMajor> Bm(B, K, N);
  Eigen::Map<EigenRowMajor> Cm(C, M, N);
  Cm.noalias() = A * Bm;
}

// Times `trials` runs of ref_eigen_spmm (after one warm-up) with
// steady_clock; writes each trial's seconds to times[].
void ref_eigen_spmm_timed(int32_t M, int32_t K, int32_t N, int32_t nnz,
                          const int32_t *pos, const int32_t *crd,
                          const float *vals, const float *B, float *C,
                          int32_t threads, int32_t trials, double *times) {
  ref_eigen_spmm(M, K, N, nnz, pos, crd, vals, B, C, threads);
  for (int32_t t = 0; t < trials; ++t) {
    auto t0 = std::chrono::steady_clock::now();
    ref_eigen_spmm(M, K, N, nnz, pos, crd, vals, B, C, threads);
    auto t1 = std::chrono::steady_clock::now();
    times[t] = std::chrono::duration<double>(t1 - t0).count();
  }
}

// COO triplets -> compressed CSR through ds.h:380-397 (to_eigen_csr:
// setFromTriplets sorts by row/column and sums duplicates; makeCompressed).
// Returns nnz after compression; *pos_out / *crd_out / *vals_out are
// malloc'd and released with ref_free.
int32_t ref_coo_to_csr(int32_t m, int32_t n, int32_t nnz, const int32_t *rows,
                       const int32_t *cols, const float *vals,
                       int32_t **pos_out, int32_t **crd_out,
                       float **vals_out) {
  COO coo;
  coo.m = m;
  coo.n = n;
  coo.nnz = nnz;
  coo.rows = const_cast<int32_t *>(rows);
  coo.cols = const_cast<int32_t *>(cols);
  coo.vals = const_cast<float *>(vals);
  EigenCSR csr = to_eigen_csr(coo);
  const int32_t out_nnz = (int32_t)csr.nonZeros();
  *pos_out = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(m + 1));
  *crd_out = (int32_t *)std::malloc(sizeof(int32_t) * (size_t)(out_nnz + 1));
  *vals_out = (float *)std::malloc(sizeof(float) * (size_t)(out_nnz + 1));
  for (int32_t i = 0; i <= m; ++i) (*pos_out)[i] = csr.outerIndexPtr()[i];
  for (int32_t p = 0; p < out_nnz; ++p) {
    (*crd_out)[p] = csr.innerIndexPtr()[p];
    (*vals_out)[p] = csr.valuePtr()[p];
  }
  return out_nnz;
}

// MatrixMarket pattern -> CSR through ds.h:183-268 (read_mtx_file: 1-based
// coordinates, symmetric expansion with de-duplication, values dropped).
// Returns nnz, or -1 if the file cannot be opened (the reference exits).
int32_t ref_read_mtx(const char *path, int32_t *nrow, int32_t *ncol,
                     int32_t **pos_out, int32_t **crd_out) {
  FILE *f = std::fopen(path, "r");
  if (!f) return -1;
  std::fclose(f);
  int nr = 0, nc = 0, nz = 0;
  std::vector<int> indptr, indices, rowind;
  read_mtx_file(std::string(path), nr, nc, nz, indptr, indices, rowind);
  *nrow = nr;
  *ncol = nc;
  *pos_out = (int32_t *)std::malloc(sizeof(int32_t) * indptr.size());
  *crd_out = (int32_t *)std::malloc(sizeof(int32_t) * (indices.size() + 1));
  for (size_t i = 0; i < indptr.size(); ++i) (*pos_out)[i] = indptr[i];
  for (size_t i = 0; i < indices.size(); ++i) (*crd_out)[i] = indices[i];
  return (int32_t)indices.size();
}

// The reference's result comparators, ds.h:463-499 (compare_array_my prints
// every mismatch to stdout, as the reference does).
float ref_compare_array(const float *x, const float *y, size_t n) {
  return compare_array(x, y, n);
}
float ref_compare_array_my(const float *x, const float *y, size_t n) {
  return compare_array_my(x, y, n);
}

void ref_free(void *p) { std::free(p); }

int32_t ref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"

// vendor.cpp — rocSPARSE CSR SpMM as the sweep's vendor-library row.
//
// The reference's sweep ends with a vendor-library row (cusparseSpMM with
// CUSPARSE_SPMM_CSR_ALG2 on row-major B and C, spmm_csr_gpu_cusparse.h:
// 171-260, logged as "alg2" by spmm_csr_gpu.h:343). On this hardware the
// vendor library is rocSPARSE; this file runs its generic SpMM on the same
// device operands the engine uses, timing only the compute stage (the
// reference's TIME_GPU brackets cusparseSpMM alone, :235-250). It is a
// comparison row: no engine schedule dispatches to it.
#include <rocsparse/rocsparse.h>

#include <mutex>

#include "engine.h"
#include "segtaco/spmm.h"

namespace segtaco {
namespace {

struct VendorState {
  std::mutex mu;
  rocsparse_handle handle = nullptr;
  void *buf = nullptr;
  size_t buf_bytes = 0;
};

VendorState &vstate() {
  static VendorState s;
  return s;
}

int32_t sparse_error(rocsparse_status st, const char *what) {
  return set_error(SEGTACO_ERR_HIP, "%s failed: rocsparse status %d", what, (int)st);
}

#define SEGTACO_SPARSE(call)                                             \
  do {                                                                   \
    rocsparse_status st_ = (call);                                       \
    if (st_ != rocsparse_status_success) return sparse_error(st_, #call); \
  } while (0)

rocsparse_spmm_alg alg_of(int32_t alg) {
  switch (alg) {
    case 1: return rocsparse_spmm_alg_csr;
    case 2: return rocsparse_spmm_alg_csr_row_split;
    case 3: return rocsparse_spmm_alg_csr_nnz_split;
    case 4: return rocsparse_spmm_alg_csr_merge_path;
    default: return rocsparse_spmm_alg_default;
  }
}

struct Descs {
  rocsparse_spmat_descr A = nullptr;
  rocsparse_dnmat_descr B = nullptr, C = nullptr;
  ~Descs() {
    if (A) rocsparse_destroy_spmat_descr(A);
    if (B) rocsparse_destroy_dnmat_descr(B);
    if (C) rocsparse_destroy_dnmat_descr(C);
  }
};

}  // namespace

// C = A * B with rocSPARSE (FP32 values, row-major B and C). ev0/ev1 (may be
// null) bracket the compute stage on `stream`.
int32_t vendor_spmm(const segtaco_csr_t *A, const float *B, int64_t ldb,
                    int32_t N, float *C, int64_t ldc, int32_t alg,
                    hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
  if (!A || A->M < 0 || A->K < 0 || A->nnz < 0 || N < 0)
    return set_error(SEGTACO_ERR_INVALID, "vendor spmm: bad operand");
  if (A->vals_dtype != SEGTACO_DT_F32 || (A->nnz > 0 && !A->vals))
    return set_error(SEGTACO_ERR_UNSUPPORTED, "vendor spmm: needs FP32 values");
  if (ldb < N || ldc < N)
    return set_error(SEGTACO_ERR_INVALID, "vendor spmm: leading dimension < N");
  VendorState &s = vstate();
  std::lock_guard<std::mutex> lock(s.mu);
  if (!s.handle) SEGTACO_SPARSE(rocsparse_create_handle(&s.handle));
  SEGTACO_SPARSE(rocsparse_set_stream(s.handle, stream));
  Descs d;
  SEGTACO_SPARSE(rocsparse_create_csr_descr(
      &d.A, A->M, A->K, A->nnz, const_cast<int32_t *>(A->pos),
      const_cast<int32_t *>(A->crd), const_cast<void *>(A->vals),
      rocsparse_indextype_i32, rocsparse_indextype_i32, rocsparse_index_base_zero,
      rocsparse_datatype_f32_r));
  SEGTACO_SPARSE(rocsparse_create_dnmat_descr(&d.B, A->K, N, ldb, const_cast<float *>(B),
                                              rocsparse_datatype_f32_r, rocsparse_order_row));
  SEGTACO_SPARSE(rocsparse_create_dnmat_descr(&d.C, A->M, N, ldc, C, rocsparse_datatype_f32_r,
                                              rocsparse_order_row));
  const float alpha = 1.0f, beta = 0.0f;
  const rocsparse_spmm_alg a = alg_of(alg);
  size_t need = 0;
  SEGTACO_SPARSE(rocsparse_spmm(s.handle, rocsparse_operation_none, rocsparse_operation_none,
                                &alpha, d.A, d.B, &beta, d.C, rocsparse_datatype_f32_r, a,
                                rocsparse_spmm_stage_buffer_size, &need, nullptr));
  if (need > s.buf_bytes) {
    if (s.buf) {
      SEGTACO_HIP(hipStreamSynchronize(stream));
      SEGTACO_HIP(hipFree(s.buf));
      s.buf = nullptr;
      s.buf_bytes = 0;
    }
    SEGTACO_HIP(hipMalloc(&s.buf, need));
    s.buf_bytes = need;
  }
  SEGTACO_SPARSE(rocsparse_spmm(s.handle, rocsparse_operation_none, rocsparse_operation_none,
                                &alpha, d.A, d.B, &beta, d.C, rocsparse_datatype_f32_r, a,
                                rocsparse_spmm_stage_preprocess, &need, s.buf));
  if (ev0) SEGTACO_HIP(hipEventRecord(ev0, stream));
  SEGTACO_SPARSE(rocsparse_spmm(s.handle, rocsparse_operation_none, rocsparse_operation_none,
                                &alpha, d.A, d.B, &beta, d.C, rocsparse_datatype_f32_r, a,
                                rocsparse_spmm_stage_compute, &need, s.buf));
  if (ev1) SEGTACO_HIP(hipEventRecord(ev1, stream));
  // descriptors are released on return; the work they describe is already
  // enqueued with the pointers it needs
  return SEGTACO_OK;
}

rocsparse_spmv_alg spmv_alg_of(int32_t alg) {
  switch (alg) {
    case 1: return rocsparse_spmv_alg_csr_adaptive;
    case 2: return rocsparse_spmv_alg_csr_rowsplit;
    case 3: return rocsparse_spmv_alg_csr_lrb;
    case 4: return rocsparse_spmv_alg_csr_nnzsplit;
    default: return rocsparse_spmv_alg_default;
  }
}

// y = A x with rocSPARSE's SpMV (FP32). ev0/ev1 (may be null) bracket the
// compute stage on `stream`; the preprocessing stage (adaptive / LRB
// analysis of A) runs before ev0, as the SpMM row's does.
int32_t vendor_spmv(const segtaco_csr_t *A, const float *x, float *y, int32_t alg,
                    hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1) {
  if (!A || A->M < 0 || A->K < 0 || A->nnz < 0 || !x || !y)
    return set_error(SEGTACO_ERR_INVALID, "vendor spmv: bad operand");
  if (A->vals_dtype != SEGTACO_DT_F32 || (A->nnz > 0 && !A->vals))
    return set_error(SEGTACO_ERR_UNSUPPORTED, "vendor spmv: needs FP32 values");
  VendorState &s = vstate();
  std::lock_guard<std::mutex> lock(s.mu);
  if (!s.handle) SEGTACO_SPARSE(rocsparse_create_handle(&s.handle));
  SEGTACO_SPARSE(rocsparse_set_stream(s.handle, stream));
  rocsparse_spmat_descr mat = nullptr;
  rocsparse_dnvec_descr vx = nullptr, vy = nullptr;
  struct Release {
    rocsparse_spmat_descr &m;
    rocsparse_dnvec_descr &a, &b;
    ~Release() {
      if (m) rocsparse_destroy_spmat_descr(m);
      if (a) rocsparse_destroy_dnvec_descr(a);
      if (b) rocsparse_destroy_dnvec_descr(b);
    }
  } release{mat, vx, vy};
  SEGTACO_SPARSE(rocsparse_create_csr_descr(
      &mat, A->M, A->K, A->nnz, const_cast<int32_t *>(A->pos),
      const_cast<int32_t *>(A->crd), const_cast<void *>(A->vals),
      rocsparse_indextype_i32, rocsparse_indextype_i32, rocsparse_index_base_zero,
      rocsparse_datatype_f32_r));
  SEGTACO_SPARSE(rocsparse_create_dnvec_descr(&vx, A->K, const_cast<float *>(x),
                                              rocsparse_datatype_f32_r));
  SEGTACO_SPARSE(rocsparse_create_dnvec_descr(&vy, A->M, y, rocsparse_datatype_f32_r));
  const float alpha = 1.0f, beta = 0.0f;
  const rocsparse_spmv_alg a = spmv_alg_of(alg);
  size_t need = 0;
  // the generic SpMV entry point of this ROCm release (its successor takes
  // a separate descriptor object; the computation is the same)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"
  auto spmv = [&](rocsparse_spmv_stage stage, void *buf) {
    return rocsparse_spmv(s.handle, rocsparse_operation_none, &alpha, mat, vx, &beta, vy,
                          rocsparse_datatype_f32_r, a, stage, &need, buf);
  };
#pragma clang diagnostic pop
  SEGTACO_SPARSE(spmv(rocsparse_spmv_stage_buffer_size, nullptr));
  if (need > s.buf_bytes) {
    if (s.buf) {
      SEGTACO_HIP(hipStreamSynchronize(stream));
      SEGTACO_HIP(hipFree(s.buf));
      s.buf = nullptr;
      s.buf_bytes = 0;
    }
    SEGTACO_HIP(hipMalloc(&s.buf, need > 0 ? need : 1));
    s.buf_bytes = need > 0 ? need : 1;
  }
  SEGTACO_SPARSE(spmv(rocsparse_spmv_stage_preprocess, s.buf));
  // y cleared first (outside the timed stage): with beta = 0 some of the
  // algorithms still read y, and 0 * NaN left there would survive
  if (A->M > 0) SEGTACO_HIP(hipMemsetAsync(y, 0, sizeof(float) * (size_t)A->M, stream));
  if (ev0) SEGTACO_HIP(hipEventRecord(ev0, stream));
  SEGTACO_SPARSE(spmv(rocsparse_spmv_stage_compute, s.buf));
  if (ev1) SEGTACO_HIP(hipEventRecord(ev1, stream));
  return SEGTACO_OK;
}

}  // namespace segtaco

extern "C" int32_t segtaco_vendor_spmv(const segtaco_csr_t *A, const float *x, float *y,
                                       int32_t alg, void *stream, float *ms_out) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!ms_out) return segtaco::vendor_spmv(A, x, y, alg, s, nullptr, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  SEGTACO_HIP(hipEventCreate(&e0));
  SEGTACO_HIP(hipEventCreate(&e1));
  int32_t st = segtaco::vendor_spmv(A, x, y, alg, s, e0, e1);
  if (st == SEGTACO_OK) {
    if (hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms_out, e0, e1) != hipSuccess)
      st = segtaco::set_error(SEGTACO_ERR_HIP, "segtaco_vendor_spmv: event timing");
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

extern "C" int32_t segtaco_vendor_spmm(const segtaco_csr_t *A, const float *B,
                                       int64_t ldb, int32_t N, float *C,
                                       int64_t ldc, int32_t alg, void *stream,
                                       float *ms_out) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!ms_out) return segtaco::vendor_spmm(A, B, ldb, N, C, ldc, alg, s, nullptr, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  SEGTACO_HIP(hipEventCreate(&e0));
  SEGTACO_HIP(hipEventCreate(&e1));
  int32_t st = segtaco::vendor_spmm(A, B, ldb, N, C, ldc, alg, s, e0, e1);
  if (st == SEGTACO_OK) {
    if (hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(ms_out, e0, e1) != hipSuccess)
      st = segtaco::set_error(SEGTACO_ERR_HIP, "segtaco_vendor_spmm: event timing");
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

/*
 * sddmm.h — sampled dense-dense matrix product on a CSR pattern.
 *
 *   out[p] = s_p * sum_{j < J} X[i][j] * Y[k][j]   for each stored entry
 *                                                  p = (i, k) of S
 *
 * The reference's SDDMM (sddmm_csr_gpu_taco / sddmm_csr_gpu_warp,
 * sddmm_csr_gpu_taco.h:5-148; disabled in its build, gpu_kernels.cuh:25-26)
 * computes the same product in FP64 with one atomic per entry. This one
 * stores S's values, X and Y as FP32, BF16 or FP16 and accumulates in FP32;
 * each entry's dot product has a fixed order (per-lane sequential over V
 * elements, then a butterfly over the lanes), so results are deterministic.
 * It reuses the SpMM path's row partition (the repaired K21 array).
 */
#ifndef SEGTACO_SDDMM_H
#define SEGTACO_SDDMM_H
#include <stddef.h>
#include <stdint.h>

#include "segtaco/spmm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device bytes of workspace segtaco_sddmm needs for S. */
size_t segtaco_sddmm_workspace_bytes(const segtaco_csr_t *S);

/* out[p] for p < S->nnz (device pointers; asynchronous on `stream`). X is
 * S->M x J with leading dimension ldx >= J, Y is S->K x J with ldy >= J,
 * both row-major in `dtype` storage (SEGTACO_DT_*); S's values (or NULL for
 * all ones) must have the same storage type. */
int32_t segtaco_sddmm(const segtaco_csr_t *S, const void *X, int64_t ldx,
                      const void *Y, int64_t ldy, int32_t J, int32_t dtype,
                      float *out, void *workspace, size_t workspace_bytes,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif

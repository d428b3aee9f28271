/*
 * taco_tensor_t — the operand container of the segTACO operator surface.
 *
 * Layout is identical to the reference struct (taco_tensor_t.h:6-15) so that
 * a caller built against the reference links against libsegtaco unchanged.
 *
 * For SpMM (C = A * B):
 *   A (CSR):   dimensions = {M, K}; indices[1][0] = pos int32[M+1];
 *              indices[1][1] = crd int32[nnz]; vals = float[nnz] (NULL means
 *              every stored value reads as 1.0, as __guard_load_default_one,
 *              gpu_library.h:154-161).
 *   B (dense): dimensions = {K, N}; vals = float[K*N], row-major, ld = N.
 *   C (dense): dimensions = {M, N}; vals = float[M*N], row-major, ld = N.
 */
#ifndef SEGTACO_TACO_TENSOR_T_H
#define SEGTACO_TACO_TENSOR_T_H
#include <stdint.h>

#ifndef TACO_TENSOR_T_DEFINED
#define TACO_TENSOR_T_DEFINED
typedef enum { taco_mode_dense, taco_mode_sparse } taco_mode_t;

typedef struct taco_tensor_t {
  int32_t      order;         /* tensor order (number of modes) */
  int32_t*     dimensions;    /* tensor dimensions */
  int32_t      csize;         /* component size */
  int32_t*     mode_ordering; /* mode storage ordering */
  taco_mode_t* mode_types;    /* mode storage types */
  int32_t***   indices;       /* tensor index data (per mode) */
  float*       vals;          /* tensor values */
  int32_t      vals_size;     /* values array size */
} taco_tensor_t;
#endif

#endif

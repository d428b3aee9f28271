/*
 * batch.h — many heterogeneous SpMMs in one persistent launch.
 *
 * No counterpart in the reference (one matrix per process, one kernel per
 * call, matrix_main.cpp:13-57). A batch fixes a set of device-resident
 * products C_i = A_i * B_i (each with its own shape, N_i and plan) and runs
 * them all with one persistent kernel whose workgroups pull work units from a
 * shared counter, followed by one batched carry fix-up. Creating the batch
 * resolves every plan (the autotuner picks one from the row-length
 * statistics when the item's schedule is negative) and runs the per-matrix
 * row partition once; running it replays only the compute. The run sequence
 * can be captured into a HIP graph and replayed with one launch.
 */
#ifndef SEGTACO_BATCH_H
#define SEGTACO_BATCH_H
#include <stddef.h>
#include <stdint.h>

#include "segtaco/spmm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SEGTACO_SCHEDULE_AUTO (-1)

typedef struct segtaco_batch_item {
  segtaco_csr_t A;      /* device CSR; vals_dtype == the batch dtype or NULL */
  const void *B;        /* device, K x N, row-major, ld = ldb */
  int64_t ldb;
  float *C;             /* device, M x N, row-major, ld = ldc */
  int64_t ldc;
  int32_t N;
  int32_t reserved;
  segtaco_plan_t plan;  /* schedule SEGTACO_SCHEDULE_AUTO: tuned per matrix */
  const int32_t *pos_host; /* host copy of A.pos for the tuner (NULL: the
                              tuner reads A.pos back from the device) */
} segtaco_batch_item_t;

typedef struct segtaco_batch *segtaco_batch_t;

/* Builds a batch over n items whose vals and B share storage b_dtype.
 * Allocates the batch's device workspace, resolves plans and computes the
 * row partitions on `stream` (synchronizes it before returning). */
int32_t segtaco_batch_create(const segtaco_batch_item_t *items, int32_t n,
                             int32_t b_dtype, void *stream,
                             segtaco_batch_t *out);
/* Runs every product of the batch on `stream` (asynchronous). */
int32_t segtaco_batch_run(segtaco_batch_t b, void *stream);
/* Captures the run sequence into a HIP graph (on an internal stream). */
int32_t segtaco_batch_capture(segtaco_batch_t b);
/* Launches the captured graph on `stream` (asynchronous). */
int32_t segtaco_batch_replay(segtaco_batch_t b, void *stream);
/* Resolved plan of item i; number of work units of the persistent launch;
 * bytes of device workspace the batch holds. */
int32_t segtaco_batch_plan(segtaco_batch_t b, int32_t i, segtaco_plan_t *out);
int64_t segtaco_batch_units(segtaco_batch_t b);
/* Item i's row partition as the batch holds it (device pointers, valid until
 * the batch is destroyed): the K21 array (gpu_library.h:206-215) in *part and
 * the repaired first row of each tile in *tile_rows (ntiles + 1 entries
 * each); *ntiles = 0 and NULL pointers for row-balanced items. */
int32_t segtaco_batch_partition(segtaco_batch_t b, int32_t i, int32_t *ntiles,
                                const int32_t **part, const int32_t **tile_rows);
size_t segtaco_batch_workspace_bytes(segtaco_batch_t b);
int32_t segtaco_batch_destroy(segtaco_batch_t b);

#ifdef __cplusplus
}
#endif
#endif

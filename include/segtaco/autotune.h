/*
 * autotune.h — plan selection from a matrix's row-length histogram.
 *
 * No counterpart in the reference, which encodes schedule parameters in 61
 * function names and leaves the choice to an offline sweep
 * (matrix_experiments/spmm_csr_gpu.h:283-343, resana.py). Here the
 * schedule (rb/eb x sr/pr), lane grouping, column tile and nnz tile are
 * chosen per matrix either by a model over the row-length statistics or by
 * timing a candidate set on the device; measured choices are cached by a
 * signature of the statistics and can be saved to / loaded from a file.
 */
#ifndef SEGTACO_AUTOTUNE_H
#define SEGTACO_AUTOTUNE_H
#include <stddef.h>
#include <stdint.h>

#include "segtaco/spmm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SEGTACO_HIST_BINS 24

/* Row-length statistics. hist[0] counts empty rows; hist[b] (b >= 1) counts
 * rows with length in [2^(b-1), 2^b); the last bin is open-ended. */
typedef struct segtaco_row_stats {
  int32_t M, K, nnz, max_len, empty_rows;
  int32_t reserved;
  double mean_len;      /* nnz / M */
  double cv;            /* standard deviation / mean of the row lengths */
  double frac_nnz_long; /* share of entries in rows longer than 1024 */
  int64_t hist[SEGTACO_HIST_BINS];
} segtaco_row_stats_t;

/* Statistics of a CSR row pointer held in host memory. */
int32_t segtaco_row_stats(const int32_t *pos_host, int32_t M, int32_t K,
                          segtaco_row_stats_t *out);

/* Model choice (deterministic, no device work): a fully resolved plan for
 * N columns of B with storage b_dtype. */
int32_t segtaco_tune_model(const segtaco_row_stats_t *st, int32_t N,
                           int32_t b_dtype, segtaco_plan_t *out);

/* Measured choice: times the model's plan and a candidate set around it,
 * SEGTACO_EB_CS included (reps timed runs each, after one warm-up), on
 * `stream` with operands already on the device, writes C with the fastest
 * plan, and returns that plan and its mean ms. The plan cache, which batches
 * and the server read, stores under the signature the fastest candidate they
 * can run (SEGTACO_RB_SR .. SEGTACO_EB_PR), which differs from the returned
 * plan when SEGTACO_EB_CS wins. */
int32_t segtaco_tune_measure(const segtaco_csr_t *A,
                             const segtaco_row_stats_t *st, const void *B,
                             int32_t b_dtype, int64_t ldb, int32_t N, float *C,
                             int64_t ldc, int32_t reps, void *stream,
                             segtaco_plan_t *out, float *out_ms);

/* Signature of (statistics, N, dtype) used as the cache key: quantized
 * log2(M), log2(mean), cv, share of long-row entries, log2(max length). */
uint64_t segtaco_tune_signature(const segtaco_row_stats_t *st, int32_t N,
                                int32_t b_dtype);
/* Cached plan for a signature: SEGTACO_OK if present, SEGTACO_ERR_STATE if
 * not. */
int32_t segtaco_tune_cache_get(uint64_t signature, segtaco_plan_t *out);
int32_t segtaco_tune_cache_put(uint64_t signature, const segtaco_plan_t *plan);
int32_t segtaco_tune_cache_clear(void);
/* Text file, one "signature schedule vec col_lanes group tile comp" line per
 * entry; load merges into the in-memory cache. */
int32_t segtaco_tune_cache_save(const char *path);
int32_t segtaco_tune_cache_load(const char *path);

/* Best plan for a matrix: the cached measurement if its signature is known,
 * else the model. */
int32_t segtaco_tune(const segtaco_row_stats_t *st, int32_t N, int32_t b_dtype,
                     segtaco_plan_t *out);

#ifdef __cplusplus
}
#endif
#endif

/*
 * serve.h — continuous-batching SpMM server over a device-resident CSR cache.
 *
 * No counterpart in the reference (a benchmark harness: one matrix per
 * process, matrix_main.cpp). A server owns:
 *   * a registry of CSR matrices (host copies kept in pinned memory);
 *   * a device cache of resident matrices (capacity cache_bytes, least
 *     recently used matrices not needed by work on the device are evicted);
 *   * a scheduler thread: requests C = A_id * B (B and C in device memory)
 *     are admitted in arrival order while their device footprint fits
 *     (inflight_bytes for per-request workspace, the cache for their matrix),
 *     packed into ticks of up to max_batch products, and each tick runs as one
 *     batched persistent launch on one of `streams` streams, so requests that
 *     arrive while earlier ticks are on the GPU are admitted into new ticks
 *     whose work overlaps the running ones.
 * Every request's result is bit-identical to running it alone: the plan of a
 * (matrix, N) pair is fixed and no schedule uses floating-point atomics.
 * Latency is measured from submission to completion of the request's tick on
 * the device (B and C are device pointers: no PCIe transfer is included).
 */
#ifndef SEGTACO_SERVE_H
#define SEGTACO_SERVE_H
#include <stddef.h>
#include <stdint.h>

#include "segtaco/spmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct segtaco_serve_config {
  size_t cache_bytes;    /* device bytes for resident CSR matrices */
  size_t inflight_bytes; /* device bytes for admitted requests' workspace */
  int32_t max_batch;     /* products per tick */
  int32_t streams;       /* ticks on the device at once */
  int32_t dtype;         /* storage of vals and B (SEGTACO_DT_*) */
  int32_t reserved;
} segtaco_serve_config_t;

typedef struct segtaco_serve_stats {
  int64_t submitted, completed, rejected, ticks;
  int64_t cache_hits, cache_misses, evictions;
  int64_t uploaded_bytes;
  double mean_batch;        /* products per tick */
  double p50_us, p99_us, max_us, mean_us; /* submit -> done */
  double gflops;            /* 2*nnz*N of completed requests / wall time */
  double wall_s;            /* first submission -> last completion */
  int32_t max_concurrent_ticks;
  int32_t reserved;
} segtaco_serve_stats_t;

typedef struct segtaco_server *segtaco_server_t;

int32_t segtaco_server_create(const segtaco_serve_config_t *cfg,
                              segtaco_server_t *out);
/* Registers a CSR matrix (host arrays, vals of the server dtype or NULL);
 * the data is copied. Returns its id in *matrix_id. */
int32_t segtaco_server_add_matrix(segtaco_server_t s, int32_t M, int32_t K,
                                  int32_t nnz, const int32_t *pos,
                                  const int32_t *crd, const void *vals,
                                  int32_t *matrix_id);
/* Queues C = A[matrix_id] * B (device pointers). *request_id identifies it.
 * A request that can never fit the configured capacities completes at once
 * with status SEGTACO_ERR_CAPACITY, and so does a request at the head of
 * the queue that cannot be admitted while nothing runs (what it needs is
 * held by nothing the server can release). After a HIP error in a tick the
 * tick's requests complete with SEGTACO_ERR_HIP, so do all requests still
 * queued, and further submissions are refused. */
int32_t segtaco_server_submit(segtaco_server_t s, int32_t matrix_id,
                              const void *B, int64_t ldb, int32_t N, float *C,
                              int64_t ldc, int64_t *request_id);
/* Waits for one request (timeout_s < 0: forever); *status gets its outcome.
 * Returns SEGTACO_ERR_STATE on timeout. */
int32_t segtaco_server_wait(segtaco_server_t s, int64_t request_id,
                            double timeout_s, int32_t *status);
/* Waits until every submitted request has completed. */
int32_t segtaco_server_drain(segtaco_server_t s, double timeout_s);
/* Submits n requests at the given arrival offsets (microseconds after the
 * call) from a timing thread, then drains. ids_out (optional) receives the
 * request ids. */
int32_t segtaco_server_replay(segtaco_server_t s, int32_t n,
                              const double *arrival_us,
                              const int32_t *matrix_ids, const void *const *B,
                              const int64_t *ldb, const int32_t *N,
                              float *const *C, const int64_t *ldc,
                              int64_t *ids_out, double timeout_s);
int32_t segtaco_server_stats(segtaco_server_t s, segtaco_serve_stats_t *out);
int32_t segtaco_server_reset_stats(segtaco_server_t s);
/* Plan the server uses for (matrix, N). */
int32_t segtaco_server_plan(segtaco_server_t s, int32_t matrix_id, int32_t N,
                            segtaco_plan_t *out);
/* Error-path testing: the launch of the tick `ticks_ahead` ticks from now
 * (1 = the next one) reports a HIP error without running any of its device
 * work, exactly as a failed launch would be handled. */
int32_t segtaco_server_debug_fail_tick(segtaco_server_t s, int64_t ticks_ahead);
int32_t segtaco_server_destroy(segtaco_server_t s);

#ifdef __cplusplus
}
#endif
#endif

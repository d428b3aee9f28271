/*
 * spmm.h — C-ABI of libsegtaco, the MI355X (gfx950) CSR SpMM engine.
 *
 * Two layers:
 *
 * 1. The segTACO operator surface (reference gpu_kernels.cuh:4-93): entry
 *    points `void f(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B)`
 *    with host operands. Each call copies A and B to the device, computes
 *    C = A * B, copies C back and frees device memory, and stores the elapsed
 *    milliseconds of the device work (partition + kernels, the reference's
 *    TIME_GPU scope, gpu_library.h:853-857) for get_gpu_timer_result().
 *    Prior contents of C are ignored. The reference declares these with C++
 *    linkage; libsegtaco exports them with C linkage (this header) and also
 *    with C++ linkage, so objects compiled against the reference header link
 *    unchanged.
 *
 * 2. A device-resident API (segtaco_*): CSR operands, B and C already in
 *    device memory, explicit plans (schedule, vector width, lane grouping,
 *    tile), caller-provided workspace and stream, integer status codes. The
 *    batched and serving layers (batch.h, serve.h) are built on it.
 *
 * Semantics shared by both layers: C (M x N, row-major, ld >= N) is
 * overwritten with A * B; A is M x K CSR with int32 pos[M+1] and crd[nnz]
 * (columns in [0, K)); B is K x N row-major; a NULL vals pointer reads every
 * stored value as 1.0 (gpu_library.h:154-161). Every schedule is free of
 * floating-point atomics, so results are bit-identical run to run.
 */
#ifndef SEGTACO_SPMM_H
#define SEGTACO_SPMM_H

#include <stddef.h>
#include <stdint.h>

#include "segtaco/taco_tensor_t.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes and error mode                                               */
/* ------------------------------------------------------------------------ */
#define SEGTACO_OK 0
#define SEGTACO_ERR_INVALID 1     /* bad argument, shape or CSR structure */
#define SEGTACO_ERR_HIP 2         /* HIP runtime error */
#define SEGTACO_ERR_NOMEM 3       /* device or host allocation failed */
#define SEGTACO_ERR_UNSUPPORTED 4 /* plan not supported for these operands */
#define SEGTACO_ERR_CAPACITY 5    /* request larger than a configured limit */
#define SEGTACO_ERR_STATE 6       /* object not in a state for this call */

/* Error behaviour of the operator-surface entry points, which return void.
 * SEGTACO_ERRMODE_EXIT (default) prints the error and exits the process with
 * a non-zero code, as the reference's gpuErrchk does (gpu_library.h:22-30).
 * SEGTACO_ERRMODE_RECORD records the error (segtaco_last_error) and returns.
 * Returns the previous mode. */
#define SEGTACO_ERRMODE_EXIT 0
#define SEGTACO_ERRMODE_RECORD 1
int32_t segtaco_set_error_mode(int32_t mode);
/* Status of the last failing call on this thread (0 if none), and its
 * message. segtaco_clear_error resets both. */
int32_t segtaco_last_error(void);
const char *segtaco_last_error_string(void);
void segtaco_clear_error(void);

/* ------------------------------------------------------------------------ */
/* Operator surface (reference gpu_kernels.cuh)                              */
/* ------------------------------------------------------------------------ */
/* gpu_kernels.cuh:4, gpu_kernels.cu:20-22: ms of the last call's device work
 * (-1 before any call). */
float get_gpu_timer_result(void);

/* The operator surface as a table, for drivers and sweeps that iterate over
 * it: segtaco_legacy_count() entry points, id -> name, and a call by id
 * (equivalent to calling the named function). */
int32_t segtaco_legacy_count(void);
const char *segtaco_legacy_name(int32_t id);
void segtaco_legacy_call(int32_t id, taco_tensor_t *C, taco_tensor_t *A,
                         taco_tensor_t *B);

/* The five primary entry points (the four schedules and the TACO kernel). */
/* spmm_rb_sr.h:220-243: row-balanced, sequential reduction. */
void spmm_rb_sr(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_rb_pr.h:333-375: row-balanced, parallel (tree) reduction. */
void spmm_rb_pr(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_eb_sr.h:335-423: element-balanced, sequential reduction. */
void spmm_eb_sr(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_eb_pr.h:435-480: element-balanced, segmented parallel reduction. */
void spmm_eb_pr(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_csr_gpu_taco.h:557-587: element-balanced sequential reduction over a
 * 256-nnz partition (the reference restricts it to N = 128; this one takes
 * any N). */
void spmm_csr_gpu_taco(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);

/* The remaining live names of gpu_kernels.cuh:12-93. Each runs one of the
 * four schedules with the (group, tile, vector) preset its name encodes;
 * unlike the reference's tuned kernels, every one is correct for any N. */
void spmm_csr_gpu_warp(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);      /* spmm_csr_gpu_taco.h:45 */
void spmm_csr_gpu_reduction(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B); /* spmm_csr_gpu_taco.h:614 */
void spmm_csr_pr_eb(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);         /* spmm_csr_gpu_taco.h:589 */
void spmm_rb_sr_taco(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);        /* spmm_rb_sr.h:44 */
void spmm_rb_pr_taco(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);        /* spmm_rb_pr.h:46 */
void spmm_eb_sr_taco(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);        /* spmm_eb_sr.h:58 */
void spmm_eb_pr_taco(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);        /* spmm_eb_pr.h:54 */
void spmm_eb_sr_search(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);      /* spmm_eb_sr.h:615 */
void spmm_eb_pr_search(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);      /* spmm_eb_pr.h:1305 */
void spmm_rb_sr_512_64(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);      /* spmm_rb_sr.h:270 */
void spmm_eb_sr_512_64(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);      /* spmm_eb_sr.h:763 */
void spmm_rb_pr_256(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);         /* spmm_rb_pr.h:412 */
/* spmm_rb_pr.h:469-948: spmm_rb_pr_256_<c>_<g>_<gs> */
void spmm_rb_pr_256_1_4_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_4_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_4_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_8_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_8_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_8_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_8_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_8_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_8_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_16_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_16_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_16_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_16_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_16_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_16_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_16_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_16_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_16_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_32_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_32_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_32_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_32_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_32_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_32_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_32_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_32_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_32_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_1_32_32(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_2_32_32(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_pr_256_4_32_32(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_eb_pr.h:526-779: spmm_eb_pr_<Nnz>_<gs> */
void spmm_eb_pr_256_32(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_256_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_256_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_256_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_128_32(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_128_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_128_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_128_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_64_32(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_64_16(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_64_8(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_pr_64_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_rb_sr.h:392-535: spmm_rb_sr_256_<r>_<c> */
void spmm_rb_sr_256_1_1(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_1_2(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_1_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_2_1(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_2_2(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_2_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_4_1(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_4_2(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_rb_sr_256_4_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
/* spmm_eb_sr.h:908-1095: spmm_eb_sr_256_<g>_<c> */
void spmm_eb_sr_256_4_1(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_4_2(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_4_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_8_1(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_8_2(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_8_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_16_1(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_16_2(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_eb_sr_256_16_4(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);

/* The vendor-library rows (spmm_csr_gpu_cusparse.h:52-170 and :171-260, the
 * sweep's "alg2" row, spmm_csr_gpu.h:343): C = A * B with rocSPARSE's CSR
 * SpMM (default algorithm, row-major B and C), the timer covering its compute
 * stage only as the reference's does. A comparison row for sweeps; values
 * must be present. */
void spmm_csr_cusparse_row(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);
void spmm_csr_gpu_cusparse(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B);

/* ------------------------------------------------------------------------ */
/* Device-resident API                                                       */
/* ------------------------------------------------------------------------ */
#define SEGTACO_DT_F32 0
#define SEGTACO_DT_BF16 1
#define SEGTACO_DT_F16 2

#define SEGTACO_RB_SR 0 /* row-balanced, sequential reduction */
#define SEGTACO_RB_PR 1 /* row-balanced, parallel (tree) reduction */
#define SEGTACO_EB_SR 2 /* element-balanced, sequential reduction */
#define SEGTACO_EB_PR 3 /* element-balanced, segmented parallel reduction */
/* eb_sr's tiles, each swept in ascending column order so that all of the
 * device's lane groups gather from a narrow band of B at a time (L2 reuse);
 * the rows of a unit of tiles accumulate in LDS. Same results as eb_sr with
 * the same tile for column-sorted rows; no compensated summation. */
#define SEGTACO_EB_CS 4
/* Block-dense, for BF16 / FP16 storage: rows in blocks of tile = 32
 * (default) or 64; a 32-row block whose entries span at most 256 columns, or
 * a 64-row block spanning at most 192 (else each of its 32-row halves as
 * above), is scattered into a dense tile and multiplied with the matrix
 * cores (MFMA, FP32 accumulation); wider blocks run a per-row vector walk.
 * For banded and block-structured A. Runs standalone only (not in batches
 * or the server). */
#define SEGTACO_BD 5

/* A schedule and its parameters. Zero fields are resolved by
 * segtaco_plan_resolve from the operands. */
typedef struct segtaco_plan {
  int32_t schedule;       /* SEGTACO_RB_SR .. SEGTACO_EB_PR */
  int32_t vec;            /* elements of B per lane per load (1,2,4; 8 for
                             16-bit storage); 0 = widest N allows */
  int32_t col_lanes;      /* lanes sharing one B row segment, power of two
                             1..64; the column tile is col_lanes * vec */
  int32_t group;          /* *_pr: nnz slots reduced together (power of two);
                             rb_sr: rows per lane group; 0 = auto */
  int32_t tile;           /* eb_*: nnz per tile (multiple of 8); 0 = auto:
                             eb_sr 128, halved (down to 8) while the product
                             would fill fewer than two workgroups per compute
                             unit; eb_pr 128 (256 at N = 1); bd: rows per
                             block, 32 (auto) or 64 */
  int32_t comp_threshold; /* rows with more nnz than this use compensated
                             (Neumaier) summation; 0 = never; -1 = default
                             (1024 for 16-bit storage, never for FP32) */
} segtaco_plan_t;

/* CSR operand in device memory. vals may be NULL (all ones). */
typedef struct segtaco_csr {
  int32_t M, K, nnz;
  int32_t vals_dtype; /* SEGTACO_DT_*; must equal the B dtype for 16-bit */
  const int32_t *pos;
  const int32_t *crd;
  const void *vals;
} segtaco_csr_t;

/* Fill zero fields of *plan for these operands; validates the rest. */
int32_t segtaco_plan_resolve(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                             int64_t ldb, int64_t ldc, segtaco_plan_t *plan);

/* Workspace bytes segtaco_spmm needs for this plan: exact for a resolved
 * plan; for a plan with open (zero) fields, enough for every resolution a
 * call with these operands can reach. */
size_t segtaco_spmm_workspace_bytes(const segtaco_csr_t *A, int32_t N,
                                    const segtaco_plan_t *plan);

/* C = A * B on `stream` (a hipStream_t; NULL = default stream). B has dtype
 * b_dtype, ldb >= N; C is FP32, ldc >= N. workspace may be NULL only when
 * segtaco_spmm_workspace_bytes is 0. Asynchronous. */
int32_t segtaco_spmm(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                     int64_t ldb, int32_t N, float *C, int64_t ldc,
                     const segtaco_plan_t *plan, void *workspace,
                     size_t workspace_bytes, void *stream);

/* Preparation of A for repeated products (A resident and unchanged between
 * calls): writes into `workspace` what segtaco_spmm computes from A alone —
 * for eb_cs the slot plan and the column-sorted records (the larger part of
 * an eb_cs call), for eb_sr / eb_pr the tile rows — on `stream`, without
 * reading B or C. The plan resolves as segtaco_plan_resolve resolves it for
 * these ld's (no operand pointers). No counterpart in the reference, whose
 * entry points redo this work on every call; the analogue of a vendor
 * library's SpMM preprocessing stage. */
int32_t segtaco_spmm_prepare(const segtaco_csr_t *A, int32_t b_dtype, int64_t ldb, int32_t N,
                             int64_t ldc, const segtaco_plan_t *plan, void *workspace,
                             size_t workspace_bytes, void *stream);

/* segtaco_spmm on a workspace whose last preparation (by
 * segtaco_spmm_prepare or by a segtaco_spmm call) was of this product: the
 * same device arrays of A with unchanged contents, the same N and storage
 * type, and a plan that resolves, for these operands, to the prepared one;
 * the workspace itself must not have been written since by anything else
 * (the caller, a batch, another library call given the same memory).
 * Skips the preparation (eb_cs); every other step runs as in segtaco_spmm,
 * and C is bit-identical to it. A workspace prepared for another product is
 * refused (SEGTACO_ERR_INVALID) for eb_cs; whether A's contents changed is
 * not checked. Calls on other streams than the preparation's must be ordered
 * after it by the caller. */
int32_t segtaco_spmm_prepared(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                              int64_t ldb, int32_t N, float *C, int64_t ldc,
                              const segtaco_plan_t *plan, void *workspace,
                              size_t workspace_bytes, void *stream);

/* segtaco_spmm without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * The workspace is a fixed array of the library in device memory,
 * SEGTACO_NOALLOC_WORKSPACE_BYTES long, one per GPU, held from the first call
 * on that GPU until the process ends; the call itself allocates nothing.
 * A product whose workspace (segtaco_spmm_workspace_bytes of its resolved
 * plan) exceeds that size is refused with SEGTACO_ERR_CAPACITY before any
 * device work. Results are identical to segtaco_spmm with the same plan.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (on different streams, or from different threads) race and must not be
 * issued; calls in one stream's order are safe. The array keeps the last
 * call's contents, and no call reads what an earlier one left there. */
#define SEGTACO_NOALLOC_WORKSPACE_BYTES ((size_t)128 << 20)
int32_t segtaco_spmm_noalloc(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                             int64_t ldb, int32_t N, float *C, int64_t ldc,
                             const segtaco_plan_t *plan, void *stream);

/* Runs segtaco_spmm `reps` times back to back on `stream` and returns the
 * mean ms of its three stages in out_ms[0..2]: row partition, main kernel,
 * carry fix-up (0 for stages a plan does not have), each timed with HIP
 * events on `stream`. Returns after the last run completed. */
int32_t segtaco_spmm_stage_times(const segtaco_csr_t *A, const void *B,
                                 int32_t b_dtype, int64_t ldb, int32_t N,
                                 float *C, int64_t ldc,
                                 const segtaco_plan_t *plan, void *workspace,
                                 size_t workspace_bytes, void *stream,
                                 int32_t reps, float *out_ms);

/* K21 row partition (gpu_library.h:206-215, 236-240), bit-exact:
 * out[b] = taco_binarySearchBefore(pos, 0, M, b * values_per_block) for
 * b = 0 .. num_blocks (num_blocks + 1 entries), device pointers. */
int32_t segtaco_block_starts(const int32_t *pos, int32_t M,
                             int32_t values_per_block, int32_t num_blocks,
                             int32_t *out, void *stream);

/* ------------------------------------------------------------------------ */
/* Partition introspection: the row partitions the schedules run on          */
/* ------------------------------------------------------------------------ */
#define SEGTACO_ROWS_SCATTER 0 /* one pass over pos (standalone calls, SDDMM) */
#define SEGTACO_ROWS_SEARCH 1  /* K21 search + repair per tile (batches, server) */

/* rows[b] = the row holding stored entry b*tile for b < ntiles =
 * ceil(nnz / tile), rows[ntiles] = M: the K21 search (gpu_library.h:185-215)
 * followed by the repair loop that steps over empty rows (spmm_eb_sr.h:
 * 39-42), i.e. the row each element-balanced tile starts at. With
 * SEGTACO_ROWS_SEARCH and part != NULL, part gets the unrepaired K21 array
 * (as segtaco_block_starts). Device pointers; asynchronous on `stream`. */
int32_t segtaco_tile_rows(const int32_t *pos, int32_t M, int32_t tile, int32_t ntiles,
                          int32_t method, int32_t *rows, int32_t *part, void *stream);

/* Where a call of segtaco_spmm with this resolved element-balanced plan
 * (eb_sr, eb_pr, eb_cs) keeps its partition in the workspace: out[0] =
 * ntiles; out[1] = byte offset of the per-tile fix-up records (int32
 * [ntiles]: the tile where the row ending in tile b started if it started
 * in an earlier tile, else -1); out[2] = byte offset of the tile rows (int32
 * [ntiles + 1], as segtaco_tile_rows); out[3], out[4] = byte offsets of the
 * carry arrays (float [ntiles][N] each: partial sums of rows continued from
 * / into a neighbouring tile). */
int32_t segtaco_spmm_workspace_layout(const segtaco_csr_t *A, int32_t N,
                                      const segtaco_plan_t *plan, int64_t *out);

/* Where an eb_cs call with this resolved plan keeps its preparation in the
 * workspace (after the element-balanced part above): out[0] = units (tiles
 * per unit: plan->group), out[1] = LDS slots per unit, out[2] = 1 for 64-bit
 * records ({column << 12 | code, value bits}), 0 for 96-bit ones ({column,
 * code, value bits}); byte offsets of out[3] the records (one per stored
 * entry: a sweeping tile's sorted by column and position with its rows' LDS
 * slots as codes, a walking tile's in CSR order with 0 head, 1 tail, 2 + k
 * its k-th interior row), out[4] the rows of each unit's slots (int32 [units]
 * [slots]), out[5] each tile's interior rows (int32 [ntiles][tile]), out[6]
 * each entry's in-tile row code (uint16 [nnz]), out[7] the slots each unit
 * uses (int32 [units]), out[8] each tile's interior row count (int32
 * [ntiles]), out[9] each tile's flags (uint8 [ntiles]: 1 head row, 2 tail
 * row, 4 walks). */
int32_t segtaco_cs_workspace_layout(const segtaco_csr_t *A, int32_t N,
                                    const segtaco_plan_t *plan, int64_t *out);

/* rows[p] for every stored entry p < nnz: its row as the eb_pr schedule
 * derives it inside tiles of `tile` entries (per 128-entry chunk, a
 * bisection in a staged window of the row pointer, continued in memory for
 * rows past the window) — the segment ids of its segmented reduction (the
 * reference computes them with binary_search_segment_number,
 * gpu_library.h:138-152). tile_rows as from segtaco_tile_rows. Device
 * pointers; asynchronous on `stream`. */
int32_t segtaco_entry_rows(const int32_t *pos, int32_t M, int32_t nnz, int32_t tile,
                           const int32_t *tile_rows, int32_t *rows, void *stream);

/* Host-side CSR structure check: pos[0] == 0, pos non-decreasing,
 * pos[M] == nnz, every crd in [0, K). Host pointers. */
int32_t segtaco_csr_validate_host(int32_t M, int32_t K, int32_t nnz,
                                  const int32_t *pos, const int32_t *crd);

/* Thin device-memory helpers for callers without HIP headers. kind:
 * 0 host->device, 1 device->host, 2 device->device. */
int32_t segtaco_device_malloc(void **ptr, size_t bytes);
int32_t segtaco_device_free(void *ptr);
int32_t segtaco_memcpy(void *dst, const void *src, size_t bytes, int32_t kind);
int32_t segtaco_memset(void *dst, int32_t value, size_t bytes);
int32_t segtaco_device_synchronize(void);
int32_t segtaco_stream_synchronize(void *stream);
const char *segtaco_version(void);
/* GPUs visible to the HIP runtime libsegtaco is bound to; a negative value is
 * minus the HIP error code of hipGetDeviceCount. */
int32_t segtaco_device_count(void);

/* rocSPARSE CSR SpMM on device operands (FP32 values, row-major B and C):
 * the in-box vendor library as a comparison for benchmarks, not an engine
 * schedule. alg: 0 default, 1 csr, 2 csr_row_split, 3 csr_nnz_split,
 * 4 csr_merge_path. With ms_out non-NULL the call is synchronous and
 * *ms_out gets the compute stage's event time. */
int32_t segtaco_vendor_spmm(const segtaco_csr_t *A, const float *B, int64_t ldb,
                            int32_t N, float *C, int64_t ldc, int32_t alg,
                            void *stream, float *ms_out);

/* rocSPARSE CSR SpMV, y = A x (FP32, contiguous x and y): the comparison
 * row for the N = 1 case (the reference's vendor row is cusparseSpMM,
 * spmm_csr_gpu_cusparse.h:171-260; rocSPARSE's dedicated SpMV is the fairer
 * opponent at one column). alg: 0 default, 1 csr_adaptive, 2 csr_rowsplit,
 * 3 csr_lrb, 4 csr_nnzsplit (which this ROCm release's generic SpMV
 * refuses for FP32 CSR). Timing as for segtaco_vendor_spmm. */
int32_t segtaco_vendor_spmv(const segtaco_csr_t *A, const float *x, float *y,
                            int32_t alg, void *stream, float *ms_out);

#ifdef __cplusplus
}
#endif
#endif

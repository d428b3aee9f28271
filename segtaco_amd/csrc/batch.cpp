// batch.cpp — host side of the batched persistent launch (segtaco/batch.h).
//
// Create: resolve each product's plan (tuned from its row-length statistics
// when asked), lay out one device workspace for all element-balanced
// products (partition, tile rows, carries), compute each product's work-unit
// range, and run the row partitions once. Run: reset the unit counter, one
// persistent launch over all units, one batched fix-up launch. The sequence
// can be captured into a HIP graph (no memset nodes: the counter is cleared
// by a kernel).
#include <climits>
#include <cstring>
#include <vector>

#include "engine.h"
#include "launch.h"
#include "spmv.h"
#include "segtaco/autotune.h"
#include "segtaco/batch.h"

namespace segtaco {
int32_t plan_resolve_impl(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                          int64_t ldb, int64_t ldc, segtaco_plan_t *plan);
}  // namespace segtaco

using namespace segtaco;

struct segtaco_batch {
  int32_t n = 0, dt = 0;
  std::vector<BatchDesc> descs;
  std::vector<segtaco_plan_t> plans;
  std::vector<int64_t> prefix, fix_prefix;
  int64_t total_units = 0, total_fix = 0;
  BatchDesc *d_desc = nullptr;
  int64_t *d_prefix = nullptr, *d_fix_prefix = nullptr;
  unsigned long long *d_counter = nullptr;
  char *ws = nullptr;
  size_t ws_bytes = 0;
  size_t lds = 0;
  int32_t grid = 0;
  bool sv = false;  // some product runs the N = 1 form of eb_pr
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void release(segtaco_batch *b) {
  if (!b) return;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  (void)hipFree(b->d_desc);
  (void)hipFree(b->d_prefix);
  (void)hipFree(b->d_fix_prefix);
  (void)hipFree(b->d_counter);
  (void)hipFree(b->ws);
  delete b;
}

int32_t tune_item(const segtaco_batch_item_t &it, int32_t dt,
                  segtaco_plan_t *plan) {
  const int32_t M = it.A.M;
  std::vector<int32_t> pos_copy;
  const int32_t *pos = it.pos_host;
  if (!pos) {
    pos_copy.resize((size_t)M + 1);
    SEGTACO_HIP(hipMemcpy(pos_copy.data(), it.A.pos, sizeof(int32_t) * (M + 1),
                          hipMemcpyDeviceToHost));
    pos = pos_copy.data();
  }
  segtaco_row_stats_t st;
  SEGTACO_RET(segtaco_row_stats(pos, M, it.A.K, &st));
  SEGTACO_RET(segtaco_tune(&st, it.N, dt, plan));
  plan->vec = 0;  // re-derive from the actual pointers and leading dims
  return SEGTACO_OK;
}

}  // namespace

namespace segtaco {

void fill_batch_desc(const segtaco_csr_t &A, const void *B, int64_t ldb,
                     float *C, int64_t ldc, int32_t N,
                     const segtaco_plan_t &plan, BatchDesc *D, int64_t *units,
                     int64_t *fix_units, size_t *lds, EbLayout *lay) {
  std::memset(D, 0, sizeof(*D));
  KArgs &a = D->a;
  a.M = A.M;
  a.K = A.K;
  a.N = N;
  a.nnz = A.nnz;
  a.pos = A.pos;
  a.crd = A.crd;
  a.vals = A.vals;
  a.B = B;
  a.ldb = ldb;
  // checked for 4-byte elements, which covers the 16-bit storage types too
  a.b_off32 = b_off32_ok(A.K, ldb, 4);
  a.C = C;
  a.ldc = ldc;
  a.lq = ilog2(plan.col_lanes);
  a.lg = ilog2(plan.group);
  a.tile = plan.tile;
  a.comp_threshold = plan.comp_threshold == 0 ? INT_MAX : plan.comp_threshold;
  a.panels = (int32_t)ceil_div(N, (int64_t)plan.col_lanes * plan.vec);
  D->schedule = plan.schedule;
  D->vec = plan.vec;
  *fix_units = 0;
  *lds = 0;
  switch (plan.schedule) {
    case SEGTACO_RB_SR:
      a.nblocks = (int32_t)ceil_div(
          A.M, (int64_t)kWavesPerBlock * (64 / plan.col_lanes) * plan.group);
      break;
    case SEGTACO_RB_PR:
      a.nblocks = (int32_t)ceil_div(
          A.M, (int64_t)(64 / (plan.col_lanes * plan.group)) * kWavesPerBlock);
      break;
    default: {
      *lay = eb_layout(A.nnz, N, plan.tile);
      a.ntiles = lay->ntiles;
      a.rows_per_slice = (int32_t)ceil_div(A.M, lay->ntiles);
      const int64_t tiles_per_group_block =
          (int64_t)kWavesPerBlock * (64 / plan.col_lanes);
      a.nblocks = plan.schedule == SEGTACO_EB_SR
                      ? (int32_t)ceil_div(lay->ntiles, tiles_per_group_block)
                      : (int32_t)ceil_div(lay->ntiles, kWavesPerBlock);
      *lds = plan.schedule == SEGTACO_EB_SR ? eb_sr_lds_bytes(a)
             : spmv_form(a, plan.vec)        ? kSvLdsBytes
                                             : eb_pr_lds_bytes(a);
      *fix_units = ceil_div(lay->ntiles, tiles_per_group_block) * a.panels;
      break;
    }
  }
  *units = (int64_t)a.nblocks * a.panels;
}

}  // namespace segtaco

extern "C" {

int32_t segtaco_batch_create(const segtaco_batch_item_t *items, int32_t n,
                             int32_t b_dtype, void *stream_v,
                             segtaco_batch_t *out) {
  if (!out || (n > 0 && !items) || n < 0)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_batch_create: bad argument");
  if (b_dtype < SEGTACO_DT_F32 || b_dtype > SEGTACO_DT_F16)
    return set_error(SEGTACO_ERR_INVALID, "unknown dtype %d", b_dtype);
  *out = nullptr;
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  segtaco_batch *b = new segtaco_batch();
  b->n = n;
  b->dt = b_dtype;
  b->descs.resize(n);
  b->plans.resize(n);
  b->prefix.resize(n + 1, 0);
  b->fix_prefix.resize(n + 1, 0);
  std::vector<size_t> ws_off(n, 0);
  std::vector<EbLayout> lay(n);
  size_t ws_total = 0;
  int32_t st = SEGTACO_OK;

  for (int32_t i = 0; i < n && st == SEGTACO_OK; ++i) {
    const segtaco_batch_item_t &it = items[i];
    const segtaco_csr_t &A = it.A;
    if (A.M < 0 || A.K < 0 || A.nnz < 0 || it.N < 0 ||
        (A.M > 0 && !A.pos) || (A.nnz > 0 && !A.crd) ||
        (A.M > 0 && it.N > 0 && (!it.C || it.ldc < it.N)) ||
        (A.nnz > 0 && it.N > 0 && (!it.B || it.ldb < it.N))) {
      st = set_error(SEGTACO_ERR_INVALID, "batch item %d: bad operands", i);
      break;
    }
    if (A.vals && A.vals_dtype != b_dtype) {
      st = set_error(SEGTACO_ERR_UNSUPPORTED, "batch item %d: vals dtype", i);
      break;
    }
    segtaco_plan_t plan = it.plan;
    if (A.nnz == 0) {
      plan = segtaco_plan_t{SEGTACO_RB_SR, 0, 0, 0, 0, -1};  // writes zero rows
    } else if (plan.schedule < 0) {
      st = tune_item(it, b_dtype, &plan);
      if (st != SEGTACO_OK) break;
    }
    BatchDesc &D = b->descs[i];
    std::memset(&D, 0, sizeof(D));
    if (A.M == 0 || it.N == 0) {
      b->plans[i] = plan;
      b->prefix[i + 1] = b->prefix[i];
      b->fix_prefix[i + 1] = b->fix_prefix[i];
      continue;
    }
    if (plan.schedule == SEGTACO_EB_CS || plan.schedule == SEGTACO_BD) {
      // their tables / tiles and LDS sizing assume one standalone launch
      st = set_error(SEGTACO_ERR_UNSUPPORTED, "batch item %d: %s runs standalone only", i,
                     plan.schedule == SEGTACO_EB_CS ? "eb_cs" : "bd");
      break;
    }
    // Alignment-aware resolution (vec from the real pointers).
    {
      segtaco_csr_t Ar = A;
      st = plan_resolve_impl(&Ar, it.N, b_dtype, it.ldb, it.ldc, &plan);
      if (st != SEGTACO_OK) break;
      const int32_t V = plan.vec;
      const size_t eb = (size_t)V * elem_bytes(b_dtype);
      if ((it.B && ((uintptr_t)it.B % eb)) || ((uintptr_t)it.C % ((size_t)V * 4))) {
        st = set_error(SEGTACO_ERR_UNSUPPORTED, "batch item %d: operand alignment", i);
        break;
      }
    }
    b->plans[i] = plan;
    int64_t units = 0, fix_units = 0;
    size_t lds = 0;
    fill_batch_desc(A, it.B, it.ldb, it.C, it.ldc, it.N, plan, &D, &units,
                    &fix_units, &lds, &lay[i]);
    if (lds > b->lds) b->lds = lds;
    b->sv = b->sv || batch_uses_spmv(D);
    if (plan.schedule >= SEGTACO_EB_SR) {
      ws_off[i] = ws_total;
      ws_total += align_up(lay[i].total, 256);
    }
    D.unit_begin = b->prefix[i];
    D.fix_begin = b->fix_prefix[i];
    b->prefix[i + 1] = b->prefix[i] + units;
    b->fix_prefix[i + 1] = b->fix_prefix[i] + fix_units;
  }
  if (st != SEGTACO_OK) {
    release(b);
    return st;
  }
  b->total_units = b->prefix[n];
  b->total_fix = b->fix_prefix[n];

  auto fail = [&](hipError_t e, const char *what) {
    const int32_t s = hip_error(e, what);
    release(b);
    return s;
  };
  hipError_t e;
  if (ws_total) {
    if ((e = hipMalloc(&b->ws, ws_total)) != hipSuccess) return fail(e, "hipMalloc(batch workspace)");
    b->ws_bytes = ws_total;
  }
  for (int32_t i = 0; i < n; ++i) {
    BatchDesc &D = b->descs[i];
    if (D.schedule < SEGTACO_EB_SR || D.a.ntiles == 0) continue;
    char *w = b->ws + ws_off[i];
    D.a.part = reinterpret_cast<int32_t *>(w + lay[i].part_off);
    D.a.tile_row = reinterpret_cast<int32_t *>(w + lay[i].rows_off);
    D.a.carry_head = reinterpret_cast<float *>(w + lay[i].head_off);
    D.a.carry_tail = reinterpret_cast<float *>(w + lay[i].tail_off);
    if ((e = launch_block_starts(D.a.pos, D.a.M, D.a.tile, D.a.ntiles,
                                 const_cast<int32_t *>(D.a.part),
                                 const_cast<int32_t *>(D.a.tile_row), stream)) != hipSuccess)
      return fail(e, "launch_block_starts");
  }
  const size_t nd = sizeof(BatchDesc) * (size_t)(n > 0 ? n : 1);
  if ((e = hipMalloc(&b->d_desc, nd)) != hipSuccess) return fail(e, "hipMalloc(descs)");
  if ((e = hipMalloc(&b->d_prefix, sizeof(int64_t) * (n + 1))) != hipSuccess)
    return fail(e, "hipMalloc(prefix)");
  if ((e = hipMalloc(&b->d_fix_prefix, sizeof(int64_t) * (n + 1))) != hipSuccess)
    return fail(e, "hipMalloc(fix prefix)");
  if ((e = hipMalloc(&b->d_counter, sizeof(unsigned long long))) != hipSuccess)
    return fail(e, "hipMalloc(counter)");
  if (n > 0 && (e = hipMemcpy(b->d_desc, b->descs.data(), sizeof(BatchDesc) * n,
                              hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy(descs)");
  if ((e = hipMemcpy(b->d_prefix, b->prefix.data(), sizeof(int64_t) * (n + 1),
                     hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy(prefix)");
  if ((e = hipMemcpy(b->d_fix_prefix, b->fix_prefix.data(), sizeof(int64_t) * (n + 1),
                     hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy(fix prefix)");
  if ((e = batch_main_grid(b_dtype, b->lds, &b->grid, b->sv)) != hipSuccess)
    return fail(e, "batch_main_grid");
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
  *out = b;
  return SEGTACO_OK;
}

int32_t segtaco_batch_run(segtaco_batch_t b, void *stream_v) {
  if (!b) return set_error(SEGTACO_ERR_INVALID, "null batch");
  hipStream_t s = static_cast<hipStream_t>(stream_v);
  if (b->n == 0) return SEGTACO_OK;
  SEGTACO_HIP(launch_batch_main(b->dt, b->d_desc, b->n, b->d_prefix, b->total_units,
                                b->d_counter, b->grid, b->lds, s, /*reset_counter=*/true,
                                b->sv));
  SEGTACO_HIP(launch_batch_fixup(b->d_desc, b->n, b->d_fix_prefix, b->total_fix, s));
  return SEGTACO_OK;
}

int32_t segtaco_batch_capture(segtaco_batch_t b) {
  if (!b) return set_error(SEGTACO_ERR_INVALID, "null batch");
  if (b->exec) return SEGTACO_OK;
  if (!b->cap_stream)
    SEGTACO_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
  SEGTACO_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
  const int32_t st = segtaco_batch_run(b, b->cap_stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
  if (st != SEGTACO_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  b->graph = g;
  SEGTACO_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return SEGTACO_OK;
}

int32_t segtaco_batch_replay(segtaco_batch_t b, void *stream_v) {
  if (!b) return set_error(SEGTACO_ERR_INVALID, "null batch");
  if (!b->exec) return set_error(SEGTACO_ERR_STATE, "batch not captured");
  SEGTACO_HIP(hipGraphLaunch(b->exec, static_cast<hipStream_t>(stream_v)));
  return SEGTACO_OK;
}

int32_t segtaco_batch_plan(segtaco_batch_t b, int32_t i, segtaco_plan_t *out) {
  if (!b || !out || i < 0 || i >= b->n)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_batch_plan: bad argument");
  *out = b->plans[i];
  return SEGTACO_OK;
}

int32_t segtaco_batch_partition(segtaco_batch_t b, int32_t i, int32_t *ntiles,
                                const int32_t **part, const int32_t **tile_rows) {
  if (!b || !ntiles || !part || !tile_rows || i < 0 || i >= b->n)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_batch_partition: bad argument");
  const BatchDesc &D = b->descs[i];
  if (D.schedule < SEGTACO_EB_SR || D.a.ntiles == 0) {
    *ntiles = 0;
    *part = nullptr;
    *tile_rows = nullptr;
    return SEGTACO_OK;
  }
  *ntiles = D.a.ntiles;
  *part = D.a.part;
  *tile_rows = D.a.tile_row;
  return SEGTACO_OK;
}

int64_t segtaco_batch_units(segtaco_batch_t b) { return b ? b->total_units : -1; }

size_t segtaco_batch_workspace_bytes(segtaco_batch_t b) {
  return b ? b->ws_bytes : 0;
}

int32_t segtaco_batch_destroy(segtaco_batch_t b) {
  if (b) (void)hipDeviceSynchronize();
  release(b);
  return SEGTACO_OK;
}

}  // extern "C"

// autotune.cpp — per-matrix plan selection from row-length statistics.
//
// Model (from measurements on MI355X, where a random B-row gather past L2 is
// the cost that dominates):
//   * near-uniform row lengths (max <= 4 x mean and cv < 0.6, e.g. banded
//     or fixed-degree matrices): rb_sr — one lane group per row, sequential
//     sums, no partition or fix-up pass; short rows get several rows per
//     lane group so a workgroup still carries ~1K entries;
//   * skewed (power-law) lengths: eb_sr — equal-entry tiles keep every lane
//     group equally busy whatever the tail; tile 128 for short mean rows,
//     256 above 32 entries per row.
// The measured mode times the model's plan and a fixed candidate set
// (eb_cs among them) and returns the fastest; the cache, keyed by a
// signature of the statistics, keeps the fastest plan that batches and the
// server can run (rb_sr .. eb_pr).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "launch.h"
#include "segtaco/autotune.h"

namespace segtaco {

namespace {

std::mutex g_cache_mu;
std::unordered_map<uint64_t, segtaco_plan_t> &cache() {
  static std::unordered_map<uint64_t, segtaco_plan_t> c;
  return c;
}

uint64_t fnv1a(const int64_t *v, int n) {
  uint64_t h = 1469598103934665603ull;
  for (int i = 0; i < n; ++i) {
    uint64_t x = (uint64_t)v[i];
    for (int b = 0; b < 8; ++b) {
      h ^= (x >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  }
  return h;
}

segtaco_plan_t make(int32_t sched, int32_t vec, int32_t q, int32_t g,
                    int32_t tile) {
  segtaco_plan_t p{};
  p.schedule = sched;
  p.vec = vec;
  p.col_lanes = q;
  p.group = g;
  p.tile = tile;
  p.comp_threshold = -1;
  return p;
}

// Schedules that batches and the server run (eb_cs and bd run standalone).
bool shared_schedule(int32_t sched) {
  return sched >= SEGTACO_RB_SR && sched <= SEGTACO_EB_PR;
}

bool same_plan(const segtaco_plan_t &a, const segtaco_plan_t &b) {
  return a.schedule == b.schedule && a.vec == b.vec &&
         a.col_lanes == b.col_lanes && a.group == b.group &&
         a.tile == b.tile && a.comp_threshold == b.comp_threshold;
}

segtaco_csr_t shape_only(const segtaco_row_stats_t *st, int32_t dt) {
  segtaco_csr_t A{};
  A.M = st->M;
  A.K = st->K;
  A.nnz = st->nnz;
  A.vals_dtype = dt;
  return A;
}

}  // namespace

int32_t tune_model_impl(const segtaco_row_stats_t *st, int32_t N,
                        int32_t b_dtype, segtaco_plan_t *out) {
  if (!st || !out) return set_error(SEGTACO_ERR_INVALID, "null argument");
  segtaco_plan_t p = make(SEGTACO_RB_SR, 0, 0, 0, 0);
  const double mean = st->mean_len;
  const bool uniform =
      st->nnz == 0 || (st->max_len <= 4.0 * (mean > 1.0 ? mean : 1.0) && st->cv < 0.6);
  segtaco_csr_t A = shape_only(st, b_dtype);
  if (uniform) {
    p.schedule = SEGTACO_RB_SR;
    p.group = 1;
    SEGTACO_RET(plan_resolve_impl(&A, N, b_dtype, N, N, &p));
    // Rows per lane group so that one workgroup walks ~1K entries.
    const double rows_per_block = (double)kWavesPerBlock * (64 / p.col_lanes);
    int32_t g = 1;
    while (g < 16 && rows_per_block * g * (mean > 0.5 ? mean : 0.5) < 1024.0) g <<= 1;
    p.group = g;
  } else {
    p.schedule = SEGTACO_EB_SR;
    p.tile = mean > 32.0 ? 256 : 128;
  }
  SEGTACO_RET(plan_resolve_impl(&A, N, b_dtype, N, N, &p));
  *out = p;
  return SEGTACO_OK;
}

uint64_t tune_signature_impl(const segtaco_row_stats_t *st, int32_t N,
                             int32_t b_dtype) {
  if (!st) return 0;
  const int64_t key[8] = {
      (int64_t)ilog2(st->M > 0 ? st->M : 1),
      (int64_t)std::lround(std::log2(st->mean_len > 1e-9 ? st->mean_len : 1e-9) * 4),
      (int64_t)std::lround(st->cv * 4),
      (int64_t)std::lround(st->frac_nnz_long * 10),
      (int64_t)ilog2((int64_t)st->max_len + 1),
      (int64_t)N,
      (int64_t)b_dtype,
      7};
  return fnv1a(key, 8);
}

}  // namespace segtaco

using namespace segtaco;

extern "C" {

int32_t segtaco_row_stats(const int32_t *pos, int32_t M, int32_t K,
                          segtaco_row_stats_t *out) {
  if (!out || M < 0 || (M > 0 && !pos))
    return set_error(SEGTACO_ERR_INVALID, "segtaco_row_stats: bad argument");
  std::memset(out, 0, sizeof(*out));
  out->M = M;
  out->K = K;
  out->nnz = M > 0 ? pos[M] : 0;
  double sum = 0.0, sum2 = 0.0;
  int64_t long_nnz = 0;
  for (int32_t r = 0; r < M; ++r) {
    const int32_t L = pos[r + 1] - pos[r];
    if (L < 0) return set_error(SEGTACO_ERR_INVALID, "pos decreases at row %d", r);
    if (L > out->max_len) out->max_len = L;
    if (L == 0) ++out->empty_rows;
    if (L > 1024) long_nnz += L;
    int b = L == 0 ? 0 : ilog2(L) + 1;
    if (b >= SEGTACO_HIST_BINS) b = SEGTACO_HIST_BINS - 1;
    ++out->hist[b];
    sum += L;
    sum2 += (double)L * L;
  }
  out->mean_len = M > 0 ? sum / M : 0.0;
  const double var = M > 0 ? sum2 / M - out->mean_len * out->mean_len : 0.0;
  out->cv = out->mean_len > 0 ? std::sqrt(var > 0 ? var : 0) / out->mean_len : 0.0;
  out->frac_nnz_long = out->nnz > 0 ? (double)long_nnz / out->nnz : 0.0;
  return SEGTACO_OK;
}

int32_t segtaco_tune_model(const segtaco_row_stats_t *st, int32_t N,
                           int32_t b_dtype, segtaco_plan_t *out) {
  return tune_model_impl(st, N, b_dtype, out);
}

uint64_t segtaco_tune_signature(const segtaco_row_stats_t *st, int32_t N,
                                int32_t b_dtype) {
  return tune_signature_impl(st, N, b_dtype);
}

int32_t segtaco_tune_cache_get(uint64_t sig, segtaco_plan_t *out) {
  if (!out) return set_error(SEGTACO_ERR_INVALID, "null plan");
  std::lock_guard<std::mutex> lock(g_cache_mu);
  auto it = cache().find(sig);
  if (it == cache().end()) return SEGTACO_ERR_STATE;
  *out = it->second;
  return SEGTACO_OK;
}

int32_t segtaco_tune_cache_put(uint64_t sig, const segtaco_plan_t *plan) {
  if (!plan) return set_error(SEGTACO_ERR_INVALID, "null plan");
  // the cache feeds batches and the server, which run rb_sr .. eb_pr
  if (!shared_schedule(plan->schedule))
    return set_error(SEGTACO_ERR_UNSUPPORTED, "tune cache: schedule %d runs standalone only",
                     plan->schedule);
  std::lock_guard<std::mutex> lock(g_cache_mu);
  cache()[sig] = *plan;
  return SEGTACO_OK;
}

int32_t segtaco_tune_cache_clear(void) {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  cache().clear();
  return SEGTACO_OK;
}

int32_t segtaco_tune_cache_save(const char *path) {
  if (!path) return set_error(SEGTACO_ERR_INVALID, "null path");
  std::lock_guard<std::mutex> lock(g_cache_mu);
  FILE *f = std::fopen(path, "w");
  if (!f) return set_error(SEGTACO_ERR_INVALID, "cannot write %s", path);
  for (const auto &kv : cache()) {
    const segtaco_plan_t &p = kv.second;
    std::fprintf(f, "%016llx %d %d %d %d %d %d\n", (unsigned long long)kv.first,
                 p.schedule, p.vec, p.col_lanes, p.group, p.tile, p.comp_threshold);
  }
  std::fclose(f);
  return SEGTACO_OK;
}

int32_t segtaco_tune_cache_load(const char *path) {
  if (!path) return set_error(SEGTACO_ERR_INVALID, "null path");
  FILE *f = std::fopen(path, "r");
  if (!f) return set_error(SEGTACO_ERR_INVALID, "cannot read %s", path);
  std::lock_guard<std::mutex> lock(g_cache_mu);
  unsigned long long sig;
  segtaco_plan_t p{};
  while (std::fscanf(f, "%llx %d %d %d %d %d %d", &sig, &p.schedule, &p.vec,
                     &p.col_lanes, &p.group, &p.tile, &p.comp_threshold) == 7) {
    if (!shared_schedule(p.schedule)) continue;
    cache()[(uint64_t)sig] = p;
  }
  std::fclose(f);
  return SEGTACO_OK;
}

int32_t segtaco_tune(const segtaco_row_stats_t *st, int32_t N, int32_t b_dtype,
                     segtaco_plan_t *out) {
  if (!st || !out) return set_error(SEGTACO_ERR_INVALID, "null argument");
  if (segtaco_tune_cache_get(tune_signature_impl(st, N, b_dtype), out) ==
      SEGTACO_OK)
    return SEGTACO_OK;
  return tune_model_impl(st, N, b_dtype, out);
}

int32_t segtaco_tune_measure(const segtaco_csr_t *A,
                             const segtaco_row_stats_t *st, const void *B,
                             int32_t b_dtype, int64_t ldb, int32_t N, float *C,
                             int64_t ldc, int32_t reps, void *stream_v,
                             segtaco_plan_t *out, float *out_ms) {
  if (!A || !st || !out) return set_error(SEGTACO_ERR_INVALID, "null argument");
  if (reps < 1) reps = 1;
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  segtaco_plan_t model{};
  SEGTACO_RET(tune_model_impl(st, N, b_dtype, &model));
  std::vector<segtaco_plan_t> cands = {
      model,
      make(SEGTACO_RB_SR, 0, 0, 0, 0),
      make(SEGTACO_RB_SR, 0, 0, 4, 0),
      make(SEGTACO_RB_PR, 0, 0, 0, 0),
      make(SEGTACO_RB_PR, 0, 0, 8, 0),
      make(SEGTACO_EB_SR, 0, 0, 0, 64),
      make(SEGTACO_EB_SR, 0, 0, 0, 128),
      make(SEGTACO_EB_SR, 0, 0, 0, 256),
      make(SEGTACO_EB_PR, 0, 0, 0, 0),
      make(SEGTACO_EB_PR, 0, 0, 0, 256),
      make(SEGTACO_EB_CS, 0, 0, 0, 0),
  };
  std::vector<segtaco_plan_t> plans;
  size_t ws_need = 0;
  for (auto p : cands) {
    if (plan_resolve_impl(A, N, b_dtype, ldb, ldc, &p) != SEGTACO_OK) continue;
    bool dup = false;
    for (const auto &q : plans) dup = dup || same_plan(p, q);
    if (dup) continue;
    plans.push_back(p);
    const size_t w = spmm_workspace_bytes_impl(A, N, &p);
    if (w > ws_need) ws_need = w;
  }
  segtaco_clear_error();
  void *ws = nullptr;
  if (ws_need) SEGTACO_HIP(hipMalloc(&ws, ws_need));
  hipEvent_t e0, e1;
  SEGTACO_HIP(hipEventCreate(&e0));
  SEGTACO_HIP(hipEventCreate(&e1));
  float best = 1e30f, best_shared = 1e30f;
  segtaco_plan_t best_plan = model, shared_plan = model;
  int32_t status = SEGTACO_OK;
  for (const auto &p : plans) {
    status = spmm_impl(A, B, b_dtype, ldb, N, C, ldc, &p, ws, ws_need, stream, nullptr);
    if (status != SEGTACO_OK) break;
    if (hipEventRecord(e0, stream) != hipSuccess) { status = SEGTACO_ERR_HIP; break; }
    for (int32_t r = 0; r < reps && status == SEGTACO_OK; ++r)
      status = spmm_impl(A, B, b_dtype, ldb, N, C, ldc, &p, ws, ws_need, stream, nullptr);
    if (status != SEGTACO_OK) break;
    if (hipEventRecord(e1, stream) != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess) {
      status = SEGTACO_ERR_HIP;
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    ms /= reps;
    if (ms < best) {
      best = ms;
      best_plan = p;
    }
    if (shared_schedule(p.schedule) && ms < best_shared) {
      best_shared = ms;
      shared_plan = p;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ws) (void)hipFree(ws);
  if (status != SEGTACO_OK) return status;
  // Leave C holding the winner's result.
  if (ws_need) SEGTACO_HIP(hipMalloc(&ws, ws_need));
  status = spmm_impl(A, B, b_dtype, ldb, N, C, ldc, &best_plan, ws, ws_need,
                     stream, nullptr);
  (void)hipStreamSynchronize(stream);
  if (ws) (void)hipFree(ws);
  if (status != SEGTACO_OK) return status;
  *out = best_plan;
  if (out_ms) *out_ms = best;
  segtaco_tune_cache_put(tune_signature_impl(st, N, b_dtype), &shared_plan);
  return SEGTACO_OK;
}

}  // extern "C"

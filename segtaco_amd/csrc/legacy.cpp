// legacy.cpp — the segTACO operator surface over the device-resident engine.
//
// Each entry point keeps the reference's call contract (gpu_library.h:474-535,
// 604-612, 818-857): host operands in, C written in place, the device work
// timed with a pair of events for get_gpu_timer_result(). It differs where the
// reference is a liability on this hardware:
//   * device buffers come from a grow-only pool reused across calls instead of
//     five hipMalloc/hipFree pairs per call;
//   * C is not copied host->device: every schedule writes all of C;
//   * the CSR structure is checked on the host before any kernel runs
//     (pos monotone, crd in [0, K)), so a malformed operand is an error
//     rather than an out-of-bounds device access;
//   * errors exit (as gpuErrchk does, gpu_library.h:22-30) only in the default
//     error mode; SEGTACO_ERRMODE_RECORD returns with segtaco_last_error set;
//   * large operands travel through a ring of pinned buffers (staging.h)
//     instead of pageable copies, about twice the PCIe rate.
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "engine.h"
#include "legacy_table.h"
#include "staging.h"

namespace segtaco {

int32_t vendor_spmm(const segtaco_csr_t *A, const float *B, int64_t ldb,
                    int32_t N, float *C, int64_t ldc, int32_t alg,
                    hipStream_t stream, hipEvent_t ev0, hipEvent_t ev1);

namespace {

struct Preset {
  int32_t schedule, vec_cap, col_lanes, group, tile;
};

const Preset kPresets[SEGTACO_LEGACY_COUNT] = {
#define SEGTACO_LEGACY_PRESET(name, s, v, q, g, t) {s, v, q, g, t},
    SEGTACO_LEGACY_TABLE(SEGTACO_LEGACY_PRESET)
#undef SEGTACO_LEGACY_PRESET
};

const char *const kNames[SEGTACO_LEGACY_COUNT] = {
#define SEGTACO_LEGACY_NAME(name, s, v, q, g, t) #name,
    SEGTACO_LEGACY_TABLE(SEGTACO_LEGACY_NAME)
#undef SEGTACO_LEGACY_NAME
};

struct LegacyState {
  std::mutex mu;
  bool ready = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  char *pool = nullptr;
  size_t pool_bytes = 0;
  Stager stager;  // used under mu only
};

LegacyState &state() {
  static LegacyState s;
  return s;
}

float g_timer_ms = -1.0f;
int32_t g_error_mode = SEGTACO_ERRMODE_EXIT;

int32_t ensure_ready(LegacyState &s) {
  if (s.ready) return SEGTACO_OK;
  SEGTACO_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
  SEGTACO_HIP(hipEventCreate(&s.ev0));
  SEGTACO_HIP(hipEventCreate(&s.ev1));
  s.ready = true;
  return SEGTACO_OK;
}

int32_t ensure_pool(LegacyState &s, size_t bytes) {
  if (bytes <= s.pool_bytes) return SEGTACO_OK;
  if (s.pool) {
    SEGTACO_HIP(hipStreamSynchronize(s.stream));
    SEGTACO_HIP(hipFree(s.pool));
    s.pool = nullptr;
    s.pool_bytes = 0;
  }
  const size_t want = align_up(bytes + bytes / 8, 1 << 20);
  SEGTACO_HIP(hipMalloc(&s.pool, want));
  s.pool_bytes = want;
  return SEGTACO_OK;
}

int32_t run_legacy(int id, taco_tensor_t *C, taco_tensor_t *A,
                   taco_tensor_t *B) {
  const char *name = kNames[id];
  if (!C || !A || !B)
    return set_error(SEGTACO_ERR_INVALID, "%s: null tensor", name);
  if (!C->dimensions || !B->dimensions || !A->indices || !A->indices[1] ||
      !A->indices[1][0])
    return set_error(SEGTACO_ERR_INVALID, "%s: tensor fields missing", name);
  const int32_t M = C->dimensions[0];
  const int32_t N = C->dimensions[1];
  const int32_t K = B->dimensions[0];
  if (M < 0 || N < 0 || K < 0)
    return set_error(SEGTACO_ERR_INVALID, "%s: negative dimension", name);
  if (A->dimensions && (A->dimensions[0] != M || A->dimensions[1] != K))
    return set_error(SEGTACO_ERR_INVALID,
                     "%s: A is %dx%d but C has %d rows and B has %d rows",
                     name, A->dimensions[0], A->dimensions[1], M, K);
  if (B->dimensions[1] != N)
    return set_error(SEGTACO_ERR_INVALID, "%s: B has %d columns, C has %d",
                     name, B->dimensions[1], N);
  const int32_t *pos = A->indices[1][0];
  const int32_t *crd = A->indices[1][1];
  const int32_t nnz = M > 0 ? pos[M] : 0;
  SEGTACO_RET(segtaco_csr_validate_host(M, K, nnz, pos, crd));
  if (M == 0 || N == 0) {
    g_timer_ms = 0.0f;
    return SEGTACO_OK;
  }
  if (!C->vals || (nnz > 0 && !B->vals))
    return set_error(SEGTACO_ERR_INVALID, "%s: null vals", name);

  const Preset &pr = kPresets[id];
  const bool vendor = pr.schedule == SEGTACO_VENDOR_ROW;
  segtaco_plan_t plan{};
  plan.schedule = pr.schedule;
  plan.col_lanes = pr.col_lanes;
  plan.group = pr.group;
  plan.tile = pr.tile;
  plan.comp_threshold = -1;
  // vec is a cap: the widest width <= cap that N allows.
  {
    int32_t v = pr.vec_cap ? pr.vec_cap : 4;
    while (v > 1 && N % v) v >>= 1;
    plan.vec = v;
  }
  segtaco_csr_t dA{M, K, nnz, SEGTACO_DT_F32, nullptr, nullptr, nullptr};
  if (!vendor) SEGTACO_RET(plan_resolve_impl(&dA, N, SEGTACO_DT_F32, N, N, &plan));
  if (vendor && nnz > 0 && !A->vals)
    return set_error(SEGTACO_ERR_UNSUPPORTED, "%s: the vendor row needs values", name);

  LegacyState &s = state();
  std::lock_guard<std::mutex> lock(s.mu);
  SEGTACO_RET(ensure_ready(s));
  const size_t b_pos = sizeof(int32_t) * ((size_t)M + 1);
  const size_t b_crd = sizeof(int32_t) * (size_t)nnz;
  const size_t b_val = A->vals ? sizeof(float) * (size_t)nnz : 0;
  const size_t b_B = sizeof(float) * (size_t)K * N;
  const size_t b_C = sizeof(float) * (size_t)M * N;
  const size_t b_ws = vendor ? 0 : spmm_workspace_bytes_impl(&dA, N, &plan);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  const size_t o_pos = take(b_pos), o_crd = take(b_crd), o_val = take(b_val),
               o_B = take(b_B), o_C = take(b_C), o_ws = take(b_ws);
  SEGTACO_RET(ensure_pool(s, off));
  char *base = s.pool;
  SEGTACO_RET(s.stager.h2d(base + o_pos, pos, b_pos, s.stream));
  if (b_crd) SEGTACO_RET(s.stager.h2d(base + o_crd, crd, b_crd, s.stream));
  if (b_val) SEGTACO_RET(s.stager.h2d(base + o_val, A->vals, b_val, s.stream));
  if (b_B && B->vals) SEGTACO_RET(s.stager.h2d(base + o_B, B->vals, b_B, s.stream));
  dA.pos = reinterpret_cast<const int32_t *>(base + o_pos);
  dA.crd = reinterpret_cast<const int32_t *>(base + o_crd);
  dA.vals = b_val ? base + o_val : nullptr;

  if (vendor) {
    // the vendor row times rocSPARSE's compute stage only
    SEGTACO_RET(vendor_spmm(&dA, reinterpret_cast<const float *>(base + o_B), N, N,
                            reinterpret_cast<float *>(base + o_C), N, 0, s.stream,
                            s.ev0, s.ev1));
  } else {
    SEGTACO_HIP(hipEventRecord(s.ev0, s.stream));
    SEGTACO_RET(spmm_impl(&dA, base + o_B, SEGTACO_DT_F32, N, N,
                          reinterpret_cast<float *>(base + o_C), N, &plan,
                          base + o_ws, b_ws, s.stream, nullptr));
    SEGTACO_HIP(hipEventRecord(s.ev1, s.stream));
  }
  SEGTACO_HIP(hipEventSynchronize(s.ev1));
  float ms = -1.0f;
  SEGTACO_HIP(hipEventElapsedTime(&ms, s.ev0, s.ev1));
  g_timer_ms = ms;
  SEGTACO_RET(s.stager.d2h(C->vals, base + o_C, b_C, s.stream));
  return SEGTACO_OK;
}

}  // namespace
}  // namespace segtaco

using namespace segtaco;

extern "C" {

int32_t segtaco_set_error_mode(int32_t mode) {
  const int32_t prev = g_error_mode;
  g_error_mode = mode == SEGTACO_ERRMODE_RECORD ? SEGTACO_ERRMODE_RECORD
                                                : SEGTACO_ERRMODE_EXIT;
  return prev;
}

float get_gpu_timer_result(void) { return g_timer_ms; }

// Dispatch by table index; also the target of the C++-linkage aliases.
void segtaco_legacy_call(int32_t id, taco_tensor_t *C, taco_tensor_t *A,
                         taco_tensor_t *B) {
  if (id < 0 || id >= SEGTACO_LEGACY_COUNT) {
    set_error(SEGTACO_ERR_INVALID, "unknown operator id %d", id);
  } else {
    const int32_t st = run_legacy(id, C, A, B);
    if (st == SEGTACO_OK) return;
  }
  if (g_error_mode == SEGTACO_ERRMODE_EXIT) {
    fprintf(stderr, "segtaco: %s\n", segtaco_last_error_string());
    exit(segtaco_last_error() ? segtaco_last_error() : 1);
  }
}

// Number of operator-surface entry points and their names (for drivers and
// sweeps that iterate over the surface).
int32_t segtaco_legacy_count(void) { return SEGTACO_LEGACY_COUNT; }
const char *segtaco_legacy_name(int32_t id) {
  return (id >= 0 && id < SEGTACO_LEGACY_COUNT) ? kNames[id] : nullptr;
}

#define SEGTACO_LEGACY_DEFINE(name, s, v, q, g, t)                   \
  void name(taco_tensor_t *C, taco_tensor_t *A, taco_tensor_t *B) { \
    segtaco_legacy_call(SEGTACO_LEGACY_##name, C, A, B);             \
  }
SEGTACO_LEGACY_TABLE(SEGTACO_LEGACY_DEFINE)
#undef SEGTACO_LEGACY_DEFINE

}  // extern "C"

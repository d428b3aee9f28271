// engine.cpp — the device-resident SpMM API: plan resolution, workspace
// layout, validation and kernel sequencing for every schedule (the
// reference's four, eb_cs and bd).
#include "engine.h"

#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "bd.h"
#include "cs.h"
#include "launch.h"
#include "spmv.h"

namespace segtaco {

namespace {
thread_local int32_t t_last_status = SEGTACO_OK;
thread_local char t_last_msg[512] = "";
}  // namespace

int32_t set_error(int32_t status, const char *fmt, ...) {
  t_last_status = status;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_last_msg, sizeof(t_last_msg), fmt, ap);
  va_end(ap);
  return status;
}

int32_t hip_error(hipError_t e, const char *what) {
  const int32_t st =
      e == hipErrorOutOfMemory ? SEGTACO_ERR_NOMEM : SEGTACO_ERR_HIP;
  return set_error(st, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

EbLayout eb_layout(int32_t nnz, int32_t N, int32_t tile) {
  EbLayout L{};
  L.ntiles = (int32_t)ceil_div(nnz, tile);
  const size_t part = sizeof(int32_t) * ((size_t)L.ntiles + 1);
  const size_t carry = sizeof(float) * (size_t)L.ntiles * (size_t)N;
  L.part_off = 0;
  L.rows_off = align_up(part, 256);
  L.head_off = L.rows_off + align_up(part, 256);
  L.tail_off = L.head_off + align_up(carry, 256);
  L.total = L.tail_off + align_up(carry, 256);
  return L;
}

// Workspace of an eb_cs plan: the eb layout (partition, tile rows, carries)
// followed by the sorted copy of A (one record per entry: column, slot and
// value, 64 bits when every column is below 2^20, else 96) and the unit
// tables the preparation kernels write.
struct CsLayout {
  EbLayout eb;
  int32_t units, cap, packed;
  size_t rec_off, rows_off, int_off, code_off, nslot_off, count_off, flag_off, total;
};

size_t device_max_lds();

CsLayout cs_layout(int32_t nnz, int32_t K, int32_t N, int32_t tile, int32_t group,
                   int32_t vec) {
  CsLayout L{};
  L.eb = eb_layout(nnz, N, tile);
  L.units = (int32_t)ceil_div(L.eb.ntiles, group);
  // a slot: kCsLanes * vec FP32 accumulators, and its row id beside them
  const size_t slot = (size_t)kCsLanes * vec * sizeof(float);
  size_t cap = kCsLdsBudget / slot;
  if (cap > device_max_lds() / (slot + sizeof(int32_t)))
    cap = device_max_lds() / (slot + sizeof(int32_t));
  const size_t max_code = (size_t)1 << kCsCodeBits;  // slots fit the code field
  L.cap = (int32_t)(cap < max_code ? cap : max_code);
  L.packed = K <= (1 << (32 - kCsCodeBits)) ? 1 : 0;
  size_t off = L.eb.total;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off = align_up(off + bytes, 256);
    return o;
  };
  L.rec_off = take((L.packed ? 8 : 12) * (size_t)nnz);
  L.rows_off = take(sizeof(int32_t) * (size_t)L.units * L.cap);
  L.int_off = take(sizeof(int32_t) * (size_t)L.eb.ntiles * tile);
  L.code_off = take(sizeof(uint16_t) * (size_t)nnz);
  L.nslot_off = take(sizeof(int32_t) * (size_t)L.units);
  L.count_off = take(sizeof(int32_t) * (size_t)L.eb.ntiles);
  L.flag_off = take((size_t)L.eb.ntiles);
  L.total = off;
  return L;
}

// LDS bytes one workgroup may use on the current device (160 KiB on gfx950;
// also the answer when no device is visible). Queried once per process.
size_t device_max_lds() {
  static const size_t cached = [] {
    int dev = 0, n = 0;
    size_t v = 160 * 1024;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
            hipSuccess &&
        n > 0)
      v = (size_t)n;
    (void)hipGetLastError();
    return v;
  }();
  return cached;
}

// Compute units of the current device (256 on MI355X; also the answer when
// no device is visible, so plans resolve the same on a host without a GPU).
// Queried once per process.
int32_t device_cus() {
  static const int32_t cached = [] {
    int dev = 0, n = 0;
    int32_t v = 256;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) ==
            hipSuccess &&
        n > 0)
      v = n;
    (void)hipGetLastError();
    return v;
  }();
  return cached;
}

int32_t latency_tile(int64_t nnz, int32_t tile, int32_t tiles_per_block) {
  const int64_t want = 2 * (int64_t)device_cus() * (tiles_per_block > 0 ? tiles_per_block : 1);
  while (tile > 8 && ceil_div(nnz, (int64_t)tile) < want) tile >>= 1;
  return tile;
}

namespace {

bool aligned(const void *p, size_t a) {
  return ((uintptr_t)p % a) == 0;
}

// Widest per-lane vector (in elements) that N, the leading dimensions and the
// base pointers allow.
int32_t widest_vec(int32_t dt, int32_t N, int64_t ldb, int64_t ldc,
                   const void *B, const void *C) {
  const int32_t maxv = dt == SEGTACO_DT_F32 ? 4 : 8;
  for (int32_t v = maxv; v > 1; v >>= 1) {
    if (N % v || ldb % v || ldc % v) continue;
    if (B && !aligned(B, (size_t)v * elem_bytes(dt))) continue;
    if (C && !aligned(C, (size_t)v * 4)) continue;
    return v;
  }
  return 1;
}

int32_t natural_lanes(int32_t N, int32_t vec) {
  int64_t q = pow2_ceil(ceil_div(N, vec));
  return (int32_t)(q > 64 ? 64 : q);
}

int32_t resolve(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                int64_t ldb, int64_t ldc, const void *B, const void *C,
                segtaco_plan_t *plan) {
  if (!A || !plan) return set_error(SEGTACO_ERR_INVALID, "null argument");
  if (A->nnz > kMaxNnz)
    return set_error(SEGTACO_ERR_UNSUPPORTED, "nnz %d above the supported %d", A->nnz,
                     kMaxNnz);
  if (plan->schedule < SEGTACO_RB_SR || plan->schedule > SEGTACO_BD)
    return set_error(SEGTACO_ERR_INVALID, "unknown schedule %d",
                     plan->schedule);
  if (b_dtype < SEGTACO_DT_F32 || b_dtype > SEGTACO_DT_F16)
    return set_error(SEGTACO_ERR_INVALID, "unknown B dtype %d", b_dtype);
  if (A->vals && A->vals_dtype != b_dtype)
    return set_error(SEGTACO_ERR_UNSUPPORTED,
                     "vals dtype %d differs from B dtype %d", A->vals_dtype,
                     b_dtype);
  if (N < 1) N = 1;
  const int32_t maxv = b_dtype == SEGTACO_DT_F32 ? 4 : 8;
  if (plan->vec == 0) {
    plan->vec = widest_vec(b_dtype, N, ldb, ldc, B, C);
    // eb_cs keeps W = 8 * vec floats per row in LDS: at most 4 per lane
    if (plan->schedule == SEGTACO_EB_CS && plan->vec > 4) plan->vec = 4;
  } else {
    const int32_t v = plan->vec;
    if (!is_pow2(v) || v > maxv)
      return set_error(SEGTACO_ERR_INVALID, "vec %d not in 1..%d", v, maxv);
    if (N % v || ldb % v || ldc % v ||
        (B && !aligned(B, (size_t)v * elem_bytes(b_dtype))) ||
        (C && !aligned(C, (size_t)v * 4)))
      return set_error(SEGTACO_ERR_UNSUPPORTED,
                       "vec %d does not divide N/ld or align the operands", v);
  }
  const int32_t V = plan->vec;
  const int32_t qn = natural_lanes(N, V);
  switch (plan->schedule) {
    case SEGTACO_RB_SR:
      if (plan->col_lanes == 0) plan->col_lanes = qn;
      if (plan->group == 0) plan->group = 1;
      break;
    case SEGTACO_RB_PR:
      // at most 16 lanes per column panel, so a row gets >= 4 entry slots
      // (power-law 262144^2, N = 128: 16 x 4 runs in 432 us, 32 x 2 in 548)
      if (plan->col_lanes == 0) {
        int32_t g = plan->group ? plan->group : 1;
        int32_t q = qn < 16 ? qn : 16;
        while (q * g > 64 && q > 1) q >>= 1;
        plan->col_lanes = q;
      }
      if (plan->group == 0) plan->group = 64 / plan->col_lanes;
      break;
    case SEGTACO_EB_SR:
      if (plan->col_lanes == 0) plan->col_lanes = qn;
      // 128-entry tiles, fewer on small products (BF16 1000^2 with 15.6 K
      // entries at N = 16: 70 us per call at 128, one workgroup)
      if (plan->tile == 0 && is_pow2(plan->col_lanes) && plan->col_lanes <= 64)
        plan->tile = latency_tile(A->nnz, 128, kWavesPerBlock * (64 / plan->col_lanes));
      if (plan->group == 0) plan->group = 1;
      break;
    case SEGTACO_EB_PR:
      // up to 32 lanes per panel and 128-entry tiles (same workload: 32 x 2
      // slots with tile 128 runs in 464 us, 8 x 8 with tile 256 in 693)
      if (plan->col_lanes == 0 && plan->group != 0)
        plan->col_lanes = 64 / plan->group;
      if (plan->col_lanes == 0) plan->col_lanes = qn < 32 ? qn : 32;
      if (plan->group == 0) plan->group = 64 / plan->col_lanes;
      // tiles of 128 entries; at one column (SpMV, k_spmv) 256, four
      // consecutive entries per lane read with one 128-bit load each for
      // crd and vals (power-law 262144^2: 29.6 us per call vs 37.9 / 30.9 /
      // 31.2 / 35.1 us with 1 / 2 / 8 / 16 per lane)
      if (plan->tile == 0) plan->tile = N == 1 ? 256 : 128;
      break;
    case SEGTACO_EB_CS: {
      if (plan->col_lanes == 0) plan->col_lanes = kCsLanes;
      if (plan->group == 0) plan->group = kCsMaxGroup;
      if (plan->tile == 0) {
        // one unit per CU: nnz / (CUs * tiles per unit), rounded up
        int64_t t = ceil_div((int64_t)A->nnz, (int64_t)device_cus() * plan->group);
        if (t < 32) t = 32;
        if (t > kCsMaxTile) t = kCsMaxTile;
        plan->tile = (int32_t)t;
      }
      if (plan->col_lanes != kCsLanes)
        return set_error(SEGTACO_ERR_INVALID, "eb_cs: col_lanes must be %d", kCsLanes);
      if (plan->group > kCsMaxGroup)
        return set_error(SEGTACO_ERR_INVALID, "eb_cs: group %d > %d", plan->group,
                         kCsMaxGroup);
      if (plan->vec > 4)
        return set_error(SEGTACO_ERR_UNSUPPORTED, "eb_cs: vec %d > 4", plan->vec);
      if (plan->tile > kCsMaxTile)
        return set_error(SEGTACO_ERR_INVALID, "eb_cs: tile %d > %d", plan->tile, kCsMaxTile);
      if (plan->comp_threshold == -1) plan->comp_threshold = 0;
      if (plan->comp_threshold != 0)
        return set_error(SEGTACO_ERR_UNSUPPORTED,
                         "eb_cs: compensated summation is not supported");
      break;
    }
    case SEGTACO_BD:
      // blocks of 32 rows, or 64 (two MFMA row tiles sharing each staged B
      // chunk: 7 % less B traffic on the banded 1M^2 BF16 workload at
      // N = 256, but fewer workgroups per CU and 590 us per call against 500),
      // on the matrix cores; lanes and groups are fixed
      if (b_dtype == SEGTACO_DT_F32)
        return set_error(SEGTACO_ERR_UNSUPPORTED, "bd: BF16 / FP16 storage only");
      if (plan->tile == 0) plan->tile = kBdRows;
      if (plan->tile != kBdRows && plan->tile != 2 * kBdRows)
        return set_error(SEGTACO_ERR_INVALID, "bd: tile (rows per block) must be %d or %d",
                         kBdRows, 2 * kBdRows);
      if (plan->col_lanes == 0) plan->col_lanes = 1;
      if (plan->group == 0) plan->group = 1;
      if (plan->col_lanes != 1 || plan->group != 1)
        return set_error(SEGTACO_ERR_INVALID, "bd: col_lanes and group must be 1");
      break;
  }
  const int32_t Q = plan->col_lanes, G = plan->group;
  if (!is_pow2(Q) || Q > 64)
    return set_error(SEGTACO_ERR_INVALID, "col_lanes %d not a power of two <= 64", Q);
  if (!is_pow2(G) || G > 64)
    return set_error(SEGTACO_ERR_INVALID, "group %d not a power of two <= 64", G);
  if (plan->schedule == SEGTACO_RB_PR && Q * G > 64)
    return set_error(SEGTACO_ERR_INVALID, "rb_pr: col_lanes*group = %d > 64", Q * G);
  if (plan->schedule == SEGTACO_EB_PR && Q * G != 64)
    return set_error(SEGTACO_ERR_INVALID, "eb_pr: col_lanes*group = %d != 64", Q * G);
  if (plan->schedule == SEGTACO_RB_SR && G > 16)
    return set_error(SEGTACO_ERR_INVALID, "rb_sr: rows per group %d > 16", G);
  if (plan->schedule >= SEGTACO_EB_SR &&
      (plan->tile < 8 || plan->tile > (1 << 22)))
    return set_error(SEGTACO_ERR_INVALID, "tile %d not in [8, 2^22]", plan->tile);
  if (plan->comp_threshold == -1)
    plan->comp_threshold = b_dtype == SEGTACO_DT_F32 ? 0 : 1024;
  if (plan->comp_threshold < 0)
    return set_error(SEGTACO_ERR_INVALID, "comp_threshold %d", plan->comp_threshold);
  return SEGTACO_OK;
}

int32_t check_csr(const segtaco_csr_t *A) {
  if (!A) return set_error(SEGTACO_ERR_INVALID, "null CSR");
  if (A->M < 0 || A->K < 0 || A->nnz < 0)
    return set_error(SEGTACO_ERR_INVALID, "negative CSR size");
  if (A->nnz > kMaxNnz)
    return set_error(SEGTACO_ERR_UNSUPPORTED, "nnz %d above the supported %d", A->nnz,
                     kMaxNnz);
  if (A->M > 0 && !A->pos) return set_error(SEGTACO_ERR_INVALID, "null pos");
  if (A->nnz > 0 && !A->crd) return set_error(SEGTACO_ERR_INVALID, "null crd");
  return SEGTACO_OK;
}

}  // namespace

namespace {

// Workspace of a plan whose tile, group and vec are all set.
size_t resolved_workspace_bytes(const segtaco_csr_t *A, int32_t N,
                                const segtaco_plan_t &p, int32_t b_dtype) {
  if (p.schedule == SEGTACO_EB_CS)
    return cs_layout(A->nnz, A->K, N, p.tile, p.group, p.vec).total;
  if (p.schedule != SEGTACO_EB_SR && p.schedule != SEGTACO_EB_PR) return 0;
  return eb_layout(A->nnz, N, p.tile).total;
}

}  // namespace

// A resolved plan gives the exact bytes. For a plan with open fields the
// answer covers every resolution a later call can reach: the plan is
// resolved as segtaco_plan_resolve does (no pointers, ld = N), for each
// storage type the call may use (the values' type, or FP32 and BF16 without
// values — a wider vector means fewer lanes per row and smaller default
// tiles, so more tiles), with eb_cs's open vector width taken as 1 (the most
// rows per unit); the largest result is returned (0 when none resolves).
size_t spmm_workspace_bytes_impl(const segtaco_csr_t *A, int32_t N,
                                 const segtaco_plan_t *plan) {
  if (!A || !plan || A->nnz == 0 || A->M == 0 || N <= 0) return 0;
  if (plan->schedule < SEGTACO_EB_SR || plan->schedule > SEGTACO_EB_CS) return 0;
  const bool open = plan->tile == 0 || plan->group == 0 || plan->vec == 0 ||
                    plan->col_lanes == 0 || plan->comp_threshold == -1;
  const int32_t own_dt = A->vals ? A->vals_dtype : SEGTACO_DT_F32;
  if (!open) return resolved_workspace_bytes(A, N, *plan, own_dt);
  const int32_t dts[2] = {own_dt, SEGTACO_DT_BF16};
  const int ndt = A->vals ? 1 : 2;
  size_t best = 0;
  for (int k = 0; k < ndt; ++k) {
    segtaco_plan_t p = *plan;
    if (p.schedule == SEGTACO_EB_CS && p.vec == 0) p.vec = 1;
    if (resolve(A, N, dts[k], N, N, nullptr, nullptr, &p) != SEGTACO_OK) continue;
    const size_t b = resolved_workspace_bytes(A, N, p, dts[k]);
    if (b > best) best = b;
  }
  segtaco_clear_error();
  return best;
}

int32_t spmm_workspace_layout_impl(const segtaco_csr_t *A, int32_t N,
                                   const segtaco_plan_t *plan, int64_t *out) {
  if (!A || !plan || !out)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_spmm_workspace_layout: null argument");
  if (plan->schedule < SEGTACO_EB_SR || plan->schedule > SEGTACO_EB_CS)
    return set_error(SEGTACO_ERR_UNSUPPORTED,
                     "segtaco_spmm_workspace_layout: schedule %d has no workspace",
                     plan->schedule);
  if (plan->tile < 8)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_spmm_workspace_layout: unresolved plan");
  const EbLayout L = eb_layout(A->nnz, N, plan->tile);
  out[0] = L.ntiles;
  out[1] = (int64_t)L.part_off;
  out[2] = (int64_t)L.rows_off;
  out[3] = (int64_t)L.head_off;
  out[4] = (int64_t)L.tail_off;
  return SEGTACO_OK;
}

int32_t cs_workspace_layout_impl(const segtaco_csr_t *A, int32_t N,
                                 const segtaco_plan_t *plan, int64_t *out) {
  if (!A || !plan || !out)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_cs_workspace_layout: null argument");
  if (plan->schedule != SEGTACO_EB_CS)
    return set_error(SEGTACO_ERR_UNSUPPORTED,
                     "segtaco_cs_workspace_layout: schedule %d is not eb_cs", plan->schedule);
  if (plan->tile < 8 || plan->vec < 1 || plan->group < 1)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_cs_workspace_layout: unresolved plan");
  const CsLayout L = cs_layout(A->nnz, A->K, N, plan->tile, plan->group, plan->vec);
  const int64_t v[10] = {L.units, L.cap, L.packed,
                         (int64_t)L.rec_off, (int64_t)L.rows_off, (int64_t)L.int_off,
                         (int64_t)L.code_off, (int64_t)L.nslot_off, (int64_t)L.count_off,
                         (int64_t)L.flag_off};
  for (int i = 0; i < 10; ++i) out[i] = v[i];
  return SEGTACO_OK;
}

namespace {

// What an eb_cs workspace holds: the product whose preparation (slot plan
// and sorted records) was last written there. Every call that prepares
// into a workspace records it (or forgets the workspace, for schedules whose
// run repeats the preparation anyway); a run-only call checks it, so that a
// workspace prepared for another matrix, width or plan is refused instead of
// read. The check compares device pointers and sizes only: that A's arrays
// still hold the same entries is the caller's promise.
struct PrepKey {
  const void *pos, *crd, *vals;
  int32_t M, K, nnz, b_dtype, N;
  segtaco_plan_t plan;
  bool operator==(const PrepKey &o) const {
    return pos == o.pos && crd == o.crd && vals == o.vals && M == o.M && K == o.K &&
           nnz == o.nnz && b_dtype == o.b_dtype && N == o.N &&
           std::memcmp(&plan, &o.plan, sizeof(plan)) == 0;
  }
};

std::mutex g_prep_mu;
std::unordered_map<const void *, PrepKey> g_prepared;
constexpr size_t kMaxPrepared = 4096;  // forgotten beyond: run-only calls then fail

PrepKey prep_key(const segtaco_csr_t *A, int32_t b_dtype, int32_t N,
                 const segtaco_plan_t &plan) {
  PrepKey k;
  std::memset(&k, 0, sizeof(k));
  k.pos = A->pos;
  k.crd = A->crd;
  k.vals = A->vals;
  k.M = A->M;
  k.K = A->K;
  k.nnz = A->nnz;
  k.b_dtype = b_dtype;
  k.N = N;
  k.plan = plan;
  return k;
}

void prep_record(const void *ws, const PrepKey *k) {
  if (!ws) return;
  std::lock_guard<std::mutex> g(g_prep_mu);
  if (!k) {
    g_prepared.erase(ws);
    return;
  }
  if (g_prepared.size() >= kMaxPrepared && !g_prepared.count(ws)) g_prepared.clear();
  g_prepared[ws] = *k;
}

bool prep_matches(const void *ws, const PrepKey &k) {
  std::lock_guard<std::mutex> g(g_prep_mu);
  const auto it = g_prepared.find(ws);
  return it != g_prepared.end() && it->second == k;
}

}  // namespace

// ev (optional, 4 events): recorded before the partition, after it, after
// the main kernel and after the fix-up (row-balanced plans have neither
// partition nor fix-up; their events bracket nothing).
int32_t spmm_impl(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                  int64_t ldb, int32_t N, float *C, int64_t ldc,
                  const segtaco_plan_t *plan_in, void *ws, size_t ws_bytes,
                  hipStream_t stream, hipEvent_t *ev, int32_t phases) {
  SEGTACO_RET(check_csr(A));
  if (!plan_in) return set_error(SEGTACO_ERR_INVALID, "null plan");
  if (phases < kPhasePrepare || phases > kPhaseAll)
    return set_error(SEGTACO_ERR_INVALID, "phases %d", phases);
  const bool prepare = (phases & kPhasePrepare) != 0;
  const bool run = (phases & kPhaseRun) != 0;
  if (N < 0) return set_error(SEGTACO_ERR_INVALID, "N < 0");
  if (A->M == 0 || N == 0) return SEGTACO_OK;
  if (run && !C) return set_error(SEGTACO_ERR_INVALID, "null C");
  if (ldc < N) return set_error(SEGTACO_ERR_INVALID, "ldc < N");
  if (A->nnz == 0) {
    if (run) SEGTACO_HIP(launch_fill_zero(C, A->M, N, ldc, stream));
    return SEGTACO_OK;
  }
  if (run && !B) return set_error(SEGTACO_ERR_INVALID, "null B");
  if (ldb < N) return set_error(SEGTACO_ERR_INVALID, "ldb < N");
  segtaco_plan_t plan = *plan_in;
  SEGTACO_RET(resolve(A, N, b_dtype, ldb, ldc, run ? B : nullptr, run ? C : nullptr, &plan));
  if (!run) {
    B = nullptr;
    C = nullptr;
  }
  const PrepKey key = prep_key(A, b_dtype, N, plan);
  if (!prepare && plan.schedule == SEGTACO_EB_CS && !prep_matches(ws, key))
    return set_error(SEGTACO_ERR_INVALID,
                     "segtaco_spmm_prepared: the workspace was not prepared for this "
                     "product (matrix, width, storage type and resolved plan)");

  KArgs a{};
  a.M = A->M;
  a.K = A->K;
  a.N = N;
  a.nnz = A->nnz;
  a.pos = A->pos;
  a.crd = A->crd;
  a.vals = A->vals;
  a.B = B;
  a.ldb = ldb;
  a.b_off32 = b_off32_ok(A->K, ldb, elem_bytes(b_dtype));
  a.C = C;
  a.ldc = ldc;
  a.lq = ilog2(plan.col_lanes);
  a.lg = ilog2(plan.group);
  a.tile = plan.tile;
  a.comp_threshold = plan.comp_threshold == 0 ? INT_MAX : plan.comp_threshold;
  const int32_t V = plan.vec;
  a.panels = (int32_t)ceil_div(N, (int64_t)plan.col_lanes * V);

  auto mark = [&](int i) -> hipError_t {
    return ev ? hipEventRecord(ev[i], stream) : hipSuccess;
  };
  switch (plan.schedule) {
    case SEGTACO_RB_SR: {
      if (!run) return SEGTACO_OK;  // nothing depends on A alone
      const int64_t rows_per_block =
          (int64_t)kWavesPerBlock * (64 / plan.col_lanes) * plan.group;
      a.nblocks = (int32_t)ceil_div(A->M, rows_per_block);
      SEGTACO_HIP(mark(0));
      SEGTACO_HIP(mark(1));
      SEGTACO_HIP(launch_rb_sr(b_dtype, V, a, stream));
      SEGTACO_HIP(mark(2));
      SEGTACO_HIP(mark(3));
      return SEGTACO_OK;
    }
    case SEGTACO_RB_PR: {
      if (!run) return SEGTACO_OK;  // nothing depends on A alone
      const int64_t rows_per_wave = 64 / (plan.col_lanes * plan.group);
      a.nblocks = (int32_t)ceil_div(A->M, rows_per_wave * kWavesPerBlock);
      SEGTACO_HIP(mark(0));
      SEGTACO_HIP(mark(1));
      SEGTACO_HIP(launch_rb_pr(b_dtype, V, a, stream));
      SEGTACO_HIP(mark(2));
      SEGTACO_HIP(mark(3));
      return SEGTACO_OK;
    }
    case SEGTACO_BD: {
      if (!run) return SEGTACO_OK;  // nothing depends on A alone
      a.nblocks = (int32_t)ceil_div(A->M, (int64_t)plan.tile);
      a.panels = (int32_t)ceil_div(N, (int64_t)bd_col_tiles(N) * kWavesPerBlock * 32);
      SEGTACO_HIP(mark(0));
      SEGTACO_HIP(mark(1));
      SEGTACO_HIP(launch_bd(b_dtype, a, stream));
      SEGTACO_HIP(mark(2));
      SEGTACO_HIP(mark(3));
      return SEGTACO_OK;
    }
    default:
      break;
  }
  if (plan.schedule == SEGTACO_EB_CS) {
    const CsLayout L = cs_layout(A->nnz, A->K, N, plan.tile, plan.group, V);
    if (!ws || ws_bytes < L.total)
      return set_error(SEGTACO_ERR_INVALID,
                       "workspace of %zu bytes needed, %zu given", L.total,
                       ws ? ws_bytes : (size_t)0);
    char *w = static_cast<char *>(ws);
    a.ntiles = L.eb.ntiles;
    a.rows_per_slice = (int32_t)ceil_div(A->M, L.eb.ntiles);
    a.fix_t0 = reinterpret_cast<int32_t *>(w + L.eb.part_off);
    a.carry_head = reinterpret_cast<float *>(w + L.eb.head_off);
    a.carry_tail = reinterpret_cast<float *>(w + L.eb.tail_off);
    a.tile_row = reinterpret_cast<int32_t *>(w + L.eb.rows_off);
    CsArgs c{};
    c.M = A->M;
    c.K = A->K;
    c.N = N;
    c.nnz = A->nnz;
    c.pos = A->pos;
    c.crd = A->crd;
    c.vals = A->vals;
    c.B = B;
    c.ldb = ldb;
    c.C = C;
    c.ldc = ldc;
    c.tile = plan.tile;
    c.ntiles = L.eb.ntiles;
    c.group = plan.group;
    c.units = L.units;
    const int32_t cus = device_cus();
    c.rounds = (int32_t)ceil_div(L.units, cus);
    c.cap = L.cap;
    c.elem = elem_bytes(b_dtype);
    c.packed = L.packed;
    c.one_bits = b_dtype == SEGTACO_DT_F32 ? 0x3f800000u
                 : b_dtype == SEGTACO_DT_BF16 ? 0x3f80u
                                              : 0x3c00u;
    c.tile_row = a.tile_row;
    c.carry_head = a.carry_head;
    c.carry_tail = a.carry_tail;
    c.s_rec = static_cast<void *>(w + L.rec_off);
    c.slot_rows = reinterpret_cast<int32_t *>(w + L.rows_off);
    c.int_rows = reinterpret_cast<int32_t *>(w + L.int_off);
    c.codes = reinterpret_cast<uint16_t *>(w + L.code_off);
    c.n_slots = reinterpret_cast<int32_t *>(w + L.nslot_off);
    c.tile_count = reinterpret_cast<int32_t *>(w + L.count_off);
    c.tile_flags = reinterpret_cast<uint8_t *>(w + L.flag_off);
    // the fix-up runs with its own natural lane count for N, and skips the
    // rows k_eb_cs wrote itself
    a.cs_flags = c.tile_flags;
    a.cs_group = c.group;
    a.lq = ilog2(natural_lanes(N, V));
    a.panels = (int32_t)ceil_div(N, (int64_t)(1 << a.lq) * V);
    // The tile rows are rewritten on every call: the same kernel zeroes C's
    // empty rows (none when C is null, in a preparation alone).
    if (prepare) prep_record(ws, nullptr);
    SEGTACO_HIP(mark(0));
    SEGTACO_HIP(launch_tile_rows_fix(V, a, stream));
    if (prepare) SEGTACO_HIP(launch_cs_prep(c, stream));
    SEGTACO_HIP(mark(1));
    if (run) {
      SEGTACO_HIP(launch_eb_cs(b_dtype, V, c, stream));
      SEGTACO_HIP(mark(2));
      SEGTACO_HIP(launch_fixup(V, a, stream));
      SEGTACO_HIP(mark(3));
    }
    if (prepare) prep_record(ws, &key);
    return SEGTACO_OK;
  }
  // Element-balanced: partition, tiles, carry fix-up.
  const EbLayout L = eb_layout(A->nnz, N, plan.tile);
  if (!ws || ws_bytes < L.total)
    return set_error(SEGTACO_ERR_INVALID,
                     "workspace of %zu bytes needed, %zu given", L.total,
                     ws ? ws_bytes : (size_t)0);
  char *w = static_cast<char *>(ws);
  a.ntiles = L.ntiles;
  a.rows_per_slice = (int32_t)ceil_div(A->M, L.ntiles);
  a.fix_t0 = reinterpret_cast<int32_t *>(w + L.part_off);
  a.carry_head = reinterpret_cast<float *>(w + L.head_off);
  a.carry_tail = reinterpret_cast<float *>(w + L.tail_off);
  a.tile_row = reinterpret_cast<int32_t *>(w + L.rows_off);
  if (prepare) prep_record(ws, nullptr);
  SEGTACO_HIP(mark(0));
  SEGTACO_HIP(launch_tile_rows_fix(V, a, stream));
  SEGTACO_HIP(mark(1));
  if (!run) return SEGTACO_OK;
  if (plan.schedule == SEGTACO_EB_SR) {
    const int64_t tiles_per_block =
        (int64_t)kWavesPerBlock * (64 / plan.col_lanes);
    a.nblocks = (int32_t)ceil_div(L.ntiles, tiles_per_block);
    SEGTACO_HIP(launch_eb_sr(b_dtype, V, a, stream));
  } else {
    a.nblocks = (int32_t)ceil_div(L.ntiles, kWavesPerBlock);
    if (spmv_applies(a, V))
      SEGTACO_HIP(launch_spmv(b_dtype, a, stream));
    else
      SEGTACO_HIP(launch_eb_pr(b_dtype, V, a, stream));
  }
  SEGTACO_HIP(mark(2));
  SEGTACO_HIP(launch_fixup(V, a, stream));
  SEGTACO_HIP(mark(3));
  return SEGTACO_OK;
}

int32_t plan_resolve_impl(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                          int64_t ldb, int64_t ldc, segtaco_plan_t *plan) {
  return resolve(A, N, b_dtype, ldb, ldc, nullptr, nullptr, plan);
}

int32_t plan_resolve_ptrs_impl(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                               int64_t ldb, int64_t ldc, const void *B, const void *C,
                               segtaco_plan_t *plan) {
  return resolve(A, N, b_dtype, ldb, ldc, B, C, plan);
}

}  // namespace segtaco

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace segtaco;

extern "C" {

int32_t segtaco_last_error(void) { return t_last_status; }
const char *segtaco_last_error_string(void) { return t_last_msg; }
void segtaco_clear_error(void) {
  t_last_status = SEGTACO_OK;
  t_last_msg[0] = '\0';
}

const char *segtaco_version(void) { return "segtaco-mi355x 0.1 (gfx950)"; }

int32_t segtaco_device_count(void) {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    hip_error(e, "hipGetDeviceCount");
    return -(int32_t)e;
  }
  return n;
}

int32_t segtaco_plan_resolve(const segtaco_csr_t *A, int32_t N,
                             int32_t b_dtype, int64_t ldb, int64_t ldc,
                             segtaco_plan_t *plan) {
  return plan_resolve_impl(A, N, b_dtype, ldb, ldc, plan);
}

size_t segtaco_spmm_workspace_bytes(const segtaco_csr_t *A, int32_t N,
                                    const segtaco_plan_t *plan) {
  return spmm_workspace_bytes_impl(A, N, plan);
}

int32_t segtaco_spmm_workspace_layout(const segtaco_csr_t *A, int32_t N,
                                      const segtaco_plan_t *plan, int64_t *out) {
  return spmm_workspace_layout_impl(A, N, plan, out);
}

int32_t segtaco_cs_workspace_layout(const segtaco_csr_t *A, int32_t N,
                                    const segtaco_plan_t *plan, int64_t *out) {
  return cs_workspace_layout_impl(A, N, plan, out);
}

int32_t segtaco_tile_rows(const int32_t *pos, int32_t M, int32_t tile, int32_t ntiles,
                          int32_t method, int32_t *rows, int32_t *part, void *stream) {
  if (!pos || !rows || M < 0 || tile < 1 || ntiles < 0)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_tile_rows: bad argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (method == SEGTACO_ROWS_SCATTER) {
    if (part) return set_error(SEGTACO_ERR_INVALID, "segtaco_tile_rows: part needs SEARCH");
    SEGTACO_HIP(launch_tile_rows(pos, M, tile, ntiles, rows, s));
  } else if (method == SEGTACO_ROWS_SEARCH) {
    SEGTACO_HIP(launch_block_starts(pos, M, tile, ntiles, part, rows, s));
  } else {
    return set_error(SEGTACO_ERR_INVALID, "segtaco_tile_rows: method %d", method);
  }
  return SEGTACO_OK;
}

int32_t segtaco_entry_rows(const int32_t *pos, int32_t M, int32_t nnz, int32_t tile,
                           const int32_t *tile_rows, int32_t *rows, void *stream) {
  if (M < 0 || nnz < 0 || tile < 8 || (nnz > 0 && (!pos || !tile_rows || !rows)))
    return set_error(SEGTACO_ERR_INVALID, "segtaco_entry_rows: bad argument");
  if (nnz == 0) return SEGTACO_OK;
  SEGTACO_HIP(launch_entry_rows(pos, M, nnz, tile, tile_rows, rows,
                                static_cast<hipStream_t>(stream)));
  return SEGTACO_OK;
}

int32_t segtaco_spmm(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                     int64_t ldb, int32_t N, float *C, int64_t ldc,
                     const segtaco_plan_t *plan, void *workspace,
                     size_t workspace_bytes, void *stream) {
  return spmm_impl(A, B, b_dtype, ldb, N, C, ldc, plan, workspace,
                   workspace_bytes, static_cast<hipStream_t>(stream), nullptr);
}

int32_t segtaco_spmm_prepare(const segtaco_csr_t *A, int32_t b_dtype, int64_t ldb, int32_t N,
                             int64_t ldc, const segtaco_plan_t *plan, void *workspace,
                             size_t workspace_bytes, void *stream) {
  return spmm_impl(A, nullptr, b_dtype, ldb, N, nullptr, ldc, plan, workspace,
                   workspace_bytes, static_cast<hipStream_t>(stream), nullptr, kPhasePrepare);
}

int32_t segtaco_spmm_prepared(const segtaco_csr_t *A, const void *B, int32_t b_dtype,
                              int64_t ldb, int32_t N, float *C, int64_t ldc,
                              const segtaco_plan_t *plan, void *workspace,
                              size_t workspace_bytes, void *stream) {
  return spmm_impl(A, B, b_dtype, ldb, N, C, ldc, plan, workspace, workspace_bytes,
                   static_cast<hipStream_t>(stream), nullptr, kPhaseRun);
}

int32_t segtaco_spmm_stage_times(const segtaco_csr_t *A, const void *B,
                                 int32_t b_dtype, int64_t ldb, int32_t N,
                                 float *C, int64_t ldc,
                                 const segtaco_plan_t *plan, void *workspace,
                                 size_t workspace_bytes, void *stream_v,
                                 int32_t reps, float *out_ms) {
  if (!out_ms || reps < 1)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_spmm_stage_times: bad argument");
  hipStream_t stream = static_cast<hipStream_t>(stream_v);
  // All reps are enqueued back to back (as in a steady loop of calls), each
  // with its own four events; the host waits once at the end. The events
  // between kernels skip the system-scope fence (an L2 writeback and
  // invalidate that a loop of untimed calls does not have); the last one
  // keeps it, so waiting on it covers all the work.
  std::vector<hipEvent_t> ev((size_t)reps * 4, nullptr);
  int32_t st = SEGTACO_OK;
  for (auto &e : ev) {
    const unsigned flags = &e == &ev.back() ? hipEventDefault : hipEventDisableSystemFence;
    const hipError_t he = hipEventCreateWithFlags(&e, flags);
    if (he != hipSuccess) {
      e = nullptr;
      st = hip_error(he, "hipEventCreate");
      break;
    }
  }
  for (int32_t r = 0; r < reps && st == SEGTACO_OK; ++r)
    st = spmm_impl(A, B, b_dtype, ldb, N, C, ldc, plan, workspace,
                   workspace_bytes, stream, ev.data() + 4 * (size_t)r);
  if (st == SEGTACO_OK && hipEventSynchronize(ev.back()) != hipSuccess)
    st = set_error(SEGTACO_ERR_HIP, "hipEventSynchronize");
  double acc[3] = {0, 0, 0};
  for (int32_t r = 0; r < reps && st == SEGTACO_OK; ++r)
    for (int i = 0; i < 3; ++i) {
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev[4 * (size_t)r + i], ev[4 * (size_t)r + i + 1]);
      acc[i] += ms;
    }
  if (st != SEGTACO_OK) (void)hipStreamSynchronize(stream);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st != SEGTACO_OK) return st;
  for (int i = 0; i < 3; ++i) out_ms[i] = (float)(acc[i] / reps);
  return SEGTACO_OK;
}

int32_t segtaco_block_starts(const int32_t *pos, int32_t M,
                             int32_t values_per_block, int32_t num_blocks,
                             int32_t *out, void *stream) {
  if (!pos || !out || M < 0 || values_per_block < 1 || num_blocks < 0)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_block_starts: bad argument");
  SEGTACO_HIP(launch_block_starts(pos, M, values_per_block, num_blocks, out,
                                  nullptr, static_cast<hipStream_t>(stream)));
  return SEGTACO_OK;
}

int32_t segtaco_csr_validate_host(int32_t M, int32_t K, int32_t nnz,
                                  const int32_t *pos, const int32_t *crd) {
  if (M < 0 || K < 0 || nnz < 0)
    return set_error(SEGTACO_ERR_INVALID, "negative CSR size");
  if (!pos) return set_error(SEGTACO_ERR_INVALID, "null pos");
  if (pos[0] != 0) return set_error(SEGTACO_ERR_INVALID, "pos[0] = %d", pos[0]);
  // Branch-free scans over blocks (they vectorise); the offending index is
  // searched for only once a block fails.
  constexpr int32_t kBlk = 4096;
  for (int32_t r0 = 0; r0 < M; r0 += kBlk) {
    const int32_t r1 = M - r0 < kBlk ? M : r0 + kBlk;
    int bad = 0;
    for (int32_t r = r0; r < r1; ++r) bad |= pos[r + 1] < pos[r];
    if (bad)
      for (int32_t r = r0; r < r1; ++r)
        if (pos[r + 1] < pos[r])
          return set_error(SEGTACO_ERR_INVALID, "pos decreases at row %d", r);
  }
  if (pos[M] != nnz)
    return set_error(SEGTACO_ERR_INVALID, "pos[M] = %d != nnz = %d", pos[M], nnz);
  if (nnz > 0 && !crd) return set_error(SEGTACO_ERR_INVALID, "null crd");
  const uint32_t k = (uint32_t)K;
  for (int32_t p0 = 0; p0 < nnz; p0 += kBlk) {
    const int32_t p1 = nnz - p0 < kBlk ? nnz : p0 + kBlk;
    int bad = 0;
    for (int32_t p = p0; p < p1; ++p) bad |= (uint32_t)crd[p] >= k;
    if (bad)
      for (int32_t p = p0; p < p1; ++p)
        if (crd[p] < 0 || crd[p] >= K)
          return set_error(SEGTACO_ERR_INVALID, "crd[%d] = %d outside [0, %d)", p,
                           crd[p], K);
  }
  return SEGTACO_OK;
}

int32_t segtaco_device_malloc(void **ptr, size_t bytes) {
  if (!ptr) return set_error(SEGTACO_ERR_INVALID, "null ptr");
  *ptr = nullptr;
  if (bytes == 0) return SEGTACO_OK;
  SEGTACO_HIP(hipMalloc(ptr, bytes));
  return SEGTACO_OK;
}

int32_t segtaco_device_free(void *ptr) {
  if (ptr) SEGTACO_HIP(hipFree(ptr));
  return SEGTACO_OK;
}

int32_t segtaco_memcpy(void *dst, const void *src, size_t bytes, int32_t kind) {
  if (bytes == 0) return SEGTACO_OK;
  hipMemcpyKind k = kind == 0   ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost
                                : hipMemcpyDeviceToDevice;
  SEGTACO_HIP(hipMemcpy(dst, src, bytes, k));
  return SEGTACO_OK;
}

int32_t segtaco_memset(void *dst, int32_t value, size_t bytes) {
  if (bytes == 0) return SEGTACO_OK;
  SEGTACO_HIP(hipMemset(dst, value, bytes));
  return SEGTACO_OK;
}

int32_t segtaco_device_synchronize(void) {
  SEGTACO_HIP(hipDeviceSynchronize());
  return SEGTACO_OK;
}

int32_t segtaco_stream_synchronize(void *stream) {
  SEGTACO_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return SEGTACO_OK;
}

}  // extern "C"

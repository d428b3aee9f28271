// serve.cpp — continuous-batching SpMM server (segtaco/serve.h).
//
// One scheduler thread per server. Each pass it (1) retires ticks whose
// completion event has fired, releasing their workspace and matrix pins and
// recording latencies; (2) if a stream slot is free, admits pending requests
// in arrival order while they fit — the request's matrix must be resident in
// the CSR cache (uploaded on a miss, evicting least-recently-used matrices
// that no running tick uses) and its workspace must fit the in-flight arena —
// and launches them as one tick: the batched persistent kernel plus the
// batched fix-up on that slot's stream. Admission stops at the first request
// that does not fit (no overtaking), which keeps arrival order. Ticks on
// different slots overlap on the device.
#include <sys/prctl.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"
#include "launch.h"
#include "spmv.h"
#include "segtaco/autotune.h"
#include "segtaco/serve.h"

namespace segtaco {
int32_t plan_resolve_impl(const segtaco_csr_t *A, int32_t N, int32_t b_dtype,
                          int64_t ldb, int64_t ldc, segtaco_plan_t *plan);
}  // namespace segtaco

using namespace segtaco;
using Clock = std::chrono::steady_clock;

namespace {

// Best-fit allocator over one device allocation, 256-byte granules.
struct Arena {
  char *base = nullptr;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes

  hipError_t init(size_t bytes) {
    cap = align_up(bytes, 256);
    if (cap == 0) return hipSuccess;
    hipError_t e = hipMalloc(&base, cap);
    if (e == hipSuccess) free_[0] = cap;
    return e;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = align_up(bytes ? bytes : 1, 256);
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes && (best == free_.end() || it->second < best->second))
        best = it;
    if (best == free_.end()) return false;
    *off = best->first;
    const size_t rest = best->second - bytes;
    const size_t at = best->first + bytes;
    free_.erase(best);
    if (rest) free_[at] = rest;
    return true;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(bytes ? bytes : 1, 256);
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  void destroy() {
    if (base) (void)hipFree(base);
    base = nullptr;
  }
};

struct TileRows {
  int32_t tile;
  size_t off, bytes;
  int32_t *part, *rows;
};

struct MatrixRec {
  int32_t M = 0, K = 0, nnz = 0;
  int32_t *h_pos = nullptr, *h_crd = nullptr;
  void *h_vals = nullptr;
  size_t val_bytes = 0;
  segtaco_row_stats_t stats{};
  bool resident = false;
  size_t off = 0, bytes = 0;
  int32_t *d_pos = nullptr, *d_crd = nullptr;
  void *d_vals = nullptr;
  std::vector<TileRows> tiles;
  int64_t last_use = 0;
  int32_t inflight = 0;
  // The tick whose stream last wrote this matrix's device data (upload or
  // tile rows); ticks on other streams wait for it while it runs.
  int32_t ready_slot = -1;
  int64_t ready_tick = -1;
  std::map<int32_t, segtaco_plan_t> plan_by_n;
};

struct Req {
  int32_t mat = 0, N = 0;
  const void *B = nullptr;
  float *C = nullptr;
  int64_t ldb = 0, ldc = 0;
  Clock::time_point t_submit, t_done;
  int32_t status = -1;  // -1 pending
  bool done = false;
  size_t ws_off = 0, ws_bytes = 0;
  segtaco_plan_t plan{};
  double flops = 0.0;
};

// A tick's launch tables travel in one copy: one pinned host buffer and its
// device twin, laid out as [work counter (zero)][unit prefix, n + 1]
// [fix-up prefix, n + 1][descriptors, n] (PackLayout). Uploading the zero
// counter with the tables replaces a reset kernel. The counter has a
// 256-byte block to itself: every workgroup increments it atomically and
// then bisects the prefix array, and with both in one cache line each
// bisection read contended with the atomics (a power-law 262144^2 BF16
// N = 128 burst ran at 5150 instead of 6570 GFLOP/s).
struct PackLayout {
  size_t prefix_off, fix_off, desc_off, total;
  explicit PackLayout(size_t n) {
    prefix_off = 256;
    fix_off = prefix_off + sizeof(int64_t) * (n + 1);
    desc_off = align_up(fix_off + sizeof(int64_t) * (n + 1), 64);
    total = desc_off + sizeof(BatchDesc) * n;
  }
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  char *d_pack = nullptr, *h_pack = nullptr;  // PackLayout(max_batch)
  BatchDesc *d_desc = nullptr, *h_desc = nullptr;
  int64_t *d_prefix = nullptr, *h_prefix = nullptr;
  int64_t *d_fix = nullptr, *h_fix = nullptr;
  unsigned long long *d_counter = nullptr, *h_counter = nullptr;
  std::vector<int64_t> reqs;
  int64_t tick_id = -1;
  bool busy = false;
  // SEGTACO_OK, or SEGTACO_ERR_HIP when a step of the running tick's launch
  // or its completion query failed: its requests then retire with that error
  int32_t status = SEGTACO_OK;
};

size_t csr_bytes(const MatrixRec &m) {
  return align_up(sizeof(int32_t) * ((size_t)m.M + 1), 256) +
         align_up(sizeof(int32_t) * (size_t)m.nnz, 256) + align_up(m.val_bytes, 256);
}

size_t tile_rows_bytes(int32_t nnz, int32_t tile) {
  const size_t nt = (size_t)ceil_div(nnz, tile) + 1;
  return 2 * align_up(sizeof(int32_t) * nt, 256);
}

// Carries of an element-balanced plan: 2 * ntiles * N floats.
size_t carry_bytes(int32_t nnz, int32_t N, const segtaco_plan_t &p) {
  if (p.schedule < SEGTACO_EB_SR || nnz == 0) return 0;
  const size_t nt = (size_t)ceil_div(nnz, p.tile);
  return 2 * align_up(sizeof(float) * nt * (size_t)N, 256);
}

}  // namespace

struct segtaco_server {
  segtaco_serve_config_t cfg{};
  int device = 0;
  Arena cache, inflight;
  std::vector<std::unique_ptr<MatrixRec>> mats;
  std::unordered_map<int64_t, Req> reqs;
  std::deque<int64_t> pending;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread th;
  bool stop = false;
  int64_t next_id = 0, use_clock = 0;
  std::map<std::pair<size_t, bool>, int32_t> grid_by_lds;  // (LDS, N = 1 form)
  // statistics
  std::vector<double> lat_us;
  int64_t submitted = 0, completed = 0, rejected = 0, ticks = 0;
  int64_t hits = 0, misses = 0, evictions = 0, uploaded = 0;
  int64_t batch_products = 0;
  double flops_done = 0.0;
  bool have_first = false;
  Clock::time_point first_submit, last_done;
  int32_t max_conc = 0;
  int32_t fatal = SEGTACO_OK;
  // error-path testing: the launch of this tick (1-based) reports a HIP
  // error without running (segtaco_server_debug_fail_tick); 0 = never
  int64_t fail_tick = 0;
};

namespace {

segtaco_plan_t plan_for(segtaco_server *s, MatrixRec &m, int32_t N) {
  auto it = m.plan_by_n.find(N);
  if (it != m.plan_by_n.end()) return it->second;
  segtaco_plan_t p{};
  if (m.nnz == 0 || segtaco_tune(&m.stats, N, s->cfg.dtype, &p) != SEGTACO_OK)
    p = segtaco_plan_t{SEGTACO_RB_SR, 0, 0, 0, 0, -1};
  p.vec = 0;  // widest the request's pointers allow
  // A tick often carries a single product: size its tiles for latency.
  if (p.schedule == SEGTACO_EB_SR && p.col_lanes > 0 && p.col_lanes <= 64 && p.tile > 0)
    p.tile = latency_tile(m.nnz, p.tile, kWavesPerBlock * (64 / p.col_lanes));
  m.plan_by_n[N] = p;
  return p;
}

void evict(segtaco_server *s, MatrixRec &m) {
  for (auto &t : m.tiles) s->cache.release(t.off, t.bytes);
  m.tiles.clear();
  s->cache.release(m.off, m.bytes);
  m.resident = false;
  m.d_pos = m.d_crd = nullptr;
  m.d_vals = nullptr;
  ++s->evictions;
}

// Allocates `bytes` in the CSR cache, evicting least-recently-used matrices
// that no tick uses and whose uploads have landed. `keep` (the matrix being
// admitted) is never evicted; its tile-row arrays for other tile sizes are
// released, oldest first, once no other matrix can go, provided only the
// request being admitted pins it (no running tick and no earlier request of
// this tick uses the matrix, so none of those arrays is in use).
bool cache_alloc(segtaco_server *s, size_t bytes, MatrixRec *keep, size_t *off) {
  while (!s->cache.alloc(bytes, off)) {
    MatrixRec *victim = nullptr;
    for (auto &mp : s->mats) {
      MatrixRec &m = *mp;
      if (!m.resident || m.inflight > 0 || &m == keep) continue;
      if (m.ready_slot >= 0) {  // its upload may still be in flight
        const Slot &w = s->slots[m.ready_slot];
        if (w.busy && w.tick_id == m.ready_tick) continue;
      }
      if (!victim || m.last_use < victim->last_use) victim = &m;
    }
    if (victim) {
      evict(s, *victim);
      continue;
    }
    if (keep && keep->inflight == 1 && !keep->tiles.empty()) {
      const TileRows t = keep->tiles.front();
      keep->tiles.erase(keep->tiles.begin());
      s->cache.release(t.off, t.bytes);
      continue;
    }
    return false;
  }
  return true;
}

// Makes m resident (and, for an element-balanced plan, its tile rows for the
// plan's tile), issuing uploads / the partition on slot `si`'s stream as part
// of tick `tick`; *issued is set when device work was queued.
bool ensure_resident(segtaco_server *s, MatrixRec &m, const segtaco_plan_t &p,
                     int32_t si, int64_t tick, TileRows **tr, bool *issued) {
  *tr = nullptr;
  hipStream_t stream = s->slots[si].stream;
  if (!m.resident) {
    size_t off;
    const size_t bytes = csr_bytes(m);
    if (!cache_alloc(s, bytes, &m, &off)) return false;
    char *base = s->cache.base + off;
    m.off = off;
    m.bytes = bytes;
    m.d_pos = reinterpret_cast<int32_t *>(base);
    m.d_crd = reinterpret_cast<int32_t *>(base + align_up(sizeof(int32_t) * ((size_t)m.M + 1), 256));
    m.d_vals = m.val_bytes ? base + align_up(sizeof(int32_t) * ((size_t)m.M + 1), 256) +
                                 align_up(sizeof(int32_t) * (size_t)m.nnz, 256)
                           : nullptr;
    if (hipMemcpyAsync(m.d_pos, m.h_pos, sizeof(int32_t) * ((size_t)m.M + 1),
                       hipMemcpyHostToDevice, stream) != hipSuccess ||
        (m.nnz && hipMemcpyAsync(m.d_crd, m.h_crd, sizeof(int32_t) * (size_t)m.nnz,
                                 hipMemcpyHostToDevice, stream) != hipSuccess) ||
        (m.val_bytes && hipMemcpyAsync(m.d_vals, m.h_vals, m.val_bytes,
                                       hipMemcpyHostToDevice, stream) != hipSuccess)) {
      s->cache.release(off, bytes);
      s->fatal = SEGTACO_ERR_HIP;
      return false;
    }
    m.resident = true;
    m.ready_slot = si;
    m.ready_tick = tick;
    *issued = true;
    ++s->misses;
    s->uploaded += (int64_t)(sizeof(int32_t) * ((size_t)m.M + 1 + m.nnz) + m.val_bytes);
  } else {
    ++s->hits;
  }
  if (p.schedule >= SEGTACO_EB_SR && m.nnz > 0) {
    for (auto &t : m.tiles)
      if (t.tile == p.tile) *tr = &t;
    if (!*tr) {
      size_t off;
      const size_t bytes = tile_rows_bytes(m.nnz, p.tile);
      if (!cache_alloc(s, bytes, &m, &off)) return false;
      const int32_t nt = (int32_t)ceil_div(m.nnz, p.tile);
      TileRows t{p.tile, off, bytes, nullptr, nullptr};
      t.part = reinterpret_cast<int32_t *>(s->cache.base + off);
      t.rows = reinterpret_cast<int32_t *>(s->cache.base + off +
                                           align_up(sizeof(int32_t) * ((size_t)nt + 1), 256));
      if (launch_block_starts(m.d_pos, m.M, p.tile, nt, t.part, t.rows, stream) != hipSuccess) {
        s->cache.release(off, bytes);
        s->fatal = SEGTACO_ERR_HIP;
        return false;
      }
      m.tiles.push_back(t);
      *tr = &m.tiles.back();
      m.ready_slot = si;
      m.ready_tick = tick;
      *issued = true;
    }
  }
  return true;
}

void finish_request(segtaco_server *s, int64_t id, int32_t status,
                    Clock::time_point now) {
  Req &r = s->reqs[id];
  r.status = status;
  r.done = true;
  r.t_done = now;
  if (status == SEGTACO_OK) {
    s->lat_us.push_back(std::chrono::duration<double, std::micro>(now - r.t_submit).count());
    s->flops_done += r.flops;
    ++s->completed;
    s->last_done = now;
  } else {
    ++s->rejected;
  }
}

// Retires the tick of `slot`: its requests finish with the tick's status. A
// tick whose launch or completion query failed may still have device work
// queued (its completion event was never recorded, or cannot be read), so
// its stream is drained before the workspace and pins are released.
void retire(segtaco_server *s, Slot &slot) {
  if (slot.status != SEGTACO_OK) (void)hipStreamSynchronize(slot.stream);
  const auto now = Clock::now();
  for (int64_t id : slot.reqs) {
    Req &r = s->reqs[id];
    if (r.ws_bytes) s->inflight.release(r.ws_off, r.ws_bytes);
    s->mats[r.mat]->inflight--;
    finish_request(s, id, slot.status, now);
  }
  slot.reqs.clear();
  slot.busy = false;
  slot.status = SEGTACO_OK;
  s->cv_done.notify_all();
}

// After a HIP error nothing more is launched: every request still waiting
// for admission finishes with SEGTACO_ERR_HIP.
void fail_pending(segtaco_server *s) {
  const auto now = Clock::now();
  while (!s->pending.empty()) {
    const int64_t id = s->pending.front();
    s->pending.pop_front();
    finish_request(s, id, SEGTACO_ERR_HIP, now);
  }
  s->cv_done.notify_all();
}

// Admits pending requests into slot `si` and launches them. Caller holds mu.
void launch_tick(segtaco_server *s, int32_t si) {
  if (s->fatal != SEGTACO_OK) return;
  Slot &slot = s->slots[si];
  const int64_t tick = s->ticks + 1;
  std::vector<int64_t> chosen;
  // Copies, not pointers into MatrixRec::tiles: a later request of this tick
  // on the same matrix with another tile size appends to that vector, which
  // may move its elements.
  std::vector<TileRows> trs;
  bool issued = false;
  while (!s->pending.empty() && (int32_t)chosen.size() < s->cfg.max_batch) {
    const int64_t id = s->pending.front();
    Req &r = s->reqs[id];
    MatrixRec &m = *s->mats[r.mat];
    segtaco_plan_t p = plan_for(s, m, r.N);
    segtaco_csr_t A{m.M, m.K, m.nnz, s->cfg.dtype, nullptr, nullptr, nullptr};
    if (plan_resolve_impl(&A, r.N, s->cfg.dtype, r.ldb, r.ldc, &p) != SEGTACO_OK) {
      s->pending.pop_front();
      finish_request(s, id, SEGTACO_ERR_INVALID, Clock::now());
      continue;
    }
    m.inflight++;  // pin while its uploads are queued
    // Data of this matrix written by a tick still running on another stream
    // (its CSR upload or a tile-row array): this stream waits for it before
    // it reads the matrix, including the tile-row pass ensure_resident may
    // launch for a tile size the matrix does not have yet.
    if (m.ready_slot >= 0 && m.ready_slot != si) {
      const Slot &w = s->slots[m.ready_slot];
      if (w.busy && w.tick_id == m.ready_tick &&
          hipStreamWaitEvent(slot.stream, w.done, 0) != hipSuccess)
        s->fatal = SEGTACO_ERR_HIP;
    }
    TileRows *tr = nullptr;
    if (!ensure_resident(s, m, p, si, tick, &tr, &issued)) {
      m.inflight--;
      break;
    }
    const size_t ws = carry_bytes(m.nnz, r.N, p);
    size_t off = 0;
    if (ws && !s->inflight.alloc(ws, &off)) {
      m.inflight--;
      break;
    }
    r.ws_off = off;
    r.ws_bytes = ws;
    r.plan = p;
    m.last_use = ++s->use_clock;
    s->pending.pop_front();
    chosen.push_back(id);
    trs.push_back(tr ? *tr : TileRows{});
  }
  if (chosen.empty() && !issued) return;
  size_t lds = 0;
  bool sv = false;  // some product runs the N = 1 form of eb_pr
  int64_t total = 0, total_fix = 0;
  slot.h_prefix[0] = slot.h_fix[0] = 0;
  for (size_t i = 0; i < chosen.size(); ++i) {
    Req &r = s->reqs[chosen[i]];
    MatrixRec &m = *s->mats[r.mat];
    segtaco_csr_t A{m.M, m.K, m.nnz, s->cfg.dtype, m.d_pos, m.d_crd, m.d_vals};
    BatchDesc &D = slot.h_desc[i];
    int64_t units = 0, fix = 0;
    size_t l = 0;
    EbLayout lay{};
    fill_batch_desc(A, r.B, r.ldb, r.C, r.ldc, r.N, r.plan, &D, &units, &fix, &l, &lay);
    if (r.plan.schedule >= SEGTACO_EB_SR && m.nnz > 0) {
      D.a.part = trs[i].part;
      D.a.tile_row = trs[i].rows;
      const size_t half = r.ws_bytes / 2;
      D.a.carry_head = reinterpret_cast<float *>(s->inflight.base + r.ws_off);
      D.a.carry_tail = reinterpret_cast<float *>(s->inflight.base + r.ws_off + half);
    }
    if (m.M == 0 || r.N == 0) units = fix = 0;
    D.unit_begin = total;
    D.fix_begin = total_fix;
    total += units;
    total_fix += fix;
    slot.h_prefix[i + 1] = total;
    slot.h_fix[i + 1] = total_fix;
    lds = std::max(lds, l);
    sv = sv || batch_uses_spmv(D);
    r.flops = 2.0 * m.nnz * r.N;
  }
  const int32_t n = (int32_t)chosen.size();
  int32_t grid = 1;
  auto g = s->grid_by_lds.find({lds, sv});
  if (g != s->grid_by_lds.end()) {
    grid = g->second;
  } else if (batch_main_grid(s->cfg.dtype, lds, &grid, sv) == hipSuccess) {
    s->grid_by_lds[{lds, sv}] = grid;
  } else {
    grid = 1024;
  }
  // a forced failure (error-path testing) skips every launch of the tick
  hipError_t e = tick == s->fail_tick ? hipErrorLaunchFailure : hipSuccess;
  // one copy: counter, both prefix arrays, the first n descriptors
  *slot.h_counter = 0ull;
  if (e == hipSuccess && n > 0)
    e = hipMemcpyAsync(slot.d_pack, slot.h_pack,
                       reinterpret_cast<char *>(slot.h_desc + n) - slot.h_pack,
                       hipMemcpyHostToDevice, slot.stream);
  if (e == hipSuccess && n > 0)
    e = launch_batch_main(s->cfg.dtype, slot.d_desc, n, slot.d_prefix, total,
                          slot.d_counter, grid, lds, slot.stream, /*reset_counter=*/false, sv);
  if (e == hipSuccess && n > 0)
    e = launch_batch_fixup(slot.d_desc, n, slot.d_fix, total_fix, slot.stream);
  if (e == hipSuccess) e = hipEventRecord(slot.done, slot.stream);
  if (e != hipSuccess) s->fatal = SEGTACO_ERR_HIP;
  // any failure while this tick was put together (an upload, a partition, a
  // cross-stream wait or a launch) fails the tick's requests
  slot.status = s->fatal != SEGTACO_OK ? SEGTACO_ERR_HIP : SEGTACO_OK;
  slot.reqs = chosen;
  slot.busy = true;
  slot.tick_id = tick;
  ++s->ticks;
  s->batch_products += n;
  int32_t busy = 0;
  for (auto &sl : s->slots) busy += sl.busy ? 1 : 0;
  s->max_conc = std::max(s->max_conc, busy);
}

// Interval at which the scheduler polls a running tick's completion event.
#ifndef SEGTACO_POLL_US
#define SEGTACO_POLL_US 2
#endif
constexpr int kPollUs = SEGTACO_POLL_US;

void scheduler(segtaco_server *s) {
  (void)hipSetDevice(s->device);
  // The waits below are a few microseconds long while a tick runs; Linux
  // otherwise rounds such a timed wait up by the default 50 us timer slack,
  // which lands directly on every request's latency.
  (void)prctl(PR_SET_TIMERSLACK, 1UL, 0UL, 0UL, 0UL);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progressed = false;
    for (auto &slot : s->slots) {
      if (slot.busy) {
        const hipError_t q = hipEventQuery(slot.done);
        if (q == hipSuccess) {
          retire(s, slot);
          progressed = true;
        } else if (q != hipErrorNotReady) {
          s->fatal = SEGTACO_ERR_HIP;
          slot.status = SEGTACO_ERR_HIP;
          retire(s, slot);
          progressed = true;
        }
      }
    }
    if (s->fatal != SEGTACO_OK && !s->pending.empty()) fail_pending(s);
    bool any_busy = false;
    for (auto &slot : s->slots) any_busy = any_busy || slot.busy;
    if (s->stop && s->pending.empty() && !any_busy) break;
    bool launched = false;
    if (!s->pending.empty()) {
      for (int32_t si = 0; si < (int32_t)s->slots.size(); ++si) {
        if (!s->slots[si].busy && !s->pending.empty()) {
          const int64_t before = s->ticks;
          launch_tick(s, si);
          launched = launched || s->ticks != before;
        }
      }
      progressed = progressed || launched;
    }
    if (!any_busy && !launched && !s->pending.empty()) {
      // Nothing runs, so the whole in-flight arena is free and every cached
      // matrix may be evicted, and still the head request cannot be
      // admitted: it never will be. It fails with a capacity error and the
      // requests behind it get their turn.
      const int64_t id = s->pending.front();
      s->pending.pop_front();
      finish_request(s, id, SEGTACO_ERR_CAPACITY, Clock::now());
      s->cv_done.notify_all();
      continue;
    }
    // A running tick's completion event is polled every few microseconds;
    // an idle scheduler sleeps until a submission (or stop) wakes it.
    if (!progressed)
      s->cv_work.wait_for(lk, std::chrono::microseconds(any_busy ? kPollUs : 500));
  }
}

}  // namespace

extern "C" {

int32_t segtaco_server_create(const segtaco_serve_config_t *cfg,
                              segtaco_server_t *out) {
  if (!cfg || !out) return set_error(SEGTACO_ERR_INVALID, "null argument");
  if (cfg->dtype < SEGTACO_DT_F32 || cfg->dtype > SEGTACO_DT_F16 ||
      cfg->max_batch < 1 || cfg->streams < 1 || cfg->streams > 8)
    return set_error(SEGTACO_ERR_INVALID, "bad server config");
  *out = nullptr;
  auto s = std::make_unique<segtaco_server>();
  s->cfg = *cfg;
  SEGTACO_HIP(hipGetDevice(&s->device));
  SEGTACO_HIP(s->cache.init(cfg->cache_bytes));
  SEGTACO_HIP(s->inflight.init(cfg->inflight_bytes));
  s->slots.resize(cfg->streams);
  const size_t n = (size_t)cfg->max_batch;
  for (auto &sl : s->slots) {
    SEGTACO_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    SEGTACO_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    const PackLayout L(n);
    SEGTACO_HIP(hipMalloc(&sl.d_pack, L.total));
    SEGTACO_HIP(hipHostMalloc(&sl.h_pack, L.total, hipHostMallocDefault));
    std::memset(sl.h_pack, 0, L.total);
    sl.d_counter = reinterpret_cast<unsigned long long *>(sl.d_pack);
    sl.h_counter = reinterpret_cast<unsigned long long *>(sl.h_pack);
    sl.d_prefix = reinterpret_cast<int64_t *>(sl.d_pack + L.prefix_off);
    sl.h_prefix = reinterpret_cast<int64_t *>(sl.h_pack + L.prefix_off);
    sl.d_fix = reinterpret_cast<int64_t *>(sl.d_pack + L.fix_off);
    sl.h_fix = reinterpret_cast<int64_t *>(sl.h_pack + L.fix_off);
    sl.d_desc = reinterpret_cast<BatchDesc *>(sl.d_pack + L.desc_off);
    sl.h_desc = reinterpret_cast<BatchDesc *>(sl.h_pack + L.desc_off);
  }
  segtaco_server *raw = s.release();
  raw->th = std::thread(scheduler, raw);
  *out = raw;
  return SEGTACO_OK;
}

int32_t segtaco_server_add_matrix(segtaco_server_t s, int32_t M, int32_t K,
                                  int32_t nnz, const int32_t *pos,
                                  const int32_t *crd, const void *vals,
                                  int32_t *matrix_id) {
  if (!s || !matrix_id) return set_error(SEGTACO_ERR_INVALID, "null argument");
  if (nnz > kMaxNnz)
    return set_error(SEGTACO_ERR_UNSUPPORTED, "nnz %d above the supported %d", nnz, kMaxNnz);
  SEGTACO_RET(segtaco_csr_validate_host(M, K, nnz, pos, crd));
  auto m = std::make_unique<MatrixRec>();
  m->M = M;
  m->K = K;
  m->nnz = nnz;
  m->val_bytes = vals ? (size_t)nnz * elem_bytes(s->cfg.dtype) : 0;
  SEGTACO_RET(segtaco_row_stats(pos, M, K, &m->stats));
  SEGTACO_HIP(hipHostMalloc(&m->h_pos, sizeof(int32_t) * ((size_t)M + 1), hipHostMallocDefault));
  std::memcpy(m->h_pos, pos, sizeof(int32_t) * ((size_t)M + 1));
  if (nnz) {
    SEGTACO_HIP(hipHostMalloc(&m->h_crd, sizeof(int32_t) * (size_t)nnz, hipHostMallocDefault));
    std::memcpy(m->h_crd, crd, sizeof(int32_t) * (size_t)nnz);
  }
  if (m->val_bytes) {
    SEGTACO_HIP(hipHostMalloc(&m->h_vals, m->val_bytes, hipHostMallocDefault));
    std::memcpy(m->h_vals, vals, m->val_bytes);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  s->mats.push_back(std::move(m));
  *matrix_id = (int32_t)s->mats.size() - 1;
  return SEGTACO_OK;
}

int32_t segtaco_server_submit(segtaco_server_t s, int32_t matrix_id,
                              const void *B, int64_t ldb, int32_t N, float *C,
                              int64_t ldc, int64_t *request_id) {
  if (!s || !request_id) return set_error(SEGTACO_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->fatal != SEGTACO_OK) return set_error(s->fatal, "server stopped after a HIP error");
  if (matrix_id < 0 || matrix_id >= (int32_t)s->mats.size())
    return set_error(SEGTACO_ERR_INVALID, "unknown matrix %d", matrix_id);
  MatrixRec &m = *s->mats[matrix_id];
  if (N < 0 || (m.M > 0 && N > 0 && (!C || ldc < N)) ||
      (m.nnz > 0 && N > 0 && (!B || ldb < N)))
    return set_error(SEGTACO_ERR_INVALID, "bad request operands");
  const int64_t id = s->next_id++;
  Req &r = s->reqs[id];
  r.mat = matrix_id;
  r.B = B;
  r.ldb = ldb;
  r.N = N;
  r.C = C;
  r.ldc = ldc;
  r.t_submit = Clock::now();
  if (!s->have_first) {
    s->have_first = true;
    s->first_submit = r.t_submit;
  }
  ++s->submitted;
  *request_id = id;
  // Requests that can never fit complete at once with a capacity error.
  const segtaco_plan_t p = plan_for(s, m, N);
  segtaco_plan_t pr = p;
  segtaco_csr_t A{m.M, m.K, m.nnz, s->cfg.dtype, nullptr, nullptr, nullptr};
  const bool resolvable = plan_resolve_impl(&A, N, s->cfg.dtype, ldb, ldc, &pr) == SEGTACO_OK;
  const size_t need_cache = csr_bytes(m) + (pr.schedule >= SEGTACO_EB_SR && m.nnz
                                                ? tile_rows_bytes(m.nnz, pr.tile) : 0);
  if (!resolvable || need_cache > s->cache.cap ||
      carry_bytes(m.nnz, N, pr) > s->inflight.cap) {
    finish_request(s, id, resolvable ? SEGTACO_ERR_CAPACITY : SEGTACO_ERR_INVALID,
                   r.t_submit);
    s->cv_done.notify_all();
    return SEGTACO_OK;
  }
  s->pending.push_back(id);
  lk.unlock();
  s->cv_work.notify_one();
  return SEGTACO_OK;
}

int32_t segtaco_server_wait(segtaco_server_t s, int64_t id, double timeout_s,
                            int32_t *status) {
  if (!s) return set_error(SEGTACO_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return set_error(SEGTACO_ERR_INVALID, "unknown request %lld", (long long)id);
  auto ready = [&] { return s->reqs[id].done; };
  if (timeout_s < 0) {
    s->cv_done.wait(lk, ready);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
    return set_error(SEGTACO_ERR_STATE, "request %lld still running", (long long)id);
  }
  if (status) *status = s->reqs[id].status;
  s->reqs.erase(id);
  return SEGTACO_OK;
}

int32_t segtaco_server_drain(segtaco_server_t s, double timeout_s) {
  if (!s) return set_error(SEGTACO_ERR_INVALID, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto idle = [&] {
    if (!s->pending.empty()) return false;
    for (auto &sl : s->slots)
      if (sl.busy) return false;
    return true;
  };
  s->cv_work.notify_one();
  if (timeout_s < 0) {
    s->cv_done.wait(lk, idle);
  } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle)) {
    return set_error(SEGTACO_ERR_STATE, "server still busy");
  }
  for (auto it = s->reqs.begin(); it != s->reqs.end();)
    it = it->second.done ? s->reqs.erase(it) : std::next(it);
  return s->fatal;
}

int32_t segtaco_server_replay(segtaco_server_t s, int32_t n,
                              const double *arrival_us,
                              const int32_t *matrix_ids, const void *const *B,
                              const int64_t *ldb, const int32_t *N,
                              float *const *C, const int64_t *ldc,
                              int64_t *ids_out, double timeout_s) {
  if (!s || n < 0 || (n > 0 && (!arrival_us || !matrix_ids || !B || !ldb || !N || !C || !ldc)))
    return set_error(SEGTACO_ERR_INVALID, "segtaco_server_replay: bad argument");
  int32_t st = SEGTACO_OK;
  std::thread timer([&] {
    const auto t0 = Clock::now();
    for (int32_t i = 0; i < n; ++i) {
      std::this_thread::sleep_until(
          t0 + std::chrono::duration_cast<Clock::duration>(
                   std::chrono::duration<double, std::micro>(arrival_us[i])));
      int64_t id = -1;
      const int32_t r = segtaco_server_submit(s, matrix_ids[i], B[i], ldb[i], N[i], C[i],
                                              ldc[i], &id);
      if (r != SEGTACO_OK && st == SEGTACO_OK) st = r;
      if (ids_out) ids_out[i] = id;
    }
  });
  timer.join();
  const int32_t d = segtaco_server_drain(s, timeout_s);
  return st != SEGTACO_OK ? st : d;
}

int32_t segtaco_server_stats(segtaco_server_t s, segtaco_serve_stats_t *out) {
  if (!s || !out) return set_error(SEGTACO_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->rejected = s->rejected;
  out->ticks = s->ticks;
  out->cache_hits = s->hits;
  out->cache_misses = s->misses;
  out->evictions = s->evictions;
  out->uploaded_bytes = s->uploaded;
  out->mean_batch = s->ticks ? (double)s->batch_products / s->ticks : 0.0;
  out->max_concurrent_ticks = s->max_conc;
  if (!s->lat_us.empty()) {
    std::vector<double> v = s->lat_us;
    std::sort(v.begin(), v.end());
    auto q = [&](double f) {
      const size_t i = (size_t)std::min<double>(v.size() - 1, std::floor(f * (v.size() - 1) + 0.5));
      return v[i];
    };
    out->p50_us = q(0.50);
    out->p99_us = q(0.99);
    out->max_us = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    out->mean_us = sum / v.size();
    out->wall_s = std::chrono::duration<double>(s->last_done - s->first_submit).count();
    out->gflops = out->wall_s > 0 ? s->flops_done / out->wall_s / 1e9 : 0.0;
  }
  return SEGTACO_OK;
}

int32_t segtaco_server_reset_stats(segtaco_server_t s) {
  if (!s) return set_error(SEGTACO_ERR_INVALID, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_us.clear();
  s->submitted = s->completed = s->rejected = s->ticks = 0;
  s->hits = s->misses = s->evictions = s->uploaded = 0;
  s->batch_products = 0;
  s->flops_done = 0.0;
  s->have_first = false;
  s->max_conc = 0;
  return SEGTACO_OK;
}

int32_t segtaco_server_debug_fail_tick(segtaco_server_t s, int64_t ticks_ahead) {
  if (!s || ticks_ahead < 1)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_server_debug_fail_tick: bad argument");
  std::lock_guard<std::mutex> lk(s->mu);
  s->fail_tick = s->ticks + ticks_ahead;
  return SEGTACO_OK;
}

int32_t segtaco_server_plan(segtaco_server_t s, int32_t matrix_id, int32_t N,
                            segtaco_plan_t *out) {
  if (!s || !out) return set_error(SEGTACO_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  if (matrix_id < 0 || matrix_id >= (int32_t)s->mats.size())
    return set_error(SEGTACO_ERR_INVALID, "unknown matrix %d", matrix_id);
  *out = plan_for(s, *s->mats[matrix_id], N);
  return SEGTACO_OK;
}

int32_t segtaco_server_destroy(segtaco_server_t s) {
  if (!s) return SEGTACO_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  for (auto &sl : s->slots) {
    (void)hipStreamSynchronize(sl.stream);
    (void)hipStreamDestroy(sl.stream);
    (void)hipEventDestroy(sl.done);
    (void)hipFree(sl.d_pack);
    (void)hipHostFree(sl.h_pack);
  }
  for (auto &m : s->mats) {
    (void)hipHostFree(m->h_pos);
    if (m->h_crd) (void)hipHostFree(m->h_crd);
    if (m->h_vals) (void)hipHostFree(m->h_vals);
  }
  s->cache.destroy();
  s->inflight.destroy();
  delete s;
  return SEGTACO_OK;
}

}  // extern "C"

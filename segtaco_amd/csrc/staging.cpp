// staging.cpp — pinned-ring copies for the operator surface (staging.h).
#include "staging.h"

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

#include "engine.h"

namespace segtaco {
namespace {

// A fixed set of worker threads that split one memcpy between them and the
// caller. Created once and never destroyed (its threads outlive static
// destruction at exit, so the pool is a leaked heap object).
class CopyPool {
 public:
  explicit CopyPool(int threads) : n_(threads) {
    for (int i = 1; i < n_; ++i) std::thread([this, i] { worker(i); }).detach();
  }

  void copy(void *dst, const void *src, size_t bytes) {
    if (n_ == 1 || bytes < (1u << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> l(mu_);
      dst_ = static_cast<char *>(dst);
      src_ = static_cast<const char *>(src);
      bytes_ = bytes;
      pending_ = n_ - 1;
      ++gen_;
    }
    work_.notify_all();
    part(dst_, src_, bytes, 0);
    std::unique_lock<std::mutex> l(mu_);
    done_.wait(l, [&] { return pending_ == 0; });
  }

 private:
  // Part i of n_: 4 KiB-aligned slices, so threads do not share pages.
  void part(char *dst, const char *src, size_t bytes, int i) const {
    const size_t per = ((bytes + n_ - 1) / n_ + 4095) & ~size_t(4095);
    const size_t b = std::min(bytes, per * (size_t)i);
    const size_t e = std::min(bytes, b + per);
    if (b < e) std::memcpy(dst + b, src + b, e - b);
  }

  void worker(int i) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> l(mu_);
      work_.wait(l, [&] { return gen_ != seen; });
      seen = gen_;
      char *d = dst_;
      const char *s = src_;
      const size_t n = bytes_;
      l.unlock();
      part(d, s, n, i);
      l.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }

  const int n_;
  std::mutex mu_;
  std::condition_variable work_, done_;
  uint64_t gen_ = 0;
  int pending_ = 0;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t bytes_ = 0;
};

CopyPool &pool() {
  static CopyPool *p = [] {
    const unsigned hw = std::thread::hardware_concurrency();
    const int t = (int)std::max(1u, std::min(8u, hw / 2));
    return new CopyPool(t);
  }();
  return *p;
}

}  // namespace

bool Stager::ready() {
  if (failed_) return false;
  if (slot_[0]) return true;
  for (int i = 0; i < kSlots; ++i) {
    if (hipHostMalloc(&slot_[i], kChunk, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&ev_[i], hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      failed_ = true;  // no pinned memory: every copy goes direct
      return false;
    }
  }
  return true;
}

int32_t Stager::claim(int slot) {
  if (busy_[slot]) {
    SEGTACO_HIP(hipEventSynchronize(ev_[slot]));
    busy_[slot] = false;
  }
  return SEGTACO_OK;
}

int32_t Stager::h2d(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes < kMinBytes || !ready()) {
    SEGTACO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    return SEGTACO_OK;
  }
  char *d = static_cast<char *>(dst);
  const char *h = static_cast<const char *>(src);
  for (size_t off = 0; off < bytes; off += kChunk) {
    const size_t n = std::min(kChunk, bytes - off);
    const int k = next_;
    next_ = (next_ + 1) % kSlots;
    SEGTACO_RET(claim(k));
    pool().copy(slot_[k], h + off, n);
    SEGTACO_HIP(hipMemcpyAsync(d + off, slot_[k], n, hipMemcpyHostToDevice, s));
    SEGTACO_HIP(hipEventRecord(ev_[k], s));
    busy_[k] = true;
  }
  return SEGTACO_OK;
}

int32_t Stager::d2h(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes < kMinBytes || !ready()) {
    SEGTACO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    SEGTACO_HIP(hipStreamSynchronize(s));
    return SEGTACO_OK;
  }
  char *h = static_cast<char *>(dst);
  const char *d = static_cast<const char *>(src);
  const size_t nchunks = (bytes + kChunk - 1) / kChunk;
  int slot_of[kSlots];
  auto issue = [&](size_t c) -> int32_t {
    const int k = next_;
    next_ = (next_ + 1) % kSlots;
    SEGTACO_RET(claim(k));
    const size_t off = c * kChunk, n = std::min(kChunk, bytes - off);
    SEGTACO_HIP(hipMemcpyAsync(slot_[k], d + off, n, hipMemcpyDeviceToHost, s));
    SEGTACO_HIP(hipEventRecord(ev_[k], s));
    busy_[k] = true;
    slot_of[c % kSlots] = k;
    return SEGTACO_OK;
  };
  for (size_t c = 0; c < std::min(nchunks, (size_t)kSlots); ++c) SEGTACO_RET(issue(c));
  for (size_t c = 0; c < nchunks; ++c) {
    const int k = slot_of[c % kSlots];
    SEGTACO_RET(claim(k));  // this chunk has landed in the slot
    const size_t off = c * kChunk, n = std::min(kChunk, bytes - off);
    pool().copy(h + off, slot_[k], n);
    if (c + kSlots < nchunks) SEGTACO_RET(issue(c + kSlots));
  }
  return SEGTACO_OK;
}

}  // namespace segtaco

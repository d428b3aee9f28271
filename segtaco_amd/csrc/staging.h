// staging.h — host <-> device copies of pageable host memory through a ring
// of pinned buffers (for the operator surface, whose operands live in the
// caller's ordinary host memory).
//
// A pageable hipMemcpy goes through the runtime's own bounce buffers at about
// half the PCIe rate of a pinned copy (measured on MI355X hosts: 28 GB/s H2D
// and 18 GB/s D2H pageable, 57 GB/s pinned, both directions). Here worker
// threads copy chunk i between the caller's memory and a pinned slot while
// the DMA of the neighbouring chunks is in flight.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace segtaco {

class Stager {
 public:
  static constexpr size_t kChunk = 8u << 20;  // bytes per pinned slot
  static constexpr int kSlots = 4;
  static constexpr size_t kMinBytes = 4u << 20;  // smaller copies go direct

  // dst (device) <- src (host); returns once src has been read in full (the
  // last chunks' DMA may still be in flight on `s`).
  int32_t h2d(void *dst, const void *src, size_t bytes, hipStream_t s);
  // dst (host) <- src (device), after the work already enqueued on `s`;
  // returns once dst holds the data.
  int32_t d2h(void *dst, const void *src, size_t bytes, hipStream_t s);

 private:
  bool ready();             // pinned slots and events allocated (lazily)
  int32_t claim(int slot);  // waits for the slot's last DMA
  bool failed_ = false;
  void *slot_[kSlots] = {};
  hipEvent_t ev_[kSlots] = {};
  bool busy_[kSlots] = {};
  int next_ = 0;
};

}  // namespace segtaco

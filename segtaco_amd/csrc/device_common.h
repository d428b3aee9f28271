// device_common.h — shared device helpers for the gfx950 SpMM kernels.
//
// Storage types: DT 0 = FP32, 1 = BF16, 2 = FP16 (vals and B share the
// storage type; accumulation is always FP32). V is the number of B elements
// one lane loads per gather (one 32/64/128-bit load), so a lane group of Q
// lanes covers a column tile of Q*V columns of one B row with a single
// coalesced request of Q*V*sizeof(elem) bytes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"

namespace segtaco {

#define SEGTACO_WAVE 64

__device__ __forceinline__ float bf16_to_f32(uint32_t h16) {
  return __uint_as_float(h16 << 16);
}
__device__ __forceinline__ float f16_to_f32(uint32_t h16) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)h16);
}

template <int DT>
__device__ __forceinline__ float half_to_f32(uint32_t h16) {
  if constexpr (DT == 1)
    return bf16_to_f32(h16 & 0xffffu);
  else
    return f16_to_f32(h16 & 0xffffu);
}

// One stored A value (index p); vals == nullptr reads as 1.0
// (the reference's __guard_load_default_one, gpu_library.h:154-161).
template <int DT>
__device__ __forceinline__ float load_val(const void *__restrict__ vals,
                                          int64_t p) {
  if (vals == nullptr) return 1.0f;
  if constexpr (DT == 0) {
    return static_cast<const float *>(vals)[p];
  } else {
    return half_to_f32<DT>(static_cast<const uint16_t *>(vals)[p]);
  }
}


// load_val for an entry stream read once (non-temporal: not kept in L2, which
// the gathers of B can use instead).
template <int DT>
__device__ __forceinline__ float load_val_stream(const void *__restrict__ vals, int64_t p) {
  if (vals == nullptr) return 1.0f;
  if constexpr (DT == 0) {
    return __builtin_nontemporal_load(static_cast<const float *>(vals) + p);
  } else {
    return half_to_f32<DT>(__builtin_nontemporal_load(static_cast<const uint16_t *>(vals) + p));
  }
}

// V consecutive elements starting at element index i (i aligned to V).
template <int DT, int V>
__device__ __forceinline__ void load_vec(const void *__restrict__ base,
                                         int64_t i, float (&out)[V]) {
  if constexpr (DT == 0) {
    const float *p = static_cast<const float *>(base) + i;
    if constexpr (V == 4) {
      float4 x = *reinterpret_cast<const float4 *>(p);
      out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = x.w;
    } else if constexpr (V == 2) {
      float2 x = *reinterpret_cast<const float2 *>(p);
      out[0] = x.x; out[1] = x.y;
    } else {
      static_assert(V == 1, "fp32 vector width");
      out[0] = *p;
    }
  } else {
    const uint16_t *p = static_cast<const uint16_t *>(base) + i;
    if constexpr (V == 8) {
      uint4 x = *reinterpret_cast<const uint4 *>(p);
      uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        out[2 * k] = half_to_f32<DT>(w[k]);
        out[2 * k + 1] = half_to_f32<DT>(w[k] >> 16);
      }
    } else if constexpr (V == 4) {
      uint2 x = *reinterpret_cast<const uint2 *>(p);
      out[0] = half_to_f32<DT>(x.x);
      out[1] = half_to_f32<DT>(x.x >> 16);
      out[2] = half_to_f32<DT>(x.y);
      out[3] = half_to_f32<DT>(x.y >> 16);
    } else if constexpr (V == 2) {
      uint32_t x = *reinterpret_cast<const uint32_t *>(p);
      out[0] = half_to_f32<DT>(x);
      out[1] = half_to_f32<DT>(x >> 16);
    } else {
      static_assert(V == 1, "16-bit vector width");
      out[0] = half_to_f32<DT>(*p);
    }
  }
}

// Packed (not yet widened) V-element slice of a B row: what a gather leaves
// in registers while in flight. unpack() widens it to FP32 at use.
template <int DT, int V>
struct RawT;
template <>
struct RawT<0, 4> { using type = float4; };
template <>
struct RawT<0, 2> { using type = float2; };
template <>
struct RawT<0, 1> { using type = float; };
template <int DT>
struct RawT<DT, 8> { using type = uint4; };
template <int DT>
struct RawT<DT, 4> { using type = uint2; };
template <int DT>
struct RawT<DT, 2> { using type = uint32_t; };
template <int DT>
struct RawT<DT, 1> { using type = uint16_t; };

template <int DT, int V>
__device__ __forceinline__ typename RawT<DT, V>::type load_raw(
    const void *__restrict__ base, int64_t i) {
  using R = typename RawT<DT, V>::type;
  if constexpr (DT == 0) {
    return *reinterpret_cast<const R *>(static_cast<const float *>(base) + i);
  } else {
    return *reinterpret_cast<const R *>(static_cast<const uint16_t *>(base) + i);
  }
}

template <int DT, int V>
__device__ __forceinline__ void unpack(const typename RawT<DT, V>::type &x,
                                       float (&out)[V]) {
  if constexpr (DT == 0) {
    if constexpr (V == 4) {
      out[0] = x.x; out[1] = x.y; out[2] = x.z; out[3] = x.w;
    } else if constexpr (V == 2) {
      out[0] = x.x; out[1] = x.y;
    } else {
      out[0] = x;
    }
  } else {
    if constexpr (V == 8) {
      const uint32_t w[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        out[2 * k] = half_to_f32<DT>(w[k]);
        out[2 * k + 1] = half_to_f32<DT>(w[k] >> 16);
      }
    } else if constexpr (V == 4) {
      out[0] = half_to_f32<DT>(x.x);
      out[1] = half_to_f32<DT>(x.x >> 16);
      out[2] = half_to_f32<DT>(x.y);
      out[3] = half_to_f32<DT>(x.y >> 16);
    } else if constexpr (V == 2) {
      out[0] = half_to_f32<DT>(x);
      out[1] = half_to_f32<DT>(x >> 16);
    } else {
      out[0] = half_to_f32<DT>(x);
    }
  }
}

// V elements of row `row` of a row-major matrix (leading dimension ld,
// elements) from column c. kOff32: row, ld < 2^24 and every byte offset of
// the matrix fits 32 bits, so the offset is one 24-bit multiply added to the
// uniform base as a 32-bit byte offset (global_load with a scalar base)
// instead of 64-bit integer arithmetic with quarter-rate multiplies.
template <int DT, int V, bool kOff32>
__device__ __forceinline__ typename RawT<DT, V>::type load_row_raw(
    const void *__restrict__ base, int32_t row, int64_t ld, int32_t c) {
  if constexpr (kOff32) {
    constexpr uint32_t kEb = DT == 0 ? 4 : 2;
    const uint32_t off = (__umul24((uint32_t)row, (uint32_t)ld) + (uint32_t)c) * kEb;
    return *reinterpret_cast<const typename RawT<DT, V>::type *>(
        static_cast<const char *>(base) + off);
  } else {
    return load_raw<DT, V>(base, (int64_t)row * ld + c);
  }
}

template <int DT, int V, bool kOff32>
__device__ __forceinline__ void load_row_vec(const void *__restrict__ base,
                                             int32_t row, int64_t ld, int32_t c,
                                             float (&out)[V]) {
  unpack<DT, V>(load_row_raw<DT, V, kOff32>(base, row, ld, c), out);
}

// Makes this wave's LDS writes visible to all of its lanes: a release/acquire
// fence at wavefront scope around a wave barrier (no block-wide s_barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// V consecutive FP32 values to C (aligned to V floats; V == 8 is two float4).
template <int V>
__device__ __forceinline__ void store_vec(float *__restrict__ p,
                                          const float (&v)[V]) {
  if constexpr (V == 8) {
    reinterpret_cast<float4 *>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4 *>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
  } else if constexpr (V == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (V == 2) {
    *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
  } else {
    p[0] = v[0];
  }
}

// store_vec for data nothing reads back soon (non-temporal: not kept in L2),
// as whole 16- / 8-byte vector stores.
template <int V>
__device__ __forceinline__ void store_vec_nt(float *__restrict__ p, const float (&v)[V]) {
  typedef float f4 __attribute__((ext_vector_type(4)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  if constexpr (V == 8) {
    __builtin_nontemporal_store(f4{v[0], v[1], v[2], v[3]}, reinterpret_cast<f4 *>(p));
    __builtin_nontemporal_store(f4{v[4], v[5], v[6], v[7]}, reinterpret_cast<f4 *>(p) + 1);
  } else if constexpr (V == 4) {
    __builtin_nontemporal_store(f4{v[0], v[1], v[2], v[3]}, reinterpret_cast<f4 *>(p));
  } else if constexpr (V == 2) {
    __builtin_nontemporal_store(f2{v[0], v[1]}, reinterpret_cast<f2 *>(p));
  } else {
    __builtin_nontemporal_store(v[0], p);
  }
}

template <int V>
__device__ __forceinline__ void load_f32vec(const float *__restrict__ p,
                                            float (&v)[V]) {
  if constexpr (V == 8) {
    float4 a = reinterpret_cast<const float4 *>(p)[0];
    float4 b = reinterpret_cast<const float4 *>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else if constexpr (V == 4) {
    float4 a = *reinterpret_cast<const float4 *>(p);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  } else if constexpr (V == 2) {
    float2 a = *reinterpret_cast<const float2 *>(p);
    v[0] = a.x; v[1] = a.y;
  } else {
    v[0] = p[0];
  }
}

template <int V>
__device__ __forceinline__ void zero_vec(float (&v)[V]) {
#pragma unroll
  for (int j = 0; j < V; ++j) v[j] = 0.0f;
}

// Neumaier (compensated) accumulation of the exact-or-rounded product v*b.
// Contraction is disabled so that t is the rounded sum of s and x.
__device__ __forceinline__ void neumaier_add(float &s, float &c, float v,
                                             float b) {
#pragma clang fp contract(off)
  const float x = v * b;
  const float t = s + x;
  if (fabsf(s) >= fabsf(x))
    c += (s - t) + x;
  else
    c += (x - t) + s;
  s = t;
}

// XCD-aware decode of a 1-D grid into (row/tile block, column panel).
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), so the column panels of one row block are given consecutive
// per-XCD slots: they then share that XCD's L2 for the block's pos/crd/vals.
// Returns false for the padding blocks of the last round.
__device__ __forceinline__ bool decode_block(int32_t panels, int32_t nblocks,
                                             int32_t &blk, int32_t &panel) {
  const int32_t b = (int32_t)blockIdx.x;
  if (panels == 1) {
    blk = b;
    panel = 0;
    return blk < nblocks;
  }
  const int32_t xcd = b & 7;
  const int32_t k = b >> 3;
  panel = k % panels;
  blk = xcd + 8 * (k / panels);
  return blk < nblocks;
}

// Grid size that decode_block expects for `nblocks` row blocks.
__host__ __device__ inline int64_t grid_for(int32_t panels, int64_t nblocks) {
  if (panels == 1) return nblocks;
  return 8 * ((nblocks + 7) / 8) * panels;
}

// The reference's taco_binarySearchBefore (gpu_library.h:185-205) — see the
// oracle restatement for the exact tie-break it keeps.
__device__ __forceinline__ int32_t search_before(const int32_t *__restrict__ a,
                                                 int32_t start, int32_t end,
                                                 int64_t target) {
  if ((int64_t)a[end] <= target) return end;
  int32_t lo = start, hi = end;
  while (hi - lo > 1) {
    const int32_t mid = (hi + lo) >> 1;
    const int64_t v = a[mid];
    if (v < target)
      lo = mid;
    else if (v > target)
      hi = mid;
    else
      return mid;
  }
  return lo;
}

// Last row r in [lo, hi] with pos[r] <= p (the non-empty row holding entry p
// when pos[lo] <= p < pos[hi + 1]).
__device__ __forceinline__ int32_t row_of(const int32_t *__restrict__ pos,
                                          int32_t lo, int32_t hi, int32_t p) {
  while (lo < hi) {
    const int32_t mid = (lo + hi + 1) >> 1;
    if (pos[mid] <= p)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

}  // namespace segtaco

// gen.cpp — seeded synthetic CSR / dense operands for the benchmark configs.
//
// Counter-based: every random draw is splitmix64 of (seed, stream, index),
// so the output does not depend on thread count and any consumer (driver,
// benchmark, tests) regenerates identical inputs from (kind, seed, sizes).
//   kind 0 (uniform):    every row has round(mean) entries, columns uniform.
//   kind 1 (power-law):  row length L = min(cap, round(exp(mu + sigma*Z))),
//                        mu = ln(mean) - sigma^2/2 (so E[L] ~= mean before the
//                        cap), Z standard normal; L = 0 (empty rows) allowed;
//                        columns uniform.
//   kind 2 (banded):     row i holds columns max(0, i-bw) .. min(K-1, i+bw).
// Columns are unique and sorted within a row; values are U[-1, 1] drawn in
// double and rounded to FP32.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "engine.h"
#include "segtaco/gen.h"

namespace segtaco {
namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

struct Rng {
  uint64_t state;
  Rng(uint64_t seed, uint64_t stream, uint64_t index)
      : state(splitmix64(splitmix64(seed ^ (stream * 0xd1b54a32d192ed03ull)) ^
                         index)) {}
  uint64_t next() {
    state += 0x9e3779b97f4a7c15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform01() { return (double)(next() >> 11) * 0x1.0p-53; }
  uint32_t below(uint32_t n) {  // unbiased enough for n << 2^64
    return (uint32_t)(((__uint128_t)next() * n) >> 64);
  }
};

constexpr uint64_t kStreamLen = 1, kStreamCol = 2, kStreamVal = 3,
                   kStreamDense = 4;

inline float u11(uint64_t seed, uint64_t stream, uint64_t i) {
  Rng r(seed, stream, i);
  return (float)(2.0 * r.uniform01() - 1.0);
}

int32_t row_len(int32_t kind, uint64_t seed, int32_t i, int32_t K,
                double mean, double sigma, int32_t cap, int32_t bw) {
  if (kind == 0) {
    int64_t L = llround(mean);
    return (int32_t)std::min<int64_t>(std::max<int64_t>(L, 0), K);
  }
  if (kind == 2) {
    const int64_t lo = std::max<int64_t>(0, (int64_t)i - bw);
    const int64_t hi = std::min<int64_t>((int64_t)K - 1, (int64_t)i + bw);
    return hi >= lo ? (int32_t)(hi - lo + 1) : 0;
  }
  Rng r(seed, kStreamLen, (uint64_t)i);
  double u1 = r.uniform01(), u2 = r.uniform01();
  if (u1 < 1e-300) u1 = 1e-300;
  const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  const double mu = std::log(mean) - 0.5 * sigma * sigma;
  double L = std::exp(mu + sigma * z);
  int64_t Li = llround(L);
  Li = std::min<int64_t>(Li, cap);
  Li = std::min<int64_t>(Li, K);
  return (int32_t)std::max<int64_t>(Li, 0);
}

template <class F>
void parallel_rows(int32_t M, F &&f) {
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(nt, 16u));
  if (M < 4096) nt = 1;
  std::vector<std::thread> th;
  const int64_t chunk = ((int64_t)M + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min<int64_t>(M, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([&, lo, hi] { f((int32_t)lo, (int32_t)hi); });
  }
  for (auto &x : th) x.join();
}

// L distinct sorted columns in [0, K) for row i.
void row_columns(uint64_t seed, int32_t i, int32_t K, int32_t L,
                 int32_t *out, std::vector<int32_t> &scratch) {
  Rng r(seed, kStreamCol, (uint64_t)i);
  if ((int64_t)L * 4 > K) {
    scratch.resize(K);
    for (int32_t k = 0; k < K; ++k) scratch[k] = k;
    for (int32_t k = 0; k < L; ++k) {
      const int32_t j = k + (int32_t)r.below((uint32_t)(K - k));
      std::swap(scratch[k], scratch[j]);
    }
    std::copy(scratch.begin(), scratch.begin() + L, out);
    std::sort(out, out + L);
    return;
  }
  int32_t have = 0;
  while (have < L) {
    for (int32_t k = have; k < L; ++k) out[k] = (int32_t)r.below((uint32_t)K);
    std::sort(out, out + L);
    have = (int32_t)(std::unique(out, out + L) - out);
  }
}

}  // namespace
}  // namespace segtaco

using namespace segtaco;

extern "C" {

int64_t segtaco_gen_pos(int32_t kind, uint64_t seed, int32_t M, int32_t K,
                        double mean, double sigma, int32_t cap, int32_t bw,
                        int32_t *pos) {
  if (M < 0 || K < 0 || !pos || kind < 0 || kind > 2) return -1;
  std::vector<int32_t> len(M);
  parallel_rows(M, [&](int32_t lo, int32_t hi) {
    for (int32_t i = lo; i < hi; ++i)
      len[i] = row_len(kind, seed, i, K, mean, sigma, cap, bw);
  });
  int64_t acc = 0;
  pos[0] = 0;
  for (int32_t i = 0; i < M; ++i) {
    acc += len[i];
    if (acc > INT32_MAX) return -1;
    pos[i + 1] = (int32_t)acc;
  }
  return acc;
}

int32_t segtaco_gen_fill(int32_t kind, uint64_t seed, int32_t M, int32_t K,
                         int32_t bw, const int32_t *pos, int32_t *crd,
                         float *vals) {
  if (M < 0 || K < 0 || !pos) return SEGTACO_ERR_INVALID;
  parallel_rows(M, [&](int32_t lo, int32_t hi) {
    std::vector<int32_t> scratch;
    for (int32_t i = lo; i < hi; ++i) {
      const int32_t s = pos[i], L = pos[i + 1] - pos[i];
      if (crd && L > 0) {
        if (kind == 2) {
          const int32_t c0 = std::max(0, i - bw);
          for (int32_t k = 0; k < L; ++k) crd[s + k] = c0 + k;
        } else {
          row_columns(seed, i, K, L, crd + s, scratch);
        }
      }
      if (vals)
        for (int32_t p = s; p < s + L; ++p)
          vals[p] = u11(seed, kStreamVal, (uint64_t)p);
    }
  });
  return SEGTACO_OK;
}

int32_t segtaco_gen_dense(uint64_t seed, int64_t n, float *out) {
  if (n < 0 || (n > 0 && !out)) return SEGTACO_ERR_INVALID;
  const int32_t rows = (int32_t)std::min<int64_t>((n + 4095) / 4096, INT32_MAX);
  parallel_rows(rows, [&](int32_t lo, int32_t hi) {
    const int64_t a = (int64_t)lo * 4096, b = std::min<int64_t>(n, (int64_t)hi * 4096);
    for (int64_t i = a; i < b; ++i) out[i] = u11(seed, kStreamDense, (uint64_t)i);
  });
  return SEGTACO_OK;
}

// FP32 -> BF16 (dtype 1) or FP16 (dtype 2), round to nearest even.
int32_t segtaco_convert_f32(int32_t dtype, int64_t n, const float *in,
                            uint16_t *out) {
  if (n < 0 || (n > 0 && (!in || !out)) || (dtype != 1 && dtype != 2))
    return SEGTACO_ERR_INVALID;
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u;
    std::memcpy(&u, &in[i], 4);
    if (dtype == 1) {
      if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) {
        out[i] = (uint16_t)((u >> 16) | 0x40u);
      } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        out[i] = (uint16_t)(u >> 16);
      }
    } else {
      _Float16 h = (_Float16)in[i];  // IEEE conversion, round to nearest even
      std::memcpy(&out[i], &h, 2);
    }
  }
  return SEGTACO_OK;
}

}  // extern "C"

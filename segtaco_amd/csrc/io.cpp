// io.cpp — MatrixMarket reading, COO -> CSR canonicalisation, the reference's
// value fill and result comparators (host code; see include/segtaco/io.h).
//
// The reader keeps the reference's observable contract (mm_read_banner /
// mm_read_mtx_crd_size, mmio.h:203-310; read_mtx_file, ds.h:183-268): the
// banner's five tokens, '%' comment lines before the size line, entries read
// as whitespace-separated tokens (row, column, then 0/1/2 value tokens by
// field), 1-based -> 0-based, symmetric expansion with de-duplication, sort by
// (row, column). It parses a memory image of the file with a hand-written
// scanner instead of fscanf per token, and reports malformed input as an
// error instead of exiting or silently dropping entries.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "engine.h"
#include "segtaco/io.h"

namespace segtaco {
namespace {

struct Scanner {
  const char *p, *end;
  bool at_end() const { return p >= end; }
  void skip_ws() {
    while (p < end && std::isspace(static_cast<unsigned char>(*p))) ++p;
  }
  // Reads one whitespace-separated token; false at end of input.
  bool token(const char **b, const char **e) {
    skip_ws();
    if (p >= end) return false;
    *b = p;
    while (p < end && !std::isspace(static_cast<unsigned char>(*p))) ++p;
    *e = p;
    return true;
  }
  // Reads a decimal integer token (optional sign); false if the next token is
  // not one or input ended.
  bool integer(int64_t *out) {
    const char *b, *e;
    if (!token(&b, &e)) return false;
    bool neg = false;
    if (b < e && (*b == '-' || *b == '+')) neg = (*b++ == '-');
    if (b == e) return false;
    int64_t v = 0;
    for (; b < e; ++b) {
      if (*b < '0' || *b > '9') return false;
      v = v * 10 + (*b - '0');
      if (v > (int64_t)std::numeric_limits<int32_t>::max() + 1) return false;
    }
    *out = neg ? -v : v;
    return true;
  }
  // Next line (without the newline); false at end of input.
  bool line(std::string *out) {
    if (p >= end) return false;
    const char *b = p;
    while (p < end && *p != '\n') ++p;
    out->assign(b, p);
    if (p < end) ++p;
    return true;
  }
};

bool read_file(const char *path, std::vector<char> *buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf->resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = sz > 0 ? std::fread(buf->data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  buf->resize(got);
  return true;
}

std::string lower(std::string s) {
  for (auto &c : s) c = (char)std::tolower(static_cast<unsigned char>(c));
  return s;
}

// Banner line: "%%MatrixMarket matrix coordinate real general".
int32_t parse_banner(const std::string &line, segtaco_mtx_info_t *info) {
  char t[5][1025];
  if (std::sscanf(line.c_str(), "%1024s %1024s %1024s %1024s %1024s", t[0], t[1],
                  t[2], t[3], t[4]) != 5)
    return set_error(SEGTACO_ERR_INVALID, "mtx: incomplete banner line");
  if (std::strncmp(t[0], "%%MatrixMarket", 14) != 0)
    return set_error(SEGTACO_ERR_INVALID, "mtx: no %%%%MatrixMarket banner");
  const std::string obj = lower(t[1]), fmt = lower(t[2]), field = lower(t[3]),
                    sym = lower(t[4]);
  if (obj != "matrix")
    return set_error(SEGTACO_ERR_INVALID, "mtx: unsupported object '%s'", obj.c_str());
  if (fmt == "coordinate")
    info->coordinate = 1;
  else if (fmt == "array")
    info->coordinate = 0;
  else
    return set_error(SEGTACO_ERR_INVALID, "mtx: unsupported format '%s'", fmt.c_str());
  if (field == "real")
    info->field = 0;
  else if (field == "complex")
    info->field = 1;
  else if (field == "pattern")
    info->field = 2;
  else if (field == "integer")
    info->field = 3;
  else
    return set_error(SEGTACO_ERR_INVALID, "mtx: unsupported field '%s'", field.c_str());
  if (sym == "general")
    info->symmetry = 0;
  else if (sym == "symmetric")
    info->symmetry = 1;
  else if (sym == "hermitian")
    info->symmetry = 2;
  else if (sym == "skew-symmetric")
    info->symmetry = 3;
  else
    return set_error(SEGTACO_ERR_INVALID, "mtx: unsupported symmetry '%s'", sym.c_str());
  if (!info->coordinate)
    return set_error(SEGTACO_ERR_INVALID, "mtx: dense array files are not sparse input");
  return SEGTACO_OK;
}

inline uint64_t key_of(int32_t r, int32_t c) {
  return ((uint64_t)(uint32_t)r << 32) | (uint32_t)c;
}

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

}  // namespace
}  // namespace segtaco

using namespace segtaco;

extern "C" {

int32_t segtaco_mtx_read(const char *path, int32_t flags, int32_t *nrow,
                         int32_t *ncol, int32_t *nnz, int32_t **rows,
                         int32_t **cols, segtaco_mtx_info_t *info_out) {
  if (!path || !nrow || !ncol || !nnz || !rows || !cols)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_mtx_read: null argument");
  *rows = *cols = nullptr;
  *nrow = *ncol = *nnz = 0;
  std::vector<char> buf;
  if (!read_file(path, &buf))
    return set_error(SEGTACO_ERR_INVALID, "mtx: file %s not found", path);
  Scanner sc{buf.data(), buf.data() + buf.size()};
  segtaco_mtx_info_t info{};
  std::string line;
  if (!sc.line(&line)) return set_error(SEGTACO_ERR_INVALID, "mtx: empty file");
  SEGTACO_RET(parse_banner(line, &info));
  // size line: skip '%' lines; "M N entries" on one line, else keep reading
  // integer triples (mm_read_mtx_crd_size's fscanf loop).
  do {
    if (!sc.line(&line)) return set_error(SEGTACO_ERR_INVALID, "mtx: no size line");
  } while (!line.empty() && line[0] == '%');
  long long m = 0, n = 0, e = 0;
  if (std::sscanf(line.c_str(), "%lld %lld %lld", &m, &n, &e) != 3) {
    int64_t v[3];
    if (!sc.integer(&v[0]) || !sc.integer(&v[1]) || !sc.integer(&v[2]))
      return set_error(SEGTACO_ERR_INVALID, "mtx: bad size line");
    m = v[0], n = v[1], e = v[2];
  }
  if (m < 0 || n < 0 || e < 0 || m > INT32_MAX || n > INT32_MAX || e > INT32_MAX)
    return set_error(SEGTACO_ERR_INVALID, "mtx: size %lld x %lld with %lld entries out of range",
                     m, n, e);
  info.entries = (int32_t)e;
  const int values = info.field == 2 ? 0 : (info.field == 1 ? 2 : 1);
  // The header's entry count is not trusted for the reservation: every entry
  // takes at least 4 bytes of text ("1 1\n"), so the bytes left bound it.
  // The vector may still grow past what the host can give: that is a status,
  // not an exception escaping the C ABI.
  std::vector<uint64_t> keys;
  const size_t left_entries = (size_t)(sc.end - sc.p) / 4 + 1;
  const size_t want = (size_t)e < left_entries ? (size_t)e : left_entries;
  try {
    keys.reserve(info.symmetry == 1 ? 2 * want : want);
  } catch (const std::exception &) {
    return set_error(SEGTACO_ERR_NOMEM, "mtx: out of host memory");
  }
  for (long long i = 0; i < e; ++i) {
    int64_t r, c;
    if (!sc.integer(&r) || !sc.integer(&c))
      return set_error(SEGTACO_ERR_INVALID, "mtx: entry %lld of %lld missing or malformed",
                       i + 1, e);
    for (int v = 0; v < values; ++v) {
      const char *b, *x;
      if (!sc.token(&b, &x))
        return set_error(SEGTACO_ERR_INVALID, "mtx: entry %lld has no value", i + 1);
    }
    if (r < 1 || r > m || c < 1 || c > n)
      return set_error(SEGTACO_ERR_INVALID, "mtx: entry %lld (%lld, %lld) outside %lld x %lld",
                       i + 1, (long long)r, (long long)c, m, n);
    try {
      keys.push_back(key_of((int32_t)(r - 1), (int32_t)(c - 1)));
      if (info.symmetry == 1 && r != c) {
        // the mirrored entry must fit too (a non-square "symmetric" file)
        if (c > m || r > n)
          return set_error(SEGTACO_ERR_INVALID,
                           "mtx: entry %lld (%lld, %lld) mirrors outside %lld x %lld", i + 1,
                           (long long)r, (long long)c, m, n);
        keys.push_back(key_of((int32_t)(c - 1), (int32_t)(r - 1)));
      }
    } catch (const std::exception &) {
      return set_error(SEGTACO_ERR_NOMEM, "mtx: out of host memory");
    }
  }
  std::sort(keys.begin(), keys.end());
  if (info.symmetry == 1) {
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    if ((flags & SEGTACO_MTX_REFERENCE_TRUNCATE) && keys.size() > (size_t)e)
      keys.resize((size_t)e);
  }
  const size_t cnt = keys.size();
  if (cnt > (size_t)INT32_MAX)
    return set_error(SEGTACO_ERR_INVALID, "mtx: %zu entries exceed int32", cnt);
  int32_t *R = (int32_t *)std::malloc(sizeof(int32_t) * (cnt + 1));
  int32_t *C = (int32_t *)std::malloc(sizeof(int32_t) * (cnt + 1));
  if (!R || !C) {
    std::free(R);
    std::free(C);
    return set_error(SEGTACO_ERR_NOMEM, "mtx: out of host memory");
  }
  for (size_t i = 0; i < cnt; ++i) {
    R[i] = (int32_t)(keys[i] >> 32);
    C[i] = (int32_t)(keys[i] & 0xffffffffu);
  }
  *nrow = (int32_t)m;
  *ncol = (int32_t)n;
  *nnz = (int32_t)cnt;
  *rows = R;
  *cols = C;
  if (info_out) *info_out = info;
  return SEGTACO_OK;
}

void segtaco_fill_reference_vals(uint64_t seed, int64_t n, float *vals) {
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t r = splitmix64(seed * 0x9e3779b97f4a7c15ull + (uint64_t)i);
    vals[i] = (float)(int)(r % 3) / 10;
  }
}

int32_t segtaco_coo_to_csr(int32_t m, int32_t n, int32_t nnz,
                           const int32_t *rows, const int32_t *cols,
                           const float *vals, int32_t *pos, int32_t *crd,
                           float *out_vals, int32_t *out_nnz) {
  if (m < 0 || n < 0 || nnz < 0 || !pos || !out_nnz ||
      (nnz > 0 && (!rows || !cols || !crd)))
    return set_error(SEGTACO_ERR_INVALID, "segtaco_coo_to_csr: bad argument");
  std::vector<int64_t> start((size_t)m + 1, 0);
  for (int32_t i = 0; i < nnz; ++i) {
    if (rows[i] < 0 || rows[i] >= m || cols[i] < 0 || cols[i] >= n)
      return set_error(SEGTACO_ERR_INVALID, "segtaco_coo_to_csr: entry %d (%d, %d) outside %dx%d",
                       i, rows[i], cols[i], m, n);
    ++start[(size_t)rows[i] + 1];
  }
  for (int32_t r = 0; r < m; ++r) start[r + 1] += start[r];
  // stable bucket by row, then stable sort each row by column
  std::vector<int32_t> order((size_t)nnz);
  {
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int32_t i = 0; i < nnz; ++i) order[(size_t)fill[rows[i]]++] = i;
  }
  int32_t out = 0;
  pos[0] = 0;
  for (int32_t r = 0; r < m; ++r) {
    auto b = order.begin() + start[r], e = order.begin() + start[r + 1];
    std::stable_sort(b, e, [&](int32_t x, int32_t y) { return cols[x] < cols[y]; });
    for (auto it = b; it != e;) {
      const int32_t c = cols[*it];
      float v = vals ? vals[*it] : 1.0f;
      for (++it; it != e && cols[*it] == c; ++it)
        if (vals) v = v + vals[*it];
      crd[out] = c;
      if (out_vals) out_vals[out] = v;
      ++out;
    }
    pos[r + 1] = out;
  }
  *out_nnz = out;
  return SEGTACO_OK;
}

int32_t segtaco_mtx_read_csr(const char *path, int32_t flags, uint64_t seed,
                             int32_t *M, int32_t *K, int32_t *nnz,
                             int32_t **pos, int32_t **crd, float **vals) {
  if (!M || !K || !nnz || !pos || !crd || !vals)
    return set_error(SEGTACO_ERR_INVALID, "segtaco_mtx_read_csr: null argument");
  *pos = *crd = nullptr;
  *vals = nullptr;
  int32_t m, n, cnt, *R, *C;
  SEGTACO_RET(segtaco_mtx_read(path, flags, &m, &n, &cnt, &R, &C, nullptr));
  std::vector<float> v((size_t)cnt);
  if (seed == UINT64_MAX)
    std::fill(v.begin(), v.end(), 1.0f);
  else
    segtaco_fill_reference_vals(seed, cnt, v.data());
  int32_t *P = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)m + 1));
  int32_t *D = (int32_t *)std::malloc(sizeof(int32_t) * ((size_t)cnt + 1));
  float *V = (float *)std::malloc(sizeof(float) * ((size_t)cnt + 1));
  int32_t st = (P && D && V) ? SEGTACO_OK
                             : set_error(SEGTACO_ERR_NOMEM, "mtx: out of host memory");
  int32_t out = 0;
  if (st == SEGTACO_OK) st = segtaco_coo_to_csr(m, n, cnt, R, C, v.data(), P, D, V, &out);
  std::free(R);
  std::free(C);
  if (st != SEGTACO_OK) {
    std::free(P);
    std::free(D);
    std::free(V);
    return st;
  }
  *M = m;
  *K = n;
  *nnz = out;
  *pos = P;
  *crd = D;
  *vals = V;
  return SEGTACO_OK;
}

void segtaco_host_free(void *p) { std::free(p); }

float segtaco_compare_array(const float *x, const float *y, size_t n) {
  float ret = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    if (x[i] != 0.0f) {
      const float diff = (float)std::fabs((double)(y[i] / x[i]) - 1.0);
      if (diff > ret) ret = diff;
    } else if (y[i] != 0.0f) {
      return std::numeric_limits<float>::infinity();
    }
  }
  return ret;
}

float segtaco_compare_array_my(const float *x, const float *y, size_t n) {
  const float thres = 0.001f;
  float count = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    if (std::fabs(x[i] - y[i]) > thres) count += 1.0f;
    if (count > 10000.0f) break;
  }
  return count;
}

float segtaco_compare_matrices(const taco_tensor_t *a, const taco_tensor_t *b) {
  if (!a || !b || !a->dimensions || !b->dimensions ||
      a->dimensions[0] != b->dimensions[0] || a->dimensions[1] != b->dimensions[1])
    return std::numeric_limits<float>::infinity();
  const size_t n = (size_t)a->dimensions[0] * (size_t)a->dimensions[1];
  return segtaco_compare_array_my(a->vals, b->vals, n);
}

}  // extern "C"

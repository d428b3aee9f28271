// Host-only fuzz driver for the MatrixMarket reader and the COO -> CSR
// canonicaliser (segtaco_amd/csrc/io.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_io_asan.py. Every input is either
// accepted with a structurally valid CSR or rejected with an error status;
// a memory or UB error aborts the run.
//
//   mtx_fuzz <iterations> <seed> <scratch dir>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "segtaco/io.h"

namespace {

const char *kBanners[] = {
    "%%MatrixMarket matrix coordinate pattern general",
    "%%MatrixMarket matrix coordinate real general",
    "%%MatrixMarket matrix coordinate integer symmetric",
    "%%MatrixMarket matrix coordinate complex hermitian",
    "%%MatrixMarket matrix coordinate real skew-symmetric",
    "%%MatrixMarket MATRIX Coordinate Real Symmetric",
    "%%MatrixMarket matrix array real general",
    "%MatrixMarket matrix coordinate real general",
    "%%MatrixMarket vector coordinate real general",
    "%%MatrixMarket matrix coordinate real",
};

std::string valid_file(std::mt19937_64 &rng) {
  const int b = (int)(rng() % 6);
  const int m = 1 + (int)(rng() % 40), n = 1 + (int)(rng() % 40);
  const int e = (int)(rng() % 120);
  std::string s = std::string(kBanners[b]) + "\n% comment\n";
  if (rng() % 4 == 0) s += "\n";
  s += std::to_string(m) + " " + std::to_string(n) + " " + std::to_string(e) + "\n";
  const int vals = b == 0 ? 0 : (b == 3 ? 2 : 1);
  const bool sym = b == 2 || b == 3 || b == 4 || b == 5;
  for (int i = 0; i < e; ++i) {
    int r = 1 + (int)(rng() % m), c = 1 + (int)(rng() % n);
    if (sym && c > r) std::swap(r, c);
    if (sym && r > n) r = n;  // symmetric files are square in practice
    s += std::to_string(r) + " " + std::to_string(c);
    for (int v = 0; v < vals; ++v) s += " " + std::to_string((int)(rng() % 2001) - 1000) + ".5";
    s += "\n";
  }
  return s;
}

std::string mutate(std::string s, std::mt19937_64 &rng) {
  const int edits = 1 + (int)(rng() % 6);
  const char alphabet[] = "0123456789 -+.\n%eE\t aZ\x01\xff";
  for (int k = 0; k < edits && !s.empty(); ++k) {
    const size_t at = rng() % (s.size() + 1);
    switch (rng() % 7) {
      case 0: s.erase(at, 1 + rng() % 8); break;
      case 1: s.insert(at, 1, alphabet[rng() % (sizeof(alphabet) - 1)]); break;
      case 2: s.insert(at, "2147483648"); break;
      case 3: s.insert(at, "-1 "); break;
      case 4: s.resize(at); break;
      case 5: s.insert(at, std::string(1100 + rng() % 200, '7')); break;
      default: s.insert(at, std::string(kBanners[rng() % 10]) + "\n"); break;
    }
  }
  return s;
}

void check_csr(int32_t M, int32_t K, int32_t nnz, const int32_t *pos, const int32_t *crd) {
  if (pos[0] != 0 || pos[M] != nnz) std::abort();
  for (int32_t r = 0; r < M; ++r) {
    if (pos[r] > pos[r + 1]) std::abort();
    for (int32_t p = pos[r]; p < pos[r + 1]; ++p) {
      if (crd[p] < 0 || crd[p] >= K) std::abort();
      if (p > pos[r] && crd[p] <= crd[p - 1]) std::abort();  // sorted, unique
    }
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  const long iters = std::atol(argv[1]);
  std::mt19937_64 rng(std::strtoull(argv[2], nullptr, 10));
  const std::string path = std::string(argv[3]) + "/fuzz.mtx";
  long accepted = 0, rejected = 0;
  for (long it = 0; it < iters; ++it) {
    std::string text = valid_file(rng);
    if (rng() % 3) text = mutate(text, rng);
    if (it % 50 == 7) {
      // a short file whose header declares about 2^31 entries: rejected
      // (entries missing) without reserving memory for the declared count
      text = std::string(kBanners[rng() % 2 ? 2 : 1]) + "\n10 10 2147483000\n1 1 0.5\n2 1 1.5\n";
    }
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return 3;
    std::fwrite(text.data(), 1, text.size(), f);
    std::fclose(f);
    const int32_t flags = (int32_t)(rng() % 2);
    int32_t m = -1, n = -1, cnt = -1, *R = nullptr, *C = nullptr;
    segtaco_mtx_info_t info;
    if (segtaco_mtx_read(path.c_str(), flags, &m, &n, &cnt, &R, &C, &info) == SEGTACO_OK) {
      ++accepted;
      for (int32_t i = 0; i < cnt; ++i)
        if (R[i] < 0 || R[i] >= m || C[i] < 0 || C[i] >= n) std::abort();
      segtaco_host_free(R);
      segtaco_host_free(C);
      int32_t M, K, nnz, *P, *D;
      float *V;
      if (segtaco_mtx_read_csr(path.c_str(), flags, it, &M, &K, &nnz, &P, &D, &V) != SEGTACO_OK)
        std::abort();
      check_csr(M, K, nnz, P, D);
      segtaco_host_free(P);
      segtaco_host_free(D);
      segtaco_host_free(V);
    } else {
      ++rejected;
    }
    // COO -> CSR on random triplets, sometimes out of range
    const int32_t mm = 1 + (int32_t)(rng() % 30), nn = 1 + (int32_t)(rng() % 30);
    const int32_t k = (int32_t)(rng() % 200);
    std::vector<int32_t> rr(k), cc(k), pos(mm + 1), crd(k + 1);
    std::vector<float> vv(k), ov(k + 1);
    for (int32_t i = 0; i < k; ++i) {
      rr[i] = (int32_t)(rng() % mm) - (rng() % 50 == 0 ? mm : 0);
      cc[i] = (int32_t)(rng() % nn) + (rng() % 50 == 0 ? nn : 0);
      vv[i] = (float)(rng() % 7) - 3.0f;
    }
    int32_t out = 0;
    if (segtaco_coo_to_csr(mm, nn, k, rr.data(), cc.data(), vv.data(), pos.data(), crd.data(),
                           ov.data(), &out) == SEGTACO_OK)
      check_csr(mm, nn, out, pos.data(), crd.data());
  }
  std::printf("accepted %ld rejected %ld\n", accepted, rejected);
  return 0;
}

// matrix_main.cpp — the sweep driver (matrix-eval-gpu).
//
// Same command line and CSV rows as the reference's driver (matrix_main.cpp:
// 13-57, spmm_csr_gpu.h:129-344, timers.h:93-137):
//
//   matrix-eval-gpu <matrix_dir> <matrix_name> <experiment> <results_dir> <hw>
//                   [--cols N] [--trials T] [--seed S] [--random-b]
//                   [--synthetic KIND:M:MEAN] [--reference-truncate] [--summary]
//
// reads <matrix_dir>/<matrix_name>.mtx, builds A (CSR, values in
// {0, 0.1, 0.2} as fill_random gives, from a seeded stream), B = ones(K x N)
// and C (M x N), and for experiment "spmm_csr_gpu" runs the reference's 61
// sweep rows (60 tuned operator presets + the vendor-library row) `trials`
// times each, appending one line per trial,
//
//   kernel,platform,format,library,tensor,time_ms
//
// to <results_dir>/<experiment>_<N>_<hw>.csv (a header line is written when
// the file is new). Differences from the reference, each deliberate:
//   * inputs are reproducible (--seed) instead of seeded from the clock;
//   * --random-b fills B with U[-1, 1] (seeded) instead of ones, so a column
//     permutation cannot hide behind identical columns;
//   * "spmm_csr_gpu_verify" runs every operator-surface entry point once and
//     prints its mismatch count against the vendor-library row
//     (compare_matrices), as the reference's disabled verify branch does
//     against cuSPARSE (spmm_csr_gpu.h:152-272);
//   * --synthetic uniform|powerlaw|banded:M:MEAN builds A with the seeded
//     generator instead of reading a file (M x M, mean entries per row;
//     banded uses MEAN as the half bandwidth; values U[-1, 1]);
//   * --summary also appends one line per matrix with each row's minimum
//     over the trials (resana.py:150-180's 61-column table) to
//     <results_dir>/<experiment>_<N>_<hw>_summary.csv, and one with each
//     row's mean over the trials rounded to 4 decimals (resana.py:28-33)
//     to <results_dir>/<experiment>_<N>_<hw>_summary_mean.csv.
// Only the C-ABI headers are included: no HIP header (reference: timers.h:90
// pulls in cuda_runtime.h).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "segtaco/gen.h"
#include "segtaco/io.h"
#include "segtaco/spmm.h"

namespace {

using EntryFn = void (*)(taco_tensor_t *, taco_tensor_t *, taco_tensor_t *);

struct SweepRow {
  EntryFn fn;
  const char *platform, *format, *library;
};

// The sweep's rows in the reference's order (spmm_csr_gpu.h:283-343).
const SweepRow kSweep[] = {
    {spmm_rb_pr_256_1_4_4, "rb", "pr", "tune0"},   {spmm_rb_pr_256_2_4_4, "rb", "pr", "tune1"},
    {spmm_rb_pr_256_4_4_4, "rb", "pr", "tune2"},   {spmm_rb_pr_256_1_8_4, "rb", "pr", "tune3"},
    {spmm_rb_pr_256_2_8_4, "rb", "pr", "tune4"},   {spmm_rb_pr_256_4_8_4, "rb", "pr", "tune5"},
    {spmm_rb_pr_256_1_8_8, "rb", "pr", "tune6"},   {spmm_rb_pr_256_2_8_8, "rb", "pr", "tune7"},
    {spmm_rb_pr_256_4_8_8, "rb", "pr", "tune8"},   {spmm_rb_pr_256_1_16_4, "rb", "pr", "tune9"},
    {spmm_rb_pr_256_2_16_4, "rb", "pr", "tune10"}, {spmm_rb_pr_256_4_16_4, "rb", "pr", "tune11"},
    {spmm_rb_pr_256_1_16_8, "rb", "pr", "tune12"}, {spmm_rb_pr_256_2_16_8, "rb", "pr", "tune13"},
    {spmm_rb_pr_256_4_16_8, "rb", "pr", "tune14"}, {spmm_rb_pr_256_1_16_16, "rb", "pr", "tune15"},
    {spmm_rb_pr_256_2_16_16, "rb", "pr", "tune16"}, {spmm_rb_pr_256_4_16_16, "rb", "pr", "tune17"},
    {spmm_rb_pr_256_1_32_4, "rb", "pr", "tune18"}, {spmm_rb_pr_256_2_32_4, "rb", "pr", "tune19"},
    {spmm_rb_pr_256_4_32_4, "rb", "pr", "tune20"}, {spmm_rb_pr_256_1_32_8, "rb", "pr", "tune21"},
    {spmm_rb_pr_256_2_32_8, "rb", "pr", "tune22"}, {spmm_rb_pr_256_4_32_8, "rb", "pr", "tune23"},
    {spmm_rb_pr_256_1_32_16, "rb", "pr", "tune24"}, {spmm_rb_pr_256_2_32_16, "rb", "pr", "tune25"},
    {spmm_rb_pr_256_4_32_16, "rb", "pr", "tune26"}, {spmm_rb_pr_256_1_32_32, "rb", "pr", "tune27"},
    {spmm_rb_pr_256_2_32_32, "rb", "pr", "tune28"}, {spmm_rb_pr_256_4_32_32, "rb", "pr", "tune29"},
    {spmm_eb_pr_256_32, "eb", "pr", "tune0"},      {spmm_eb_pr_256_16, "eb", "pr", "tune1"},
    {spmm_eb_pr_256_8, "eb", "pr", "tune2"},       {spmm_eb_pr_256_4, "eb", "pr", "tune3"},
    {spmm_eb_pr_128_32, "eb", "pr", "tune4"},      {spmm_eb_pr_128_16, "eb", "pr", "tune5"},
    {spmm_eb_pr_128_8, "eb", "pr", "tune6"},       {spmm_eb_pr_128_4, "eb", "pr", "tune7"},
    {spmm_eb_pr_64_32, "eb", "pr", "tune8"},       {spmm_eb_pr_64_16, "eb", "pr", "tune9"},
    {spmm_eb_pr_64_8, "eb", "pr", "tune10"},       {spmm_eb_pr_64_4, "eb", "pr", "tune11"},
    {spmm_rb_sr_256_1_1, "rb", "sr", "tune0"},     {spmm_rb_sr_256_1_2, "rb", "sr", "tune1"},
    {spmm_rb_sr_256_1_4, "rb", "sr", "tune2"},     {spmm_rb_sr_256_2_1, "rb", "sr", "tune3"},
    {spmm_rb_sr_256_2_2, "rb", "sr", "tune4"},     {spmm_rb_sr_256_2_4, "rb", "sr", "tune5"},
    {spmm_rb_sr_256_4_1, "rb", "sr", "tune6"},     {spmm_rb_sr_256_4_2, "rb", "sr", "tune7"},
    {spmm_rb_sr_256_4_4, "rb", "sr", "tune8"},     {spmm_eb_sr_256_4_1, "eb", "sr", "tune0"},
    {spmm_eb_sr_256_4_2, "eb", "sr", "tune1"},     {spmm_eb_sr_256_4_4, "eb", "sr", "tune2"},
    {spmm_eb_sr_256_8_1, "eb", "sr", "tune3"},     {spmm_eb_sr_256_8_2, "eb", "sr", "tune4"},
    {spmm_eb_sr_256_8_4, "eb", "sr", "tune5"},     {spmm_eb_sr_256_16_1, "eb", "sr", "tune6"},
    {spmm_eb_sr_256_16_2, "eb", "sr", "tune7"},    {spmm_eb_sr_256_16_4, "eb", "sr", "tune8"},
    {spmm_csr_cusparse_row, "eb", "sr", "alg2"},
};
constexpr int kSweepRows = sizeof(kSweep) / sizeof(kSweep[0]);
static_assert(kSweepRows == 61, "the reference's sweep has 61 rows");

struct Options {
  std::string matrix_dir = ".", matrix_name = "192bit", experiment = "spmm_csr_gpu";
  std::string results_dir = ".", hardware = "bench";
  int32_t cols = 4, trials = 25;
  uint64_t seed = 1;
  bool random_b = false, truncate = false, summary = false;
  std::string synthetic;
};

[[noreturn]] void usage(const char *msg) {
  std::fprintf(stderr,
               "error: %s\nusage: matrix-eval-gpu <matrix_dir> <matrix_name> <experiment> "
               "<results_dir> <hw> [--cols N] [--trials T] [--seed S] [--random-b] "
               "[--synthetic KIND:M:MEAN] [--reference-truncate] [--summary]\n",
               msg);
  std::exit(2);
}

Options parse(int argc, char **argv) {
  Options o;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto value = [&](const char *flag) -> std::string {
      if (i + 1 >= argc) usage((std::string(flag) + " needs a value").c_str());
      return argv[++i];
    };
    if (a == "--cols") o.cols = std::atoi(value("--cols").c_str());
    else if (a == "--trials") o.trials = std::atoi(value("--trials").c_str());
    else if (a == "--seed") o.seed = std::strtoull(value("--seed").c_str(), nullptr, 10);
    else if (a == "--random-b") o.random_b = true;
    else if (a == "--reference-truncate") o.truncate = true;
    else if (a == "--summary") o.summary = true;
    else if (a == "--synthetic") o.synthetic = value("--synthetic");
    else if (a.rfind("--", 0) == 0) usage(("unknown option " + a).c_str());
    else pos.push_back(a);
  }
  if (pos.size() > 0) o.matrix_dir = pos[0];
  if (pos.size() > 1) o.matrix_name = pos[1];
  if (pos.size() > 2) o.experiment = pos[2];
  if (pos.size() > 3) o.results_dir = pos[3];
  if (pos.size() > 4) o.hardware = pos[4];
  if (o.cols < 1 || o.trials < 1) usage("--cols and --trials must be >= 1");
  return o;
}

// Operands with the taco_tensor_t views the entry points take
// (to_taco_tensor, ds.h:411-444).
struct Operands {
  int32_t M = 0, K = 0, nnz = 0, N = 0;
  std::vector<int32_t> pos, crd;
  std::vector<float> vals, B, C;
  int32_t dimA[2], dimB[2], dimC[2], ord[2] = {0, 1};
  int32_t *idx1[2];
  int32_t **idx[2];
  taco_tensor_t tA{}, tB{}, tC{};

  void views() {
    dimA[0] = M, dimA[1] = K;
    dimB[0] = K, dimB[1] = N;
    dimC[0] = M, dimC[1] = N;
    idx1[0] = pos.data();
    idx1[1] = crd.data();
    idx[0] = nullptr;
    idx[1] = idx1;
    tA.order = tB.order = tC.order = 2;
    tA.dimensions = dimA;
    tA.indices = idx;
    tA.vals = vals.data();
    tA.vals_size = nnz;
    tB.dimensions = dimB;
    tB.mode_ordering = ord;
    tB.vals = B.data();
    tB.vals_size = (int32_t)B.size();
    tC.dimensions = dimC;
    tC.mode_ordering = ord;
    tC.vals = C.data();
    tC.vals_size = (int32_t)C.size();
  }
};

bool load_operands(const Options &o, Operands *op) {
  if (!o.synthetic.empty()) {
    char kind[32] = {0};
    long long m = 0;
    double mean = 0;
    if (std::sscanf(o.synthetic.c_str(), "%31[^:]:%lld:%lf", kind, &m, &mean) != 3 || m < 1 ||
        m > INT32_MAX)
      usage("--synthetic expects KIND:M:MEAN");
    int32_t k = !std::strcmp(kind, "uniform")    ? SEGTACO_GEN_UNIFORM
                : !std::strcmp(kind, "powerlaw") ? SEGTACO_GEN_POWERLAW
                : !std::strcmp(kind, "banded")   ? SEGTACO_GEN_BANDED
                                                 : -1;
    if (k < 0) usage("--synthetic KIND is uniform, powerlaw or banded");
    op->M = op->K = (int32_t)m;
    op->pos.resize((size_t)m + 1);
    const int32_t bw = k == SEGTACO_GEN_BANDED ? (int32_t)mean : 0;
    const int64_t nnz = segtaco_gen_pos(k, o.seed, op->M, op->K, mean, 1.5, 8192, bw,
                                        op->pos.data());
    if (nnz < 0 || nnz > INT32_MAX) {
      std::fprintf(stderr, "synthetic generation failed\n");
      return false;
    }
    op->nnz = (int32_t)nnz;
    op->crd.resize((size_t)nnz + 1);
    op->vals.resize((size_t)nnz + 1);
    if (segtaco_gen_fill(k, o.seed, op->M, op->K, bw, op->pos.data(), op->crd.data(),
                         op->vals.data()) != SEGTACO_OK) {
      std::fprintf(stderr, "synthetic fill failed: %s\n", segtaco_last_error_string());
      return false;
    }
    return true;
  }
  const std::string path = o.matrix_dir + "/" + o.matrix_name + ".mtx";
  int32_t *P = nullptr, *D = nullptr;
  float *V = nullptr;
  if (segtaco_mtx_read_csr(path.c_str(), o.truncate ? SEGTACO_MTX_REFERENCE_TRUNCATE : 0,
                           o.seed, &op->M, &op->K, &op->nnz, &P, &D, &V) != SEGTACO_OK) {
    std::fprintf(stderr, "%s\n", segtaco_last_error_string());
    return false;
  }
  op->pos.assign(P, P + op->M + 1);
  op->crd.assign(D, D + op->nnz);
  op->vals.assign(V, V + op->nnz);
  op->crd.push_back(0);  // keep data() valid for an empty matrix
  op->vals.push_back(0);
  segtaco_host_free(P);
  segtaco_host_free(D);
  segtaco_host_free(V);
  return true;
}

bool file_is_empty(const std::string &path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  return !f || f.tellg() == 0;
}

}  // namespace

int main(int argc, char **argv) {
  const Options o = parse(argc, argv);
  segtaco_set_error_mode(SEGTACO_ERRMODE_EXIT);
  Operands op;
  if (!load_operands(o, &op)) return 1;
  op.N = o.cols;
  op.B.assign((size_t)op.K * op.N, 1.0f);  // gen_row_major_matrix: ones (ds.h:282-285)
  if (o.random_b) {
    if (segtaco_gen_dense(o.seed ^ 0x5bd1e995ull, (int64_t)op.B.size(), op.B.data()) != SEGTACO_OK)
      return 1;
  }
  op.C.assign((size_t)op.M * op.N, 1.0f);
  op.views();
  std::cout << "nnz: " << op.nnz << " K: " << op.N << " nnz*K: " << (long long)op.nnz * op.N
            << std::endl;

  if (o.experiment == "spmm_csr_gpu_verify") {
    std::vector<float> ref((size_t)op.M * op.N);
    op.tC.vals = ref.data();
    spmm_csr_cusparse_row(&op.tC, &op.tA, &op.tB);
    taco_tensor_t tref = op.tC;
    op.tC.vals = op.C.data();
    int bad = 0;
    for (int32_t id = 0; id < segtaco_legacy_count(); ++id) {
      std::fill(op.C.begin(), op.C.end(), -7.0f);
      segtaco_legacy_call(id, &op.tC, &op.tA, &op.tB);
      const float mism = segtaco_compare_matrices(&tref, &op.tC);
      std::cout << segtaco_legacy_name(id) << " vs rocsparse: " << mism << std::endl;
      bad += mism != 0.0f;
    }
    return bad ? 3 : 0;
  }
  if (o.experiment != "spmm_csr_gpu") {
    // the reference's other experiments are disabled (matrix_main.cpp:38-49)
    std::fprintf(stderr, "experiment %s is not part of this driver\n", o.experiment.c_str());
    return 0;
  }

  const std::string log_path = o.results_dir + "/" + o.experiment + "_" +
                               std::to_string(o.cols) + "_" + o.hardware + ".csv";
  const bool fresh = file_is_empty(log_path);
  std::ofstream log(log_path, std::ofstream::app);
  if (!log) {
    std::fprintf(stderr, "cannot open %s\n", log_path.c_str());
    return 1;
  }
  std::cout << log_path + " has been opened" << std::endl;
  if (fresh) log << "kernel,platform,format,library,tensor,time\n";
  std::vector<float> best(kSweepRows, 0.0f);
  std::vector<double> mean(kSweepRows, 0.0);
  for (int r = 0; r < kSweepRows; ++r) {
    const SweepRow &row = kSweep[r];
    float mn = 0.0f;
    double sum = 0.0;
    for (int t = 0; t < o.trials; ++t) {
      row.fn(&op.tC, &op.tA, &op.tB);
      const float ms = get_gpu_timer_result();
      log << "spmm," << row.platform << "," << row.format << "," << row.library << ","
          << o.matrix_name << "," << ms << "\n";
      mn = t == 0 ? ms : std::min(mn, ms);
      sum += ms;
    }
    best[r] = mn;
    mean[r] = o.trials > 0 ? sum / o.trials : 0.0;
  }
  log.close();
  if (o.summary) {
    // resana.py's two tables: each row's minimum over the trials
    // (resana.py:150-180) and its mean rounded to 4 decimals
    // (resana.py:28-33), one line per matrix in each file
    auto table = [&](const std::string &suffix, auto value) {
      const std::string path = o.results_dir + "/" + o.experiment + "_" +
                               std::to_string(o.cols) + "_" + o.hardware + suffix;
      const bool sfresh = file_is_empty(path);
      std::ofstream s(path, std::ofstream::app);
      if (sfresh) {
        s << "dataset";
        for (int r = 0; r < kSweepRows; ++r) {
          if (!std::strcmp(kSweep[r].library, "alg2"))
            s << ",alg2";
          else
            s << "," << kSweep[r].platform << kSweep[r].format << kSweep[r].library;
        }
        s << "\n";
      }
      s << o.matrix_name;
      for (int r = 0; r < kSweepRows; ++r) s << "," << value(r);
      s << "\n";
    };
    table("_summary.csv", [&](int r) { return best[r]; });
    table("_summary_mean.csv",
          [&](int r) { return std::round(mean[r] * 1e4) / 1e4; });
  }
  return 0;
}

// runals.cpp -- standalone ALS driver (the reference's runALS.sh, without Spark).
//
// Reference: als/runALS.sh:18-37 takes <cpu|gpu> <master> <rank> <iterations>
// <lambda> <file> and launches Spark's MovieLensALS, which prints the
// held-out RMSE.  This driver takes the same knobs minus the Spark master:
//
//   runals <cpu|gpu> <rank> <iterations> <lambda> <dataset> [--seed S]
//          [--test <count>:<prefix> | --holdout <fraction>] [--threads N]
//          gpu only: [--split N] [--precision fp32|bf16] [--solver cholesky|cg]
//                    [--graph 0|1] [--tiles <file>] [--autotune <file>]
//
// <dataset> is a generator preset name (see cumf_gen_preset) or
//   csr:<m>:<n>:<nnz>:<prefix>   raw CSR files <prefix>.row.bin, .col.bin,
//                                 .data.bin (the reference's binary format)
// with an optional held-out COO set  --test <count>:<prefix>  (<prefix>.u.bin,
// .v.bin int32 and .r.bin fp32), or a seeded split of the file's ratings
// --holdout <fraction> (cumf_split_holdout: every row and column keeps a
// training rating).  The "cpu" mode runs the host half-sweep of
// libcumf_cpu.so (FP64 normal equations + Cholesky, OpenMP over rows)
// where runALS.sh runs Spark's JVM ALS; both modes start from the same
// seeded factors, so their RMSEs are comparable.
//
// Output: one line per iteration and a final JSON summary on stdout.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/als.h"
#include "../../include/cumf_cpu.h"
#include "../../include/cumf_data.h"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: runals <cpu|gpu> <rank> <iterations> <lambda> <dataset> [--seed S]\n"
               "             [--test <count>:<prefix> | --holdout <fraction>] [--threads N]\n"
               "  gpu only:  [--split N] [--precision fp32|bf16] [--solver cholesky|cg]\n"
               "             [--graph 0|1] [--tiles <tile table>] [--autotune <tile table to write>]\n"
               "  dataset: preset name (zipf_10k_2k_500k, netflix_480k_17770_99m, zipf_4m_500k_1b)\n"
               "           or csr:<m>:<n>:<nnz>:<prefix> (raw <prefix>.row.bin/.col.bin/.data.bin)\n");
  return 2;
}

bool read_file(const std::string& path, void* dst, size_t bytes) {
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) return false;
  const bool ok = std::fread(dst, 1, bytes, fp) == bytes && std::fgetc(fp) == EOF;
  std::fclose(fp);
  return ok;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// The cpu mode: X half-sweep over the CSR, Theta half-sweep over the CSC
// (made here when the dataset did not bring one), held-out RMSE after each
// iteration; the same output as the gpu mode.
int run_cpu(int f, int iters, float lambda, uint64_t seed, int threads, int64_t m, int64_t n, int64_t nnz,
            const std::vector<int32_t>& rp, const std::vector<int32_t>& ci, const std::vector<float>& rv,
            std::vector<int32_t>& cp, std::vector<int32_t>& ri, std::vector<float>& cv,
            const std::vector<int32_t>& tu, const std::vector<int32_t>& tv, const std::vector<float>& tr) {
  if (cp.empty()) {
    cp.resize(n + 1); ri.resize(nnz); cv.resize(nnz);
    if (cumf_cpu_transpose(m, n, nnz, rp.data(), ci.data(), rv.data(), cp.data(), ri.data(), cv.data()) != 0) {
      std::fprintf(stderr, "runals: bad CSR (column index out of range)\n");
      return 1;
    }
  }
  std::vector<float> X((size_t)m * f), theta((size_t)n * f);
  cumf_gen_init_factors(n, f, seed, theta.data());
  const int64_t ntest = (int64_t)tu.size();
  double total = 0.0, rmse = 0.0;
  int failed = 0;
  for (int it = 0; it < iters; ++it) {
    const double t0 = now_s();
    const int a = cumf_cpu_half_sweep(m, n, f, rp.data(), ci.data(), rv.data(), theta.data(), lambda, X.data(),
                                      threads);
    const int b = cumf_cpu_half_sweep(n, m, f, cp.data(), ri.data(), cv.data(), X.data(), lambda, theta.data(),
                                      threads);
    const double dt = now_s() - t0;
    if (a == -1 || b == -1) {
      std::fprintf(stderr, "runals: bad arguments to the cpu half-sweep\n");
      return 1;
    }
    failed += (a == -2) + (b == -2);
    total += dt;
    if (ntest > 0) rmse = cumf_cpu_rmse(ntest, tu.data(), tv.data(), tr.data(), f, X.data(), theta.data(), threads);
    std::printf("iteration %d: %.4f s, held-out RMSE %.6f\n", it + 1, dt, rmse);
  }
  std::printf("{\"mode\": \"cpu\", \"rank\": %d, \"iterations\": %d, \"lambda\": %g, \"m\": %lld, \"n\": %lld, "
              "\"nnz\": %lld, \"s_per_iteration\": %.6f, \"heldout_rmse\": %.6f, \"sweeps_with_failed_rows\": %d}\n",
              f, iters, (double)lambda, (long long)m, (long long)n, (long long)nnz, iters ? total / iters : 0.0, rmse,
              failed);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) return usage();
  const std::string mode = argv[1];
  if (mode != "gpu" && mode != "cpu") return usage();
  const bool cpu = mode == "cpu";
  const int f = std::atoi(argv[2]);
  const int iters = std::atoi(argv[3]);
  const float lambda = (float)std::atof(argv[4]);
  const std::string dataset = argv[5];
  int split = 0, precision = CUMF_PREC_FP32, solver = CUMF_SOLVER_CHOLESKY, graph = 0, threads = 0;
  bool gpu_option = false;
  double holdout = -1.0;
  uint64_t seed = 0;
  std::string test_spec, tiles, autotune;
  if ((argc - 6) % 2) return usage();
  for (int i = 6; i + 1 < argc; i += 2) {
    const std::string v = argv[i + 1];
    const bool g = std::strcmp(argv[i], "--seed") && std::strcmp(argv[i], "--test") &&
                   std::strcmp(argv[i], "--holdout") && std::strcmp(argv[i], "--threads");
    gpu_option = gpu_option || g;
    if (!std::strcmp(argv[i], "--split")) split = std::atoi(argv[i + 1]);
    else if (!std::strcmp(argv[i], "--seed")) seed = std::strtoull(argv[i + 1], nullptr, 10);
    else if (!std::strcmp(argv[i], "--test")) test_spec = v;
    else if (!std::strcmp(argv[i], "--holdout")) holdout = std::atof(argv[i + 1]);
    else if (!std::strcmp(argv[i], "--threads")) threads = std::atoi(argv[i + 1]);
    else if (!std::strcmp(argv[i], "--precision") && (v == "fp32" || v == "bf16"))
      precision = v == "bf16" ? CUMF_PREC_BF16_GRAM : CUMF_PREC_FP32;
    else if (!std::strcmp(argv[i], "--solver") && (v == "cholesky" || v == "cg"))
      solver = v == "cg" ? CUMF_SOLVER_CG : CUMF_SOLVER_CHOLESKY;
    else if (!std::strcmp(argv[i], "--graph")) graph = std::atoi(argv[i + 1]) != 0;
    else if (!std::strcmp(argv[i], "--tiles")) tiles = v;
    else if (!std::strcmp(argv[i], "--autotune")) autotune = v;
    else return usage();
  }
  if (f <= 0 || iters < 0) return usage();
  if (cpu && gpu_option) {
    std::fprintf(stderr, "runals: --split, --precision, --solver, --graph, --tiles and --autotune are gpu options\n");
    return 2;
  }
  if (holdout >= 0.0 && (!test_spec.empty() || holdout > 1.0 || dataset.rfind("csr:", 0) != 0)) {
    std::fprintf(stderr, "runals: --holdout <fraction in [0,1]> splits a csr: dataset; it excludes --test\n");
    return 2;
  }

  int64_t m = 0, n = 0, nnz = 0, ntest = 0;
  std::vector<int32_t> rp, ci, tu, tv, cp, ri;
  std::vector<float> rv, tr, cv;
  const double t_load = now_s();
  if (dataset.rfind("csr:", 0) == 0) {
    long long mm, nn, zz;
    char prefix[4096];
    if (std::sscanf(dataset.c_str(), "csr:%lld:%lld:%lld:%4095s", &mm, &nn, &zz, prefix) != 4) return usage();
    m = mm; n = nn; nnz = zz;
    rp.resize(m + 1); ci.resize(nnz); rv.resize(nnz);
    const std::string p = prefix;
    if (cumf_load_csr_bin((p + ".data.bin").c_str(), (p + ".row.bin").c_str(), (p + ".col.bin").c_str(),
                          rv.data(), rp.data(), ci.data(), (int)m, (long)nnz) != CUMF_OK) {
      std::fprintf(stderr, "runals: %s\n", cumf_last_error());
      return 1;
    }
    if (holdout >= 0.0) {
      std::vector<int32_t> trp(m + 1), tci(nnz);
      std::vector<float> trv(nnz);
      tu.resize(nnz); tv.resize(nnz); tr.resize(nnz);
      int64_t ntrain = 0;
      if (cumf_split_holdout(m, n, nnz, rp.data(), ci.data(), rv.data(), holdout, seed, trp.data(), tci.data(),
                             trv.data(), &ntrain, tu.data(), tv.data(), tr.data(), &ntest) != CUMF_OK) {
        std::fprintf(stderr, "runals: %s\n", cumf_last_error());
        return 1;
      }
      tci.resize(ntrain); trv.resize(ntrain);
      tu.resize(ntest); tv.resize(ntest); tr.resize(ntest);
      rp.swap(trp); ci.swap(tci); rv.swap(trv);
      nnz = ntrain;
    }
    if (!test_spec.empty()) {
      long long cnt;
      char tp[4096];
      if (std::sscanf(test_spec.c_str(), "%lld:%4095s", &cnt, tp) != 2) return usage();
      ntest = cnt;
      tu.resize(ntest); tv.resize(ntest); tr.resize(ntest);
      const std::string q = tp;
      if (!read_file(q + ".u.bin", tu.data(), 4 * ntest) || !read_file(q + ".v.bin", tv.data(), 4 * ntest) ||
          !read_file(q + ".r.bin", tr.data(), 4 * ntest)) {
        std::fprintf(stderr, "runals: cannot read held-out files %s.{u,v,r}.bin\n", tp);
        return 1;
      }
    }
  } else {
    cumf_gen_params gp;
    if (cumf_gen_preset(dataset.c_str(), &gp) != 0) return usage();
    gp.seed = seed;
    cumf_dataset* ds = nullptr;
    if (cumf_gen_create(&gp, &ds) != 0) {
      std::fprintf(stderr, "runals: generator rejected %s\n", dataset.c_str());
      return 1;
    }
    cumf_gen_info_t info;
    cumf_gen_info(ds, &info);
    m = info.m; n = info.n; nnz = info.nnz; ntest = info.n_test;
    rp.resize(m + 1); ci.resize(nnz); rv.resize(nnz);
    cp.resize(n + 1); ri.resize(nnz); cv.resize(nnz);
    tu.resize(ntest); tv.resize(ntest); tr.resize(ntest);
    cumf_gen_copy(ds, rp.data(), ci.data(), rv.data(), cp.data(), ri.data(), cv.data(), tu.data(), tv.data(),
                  tr.data());
    cumf_gen_free(ds);
  }
  std::fprintf(stderr, "runals: %s m=%lld n=%lld nnz=%lld test=%lld loaded in %.1fs\n", dataset.c_str(),
               (long long)m, (long long)n, (long long)nnz, (long long)ntest, now_s() - t_load);
  if (cpu) return run_cpu(f, iters, lambda, seed, threads, m, n, nnz, rp, ci, rv, cp, ri, cv, tu, tv, tr);

  cumf_options opt;
  cumf_default_options(&opt);
  opt.split = split;
  opt.timing = 1;
  opt.precision = precision;
  opt.solver = solver;
  opt.use_graph = graph;
  cumf_handle* h = nullptr;
  auto check = [&](int rc, const char* what) {
    if (rc != CUMF_OK) {
      std::fprintf(stderr, "runals: %s failed: %s\n", what, cumf_last_error());
      if (h) cumf_destroy(h);
      std::exit(1);
    }
  };
  check(cumf_create(-1, f, lambda, &opt, &h), "cumf_create");
  check(cumf_set_ratings(h, m, n, nnz, rp.data(), ci.data(), rv.data(), cp.empty() ? nullptr : cp.data(),
                         ri.empty() ? nullptr : ri.data(), cv.empty() ? nullptr : cv.data()),
        "cumf_set_ratings");
  if (ntest > 0) check(cumf_set_heldout(h, ntest, tu.data(), tv.data(), tr.data()), "cumf_set_heldout");
  std::vector<float> theta((size_t)n * f);
  cumf_gen_init_factors(n, f, seed, theta.data());
  check(cumf_set_factors(h, nullptr, theta.data()), "cumf_set_factors");
  if (!tiles.empty()) check(cumf_load_tile_table(h, tiles.c_str()), "cumf_load_tile_table");
  if (!autotune.empty()) {
    const double t0 = now_s();
    check(cumf_autotune(h, 2, autotune.c_str()), "cumf_autotune");
    std::fprintf(stderr, "runals: autotuned in %.1fs, table written to %s\n", now_s() - t0, autotune.c_str());
  }

  double total = 0.0, rmse = 0.0;
  for (int it = 0; it < iters; ++it) {
    const double t0 = now_s();
    check(cumf_iterate(h, 1), "cumf_iterate");
    check(cumf_synchronize(h), "cumf_synchronize");
    const double dt = now_s() - t0;
    total += dt;
    if (ntest > 0) check(cumf_heldout_rmse(h, &rmse), "cumf_heldout_rmse");
    std::printf("iteration %d: %.4f s, held-out RMSE %.6f\n", it + 1, dt, rmse);
  }
  cumf_stats st;
  check(cumf_get_stats(h, &st), "cumf_get_stats");
  std::printf("{\"mode\": \"gpu\", \"rank\": %d, \"iterations\": %d, \"lambda\": %g, \"m\": %lld, \"n\": %lld, \"nnz\": %lld, "
              "\"precision\": \"%s\", \"solver\": \"%s\", "
              "\"s_per_iteration\": %.6f, \"heldout_rmse\": %.6f, \"failed_rows\": %lld, "
              "\"cg_fallback_rows\": %lld}\n",
              f, iters, (double)lambda, (long long)m, (long long)n, (long long)nnz,
              precision == CUMF_PREC_BF16_GRAM ? "bf16" : "fp32", solver == CUMF_SOLVER_CG ? "cg" : "cholesky",
              iters ? total / iters : 0.0, rmse, (long long)st.failed_rows, (long long)st.cg_fallback_rows);
  cumf_destroy(h);
  return 0;
}

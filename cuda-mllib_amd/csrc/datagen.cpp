// datagen.cpp -- seeded synthetic rating matrices with power-law degrees.
//
// The reference ships no data and no generator (its users feed Netflix /
// MovieLens through Spark, als/runALS.sh:33-37).  This generator produces
// rating matrices of a stated shape for benchmarks and tests:
//
//   * Row degrees follow a Zipf-Mandelbrot law d_i = C / (i + q)^s with
//     d_0 = cap and sum d_i = nnz_total (q solved by bisection), rounded to
//     integers by largest remainder, each >= 1, then assigned to rows in a
//     seeded random order.
//   * Column weights follow the same law with their own (s, cap); each row
//     draws its d_i distinct columns with probability proportional to the
//     weights (alias-table rejection sampling for light rows, Efraimidis-
//     Spirakis keys for rows that cover more than 1/8 of the columns).
//   * Empty columns are repaired by moving one rating from a row that
//     keeps >= 1 and a column that keeps >= 1.
//   * Ratings = clip(round(mu + u_i . v_j + N(0, noise^2)), 1, 5) with latent
//     u_i, v_j ~ N(0, latent_var) per entry (or the raw value when
//     round_clip == 0).
//   * n_holdout pairs are moved to a held-out COO set, uniformly at random
//     among those whose row and column keep >= 1 training rating.
//
// Every random draw comes from a counter-based stream keyed by (seed, tag,
// index), so the output does not depend on the thread count.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string_view>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../include/cumf_data.h"

namespace {

inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

inline uint64_t key3(uint64_t seed, uint64_t tag, uint64_t idx) {
  return mix64(mix64(seed ^ (tag * 0xD1B54A32D192ED03ull)) ^ idx);
}

struct Stream {
  uint64_t s;
  explicit Stream(uint64_t k) : s(k) {}
  uint64_t next() { s += 0x9E3779B97F4A7C15ull; return mix64(s); }
  // (0, 1]
  double uniform() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

enum : uint64_t {
  TAG_ROWPERM = 1, TAG_COLPERM, TAG_SAMPLE, TAG_LAT_U, TAG_LAT_V, TAG_NOISE,
  TAG_HOLDOUT, TAG_REPAIR, TAG_INIT
};

// sum_{i<count} cap * (q / (i + q))^s, exact for the first terms and the
// midpoint-rule integral for the smooth tail.
double law_sum(double q, double s, double cap, int64_t count) {
  const int64_t head = std::min<int64_t>(count, 20000);
  double acc = 0.0;
  for (int64_t i = 0; i < head; ++i) acc += std::pow(q / ((double)i + q), s);
  if (count > head) {
    double a = (double)head - 0.5 + q, b = (double)count - 0.5 + q;
    double integral = (std::fabs(s - 1.0) < 1e-12)
                          ? std::log(b / a)
                          : (std::pow(b, 1.0 - s) - std::pow(a, 1.0 - s)) / (1.0 - s);
    acc += std::pow(q, s) * integral;
  }
  return cap * acc;
}

// q such that law_sum(q) == total (law_sum increases with q).
double solve_q(double s, double cap, int64_t count, double total) {
  double lo = 1e-9, hi = 1.0;
  while (law_sum(hi, s, cap, count) < total && hi < 1e15) hi *= 2.0;
  for (int it = 0; it < 200; ++it) {
    double mid = std::sqrt(lo * hi);
    if (law_sum(mid, s, cap, count) < total) lo = mid; else hi = mid;
  }
  return std::sqrt(lo * hi);
}

// Integer degrees summing exactly to total, each in [1, maxdeg].
std::vector<int64_t> integer_degrees(int64_t count, double s, int64_t cap, int64_t total,
                                     int64_t maxdeg, double* q_out) {
  double q = solve_q(s, (double)cap, count, (double)total);
  if (q_out) *q_out = q;
  std::vector<int64_t> d(count);
  std::vector<double> frac(count);
  int64_t sum = 0;
  for (int64_t i = 0; i < count; ++i) {
    double x = (double)cap * std::pow(q / ((double)i + q), s);
    int64_t k = (int64_t)std::floor(x);
    k = std::max<int64_t>(1, std::min<int64_t>(k, maxdeg));
    d[i] = k;
    frac[i] = x - std::floor(x);
    sum += k;
  }
  // Distribute the remainder to the largest fractional parts (then round
  // robin), or take the excess from the largest degrees.
  if (sum < total) {
    std::vector<int64_t> idx(count);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return frac[a] > frac[b]; });
    int64_t need = total - sum;
    for (int pass = 0; need > 0 && pass < 1 << 20; ++pass)
      for (int64_t t = 0; t < count && need > 0; ++t)
        if (d[idx[t]] < maxdeg) { d[idx[t]]++; need--; }
  } else if (sum > total) {
    int64_t excess = sum - total;
    for (int64_t i = 0; i < count && excess > 0; ++i) {
      int64_t take = std::min<int64_t>(excess, d[i] - 1);
      d[i] -= take;
      excess -= take;
    }
  }
  return d;
}

std::vector<int64_t> seeded_perm(int64_t count, uint64_t seed, uint64_t tag) {
  std::vector<int64_t> p(count);
  std::iota(p.begin(), p.end(), 0);
  Stream st(key3(seed, tag, 0));
  for (int64_t i = count - 1; i > 0; --i) {
    int64_t j = (int64_t)(st.next() % (uint64_t)(i + 1));
    std::swap(p[i], p[j]);
  }
  return p;
}

struct Alias {
  std::vector<double> prob;
  std::vector<int64_t> alias;
  void build(const std::vector<double>& w) {
    int64_t n = (int64_t)w.size();
    double tot = 0.0;
    for (double x : w) tot += x;
    prob.assign(n, 0.0);
    alias.assign(n, 0);
    std::vector<int64_t> small, large;
    std::vector<double> p(n);
    for (int64_t i = 0; i < n; ++i) {
      p[i] = w[i] * (double)n / tot;
      (p[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
      int64_t a = small.back(); small.pop_back();
      int64_t b = large.back();
      prob[a] = p[a];
      alias[a] = b;
      p[b] -= 1.0 - p[a];
      if (p[b] < 1.0) { large.pop_back(); small.push_back(b); }
    }
    for (int64_t i : large) prob[i] = 1.0;
    for (int64_t i : small) prob[i] = 1.0;
  }
  int64_t draw(Stream& st) const {
    uint64_t r = st.next();
    int64_t i = (int64_t)((r >> 20) % (uint64_t)prob.size());
    double u = (double)(r & 0xFFFFF) * (1.0 / 1048576.0);
    return u < prob[i] ? i : alias[i];
  }
};

}  // namespace

struct cumf_dataset {
  cumf_gen_params p{};
  double q_row = 0, q_col = 0;
  std::vector<int32_t> rp, ci;    // CSR of the training ratings
  std::vector<float> rv;
  std::vector<int32_t> cp, ri;    // CSC (= CSR of R^T)
  std::vector<float> cv;
  std::vector<int32_t> tu, tv;    // held-out COO
  std::vector<float> tr;
  int64_t hist[5] = {0, 0, 0, 0, 0};
};

namespace {

float rating_value(const cumf_gen_params& p, const std::vector<float>& U, const std::vector<float>& V,
                   int64_t i, int64_t j) {
  double dot = 0.0;
  for (int k = 0; k < p.rank; ++k) dot += (double)U[i * p.rank + k] * (double)V[j * p.rank + k];
  Stream st(key3(p.seed, TAG_NOISE, (uint64_t)i * 0x100000001ull + (uint64_t)j));
  double x = p.mu + dot + p.noise_std * st.normal();
  if (p.round_clip) x = std::min(5.0, std::max(1.0, std::nearbyint(x)));
  return (float)x;
}

void transpose_csr(int64_t m, int64_t n, const std::vector<int32_t>& rp, const std::vector<int32_t>& ci,
                   const std::vector<float>& rv, std::vector<int32_t>& cp, std::vector<int32_t>& ri,
                   std::vector<float>& cv) {
  int64_t nnz = rp[m];
  cp.assign(n + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) cp[ci[k] + 1]++;
  for (int64_t j = 0; j < n; ++j) cp[j + 1] += cp[j];
  std::vector<int64_t> pos(cp.begin(), cp.end() - 1);
  ri.resize(nnz);
  cv.resize(nnz);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      int64_t d = pos[ci[k]]++;
      ri[d] = (int32_t)i;
      cv[d] = rv[k];
    }
}

}  // namespace

extern "C" {

int cumf_gen_create(const cumf_gen_params* params, cumf_dataset** out) {
  if (!params || !out) return -1;
  const cumf_gen_params p = *params;
  if (p.m <= 0 || p.n <= 0 || p.nnz_total < std::max(p.m, p.n) || p.n_holdout < 0 ||
      p.n_holdout >= p.nnz_total || p.nnz_total > (int64_t)p.m * p.n || p.rank <= 0 ||
      p.nnz_total - p.n_holdout > 2147483647ll || p.m > 2147483647ll || p.n > 2147483647ll)
    return -1;
  cumf_dataset* ds = new cumf_dataset();
  ds->p = p;
  const int64_t m = p.m, n = p.n;

  // Degrees and column weights.
  std::vector<int64_t> deg_sorted =
      integer_degrees(m, p.row_s, std::min<int64_t>(p.row_cap, n), p.nnz_total, n, &ds->q_row);
  std::vector<int64_t> rperm = seeded_perm(m, p.seed, TAG_ROWPERM);
  std::vector<int64_t> deg(m);
  for (int64_t i = 0; i < m; ++i) deg[rperm[i]] = deg_sorted[i];
  ds->q_col = solve_q(p.col_s, (double)p.col_cap, n, (double)p.nnz_total);
  std::vector<int64_t> cperm = seeded_perm(n, p.seed, TAG_COLPERM);
  std::vector<double> w(n);
  for (int64_t j = 0; j < n; ++j) w[cperm[j]] = std::pow(ds->q_col / ((double)j + ds->q_col), p.col_s);
  Alias alias;
  alias.build(w);

  // Sample distinct columns per row.
  std::vector<int64_t> rp(m + 1, 0);
  for (int64_t i = 0; i < m; ++i) rp[i + 1] = rp[i] + deg[i];
  std::vector<int32_t> cols(p.nnz_total);
#pragma omp parallel
  {
    std::vector<int64_t> stamp(n, -1);
    std::vector<std::pair<double, int32_t>> keys;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < m; ++i) {
      Stream st(key3(p.seed, TAG_SAMPLE, (uint64_t)i));
      int64_t d = deg[i];
      int32_t* out_c = cols.data() + rp[i];
      if (d * 8 <= n) {
        int64_t got = 0;
        while (got < d) {
          int64_t j = alias.draw(st);
          if (stamp[j] != i) { stamp[j] = i; out_c[got++] = (int32_t)j; }
        }
      } else {
        keys.resize(n);
        for (int64_t j = 0; j < n; ++j) keys[j] = {std::log(st.uniform()) / w[j], (int32_t)j};
        std::nth_element(keys.begin(), keys.begin() + (d - 1), keys.end(),
                         [](const std::pair<double, int32_t>& a, const std::pair<double, int32_t>& b) {
                           return a.first > b.first;
                         });
        for (int64_t t = 0; t < d; ++t) out_c[t] = keys[t].second;
      }
      std::sort(out_c, out_c + d);
    }
  }

  // Repair empty columns.
  std::vector<int64_t> ccount(n, 0);
  for (int64_t k = 0; k < p.nnz_total; ++k) ccount[cols[k]]++;
  {
    Stream st(key3(p.seed, TAG_REPAIR, 0));
    for (int64_t j = 0; j < n; ++j) {
      if (ccount[j] > 0) continue;
      for (int tries = 0; tries < 1 << 24; ++tries) {
        int64_t i = (int64_t)(st.next() % (uint64_t)m);
        if (deg[i] < 2) continue;
        int32_t* c = cols.data() + rp[i];
        if (std::binary_search(c, c + deg[i], (int32_t)j)) continue;
        int64_t t = (int64_t)(st.next() % (uint64_t)deg[i]);
        if (ccount[c[t]] < 2) continue;
        ccount[c[t]]--;
        c[t] = (int32_t)j;
        ccount[j]++;
        std::sort(c, c + deg[i]);
        break;
      }
    }
  }

  // Latent factors and rating values.
  std::vector<float> U((size_t)m * p.rank), V((size_t)n * p.rank);
  const double lstd = std::sqrt(p.latent_var);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < m; ++i) {
    Stream st(key3(p.seed, TAG_LAT_U, (uint64_t)i));
    for (int k = 0; k < p.rank; ++k) U[i * p.rank + k] = (float)(lstd * st.normal());
  }
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) {
    Stream st(key3(p.seed, TAG_LAT_V, (uint64_t)j));
    for (int k = 0; k < p.rank; ++k) V[j * p.rank + k] = (float)(lstd * st.normal());
  }

  // Held-out selection: candidates with the smallest seeded keys, accepted
  // greedily while their row and column keep >= 1 training rating.
  std::vector<uint8_t> held(p.nnz_total, 0);
  int64_t n_held = 0;
  if (p.n_holdout > 0) {
    std::vector<int64_t> rcount(deg.begin(), deg.end());
    double frac = std::min(1.0, 1.5 * (double)p.n_holdout / (double)p.nnz_total + 1e-6);
    while (n_held < p.n_holdout) {
      uint64_t thr = frac >= 1.0 ? ~0ull : (uint64_t)(frac * 18446744073709551616.0);
      std::vector<std::pair<uint64_t, int64_t>> cand;
#pragma omp parallel
      {
        std::vector<std::pair<uint64_t, int64_t>> loc;
#pragma omp for schedule(static) nowait
        for (int64_t k = 0; k < p.nnz_total; ++k) {
          uint64_t key = key3(p.seed, TAG_HOLDOUT, (uint64_t)k);
          if (key <= thr && !held[k]) loc.push_back({key, k});
        }
#pragma omp critical
        cand.insert(cand.end(), loc.begin(), loc.end());
      }
      std::sort(cand.begin(), cand.end());
      std::vector<int64_t> row_of;  // lazily computed row of a rating index
      for (auto& c : cand) {
        if (n_held >= p.n_holdout) break;
        int64_t k = c.second;
        int64_t i = (int64_t)(std::upper_bound(rp.begin(), rp.end(), k) - rp.begin()) - 1;
        int64_t j = cols[k];
        if (rcount[i] > 1 && ccount[j] > 1) {
          held[k] = 1;
          rcount[i]--;
          ccount[j]--;
          n_held++;
        }
      }
      if (frac >= 1.0) break;
      frac = std::min(1.0, frac * 2.0);
    }
  }

  // Emit CSR (training), COO (held out), values.
  const int64_t nnz_train = p.nnz_total - n_held;
  ds->rp.assign(m + 1, 0);
  for (int64_t i = 0; i < m; ++i) {
    int64_t c = 0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) c += !held[k];
    ds->rp[i + 1] = (int32_t)c;
  }
  for (int64_t i = 0; i < m; ++i) ds->rp[i + 1] += ds->rp[i];
  ds->ci.resize(nnz_train);
  ds->rv.resize(nnz_train);
  std::vector<int64_t> hoff(m + 1, 0);
  for (int64_t i = 0; i < m; ++i) hoff[i + 1] = hoff[i] + (rp[i + 1] - rp[i]) - (ds->rp[i + 1] - ds->rp[i]);
  ds->tu.resize(n_held);
  ds->tv.resize(n_held);
  ds->tr.resize(n_held);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < m; ++i) {
    int64_t t = ds->rp[i], h = hoff[i];
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      float r = rating_value(p, U, V, i, cols[k]);
      if (held[k]) {
        ds->tu[h] = (int32_t)i; ds->tv[h] = cols[k]; ds->tr[h] = r; ++h;
      } else {
        ds->ci[t] = cols[k]; ds->rv[t] = r; ++t;
      }
    }
  }
  for (int64_t k = 0; k < nnz_train; ++k) {
    int b = (int)ds->rv[k];
    if (p.round_clip && b >= 1 && b <= 5) ds->hist[b - 1]++;
  }
  transpose_csr(m, n, ds->rp, ds->ci, ds->rv, ds->cp, ds->ri, ds->cv);
  *out = ds;
  return 0;
}

int cumf_gen_info(const cumf_dataset* ds, cumf_gen_info_t* info) {
  if (!ds || !info) return -1;
  info->m = ds->p.m;
  info->n = ds->p.n;
  info->nnz = (int64_t)ds->ci.size();
  info->n_test = (int64_t)ds->tu.size();
  info->q_row = ds->q_row;
  info->q_col = ds->q_col;
  int64_t maxr = 0, minr = INT64_MAX, maxc = 0, minc = INT64_MAX;
  for (int64_t i = 0; i < ds->p.m; ++i) {
    int64_t d = ds->rp[i + 1] - ds->rp[i];
    maxr = std::max(maxr, d);
    minr = std::min(minr, d);
  }
  for (int64_t j = 0; j < ds->p.n; ++j) {
    int64_t d = ds->cp[j + 1] - ds->cp[j];
    maxc = std::max(maxc, d);
    minc = std::min(minc, d);
  }
  info->max_row_deg = maxr;
  info->min_row_deg = minr;
  info->max_col_deg = maxc;
  info->min_col_deg = minc;
  for (int b = 0; b < 5; ++b) info->hist[b] = ds->hist[b];
  return 0;
}

int cumf_gen_copy(const cumf_dataset* ds, int32_t* csr_row_ptr, int32_t* csr_col, float* csr_val,
                  int32_t* csc_col_ptr, int32_t* csc_row, float* csc_val, int32_t* test_u,
                  int32_t* test_v, float* test_r) {
  if (!ds) return -1;
  auto cp = [](auto* dst, const auto& src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
  };
  cp(csr_row_ptr, ds->rp);
  cp(csr_col, ds->ci);
  cp(csr_val, ds->rv);
  cp(csc_col_ptr, ds->cp);
  cp(csc_row, ds->ri);
  cp(csc_val, ds->cv);
  cp(test_u, ds->tu);
  cp(test_v, ds->tv);
  cp(test_r, ds->tr);
  return 0;
}

void cumf_gen_free(cumf_dataset* ds) { delete ds; }

// Seeded initial factors: each row is N(0,1)^f scaled to unit L2 norm.
int cumf_gen_init_factors(int64_t rows, int f, uint64_t seed, float* out) {
  if (rows < 0 || f <= 0 || !out) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < rows; ++i) {
    Stream st(key3(seed, TAG_INIT, (uint64_t)i));
    double nrm = 0.0;
    for (int k = 0; k < f; ++k) {
      double g = st.normal();
      out[i * f + k] = (float)g;
      nrm += g * g;
    }
    float s = (float)(1.0 / std::sqrt(nrm));
    for (int k = 0; k < f; ++k) out[i * f + k] *= s;
  }
  return 0;
}

int cumf_gen_preset(const char* name, cumf_gen_params* p) {
  if (!name || !p) return -1;
  std::memset(p, 0, sizeof(*p));
  std::string_view nm(name);
  if (nm == "zipf_10k_2k_500k") {
    // m=10,000, n=2,000, 500,000 ratings of which 10% held out; Zipf(1.1)
    // rows and columns; rank-10 latent with per-entry variance 1/sqrt(10);
    // noise sd 0.5; no offset.
    *p = cumf_gen_params{10000, 2000, 500000, 50000, 1.1, 1.1, 2000, 10000, 10,
                         1.0 / std::sqrt(10.0), 0.5, 0.0, 1, 0};
    return 0;
  }
  if (nm == "netflix_480k_17770_99m") {
    // m=480,189, n=17,770, 99,072,112 training ratings plus a 1% held-out
    // set (990,721) drawn from the same law; rows Zipf(1.0) capped at
    // 17,000, columns Zipf(0.8) capped at 232,944; rank-20 latent with
    // u.v of unit variance, offset 3.6, noise sd 0.7.
    *p = cumf_gen_params{480189, 17770, 99072112 + 990721, 990721, 1.0, 0.8, 17000, 232944, 20,
                         1.0 / std::sqrt(20.0), 0.7, 3.6, 1, 0};
    return 0;
  }
  if (nm == "zipf_4m_500k_1b") {
    // m=4,000,000, n=500,000, 1.0e9 training ratings plus 0.1% held out;
    // rows Zipf(1.2) capped at 200,000, columns Zipf(0.8) capped at
    // 1,000,000; latent model as in the Netflix-shaped workload.
    *p = cumf_gen_params{4000000, 500000, 1000000000 + 1000000, 1000000, 1.2, 0.8, 200000, 1000000,
                         20, 1.0 / std::sqrt(20.0), 0.7, 3.6, 1, 0};
    return 0;
  }
  return -1;
}

}  // extern "C"

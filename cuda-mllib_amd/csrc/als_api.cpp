// als_api.cpp -- host driver of the MI355X ALS update.
//
// Reference: the half-sweep driver doALSWithCSR (als/src/cuda/als.cu:1000-
// 1115) and its batching loop (:1061-1100).  The reference allocates,
// creates BLAS handles and synchronises around every step; here a half-
// sweep is two kernel launches on one stream (fused Gram+solve over
// wavefront tasks, then the reduction of split rows), with the ratings and
// both factor matrices resident on the device between sweeps.
//
// Task planning (host, once per rating matrix): every row with at most
// `split` ratings becomes one task; longer rows are cut into `split`-sized
// chunks whose partial Grams land in scratch slots and are summed in chunk
// order by the finish kernel.  Tasks are ordered by decreasing length so
// the longest start first (longest-processing-time-first scheduling).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/als.h"
#include "../../include/cumf_mlp.h"
#include "../../include/utilities.h"
#include "als_kernels.h"
#include "cumf_internal.h"
#include "mlp_internal.h"
#include "mlp_kernels.h"

using cumf::DevBuf;

namespace {

constexpr int kMaxF = 256;
// Largest f of the fused register-tile kernels (9 column blocks); above it
// the half-sweep runs the work-group-per-row FP32 Cholesky kernel.
constexpr int kMaxFusedF = 140;
constexpr int kDefaultSplit = 16384;
constexpr int kDefaultCgIters = 6;
constexpr int kDefaultCgRounds = 6;
constexpr float kDefaultCgTol = 1e-5f;  // the per-row relative residual bar

int64_t round_up4(int64_t f) { return (f + 3) & ~int64_t(3); }

// Degree buckets of the batching form: tasks whose rating count falls in
// [kBucketEdge[b], kBucketEdge[b + 1]) form bucket b.
constexpr int kBuckets = 8;
constexpr int64_t kBucketEdge[kBuckets + 1] = {0, 32, 64, 96, 256, 1024, 4096, 16384, INT64_MAX};
constexpr int kAuxStreams = 3;

// Above kMaxFusedF, lighter rows are solved in the dual forms (see
// large_launches) unless a test switched them off (cumf_debug_large_dual).
std::atomic<int> g_large_dual{1};

// Launch shape of one bucket: register budget (wavefronts per SIMD the
// kernel was compiled for, 1 or 2), persistent grid (0: one work-group per
// task; > 0: that many wavefronts taking tasks from a counter) and solver
// (-1: the handle's; CUMF_SOLVER_CHOLESKY or CUMF_SOLVER_CG, which the
// autotuner may pick per bucket when the handle's solver is CG).
struct BucketCfg {
  int variant = 2;
  int persistent = 0;
  int solver = -1;
  int occupancy = 0;  // at most this many wavefronts per CU (0: as registers and LDS allow)
};

// Host half of the planner: every row with at most `split` ratings is one
// task; a longer row is cut into ceil(n / split) chunks whose sizes differ
// by at most one rating (so none is short while split >= 2 x 96: no chunk
// lands in the dual form's degree buckets), each a task writing its
// partial Gram to a scratch slot (consecutive slots per row), and the row
// is listed as heavy.  Above kMaxFusedF the work-group-per-row kernel
// solves whole rows: nothing is split.  Tasks are ordered by decreasing
// rating count (stable), so every degree bucket is a contiguous range.
int plan_tasks(const int32_t* row_ptr, int64_t m, int split_, int f, std::vector<cumf::Task>& t,
               std::vector<cumf::HeavyRow>& hv, int64_t& nslots) {
  const int64_t split = f <= kMaxFusedF ? std::max(split_, 1) : INT64_MAX;
  t.clear();
  hv.clear();
  t.reserve((size_t)m);
  int64_t slot = 0;
  for (int64_t u = 0; u < m; ++u) {
    const int64_t n = (int64_t)row_ptr[u + 1] - row_ptr[u];
    if (n < 0) {
      cumf::set_last_error("row pointers decrease at row " + std::to_string(u));
      return CUMF_ERR_INVALID;
    }
    if (n <= split) {
      t.push_back({(int32_t)u, row_ptr[u], (int32_t)n, -1});
    } else {
      const int64_t nc = (n + split - 1) / split;
      const int64_t base = n / nc, rem = n % nc;  // the first rem chunks take base + 1 ratings
      hv.push_back({(int32_t)u, (int32_t)slot, (int32_t)nc, 0});
      int64_t off = row_ptr[u];
      for (int64_t c = 0; c < nc; ++c) {
        const int64_t len = base + (c < rem ? 1 : 0);
        t.push_back({(int32_t)u, (int32_t)off, (int32_t)len, (int32_t)(slot + c)});
        off += len;
      }
      slot += nc;
    }
  }
  std::stable_sort(t.begin(), t.end(), [](const cumf::Task& a, const cumf::Task& b) { return a.count > b.count; });
  nslots = slot;
  return CUMF_OK;
}

// Wavefront task list of one rating matrix side on the device.
struct Plan {
  int64_t rows = 0, nnz = 0, ntasks = 0, nheavy = 0, nslots = 0;
  int64_t light0 = 0;  // first task of at most kDualMaxRatingsHost ratings (tasks light0 .. ntasks-1)
  // Tasks [wg_t0[d], wg_t1[d]) have more than kDualMaxRatingsHost ratings
  // and d = (count + 16) / 16 <= kLargeDualWgMaxBlocks row blocks.
  int64_t wg_t0[cumf::kLargeDualWgMaxBlocks + 1] = {}, wg_t1[cumf::kLargeDualWgMaxBlocks + 1] = {};
  int split = 0;
  int64_t bucket_t0[kBuckets] = {}, bucket_t1[kBuckets] = {};
  bool bucket_partial[kBuckets] = {};  // bucket b holds a split row's chunk
  DevBuf tasks, heavy, partial;

  int build(const int32_t* row_ptr, int64_t m, int split_, int f) {
    std::vector<cumf::Task> t;
    std::vector<cumf::HeavyRow> hv;
    int64_t slots = 0;
    CUMF_TRY(plan_tasks(row_ptr, m, split_, f, t, hv, slots));
    rows = m;
    nnz = row_ptr[m];
    split = f <= kMaxFusedF ? split_ : 0;
    ntasks = (int64_t)t.size();
    for (int b = 0; b < kBuckets; ++b) {
      bucket_t0[b] = bucket_t1[b] = 0;
      bucket_partial[b] = false;
    }
    for (int64_t i = 0; i < ntasks; ++i) {
      int b = 0;
      while (b + 1 < kBuckets && t[(size_t)i].count >= kBucketEdge[b + 1]) ++b;
      if (bucket_t1[b] == 0) bucket_t0[b] = i;
      bucket_t1[b] = i + 1;
      bucket_partial[b] = bucket_partial[b] || t[(size_t)i].slot >= 0;
    }
    light0 = ntasks;
    for (int b = 0; b < kBuckets && kBucketEdge[b + 1] <= cumf::kDualMaxRatingsHost + 1; ++b)
      if (bucket_t1[b] > bucket_t0[b]) light0 = std::min(light0, bucket_t0[b]);
    for (int d = 0; d <= cumf::kLargeDualWgMaxBlocks; ++d) wg_t0[d] = wg_t1[d] = 0;
    for (int64_t i = 0; i < ntasks; ++i) {
      const int cnt = t[(size_t)i].count;
      const int d = (cnt + 16) / 16;
      if (cnt <= cumf::kDualMaxRatingsHost || d > cumf::kLargeDualWgMaxBlocks) continue;
      if (wg_t1[d] == 0) wg_t0[d] = i;
      wg_t1[d] = i + 1;
    }
    nheavy = (int64_t)hv.size();
    nslots = slots;
    CUMF_TRY(tasks.alloc(sizeof(cumf::Task) * t.size()));
    CUMF_TRY(heavy.alloc(sizeof(cumf::HeavyRow) * hv.size()));
    CUMF_TRY(partial.alloc(sizeof(float) * cumf::partial_floats_per_slot(f) * (size_t)nslots));
    if (!t.empty()) CUMF_HIP_TRY(hipMemcpy(tasks.p, t.data(), tasks.bytes, hipMemcpyHostToDevice));
    if (!hv.empty()) CUMF_HIP_TRY(hipMemcpy(heavy.p, hv.data(), heavy.bytes, hipMemcpyHostToDevice));
    return CUMF_OK;
  }
};

int check_f(int f) {
  if (f < 1 || f > kMaxF) {
    cumf::set_last_error("f must be in [1, " + std::to_string(kMaxF) + "], got " + std::to_string(f));
    return CUMF_ERR_INVALID;
  }
  return CUMF_OK;
}

// Kernel-time bookkeeping with HIP events.
struct Timer {
  enum Kind { TASK = 0, FINISH, RMSE, SWEEP, ITER, WARM, FALLBACK, NKINDS };
  struct Pending { hipEvent_t a, b; int kind; int launches; };
  std::vector<hipEvent_t> pool;
  std::vector<Pending> pending;
  double ms[NKINDS] = {};
  int64_t count[NKINDS] = {};

  ~Timer() {
    for (auto& p : pending) { pool.push_back(p.a); pool.push_back(p.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    return e;
  }
  // Must be called after the work the events bracket has completed.
  void resolve() {
    for (auto& p : pending) {
      float t = 0.f;
      if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) { ms[p.kind] += t; count[p.kind] += p.launches; }
      else (void)hipGetLastError();
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
};

}  // namespace

struct cumf_handle {
  int device = 0;
  int f = 0;
  int64_t ldf = 0;
  float lambda = 0.f;
  cumf_options opt{};
  hipStream_t stream = nullptr;
  // Streams (default priority) on which the launches of a half-sweep's
  // degree buckets run beside the first one, forked from and joined back
  // into `stream` with the events below (also under graph capture).
  hipStream_t aux[kAuxStreams] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kAuxStreams] = {};
  int64_t m = 0, n = 0, nnz = 0, ntest = 0;
  DevBuf rp, ci, rv, cp, ri, cv;  // CSR and CSC
  DevBuf X, T;                    // factors, row stride ldf
  DevBuf Xb, Tb;                  // BF16 shadow factors (precision mode), row stride ldb
  int64_t ldb = 0;
  DevBuf tu, tv, tr;              // held-out COO
  DevBuf rmse_partial, rmse_out, fail;
  Plan plan_x, plan_t;
  Timer timer;
  int64_t gram_flops = 0;
  bool have_ratings = false;
  // Learned warm start: attached network, features and predictions of the
  // X side's rows.
  cumf_mlp* mlp = nullptr;
  DevBuf feat, x0;
  int64_t x_sweeps = 0;           // X half-sweeps since the factors were last set
  bool graph_net[2] = {false, false};  // the captured half-sweep runs the network
  // Rows the CG pass hands to the Cholesky fallback, and their count.
  DevBuf fb_list, fb_count;
  // Host copies of the row / column pointers (plans are rebuilt when the
  // tile table changes the split size).
  std::vector<int32_t> rp_host, cp_host;
  // Tile table (batching form): per side, the split size, one launch shape
  // per degree bucket and the register budget of the finish and fallback
  // kernels.  Without a table: one launch of all tasks, two-wave budget.
  bool tiled = false;
  int tile_split[2] = {0, 0};
  BucketCfg tiles[2][kBuckets];
  int finish_variant[2] = {2, 2}, fallback_variant[2] = {2, 2};
  DevBuf counters;  // persistent task counters, kBuckets per side
  int num_cus = 0;
  // One captured half-sweep per side (use_graph), rebuilt after anything
  // that changes its launches.
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  bool capturing = false;
  void drop_graph() {
    for (auto& g : graph) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }
  void destroy_streams() {
    for (int i = 0; i < kAuxStreams; ++i) {
      if (aux[i]) (void)hipStreamDestroy(aux[i]);
      if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
      aux[i] = nullptr;
      join_ev[i] = nullptr;
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    fork_ev = nullptr;
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

namespace {

constexpr int kRmseBlocks = 1024;

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) { (void)hipGetLastError(); prev = -1; }
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

template <class F>
int timed(cumf_handle* h, int kind, F&& launch, int launches = 1) {
  hipEvent_t a = nullptr, b = nullptr;
  if (h->opt.timing && !h->capturing) {
    a = h->timer.get();
    b = h->timer.get();
    if (a) (void)hipEventRecord(a, h->stream);
  }
  hipError_t e = launch();
  if (e != hipSuccess) {
    cumf::set_last_error(std::string("kernel launch failed: ") + hipGetErrorString(e));
    if (a) h->timer.pool.push_back(a);
    if (b) h->timer.pool.push_back(b);
    return CUMF_ERR_HIP;
  }
  if (a && b) {
    (void)hipEventRecord(b, h->stream);
    h->timer.pending.push_back({a, b, kind, launches});
  }
  return CUMF_OK;
}

// Row stride (in bf16) of the BF16 factor copies: f rounded up to 8, then
// the row to a power of two up to 128 bytes or a multiple of 128 bytes, so
// no row straddles more 128-byte lines than it fills (at f = 100: 256-byte
// rows instead of 208; the Theta half-sweep's gathers of X rows fetched
// 13% less time with them).
int64_t bf16_stride(int f) {
  const int64_t b = 2 * (int64_t)((f + 7) & ~7);
  int64_t r = 16;
  while (r < b && r < 128) r *= 2;
  if (b > 128) r = (b + 127) & ~int64_t(127);
  return r / 2;
}

// Register budget of the launches without a tile table: two wavefronts
// per SIMD while the accumulator tiles leave room for the solver's
// registers (up to 8 column blocks, f <= 124), one wavefront per SIMD
// beyond (the two-wavefront build spills several hundred bytes per lane).
int default_variant(int f) { return cumf::sweep_nb(f) >= 9 ? 1 : 2; }

hipError_t task_launch(const cumf::SweepArgs& a, int variant, hipStream_t s) {
  if (a.solver == CUMF_SOLVER_CG)
    return variant == 1 ? cumf::w1::cg::launch_task_kernel(a, s) : cumf::w2::cg::launch_task_kernel(a, s);
  return variant == 1 ? cumf::w1::chol::launch_task_kernel(a, s) : cumf::w2::chol::launch_task_kernel(a, s);
}
hipError_t finish_launch(const cumf::SweepArgs& a, int variant, hipStream_t s) {
  if (a.solver == CUMF_SOLVER_CG)
    return variant == 1 ? cumf::w1::cg::launch_finish_kernel(a, s) : cumf::w2::cg::launch_finish_kernel(a, s);
  return variant == 1 ? cumf::w1::chol::launch_finish_kernel(a, s) : cumf::w2::chol::launch_finish_kernel(a, s);
}
hipError_t fallback_launch(const cumf::SweepArgs& a, int variant, hipStream_t s) {
  return variant == 1 ? cumf::w1::chol::launch_fallback_kernel(a, s) : cumf::w2::chol::launch_fallback_kernel(a, s);
}
hipError_t dual_launch(const cumf::SweepArgs& a, int variant, hipStream_t s) {
  return variant == 1 ? cumf::w1::chol::launch_dual_kernel(a, s) : cumf::w2::chol::launch_dual_kernel(a, s);
}

// Degree buckets whose rows may take the dual form (every task at most 95
// ratings): 0, 1 and 2.
constexpr int kDualBuckets = 3;
static_assert(kBucketEdge[kDualBuckets] == cumf::kDualMaxRatingsHost + 1, "dual-form buckets hold at most 95 ratings");
// 32-rating chunks the dual form stages for bucket b: one up to 31 ratings,
// two up to 63, three up to 95.
int dual_chunks_of(int b) { return kBucketEdge[b + 1] <= 32 ? 1 : kBucketEdge[b + 1] <= 64 ? 2 : 3; }

// Launch of the tasks of degree buckets b_lo .. b (a contiguous range of the
// task list: tasks are ordered by decreasing rating count) with one shape.
hipError_t bucket_launch(cumf_handle* h, int side, const Plan& p, const cumf::SweepArgs& a, int b,
                         const BucketCfg& cfg, int b_lo = -1, hipStream_t stream = nullptr) {
  if (b_lo < 0) b_lo = b;
  if (!stream) stream = h->stream;
  cumf::SweepArgs x = a;
  x.tasks = a.tasks + p.bucket_t0[b];
  x.ntasks = p.bucket_t1[b_lo] - p.bucket_t0[b];
  x.persistent = cfg.persistent;
  x.occupancy = cfg.persistent ? 0 : cfg.occupancy;
  x.task_counter = h->counters.as<int>() + side * kBuckets + b;
  if (cfg.solver == CUMF_SOLVER_DUAL) {
    x.dual_partial = 0;
    for (int bb = b_lo; bb <= b; ++bb) x.dual_partial |= p.bucket_partial[bb] ? 1 : 0;
    if (x.precision == CUMF_PREC_FP32) {
      // FP32 factors: the wavefront dual form reads the rows in place (any
      // ldf <= 256) and solves whole rows only; a launch holding a split
      // row's chunk (a split size under 2 x 96) takes the primal kernel.
      if (x.dual_partial) {
        x.solver = CUMF_SOLVER_CHOLESKY;
        return task_launch(x, cfg.variant, stream);
      }
      return cumf::launch_large_dual_kernel(x, x.ntasks, stream);
    }
    // (the one-chunk kernel has no partial-Gram branch: a bucket 0 with a
    // split row's chunk takes the two-chunk one)
    x.dual_chunks = std::max(dual_chunks_of(b), x.dual_partial ? 2 : 1);
    return dual_launch(x, cfg.variant, stream);
  }
  if (cfg.solver >= 0) x.solver = cfg.solver;
  return task_launch(x, cfg.variant, stream);
}

// The half-sweep above kMaxFusedF over a plan's whole-row tasks (heaviest
// first), with the dual forms on (g_large_dual):
//   * rows of at most kDualMaxRatingsHost ratings: the wavefront dual form;
//   * rows whose dual system has at most NB - 2 row blocks (NB those of the
//     f x f system, and at most kLargeDualWgMaxBlocks): the work-group dual
//     form, one launch per block count;
//   * the rest: the work-group primal kernel.
// Side by side when auxiliary streams are given (forked from s, joined
// back); launches counts the kernel launches.
hipError_t large_launches(const cumf::SweepArgs& a, const Plan& p, int* launches, hipStream_t s,
                          hipStream_t aux0 = nullptr, hipStream_t aux1 = nullptr, hipEvent_t fork = nullptr,
                          hipEvent_t join0 = nullptr, hipEvent_t join1 = nullptr) {
  const bool dual = g_large_dual.load() != 0;
  const int nb = cumf::sweep_nb(a.f);
  int64_t primal_end = dual ? p.light0 : p.ntasks;  // tasks [0, primal_end) take the primal kernel
  int dmax = 0;                                     // the work-group dual form's largest block count
  if (dual)
    for (int d = std::min(nb - 2, cumf::kLargeDualWgMaxBlocks); d >= 1; --d)
      if (p.wg_t1[d] > p.wg_t0[d]) {
        dmax = std::max(dmax, d);
        primal_end = std::min(primal_end, p.wg_t0[d]);
      }
  const int64_t nl = dual ? p.ntasks - p.light0 : 0;
  const bool side = aux0 && aux1;
  hipError_t e = hipSuccess;
  if (side) {
    e = hipEventRecord(fork, s);
    if (e == hipSuccess) e = hipStreamWaitEvent(aux0, fork, 0);
    if (e == hipSuccess) e = hipStreamWaitEvent(aux1, fork, 0);
  }
  *launches = 0;
  if (e == hipSuccess && primal_end > 0) {
    e = cumf::launch_large_kernel(a, primal_end, s);
    ++*launches;
  }
  for (int d = dmax; e == hipSuccess && d >= 1; --d) {
    if (p.wg_t1[d] <= p.wg_t0[d]) continue;
    cumf::SweepArgs w = a;
    w.tasks = a.tasks + p.wg_t0[d];
    e = cumf::launch_large_dual_wg_kernel(w, p.wg_t1[d] - p.wg_t0[d], d, side ? aux1 : s);
    ++*launches;
  }
  if (e == hipSuccess && nl > 0) {
    cumf::SweepArgs l = a;
    l.tasks = a.tasks + p.light0;
    e = cumf::launch_large_dual_kernel(l, nl, side ? aux0 : s);
    ++*launches;
  }
  if (e == hipSuccess && side) {
    e = hipEventRecord(join0, aux0);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, join0, 0);
    if (e == hipSuccess) e = hipEventRecord(join1, aux1);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, join1, 0);
  }
  return e;
}

// All tasks of a half-sweep: one launch, or with a tile table one launch
// per non-empty degree bucket (heaviest first), persistent ones after their
// counters are cleared.
int launch_tasks(cumf_handle* h, int side, const Plan& p, const cumf::SweepArgs& a) {
  if (!h->tiled) return timed(h, Timer::TASK, [&] { return task_launch(a, default_variant(h->f), h->stream); });
  // A side whose non-empty buckets share one shape and no persistent grid is
  // one launch (launches in sequence drain the GPU between buckets).
  int v0 = 0, s0 = -2, o0 = 0;
  bool uniform = true;
  for (int b = 0; b < kBuckets; ++b) {
    if (p.bucket_t1[b] <= p.bucket_t0[b]) continue;
    const BucketCfg& c = h->tiles[side][b];
    const int sv = c.solver >= 0 ? c.solver : a.solver;
    if (v0 == 0) {
      v0 = c.variant;
      s0 = sv;
      o0 = c.occupancy;
    }
    uniform = uniform && c.persistent == 0 && c.variant == v0 && sv == s0 && c.occupancy == o0;
  }
  if (uniform && v0 != 0 && s0 != CUMF_SOLVER_DUAL) {
    cumf::SweepArgs x = a;
    x.solver = s0;
    x.occupancy = o0;
    return timed(h, Timer::TASK, [&] { return task_launch(x, v0, h->stream); });
  }
  bool persistent = false;
  for (int b = 0; b < kBuckets; ++b) persistent = persistent || h->tiles[side][b].persistent > 0;
  if (persistent)
    CUMF_HIP_TRY(cumf::launch_fallback_reset(h->counters.as<int>() + side * kBuckets, kBuckets, h->stream));
  // Launch groups, heaviest first: neighbouring non-empty buckets with one
  // shape (and no persistent grid) share a launch.
  auto same = [](const BucketCfg& x, const BucketCfg& y) {
    return x.persistent == 0 && y.persistent == 0 && x.variant == y.variant && x.solver == y.solver &&
           x.occupancy == y.occupancy;
  };
  // (the dual form stages one, two or three chunks by bucket: buckets with
  // different counts do not share a launch)
  auto dual_ch = [&](int b) { return dual_chunks_of(b); };
  int grp_hi[kBuckets], grp_lo[kBuckets], ng = 0;
  for (int b = kBuckets - 1; b >= 0; --b) {
    if (p.bucket_t1[b] <= p.bucket_t0[b]) continue;
    int lo = b;
    while (lo > 0 && (p.bucket_t1[lo - 1] <= p.bucket_t0[lo - 1] ||
                      (same(h->tiles[side][lo - 1], h->tiles[side][b]) &&
                       (h->tiles[side][b].solver != CUMF_SOLVER_DUAL || dual_ch(lo - 1) == dual_ch(b)))))
      --lo;
    while (p.bucket_t1[lo] <= p.bucket_t0[lo]) ++lo;  // the range ends at a non-empty bucket
    grp_hi[ng] = b;
    grp_lo[ng++] = lo;
    b = lo;
  }
  // The groups run side by side: the first on the handle's stream, the
  // others on the auxiliary streams forked from it and joined back (the
  // heaviest rows no longer run alone at the end of a launch of their own).
  return timed(h, Timer::TASK, [&]() -> hipError_t {
    const int nstreams = std::min(ng, kAuxStreams + 1);
    if (nstreams > 1) {
      hipError_t e = hipEventRecord(h->fork_ev, h->stream);
      for (int i = 0; e == hipSuccess && i + 1 < nstreams; ++i) e = hipStreamWaitEvent(h->aux[i], h->fork_ev, 0);
      if (e != hipSuccess) return e;
    }
    for (int g = 0; g < ng; ++g) {
      const int si = g % nstreams;
      const hipStream_t st = si == 0 ? h->stream : h->aux[si - 1];
      const hipError_t e = bucket_launch(h, side, p, a, grp_hi[g], h->tiles[side][grp_hi[g]], grp_lo[g], st);
      if (e != hipSuccess) return e;
    }
    for (int i = 0; i + 1 < nstreams; ++i) {
      hipError_t e = hipEventRecord(h->join_ev[i], h->aux[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(h->stream, h->join_ev[i], 0);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }, ng);
}

bool uses_network(const cumf_handle* h) {
  return h->opt.solver == CUMF_SOLVER_CG &&
         (h->opt.warm_start == CUMF_WARM_MLP || h->opt.warm_start == CUMF_WARM_BEST);
}

// The network starts this X half-sweep (warm_sweeps limits it to the first
// ones after the factors were set).
bool network_now(const cumf_handle* h) {
  return uses_network(h) && (h->opt.warm_sweeps <= 0 || h->x_sweeps < h->opt.warm_sweeps);
}

// Buffers of the learned warm start (allocated outside any capture).
int prepare_network(cumf_handle* h) {
  if (!uses_network(h)) return CUMF_OK;
  if (!h->mlp) {
    cumf::set_last_error("warm_start CUMF_WARM_MLP / CUMF_WARM_BEST needs a network (cumf_set_mlp)");
    return CUMF_ERR_STATE;
  }
  const size_t fb = 2 * (size_t)h->m * cumf::mlp_k0(h->mlp);
  if (h->feat.bytes < fb) CUMF_TRY(h->feat.alloc(fb));
  const size_t xb = sizeof(float) * (size_t)h->m * h->ldf;
  if (h->x0.bytes < xb) CUMF_TRY(h->x0.alloc(xb));
  return cumf::mlp_prepare_predict(h->mlp);
}

// Kernel arguments of a half-sweep over plan p (the CG fallback list is
// allocated here when missing, which must not happen during a capture).
int fill_args(cumf_handle* h, const Plan& p, const DevBuf& rowp, const DevBuf& col, const DevBuf& val,
              const float* theta, float* x, const DevBuf& theta_b, DevBuf& x_b, float* trace,
              cumf::SweepArgs& a) {
  a.row_ptr = rowp.as<int32_t>();
  a.col = col.as<int32_t>();
  a.val = val.as<float>();
  a.theta = theta;
  a.x = x;
  a.ldf = h->ldf;
  a.tasks = p.tasks.as<cumf::Task>();
  a.ntasks = p.ntasks;
  a.heavy = p.heavy.as<cumf::HeavyRow>();
  a.nheavy = p.nheavy;
  a.partial = p.partial.as<float>();
  a.fail_counter = h->fail.as<int>();
  a.f = h->f;
  a.lambda = h->lambda;
  a.precision = h->opt.precision;
  a.ldb = h->ldb;
  a.theta_b = theta_b.as<uint16_t>();
  a.x_b = x_b.as<uint16_t>();
  a.solver = h->opt.solver;
  a.cg_iters = h->opt.cg_iters;
  a.cg_rounds = h->opt.cg_rounds;
  a.cg_tol = h->opt.cg_tol;
  a.cg_fallback = h->fail.as<int>() + 1;
  a.cg_trace = trace;
  if (h->opt.solver == CUMF_SOLVER_CG) {
    const size_t need = sizeof(int) * (size_t)std::max<int64_t>(p.rows, 1);
    if (h->fb_list.bytes < need) {
      if (h->capturing) {
        cumf::set_last_error("fallback list not allocated before capture");
        return CUMF_ERR_STATE;
      }
      CUMF_TRY(h->fb_list.alloc(need));
    }
    if (!h->fb_count.p) CUMF_TRY(h->fb_count.alloc(sizeof(int)));
    a.fb_list = h->fb_list.as<int>();
    a.fb_count = h->fb_count.as<int>();
  }
  return CUMF_OK;
}

int run_half_sweep(cumf_handle* h, int side, const Plan& p, const DevBuf& rowp, const DevBuf& col,
                   const DevBuf& val, const float* theta, float* x, const DevBuf& theta_b, DevBuf& x_b,
                   float* trace = nullptr) {
  if (h->opt.compat_hermitian_bug) {
    cumf::CompatArgs c{};
    c.row_ptr = rowp.as<int32_t>();
    c.col = col.as<int32_t>();
    c.val = val.as<float>();
    c.theta = theta;
    c.x = x;
    c.ldf = h->ldf;
    c.fail_counter = h->fail.as<int>();
    c.f = h->f;
    c.lambda = h->lambda;
    CUMF_TRY(timed(h, Timer::TASK, [&] { return cumf::launch_compat(c, p.rows, h->stream); }));
    h->gram_flops += p.nnz * (int64_t)h->f * (h->f + 1);
    return CUMF_OK;
  }
  cumf::SweepArgs a{};
  CUMF_TRY(fill_args(h, p, rowp, col, val, theta, x, theta_b, x_b, trace, a));
  if (h->f > kMaxFusedF) {
    hipEvent_t s0 = nullptr, s1 = nullptr;
    if (h->opt.timing && !h->capturing) {
      s0 = h->timer.get();
      s1 = h->timer.get();
      if (s0) (void)hipEventRecord(s0, h->stream);
    }
    int nlaunch = 0;
    const size_t npend = h->timer.pending.size();
    CUMF_TRY(timed(h, Timer::TASK, [&] {
      return large_launches(a, p, &nlaunch, h->stream, h->aux[0], h->aux[1], h->fork_ev, h->join_ev[0],
                            h->join_ev[1]);
    }));
    if (h->timer.pending.size() > npend) h->timer.pending.back().launches = std::max(1, nlaunch);
    if (s0 && s1) {
      (void)hipEventRecord(s1, h->stream);
      h->timer.pending.push_back({s0, s1, Timer::SWEEP, 1});
    }
    h->gram_flops += p.nnz * (int64_t)h->f * (h->f + 1);
    return CUMF_OK;
  }
  const bool cg = h->opt.solver == CUMF_SOLVER_CG;
  a.warm = h->opt.warm_start;
  // the network predicts X rows, and only while network_now() holds
  if (a.warm >= CUMF_WARM_MLP && (side == 1 || !network_now(h))) a.warm = CUMF_WARM_PREV;
  if (side == 0 && network_now(h)) {
    CUMF_TRY(prepare_network(h));
    const int K0 = cumf::mlp_k0(h->mlp);
    CUMF_TRY(timed(h, Timer::WARM, [&] {
      hipError_t e = cumf::launch_mlp_features(rowp.as<int32_t>(), col.as<int32_t>(), val.as<float>(), theta,
                                               h->ldf, h->f, h->lambda, 0, p.rows, K0, h->feat.as<__bf16>(),
                                               h->stream);
      if (e != hipSuccess) return e;
      return cumf::mlp_predict_device(h->mlp, h->feat.as<__bf16>(), p.rows, h->x0.as<float>(), h->ldf,
                                      h->stream) == CUMF_OK
                 ? hipSuccess
                 : hipErrorLaunchFailure;
    }));
    a.x0 = h->x0.as<float>();
  }
  hipEvent_t s0 = nullptr, s1 = nullptr;
  if (h->opt.timing && !h->capturing) {
    s0 = h->timer.get();
    s1 = h->timer.get();
    if (s0) (void)hipEventRecord(s0, h->stream);
  }
  if (cg) CUMF_HIP_TRY(cumf::launch_fallback_reset(a.fb_count, 1, h->stream));
  CUMF_TRY(launch_tasks(h, side, p, a));
  const int dv = default_variant(h->f);
  const int fv = h->tiled ? h->finish_variant[side] : dv, bv = h->tiled ? h->fallback_variant[side] : dv;
  if (p.nheavy > 0)
    CUMF_TRY(timed(h, Timer::FINISH, [&] { return finish_launch(a, fv, h->stream); }));
  if (cg) CUMF_TRY(timed(h, Timer::FALLBACK, [&] { return fallback_launch(a, bv, h->stream); }));
  if (s0 && s1) {
    (void)hipEventRecord(s1, h->stream);
    h->timer.pending.push_back({s0, s1, Timer::SWEEP, 1});
  }
  h->gram_flops += p.nnz * (int64_t)h->f * (h->f + 1);
  return CUMF_OK;
}

int upload_factors(cumf_handle* h, DevBuf& dst, DevBuf& dst_b, const float* src, int64_t rows) {
  CUMF_HIP_TRY(hipMemsetAsync(dst.p, 0, dst.bytes, h->stream));
  CUMF_HIP_TRY(hipMemcpy2DAsync(dst.p, sizeof(float) * h->ldf, src, sizeof(float) * h->f,
                                sizeof(float) * h->f, rows, hipMemcpyHostToDevice, h->stream));
  if (h->opt.precision != CUMF_PREC_FP32)
    CUMF_HIP_TRY(cumf::launch_to_bf16(dst.as<float>(), h->ldf, dst_b.p, h->ldb, rows, h->f, h->stream));
  return CUMF_OK;
}

}  // namespace

extern "C" {

int cumf_default_options(cumf_options* opt) {
  if (!opt) return CUMF_ERR_INVALID;
  std::memset(opt, 0, sizeof(*opt));
  opt->version = CUMF_OPTIONS_VERSION;
  opt->precision = CUMF_PREC_FP32;
  opt->split = kDefaultSplit;
  opt->timing = 0;
  opt->solver = CUMF_SOLVER_CHOLESKY;
  opt->cg_iters = kDefaultCgIters;
  opt->cg_rounds = kDefaultCgRounds;
  opt->cg_tol = kDefaultCgTol;
  return CUMF_OK;
}

int cumf_create(int device, int f, float lambda, const cumf_options* opt, cumf_handle** out) {
  cumf::clear_last_error();
  if (!out) return CUMF_ERR_INVALID;
  *out = nullptr;
  CUMF_TRY(check_f(f));
  if (!(lambda >= 0.f)) {
    cumf::set_last_error("lambda must be >= 0");
    return CUMF_ERR_INVALID;
  }
  cumf_options o;
  cumf_default_options(&o);
  if (opt) {
    if (opt->version != CUMF_OPTIONS_VERSION) {
      cumf::set_last_error("cumf_options.version mismatch");
      return CUMF_ERR_INVALID;
    }
    o = *opt;
    if (o.split <= 0) o.split = kDefaultSplit;
  }
  if (o.cg_iters <= 0) o.cg_iters = kDefaultCgIters;
  if (o.cg_rounds <= 0) o.cg_rounds = kDefaultCgRounds;
  if (!(o.cg_tol > 0.f)) o.cg_tol = kDefaultCgTol;
  if (o.solver != CUMF_SOLVER_CHOLESKY && o.solver != CUMF_SOLVER_CG) {
    cumf::set_last_error("solver not built");
    return CUMF_ERR_UNSUPPORTED;
  }
  if (o.solver != CUMF_SOLVER_CHOLESKY && o.compat_hermitian_bug) {
    cumf::set_last_error("compat_hermitian_bug reproduces the reference's LU solve; use CUMF_SOLVER_CHOLESKY");
    return CUMF_ERR_INVALID;
  }
  if (o.warm_start < CUMF_WARM_PREV || o.warm_start > CUMF_WARM_BEST ||
      (o.warm_start != CUMF_WARM_PREV && o.solver != CUMF_SOLVER_CG)) {
    cumf::set_last_error("warm_start must be a CUMF_WARM_* value; only CUMF_SOLVER_CG has a start");
    return CUMF_ERR_INVALID;
  }
  if (f > kMaxFusedF && (o.precision != CUMF_PREC_FP32 || o.solver != CUMF_SOLVER_CHOLESKY ||
                         o.compat_hermitian_bug)) {
    cumf::set_last_error("f > " + std::to_string(kMaxFusedF) +
                         " runs the FP32 Gram + FP32 Cholesky path only");
    return CUMF_ERR_UNSUPPORTED;
  }
  if (o.warm_sweeps < 0) {
    cumf::set_last_error("warm_sweeps must be >= 0");
    return CUMF_ERR_INVALID;
  }
  if (o.cg_iters > 64 || o.cg_rounds > 16) {
    cumf::set_last_error("cg_iters must be <= 64 and cg_rounds <= 16");
    return CUMF_ERR_INVALID;
  }
  if (o.precision != CUMF_PREC_FP32 && o.precision != CUMF_PREC_BF16_GRAM) {
    cumf::set_last_error("precision mode not built");
    return CUMF_ERR_UNSUPPORTED;
  }
  if (o.precision != CUMF_PREC_FP32 && o.compat_hermitian_bug) {
    cumf::set_last_error("compat_hermitian_bug reproduces the FP32 reference; use CUMF_PREC_FP32");
    return CUMF_ERR_INVALID;
  }
  if (o.compat_hermitian_bug && (f % 10 != 0 || f < 10)) {
    cumf::set_last_error("compat_hermitian_bug reproduces the reference, which needs f % 10 == 0");
    return CUMF_ERR_INVALID;
  }
  if (device < 0) device = get_gpu();
  int count = 0;
  if (device < 0 || hipGetDeviceCount(&count) != hipSuccess || device >= count) {
    (void)hipGetLastError();
    cumf::set_last_error("no usable GPU device");
    return CUMF_ERR_NODEVICE;
  }
  DeviceGuard dg(device);
  if (o.compat_hermitian_bug && cumf::prepare_compat(f) != hipSuccess) {
    (void)hipGetLastError();
    cumf::set_last_error("compat kernel: cannot raise its LDS limit");
    return CUMF_ERR_HIP;
  }
  auto* h = new cumf_handle();
  h->device = device;
  h->f = f;
  h->ldf = round_up4(f);
  h->ldb = bf16_stride(f);
  h->lambda = lambda;
  h->opt = o;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete h;
    cumf::set_last_error("hipStreamCreate failed");
    return CUMF_ERR_HIP;
  }
  int rc = CUMF_OK;
  for (int i = 0; i < kAuxStreams && rc == CUMF_OK; ++i)
    if (hipStreamCreateWithFlags(&h->aux[i], hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->join_ev[i], hipEventDisableTiming) != hipSuccess)
      rc = CUMF_ERR_HIP;
  if (rc == CUMF_OK && hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming) != hipSuccess) rc = CUMF_ERR_HIP;
  if (rc != CUMF_OK) {
    (void)hipGetLastError();
    cumf::set_last_error("hipStreamCreate / hipEventCreate failed");
  }
  if (rc == CUMF_OK) rc = h->fail.alloc(2 * sizeof(int));  // failed rows, CG fallbacks
  if (rc == CUMF_OK) rc = h->counters.alloc(sizeof(int) * 2 * kBuckets);
  if (rc == CUMF_OK && hipDeviceGetAttribute(&h->num_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    (void)hipGetLastError();
    h->num_cus = 256;
  }
  if (rc == CUMF_OK) rc = h->rmse_partial.alloc(sizeof(double) * kRmseBlocks);
  if (rc == CUMF_OK) rc = h->rmse_out.alloc(sizeof(double));
  if (rc == CUMF_OK && hipMemset(h->fail.p, 0, 2 * sizeof(int)) != hipSuccess) rc = CUMF_ERR_HIP;
  if (rc == CUMF_OK && f > kMaxFusedF && cumf::prepare_large_kernel(f) != hipSuccess) {
    (void)hipGetLastError();
    cumf::set_last_error("large-f kernel: cannot raise its LDS limit");
    rc = CUMF_ERR_HIP;
  }
  if (rc != CUMF_OK) {
    h->destroy_streams();
    delete h;
    return rc;
  }
  *out = h;
  return CUMF_OK;
}

int cumf_destroy(cumf_handle* h) {
  if (!h) return CUMF_OK;
  {
    DeviceGuard dg(h->device);
    (void)hipStreamSynchronize(h->stream);
    h->timer.resolve();
    h->drop_graph();
    h->destroy_streams();
    delete h;  // DevBuf destructors free on the handle's device
  }
  return CUMF_OK;
}

int cumf_set_ratings(cumf_handle* h, int64_t m, int64_t n, int64_t nnz, const int32_t* csr_row_ptr,
                     const int32_t* csr_col, const float* csr_val, const int32_t* csc_col_ptr,
                     const int32_t* csc_row, const float* csc_val) {
  cumf::clear_last_error();
  if (!h || m <= 0 || n <= 0 || nnz < 0 || nnz > INT32_MAX || m > INT32_MAX || n > INT32_MAX ||
      !csr_row_ptr || (nnz > 0 && (!csr_col || !csr_val))) {
    cumf::set_last_error("cumf_set_ratings: bad arguments");
    return CUMF_ERR_INVALID;
  }
  if (csr_row_ptr[0] != 0 || csr_row_ptr[m] != nnz) {
    cumf::set_last_error("cumf_set_ratings: row pointers must start at 0 and end at nnz");
    return CUMF_ERR_INVALID;
  }
  for (int64_t k = 0; k < nnz; ++k)
    if (csr_col[k] < 0 || csr_col[k] >= n) {
      cumf::set_last_error("cumf_set_ratings: column index out of range at " + std::to_string(k));
      return CUMF_ERR_INVALID;
    }
  if (csc_col_ptr) {
    if (nnz > 0 && (!csc_row || !csc_val)) {
      cumf::set_last_error("cumf_set_ratings: CSC arrays missing");
      return CUMF_ERR_INVALID;
    }
    if (csc_col_ptr[0] != 0 || csc_col_ptr[n] != nnz) {
      cumf::set_last_error("cumf_set_ratings: CSC column pointers must start at 0 and end at nnz");
      return CUMF_ERR_INVALID;
    }
    for (int64_t k = 0; k < nnz; ++k)
      if (csc_row[k] < 0 || csc_row[k] >= m) {
        cumf::set_last_error("cumf_set_ratings: CSC row index out of range at " + std::to_string(k));
        return CUMF_ERR_INVALID;
      }
  }
  DeviceGuard dg(h->device);
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  h->have_ratings = false;
  h->drop_graph();
  CUMF_TRY(h->rp.alloc(sizeof(int32_t) * (m + 1)));
  CUMF_TRY(h->ci.alloc(sizeof(int32_t) * nnz));
  CUMF_TRY(h->rv.alloc(sizeof(float) * nnz));
  CUMF_TRY(h->cp.alloc(sizeof(int32_t) * (n + 1)));
  CUMF_TRY(h->ri.alloc(sizeof(int32_t) * nnz));
  CUMF_TRY(h->cv.alloc(sizeof(float) * nnz));
  CUMF_TRY(h->X.alloc(sizeof(float) * m * h->ldf));
  CUMF_TRY(h->T.alloc(sizeof(float) * n * h->ldf));
  if (h->opt.precision != CUMF_PREC_FP32) {
    CUMF_TRY(h->Xb.alloc(sizeof(uint16_t) * m * h->ldb));
    CUMF_TRY(h->Tb.alloc(sizeof(uint16_t) * n * h->ldb));
    CUMF_HIP_TRY(hipMemset(h->Xb.p, 0, h->Xb.bytes));
    CUMF_HIP_TRY(hipMemset(h->Tb.p, 0, h->Tb.bytes));
  }
  CUMF_HIP_TRY(hipMemcpy(h->rp.p, csr_row_ptr, h->rp.bytes, hipMemcpyHostToDevice));
  if (nnz > 0) {
    CUMF_HIP_TRY(hipMemcpy(h->ci.p, csr_col, h->ci.bytes, hipMemcpyHostToDevice));
    CUMF_HIP_TRY(hipMemcpy(h->rv.p, csr_val, h->rv.bytes, hipMemcpyHostToDevice));
  }
  std::vector<int32_t> cp_host;
  if (csc_col_ptr) {
    CUMF_HIP_TRY(hipMemcpy(h->cp.p, csc_col_ptr, h->cp.bytes, hipMemcpyHostToDevice));
    if (nnz > 0) {
      CUMF_HIP_TRY(hipMemcpy(h->ri.p, csc_row, h->ri.bytes, hipMemcpyHostToDevice));
      CUMF_HIP_TRY(hipMemcpy(h->cv.p, csc_val, h->cv.bytes, hipMemcpyHostToDevice));
    }
  } else {
    // Stable device transpose; the column pointers come back for planning.
    CUMF_TRY(cumf_csr_to_csc_device(h->stream, m, n, nnz, h->rp.as<int32_t>(), h->ci.as<int32_t>(),
                                    h->rv.as<float>(), h->cp.as<int32_t>(), h->ri.as<int32_t>(),
                                    h->cv.as<float>()));
    cp_host.resize(n + 1);
    CUMF_HIP_TRY(hipMemcpy(cp_host.data(), h->cp.p, h->cp.bytes, hipMemcpyDeviceToHost));
    csc_col_ptr = cp_host.data();
  }
  CUMF_HIP_TRY(hipMemset(h->X.p, 0, h->X.bytes));
  CUMF_HIP_TRY(hipMemset(h->T.p, 0, h->T.bytes));
  h->rp_host.assign(csr_row_ptr, csr_row_ptr + m + 1);
  h->cp_host.assign(csc_col_ptr, csc_col_ptr + n + 1);
  CUMF_TRY(h->plan_x.build(csr_row_ptr, m, h->tile_split[0] > 0 ? h->tile_split[0] : h->opt.split, h->f));
  CUMF_TRY(h->plan_t.build(csc_col_ptr, n, h->tile_split[1] > 0 ? h->tile_split[1] : h->opt.split, h->f));
  h->m = m;
  h->n = n;
  h->nnz = nnz;
  h->have_ratings = true;
  return CUMF_OK;
}

int cumf_set_heldout(cumf_handle* h, int64_t count, const int32_t* u, const int32_t* v,
                     const float* r) {
  cumf::clear_last_error();
  if (!h || !h->have_ratings || count < 0 || (count > 0 && (!u || !v || !r))) {
    cumf::set_last_error("cumf_set_heldout: bad arguments or no ratings yet");
    return CUMF_ERR_INVALID;
  }
  for (int64_t k = 0; k < count; ++k)
    if (u[k] < 0 || u[k] >= h->m || v[k] < 0 || v[k] >= h->n) {
      cumf::set_last_error("cumf_set_heldout: index out of range at " + std::to_string(k));
      return CUMF_ERR_INVALID;
    }
  DeviceGuard dg(h->device);
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  CUMF_TRY(h->tu.alloc(sizeof(int32_t) * count));
  CUMF_TRY(h->tv.alloc(sizeof(int32_t) * count));
  CUMF_TRY(h->tr.alloc(sizeof(float) * count));
  if (count > 0) {
    CUMF_HIP_TRY(hipMemcpy(h->tu.p, u, h->tu.bytes, hipMemcpyHostToDevice));
    CUMF_HIP_TRY(hipMemcpy(h->tv.p, v, h->tv.bytes, hipMemcpyHostToDevice));
    CUMF_HIP_TRY(hipMemcpy(h->tr.p, r, h->tr.bytes, hipMemcpyHostToDevice));
  }
  h->ntest = count;
  return CUMF_OK;
}

int cumf_set_factors(cumf_handle* h, const float* X, const float* theta) {
  cumf::clear_last_error();
  if (!h || !h->have_ratings) {
    cumf::set_last_error("cumf_set_factors: no ratings yet");
    return CUMF_ERR_STATE;
  }
  DeviceGuard dg(h->device);
  if (X) CUMF_TRY(upload_factors(h, h->X, h->Xb, X, h->m));
  if (theta) CUMF_TRY(upload_factors(h, h->T, h->Tb, theta, h->n));
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  h->x_sweeps = 0;  // a new start: the transient sweeps begin again
  return CUMF_OK;
}

int cumf_get_factors(cumf_handle* h, float* X, float* theta) {
  cumf::clear_last_error();
  if (!h || !h->have_ratings) {
    cumf::set_last_error("cumf_get_factors: no ratings yet");
    return CUMF_ERR_STATE;
  }
  DeviceGuard dg(h->device);
  if (X)
    CUMF_HIP_TRY(hipMemcpy2DAsync(X, sizeof(float) * h->f, h->X.p, sizeof(float) * h->ldf,
                                  sizeof(float) * h->f, h->m, hipMemcpyDeviceToHost, h->stream));
  if (theta)
    CUMF_HIP_TRY(hipMemcpy2DAsync(theta, sizeof(float) * h->f, h->T.p, sizeof(float) * h->ldf,
                                  sizeof(float) * h->f, h->n, hipMemcpyDeviceToHost, h->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  return CUMF_OK;
}

namespace {
int half_sweep(cumf_handle* h, int side);
}  // namespace

int cumf_update_x(cumf_handle* h) {
  if (!h || !h->have_ratings) {
    cumf::set_last_error("cumf_update_x: no ratings yet");
    return CUMF_ERR_STATE;
  }
  DeviceGuard dg(h->device);
  return half_sweep(h, 0);
}

int cumf_update_theta(cumf_handle* h) {
  if (!h || !h->have_ratings) {
    cumf::set_last_error("cumf_update_theta: no ratings yet");
    return CUMF_ERR_STATE;
  }
  DeviceGuard dg(h->device);
  return half_sweep(h, 1);
}

namespace {

// Captures one half-sweep (all its kernel launches, including the warm
// start's features and network; no memsets or host work) on the handle's
// stream into an executable graph.
int capture_half_sweep(cumf_handle* h, int side) {
  if (h->graph[side]) {
    (void)hipGraphExecDestroy(h->graph[side]);
    h->graph[side] = nullptr;
  }
  CUMF_TRY(prepare_network(h));
  if (h->opt.solver == CUMF_SOLVER_CG) {
    const size_t need = sizeof(int) * (size_t)std::max<int64_t>(std::max(h->m, h->n), 1);
    if (h->fb_list.bytes < need) CUMF_TRY(h->fb_list.alloc(need));
    if (!h->fb_count.p) CUMF_TRY(h->fb_count.alloc(sizeof(int)));
  }
  hipGraph_t g = nullptr;
  CUMF_HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  h->capturing = true;
  const int64_t flops = h->gram_flops;
  const int rc = side == 0 ? run_half_sweep(h, 0, h->plan_x, h->rp, h->ci, h->rv, h->T.as<float>(),
                                            h->X.as<float>(), h->Tb, h->Xb)
                           : run_half_sweep(h, 1, h->plan_t, h->cp, h->ri, h->cv, h->X.as<float>(),
                                            h->T.as<float>(), h->Xb, h->Tb);
  h->capturing = false;
  h->gram_flops = flops;  // counted again at each replay
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc != CUMF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    cumf::set_last_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return CUMF_ERR_HIP;
  }
  e = hipGraphInstantiate(&h->graph[side], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    h->graph[side] = nullptr;
    cumf::set_last_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return CUMF_ERR_HIP;
  }
  return CUMF_OK;
}

// One half-sweep: replayed from its graph with use_graph, else launched.
int half_sweep(cumf_handle* h, int side) {
  if (h->opt.use_graph) {
    // a captured X half-sweep that runs the network (or not) is recaptured
    // when network_now() changes
    const bool net = side == 0 && network_now(h);
    if (!h->graph[side] || h->graph_net[side] != net) {
      CUMF_TRY(capture_half_sweep(h, side));
      h->graph_net[side] = net;
    }
    CUMF_TRY(timed(h, Timer::ITER, [&] { return hipGraphLaunch(h->graph[side], h->stream); }));
    h->gram_flops += (side == 0 ? h->plan_x.nnz : h->plan_t.nnz) * (int64_t)h->f * (h->f + 1);
    if (side == 0) ++h->x_sweeps;
    return CUMF_OK;
  }
  const int rc = side == 0 ? run_half_sweep(h, 0, h->plan_x, h->rp, h->ci, h->rv, h->T.as<float>(),
                                            h->X.as<float>(), h->Tb, h->Xb)
                           : run_half_sweep(h, 1, h->plan_t, h->cp, h->ri, h->cv, h->X.as<float>(),
                                            h->T.as<float>(), h->Xb, h->Tb);
  if (rc == CUMF_OK && side == 0) ++h->x_sweeps;
  return rc;
}

}  // namespace

int cumf_iterate(cumf_handle* h, int iters) {
  if (!h || !h->have_ratings) {
    cumf::set_last_error("cumf_iterate: no ratings yet");
    return CUMF_ERR_STATE;
  }
  if (iters < 0) return CUMF_ERR_INVALID;
  for (int it = 0; it < iters; ++it) {
    CUMF_TRY(cumf_update_x(h));
    CUMF_TRY(cumf_update_theta(h));
  }
  return CUMF_OK;
}

int cumf_heldout_rmse(cumf_handle* h, double* rmse) {
  cumf::clear_last_error();
  if (!h || !rmse || !h->have_ratings) {
    cumf::set_last_error("cumf_heldout_rmse: no ratings yet");
    return CUMF_ERR_STATE;
  }
  if (h->ntest == 0) {
    *rmse = 0.0;
    return CUMF_OK;
  }
  DeviceGuard dg(h->device);
  CUMF_TRY(timed(h, Timer::RMSE, [&] {
    return cumf::launch_rmse(h->tu.as<int32_t>(), h->tv.as<int32_t>(), h->tr.as<float>(), h->ntest,
                             h->X.as<float>(), h->T.as<float>(), h->ldf, h->f,
                             h->rmse_partial.as<double>(), kRmseBlocks, h->rmse_out.as<double>(),
                             h->stream);
  }));
  double se = 0.0;
  CUMF_HIP_TRY(hipMemcpyAsync(&se, h->rmse_out.p, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  *rmse = std::sqrt(se / (double)h->ntest);
  return CUMF_OK;
}

int cumf_synchronize(cumf_handle* h) {
  if (!h) return CUMF_ERR_INVALID;
  DeviceGuard dg(h->device);
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  h->timer.resolve();
  return CUMF_OK;
}

int cumf_get_stats(cumf_handle* h, cumf_stats* s) {
  if (!h || !s) return CUMF_ERR_INVALID;
  CUMF_TRY(cumf_synchronize(h));
  DeviceGuard dg(h->device);
  int fails[2] = {0, 0};
  CUMF_HIP_TRY(hipMemcpy(fails, h->fail.p, 2 * sizeof(int), hipMemcpyDeviceToHost));
  std::memset(s, 0, sizeof(*s));
  s->task_ms = h->timer.ms[Timer::TASK];
  s->task_launches = h->timer.count[Timer::TASK];
  s->finish_ms = h->timer.ms[Timer::FINISH];
  s->finish_launches = h->timer.count[Timer::FINISH];
  s->rmse_ms = h->timer.ms[Timer::RMSE];
  s->rmse_launches = h->timer.count[Timer::RMSE];
  s->sweep_ms = h->timer.ms[Timer::SWEEP];
  s->sweeps = h->timer.count[Timer::SWEEP];
  s->failed_rows = fails[0];
  s->cg_fallback_rows = fails[1];
  s->warm_ms = h->timer.ms[Timer::WARM];
  s->warm_launches = h->timer.count[Timer::WARM];
  s->fallback_ms = h->timer.ms[Timer::FALLBACK];
  s->fallback_launches = h->timer.count[Timer::FALLBACK];
  s->gram_flops = h->gram_flops;
  s->iter_ms = h->timer.ms[Timer::ITER];
  s->iters = h->timer.count[Timer::ITER];
  return CUMF_OK;
}

int cumf_reset_stats(cumf_handle* h) {
  if (!h) return CUMF_ERR_INVALID;
  CUMF_TRY(cumf_synchronize(h));
  DeviceGuard dg(h->device);
  for (int k = 0; k < Timer::NKINDS; ++k) {
    h->timer.ms[k] = 0.0;
    h->timer.count[k] = 0;
  }
  h->gram_flops = 0;
  CUMF_HIP_TRY(hipMemset(h->fail.p, 0, 2 * sizeof(int)));
  return CUMF_OK;
}

hipStream_t cumf_get_stream(cumf_handle* h) { return h ? h->stream : nullptr; }

int cumf_debug_large_dual(int on) { return g_large_dual.exchange(on ? 1 : 0); }

int cumf_debug_gram_dim(int f) { return (f >= 1 && f <= kMaxFusedF) ? cumf::dump_dim(f) : -1; }

int cumf_debug_plan(const int32_t* row_ptr, int64_t m, int split, int f, int32_t* tasks, int64_t task_cap,
                    int64_t* ntasks, int32_t* heavy, int64_t heavy_cap, int64_t* nheavy) {
  cumf::clear_last_error();
  if (!row_ptr || m < 0 || !ntasks || !nheavy || check_f(f) != CUMF_OK || split < 1) {
    cumf::set_last_error("cumf_debug_plan: bad arguments");
    return CUMF_ERR_INVALID;
  }
  std::vector<cumf::Task> t;
  std::vector<cumf::HeavyRow> hv;
  int64_t slots = 0;
  CUMF_TRY(plan_tasks(row_ptr, m, split, f, t, hv, slots));
  *ntasks = (int64_t)t.size();
  *nheavy = (int64_t)hv.size();
  if ((int64_t)t.size() > task_cap || (int64_t)hv.size() > heavy_cap || (!t.empty() && !tasks) ||
      (!hv.empty() && !heavy)) {
    cumf::set_last_error("cumf_debug_plan: output arrays too small");
    return CUMF_ERR_INVALID;
  }
  if (!t.empty()) std::memcpy(tasks, t.data(), sizeof(cumf::Task) * t.size());
  if (!hv.empty()) std::memcpy(heavy, hv.data(), sizeof(cumf::HeavyRow) * hv.size());
  return CUMF_OK;
}

int cumf_debug_row_gram(cumf_handle* h, int side, const int32_t* rows, int64_t nrows, float* out) {
  cumf::clear_last_error();
  if (!h || !h->have_ratings || (side != 0 && side != 1) || nrows < 0 || (nrows > 0 && (!rows || !out))) {
    cumf::set_last_error("cumf_debug_row_gram: bad arguments");
    return CUMF_ERR_INVALID;
  }
  const int64_t lim = side == 0 ? h->m : h->n;
  for (int64_t i = 0; i < nrows; ++i)
    if (rows[i] < 0 || rows[i] >= lim) {
      cumf::set_last_error("cumf_debug_row_gram: row out of range");
      return CUMF_ERR_INVALID;
    }
  if (nrows == 0) return CUMF_OK;
  DeviceGuard dg(h->device);
  const int D = cumf::dump_dim(h->f);
  DevBuf drows, dout;
  CUMF_TRY(drows.alloc(sizeof(int32_t) * nrows));
  CUMF_TRY(dout.alloc(sizeof(float) * nrows * D * D));
  CUMF_HIP_TRY(hipMemcpy(drows.p, rows, drows.bytes, hipMemcpyHostToDevice));
  cumf::SweepArgs a{};
  a.row_ptr = side == 0 ? h->rp.as<int32_t>() : h->cp.as<int32_t>();
  a.col = side == 0 ? h->ci.as<int32_t>() : h->ri.as<int32_t>();
  a.val = side == 0 ? h->rv.as<float>() : h->cv.as<float>();
  a.theta = side == 0 ? h->T.as<float>() : h->X.as<float>();
  a.ldf = h->ldf;
  a.f = h->f;
  a.lambda = h->lambda;
  a.precision = h->opt.precision;
  a.ldb = h->ldb;
  a.theta_b = side == 0 ? h->Tb.as<uint16_t>() : h->Xb.as<uint16_t>();
  hipError_t e = cumf::launch_dump_kernel(a, drows.as<int32_t>(), nrows, dout.as<float>(), h->stream);
  if (e != hipSuccess) {
    cumf::set_last_error(std::string("dump kernel: ") + hipGetErrorString(e));
    return CUMF_ERR_HIP;
  }
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  CUMF_HIP_TRY(hipMemcpy(out, dout.p, dout.bytes, hipMemcpyDeviceToHost));
  return CUMF_OK;
}

}  // extern "C"

namespace {

int rebuild_plans(cumf_handle* h) {
  h->drop_graph();
  if (!h->have_ratings) return CUMF_OK;
  CUMF_TRY(h->plan_x.build(h->rp_host.data(), h->m, h->tile_split[0] > 0 ? h->tile_split[0] : h->opt.split, h->f));
  CUMF_TRY(h->plan_t.build(h->cp_host.data(), h->n, h->tile_split[1] > 0 ? h->tile_split[1] : h->opt.split, h->f));
  return CUMF_OK;
}

// Minimum over reps of the stream time of launch() (milliseconds); prep()
// runs before each rep, outside the timed interval (it puts back the
// factors the timed sweeps overwrite, so that every rep solves from the
// same start).
template <class P, class F>
int time_min_ms(cumf_handle* h, int reps, P&& prep, F&& launch, double* out) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CUMF_HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return CUMF_ERR_HIP;
  }
  double best = 1e300;
  int rc = CUMF_OK;
  for (int r = 0; r < reps && rc == CUMF_OK; ++r) {
    rc = prep();
    if (rc != CUMF_OK) break;
    (void)hipEventRecord(e0, h->stream);
    rc = launch();
    (void)hipEventRecord(e1, h->stream);
    if (rc == CUMF_OK && hipEventSynchronize(e1) != hipSuccess) rc = CUMF_ERR_HIP;
    float ms = 0.f;
    if (rc == CUMF_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) best = std::min(best, (double)ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *out = best;
  return rc;
}

struct SideBufs {
  const Plan* p;
  const DevBuf *rowp, *col, *val;
  const float* theta;
  float* x;
  const DevBuf* theta_b;
  DevBuf* x_b;
};

SideBufs side_bufs(cumf_handle* h, int side) {
  if (side == 0) return {&h->plan_x, &h->rp, &h->ci, &h->rv, h->T.as<float>(), h->X.as<float>(), &h->Tb, &h->Xb};
  return {&h->plan_t, &h->cp, &h->ri, &h->cv, h->X.as<float>(), h->T.as<float>(), &h->Xb, &h->Tb};
}

int copy_dev(DevBuf& dst, const DevBuf& src, hipStream_t s);

// The factors a side's timed sweeps overwrite, as they were when tuning
// started (FP32 and, in precision mode, BF16).
struct TuneStart {
  const DevBuf* x;
  const DevBuf* xb;
};

int autotune_side(cumf_handle* h, int side, int reps, const TuneStart& start) {
  const std::vector<int32_t>& rp = side == 0 ? h->rp_host : h->cp_host;
  const int64_t rows = side == 0 ? h->m : h->n;
  int64_t maxdeg = 0;
  for (int64_t u = 0; u < rows; ++u) maxdeg = std::max<int64_t>(maxdeg, (int64_t)rp[u + 1] - rp[u]);
  DevBuf& xo = side == 0 ? h->X : h->T;
  DevBuf& xbo = side == 0 ? h->Xb : h->Tb;
  auto restore = [&] {
    CUMF_TRY(copy_dev(xo, *start.x, h->stream));
    if (start.xb->p) CUMF_TRY(copy_dev(xbo, *start.xb, h->stream));
    return CUMF_OK;
  };
  // 1. Split size: whole half-sweeps, one launch per kernel.
  h->tiled = false;
  static const int kSplits[] = {1024, 2048, 4096, 8192, 16384, 32768};
  double best = 1e300;
  int best_split = h->opt.split;
  for (int sp : kSplits) {
    if (sp > 2 * maxdeg && sp != kSplits[0]) break;
    h->tile_split[side] = sp;
    CUMF_TRY(rebuild_plans(h));
    double ms = 0;
    CUMF_TRY(time_min_ms(h, reps, restore, [&] {
      SideBufs b = side_bufs(h, side);
      return run_half_sweep(h, side, *b.p, *b.rowp, *b.col, *b.val, b.theta, b.x, *b.theta_b, *b.x_b);
    }, &ms));
    if (ms < best) {
      best = ms;
      best_split = sp;
    }
  }
  h->tile_split[side] = best_split;
  CUMF_TRY(rebuild_plans(h));
  // 2. Launch shape per degree bucket: register budget x (one task per
  //    group | persistent grid | occupancy cap), and on a CG handle the
  //    solver (a CG bucket is timed with the Cholesky re-solve of the rows
  //    it hands back).
  SideBufs b = side_bufs(h, side);
  cumf::SweepArgs a{};
  CUMF_TRY(fill_args(h, *b.p, *b.rowp, *b.col, *b.val, b.theta, b.x, *b.theta_b, *b.x_b, nullptr, a));
  a.warm = side == 0 ? std::min(h->opt.warm_start, (int)CUMF_WARM_ZERO) : CUMF_WARM_PREV;
  const bool cg = h->opt.solver == CUMF_SOLVER_CG;
  const int dv = default_variant(h->f);
  std::vector<int> solvers = cg ? std::vector<int>{CUMF_SOLVER_CG, CUMF_SOLVER_CHOLESKY} : std::vector<int>{-1};
  for (int k = 0; k < kBuckets; ++k) {
    if (b.p->bucket_t1[k] <= b.p->bucket_t0[k]) continue;
    double bk = 1e300;
    BucketCfg bc;
    std::vector<int> sols = solvers;
    if (k < kDualBuckets) sols.push_back(CUMF_SOLVER_DUAL);
    for (int sol : sols)
      for (int variant : {2, 1})
        for (int pers : {0, h->num_cus * (variant == 1 ? 4 : 8), -7, -6, -3}) {
          // pers < 0: a one-task-per-group launch capped at -pers wavefronts
          // per CU (7 or 6 of the eight a two-wavefront budget allows, 3 of
          // the four of a one-wavefront budget)
          if (pers < 0 && (variant == 1) != (pers == -3)) continue;
          if (sol == CUMF_SOLVER_DUAL && pers != 0) continue;  // one task per group, uncapped
          BucketCfg cfg{variant, std::max(pers, 0), sol, std::max(-pers, 0)};
          double ms = 0;
          CUMF_TRY(time_min_ms(h, reps, restore, [&] {
            if (cg) CUMF_HIP_TRY(cumf::launch_fallback_reset(a.fb_count, 1, h->stream));
            if (cfg.persistent)
              CUMF_HIP_TRY(cumf::launch_fallback_reset(h->counters.as<int>() + side * kBuckets + k, 1, h->stream));
            CUMF_HIP_TRY(bucket_launch(h, side, *b.p, a, k, cfg));
            if (cg) CUMF_HIP_TRY(fallback_launch(a, dv, h->stream));
            return CUMF_OK;
          }, &ms));
          if (ms < bk) {
            bk = ms;
            bc = cfg;
          }
        }
    h->tiles[side][k] = bc;
  }
  // 3. Register budget of the split-row finish and of the Cholesky fallback
  //    (timed on a list of up to 8,192 light rows).
  h->finish_variant[side] = 2;
  if (b.p->nheavy > 0) {
    double t2 = 0, t1 = 0;
    auto fin = [&](int v) {
      return [&, v] {
        if (cg) CUMF_HIP_TRY(cumf::launch_fallback_reset(a.fb_count, 1, h->stream));
        CUMF_HIP_TRY(finish_launch(a, v, h->stream));
        return CUMF_OK;
      };
    };
    CUMF_TRY(time_min_ms(h, reps, restore, fin(2), &t2));
    CUMF_TRY(time_min_ms(h, reps, restore, fin(1), &t1));
    h->finish_variant[side] = t1 < t2 ? 1 : 2;
  }
  h->fallback_variant[side] = 2;
  if (cg) {
    std::vector<int> list;
    const int64_t step = std::max<int64_t>(1, rows / 8192);
    for (int64_t u = 0; u < rows && (int64_t)list.size() < 8192; u += step)
      if ((int64_t)rp[u + 1] - rp[u] <= b.p->split) list.push_back((int)u);
    const int cnt = (int)list.size();
    if (cnt > 0) {
      double t2 = 0, t1 = 0;
      auto fb = [&](int v) {
        return [&, v] {
          CUMF_HIP_TRY(hipMemcpyAsync(h->fb_list.p, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice,
                                      h->stream));
          CUMF_HIP_TRY(hipMemcpyAsync(h->fb_count.p, &cnt, sizeof(int), hipMemcpyHostToDevice, h->stream));
          CUMF_HIP_TRY(fallback_launch(a, v, h->stream));
          return CUMF_OK;
        };
      };
      CUMF_TRY(time_min_ms(h, reps, restore, fb(2), &t2));
      CUMF_TRY(time_min_ms(h, reps, restore, fb(1), &t1));
      h->fallback_variant[side] = t1 < t2 ? 1 : 2;
      CUMF_HIP_TRY(cumf::launch_fallback_reset(a.fb_count, 1, h->stream));
    }
  }
  // 4. Buckets launched one after another drain the GPU between launches:
  //    keep the per-bucket shapes only if a whole half-sweep with them beats
  //    one launch of every task with the default shape -- with the handle's
  //    solver, or on a CG handle with the Cholesky solver for the whole side.
  {
    BucketCfg tuned[kBuckets];
    std::copy(h->tiles[side], h->tiles[side] + kBuckets, tuned);
    const bool was_tiled = h->tiled;
    h->tiled = true;
    auto sweep = [&] {
      SideBufs sb = side_bufs(h, side);
      return run_half_sweep(h, side, *sb.p, *sb.rowp, *sb.col, *sb.val, sb.theta, sb.x, *sb.theta_b, *sb.x_b);
    };
    double t_best = 0;
    CUMF_TRY(time_min_ms(h, reps, restore, sweep, &t_best));
    for (int sol : solvers) {
      for (int k = 0; k < kBuckets; ++k) h->tiles[side][k] = BucketCfg{dv, 0, sol};
      double t = 0;
      CUMF_TRY(time_min_ms(h, reps, restore, sweep, &t));
      if (t < t_best) {
        t_best = t;
        for (int k = 0; k < kBuckets; ++k) tuned[k] = BucketCfg{dv, 0, sol};
      }
    }
    std::copy(tuned, tuned + kBuckets, h->tiles[side]);
    h->tiled = was_tiled;
  }
  return CUMF_OK;
}

int copy_dev(DevBuf& dst, const DevBuf& src, hipStream_t s) {
  if (!src.p) return CUMF_OK;
  if (dst.bytes != src.bytes) CUMF_TRY(dst.alloc(src.bytes));
  CUMF_HIP_TRY(hipMemcpyAsync(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice, s));
  return CUMF_OK;
}

}  // namespace

extern "C" {

int cumf_save_tile_table(cumf_handle* h, const char* path) {
  cumf::clear_last_error();
  if (!h || !path) return CUMF_ERR_INVALID;
  FILE* fp = std::fopen(path, "w");
  if (!fp) {
    cumf::set_last_error(std::string("cannot write ") + path);
    return CUMF_ERR_INVALID;
  }
  std::fprintf(fp, "# cumf tile table: f %d precision %d solver %d\n", h->f, h->opt.precision, h->opt.solver);
  std::fprintf(fp, "# side <0 X | 1 Theta> split <ratings per task> finish <1|2> fallback <1|2>\n");
  std::fprintf(fp, "# bucket <side> <b> <min ratings> <max ratings> variant <1|2> persistent <wavefronts>"
                  " solver <-1 handle's | 0 Cholesky | 1 CG | 2 dual> occupancy <wavefronts per CU, 0 any>\n");
  for (int side = 0; side < 2; ++side) {
    const Plan& p = side == 0 ? h->plan_x : h->plan_t;
    std::fprintf(fp, "side %d split %d finish %d fallback %d\n", side, p.split > 0 ? p.split : h->opt.split,
                 h->finish_variant[side], h->fallback_variant[side]);
    for (int b = 0; b < kBuckets; ++b)
      std::fprintf(fp, "bucket %d %d %lld %lld variant %d persistent %d solver %d occupancy %d\n", side, b,
                   (long long)kBucketEdge[b], (long long)(b + 1 < kBuckets ? kBucketEdge[b + 1] - 1 : -1),
                   h->tiles[side][b].variant, h->tiles[side][b].persistent, h->tiles[side][b].solver,
                   h->tiles[side][b].occupancy);
  }
  const bool ok = std::fclose(fp) == 0;
  if (!ok) cumf::set_last_error(std::string("cannot write ") + path);
  return ok ? CUMF_OK : CUMF_ERR_INVALID;
}

int cumf_load_tile_table(cumf_handle* h, const char* path) {
  cumf::clear_last_error();
  if (!h || !path) return CUMF_ERR_INVALID;
  FILE* fp = std::fopen(path, "r");
  if (!fp) {
    cumf::set_last_error(std::string("cannot read ") + path);
    return CUMF_ERR_INVALID;
  }
  int split[2] = {0, 0}, fin[2] = {2, 2}, fb[2] = {2, 2};
  BucketCfg tiles[2][kBuckets];
  char line[512];
  int rc = CUMF_OK, lineno = 0;
  while (rc == CUMF_OK && std::fgets(line, sizeof line, fp)) {
    ++lineno;
    if (line[0] == '#' || line[0] == '\n') continue;
    int side, sp, fv, bv, b, v, pers, sol = -1, occ = 0, nf = 0;
    long long lo, hi;
    if (std::sscanf(line, "side %d split %d finish %d fallback %d", &side, &sp, &fv, &bv) == 4) {
      if (side < 0 || side > 1 || sp < 1 || (fv != 1 && fv != 2) || (bv != 1 && bv != 2)) rc = CUMF_ERR_INVALID;
      else { split[side] = sp; fin[side] = fv; fb[side] = bv; }
    } else if ((nf = std::sscanf(line, "bucket %d %d %lld %lld variant %d persistent %d solver %d occupancy %d",
                                 &side, &b, &lo, &hi, &v, &pers, &sol, &occ)) >= 6) {
      // A bucket may run the CG solver only on a handle built for CG (its
      // fallback list and kernel); Cholesky is always available.
      if (nf == 6) sol = -1;
      if (nf <= 7) occ = 0;
      // The dual form needs rows of at most 95 ratings (buckets 0 to 2),
      // one task per group.
      if (side < 0 || side > 1 || b < 0 || b >= kBuckets || (v != 1 && v != 2) || pers < 0 || pers > (1 << 20) ||
          sol < -1 || sol > 2 || (sol == CUMF_SOLVER_CG && h->opt.solver != CUMF_SOLVER_CG) ||
          (sol == CUMF_SOLVER_DUAL && (b >= kDualBuckets || pers > 0)) ||
          occ < 0 || occ > cumf::kMaxOccupancy || (occ > 0 && occ < cumf::kMinOccupancy))
        rc = CUMF_ERR_INVALID;
      else tiles[side][b] = {v, pers, sol, occ};
    } else {
      rc = CUMF_ERR_INVALID;
    }
  }
  std::fclose(fp);
  if (rc != CUMF_OK) {
    cumf::set_last_error(std::string(path) + ": bad tile table line " + std::to_string(lineno));
    return rc;
  }
  DeviceGuard dg(h->device);
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  for (int side = 0; side < 2; ++side) {
    h->tile_split[side] = split[side];
    h->finish_variant[side] = fin[side];
    h->fallback_variant[side] = fb[side];
    for (int b = 0; b < kBuckets; ++b) h->tiles[side][b] = tiles[side][b];
  }
  h->tiled = true;
  return rebuild_plans(h);
}

int cumf_autotune(cumf_handle* h, int reps, const char* path) {
  cumf::clear_last_error();
  if (!h) return CUMF_ERR_INVALID;
  if (!h->have_ratings) {
    cumf::set_last_error("cumf_autotune: no ratings yet");
    return CUMF_ERR_STATE;
  }
  if (h->f > kMaxFusedF) {
    cumf::set_last_error("cumf_autotune: f > " + std::to_string(kMaxFusedF) +
                         " runs the work-group-per-row kernel, which has a single launch shape");
    return CUMF_ERR_UNSUPPORTED;
  }
  if (h->opt.compat_hermitian_bug) {
    cumf::set_last_error("cumf_autotune: the compat path has a single launch shape");
    return CUMF_ERR_INVALID;
  }
  reps = std::max(1, std::min(reps, 20));
  DeviceGuard dg(h->device);
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  // The timed sweeps solve real rows: keep the factors and put them back.
  DevBuf sx, st, sxb, stb;
  CUMF_TRY(copy_dev(sx, h->X, h->stream));
  CUMF_TRY(copy_dev(st, h->T, h->stream));
  CUMF_TRY(copy_dev(sxb, h->Xb, h->stream));
  CUMF_TRY(copy_dev(stb, h->Tb, h->stream));
  const bool timing = h->opt.timing;
  const int warm = h->opt.warm_start;
  h->opt.timing = 0;
  if (warm >= CUMF_WARM_MLP) h->opt.warm_start = CUMF_WARM_PREV;  // the network is not part of the tile choice
  int rc = autotune_side(h, 0, reps, TuneStart{&sx, &sxb});
  if (rc == CUMF_OK) rc = autotune_side(h, 1, reps, TuneStart{&st, &stb});
  h->opt.timing = timing;
  h->opt.warm_start = warm;
  // The factors are put back whether or not tuning succeeded; the first
  // error is the one reported.
  auto keep_first = [&rc](int r) {
    if (rc == CUMF_OK) rc = r;
  };
  keep_first(copy_dev(h->X, sx, h->stream));
  keep_first(copy_dev(h->T, st, h->stream));
  if (sxb.p) keep_first(copy_dev(h->Xb, sxb, h->stream));
  if (stb.p) keep_first(copy_dev(h->Tb, stb, h->stream));
  if (hipStreamSynchronize(h->stream) != hipSuccess) keep_first(CUMF_ERR_HIP);
  // Counters of the tuning runs are not statistics of the user's sweeps.
  if (rc == CUMF_OK && hipMemset(h->fail.p, 0, 2 * sizeof(int)) != hipSuccess) rc = CUMF_ERR_HIP;
  h->tiled = rc == CUMF_OK;
  h->drop_graph();
  if (rc != CUMF_OK) {
    h->tile_split[0] = h->tile_split[1] = 0;
    (void)rebuild_plans(h);
    return rc;
  }
  if (path) CUMF_TRY(cumf_save_tile_table(h, path));
  return CUMF_OK;
}

int cumf_debug_cg_trace(cumf_handle* h, int side, float* out) {
  cumf::clear_last_error();
  if (!h || !out || (side != 0 && side != 1)) return CUMF_ERR_INVALID;
  if (!h->have_ratings) {
    cumf::set_last_error("cumf_debug_cg_trace: no ratings yet");
    return CUMF_ERR_STATE;
  }
  DeviceGuard dg(h->device);
  const int64_t rows = side == 0 ? h->m : h->n;
  DevBuf tr;
  CUMF_TRY(tr.alloc(sizeof(float) * 4 * (size_t)rows));
  CUMF_HIP_TRY(hipMemsetAsync(tr.p, 0, tr.bytes, h->stream));
  if (side == 0)
    CUMF_TRY(run_half_sweep(h, 0, h->plan_x, h->rp, h->ci, h->rv, h->T.as<float>(), h->X.as<float>(), h->Tb,
                            h->Xb, tr.as<float>()));
  else
    CUMF_TRY(run_half_sweep(h, 1, h->plan_t, h->cp, h->ri, h->cv, h->X.as<float>(), h->T.as<float>(), h->Xb,
                            h->Tb, tr.as<float>()));
  CUMF_HIP_TRY(hipMemcpyAsync(out, tr.p, tr.bytes, hipMemcpyDeviceToHost, h->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  return CUMF_OK;
}

int cumf_set_mlp(cumf_handle* h, cumf_mlp* mlp) {
  cumf::clear_last_error();
  if (!h) return CUMF_ERR_INVALID;
  if (mlp && (cumf::mlp_f(mlp) != h->f || cumf::mlp_device(mlp) != h->device)) {
    cumf::set_last_error("cumf_set_mlp: network built for another rank or device");
    return CUMF_ERR_INVALID;
  }
  DeviceGuard dg(h->device);
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  h->mlp = mlp;
  h->drop_graph();
  return CUMF_OK;
}

int cumf_mlp_harvest(cumf_handle* h, cumf_mlp* mlp, int iters) {
  cumf::clear_last_error();
  if (!h || !mlp || iters < 0) return CUMF_ERR_INVALID;
  if (!h->have_ratings) {
    cumf::set_last_error("cumf_mlp_harvest: no ratings yet");
    return CUMF_ERR_STATE;
  }
  if (cumf::mlp_f(mlp) != h->f || cumf::mlp_device(mlp) != h->device) {
    cumf::set_last_error("cumf_mlp_harvest: network built for another rank or device");
    return CUMF_ERR_INVALID;
  }
  DeviceGuard dg(h->device);
  // The plain solver produces the targets.
  const cumf_options saved = h->opt;
  h->opt.solver = CUMF_SOLVER_CHOLESKY;
  h->opt.warm_start = CUMF_WARM_PREV;
  int rc = CUMF_OK;
  for (int it = 0; it < iters && rc == CUMF_OK; ++it) {
    __bf16* feats = nullptr;
    float* targets = nullptr;
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
      rc = CUMF_ERR_HIP;
      break;
    }
    rc = cumf::mlp_harvest_reserve(mlp, h->m, &feats, &targets);
    if (rc != CUMF_OK) break;
    if (cumf::launch_mlp_features(h->rp.as<int32_t>(), h->ci.as<int32_t>(), h->rv.as<float>(), h->T.as<float>(),
                                  h->ldf, h->f, h->lambda, 0, h->m, cumf::mlp_k0(mlp), feats,
                                  h->stream) != hipSuccess) {
      rc = CUMF_ERR_HIP;
      break;
    }
    rc = run_half_sweep(h, 0, h->plan_x, h->rp, h->ci, h->rv, h->T.as<float>(), h->X.as<float>(), h->Tb, h->Xb);
    if (rc != CUMF_OK) break;
    if (hipMemcpy2DAsync(targets, sizeof(float) * h->f, h->X.p, sizeof(float) * h->ldf, sizeof(float) * h->f,
                         h->m, hipMemcpyDeviceToDevice, h->stream) != hipSuccess) {
      rc = CUMF_ERR_HIP;
      break;
    }
    rc = run_half_sweep(h, 1, h->plan_t, h->cp, h->ri, h->cv, h->X.as<float>(), h->T.as<float>(), h->Xb, h->Tb);
  }
  h->opt = saved;
  if (rc == CUMF_ERR_HIP) cumf::set_last_error("cumf_mlp_harvest: HIP error");
  if (rc == CUMF_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = CUMF_ERR_HIP;
  return rc;
}

int cumf_debug_features(cumf_handle* h, int side, const int32_t* rows, int64_t nrows, float* out) {
  cumf::clear_last_error();
  if (!h || !rows || !out || nrows < 0 || (side != 0 && side != 1)) return CUMF_ERR_INVALID;
  if (!h->have_ratings) {
    cumf::set_last_error("cumf_debug_features: no ratings yet");
    return CUMF_ERR_STATE;
  }
  DeviceGuard dg(h->device);
  const int K0 = (2 * h->f + 1 + 31) / 32 * 32;
  const int64_t total = side == 0 ? h->m : h->n;
  for (int64_t k = 0; k < nrows; ++k)
    if (rows[k] < 0 || rows[k] >= total) return CUMF_ERR_INVALID;
  DevBuf fb;
  CUMF_TRY(fb.alloc(2 * (size_t)total * K0));
  const DevBuf& rp = side == 0 ? h->rp : h->cp;
  const DevBuf& ci = side == 0 ? h->ci : h->ri;
  const DevBuf& cv = side == 0 ? h->rv : h->cv;
  const float* th = side == 0 ? h->T.as<float>() : h->X.as<float>();
  CUMF_HIP_TRY(cumf::launch_mlp_features(rp.as<int32_t>(), ci.as<int32_t>(), cv.as<float>(), th, h->ldf, h->f,
                                         h->lambda, 0, total, K0, fb.as<__bf16>(), h->stream));
  std::vector<uint16_t> host((size_t)total * K0);
  CUMF_HIP_TRY(hipMemcpyAsync(host.data(), fb.p, fb.bytes, hipMemcpyDeviceToHost, h->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  for (int64_t k = 0; k < nrows; ++k)
    for (int i = 0; i < K0; ++i) {
      const uint32_t bits = (uint32_t)host[(size_t)rows[k] * K0 + i] << 16;
      float v;
      std::memcpy(&v, &bits, 4);
      out[k * K0 + i] = v;
    }
  return CUMF_OK;
}

int cumf_debug_compat_system(cumf_handle* h, int side, const int32_t* rows, int64_t nrows, double* out) {
  cumf::clear_last_error();
  if (!h || !h->have_ratings || (side != 0 && side != 1) || nrows < 0 || (nrows > 0 && (!rows || !out)) ||
      h->f % 10 != 0) {
    cumf::set_last_error("cumf_debug_compat_system: bad arguments (f must be a multiple of 10)");
    return CUMF_ERR_INVALID;
  }
  const int64_t lim = side == 0 ? h->m : h->n;
  for (int64_t i = 0; i < nrows; ++i)
    if (rows[i] < 0 || rows[i] >= lim) {
      cumf::set_last_error("cumf_debug_compat_system: row out of range");
      return CUMF_ERR_INVALID;
    }
  if (nrows == 0) return CUMF_OK;
  DeviceGuard dg(h->device);
  if (cumf::prepare_compat(h->f) != hipSuccess) {
    (void)hipGetLastError();
    cumf::set_last_error("compat kernel: cannot raise its LDS limit");
    return CUMF_ERR_HIP;
  }
  DevBuf drows, dout;
  CUMF_TRY(drows.alloc(sizeof(int32_t) * nrows));
  CUMF_TRY(dout.alloc(sizeof(double) * nrows * h->f * (h->f + 1)));
  CUMF_HIP_TRY(hipMemcpy(drows.p, rows, drows.bytes, hipMemcpyHostToDevice));
  cumf::CompatArgs c{};
  c.row_ptr = side == 0 ? h->rp.as<int32_t>() : h->cp.as<int32_t>();
  c.col = side == 0 ? h->ci.as<int32_t>() : h->ri.as<int32_t>();
  c.val = side == 0 ? h->rv.as<float>() : h->cv.as<float>();
  c.theta = side == 0 ? h->T.as<float>() : h->X.as<float>();
  c.ldf = h->ldf;
  c.rows = drows.as<int32_t>();
  c.dump = dout.as<double>();
  c.fail_counter = h->fail.as<int>();
  c.f = h->f;
  c.lambda = h->lambda;
  CUMF_HIP_TRY(cumf::launch_compat(c, nrows, h->stream));
  CUMF_HIP_TRY(hipStreamSynchronize(h->stream));
  CUMF_HIP_TRY(hipMemcpy(out, dout.p, dout.bytes, hipMemcpyDeviceToHost));
  return CUMF_OK;
}

int cumf_debug_mfma_layout(int device, float* out) {
  cumf::clear_last_error();
  if (!out) return CUMF_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    cumf::set_last_error("no such device");
    return CUMF_ERR_NODEVICE;
  }
  DeviceGuard dg(device);
  DevBuf d;
  CUMF_TRY(d.alloc(sizeof(float) * 256));
  CUMF_HIP_TRY(cumf::launch_mfma_layout(d.as<float>(), nullptr));
  CUMF_HIP_TRY(hipDeviceSynchronize());
  CUMF_HIP_TRY(hipMemcpy(out, d.p, sizeof(float) * 256, hipMemcpyDeviceToHost));
  return CUMF_OK;
}

int cumf_debug_mfma_bf16_layout(int device, float* out) {
  cumf::clear_last_error();
  if (!out) return CUMF_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    cumf::set_last_error("no such device");
    return CUMF_ERR_NODEVICE;
  }
  DeviceGuard dg(device);
  DevBuf d;
  CUMF_TRY(d.alloc(sizeof(float) * 256));
  CUMF_HIP_TRY(cumf::launch_mfma_bf16_layout(d.as<float>(), nullptr));
  CUMF_HIP_TRY(hipDeviceSynchronize());
  CUMF_HIP_TRY(hipMemcpy(out, d.p, sizeof(float) * 256, hipMemcpyDeviceToHost));
  return CUMF_OK;
}

int cumf_debug_copy(int device, int64_t bytes, int iters, double* ms_per_launch) {
  cumf::clear_last_error();
  if (bytes <= 0 || iters <= 0 || !ms_per_launch) return CUMF_ERR_INVALID;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    cumf::set_last_error("no such device");
    return CUMF_ERR_NODEVICE;
  }
  DeviceGuard dg(device);
  bytes &= ~int64_t(15);
  DevBuf src, dst;
  CUMF_TRY(src.alloc(bytes));
  CUMF_TRY(dst.alloc(bytes));
  CUMF_HIP_TRY(hipMemset(src.p, 1, bytes));
  hipEvent_t a, b;
  CUMF_HIP_TRY(hipEventCreate(&a));
  CUMF_HIP_TRY(hipEventCreate(&b));
  CUMF_HIP_TRY(cumf::launch_copy(src.p, dst.p, bytes, nullptr));  // warm-up
  CUMF_HIP_TRY(hipEventRecord(a, nullptr));
  for (int i = 0; i < iters; ++i) CUMF_HIP_TRY(cumf::launch_copy(src.p, dst.p, bytes, nullptr));
  CUMF_HIP_TRY(hipEventRecord(b, nullptr));
  CUMF_HIP_TRY(hipEventSynchronize(b));
  float ms = 0.f;
  CUMF_HIP_TRY(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms_per_launch = ms / iters;
  return CUMF_OK;
}

}  // extern "C"

namespace {

// Per-device state the device-pointer half-sweep keeps between calls: the
// task plan of the last rating matrix, reused when the next call brings the
// same row pointers (Spark calls the drop-in once per half-sweep on the
// same blocks), and device buffers kept at the largest size asked for so
// far.  A call holds its device's cache for its whole duration; a
// concurrent call on the same device that finds it held works with buffers
// of its own.  The caches live until the process ends.
struct SweepCache {
  std::mutex mu;
  std::vector<int32_t> rp;  // row pointers the plan was built from
  int f = 0, split = 0;
  Plan plan;
  DevBuf theta, x, tpad, xpad, fail;
};

SweepCache* sweep_cache(int dev) {
  static std::mutex mu;
  static std::vector<SweepCache*>* caches = new std::vector<SweepCache*>();
  if (dev < 0) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  if ((size_t)dev >= caches->size()) caches->resize((size_t)dev + 1, nullptr);
  if (!(*caches)[dev]) (*caches)[dev] = new SweepCache();
  return (*caches)[dev];
}

// Holds the current device's cache if it is free.
struct CacheLease {
  SweepCache* c = nullptr;
  std::unique_lock<std::mutex> lk;
  CacheLease() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    SweepCache* sc = sweep_cache(dev);
    if (!sc) return;
    lk = std::unique_lock<std::mutex>(sc->mu, std::try_to_lock);
    if (lk.owns_lock()) c = sc;
  }
};

int ensure(DevBuf& b, size_t bytes) { return b.bytes >= bytes ? CUMF_OK : b.alloc(bytes); }

// Device-pointer half-sweep.  The CSR row pointers are read back to plan
// the tasks (or to recognise the cached plan); factor buffers with
// f % 4 != 0 are re-strided through padded temporaries (the kernels read
// factor rows with 16-byte loads).
int half_sweep_device(hipStream_t stream, const int32_t* d_row_ptr, const int32_t* d_col, const float* d_val,
                      const float* d_theta, float* d_x, int64_t m, int64_t n, int f, int64_t nnz, float lambda,
                      const cumf_options* opt, SweepCache* cache) {
  CUMF_TRY(check_f(f));
  if (m <= 0 || n <= 0 || nnz < 0 || nnz > INT32_MAX || m > INT32_MAX || !d_row_ptr || !d_theta ||
      !d_x || (nnz > 0 && (!d_col || !d_val)) || !(lambda >= 0.f)) {
    cumf::set_last_error("cumf_half_sweep_device: bad arguments");
    return CUMF_ERR_INVALID;
  }
  cumf_options o;
  cumf_default_options(&o);
  if (opt && opt->split > 0) o.split = opt->split;
  if (opt) o.compat_hermitian_bug = opt->compat_hermitian_bug;
  if (o.compat_hermitian_bug && f % 10 != 0) {
    cumf::set_last_error("compat_hermitian_bug reproduces the reference, which needs f % 10 == 0");
    return CUMF_ERR_INVALID;
  }
  if (o.compat_hermitian_bug && f > kMaxFusedF) {
    cumf::set_last_error("compat_hermitian_bug is built for f <= " + std::to_string(kMaxFusedF));
    return CUMF_ERR_UNSUPPORTED;
  }
  std::vector<int32_t> rp(m + 1);
  CUMF_HIP_TRY(hipMemcpyAsync(rp.data(), d_row_ptr, sizeof(int32_t) * (m + 1), hipMemcpyDeviceToHost, stream));
  CUMF_HIP_TRY(hipStreamSynchronize(stream));
  if (rp[0] != 0 || rp[m] != nnz) {
    cumf::set_last_error("row pointers must start at 0 and end at nnz");
    return CUMF_ERR_INVALID;
  }
  Plan local_plan;
  DevBuf local_tpad, local_xpad, local_fail;
  Plan* plan = &local_plan;
  DevBuf* tpad = &local_tpad;
  DevBuf* xpad = &local_xpad;
  DevBuf* fail = &local_fail;
  if (cache) {
    plan = &cache->plan;
    tpad = &cache->tpad;
    xpad = &cache->xpad;
    fail = &cache->fail;
    if (!(cache->f == f && cache->split == o.split && cache->rp == rp)) {
      cache->rp.clear();  // no stale match if the build fails
      CUMF_TRY(cache->plan.build(rp.data(), m, o.split, f));
      cache->rp.swap(rp);
      cache->f = f;
      cache->split = o.split;
    }
  } else {
    CUMF_TRY(local_plan.build(rp.data(), m, o.split, f));
  }
  const int64_t ldf = round_up4(f);
  const float* theta = d_theta;
  float* x = d_x;
  if (ldf != f) {
    CUMF_TRY(ensure(*tpad, sizeof(float) * n * ldf));
    CUMF_TRY(ensure(*xpad, sizeof(float) * m * ldf));
    CUMF_HIP_TRY(hipMemsetAsync(tpad->p, 0, sizeof(float) * n * ldf, stream));
    CUMF_HIP_TRY(hipMemcpy2DAsync(tpad->p, sizeof(float) * ldf, d_theta, sizeof(float) * f,
                                  sizeof(float) * f, n, hipMemcpyDeviceToDevice, stream));
    theta = tpad->as<float>();
    x = xpad->as<float>();
  }
  CUMF_TRY(ensure(*fail, sizeof(int)));
  CUMF_HIP_TRY(hipMemsetAsync(fail->p, 0, sizeof(int), stream));
  cumf::SweepArgs a{};
  a.row_ptr = d_row_ptr;
  a.col = d_col;
  a.val = d_val;
  a.theta = theta;
  a.x = x;
  a.ldf = ldf;
  a.tasks = plan->tasks.as<cumf::Task>();
  a.ntasks = plan->ntasks;
  a.heavy = plan->heavy.as<cumf::HeavyRow>();
  a.nheavy = plan->nheavy;
  a.partial = plan->partial.as<float>();
  a.fail_counter = fail->as<int>();
  a.f = f;
  a.lambda = lambda;
  if (o.compat_hermitian_bug) {
    cumf::CompatArgs c{};
    c.row_ptr = d_row_ptr;
    c.col = d_col;
    c.val = d_val;
    c.theta = theta;
    c.x = x;
    c.ldf = ldf;
    c.fail_counter = fail->as<int>();
    c.f = f;
    c.lambda = lambda;
    CUMF_HIP_TRY(cumf::prepare_compat(f));
    CUMF_HIP_TRY(cumf::launch_compat(c, m, stream));
  } else if (f > kMaxFusedF) {
    CUMF_HIP_TRY(cumf::prepare_large_kernel(f));
    int nlaunch = 0;
    CUMF_HIP_TRY(large_launches(a, *plan, &nlaunch, stream));
  } else {
    CUMF_HIP_TRY(task_launch(a, default_variant(f), stream));
    CUMF_HIP_TRY(finish_launch(a, default_variant(f), stream));
  }
  if (ldf != f)
    CUMF_HIP_TRY(hipMemcpy2DAsync(d_x, sizeof(float) * f, x, sizeof(float) * ldf, sizeof(float) * f, m,
                                  hipMemcpyDeviceToDevice, stream));
  CUMF_HIP_TRY(hipStreamSynchronize(stream));
  return CUMF_OK;
}

}  // namespace

extern "C" {

int cumf_half_sweep_device(hipStream_t stream, const int32_t* d_row_ptr, const int32_t* d_col,
                           const float* d_val, const float* d_theta, float* d_x, int64_t m,
                           int64_t n, int f, int64_t nnz, float lambda, const cumf_options* opt) {
  cumf::clear_last_error();
  CacheLease lease;
  return half_sweep_device(stream, d_row_ptr, d_col, d_val, d_theta, d_x, m, n, f, nnz, lambda, opt, lease.c);
}

// Drop-in for als/src/cuda/als.cu:1000-1115 (see include/als.h).
void doALSWithCSR(hipStream_t cuda_stream, int* csrRowIndex, int* csrColIndex, float* csrVal,
                  float* thetaTHost, float* XTHost, const int m, const int n, const int f,
                  const long nnz, const float lambda, const int X_BATCH) {
  cumf::clear_last_error();
  auto run = [&]() -> int {
    if (!csrRowIndex || !thetaTHost || !XTHost || m <= 0 || n <= 0 || nnz < 0 || X_BATCH < 1 ||
        (nnz > 0 && (!csrColIndex || !csrVal))) {
      cumf::set_last_error("doALSWithCSR: bad arguments");
      return CUMF_ERR_INVALID;
    }
    CUMF_TRY(check_f(f));
    // The factor buffers and the task plan come from the device's cache
    // when it is free (see SweepCache), else from allocations of this call.
    CacheLease lease;
    DevBuf local_theta, local_x;
    DevBuf& theta = lease.c ? lease.c->theta : local_theta;
    DevBuf& x = lease.c ? lease.c->x : local_x;
    const size_t tb = sizeof(float) * (size_t)n * f, xb = sizeof(float) * (size_t)m * f;
    CUMF_TRY(ensure(theta, tb));
    CUMF_TRY(ensure(x, xb));
    CUMF_HIP_TRY(hipMemcpyAsync(theta.p, thetaTHost, tb, hipMemcpyHostToDevice, cuda_stream));
    cumf_options o;
    cumf_default_options(&o);
    const char* env = std::getenv("CUMF_COMPAT_HERMITIAN_BUG");
    o.compat_hermitian_bug = (env && std::strcmp(env, "1") == 0) ? 1 : 0;
    CUMF_TRY(half_sweep_device(cuda_stream, csrRowIndex, csrColIndex, csrVal, theta.as<float>(), x.as<float>(),
                               m, n, f, nnz, lambda, &o, lease.c));
    CUMF_HIP_TRY(hipMemcpyAsync(XTHost, x.p, xb, hipMemcpyDeviceToHost, cuda_stream));
    CUMF_HIP_TRY(hipStreamSynchronize(cuda_stream));
    return CUMF_OK;
  };
  int rc = run();
  std::string err = rc == CUMF_OK ? std::string() : std::string(cumf_last_error());
  // The reference consumes the CSR arrays whatever happens to the sweep
  // (als/src/cuda/als.cu:1058, 1101-1102).
  (void)hipStreamSynchronize(cuda_stream);
  if (csrVal) (void)freeBest(csrVal);
  if (csrRowIndex) (void)freeBest(csrRowIndex);
  if (csrColIndex) (void)freeBest(csrColIndex);
  if (!err.empty()) cumf::set_last_error(err);
}

}  // extern "C"

// als_sweep.hip -- the half-sweep kernels proper: fused Gram + solve per
// task, the split-row finish and the Cholesky fallback of the CG path.
//
// Compiled four times from this one source:
//   * per register budget (the CUMF_TASK_WAVES wavefronts per SIMD the
//     compiler may assume when it allocates registers): 2 (namespace
//     cumf::w2) lets two solves share a SIMD; 1 (cumf::w1) gives a wavefront
//     the whole register file, which the larger tile grids (f > 124) need
//     to keep the Gram out of scratch memory.  The tile table picks one per
//     degree bucket;
//   * per solver family (CUMF_SWEEP_SOLVER): 0 (namespace ...::chol) holds
//     the Cholesky task and finish kernels and the CG path's Cholesky
//     fallback, 1 (...::cg) the CG task and finish kernels.  The CG object
//     is built with the scheduler's register-pressure trackers, which keep
//     its FP64 residual out of scratch memory; the Cholesky kernels run
//     faster without them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "als_device.h"

#ifndef CUMF_TASK_WAVES
#define CUMF_TASK_WAVES 2
#endif
#if CUMF_TASK_WAVES == 1
#define CUMF_VNS w1
#elif CUMF_TASK_WAVES == 2
#define CUMF_VNS w2
#else
#error "CUMF_TASK_WAVES must be 1 or 2"
#endif
#ifndef CUMF_SWEEP_SOLVER
#define CUMF_SWEEP_SOLVER 0
#endif
#if CUMF_SWEEP_SOLVER == 0
#define CUMF_SNS chol
#elif CUMF_SWEEP_SOLVER == 1
#define CUMF_SNS cg
#else
#error "CUMF_SWEEP_SOLVER must be 0 (Cholesky) or 1 (CG)"
#endif

namespace cumf {
namespace CUMF_VNS {
namespace CUMF_SNS {

// One task: Gram of its ratings, then either the partial Gram of a split
// row goes to scratch, or the row is solved.
template <int NB, bool BF, bool CG>
__device__ __forceinline__ void run_task(const SweepArgs& a, const Task& t, float* lds, int lane) {
  float* xrow = a.x + (int64_t)t.row * a.ldf;
  __bf16* xb = BF ? reinterpret_cast<__bf16*>(a.x_b) + (int64_t)t.row * a.ldb : nullptr;
  if (t.count <= 0) {
    if (t.slot >= 0) {  // (the planner never makes an empty chunk of a heavy row)
      float* dst = a.partial + (int64_t)t.slot * num_tiles(NB) * 256;
      for (int i = lane; i < num_tiles(NB) * 256; i += 64) dst[i] = 0.0f;
      return;
    }
    // no ratings: x_u = 0
    for (int i = lane; i < a.ldf; i += 64) xrow[i] = 0.0f;
    if (xb)
      for (int i = lane; i < a.ldb; i += 64) xb[i] = (__bf16)0.0f;
    return;
  }

  // The row's current factor (the CG start) is fetched before the Gram so
  // that its latency hides behind the gathers.
  float xpre[rows_per_lane(NB)];
  if constexpr (CG) {
#pragma unroll
    for (int q = 0; q < rows_per_lane(NB); ++q) {
      const int i = lane + 64 * q;
      xpre[q] = (t.slot < 0 && i < a.f) ? xrow[i] : 0.0f;
    }
  }
  f4 acc[num_tiles(NB)];  // set by the Gram
  gram_any<NB, BF>(acc, a, (int64_t)t.start, t.count, lds, lane);

  if (t.slot >= 0) {  // partial Gram of a heavy row
    float* dst = a.partial + (int64_t)t.slot * num_tiles(NB) * 256;
#pragma unroll
    for (int i = 0; i < num_tiles(NB); ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(i * 4 + r) * 64 + lane] = acc[i][r];
    return;
  }
  add_ridge<NB>(acc, a.f, a.lambda * (float)t.count, lane);
  lds_order();
  solve_row<NB, CG>(acc, a, BF, lds, t.row, xb, lane, t.row, CG ? xpre : nullptr);
}

// tasks[0 .. ntasks): one work-group per task, or (PERSISTENT) a grid of
// wavefronts that take tasks from a counter until none are left (with no
// task list, PERSISTENT takes the rows 0 .. ntasks themselves).  Separate
// instantiations: the loop changes register allocation (more spill), so the
// one-task-per-group kernel keeps its own.
template <int NB, bool BF, bool CG, bool PERSISTENT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CUMF_TASK_WAVES))) void als_task_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  if constexpr (!PERSISTENT) {
    run_task<NB, BF, CG>(a, a.tasks[blockIdx.x], lds, lane);
  } else {
    for (;;) {
      int idx = 0;
      if (lane == 0) idx = atomicAdd(a.task_counter, 1);
      idx = __builtin_amdgcn_readfirstlane(idx);
      if (idx >= a.ntasks) break;
      lds_order();
      // Without a task list the tasks are the rows themselves, whole.
      const Task t = a.tasks ? a.tasks[idx] : Task{idx, a.row_ptr[idx], a.row_ptr[idx + 1] - a.row_ptr[idx], -1};
      run_task<NB, BF, CG>(a, t, lds, lane);
    }
  }
}

template <int NB, bool CG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CUMF_TASK_WAVES))) void als_finish_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  const HeavyRow h = a.heavy[blockIdx.x];
  f4 acc[num_tiles(NB)];
#pragma unroll
  for (int i = 0; i < num_tiles(NB); ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
  for (int p = 0; p < h.nchunks; ++p) {
    const float* src = a.partial + (int64_t)(h.slot + p) * num_tiles(NB) * 256;
#pragma unroll
    for (int i = 0; i < num_tiles(NB); ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][r] += src[(i * 4 + r) * 64 + lane];
  }
  const int n_u = a.row_ptr[h.row + 1] - a.row_ptr[h.row];
  add_ridge<NB>(acc, a.f, a.lambda * (float)n_u, lane);
  const bool bf = a.precision != 0;
  __bf16* xb = bf ? reinterpret_cast<__bf16*>(a.x_b) + (int64_t)h.row * a.ldb : nullptr;
  lds_order();
  solve_row<NB, CG>(acc, a, bf, lds, h.row, xb, lane, -1 - (int)blockIdx.x);
}

// Cholesky solve of the rows the CG pass did not accept: a light row's Gram
// is formed again from its ratings, a split row's from its partial Grams
// (still in the scratch buffer).  A fixed grid walks the list, whose length
// the CG pass left in fb_count.
template <int NB, bool BF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CUMF_TASK_WAVES))) void als_fallback_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  const int total = *a.fb_count;
  for (int k = blockIdx.x; k < total; k += gridDim.x) {
    const int e = a.fb_list[k];
    f4 acc[num_tiles(NB)];
#pragma unroll
    for (int i = 0; i < num_tiles(NB); ++i) acc[i] = f4{0.f, 0.f, 0.f, 0.f};
    int64_t row;
    if (e >= 0) {
      row = e;
      const int64_t s = a.row_ptr[row];
      const int cnt = a.row_ptr[row + 1] - a.row_ptr[row];
      lds_order();
      if (cnt > 0) gram_any<NB, BF>(acc, a, s, cnt, lds, lane);
    } else {
      const HeavyRow h = a.heavy[-1 - e];
      row = h.row;
      for (int p = 0; p < h.nchunks; ++p) {
        const float* src = a.partial + (int64_t)(h.slot + p) * num_tiles(NB) * 256;
#pragma unroll
        for (int i = 0; i < num_tiles(NB); ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][r] += src[(i * 4 + r) * 64 + lane];
      }
    }
    const int n_u = a.row_ptr[row + 1] - a.row_ptr[row];
    add_ridge<NB>(acc, a.f, a.lambda * (float)n_u, lane);
    __bf16* xb = BF ? reinterpret_cast<__bf16*>(a.x_b) + row * a.ldb : nullptr;
    lds_order();
    solve_row<NB, false>(acc, a, BF, lds, row, xb, lane, 0);
  }
}

#if CUMF_SWEEP_SOLVER == 0
// Dual form of the tasks of one launch, each of at most kDualMaxRatings
// ratings (BF16 factors), the rows staged in LDS.  The host launches it
// (with two or three chunks) only where a task is a split row's chunk,
// which a small split size can make that short: such a task still forms
// its partial Gram.
template <int NB, int CH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CUMF_TASK_WAVES))) void als_dual_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  const Task t = a.tasks[blockIdx.x];
  if (t.slot >= 0) {
    f4 acc[num_tiles(NB)];
    gram_any<NB, true>(acc, a, (int64_t)t.start, t.count, lds, lane);
    float* dst = a.partial + (int64_t)t.slot * num_tiles(NB) * 256;
#pragma unroll
    for (int i = 0; i < num_tiles(NB); ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) dst[(i * 4 + r) * 64 + lane] = acc[i][r];
    return;
  }
  dual_solve_row<NB, CH>(a, t, lds, lane);
}
// The dual form of whole rows of at most 32 CH - 1 ratings with the MFMA
// operands read straight from the factor rows (no staging, no partial-Gram
// branch: fewer registers and about 10 KB of LDS per wavefront).  The host
// launches it where no task is a split row's chunk.
template <int NB, int CH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(CUMF_TASK_WAVES))) void als_dual_row_kernel(SweepArgs a) {
  extern __shared__ float lds[];
  dual_solve_row_g<NB, CH>(a, a.tasks[blockIdx.x], lds, threadIdx.x);
}
#endif

constexpr int kFallbackGrid = 2048;  // work-groups walking the fallback list

template <int NB>
static hipError_t launch_nb(const SweepArgs& a, hipStream_t s, int which) {
  constexpr bool CG = CUMF_SWEEP_SOLVER == 1;
  const bool bf = a.precision != 0;
  size_t shmem = sizeof(float) * (size_t)cmax(lds_floats_nb(NB, bf), CG ? cg_lds_floats(NB, a.f) : 0);
  if (which == 0 && a.persistent == 0 && a.occupancy > 0) {
    if (a.occupancy < kMinOccupancy || a.occupancy > kMaxOccupancy) return hipErrorInvalidValue;
    if (occupancy_lds_bytes(a.occupancy) > shmem) shmem = occupancy_lds_bytes(a.occupancy);
  }
  if (which == 0) {
    if (a.ntasks > 0) {
      if (a.persistent > 0) {
        const dim3 grid((unsigned)(a.ntasks < a.persistent ? a.ntasks : a.persistent));
        if (bf) hipLaunchKernelGGL((als_task_kernel<NB, true, CG, true>), grid, dim3(64), shmem, s, a);
        else hipLaunchKernelGGL((als_task_kernel<NB, false, CG, true>), grid, dim3(64), shmem, s, a);
      } else {
        const dim3 grid((unsigned)a.ntasks);
        if (bf) hipLaunchKernelGGL((als_task_kernel<NB, true, CG, false>), grid, dim3(64), shmem, s, a);
        else hipLaunchKernelGGL((als_task_kernel<NB, false, CG, false>), grid, dim3(64), shmem, s, a);
      }
    }
  } else if (which == 1) {
    if (a.nheavy > 0)
      hipLaunchKernelGGL((als_finish_kernel<NB, CG>), dim3((unsigned)a.nheavy), dim3(64), shmem, s, a);
  } else if constexpr (!CG) {
    if (which == 3) {
      // the CG path's Cholesky fallback (lives with the Cholesky kernels)
      const size_t shc = sizeof(float) * (size_t)lds_floats_nb(NB, bf);
      if (bf) hipLaunchKernelGGL((als_fallback_kernel<NB, true>), dim3(kFallbackGrid), dim3(64), shc, s, a);
      else hipLaunchKernelGGL((als_fallback_kernel<NB, false>), dim3(kFallbackGrid), dim3(64), shc, s, a);
    } else {
#if CUMF_SWEEP_SOLVER == 0
      // which == 4: the dual form (BF16 factors only)
      if (!bf) return hipErrorInvalidValue;
      if (a.ntasks > 0) {
        const dim3 grid((unsigned)a.ntasks), block(64);
        const size_t gram = sizeof(float) * (size_t)lds_floats_nb(NB, true);  // the partial-Gram branch
        if (!a.dual_partial && a.dual_chunks >= 1 && a.dual_chunks <= 3) {
          if (a.dual_chunks == 1)
            hipLaunchKernelGGL((als_dual_row_kernel<NB, 1>), grid, block, sizeof(float) * dual_g_lds_floats<1>(), s, a);
          else if (a.dual_chunks == 2)
            hipLaunchKernelGGL((als_dual_row_kernel<NB, 2>), grid, block, sizeof(float) * dual_g_lds_floats<2>(), s, a);
          else
            hipLaunchKernelGGL((als_dual_row_kernel<NB, 3>), grid, block, sizeof(float) * dual_g_lds_floats<3>(), s, a);
        } else if (a.dual_chunks == 2) {
          const size_t shd = sizeof(float) * (size_t)dual_lds_floats<NB, 2>();
          hipLaunchKernelGGL((als_dual_kernel<NB, 2>), grid, block, shd > gram ? shd : gram, s, a);
        } else if (a.dual_chunks == 3) {
          const size_t shd = sizeof(float) * (size_t)dual_lds_floats<NB, 3>();
          hipLaunchKernelGGL((als_dual_kernel<NB, 3>), grid, block, shd > gram ? shd : gram, s, a);
        } else {
          return hipErrorInvalidValue;
        }
      }
#endif
    }
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static hipError_t dispatch(const SweepArgs& a, hipStream_t s, int which) {
  switch (sweep_nb(a.f)) {
    case 1: return launch_nb<1>(a, s, which);
    case 2: return launch_nb<2>(a, s, which);
    case 3: return launch_nb<3>(a, s, which);
    case 4: return launch_nb<4>(a, s, which);
    case 5: return launch_nb<5>(a, s, which);
    case 6: return launch_nb<6>(a, s, which);
    case 7: return launch_nb<7>(a, s, which);
    case 8: return launch_nb<8>(a, s, which);
    case 9: return launch_nb<9>(a, s, which);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_task_kernel(const SweepArgs& a, hipStream_t s) { return dispatch(a, s, 0); }
hipError_t launch_finish_kernel(const SweepArgs& a, hipStream_t s) { return dispatch(a, s, 1); }
#if CUMF_SWEEP_SOLVER == 0
hipError_t launch_fallback_kernel(const SweepArgs& a, hipStream_t s) { return dispatch(a, s, 3); }
hipError_t launch_dual_kernel(const SweepArgs& a, hipStream_t s) { return dispatch(a, s, 4); }
#endif

}  // namespace CUMF_SNS
}  // namespace CUMF_VNS
}  // namespace cumf

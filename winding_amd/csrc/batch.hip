// The batch kernel: many (cloud, view) sorts in one persistent launch.
//
// Work is a single ticketed queue of items in four runs:
//   [projection chunks of every job][scan items][low-byte tiles][high-byte tiles]
// A workgroup takes the next ticket, does that item, and takes another, until
// the queue is empty. An item only ever waits on items with smaller tickets:
// a job's scan items wait for its projection chunks, its low-byte tiles for
// its scan (the digit offsets), its high-byte tiles for all its low-byte tiles,
// and a high-byte look-back on earlier tiles of the same job. Smaller tickets were taken by workgroups
// already running, so every wait ends, whatever the grid size. Jobs of 10k and
// 2M splats share the machine tile by tile instead of launch by launch.
//
// Nothing is reset from the host between runs: the last workgroup to leave
// zeroes the tickets and flags, and the run counter tags this run's look-back
// words (in their own range, above the single-view sorts' epochs), so the same
// launch can be replayed from a hipGraph.
#include "splat_device.h"

namespace wd {
namespace {

constexpr uint32_t kBatchEpochBase = 1u << 29;

__device__ __forceinline__ uint32_t find_job(const uint32_t *__restrict__ starts, uint32_t nj, uint32_t item) {
  // largest j with starts[j] <= item (starts[0] = 0, nondecreasing)
  uint32_t lo = 0, hi = nj;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (starts[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// Bounded: a wait that runs out flags the batch as failed and goes on, so a
// bug shows as an error from the host, never as a hung GPU.
__device__ __forceinline__ void wait_value(const uint32_t *p, uint32_t want, uint32_t *stuck) {
  for (uint32_t spins = 0; __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != want; ++spins) {
    if (spins == kSpinLimit || ((spins & 1023u) == 1023u && __hip_atomic_load(stuck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicOr(stuck, 1u);
      return;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

__global__ __launch_bounds__(256) void k_batch(const BatchJob *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                               uint32_t nj, uint32_t max_jobs, BatchCtl *__restrict__ bc) {
  __shared__ union {
    ProjectSmem p;
    ScatterSmem s;
  } sm;
  __shared__ uint32_t s_item, s_runs, s_w[4], s_last;
  uint32_t *proj_done = reinterpret_cast<uint32_t *>(bc + 1);
  uint32_t *ready = proj_done + max_jobs;
  uint32_t *lo_done = ready + max_jobs;
  // starts: tile starts (chunks = low tiles = high tiles) then scan-item starts, nj + 1 each
  const uint32_t *t_start = starts, *s_start = starts + (nj + 1);
  const uint32_t t_total = t_start[nj], s_total = s_start[nj];
  const uint32_t e_proj = t_total, e_scan = e_proj + s_total, e_lo = e_scan + t_total, total = e_lo + t_total;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) s_runs = __hip_atomic_load(&bc->runs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const uint32_t epoch = kBatchEpochBase | ((s_runs + 1u) & (kBatchEpochBase - 1u));
  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(&bc->ticket, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= total) break;
    if (item < e_proj) {
      const uint32_t j = find_job(t_start, nj, item);
      const BatchJob &J = jobs[j];
      if (J.std_form)
        project_chunk<true>(sm.p, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov, J.tilehist,
                            &J.ctl->recomputed_acc, item - t_start[j]);
      else
        project_chunk<false>(sm.p, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov, J.tilehist,
                             &J.ctl->recomputed_acc, item - t_start[j]);
      __threadfence();
      __syncthreads();
      if (tid == 0) atomicAdd(&proj_done[j], 1u);
    } else if (item < e_scan) {
      const uint32_t it = item - e_proj;
      const uint32_t j = find_job(s_start, nj, it);
      const BatchJob &J = jobs[j];
      if (tid == 0) wait_value(&proj_done[j], J.tiles, &bc->stuck);
      __syncthreads();
      if (scan_digits<kBatchScanDigits>(it - s_start[j], 256 / kBatchScanDigits, J.tiles, J.tilehist, J.prefix,
                                        J.ctl, s_w, &s_last) && tid == 0)
        __hip_atomic_store(&ready[j], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (item < e_lo) {
      const uint32_t it = item - e_scan;
      const uint32_t j = find_job(t_start, nj, it);
      const BatchJob &J = jobs[j];
      if (tid == 0) wait_value(&ready[j], 1u, &bc->stuck);
      __syncthreads();
      scatter_tile<false>(sm.s, it - t_start[j], J.keys, J.mask, nullptr, J.n, J.tiles, J.prefix, J.ctl, nullptr,
                          epoch, J.packed, &bc->stuck);
      __threadfence();
      __syncthreads();
      if (tid == 0) atomicAdd(&lo_done[j], 1u);
    } else {
      const uint32_t it = item - e_lo;
      const uint32_t j = find_job(t_start, nj, it);
      const BatchJob &J = jobs[j];
      if (tid == 0) wait_value(&lo_done[j], J.tiles, &bc->stuck);
      __syncthreads();
      scatter_tile<true>(sm.s, it - t_start[j], nullptr, nullptr, J.packed, J.n, J.tiles, nullptr, J.ctl, J.status_hi,
                         epoch, J.order, &bc->stuck);
    }
  }
  // The last workgroup out resets the queue for the next run.
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&bc->exited, 1u) == gridDim.x - 1) {
      __threadfence();
      for (uint32_t j = 0; j < nj; ++j) {
        proj_done[j] = 0;
        ready[j] = 0;
        lo_done[j] = 0;
      }
      bc->ticket = 0;
      bc->exited = 0;
      __hip_atomic_store(&bc->runs, s_runs + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace

int batch_grid(wd_context *ctx, uint32_t items, uint32_t *grid) {
  int per_cu = 0;
  WD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, 256, 0), "batch: occupancy");
  const uint32_t resident = (uint32_t)std::max(per_cu, 1) * (uint32_t)ctx->cus;
  *grid = std::max<uint32_t>(1, std::min(items, resident));
  return WD_OK;
}

int launch_batch(wd_context *ctx, wd_batch *b) {
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKBatch, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(k_batch, dim3(b->grid), dim3(256), 0, ctx->stream, b->jobs, b->starts, b->n, b->max_jobs,
                     b->ctl);
  WD_HIP(hipGetLastError(), "batch: launch");
  return ctx->prof_end(kKBatch, ev);
}

}  // namespace wd

// The batch kernel: many (cloud, view) sorts in one persistent launch.
//
// Work is a single ticketed queue of items in six runs, every job in each:
//   [projection chunks][scans][low-byte super-tiles][high-byte histograms]
//   [second scans][high-byte super-tiles]
// A workgroup takes the next ticket, does that item, and takes another, until
// the queue is empty. An item only ever waits on items with smaller tickets --
// a job's scan on its chunks, its low-byte pass on its scan, its histograms on
// its low-byte pass, its second scan on its histograms, its high-byte pass on
// its second scan -- and smaller tickets were taken by workgroups already
// running, so every wait ends, whatever the grid size. Jobs of 10k and 2M
// splats share the machine tile by tile instead of launch by launch.
//
// Nothing is reset from the host between runs: the last workgroup to leave
// zeroes the tickets and the per-job counters, so the same launch can be
// replayed from a hipGraph.
#include "splat_device.h"

namespace wd {
namespace {

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
    if (spins == kSpinLimit ||
        ((spins & 1023u) == 1023u && __hip_atomic_load(stuck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicOr(stuck, 1u);
      return;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

__device__ __forceinline__ void wait_for(const uint32_t *p, uint32_t want, uint32_t *stuck) {
  if (threadIdx.x == 0) wait_value(p, want, stuck);
  __syncthreads();
}

// This workgroup's item is done: its writes are visible before the count is.
__device__ __forceinline__ void count_done(uint32_t *p) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(p, 1u);
}

__global__ __launch_bounds__(256) void k_batch(const BatchJob *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                               uint32_t nj, uint32_t max_jobs, BatchCtl *__restrict__ bc) {
  __shared__ union {
    ProjectSmem p;
    HistSmem h;
    ScatterSmem<kSuper> s;
  } sm;
  __shared__ uint32_t s_item, s_w[4], s_last;
  uint32_t *proj_done = reinterpret_cast<uint32_t *>(bc + 1);
  uint32_t *scan_done = proj_done + max_jobs;
  uint32_t *lo_done = scan_done + max_jobs;
  uint32_t *hist_done = lo_done + max_jobs;
  uint32_t *scan2_done = hist_done + max_jobs;
  // starts: per job item starts of tiles, scan blocks and super-tiles, nj + 1 each
  const uint32_t *t_start = starts, *b_start = starts + (nj + 1), *u_start = starts + 2 * (nj + 1);
  const uint32_t nt = t_start[nj], nb = b_start[nj], nu = u_start[nj];
  const uint32_t e0 = nt, e1 = e0 + nb, e2 = e1 + nu, e3 = e2 + nt, e4 = e3 + nb, total = e4 + nu;
  const uint32_t tid = threadIdx.x;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(&bc->ticket, 1u);
    __syncthreads();
    const uint32_t item = s_item;
    if (item >= total) break;
    if (item < e0) {
      const uint32_t j = find_job(t_start, nj, item);
      const BatchJob &J = jobs[j];
      if (J.std_form)
        project_chunk<true>(sm.p, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov, J.tilehist,
                            &J.ctl->recomputed_acc, item - t_start[j]);
      else
        project_chunk<false>(sm.p, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov, J.tilehist,
                             &J.ctl->recomputed_acc, item - t_start[j]);
      count_done(&proj_done[j]);
    } else if (item < e1) {
      const uint32_t it = item - e0, j = find_job(b_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&proj_done[j], J.tiles, &bc->stuck);
      if (scan_after_project(it - b_start[j], J.blocks, J.tiles, J.tilehist, J.prefix_lo, J.bsum_lo, J.bsum_hi, J.ctl,
                             s_w, &s_last) && tid == 0)
        __hip_atomic_store(&scan_done[j], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (item < e2) {
      const uint32_t it = item - e1, j = find_job(u_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&scan_done[j], 1u, &bc->stuck);
      scatter_super<false, kSuper>(sm.s, it - u_start[j], J.keys, J.mask, nullptr, J.n, J.tilehist, 512, 0,
                                   J.prefix_lo, J.bsum_lo, J.ctl, J.packed);
      count_done(&lo_done[j]);
    } else if (item < e3) {
      const uint32_t it = item - e2, j = find_job(t_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&lo_done[j], u_start[j + 1] - u_start[j], &bc->stuck);
      hist_tile(sm.h, it - t_start[j], J.packed, J.ctl, J.tilehist3);
      count_done(&hist_done[j]);
    } else if (item < e4) {
      const uint32_t it = item - e3, j = find_job(b_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&hist_done[j], J.tiles, &bc->stuck);
      if (scan_after_hist(it - b_start[j], J.blocks, J.tilehist3, J.prefix_hi, J.bsum3, J.ctl, &s_last) && tid == 0)
        __hip_atomic_store(&scan2_done[j], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const uint32_t it = item - e4, j = find_job(u_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&scan2_done[j], 1u, &bc->stuck);
      scatter_super<true, kSuper>(sm.s, it - u_start[j], nullptr, nullptr, J.packed, J.n, J.tilehist3, 256, 0,
                                  J.prefix_hi, J.bsum3, J.ctl, J.order);
    }
  }
  // The last workgroup out resets the queue for the next run.
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&bc->exited, 1u) == gridDim.x - 1) {
      __threadfence();
      for (uint32_t c = 0; c < (uint32_t)kBatchCounters; ++c)
        for (uint32_t j = 0; j < nj; ++j) proj_done[c * max_jobs + j] = 0;
      bc->ticket = 0;
      bc->exited = 0;
      bc->runs = bc->runs + 1u;
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

// Batches: many (cloud, view) sorts in one graph launch, two schedules.
//
// Persistent (one kernel): a single ticketed queue of items in six runs,
// every job in each:
//   [projection chunks][scan column groups][low-byte super-tiles]
//   [high-byte histograms][second-scan column groups][high-byte super-tiles]
// An item is a run of a few tiles of one job (a scan item: a group of
// columns). A workgroup takes the next ticket, does that item, and takes
// another, until the queue is empty. An item only ever waits on items with smaller
// tickets -- a job's scan on its chunks, its low-byte pass on its scan, its
// histograms on its low-byte pass, its second scan on its histograms, its
// high-byte pass on its second scan -- and smaller tickets were taken by
// workgroups already running, so every wait ends, whatever the grid size.
// Jobs of 10k and 2M splats share the machine run by run instead of launch
// by launch. The price is a release after every item and an acquire before
// every wait: on a chip whose L2 is one per XCD, a write-back and an
// invalidation of this XCD's L2. Runs of tiles and one wave's fence an item
// bring that down, but it still costs more than the phased schedule's five
// launch boundaries.
//
// Phased (six kernels): each of the six runs is one launch over every job's
// items; the launch boundaries are the only barriers, so no fence runs.
//
// Nothing is reset from the host between runs: the persistent kernel's last
// workgroup out zeroes the tickets and the per-job counters, so either
// schedule replays from a hipGraph.
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
// bug shows as an error from the host, never as a hung GPU. Polls relaxed;
// one acquire fence once the count is in.
__device__ __forceinline__ void wait_value(const uint32_t *p, uint32_t want, uint32_t *stuck) {
  for (uint32_t spins = 0; __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want; ++spins) {
    if (spins == kSpinLimit ||
        ((spins & 1023u) == 1023u && __hip_atomic_load(stuck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      atomicOr(stuck, 1u);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ __forceinline__ void wait_for(const uint32_t *p, uint32_t want, uint32_t *stuck) {
  if (threadIdx.x == 0) wait_value(p, want, stuck);
  __syncthreads();
}

// This workgroup's item is done: every thread's writes are released before
// the count is. The barrier orders the workgroup's writes before thread 0's
// fence, and a release is cumulative, so one wave's fence (one write-back of
// this XCD's L2) releases them all.
__device__ __forceinline__ void count_done(uint32_t *p) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Item bodies shared by both schedules.
// The projection's form by the job: .splat records or centres and Sigma, the
// harmonics' degree, and whether the state takes the draw's ellipses (the host
// takes the last three only with the reverse-Z perspective, so the general
// projection has one form here).
template <bool SIG, int SHD, bool ELL>
__device__ __forceinline__ void project_form(ProjectSmem &sm, const BatchJob &J, uint32_t t) {
  float4 *ell = (float4 *)J.ellipse;
  project_chunk<true, false, SHD, ELL, SIG, false>(sm, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov,
                                            J.tilehist, &J.ctl->recomputed_acc, t, nullptr, (const uint32_t *)J.sh,
                                            ell, ELL ? reinterpret_cast<const float *>(ell + 2 * J.capacity) : nullptr,
                                            (const uint2 *)J.sig, J.capacity);
}

template <bool SIG, bool ELL>
__device__ __forceinline__ void project_degree(ProjectSmem &sm, const BatchJob &J, uint32_t t) {
  switch (J.degree) {
    case 1: project_form<SIG, 1, ELL>(sm, J, t); break;
    case 2: project_form<SIG, 2, ELL>(sm, J, t); break;
    case 3: project_form<SIG, 3, ELL>(sm, J, t); break;
    default: project_form<SIG, 0, ELL>(sm, J, t); break;
  }
}

__device__ __forceinline__ void item_project(ProjectSmem &sm, const BatchJob &J, uint32_t t) {
  if (J.std_form) {
    if (J.ellipse) {
      if (J.sig)
        project_degree<true, true>(sm, J, t);
      else
        project_degree<false, true>(sm, J, t);
    } else if (J.sig) {
      project_degree<true, false>(sm, J, t);
    } else {
      project_degree<false, false>(sm, J, t);
    }
  } else {
    project_chunk<false>(sm, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov, J.tilehist,
                         &J.ctl->recomputed_acc, t);
  }
}

__device__ __forceinline__ void item_scan(const BatchJob &J, uint32_t cg, uint32_t (*s_w)[kBatchScanCols]) {
  scan_cols<kBatchScanCols>(cg * kBatchScanCols, J.tiles, J.tilehist, 512, J.prefix_lo, 256, 256, J.ctl->total, s_w);
  if (cg == 0 && threadIdx.x == 0) {
    J.ctl->recomputed = J.ctl->recomputed_acc;
    J.ctl->recomputed_acc = 0;
  }
}

__device__ __forceinline__ void item_scan2(const BatchJob &J, uint32_t cg, uint32_t (*s_w)[kBatchScanCols]) {
  const uint32_t T3 = (J.ctl->count + kTile - 1) / kTile;
  scan_cols<kBatchScanCols>(cg * kBatchScanCols, T3, J.tilehist3, 256, J.prefix_hi, 256, 256, J.ctl->total3, s_w);
}

constexpr uint32_t kScanItems = 512 / kBatchScanCols, kScan2Items = 256 / kBatchScanCols;
static_assert(kScanItems % kXcds == 0 && kScan2Items % kXcds == 0, "scan items a multiple of the XCDs");

// gate: when given, the whole launch is skipped unless *gate is nonzero (the
// learned sort's fallback: it runs, for every job, when any view failed its
// check).
// run: tiles (or super-tiles) an item takes, one after the other, so that a
// run of them pays one release fence and one count instead of one each.
__global__ __launch_bounds__(256) void k_batch(const BatchJob *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                               uint32_t nj, uint32_t max_jobs, BatchCtl *__restrict__ bc,
                                               const uint32_t *__restrict__ gate, uint32_t run) {
  if (gate && __hip_atomic_load(gate, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
  __shared__ union {
    ProjectSmem p;
    HistSmem h;
    ScatterSmem<kSuper> s;
    uint32_t part[4][kBatchScanCols];
  } sm;
  __shared__ uint32_t s_item;
  uint32_t *proj_done = reinterpret_cast<uint32_t *>(bc + 1);
  uint32_t *scan_done = proj_done + max_jobs;
  uint32_t *lo_done = scan_done + max_jobs;
  uint32_t *hist_done = lo_done + max_jobs;
  uint32_t *scan2_done = hist_done + max_jobs;
  // starts: per job item starts of runs of tiles and of super-tiles, nj + 1 each
  const uint32_t *t_start = starts, *u_start = starts + (nj + 1);
  const uint32_t nt = t_start[nj], nu = u_start[nj];
  const uint32_t e0 = nt, e1 = e0 + kScanItems * nj, e2 = e1 + nu, e3 = e2 + nt, e4 = e3 + kScan2Items * nj,
                 total = e4 + nu;
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
      const uint32_t t0 = (item - t_start[j]) * run, t1 = min(t0 + run, J.tiles);
      for (uint32_t t = t0; t < t1; ++t) {
        if (t != t0) __syncthreads();  // the tile before is done with the shared memory
        item_project(sm.p, J, t);
      }
      count_done(&proj_done[j]);
    } else if (item < e1) {
      const uint32_t j = (item - e0) / kScanItems;
      const BatchJob &J = jobs[j];
      wait_for(&proj_done[j], t_start[j + 1] - t_start[j], &bc->stuck);
      item_scan(J, (item - e0) % kScanItems, sm.part);
      count_done(&scan_done[j]);
    } else if (item < e2) {
      const uint32_t it = item - e1, j = find_job(u_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&scan_done[j], kScanItems, &bc->stuck);
      const uint32_t supers = (J.tiles + kSuper - 1) / kSuper;
      const uint32_t s0 = (it - u_start[j]) * run, s1 = min(s0 + run, supers);
      for (uint32_t st = s0; st < s1; ++st) {
        if (st != s0) __syncthreads();
        scatter_super<false, kSuper>(sm.s, st, J.keys, J.mask, nullptr, J.n, J.tilehist, 512, J.prefix_lo,
                                     J.ctl->total, J.ctl, J.packed);
      }
      count_done(&lo_done[j]);
    } else if (item < e3) {
      const uint32_t it = item - e2, j = find_job(t_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&lo_done[j], u_start[j + 1] - u_start[j], &bc->stuck);
      const uint32_t t0 = (it - t_start[j]) * run, t1 = min(t0 + run, J.tiles);
      for (uint32_t t = t0; t < t1; ++t) {
        if (t != t0) __syncthreads();
        hist_tile(sm.h, t, J.packed, J.ctl, J.tilehist3);
      }
      count_done(&hist_done[j]);
    } else if (item < e4) {
      const uint32_t j = (item - e3) / kScan2Items;
      const BatchJob &J = jobs[j];
      wait_for(&hist_done[j], t_start[j + 1] - t_start[j], &bc->stuck);
      item_scan2(J, (item - e3) % kScan2Items, sm.part);
      count_done(&scan2_done[j]);
    } else {
      const uint32_t it = item - e4, j = find_job(u_start, nj, it);
      const BatchJob &J = jobs[j];
      wait_for(&scan2_done[j], kScan2Items, &bc->stuck);
      const uint32_t supers = (J.tiles + kSuper - 1) / kSuper;
      const uint32_t s0 = (it - u_start[j]) * run, s1 = min(s0 + run, supers);
      for (uint32_t st = s0; st < s1; ++st) {
        if (st != s0) __syncthreads();
        scatter_super<true, kSuper>(sm.s, st, nullptr, nullptr, J.packed, J.n, J.tilehist3, 256, J.prefix_hi,
                                    J.ctl->total + 256, J.ctl, J.order);
      }
    }
  }
  // The last workgroup out resets the queue for the next run. A run whose wait
  // ran out marks every job's sort as not valid (the state's stats then fail on
  // the host), keeps the fact for wd_batch_stuck, and clears the live flag, so
  // the next run waits in full again.
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&bc->exited, 1u) == gridDim.x - 1) {
      __threadfence();
      if (__hip_atomic_load(&bc->stuck, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        for (uint32_t j = 0; j < nj; ++j) jobs[j].ctl->stuck = 1u;
        bc->stuck_seen = 1u;
        bc->stuck = 0u;
      }
      for (uint32_t c = 0; c < (uint32_t)kBatchCounters; ++c)
        for (uint32_t j = 0; j < nj; ++j) proj_done[c * max_jobs + j] = 0;
      bc->ticket = 0;
      bc->exited = 0;
      bc->runs = bc->runs + 1u;
    }
  }
}

// The phased schedule: one launch a run, a workgroup an item.
__global__ __launch_bounds__(256) void kb_project(const BatchJob *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                                  uint32_t nj) {
  __shared__ ProjectSmem sm;
  const uint32_t j = find_job(starts, nj, blockIdx.x);
  item_project(sm, jobs[j], blockIdx.x - starts[j]);
}

// (kScanItems and kScan2Items are multiples of 8, so workgroup b's XCD is
// b mod 8 within every job too: the column groups are made XCD-contiguous.)
__global__ __launch_bounds__(256) void kb_scan(const BatchJob *__restrict__ jobs) {
  __shared__ uint32_t part[4][kBatchScanCols];
  item_scan(jobs[blockIdx.x / kScanItems], xcd_item(blockIdx.x % kScanItems, kScanItems), part);
}

template <bool HI, int S>
__global__ __launch_bounds__(256) void kb_scatter(const BatchJob *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                                  uint32_t nj) {
  __shared__ ScatterSmem<S> sm;
  const uint32_t *u_start = starts + (nj + 1);
  const uint32_t item = xcd_item(blockIdx.x, gridDim.x), j = find_job(u_start, nj, item);
  const BatchJob &J = jobs[j];
  if (HI)
    scatter_super<true, S>(sm, item - u_start[j], nullptr, nullptr, J.packed, J.n, J.tilehist3, 256, J.prefix_hi,
                           J.ctl->total + 256, J.ctl, J.order);
  else
    scatter_super<false, S>(sm, item - u_start[j], J.keys, J.mask, nullptr, J.n, J.tilehist, 512, J.prefix_lo,
                            J.ctl->total, J.ctl, J.packed);
}

__global__ __launch_bounds__(256) void kb_hist(const BatchJob *__restrict__ jobs, const uint32_t *__restrict__ starts,
                                               uint32_t nj) {
  __shared__ HistSmem sm;
  const uint32_t j = find_job(starts, nj, blockIdx.x);
  const BatchJob &J = jobs[j];
  hist_tile(sm, blockIdx.x - starts[j], J.packed, J.ctl, J.tilehist3);
}

__global__ __launch_bounds__(256) void kb_scan2(const BatchJob *__restrict__ jobs) {
  __shared__ uint32_t part[4][kBatchScanCols];
  item_scan2(jobs[blockIdx.x / kScan2Items], xcd_item(blockIdx.x % kScan2Items, kScan2Items), part);
}

}  // namespace

// This build's launchers: wd::kpt8 or wd::kpt16 by its keys a thread.
#define WD_NS_CAT2(a, b) a##b
#define WD_NS_CAT(a, b) WD_NS_CAT2(a, b)
namespace WD_NS_CAT(kpt, WD_KEYS_PER_THREAD) {

int batch_grid(wd_context *ctx, uint32_t items, uint32_t *grid) {
  int per_cu = 0;
  WD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch, 256, 0), "batch: occupancy");
  const uint32_t resident = (uint32_t)std::max(per_cu, 1) * (uint32_t)ctx->cus;
  *grid = std::max<uint32_t>(1, std::min(items, resident));
  return WD_OK;
}

int launch_persistent(wd_context *ctx, const BatchJob *jobs, const uint32_t *starts, uint32_t nj, uint32_t max_jobs,
                      BatchCtl *bctl, uint32_t grid, const uint32_t *gate, uint32_t run) {
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKBatch, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(k_batch, dim3(grid), dim3(256), 0, ctx->stream, jobs, starts, nj, max_jobs, bctl, gate, run);
  WD_HIP(hipGetLastError(), "batch: launch");
  return ctx->prof_end(kKBatch, ev);
}

int launch_batch(wd_context *ctx, wd_batch *b) {
  const uint32_t nj = b->n, nt = b->host_starts[nj], nu = b->host_starts[2 * nj + 1];
  if (b->schedule == WD_SCHEDULE_PERSISTENT)
    return launch_persistent(ctx, b->jobs, b->starts, nj, b->max_jobs, b->ctl, b->grid, nullptr, b->super_tiles);
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKBatch, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(kb_project, dim3(nt), dim3(256), 0, ctx->stream, b->jobs, b->starts, nj);
  hipLaunchKernelGGL(kb_scan, dim3(kScanItems * nj), dim3(256), 0, ctx->stream, b->jobs);
  if (b->super_tiles == 1)
    hipLaunchKernelGGL((kb_scatter<false, 1>), dim3(nu), dim3(256), 0, ctx->stream, b->jobs, b->starts, nj);
  else
    hipLaunchKernelGGL((kb_scatter<false, 2>), dim3(nu), dim3(256), 0, ctx->stream, b->jobs, b->starts, nj);
  hipLaunchKernelGGL(kb_hist, dim3(nt), dim3(256), 0, ctx->stream, b->jobs, b->starts, nj);
  hipLaunchKernelGGL(kb_scan2, dim3(kScan2Items * nj), dim3(256), 0, ctx->stream, b->jobs);
  if (b->super_tiles == 1)
    hipLaunchKernelGGL((kb_scatter<true, 1>), dim3(nu), dim3(256), 0, ctx->stream, b->jobs, b->starts, nj);
  else
    hipLaunchKernelGGL((kb_scatter<true, 2>), dim3(nu), dim3(256), 0, ctx->stream, b->jobs, b->starts, nj);
  WD_HIP(hipGetLastError(), "batch: launch");
  return ctx->prof_end(kKBatch, ev);
}

}  // namespace kpt8 / kpt16
}  // namespace wd

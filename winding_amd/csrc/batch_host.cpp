// Host side of batches (wd_batch_*): the job table, the item starts, the
// hipGraph a schedule is replayed from, and the autotuner.
//
// set() uploads the tables from pinned staging (two buffers used in turn; one
// is refilled only once the copy out of it has finished), in stream order
// behind the runs already enqueued, so a new frame's views cost no host wait.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>

#include "wd_internal.h"

using wd::fail;

extern "C" {

int wd_batch_create(wd_context *ctx, uint32_t max_jobs, wd_batch **out) {
  if (!ctx || !out || max_jobs == 0) return fail(WD_E_ARG, "batch: null argument or no jobs");
  *out = nullptr;
  WD_HIP(hipSetDevice(ctx->device), "batch: set device");
  wd_batch *b = new wd_batch();
  b->ctx = ctx;
  b->max_jobs = max_jobs;
  const size_t ctl_bytes = sizeof(wd::BatchCtl) + wd::kBatchCounters * sizeof(uint32_t) * max_jobs;
  hipError_t e = hipMalloc(&b->jobs, sizeof(wd::BatchJob) * max_jobs);
  if (e == hipSuccess) e = hipMalloc(&b->starts, sizeof(uint32_t) * 2 * (max_jobs + 1));
  if (e == hipSuccess) e = hipMalloc(&b->ctl, ctl_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(b->ctl, 0, ctl_bytes, ctx->stream);
  b->stage_bytes = sizeof(wd::BatchJob) * max_jobs + sizeof(uint32_t) * 2 * (max_jobs + 1);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    e = hipHostMalloc(&b->stage[i], b->stage_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->staged[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    wd_batch_destroy(b);
    return wd::hip_fail(e, "batch: allocation");
  }
  *out = b;
  return WD_OK;
}

int wd_batch_set(wd_batch *b, const wd_job *jobs, uint32_t n) {
  if (!b || b->magic != 0x42415443) return fail(WD_E_STATE, "batch: destroyed");
  if (!jobs || n == 0 || n > b->max_jobs) return fail(WD_E_ARG, "batch: %u jobs, this batch holds 1 to %u", n, b->max_jobs);
  std::vector<wd::BatchJob> hj(n);
  std::vector<uint32_t> t(n + 1, 0), u(n + 1, 0);  // item starts: tiles, super-tiles
  std::vector<std::pair<const void *, uint64_t>> shape;
  for (uint32_t i = 0; i < n; ++i) {
    const wd_job &J = jobs[i];
    if (!J.cloud || !b->ctx->live.count(J.cloud)) return fail(WD_E_STATE, "batch: job %u: its splats were unloaded", i);
    if (!J.state || !b->ctx->live.count(J.state)) return fail(WD_E_STATE, "batch: job %u: its sort state was destroyed", i);
    if (J.cloud->ctx != b->ctx || J.state->ctx != b->ctx) return fail(WD_E_STATE, "batch: job %u is of another device", i);
    if ((J.cloud->degree || !J.cloud->records || J.state->ellipse) && !wd::is_std(J.view))
      return fail(WD_E_ARG, "batch: job %u: splats with harmonics or given as centres and covariances, and states "
                  "that take the draw's ellipses, are batched under the reverse-Z perspective only; sort them with "
                  "wd_sort", i);
    if (J.cloud->n > wd::kNarrowMax)
      return fail(WD_E_CAPACITY, "batch: job %u has %llu splats; a batch job takes up to %u, sort it with wd_sort", i,
                  (unsigned long long)J.cloud->n, wd::kNarrowMax);
    if (J.cloud->n > J.state->capacity)
      return fail(WD_E_CAPACITY, "batch: job %u: %llu splats are more than its state holds, %llu", i,
                  (unsigned long long)J.cloud->n, (unsigned long long)J.state->capacity);
    if (!(J.view.cull >= 0 && J.view.cull <= 3.4e38f)) return fail(WD_E_ARG, "splatCull must be 0 or more, got %g", (double)J.view.cull);
    for (uint32_t k = 0; k < i; ++k)
      if (jobs[k].state == J.state) return fail(WD_E_ARG, "batch: jobs %u and %u share a sort state", k, i);
    wd::BatchJob &B = hj[i];
    B.view = wd::make_viewk(J.view);
    B.std_form = wd::is_std(J.view) ? 1u : 0u;
    B.rec = J.cloud->records ? (const void *)J.cloud->records : (const void *)J.cloud->centers;
    B.sig = J.cloud->records ? nullptr : (const uint2 *)J.cloud->covs;
    B.sh = J.cloud->degree ? J.cloud->sh : nullptr;
    B.degree = J.cloud->degree;
    B.ellipse = J.state->ellipse;
    B.capacity = J.state->capacity;
    B.n = (uint32_t)J.cloud->n;
    wd_state *s = J.state;
    B.keys = s->keys;
    B.mask = s->mask;
    B.cov = s->cov;
    B.packed = s->packed;
    B.order = s->order;
    B.ctl = s->ctl;
    B.tilehist = s->tilehist;
    B.tilehist3 = s->tilehist3;
    B.prefix_lo = s->prefix_lo;
    B.prefix_hi = s->prefix_hi;
    const uint32_t tile = 256u * b->kpt;
    B.tiles = (B.n + tile - 1) / tile;
    if (b->schedule == WD_SCHEDULE_PERSISTENT) {
      // a persistent item is a run of super_tiles tiles (scatters: super-tiles of kSuper)
      const uint32_t R = b->super_tiles, supers = (B.tiles + wd::kSuper - 1) / wd::kSuper;
      t[i + 1] = t[i] + (B.tiles + R - 1) / R;
      u[i + 1] = u[i] + (supers + R - 1) / R;
    } else {
      t[i + 1] = t[i] + B.tiles;
      u[i + 1] = u[i] + (B.tiles + b->super_tiles - 1) / b->super_tiles;
    }
    s->last_n = B.n;
    shape.emplace_back(J.cloud, b->ctx->serial(J.cloud));
    shape.emplace_back(J.state, b->ctx->serial(J.state));
  }
  std::vector<uint32_t> starts(2 * (n + 1), 0);
  std::copy(t.begin(), t.end(), starts.begin());
  std::copy(u.begin(), u.end(), starts.begin() + (n + 1));
  wd_context *ctx = b->ctx;
  WD_HIP(hipSetDevice(ctx->device), "batch: set device");
  // The table is read by the kernels at run time: a later set() with new views
  // and the same clouds and states replays the same graph.
  const int f = b->flip;
  if (b->staged_used[f]) WD_HIP(hipEventSynchronize(b->staged[f]), "batch: staging");
  uint8_t *stage = (uint8_t *)b->stage[f];
  std::memcpy(stage, hj.data(), sizeof(wd::BatchJob) * n);
  std::memcpy(stage + sizeof(wd::BatchJob) * b->max_jobs, starts.data(), sizeof(uint32_t) * starts.size());
  WD_HIP(hipMemcpyAsync(b->jobs, stage, sizeof(wd::BatchJob) * n, hipMemcpyHostToDevice, ctx->stream),
         "batch: job table");
  WD_HIP(hipMemcpyAsync(b->starts, stage + sizeof(wd::BatchJob) * b->max_jobs, sizeof(uint32_t) * starts.size(),
                        hipMemcpyHostToDevice, ctx->stream), "batch: item table");
  WD_HIP(hipEventRecord(b->staged[f], ctx->stream), "batch: staging");
  b->staged_used[f] = true;
  b->flip ^= 1;
  const uint32_t total = 2 * t[n] + 2 * u[n] + (512 / wd::kBatchScanCols + 256 / wd::kBatchScanCols) * n;
  const bool same_total = total == b->total;
  if (shape != b->shape || n != b->n || total != b->total) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  b->host_jobs.swap(hj);
  b->host_starts.swap(starts);
  b->shape.swap(shape);
  if (b->api_jobs.data() != jobs) b->api_jobs.assign(jobs, jobs + n);
  b->sh_set.clear();
  b->ellipse_set.clear();
  for (uint32_t i = 0; i < n; ++i) {
    b->sh_set.emplace_back(jobs[i].cloud->sh, jobs[i].cloud->degree);
    b->ellipse_set.push_back(jobs[i].state->ellipse);
  }
  b->n = n;
  b->total = total;
  if (same_total && b->grid) return WD_OK;
  return b->kpt == 8 ? wd::kpt8::batch_grid(ctx, total, &b->grid) : wd::kpt16::batch_grid(ctx, total, &b->grid);
}

namespace {
int launch_batch(wd_context *ctx, wd_batch *b) {
  return b->kpt == 8 ? wd::kpt8::launch_batch(ctx, b) : wd::kpt16::launch_batch(ctx, b);
}
}  // namespace

int wd_batch_run(wd_context *ctx, wd_batch *b) {
  if (!ctx || !b || b->magic != 0x42415443 || b->ctx != ctx) return fail(WD_E_STATE, "batch: destroyed or of another device");
  if (b->n == 0) return fail(WD_E_STATE, "batch: no jobs set");
  // by serial, not address: an object made after a destroy may sit at the same address
  for (const auto &p : b->shape)
    if (ctx->serial(p.first) != p.second)
      return fail(WD_E_STATE, "batch: a job's splats or sort state was destroyed; set the jobs again");
  for (size_t i = 0; i < b->api_jobs.size(); ++i) {
    const wd_cloud *c = b->api_jobs[i].cloud;
    if (c->sh != b->sh_set[i].first || c->degree != b->sh_set[i].second)
      return fail(WD_E_STATE, "batch: a job's harmonics changed since the jobs were set; set the jobs again");
  }
  for (size_t i = 0; i < b->api_jobs.size(); ++i)
    if (b->api_jobs[i].state->ellipse != b->ellipse_set[i])
      return fail(WD_E_STATE, "batch: a job's ellipses were turned on or off since the jobs were set; set the jobs again");
  WD_HIP(hipSetDevice(ctx->device), "batch: set device");
  if (ctx->profiling) return launch_batch(ctx, b);  // timed launches go straight to the stream
  if (!b->exec) {
    WD_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "batch: capture");
    int rc = launch_batch(ctx, b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    WD_HIP(e, "batch: end capture");
    b->graph = g;
    WD_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "batch: instantiate");
  }
  WD_HIP(hipGraphLaunch(b->exec, ctx->stream), "batch: graph launch");
  return WD_OK;
}

int wd_batch_configure(wd_batch *b, int schedule, uint32_t super_tiles, uint32_t keys_per_thread) {
  if (!b || b->magic != 0x42415443) return fail(WD_E_STATE, "batch: destroyed");
  if (schedule != WD_SCHEDULE_PHASED && schedule != WD_SCHEDULE_PERSISTENT)
    return fail(WD_E_ARG, "batch: schedule %d is neither phased (0) nor persistent (1)", schedule);
  if (schedule == WD_SCHEDULE_PHASED && super_tiles != 1 && super_tiles != 2)
    return fail(WD_E_ARG, "batch: phased super-tiles of %u; 1 or 2", super_tiles);
  if (schedule == WD_SCHEDULE_PERSISTENT && (super_tiles < 1 || super_tiles > wd::kMaxRun))
    return fail(WD_E_ARG, "batch: persistent runs of %u tiles; 1 to %u", super_tiles, wd::kMaxRun);
  if (keys_per_thread != 8 && keys_per_thread != 16)
    return fail(WD_E_ARG, "batch: %u keys a thread; 8 or 16", keys_per_thread);
  if (schedule == b->schedule && super_tiles == b->super_tiles && keys_per_thread == b->kpt) return WD_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->grid = 0;  // the persistent kernel's occupancy is the build's
  b->schedule = schedule;
  b->super_tiles = super_tiles;
  b->kpt = keys_per_thread;
  if (b->n == 0) return WD_OK;
  // the item starts depend on the work item and the tile size: set the same jobs again
  const std::vector<wd_job> jobs = b->api_jobs;
  return wd_batch_set(b, jobs.data(), (uint32_t)jobs.size());
}

int wd_batch_schedule(wd_batch *b, int schedule, uint32_t super_tiles) {
  if (!b || b->magic != 0x42415443) return fail(WD_E_STATE, "batch: destroyed");
  return wd_batch_configure(b, schedule, super_tiles, b->kpt);
}

int wd_batch_tune(wd_context *ctx, wd_batch *b, uint32_t reps, wd_tuning *out) {
  if (!ctx || !b || b->magic != 0x42415443) return fail(WD_E_STATE, "batch: destroyed");
  if (b->n == 0) return fail(WD_E_STATE, "batch: no jobs set");
  if (ctx->profiling) return fail(WD_E_STATE, "batch: tuning times whole runs; switch kernel profiling off first");
  reps = std::max(reps, 1u);
  // schedule, work item, keys a thread (tiles of 256 x keys)
  struct Config {
    int schedule;
    uint32_t item, kpt;
  };
  static const Config cfg[WD_TUNE_CONFIGS] = {
      {WD_SCHEDULE_PHASED, 1, 16}, {WD_SCHEDULE_PHASED, 2, 16}, {WD_SCHEDULE_PHASED, 1, 8},
      {WD_SCHEDULE_PHASED, 2, 8},  {WD_SCHEDULE_PERSISTENT, wd::kPersistentRun, 16},
      {WD_SCHEDULE_PERSISTENT, wd::kPersistentRun, 8}};
  double ms[WD_TUNE_CONFIGS] = {};
  hipEvent_t a = nullptr, z = nullptr;
  WD_HIP(hipEventCreate(&a), "batch: tune events");
  hipError_t e = hipEventCreate(&z);
  int rc = WD_OK;
  for (int c = 0; c < WD_TUNE_CONFIGS && e == hipSuccess && rc == WD_OK; ++c) {
    rc = wd_batch_configure(b, cfg[c].schedule, cfg[c].item, cfg[c].kpt);
    if (!rc) rc = wd_batch_run(ctx, b);  // capture and warm
    if (!rc) e = hipEventRecord(a, ctx->stream);
    for (uint32_t r = 0; r < reps && !rc && e == hipSuccess; ++r) rc = wd_batch_run(ctx, b);
    if (!rc && e == hipSuccess) e = hipEventRecord(z, ctx->stream);
    if (!rc && e == hipSuccess) e = hipEventSynchronize(z);
    float t = 0.f;
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, a, z);
    ms[c] = (double)t / reps;
  }
  (void)hipEventDestroy(a);
  if (z) (void)hipEventDestroy(z);
  if (rc) return rc;
  if (e != hipSuccess) return wd::hip_fail(e, "batch: tune");
  int best = 0;
  for (int c = 1; c < WD_TUNE_CONFIGS; ++c)
    if (ms[c] < ms[best]) best = c;
  if ((rc = wd_batch_configure(b, cfg[best].schedule, cfg[best].item, cfg[best].kpt))) return rc;
  if (out) {
    out->schedule = cfg[best].schedule;
    out->super_tiles = cfg[best].item;
    out->keys_per_thread = cfg[best].kpt;
    out->tile = 256u * cfg[best].kpt;
    out->radix_bits = 8;
    out->passes = 2;
    for (int c = 0; c < WD_TUNE_CONFIGS; ++c) {
      out->timed[c].schedule = cfg[c].schedule;
      out->timed[c].super_tiles = cfg[c].item;
      out->timed[c].keys_per_thread = cfg[c].kpt;
      out->timed[c].ms = ms[c];
    }
  }
  return WD_OK;
}

int wd_batch_stuck(wd_context *ctx, wd_batch *b, uint32_t *stuck) {
  if (!ctx || !b || b->magic != 0x42415443 || !stuck) return fail(WD_E_ARG, "batch: null argument or destroyed");
  WD_HIP(hipStreamSynchronize(ctx->stream), "batch: sync");
  uint32_t flags[2] = {0, 0};  // stuck, stuck_seen
  WD_HIP(hipMemcpy(flags, &b->ctl->stuck, sizeof flags, hipMemcpyDeviceToHost), "batch: read");
  *stuck = (flags[0] | flags[1]) ? 1u : 0u;
  if (flags[1]) {
    // reported once: the next call tells of runs after this one
    const uint32_t zero = 0;
    WD_HIP(hipMemcpy(&b->ctl->stuck_seen, &zero, sizeof zero, hipMemcpyHostToDevice), "batch: clear");
  }
  return WD_OK;
}

int wd_debug_batch_force_stuck(wd_context *ctx, wd_batch *b) {
  if (!ctx || !b || b->magic != 0x42415443 || b->ctx != ctx) return fail(WD_E_ARG, "batch: null argument or destroyed");
  WD_HIP(hipSetDevice(ctx->device), "batch: set device");
  static const uint32_t one = 1;
  WD_HIP(hipMemcpyAsync(&b->ctl->stuck, &one, sizeof one, hipMemcpyHostToDevice, ctx->stream), "batch: force");
  WD_HIP(hipStreamSynchronize(ctx->stream), "batch: force");
  return WD_OK;
}

int wd_batch_destroy(wd_batch *b) {
  if (!b || b->magic != 0x42415443) return fail(WD_E_STATE, "batch: destroyed twice");
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->jobs) (void)hipFree(b->jobs);
  if (b->starts) (void)hipFree(b->starts);
  if (b->ctl) (void)hipFree(b->ctl);
  for (int i = 0; i < 2; ++i) {
    if (b->stage[i]) (void)hipHostFree(b->stage[i]);
    if (b->staged[i]) (void)hipEventDestroy(b->staged[i]);
  }
  b->magic = 0;
  delete b;
  return WD_OK;
}

}  // extern "C"

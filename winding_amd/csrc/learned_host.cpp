// Host side of the learned sort (wd_learned_*): the per-run job table, the
// predictor's forward pass, the learned kernels, and the gated exact fallback.
//
// A run takes up to kMaxViews views of one cloud. Everything is enqueued on
// the context's stream with no host wait: the poses and the job table go up
// from pinned staging buffers (two, used in turn; a buffer is refilled only
// after the copy out of it two runs ago has finished), the predictor and the
// learned kernels run, and the persistent batch kernel is launched behind a
// device gate that the learned kernels set when any view fails its check.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "wd_internal.h"

using wd::fail;

struct wd_learned {
  wd_context *ctx = nullptr;
  uint32_t max_views = 0;
  wd::BatchJob *jobs = nullptr;          // device [max_views]
  float *poses = nullptr;                // device [max_views][12]
  wd::BatchJob *host_jobs[2] = {nullptr, nullptr};  // pinned staging
  float *host_poses[2] = {nullptr, nullptr};
  hipEvent_t staged[2] = {nullptr, nullptr};
  bool staged_used[2] = {false, false};
  int flip = 0;
  uint32_t *gate = nullptr;              // device: some view of the run failed its check
  // the fallback: the persistent batch kernel over the same job table
  uint32_t *starts = nullptr;            // device: tile and super-tile starts, 2 x (max_views + 1)
  wd::BatchCtl *bctl = nullptr;
  uint32_t grid = 0, last_views = 0, last_tiles = 0;
  uint32_t magic = 0x4c524e44;
};

namespace {

bool alive(const wd_learned *l) { return l && l->magic == 0x4c524e44; }

// Learned-sort buffers of a state, made on its first learned run.
int ensure_learned(wd_context *ctx, wd_state *s) {
  // each buffer made once; a failed allocation leaves the state to try again on its next run
  hipError_t e = hipSuccess;
  if (!s->lut) e = hipMalloc(&s->lut, wd::kBinBlocks * sizeof(uint16_t));
  if (e == hipSuccess && !s->prefix_bin) e = hipMalloc(&s->prefix_bin, s->tiles * 512 * sizeof(uint32_t));
  if (e != hipSuccess) return wd::hip_fail(e, "learned: state buffers");
  return WD_OK;
}

}  // namespace

extern "C" {

int wd_learned_create(wd_context *ctx, uint32_t max_views, wd_learned **out) {
  if (!ctx || !out) return fail(WD_E_ARG, "learned: null argument");
  *out = nullptr;
  if (max_views == 0 || max_views > wd::kMaxViews)
    return fail(WD_E_ARG, "learned: %u views a run; one run takes 1 to %u", max_views, wd::kMaxViews);
  WD_HIP(hipSetDevice(ctx->device), "learned: set device");
  wd_learned *l = new wd_learned();
  l->ctx = ctx;
  l->max_views = max_views;
  const size_t ctl_bytes = sizeof(wd::BatchCtl) + wd::kBatchCounters * sizeof(uint32_t) * max_views;
  hipError_t e = hipMalloc(&l->jobs, sizeof(wd::BatchJob) * max_views);
  if (e == hipSuccess) e = hipMalloc(&l->poses, sizeof(float) * WD_POSE_FLOATS * max_views);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    e = hipHostMalloc(&l->host_jobs[i], sizeof(wd::BatchJob) * max_views, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&l->host_poses[i], sizeof(float) * WD_POSE_FLOATS * max_views, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&l->staged[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipMalloc(&l->gate, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&l->starts, sizeof(uint32_t) * 2 * (max_views + 1));
  if (e == hipSuccess) e = hipMalloc(&l->bctl, ctl_bytes);
  if (e == hipSuccess) e = hipMemsetAsync(l->bctl, 0, ctl_bytes, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(l->gate, 0, sizeof(uint32_t), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    wd_learned_destroy(l);
    return wd::hip_fail(e, "learned: allocation");
  }
  *out = l;
  return WD_OK;
}

int wd_learned_run(wd_context *ctx, wd_learned *l, wd_mlp *mlp, const wd_cloud *cloud, const wd_view *views,
                   wd_state *const *states, uint32_t n_views, uint32_t cap) {
  if (!ctx || !alive(l) || !mlp || !views || !states) return fail(WD_E_ARG, "learned: null argument or destroyed");
  if (l->ctx != ctx) return fail(WD_E_STATE, "learned: of another device");
  if (n_views == 0 || n_views > l->max_views)
    return fail(WD_E_ARG, "learned: %u views; this one takes 1 to %u", n_views, l->max_views);
  if (!cloud || !ctx->live.count(cloud)) return fail(WD_E_STATE, "learned: the splats were unloaded");
  if (cloud->ctx != ctx) return fail(WD_E_STATE, "learned: the splats are of another device");
  if (cloud->degree) return fail(WD_E_ARG, "learned: the splats have harmonics, which only a single sort (wd_sort) shades");
  if (!cloud->records)
    return fail(WD_E_ARG, "learned: the splats are centres and covariances, which only a single sort (wd_sort) takes");
  const uint32_t n = (uint32_t)cloud->n;
  if (cloud->n > wd::kLearnedMaxSplats)
    return fail(WD_E_CAPACITY, "learned: %llu splats are more than a learned sort takes, %u",
                (unsigned long long)cloud->n, wd::kLearnedMaxSplats);
  for (uint32_t v = 0; v < n_views; ++v) {
    wd_state *s = states[v];
    if (!s || !ctx->live.count(s)) return fail(WD_E_STATE, "learned: view %u: its sort state was destroyed", v);
    if (s->ctx != ctx) return fail(WD_E_STATE, "learned: view %u: its sort state is of another device", v);
    if (s->ellipse)
      return fail(WD_E_ARG, "learned: view %u: its state takes the draw's ellipses, which only a single sort (wd_sort) writes", v);
    if (cloud->n > s->capacity)
      return fail(WD_E_CAPACITY, "learned: view %u: %u splats are more than its state holds, %llu", v, n,
                  (unsigned long long)s->capacity);
    if (!(views[v].cull >= 0 && views[v].cull <= 3.4e38f))
      return fail(WD_E_ARG, "splatCull must be 0 or more, got %g", (double)views[v].cull);
    for (uint32_t k = 0; k < v; ++k)
      if (states[k] == s) return fail(WD_E_ARG, "learned: views %u and %u share a sort state", k, v);
  }
  WD_HIP(hipSetDevice(ctx->device), "learned: set device");
  for (uint32_t v = 0; v < n_views; ++v) {
    const int rc = ensure_learned(ctx, states[v]);
    if (rc) return rc;
  }
  const uint32_t tiles = (n + wd::kTile - 1) / wd::kTile, supers = (tiles + wd::kSuper - 1) / wd::kSuper;
  if (n_views != l->last_views || tiles != l->last_tiles) {
    // the fallback's item table: every view has the same tiles
    std::vector<uint32_t> st(2 * (n_views + 1));
    const uint32_t R = wd::kPersistentRun, runs = (tiles + R - 1) / R, sruns = (supers + R - 1) / R;
    for (uint32_t v = 0; v <= n_views; ++v) {
      st[v] = v * runs;
      st[n_views + 1 + v] = v * sruns;
    }
    WD_HIP(hipMemcpyAsync(l->starts, st.data(), st.size() * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream),
           "learned: item table");
    WD_HIP(hipStreamSynchronize(ctx->stream), "learned: item table");
    const uint32_t items = n_views * (2 * runs + 2 * sruns + 512 / wd::kBatchScanCols + 256 / wd::kBatchScanCols);
    const int rc = wd::kpt16::batch_grid(ctx, items, &l->grid);
    if (rc) return rc;
    l->last_views = n_views;
    l->last_tiles = tiles;
  }
  const int f = l->flip;
  if (l->staged_used[f]) WD_HIP(hipEventSynchronize(l->staged[f]), "learned: staging");
  bool all_std = true;
  for (uint32_t v = 0; v < n_views; ++v) {
    const int rc = wd_pose_features(&views[v], l->host_poses[f] + WD_POSE_FLOATS * v);
    if (rc) return rc;
    all_std = all_std && wd::is_std(views[v]);
  }
  WD_HIP(hipMemcpyAsync(l->poses, l->host_poses[f], sizeof(float) * WD_POSE_FLOATS * n_views, hipMemcpyHostToDevice,
                        ctx->stream), "learned: poses");
  const float *probs = nullptr;
  int rc = wd::mlp_infer(ctx, mlp, l->poses, n_views, &probs);
  if (rc) return rc;
  const uint32_t bin_cap = cap ? cap : 8u * ((n + 255u) / 256u) + 4096u;
  for (uint32_t v = 0; v < n_views; ++v) {
    wd::BatchJob &B = l->host_jobs[f][v];
    std::memset(&B, 0, sizeof B);
    wd_state *s = states[v];
    B.view = wd::make_viewk(views[v]);
    B.std_form = wd::is_std(views[v]) ? 1u : 0u;
    B.rec = cloud->records;
    B.n = n;
    B.tiles = tiles;
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
    B.lut = s->lut;
    B.prefix_bin = s->prefix_bin;
    B.probs = probs + 256 * v;
    B.cap = bin_cap;
    s->last_n = n;
  }
  WD_HIP(hipMemcpyAsync(l->jobs, l->host_jobs[f], sizeof(wd::BatchJob) * n_views, hipMemcpyHostToDevice, ctx->stream),
         "learned: job table");
  WD_HIP(hipEventRecord(l->staged[f], ctx->stream), "learned: staging");
  l->staged_used[f] = true;
  l->flip ^= 1;
  if ((rc = wd::launch_learned(ctx, l->jobs, n_views, n, all_std, l->gate))) return rc;
  return wd::kpt16::launch_persistent(ctx, l->jobs, l->starts, n_views, l->max_views, l->bctl, l->grid, l->gate,
                                      wd::kPersistentRun);
}

int wd_learned_destroy(wd_learned *l) {
  if (!alive(l)) return fail(WD_E_STATE, "learned: destroyed twice");
  (void)hipSetDevice(l->ctx->device);
  (void)hipStreamSynchronize(l->ctx->stream);
  for (int i = 0; i < 2; ++i) {
    if (l->host_jobs[i]) (void)hipHostFree(l->host_jobs[i]);
    if (l->host_poses[i]) (void)hipHostFree(l->host_poses[i]);
    if (l->staged[i]) (void)hipEventDestroy(l->staged[i]);
  }
  for (void *p : {(void *)l->jobs, (void *)l->poses, (void *)l->gate, (void *)l->starts, (void *)l->bctl})
    if (p) (void)hipFree(p);
  l->magic = 0;
  delete l;
  return WD_OK;
}

int wd_state_quantiles(wd_context *ctx, wd_state *s, uint32_t edges[257]) {
  if (!ctx || !s || s->magic != 0x53544154 || !edges) return fail(WD_E_ARG, "quantiles: null argument or destroyed");
  WD_HIP(hipSetDevice(ctx->device), "quantiles: set device");
  uint32_t *dev = nullptr;
  WD_HIP(hipMalloc(&dev, 257 * sizeof(uint32_t)), "quantiles: allocation");
  int rc = wd::launch_quantiles(ctx, s, dev);
  hipError_t e = rc ? hipSuccess : hipStreamSynchronize(ctx->stream);
  if (!rc && e == hipSuccess) e = hipMemcpy(edges, dev, 257 * sizeof(uint32_t), hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  if (rc) return rc;
  if (e != hipSuccess) return wd::hip_fail(e, "quantiles: read");
  return WD_OK;
}

}  // extern "C"

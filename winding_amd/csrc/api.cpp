// Host side of the C ABI (include/winding_amd.h): contexts, clouds, sort
// states, reads, timing. Kernels and their launchers are in splat_sort.hip.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "wd_internal.h"

namespace {
thread_local char g_error[512] = "";
}

namespace wd {

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char *what) { return fail(WD_E_HIP, "%s: %s", what, hipGetErrorString(e)); }

void box_from_ordered(const uint32_t lo[3], const uint32_t hi[3], const uint32_t zero_index[3],
                      float (*value_at)(const void *, uint32_t, int), const void *arg, float mn[3], float mx[3]) {
  for (int k = 0; k < 3; ++k) {
    uint32_t a = lo[k], b = hi[k];
    a = (a & 0x80000000u) ? (a & 0x7FFFFFFFu) : ~a;
    b = (b & 0x80000000u) ? (b & 0x7FFFFFFFu) : ~b;
    std::memcpy(&mn[k], &a, 4);
    std::memcpy(&mx[k], &b, 4);
    if ((mn[k] == 0.0f || mx[k] == 0.0f) && zero_index[k] != 0xFFFFFFFFu) {
      const float z = value_at(arg, zero_index[k], k);
      if (mn[k] == 0.0f) mn[k] = z;
      if (mx[k] == 0.0f) mx[k] = z;
    }
  }
}

int launch_selftest_kernel(wd_context *ctx, const double *args, const double *expect, uint32_t count,
                           uint32_t *mismatch3);

}  // namespace wd

using wd::fail;

int wd_context::prof_begin(int kernel, hipEvent_t *a) {
  *a = nullptr;
  if (!profiling) return WD_OK;
  if (spare.empty()) {
    hipEvent_t e;
    WD_HIP(hipEventCreate(&e), "profile: event");
    spare.push_back(e);
  }
  *a = spare.back();
  spare.pop_back();
  WD_HIP(hipEventRecord(*a, stream), "profile: record");
  (void)kernel;
  return WD_OK;
}

int wd_context::prof_end(int kernel, hipEvent_t a) {
  if (!profiling || !a) return WD_OK;
  if (spare.empty()) {
    hipEvent_t e;
    WD_HIP(hipEventCreate(&e), "profile: event");
    spare.push_back(e);
  }
  hipEvent_t b = spare.back();
  spare.pop_back();
  WD_HIP(hipEventRecord(b, stream), "profile: record");
  pending.push_back({kernel, a, b});
  return WD_OK;
}

int wd_context::prof_collect() {
  if (pending.empty()) return WD_OK;
  WD_HIP(hipStreamSynchronize(stream), "profile: sync");
  for (auto &p : pending) {
    float ms = 0;
    WD_HIP(hipEventElapsedTime(&ms, p.a, p.b), "profile: elapsed");
    prof_ms[p.kernel] += ms;
    prof_launches[p.kernel] += 1;
    spare.push_back(p.a);
    spare.push_back(p.b);
  }
  pending.clear();
  return WD_OK;
}

extern "C" {

const char *wd_last_error(void) { return g_error; }

int wd_open(int device, wd_context **out) {
  if (!out) return fail(WD_E_ARG, "open: out is null");
  *out = nullptr;
  int count = 0;
  WD_HIP(hipGetDeviceCount(&count), "open: no HIP device");
  if (device < 0 || device >= count) return fail(WD_E_ARG, "open: device %d, of %d", device, count);
  WD_HIP(hipSetDevice(device), "open: set device");
  hipDeviceProp_t prop;
  WD_HIP(hipGetDeviceProperties(&prop, device), "open: properties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(WD_E_ARG, "open: device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
  wd_context *ctx = new wd_context();
  ctx->device = device;
  ctx->cus = prop.multiProcessorCount;
  size_t free_b = 0, total_b = 0;
  hipError_t e = hipMemGetInfo(&free_b, &total_b);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return wd::hip_fail(e, "open: stream");
  }
  ctx->mem_total = total_b;
  ctx->mem_free_at_open = free_b;
  // the no-allocation sort's array: its address on this device, taken once here
  void *scratch = nullptr;
  const int rc = wd::noalloc_scratch(&scratch);
  if (rc != WD_OK) {
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return rc;
  }
  ctx->noalloc = (uint8_t *)scratch;
  *out = ctx;
  return WD_OK;
}

int wd_close(wd_context *ctx) {
  if (!ctx || ctx->closed) return fail(WD_E_STATE, "close: no context, or closed already");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &p : ctx->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : ctx->spare) (void)hipEventDestroy(e);
  if (ctx->particle_edges) (void)hipFree(ctx->particle_edges);
  (void)hipStreamDestroy(ctx->stream);
  ctx->closed = true;
  delete ctx;
  return WD_OK;
}

int wd_sync(wd_context *ctx) {
  if (!ctx) return fail(WD_E_ARG, "sync: no context");
  WD_HIP(hipStreamSynchronize(ctx->stream), "sync");
  return WD_OK;
}

int wd_stream(wd_context *ctx, void **stream) {
  if (!ctx || !stream) return fail(WD_E_ARG, "stream: null argument");
  *stream = (void *)ctx->stream;
  return WD_OK;
}

int wd_capacity(wd_context *ctx, uint64_t bytes_per_item, uint64_t *out) {
  if (!ctx || !out || bytes_per_item == 0) return fail(WD_E_ARG, "capacity: null argument or zero item size");
  *out = ctx->mem_free_at_open / bytes_per_item;
  return WD_OK;
}

static int cloud_from(wd_context *ctx, const void *src, uint64_t bytes, bool device_src, wd_cloud **out,
                      float *min_out, float *max_out) {
  if (!ctx || !src || !out) return fail(WD_E_ARG, "splats: null argument");
  *out = nullptr;
  if (bytes == 0 || bytes % WD_RECORD_BYTES != 0)
    return fail(WD_E_DATA, "splats: not a .ply, and %llu bytes is not a whole number of 32-byte .splat records",
                (unsigned long long)bytes);
  const uint64_t n = bytes / WD_RECORD_BYTES;
  if (n > WD_MAX_SORT)
    return fail(WD_E_CAPACITY, "splats: %llu splats are more than this device holds in one buffer, %u",
                (unsigned long long)n, WD_MAX_SORT);
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  wd_cloud *c = new wd_cloud();
  c->ctx = ctx;
  c->n = n;
  uint32_t *scratch = nullptr;
  uint32_t res[12];
  hipError_t e = hipMalloc(&c->records, bytes);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->records, src, bytes, device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess)
    e = hipMalloc(&scratch, (12 + (size_t)wd::kBoxWords * wd::bounds_partials(ctx)) * sizeof(uint32_t));
  if (e != hipSuccess) {
    if (scratch) (void)hipFree(scratch);
    (void)hipFree(c->records);
    delete c;
    return wd::hip_fail(e, "splats: upload");
  }
  int rc = wd::launch_bounds(ctx, c->records, n, scratch);
  if (rc == WD_OK) {
    e = hipMemcpyAsync(res, scratch, sizeof res, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = wd::hip_fail(e, "splats: check");
  }
  (void)hipFree(scratch);
  if (rc == WD_OK && res[6] != 0xFFFFFFFFu) rc = fail(WD_E_DATA, "splats: splat %u has a position that is not a number", res[6]);
  if (rc != WD_OK) {
    (void)hipFree(c->records);
    delete c;
    return rc;
  }
  // a zero end of the box takes its sign from the first zero, read from the records as uploaded
  auto at = [](const void *recs, uint32_t i, int k) -> float {
    float v = 0.f;
    if (hipMemcpy(&v, (const uint8_t *)recs + (size_t)i * WD_RECORD_BYTES + 4 * k, 4, hipMemcpyDeviceToHost) != hipSuccess)
      return 0.f;
    return v;
  };
  wd::box_from_ordered(res, res + 3, res + 7, at, c->records, c->min, c->max);
  if (min_out) std::memcpy(min_out, c->min, sizeof c->min);
  if (max_out) std::memcpy(max_out, c->max, sizeof c->max);
  ctx->adopt(c);
  *out = c;
  return WD_OK;
}

int wd_cloud_create(wd_context *ctx, const void *records, uint64_t bytes, wd_cloud **out, float min_out[3],
                    float max_out[3]) {
  return cloud_from(ctx, records, bytes, false, out, min_out, max_out);
}

int wd_cloud_create_device(wd_context *ctx, const void *device_records, uint64_t bytes, wd_cloud **out,
                           float min_out[3], float max_out[3]) {
  return cloud_from(ctx, device_records, bytes, true, out, min_out, max_out);
}

int wd_cloud_create_splats(wd_context *ctx, const float *centers, const float *covariances, uint64_t count,
                           wd_cloud **out, float min_out[3], float max_out[3]) {
  if (!ctx || !centers || !covariances || !out) return fail(WD_E_ARG, "splats: null argument");
  *out = nullptr;
  if (count == 0) return fail(WD_E_DATA, "splats: no splats");
  if (count > WD_MAX_SORT)
    return fail(WD_E_CAPACITY, "splats: %llu splats are more than this device holds in one buffer, %u",
                (unsigned long long)count, WD_MAX_SORT);
  // put()'s refusal and box (src/scene/splats.js:336-348), over the f32 centres
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = 0; i < count; ++i)
    for (int k = 0; k < 3; ++k) {
      const float x = centers[4 * i + k];
      if (!std::isfinite(x))
        return fail(WD_E_DATA, "splats: splat %llu has a position that is not a number", (unsigned long long)i);
      mn[k] = std::min(mn[k], x);
      mx[k] = std::max(mx[k], x);
    }
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  wd_cloud *c = new wd_cloud();
  c->ctx = ctx;
  c->n = count;
  hipError_t e = hipMalloc(&c->centers, count * 16);
  if (e == hipSuccess) e = hipMalloc(&c->covs, count * 24);
  if (e == hipSuccess) e = hipMemcpy(c->centers, centers, count * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->covs, covariances, count * 24, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (c->centers) (void)hipFree(c->centers);
    if (c->covs) (void)hipFree(c->covs);
    delete c;
    return wd::hip_fail(e, "splats: upload");
  }
  std::memcpy(c->min, mn, sizeof mn);
  std::memcpy(c->max, mx, sizeof mx);
  if (min_out) std::memcpy(min_out, mn, sizeof mn);
  if (max_out) std::memcpy(max_out, mx, sizeof mx);
  ctx->adopt(c);
  *out = c;
  return WD_OK;
}

int wd_cloud_info(const wd_cloud *cloud, uint64_t *count, float min_out[3], float max_out[3]) {
  if (!cloud || cloud->magic != 0x53504c54) return fail(WD_E_STATE, "splats: these splats were unloaded");
  if (count) *count = cloud->n;
  if (min_out) std::memcpy(min_out, cloud->min, sizeof cloud->min);
  if (max_out) std::memcpy(max_out, cloud->max, sizeof cloud->max);
  return WD_OK;
}

int wd_cloud_sh(wd_context *ctx, wd_cloud *cloud, const uint32_t *words, uint64_t n_words, uint32_t degree) {
  if (!ctx || !cloud || cloud->magic != 0x53504c54) return fail(WD_E_STATE, "splats: these splats were unloaded");
  if (cloud->ctx != ctx) return fail(WD_E_STATE, "splats: of another device");
  if (degree > 3) return fail(WD_E_ARG, "splats: harmonics of degree %u; this path takes 0 to 3", degree);
  const uint64_t need = degree ? cloud->n * wd::sh_words(degree) : 0;
  if (degree && (!words || n_words < need))
    return fail(WD_E_ARG, "splats: harmonics of degree %u take %llu words for %llu splats, got %llu", degree,
                (unsigned long long)need, (unsigned long long)cloud->n, (unsigned long long)n_words);
  // src/render/splats.js:398-399: the harmonics must fit one buffer of this device
  if (degree && cloud->n > ctx->mem_free_at_open / (wd::sh_words(degree) * sizeof(uint32_t)))
    return fail(WD_E_CAPACITY, "splats: %llu splats with harmonics of degree %u are more than this device holds in one buffer",
                (unsigned long long)cloud->n, degree);
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  WD_HIP(hipStreamSynchronize(ctx->stream), "splats: harmonics");  // no sort in flight reads the old ones
  if (cloud->sh) (void)hipFree(cloud->sh);
  cloud->sh = nullptr;
  cloud->degree = 0;
  if (!degree) return WD_OK;
  // one word over, as the reference allocates: a two-word read of the last coefficient never falls off
  WD_HIP(hipMalloc(&cloud->sh, (need + 1) * sizeof(uint32_t)), "splats: harmonics allocation");
  WD_HIP(hipMemcpy(cloud->sh, words, need * sizeof(uint32_t), hipMemcpyHostToDevice), "splats: harmonics upload");
  WD_HIP(hipMemset(cloud->sh + need, 0, sizeof(uint32_t)), "splats: harmonics upload");
  cloud->degree = degree;
  return WD_OK;
}

int wd_cloud_records(const wd_cloud *cloud, const void **device_records) {
  if (!cloud || cloud->magic != 0x53504c54 || !device_records)
    return fail(WD_E_STATE, "splats: these splats were unloaded");
  *device_records = cloud->records;
  return WD_OK;
}

int wd_cloud_destroy(wd_cloud *cloud) {
  if (!cloud || cloud->magic != 0x53504c54) return fail(WD_E_STATE, "splats: destroyed twice");
  (void)hipSetDevice(cloud->ctx->device);
  (void)hipStreamSynchronize(cloud->ctx->stream);
  if (cloud->records) (void)hipFree(cloud->records);
  if (cloud->centers) (void)hipFree(cloud->centers);
  if (cloud->covs) (void)hipFree(cloud->covs);
  if (cloud->sh) (void)hipFree(cloud->sh);
  cloud->ctx->live.erase(cloud);
  cloud->magic = 0;
  delete cloud;
  return WD_OK;
}

int wd_view_prepare(const float view[16], const float world[16], const float projection[16], float width,
                    float height, const float min[3], const float max[3], float cull, wd_view *out) {
  if (!view || !projection || !min || !max || !out) return fail(WD_E_ARG, "prepare: null argument");
  if (!(cull >= 0 && std::isfinite(cull))) return fail(WD_E_ARG, "splatCull must be 0 or more, got %g", (double)cull);
  static const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const float *b = world ? world : identity;
  float mv[16];
  // mat4Multiply(out, view, world): each entry summed in f64, stored f32.
  for (int c = 0; c < 4; ++c) {
    const double b0 = b[c * 4], b1 = b[c * 4 + 1], b2 = b[c * 4 + 2], b3 = b[c * 4 + 3];
    for (int r = 0; r < 4; ++r)
      mv[c * 4 + r] = (float)(b0 * view[r] + b1 * view[4 + r] + b2 * view[8 + r] + b3 * view[12 + r]);
  }
  double near = INFINITY, far = -INFINITY;
  for (int k = 0; k < 8; ++k) {
    const double x = (k & 1) ? max[0] : min[0];
    const double y = (k & 2) ? max[1] : min[1];
    const double z = (k & 4) ? max[2] : min[2];
    const double depth = -((double)mv[2] * x + (double)mv[6] * y + (double)mv[10] * z + (double)mv[14]);
    if (depth < near) near = depth;
    if (depth > far) far = depth;
  }
  if (far <= 0) return 1;
  near = std::max(near, 0.0);
  std::memcpy(out->model_view, mv, sizeof mv);
  std::memcpy(out->projection, projection, 16 * sizeof(float));
  out->viewport[0] = width;
  out->viewport[1] = height;
  out->depth_min = (float)near;
  out->depth_scale = (float)(65535.0 / std::max(far - near, 1e-6));
  out->cull = cull;
  return WD_OK;
}

int wd_state_create(wd_context *ctx, uint64_t capacity, wd_state **out) {
  if (!ctx || !out) return fail(WD_E_ARG, "state: null argument");
  *out = nullptr;
  if (capacity == 0 || capacity > WD_MAX_SORT)
    return fail(WD_E_CAPACITY, "splats: %llu splats are more than one sort holds, %u", (unsigned long long)capacity,
                WD_MAX_SORT);
  WD_HIP(hipSetDevice(ctx->device), "state: set device");
  wd_state *s = new wd_state();
  s->ctx = ctx;
  s->capacity = capacity;
  s->tiles = wd::sort_tiles_capacity(capacity);
  const uint64_t words = (capacity + 63) / 64;
  hipError_t e = hipSuccess;
  auto alloc = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(*p, 0, bytes, ctx->stream);
  };
  alloc((void **)&s->ctl, sizeof(wd::Ctl));
  alloc((void **)&s->keys, capacity * sizeof(uint16_t));
  alloc((void **)&s->mask, words * sizeof(uint64_t));
  alloc((void **)&s->packed, capacity * (capacity > wd::kNarrowMax ? sizeof(uint64_t) : sizeof(uint32_t)));
  alloc((void **)&s->order, capacity * sizeof(uint32_t));
  alloc((void **)&s->cov, capacity * 4 * sizeof(float));
  alloc((void **)&s->tilehist, s->tiles * 512 * sizeof(uint16_t));
  alloc((void **)&s->tilehist3, s->tiles * 256 * sizeof(uint16_t));
  alloc((void **)&s->prefix_lo, s->tiles * 256 * sizeof(uint32_t));
  alloc((void **)&s->prefix_hi, s->tiles * 256 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    wd_state_destroy(s);
    return wd::hip_fail(e, "state: allocation");
  }
  ctx->adopt(s);
  *out = s;
  return WD_OK;
}

int wd_state_destroy(wd_state *s) {
  if (!s || s->magic != 0x53544154) return fail(WD_E_STATE, "state: destroyed twice");
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  for (void *p : {(void *)s->ctl, (void *)s->keys, (void *)s->mask, (void *)s->packed, (void *)s->order,
                  (void *)s->cov, (void *)s->tilehist, (void *)s->tilehist3, (void *)s->prefix_lo,
                  (void *)s->prefix_hi, (void *)s->lut, (void *)s->prefix_bin, (void *)s->ellipse})
    if (p) (void)hipFree(p);
  s->ctx->live.erase(s);
  s->magic = 0;
  delete s;
  return WD_OK;
}

int wd_state_buffers(wd_state *s, void **order, void **count, void **cov2d) {
  if (!s || s->magic != 0x53544154) return fail(WD_E_STATE, "state: destroyed");
  if (order) *order = s->order;
  if (count) *count = &s->ctl->count;
  if (cov2d) *cov2d = s->cov;
  return WD_OK;
}

// ---- particles: the particle sort (src/render/particles.js:169-240) ----

int wd_pool_create(wd_context *ctx, const void *particles, uint64_t count, wd_pool **out) {
  if (!ctx || !particles || !out) return fail(WD_E_ARG, "particles: null argument");
  *out = nullptr;
  if (count == 0 || count > WD_MAX_NARROW)
    return fail(WD_E_CAPACITY, "particles: a pool of %llu particles; this sort takes 1 to %u", (unsigned long long)count,
                WD_MAX_NARROW);
  WD_HIP(hipSetDevice(ctx->device), "particles: set device");
  wd_pool *p = new wd_pool();
  p->ctx = ctx;
  p->n = count;
  hipError_t e = hipMalloc(&p->particles, count * WD_PARTICLE_BYTES);
  if (e == hipSuccess) e = hipMemcpy(p->particles, particles, count * WD_PARTICLE_BYTES, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (p->particles) (void)hipFree(p->particles);
    delete p;
    return wd::hip_fail(e, "particles: upload");
  }
  ctx->adopt(p);
  *out = p;
  return WD_OK;
}

int wd_pool_device(const wd_pool *pool, void **device_particles) {
  if (!pool || pool->magic != 0x504f4f4c || !device_particles) return fail(WD_E_STATE, "particles: the pool was destroyed");
  *device_particles = pool->particles;
  return WD_OK;
}

int wd_pool_destroy(wd_pool *pool) {
  if (!pool || pool->magic != 0x504f4f4c) return fail(WD_E_STATE, "particles: destroyed twice");
  (void)hipSetDevice(pool->ctx->device);
  (void)hipStreamSynchronize(pool->ctx->stream);
  (void)hipFree(pool->particles);
  pool->ctx->live.erase(pool);
  pool->magic = 0;
  delete pool;
  return WD_OK;
}

static int check_tiles(const wd_state *s, uint64_t n, const char *who);

int wd_particles_sort(wd_context *ctx, const wd_pool *pool, uint64_t first, uint64_t count,
                      const wd_particle_view *view, wd_state *s) {
  if (!ctx || !view) return fail(WD_E_ARG, "particles: null argument");
  if (!pool || !ctx->live.count(pool)) return fail(WD_E_STATE, "particles: the pool was destroyed");
  if (!s || !ctx->live.count(s)) return fail(WD_E_STATE, "particles: the sort state was destroyed");
  if (pool->ctx != ctx || s->ctx != ctx) return fail(WD_E_STATE, "particles: pool, state and context are of different devices");
  if (count == 0 || first > pool->n || count > pool->n - first)
    return fail(WD_E_ARG, "particles: slots %llu to %llu of a pool of %llu", (unsigned long long)first,
                (unsigned long long)(first + count), (unsigned long long)pool->n);
  if (count > s->capacity)
    return fail(WD_E_CAPACITY, "particles: %llu particles are more than this state holds, %llu",
                (unsigned long long)count, (unsigned long long)s->capacity);
  if (first + count > 0xFFFFFFFFull) return fail(WD_E_ARG, "particles: slots past 2^32 are not addressed");
  if (const int rc = check_tiles(s, count, "particles")) return rc;
  WD_HIP(hipSetDevice(ctx->device), "particles: set device");
  if (!ctx->particle_edges) {
    // bucket b's lower edge, 2^(b 30 / 4095 - 10) (LOG_NEAR = -10, LOG_FAR = 20, 4096 buckets)
    double edges[WD_PARTICLE_BUCKETS];
    edges[0] = 0;
    for (int b = 1; b < WD_PARTICLE_BUCKETS; ++b) edges[b] = std::exp2((30.0 * b - 40950.0) / 4095.0);
    WD_HIP(hipMalloc(&ctx->particle_edges, sizeof edges), "particles: edges");
    WD_HIP(hipMemcpy(ctx->particle_edges, edges, sizeof edges, hipMemcpyHostToDevice), "particles: edges");
  }
  const double v6[6] = {view->eye[0], view->eye[1], view->eye[2], view->forward[0], view->forward[1], view->forward[2]};
  return wd::launch_particle_sort(ctx, pool->particles, first, (uint32_t)count, v6, ctx->particle_edges, s);
}

int wd_state_ellipses(wd_context *ctx, wd_state *s, int on) {
  if (!ctx || !s || !ctx->live.count(s)) return fail(WD_E_STATE, "ellipses: the sort state was destroyed");
  if (s->ctx != ctx) return fail(WD_E_STATE, "ellipses: state and context are of different devices");
  WD_HIP(hipSetDevice(ctx->device), "ellipses: set device");
  if (on && !s->ellipse) {
    // 8 floats a splat, then the reach per alpha byte: sqrt(min(log(255 (k / 255)), 4)), 0 where not above zero
    float reach[256];
    for (int k = 0; k < 256; ++k) {
      const double r = std::log(255.0 * (k / 255.0));
      reach[k] = r > 0 ? (float)std::sqrt(r < 4 ? r : 4.0) : 0.0f;
    }
    WD_HIP(hipMalloc(&s->ellipse, (size_t)s->capacity * 8 * sizeof(float) + sizeof reach), "ellipses: allocation");
    WD_HIP(hipMemcpy(s->ellipse + (size_t)s->capacity * 8, reach, sizeof reach, hipMemcpyHostToDevice),
           "ellipses: reach table");
  } else if (!on && s->ellipse) {
    WD_HIP(hipStreamSynchronize(ctx->stream), "ellipses: sync");  // no sort in flight writes them
    (void)hipFree(s->ellipse);
    s->ellipse = nullptr;
  }
  return WD_OK;
}

int wd_state_ellipse_buffer(wd_state *s, void **ellipses) {
  if (!s || s->magic != 0x53544154 || !ellipses) return fail(WD_E_STATE, "ellipses: destroyed state or null argument");
  *ellipses = s->ellipse;
  return WD_OK;
}

int wd_state_read_ellipses(wd_context *ctx, wd_state *s, float *out, uint64_t n) {
  if (!ctx || !s || s->magic != 0x53544154 || !out) return fail(WD_E_ARG, "ellipses: null argument or destroyed state");
  if (!s->ellipse) return fail(WD_E_STATE, "ellipses: this state does not take them; turn them on first");
  if (n > s->last_n)
    return fail(WD_E_ARG, "ellipses: %llu asked, the last sort had %llu splats", (unsigned long long)n,
                (unsigned long long)s->last_n);
  WD_HIP(hipStreamSynchronize(ctx->stream), "ellipses: sync");
  // two planes on the device (clip positions, then axes); 8 floats a splat on the host
  WD_HIP(hipMemcpy2D(out, 8 * sizeof(float), s->ellipse, 4 * sizeof(float), 4 * sizeof(float), (size_t)n,
                     hipMemcpyDeviceToHost), "ellipses: read");
  WD_HIP(hipMemcpy2D(out + 4, 8 * sizeof(float), s->ellipse + (size_t)s->capacity * 4, 4 * sizeof(float),
                     4 * sizeof(float), (size_t)n, hipMemcpyDeviceToHost), "ellipses: read");
  return WD_OK;
}

// The tiles a single sort of n splats takes fit the state's per-tile buffers.
static int check_tiles(const wd_state *s, uint64_t n, const char *who) {
  const uint32_t kpt = wd::sort_keys_per_thread((uint32_t)n, s->kpt);
  const uint64_t tile = 256ull * kpt, need = (n + tile - 1) / tile;
  if (need > s->tiles)
    return fail(WD_E_CAPACITY, "%s: %llu splats in tiles of %llu are %llu tiles; this state's buffers hold %llu", who,
                (unsigned long long)n, (unsigned long long)tile, (unsigned long long)need,
                (unsigned long long)s->tiles);
  return WD_OK;
}

int wd_state_keys_per_thread(wd_state *s, uint32_t kpt) {
  if (!s || s->magic != 0x53544154) return fail(WD_E_STATE, "state: destroyed");
  if (kpt != 0 && kpt != 4 && kpt != 8 && kpt != 16)
    return fail(WD_E_ARG, "state: %u keys a thread; 0 (by size), 4, 8 or 16", kpt);
  s->kpt = kpt;
  return WD_OK;
}

static int check_sort_args(wd_context *ctx, const wd_cloud *cloud, const wd_view *view, wd_state *s) {
  if (!ctx || !view) return fail(WD_E_ARG, "sort: null argument");
  if (!cloud || !ctx->live.count(cloud))
    return fail(WD_E_STATE, "addSplats: these splats were unloaded; remove the node first");
  if (!s || !ctx->live.count(s)) return fail(WD_E_STATE, "sort: the sort state was destroyed");
  if (cloud->ctx != ctx || s->ctx != ctx) return fail(WD_E_STATE, "sort: splats, state and context are of different devices");
  if (cloud->n > s->capacity)
    return fail(WD_E_CAPACITY, "sort: %llu splats are more than this state holds, %llu", (unsigned long long)cloud->n,
                (unsigned long long)s->capacity);
  if (!(view->cull >= 0 && std::isfinite(view->cull)))
    return fail(WD_E_ARG, "splatCull must be 0 or more, got %g", (double)view->cull);
  return check_tiles(s, cloud->n, "sort");
}

int wd_sort(wd_context *ctx, const wd_cloud *cloud, const wd_view *view, wd_state *s) {
  int rc = check_sort_args(ctx, cloud, view, s);
  if (rc) return rc;
  WD_HIP(hipSetDevice(ctx->device), "sort: set device");
  return wd::launch_sort(ctx, cloud, *view, s);
}

int wd_sort_noalloc(wd_context *ctx, const void *records, uint64_t n, const wd_view *view, uint32_t *order,
                    uint32_t *count, float *cov2d) {
  if (!ctx || !records || !view || !order || !count || !cov2d) return fail(WD_E_ARG, "sort: null argument");
  if (n == 0 || n > WD_NOALLOC_MAX_SPLATS)
    return fail(WD_E_CAPACITY, "sort: %llu splats; a sort without allocation takes 1 to %u", (unsigned long long)n,
                WD_NOALLOC_MAX_SPLATS);
  if (!(view->cull >= 0 && std::isfinite(view->cull)))
    return fail(WD_E_ARG, "splatCull must be 0 or more, got %g", (double)view->cull);
  if (!ctx->noalloc) return fail(WD_E_STATE, "sort: the context has no scratch for a sort without allocation");
  WD_HIP(hipSetDevice(ctx->device), "sort: set device");
  constexpr wd::NoallocLayout L = wd::noalloc_layout();
  uint8_t *base = ctx->noalloc;
  const wd::SortInput in{records, nullptr, nullptr, 0, (uint32_t)n};
  const wd::SortBuffers b{(wd::Ctl *)(base + L.ctl),        (uint16_t *)(base + L.keys),
                          (uint64_t *)(base + L.mask),      (uint32_t *)(base + L.packed),
                          order,                            cov2d,
                          (uint16_t *)(base + L.tilehist),  (uint16_t *)(base + L.tilehist3),
                          (uint32_t *)(base + L.prefix_lo), (uint32_t *)(base + L.prefix_hi),
                          nullptr,                          n,
                          count,                            false,
                          0};
  return wd::launch_sort_into(ctx, in, *view, b);
}

int wd_state_stats(wd_context *ctx, wd_state *s, wd_stats *stats) {
  if (!ctx || !s || s->magic != 0x53544154 || !stats) return fail(WD_E_ARG, "stats: null argument or destroyed state");
  wd::Ctl head;
  WD_HIP(hipMemcpyAsync(&head.count, &s->ctl->count, 6 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream),
         "stats: read");
  WD_HIP(hipStreamSynchronize(ctx->stream), "stats: sync");
  if (head.stuck) {
    const uint32_t zero = 0;
    (void)hipMemcpy(&s->ctl->stuck, &zero, sizeof zero, hipMemcpyHostToDevice);
    return fail(WD_E_STATE, "sort: a bounded wait ran out; the result is not valid");
  }
  stats->count = head.count;
  stats->recomputed = head.recomputed;
  stats->fallback = head.fallback;
  stats->overflow_bins = head.overflow_bins;
  return WD_OK;
}

int wd_state_read(wd_context *ctx, wd_state *s, uint32_t *order, uint64_t order_capacity, float *cov2d,
                  uint64_t cov2d_capacity, wd_stats *stats) {
  wd_stats st;
  int rc = wd_state_stats(ctx, s, &st);
  if (rc) return rc;
  if (order) {
    if (order_capacity < st.count)
      return fail(WD_E_CAPACITY, "read: order holds %llu, the sort has %u", (unsigned long long)order_capacity, st.count);
    WD_HIP(hipMemcpy(order, s->order, (size_t)st.count * 4, hipMemcpyDeviceToHost), "read: order");
  }
  if (cov2d) {
    const uint64_t n = std::min<uint64_t>(cov2d_capacity, s->last_n);
    WD_HIP(hipMemcpy(cov2d, s->cov, (size_t)n * 16, hipMemcpyDeviceToHost), "read: cov2d");
  }
  if (stats) *stats = st;
  return WD_OK;
}

int wd_state_read_keys(wd_context *ctx, wd_state *s, uint16_t *keys, uint64_t *mask, uint64_t n) {
  if (!ctx || !s || s->magic != 0x53544154) return fail(WD_E_ARG, "read keys: null argument or destroyed state");
  if (n > s->last_n) return fail(WD_E_ARG, "read keys: %llu asked, last sort had %llu", (unsigned long long)n,
                                 (unsigned long long)s->last_n);
  WD_HIP(hipStreamSynchronize(ctx->stream), "read keys: sync");
  if (keys) WD_HIP(hipMemcpy(keys, s->keys, (size_t)n * 2, hipMemcpyDeviceToHost), "read keys: keys");
  if (mask) WD_HIP(hipMemcpy(mask, s->mask, (size_t)((n + 63) / 64) * 8, hipMemcpyDeviceToHost), "read keys: mask");
  return WD_OK;
}

int wd_profile(wd_context *ctx, int on) {
  if (!ctx) return fail(WD_E_ARG, "profile: no context");
  int rc = ctx->prof_collect();
  if (rc) return rc;
  ctx->profiling = on != 0;
  for (int k = 0; k < 16; ++k) {
    ctx->prof_ms[k] = 0;
    ctx->prof_launches[k] = 0;
  }
  return WD_OK;
}

int wd_profile_read(wd_context *ctx, const char **names, double *ms, uint64_t *launches, int capacity, int *n) {
  if (!ctx || !n) return fail(WD_E_ARG, "profile read: null argument");
  int rc = ctx->prof_collect();
  if (rc) return rc;
  int k = 0;
  for (int i = 0; i < wd::kKernelCount && k < capacity; ++i) {
    if (names) names[k] = wd::kKernelNames[i];
    if (ms) ms[k] = ctx->prof_ms[i];
    if (launches) launches[k] = ctx->prof_launches[i];
    ++k;
  }
  *n = k;
  return WD_OK;
}

int wd_time_sorts(wd_context *ctx, const wd_cloud *cloud, const wd_view *view, wd_state *s, uint32_t reps,
                  double *ms) {
  if (!ms || reps == 0) return fail(WD_E_ARG, "time sorts: no repetitions or no output");
  int rc = check_sort_args(ctx, cloud, view, s);
  if (rc) return rc;
  WD_HIP(hipSetDevice(ctx->device), "time sorts: set device");
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < ev.size() && e == hipSuccess; ++k) e = hipEventCreate(&ev[k]);
  for (uint32_t r = 0; r < reps && e == hipSuccess && rc == WD_OK; ++r) {
    e = hipEventRecord(ev[2 * r], ctx->stream);
    if (e == hipSuccess) rc = wd::launch_sort(ctx, cloud, *view, s);
    if (e == hipSuccess && rc == WD_OK) e = hipEventRecord(ev[2 * r + 1], ctx->stream);
  }
  if (e == hipSuccess && rc == WD_OK) e = hipStreamSynchronize(ctx->stream);
  for (uint32_t r = 0; r < reps && e == hipSuccess && rc == WD_OK; ++r) {
    float t = 0;
    e = hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]);
    ms[r] = t;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc) return rc;
  if (e != hipSuccess) return wd::hip_fail(e, "time sorts");
  return WD_OK;
}

// ---- debugging entry points used by the exactness tests ----

int wd_debug_canonical(wd_context *ctx, const wd_cloud *cloud, const wd_view *view, double *out5_host) {
  if (!ctx || !cloud || !view || !out5_host) return fail(WD_E_ARG, "canonical debug: null argument");
  if (!cloud->records) return fail(WD_E_ARG, "canonical debug: takes a cloud of .splat records");
  double *dev = nullptr;
  const size_t bytes = cloud->n * 5 * sizeof(double);
  WD_HIP(hipMalloc(&dev, bytes), "canonical debug: alloc");
  int rc = wd::launch_canonical_debug(ctx, cloud, *view, dev);
  if (rc == WD_OK) {
    hipError_t e = hipMemcpyAsync(out5_host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = wd::hip_fail(e, "canonical debug: read");
  }
  (void)hipFree(dev);
  return rc;
}

int wd_debug_copy(wd_context *ctx, uint64_t bytes, int reps, double *ms_mean) {
  if (!ctx || !ms_mean || bytes < 16 || reps < 1) return fail(WD_E_ARG, "copy: bad argument");
  void *src = nullptr, *dst = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipMalloc(&src, bytes);
  if (e == hipSuccess) e = hipMalloc(&dst, bytes);
  if (e == hipSuccess) e = hipMemsetAsync(src, 1, bytes, ctx->stream);
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  int rc = e == hipSuccess ? wd::launch_copy(ctx, src, dst, bytes) : wd::hip_fail(e, "copy: setup");
  if (rc == WD_OK) {
    e = hipEventRecord(a, ctx->stream);
    for (int r = 0; r < reps && rc == WD_OK; ++r) rc = wd::launch_copy(ctx, src, dst, bytes);
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (e != hipSuccess && rc == WD_OK) rc = wd::hip_fail(e, "copy: timing");
    *ms_mean = ms / reps;
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}

int wd_debug_selftest(wd_context *ctx, const double *args, const double *expect, uint32_t count, uint32_t mismatch3[3]) {
  if (!ctx || !mismatch3 || (count && (!args || !expect))) return fail(WD_E_ARG, "selftest: null argument");
  double *da = nullptr, *de = nullptr;
  uint32_t *dm = nullptr;
  const size_t bytes = (size_t)std::max<uint32_t>(count, 1) * 2 * sizeof(double);
  hipError_t e = hipMalloc(&da, bytes);
  if (e == hipSuccess) e = hipMalloc(&de, bytes);
  if (e == hipSuccess) e = hipMalloc(&dm, 3 * sizeof(uint32_t));
  if (e == hipSuccess && count) e = hipMemcpy(da, args, (size_t)count * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess && count) e = hipMemcpy(de, expect, (size_t)count * 16, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(dm, 0, 3 * sizeof(uint32_t));
  int rc = e == hipSuccess ? wd::launch_selftest_kernel(ctx, da, de, count, dm) : wd::hip_fail(e, "selftest: setup");
  if (rc == WD_OK) {
    e = hipMemcpyAsync(mismatch3, dm, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = wd::hip_fail(e, "selftest: read");
  }
  (void)hipFree(da);
  (void)hipFree(de);
  (void)hipFree(dm);
  return rc;
}

}  // extern "C"

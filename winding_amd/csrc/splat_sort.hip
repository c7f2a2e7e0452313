// gfx950 kernels and launchers of the splat hot path beside the sort's own
// passes (sort_passes.hip, built once a tile size): the cloud check at load
// (k_bounds: first non-finite position and the centres' box), the plain copy
// that sizes HBM, the FP64 debug and self-test kernels, the no-allocation
// sort's scratch array, and the choice of tile for each sort.
#include <hip/hip_runtime.h>

#include "splat_device.h"

namespace wd {
namespace {
// ---------------------------------------------------------------------------
// Cloud check at load: first non-finite position and the centres' box.
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// A box of kBoxWords words: ordered min xyz, ordered max xyz (words 3..5, the
// maxima), the first index whose coordinate k is a zero of either sign (put()'s
// strict compares keep the first of equal values, so a zero minimum or maximum
// takes that zero's sign), the first non-finite index.
__device__ __forceinline__ bool box_is_max(uint32_t k) { return k >= 3 && k < 6; }
__device__ __forceinline__ uint32_t box_fold(uint32_t k, uint32_t a, uint32_t b) {
  return box_is_max(k) ? max(a, b) : min(a, b);
}
__device__ __forceinline__ void box_empty(uint32_t (&v)[kBoxWords]) {
#pragma unroll
  for (uint32_t k = 0; k < kBoxWords; ++k) v[k] = box_is_max(k) ? 0u : 0xFFFFFFFFu;
}
// Folds a wave's boxes into every lane's, with shuffles (no LDS atomics, whose
// same-address form the compiler turns into a lane-by-lane loop).
__device__ __forceinline__ void box_wave(uint32_t (&v)[kBoxWords]) {
#pragma unroll
  for (uint32_t m = 32; m >= 1; m >>= 1)
#pragma unroll
    for (uint32_t k = 0; k < kBoxWords; ++k) v[k] = box_fold(k, v[k], (uint32_t)__shfl_xor((int)v[k], (int)m, 64));
}

// A workgroup's share of the cloud, its box folded in the waves and then
// across them, written to its own column of partials (word k of column g at
// k * gridDim.x + g, the columns in XCD-contiguous order, so that neighbouring
// words of a line come from workgroups that share an L2). Each thread has four
// records' loads in flight.
__global__ __launch_bounds__(256) void k_bounds(const uint4 *__restrict__ rec, uint32_t n, uint32_t *partials) {
  constexpr uint32_t kIn = 4;
  __shared__ uint32_t waves[4][kBoxWords];
  const uint32_t t = threadIdx.x;
  uint32_t v[kBoxWords];
  box_empty(v);
  const uint32_t stride = gridDim.x * 256u * kIn;
  for (uint32_t i0 = blockIdx.x * 256u * kIn + t; i0 < n; i0 += stride) {
    uint4 a[kIn];
#pragma unroll
    for (uint32_t j = 0; j < kIn; ++j) {
      const uint32_t i = i0 + j * 256u;
      a[j] = i < n ? rec[2 * (size_t)i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (uint32_t j = 0; j < kIn; ++j) {
      const uint32_t i = i0 + j * 256u;
      const float p[3] = {f32_of(a[j].x), f32_of(a[j].y), f32_of(a[j].z)};
      const bool finite = isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]);
      if (i >= n) continue;
      if (!finite) {
        v[9] = min(v[9], i);
        continue;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const uint32_t o = ordered_bits(p[k]);
        v[k] = min(v[k], o);
        v[3 + k] = max(v[3 + k], o);
        if (p[k] == 0.0f) v[6 + k] = min(v[6 + k], i);
      }
    }
  }
  box_wave(v);
  if ((t & 63u) == 0)
#pragma unroll
    for (uint32_t k = 0; k < kBoxWords; ++k) waves[t >> 6][k] = v[k];
  __syncthreads();
  if (t < kBoxWords) {
    uint32_t r = waves[0][t];
#pragma unroll
    for (uint32_t q = 1; q < 4; ++q) r = box_fold(t, r, waves[q][t]);
    partials[(size_t)t * gridDim.x + xcd_item(blockIdx.x, gridDim.x)] = r;  // XCD-contiguous columns
  }
}

// The boxes of count workgroups (k_bounds, the decoders) reduced by one
// workgroup: the box to lo3 and hi3, the first zeros to zero3 and the first
// non-finite index to bad. The boxes lie word-major (word k of workgroup g at
// k * count + g), so each of the ten loads of four slots a thread keeps in
// flight is coalesced; each wave folds its lanes, and ten threads fold the
// sixteen waves.
__global__ __launch_bounds__(1024) void k_box_reduce(const uint32_t *__restrict__ partials, uint32_t count,
                                                     uint32_t *lo3, uint32_t *hi3, uint32_t *zero3, uint32_t *bad) {
  constexpr uint32_t kIn = 4;  // slots a thread has in flight
  __shared__ uint32_t waves[16][kBoxWords];
  const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
  uint32_t v[kBoxWords];
  box_empty(v);
  for (uint32_t s0 = 0; s0 < count; s0 += 1024u * kIn) {
    uint32_t x[kIn][kBoxWords];
#pragma unroll
    for (uint32_t j = 0; j < kIn; ++j) {
      const uint32_t slot = s0 + j * 1024u + t;
#pragma unroll
      for (uint32_t k = 0; k < kBoxWords; ++k)
        x[j][k] = slot < count ? partials[(size_t)k * count + slot] : (box_is_max(k) ? 0u : 0xFFFFFFFFu);
    }
#pragma unroll
    for (uint32_t j = 0; j < kIn; ++j)
#pragma unroll
      for (uint32_t k = 0; k < kBoxWords; ++k) v[k] = box_fold(k, v[k], x[j][k]);
  }
  box_wave(v);
  if (lane == 0)
#pragma unroll
    for (uint32_t k = 0; k < kBoxWords; ++k) waves[w][k] = v[k];
  __syncthreads();
  if (t < kBoxWords) {
    uint32_t r = waves[0][t];
    for (uint32_t q = 1; q < 16; ++q) r = box_fold(t, r, waves[q][t]);
    if (t < 3) lo3[t] = r;
    else if (t < 6) hi3[t - 3] = r;
    else if (t < 9) zero3[t - 6] = r;
    else *bad = r;
  }
}

// ---------------------------------------------------------------------------
// Debug: the FP64 definition for every splat, for the exactness tests.
// out5 per splat: frustum+alpha pass (0/1), alpha*area, key (as computed for
// any splat passing frustum+alpha), and spare.
template <bool STD>
__global__ __launch_bounds__(256) void k_canonical(const uint4 *__restrict__ rec, uint32_t n, ViewK v, double *out5) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint4 a = rec[2 * (size_t)i], b = rec[2 * (size_t)i + 1];
  const double x = f32_of(a.x), y = f32_of(a.y), z = f32_of(a.z);
  const double vx = fma(v.dmv[8], z, fma(v.dmv[4], y, v.dmv[0] * x)) + v.dmv[12];
  const double vy = fma(v.dmv[9], z, fma(v.dmv[5], y, v.dmv[1] * x)) + v.dmv[13];
  const double vz = fma(v.dmv[10], z, fma(v.dmv[6], y, v.dmv[2] * x)) + v.dmv[14];
  double *o = out5 + 5 * (size_t)i;
  o[0] = 0;
  o[1] = __longlong_as_double(0x7FF8000000000000ll);
  o[2] = 0;
  o[3] = 0;
  o[4] = 0;
  const double cw = -vz;
  const double cx = v.dP[0] * vx + v.dP[8] * vz;
  const double cy = v.dP[5] * vy + v.dP[9] * vz;
  const double cz = v.dP[14];
  if (vz >= 0 || cz > cw || fabs(cx) > kC13 * cw || fabs(cy) > kC13 * cw || (b.z >> 24) <= 1u) return;
  o[0] = 1;
  o[1] = canonical_value(v, a, b, vx, vy, vz);
  double t = (-vz - v.ddmin) * v.ddscale;
  t = t < 0.0 ? 0.0 : t;
  t = t > 65535.0 ? 65535.0 : t;
  o[2] = (double)(65535u - (uint32_t)t);
}

// Self-test of the primitives the exact path leans on: DPP wave scan, f64
// sqrt and division rounding (against values the host computed).
__global__ __launch_bounds__(256) void k_selftest(const double *args, const double *expect, uint32_t count,
                                                  uint32_t *mismatch3) {
  // wave scan of lane+1 over each wave: inclusive sum (l+1)(l+2)/2
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t s = wave_scan_incl(lane + 1u);
  if (s != (lane + 1u) * (lane + 2u) / 2u) atomicAdd(&mismatch3[0], 1u);
  uint32_t tot;
  __shared__ uint32_t s_w[4];
  const uint32_t e = block_scan_excl(threadIdx.x + 1u, s_w, &tot);
  if (e != threadIdx.x * (threadIdx.x + 1u) / 2u || tot != 256u * 257u / 2u) atomicAdd(&mismatch3[0], 1u);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
    const double x = args[2 * i], y = args[2 * i + 1];
    if (__double_as_longlong(sqrt(x)) != __double_as_longlong(expect[2 * i])) atomicAdd(&mismatch3[1], 1u);
    if (__double_as_longlong(x / y) != __double_as_longlong(expect[2 * i + 1])) atomicAdd(&mismatch3[2], 1u);
  }
}


}  // namespace

// A plain copy of known size: what HBM sustains for one read and one write
// stream, and the yardstick the traffic counters are scaled by.
// One 16-byte load and store a thread over a grid covering the buffer: of
// the copy shapes tried on this chip, the one that moves the most (6.2 TB/s
// for 1 GiB each way, against 4.9 for a grid-stride loop over 8 workgroups a
// CU), so it is the ceiling the path's kernels are held against.
__global__ __launch_bounds__(256) void k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
  const size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < n16) dst[i] = src[i];
}

bool is_std(const wd_view &v) {
  const float *P = v.projection, *m = v.model_view;
  return P[1] == 0 && P[2] == 0 && P[3] == 0 && P[4] == 0 && P[6] == 0 && P[7] == 0 && P[10] == 0 &&
         P[11] == -1.f && P[12] == 0 && P[13] == 0 && P[15] == 0 && m[3] == 0 && m[7] == 0 && m[11] == 0 &&
         m[15] == 1.f;
}

ViewK make_viewk(const wd_view &v) {
  ViewK k;
  for (int i = 0; i < 16; ++i) {
    k.mv[i] = v.model_view[i];
    k.P[i] = v.projection[i];
  }
  k.vpw = v.viewport[0];
  k.vph = v.viewport[1];
  k.dmin = v.depth_min;
  k.dscale = v.depth_scale;
  k.cull = v.cull;
  k.fx = (float)((double)v.projection[0] * (double)v.viewport[0] * 0.5);
  k.fy = (float)((double)v.projection[5] * (double)v.viewport[1] * 0.5);
  k.limx = (float)((double)1.3f / (double)v.projection[0]);
  k.limy = (float)((double)1.3f / (double)v.projection[5]);
  k.n0 = 0;
  k.n1 = 0;
  k.persp = v.projection[15] == 0.0f ? 1 : 0;
  for (int i = 0; i < 16; ++i) {
    k.dmv[i] = (double)v.model_view[i];
    k.dP[i] = (double)v.projection[i];
  }
  k.ddmin = (double)v.depth_min;
  k.ddscale = (double)v.depth_scale;
  k.dcull = (double)v.cull;
  // The eye in the capture's space: the translation of the model-view's
  // inverse, in FP64 from the f32 matrix and stored as f32, as
  // SplatPass.prepare writes it (src/render/splats.js:585-586) with mat4Invert's
  // cofactor expansion (src/core/math/mat4.js:186-228). A singular model-view
  // leaves it at the origin.
  const float *m = v.model_view;
  const double a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3], a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
  const double a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11], a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
  const double b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
  const double b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
  const double b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
  const double b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
  double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  k.eye[0] = k.eye[1] = k.eye[2] = 0.f;
  if (det != 0) {
    det = 1 / det;
    const double e0 = (a11 * b07 - a10 * b09 - a12 * b06) * det;
    const double e1 = (a00 * b09 - a01 * b07 + a02 * b06) * det;
    const double e2 = (a31 * b01 - a30 * b03 - a32 * b00) * det;
    if (std::isfinite(e0) && std::isfinite(e1) && std::isfinite(e2)) {
      k.eye[0] = (float)e0;
      k.eye[1] = (float)e1;
      k.eye[2] = (float)e2;
    }
  }
  return k;
}

const char *const kKernelNames[kKernelCount] = {
    "k_bounds", "k_project", "k_scan", "k_scatter_lo", "k_hist",        "k_scan2",
    "k_scatter_hi", "k_batch", "k_mlp", "k_bin_table", "k_bin_scatter", "k_bin_sort", "k_particle_keys", "k_decode"};

uint32_t bounds_partials(const wd_context *ctx) { return 4u * (uint32_t)(ctx->cus > 0 ? ctx->cus : 1); }

int launch_box_reduce(wd_context *ctx, const uint32_t *partials, uint32_t count, uint32_t *lo3, uint32_t *hi3,
                      uint32_t *zero3, uint32_t *bad) {
  hipLaunchKernelGGL(k_box_reduce, dim3(1), dim3(1024), 0, ctx->stream, partials, count, lo3, hi3, zero3, bad);
  WD_HIP(hipGetLastError(), "cloud check: box launch");
  return WD_OK;
}

int launch_bounds(wd_context *ctx, const uint8_t *records, uint64_t n, uint32_t *scratch) {
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>((n + 1023) / 1024, (uint64_t)bounds_partials(ctx)));  // 1024 records a workgroup a pass
  uint32_t *partials = scratch + 12;
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKBounds, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)records, (uint32_t)n, partials);
  WD_HIP(hipGetLastError(), "cloud check: launch");
  int rc = launch_box_reduce(ctx, partials, blocks, scratch, scratch + 3, scratch + 7, scratch + 6);
  if (rc) return rc;
  return ctx->prof_end(kKBounds, ev);
}


// The tile a sort of n splats takes, measured on MI355X (single sorts of the
// config-0 generator, median of 50, µs at 4 / 8 / 16 keys a thread): 256K
// 28.3 / 33.0 / 42.7, 1M 40.6 / 39.4 / 45.4, 2M 57.2 / 53.1 / 56.1, 4M
// 98.6 / 82.0 / 81.7, slab-8m 186 / 152 / 146. A small cloud wants more
// workgroups than one tile of 4096 a workgroup gives it.
uint32_t sort_keys_per_thread(uint32_t n, uint32_t forced) {
  if (forced) return forced;
  return n <= kSortSmallMax ? 4u : n <= kSortMidMax ? 8u : 16u;
}

int launch_sort_into(wd_context *ctx, const SortInput &in, const wd_view &view, const SortBuffers &b) {
  switch (sort_keys_per_thread(in.n, b.kpt)) {
    case 4: return passes4::launch_sort_passes(ctx, in, view, b);
    case 8: return passes8::launch_sort_passes(ctx, in, view, b);
    default: return passes16::launch_sort_passes(ctx, in, view, b);
  }
}

int launch_particle_sort(wd_context *ctx, const uint8_t *pool, uint64_t first, uint32_t n, const double *view6,
                         const double *edges, wd_state *st) {
  switch (sort_keys_per_thread(n, st->kpt)) {
    case 4: return passes4::launch_particle_passes(ctx, pool, first, n, view6, edges, st);
    case 8: return passes8::launch_particle_passes(ctx, pool, first, n, view6, edges, st);
    default: return passes16::launch_particle_passes(ctx, pool, first, n, view6, edges, st);
  }
}

int launch_sort(wd_context *ctx, const wd_cloud *cloud, const wd_view &view, wd_state *st) {
  st->last_n = cloud->n;
  const bool sig = cloud->covs != nullptr;
  const SortInput in{sig ? (const uint4 *)cloud->centers : (const uint4 *)cloud->records, (const uint2 *)cloud->covs,
                     (const uint32_t *)cloud->sh, cloud->degree, (uint32_t)cloud->n};
  const SortBuffers b{st->ctl,       st->keys,      st->mask,      st->packed, st->order, st->cov, st->tilehist,
                      st->tilehist3, st->prefix_lo, st->prefix_hi, st->ellipse, st->capacity, nullptr, true,
                      st->kpt};
  return launch_sort_into(ctx, in, view, b);
}

// ---------------------------------------------------------------------------
// The sort's intermediates for callers that cannot allocate (wd_sort_noalloc):
// one array per device, of the size kNoallocBytes gives for up to
// WD_NOALLOC_MAX_SPLATS splats, laid out as noalloc_layout cuts it. Host code
// takes its device address from hipGetSymbolAddress (noalloc_scratch), never
// from the name, which on the host only stands in for the symbol.
static __device__ __align__(256) uint8_t scratch[kNoallocBytes];

int noalloc_scratch(void **ptr) {
  *ptr = nullptr;
  WD_HIP(hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)), "open: the no-allocation sort's scratch");
  return WD_OK;
}


int launch_copy(wd_context *ctx, const void *src, void *dst, size_t bytes) {
  const size_t n16 = bytes / 16;
  if ((n16 + 255) / 256 > 0x7FFFFFFFull) return fail(WD_E_ARG, "copy: %llu bytes are more than one launch copies", (unsigned long long)bytes);
  const uint32_t blocks = (uint32_t)((n16 + 255) / 256);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)src, (uint4 *)dst, n16);
  WD_HIP(hipGetLastError(), "copy: launch");
  return WD_OK;
}

int launch_canonical_debug(wd_context *ctx, const wd_cloud *cloud, const wd_view &view, double *out5) {
  if (!is_std(view)) return fail(WD_E_ARG, "canonical debug: needs the reverse-Z perspective and an affine model-view");
  const uint32_t n = (uint32_t)cloud->n;
  hipLaunchKernelGGL(k_canonical<true>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                     (const uint4 *)cloud->records, n, make_viewk(view), out5);
  WD_HIP(hipGetLastError(), "canonical debug: launch");
  return WD_OK;
}

}  // namespace wd

// Launch used by the self-test entry point (args/expect are device arrays of 2*count doubles).
namespace wd {
int launch_selftest_kernel(wd_context *ctx, const double *args, const double *expect, uint32_t count,
                           uint32_t *mismatch3) {
  hipLaunchKernelGGL(k_selftest, dim3(64), dim3(256), 0, ctx->stream, args, expect, count, mismatch3);
  WD_HIP(hipGetLastError(), "selftest: launch");
  return WD_OK;
}
}  // namespace wd

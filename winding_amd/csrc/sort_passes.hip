// The splat sort's passes for one tile size: project + key, then a two-pass
// stable LSD scatter over the 16-bit key, with reduce-then-scan offsets. This
// file is built once for each tile the sort takes (256 threads x 4, 8 or 16
// keys, WD_KEYS_PER_THREAD); its launchers live in wd::passes4, wd::passes8
// and wd::passes16, and the sort picks one by the cloud's size.
//
//   k_project   a workgroup a chunk of one tile of splats, each 32-byte record
//               read once: the reference's makeKeys (src/render/splats.js:
//               155-186) with the footprint (:102-125) fused in. Writes the
//               16-bit key of every splat, a survivor bit per splat (a wave
//               ballot, one u64 per 64 splats), (xx, xy, yy, rgba) for every
//               splat (zeros when culled), and the chunk's two 256-bin digit
//               histograms.
//   k_scan      the reference's scanCounts (:188-212) for pass 2: per digit
//               column, each tile's exclusive prefix, and the digit totals
//               (the scatter turns those into digit bases itself).
//   k_scatter   the reference's place (:214-219), made stable: each tile
//               ranked within each wave by a ballot match on the 8-bit digit,
//               staged through LDS in digit order and written out in digit
//               runs. Run once on the low byte (splats in index order in,
//               high byte << 24 | index out) and once on the high byte (order
//               out). LSD passes are stable: key ascending, then index.
//   k_hist      the high-byte histograms of pass 2's output tiles (pass 3's
//               upsweep), and k_scan2 their column prefixes.
// No kernel waits on another workgroup of the same launch: the launches are
// the only grid-wide barriers, so no fence or look-back is ever executed.
//
// Numerics. Every decision equals the FP64 evaluation of the reference's
// formulas on the f32 inputs (the definition the CPU oracle restates):
//   - view and clip coordinates and the key are computed in FP64 here too. The
//     view's products are of two f32 values, exact in f64, so fma and mul+add
//     agree; everything else is built with -ffp-contract=off.
//   - the area cull runs in f32 with a proven bound on its error against the
//     FP64 value; only a splat whose value lies inside the bound is decided
//     again in FP64 (canonical_value, rare), so the f32 fast path never
//     changes a decision.
#include <hip/hip_runtime.h>

#include "splat_device.h"

namespace wd {
namespace {

// Pass 1: a workgroup a chunk of kTile splats. SHD > 0: a cloud with
// harmonics of that degree, each survivor's colour shaded. ELL: the draw's
// ellipse set-up is written beside cov2d.
// SIG: a cloud given as centres and Sigma (rec = the centres, sig = the
// covariances), its footprint and cull in FP64.
template <bool STD, int SHD, bool ELL, bool SIG>
__global__ __launch_bounds__(256) void k_project(const uint4 *__restrict__ rec, uint32_t n, ViewK v,
                                                 uint16_t *__restrict__ keys, uint64_t *__restrict__ mask,
                                                 float4 *__restrict__ cov, uint16_t *__restrict__ tilehist,
                                                 Ctl *__restrict__ ctl, const uint32_t *__restrict__ sh,
                                                 float4 *__restrict__ ell, uint64_t n_cap,
                                                 const uint2 *__restrict__ sig) {
  __shared__ ProjectSmem sm;
  project_chunk<STD, false, SHD, ELL, SIG>(
      sm, rec, n, v, keys, mask, cov, tilehist, &ctl->recomputed_acc, blockIdx.x, nullptr, sh, ell,
      ELL ? reinterpret_cast<const float *>(ell + 2 * (size_t)n_cap) : nullptr, sig, n_cap);
}

// The particle sort's key pass (makeKeys, src/render/particles.js:198-211) over
// a ring of n particles (32 bytes each: position, age, velocity, lifetime), a
// workgroup a chunk of kTile, in the layout the splat passes after it take:
// every particle's key and survivor bit, and the chunk's two digit histograms.
// A live particle (!(age >= lifetime)) gets 4095 - bucket, far first. The
// depth is FP64 from the f32 inputs, clamped below at 1e-6f (a NaN depth
// takes the clamp); the bucket is the number of the 4095 FP64 edges at or
// below it (edges[b] = 2^((30 b - 40950) / 4095), edges[0] = 0), so no key
// depends on how a log2 rounds.
struct ParticleK {
  double eye[3], fwd[3];
};

__global__ __launch_bounds__(256) void k_particle_keys(const uint4 *__restrict__ pool, uint32_t n, ParticleK v,
                                                       const double *__restrict__ edges, uint16_t *__restrict__ keys,
                                                       uint64_t *__restrict__ mask, uint16_t *__restrict__ tilehist) {
  __shared__ uint32_t h[512];
  const uint32_t tid = threadIdx.x, chunk = blockIdx.x;
  for (uint32_t k = tid; k < 512; k += 256) h[k] = 0;
  __syncthreads();
  const uint32_t first = chunk * (kTile / 256), end = min(first + kTile / 256, (n + 255u) >> 8);
  auto load = [&](uint32_t tile, uint4 &a, uint32_t &life) {
    const uint32_t i = (tile << 8) + tid;
    if (tile < end && i < n) {
      a = pool[2 * (size_t)i];
      life = reinterpret_cast<const uint32_t *>(pool + 2 * (size_t)i + 1)[3];
    }
  };
  auto work = [&](uint32_t tile, const uint4 &a, uint32_t life) {
    const uint32_t i = (tile << 8) + tid;
    const bool in = i < n;
    const bool live = in && !(__uint_as_float(a.w) >= __uint_as_float(life));
    uint32_t key = 0;
    if (live) {
      double d = ((double)__uint_as_float(a.x) - v.eye[0]) * v.fwd[0];
      d = d + ((double)__uint_as_float(a.y) - v.eye[1]) * v.fwd[1];
      d = d + ((double)__uint_as_float(a.z) - v.eye[2]) * v.fwd[2];
      if (!(d >= (double)1e-6f)) d = (double)1e-6f;
      // The bucket from the FP64 log2 (within a few ulps): it equals the edges'
      // answer unless t lies within 1e-9 of a whole number, far beyond both the
      // log2's error and the edges' (8 ulps of each true edge); only there are
      // the edges read. The loops end by 4095 steps in any case.
      const double t = (log2(d) + 10.0) * 4095.0 / 30.0;
      const double f = floor(t);
      int b = (int)fmin(fmax(f, 0.0), 4095.0);
      if (t - f < 1e-9 || t - f > 1.0 - 1e-9) {
        while (b < 4095 && d >= edges[b + 1]) ++b;
        while (b > 0 && d < edges[b]) --b;
      }
      key = 4095u - (uint32_t)b;
    }
    const uint64_t bal = __ballot(live);
    if ((tid & 63u) == 0 && in) mask[i >> 6] = bal;
    if (in) __builtin_nontemporal_store((uint16_t)key, &keys[i]);
    if (live) {
      atomicAdd(&h[key & 255u], 1u);
      atomicAdd(&h[256u + (key >> 8)], 1u);
    }
  };
  uint4 a0 = make_uint4(0, 0, 0, 0), a1 = a0, a2 = a0;
  uint32_t l0 = 0, l1 = 0, l2 = 0;
  load(first, a0, l0);
  load(first + 1, a1, l1);
  for (uint32_t tile = first; tile < end; tile += 3) {
    load(tile + 2, a2, l2);
    work(tile, a0, l0);
    load(tile + 3, a0, l0);
    if (tile + 1 < end) work(tile + 1, a1, l1);
    load(tile + 4, a1, l1);
    if (tile + 2 < end) work(tile + 2, a2, l2);
  }
  __syncthreads();
  for (uint32_t k = tid; k < 512; k += 256) tilehist[(size_t)chunk * 512 + k] = (uint16_t)h[k];
}

// Scan of pass 2: low-byte tile prefixes, and both digits' totals. Also moves
// the FP64 recompute count of pass 1 to where the host reads it.
__global__ __launch_bounds__(256) void k_scan(uint32_t T, const uint16_t *__restrict__ tilehist,
                                              uint32_t *__restrict__ prefix_lo, Ctl *__restrict__ ctl) {
  __shared__ uint32_t s_w[4][kScanCols];
  // XCD-contiguous column groups: the 4-column pieces of one prefix row's cache
  // lines are written from one L2
  scan_cols<kScanCols>(xcd_item(blockIdx.x, gridDim.x) * kScanCols, T, tilehist, 512, prefix_lo, 256, 256, ctl->total, s_w);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->recomputed = ctl->recomputed_acc;
    ctl->recomputed_acc = 0;
  }
}

// Pass 2 and 3: a workgroup a super-tile. count_out (the low-byte pass, when
// given) gets V as well as the control block does.
// WIDE (clouds past 2^24 splats): the elements between the passes are 64-bit.
template <bool HI, bool WIDE>
__global__ __launch_bounds__(256) void k_scatter(const uint16_t *__restrict__ keys, const uint64_t *__restrict__ mask,
                                                 const void *__restrict__ in_packed, uint32_t n,
                                                 const uint16_t *__restrict__ hist, uint32_t hstride,
                                                 const uint32_t *__restrict__ prefix, Ctl *__restrict__ ctl,
                                                 void *__restrict__ out, uint32_t out_base,
                                                 uint32_t *__restrict__ count_out) {
  using Pack = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
  __shared__ ScatterSmem<kSuper, WIDE> sm;
  const uint32_t st = xcd_item(blockIdx.x, gridDim.x);
  scatter_super<HI, kSuper, WIDE>(sm, st, keys, mask, static_cast<const Pack *>(in_packed), n, hist, hstride, prefix,
                                  HI ? ctl->total + 256 : ctl->total, ctl, out, out_base);
  if (!HI && count_out && st == 0 && threadIdx.x == 0) *count_out = ctl->count;  // this thread wrote it
}

// Upsweep of pass 3: a workgroup a tile of pass 2's output.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_hist(const void *__restrict__ packed, const Ctl *__restrict__ ctl,
                                              uint16_t *__restrict__ tilehist3) {
  using Pack = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
  __shared__ HistSmem sm;
  hist_tile<Pack, true>(sm, blockIdx.x, static_cast<const Pack *>(packed), ctl, tilehist3);
}

// Scan of pass 3: high-byte tile prefixes over the tiles of V.
__global__ __launch_bounds__(256) void k_scan2(const uint16_t *__restrict__ tilehist3, uint32_t *__restrict__ prefix_hi,
                                               Ctl *__restrict__ ctl) {
  __shared__ uint32_t s_w[4][kScanCols];
  const uint32_t T3 = (ctl->count + kTile - 1) / kTile;
  scan_cols<kScanCols>(xcd_item(blockIdx.x, gridDim.x) * kScanCols, T3, tilehist3, 256, prefix_hi, 256, 256, ctl->total3, s_w);
}


}  // namespace

// This build's launchers: wd::passes4, wd::passes8 or wd::passes16.
#define WD_NS_CAT2(a, b) a##b
#define WD_NS_CAT(a, b) WD_NS_CAT2(a, b)
namespace WD_NS_CAT(passes, WD_KEYS_PER_THREAD) {

int launch_sort_passes(wd_context *ctx, const SortInput &in, const wd_view &view, const SortBuffers &b) {
  const uint32_t n = in.n;
  const ViewK vk = make_viewk(view);
  const bool std_form = is_std(view);
  const uint32_t tiles = (n + kTile - 1) / kTile;
  const uint32_t supers = (tiles + kSuper - 1) / kSuper;
  hipEvent_t ev = nullptr;
#define WD_LAUNCH(id, what, ...)                                   \
  do {                                                             \
    if (b.profile && ctx->prof_begin(id, &ev)) return WD_E_HIP;    \
    hipLaunchKernelGGL(__VA_ARGS__);                               \
    WD_HIP(hipGetLastError(), what);                               \
    if (b.profile && ctx->prof_end(id, ev)) return WD_E_HIP;       \
  } while (0)
  // the projection's form: the fast reverse-Z one or the general one, times the
  // harmonics' degree, times whether the state takes the draw's ellipses
  using ProjectKernel = void (*)(const uint4 *, uint32_t, ViewK, uint16_t *, uint64_t *, float4 *, uint16_t *, Ctl *,
                                 const uint32_t *, float4 *, uint64_t, const uint2 *);
#define WD_FORMS(STD, SHD) \
  {{k_project<STD, SHD, false, false>, k_project<STD, SHD, false, true>}, \
   {k_project<STD, SHD, true, false>, k_project<STD, SHD, true, true>}}
  // [reverse-Z form][harmonics degree][ellipses][centres and Sigma]
  static const ProjectKernel kForms[2][4][2][2] = {
      {WD_FORMS(false, 0), WD_FORMS(false, 1), WD_FORMS(false, 2), WD_FORMS(false, 3)},
      {WD_FORMS(true, 0), WD_FORMS(true, 1), WD_FORMS(true, 2), WD_FORMS(true, 3)}};
#undef WD_FORMS
  const bool sig = in.covs != nullptr;
  const ProjectKernel project =
      kForms[std_form ? 1 : 0][in.degree <= 3 ? in.degree : 0][b.ellipse ? 1 : 0][sig ? 1 : 0];
  WD_LAUNCH(kKProject, "sort: project launch", project, dim3(tiles), dim3(256), 0, ctx->stream, (const uint4 *)in.first,
            n, vk, b.keys, b.mask, (float4 *)b.cov, b.tilehist, b.ctl, (const uint32_t *)in.sh, (float4 *)b.ellipse,
            b.capacity, (const uint2 *)in.covs);
  WD_LAUNCH(kKScan, "sort: scan launch", k_scan, dim3(512 / kScanCols), dim3(256), 0, ctx->stream, tiles, b.tilehist,
            b.prefix_lo, b.ctl);
  // past 2^24 splats the elements between the passes carry the index in 32 bits of their own
  const bool wide = n > kNarrowMax;
  WD_LAUNCH(kKScatterLo, "sort: low-byte pass launch", (wide ? k_scatter<false, true> : k_scatter<false, false>),
            dim3(supers), dim3(256), 0, ctx->stream, b.keys, b.mask, (const void *)nullptr, n, b.tilehist, 512u,
            b.prefix_lo, b.ctl, (void *)b.packed, 0u, b.count_out);
  WD_LAUNCH(kKHist, "sort: histogram launch", (wide ? k_hist<true> : k_hist<false>), dim3(tiles), dim3(256), 0,
            ctx->stream, (const void *)b.packed, b.ctl, b.tilehist3);
  WD_LAUNCH(kKScan2, "sort: second scan launch", k_scan2, dim3(256 / kScanCols), dim3(256), 0, ctx->stream,
            b.tilehist3, b.prefix_hi, b.ctl);
  WD_LAUNCH(kKScatterHi, "sort: high-byte pass launch", (wide ? k_scatter<true, true> : k_scatter<true, false>),
            dim3(supers), dim3(256), 0, ctx->stream, (const uint16_t *)nullptr, (const uint64_t *)nullptr,
            (const void *)b.packed, n, b.tilehist3, 256u, b.prefix_hi, b.ctl, (void *)b.order, 0u,
            (uint32_t *)nullptr);
#undef WD_LAUNCH
  return WD_OK;
}

// The particle sort: the key pass above, then the splat sort's five passes
// unchanged (a 12-bit key is a 16-bit key whose high byte is below 16). The
// draw order holds slots: first + the particle's index in the ring.
int launch_particle_passes(wd_context *ctx, const uint8_t *pool, uint64_t first, uint32_t n, const double *view6,
                         const double *edges, wd_state *st) {
  st->last_n = n;
  const uint32_t tiles = (n + kTile - 1) / kTile;
  const uint32_t supers = (tiles + kSuper - 1) / kSuper;
  ParticleK v;
  for (int k = 0; k < 3; ++k) {
    v.eye[k] = view6[k];
    v.fwd[k] = view6[3 + k];
  }
  hipEvent_t ev = nullptr;
#define WD_LAUNCH(id, what, ...)                  \
  do {                                            \
    if (ctx->prof_begin(id, &ev)) return WD_E_HIP; \
    hipLaunchKernelGGL(__VA_ARGS__);              \
    WD_HIP(hipGetLastError(), what);              \
    if (ctx->prof_end(id, ev)) return WD_E_HIP;   \
  } while (0)
  WD_LAUNCH(kKParticleKeys, "particles: key launch", k_particle_keys, dim3(tiles), dim3(256), 0, ctx->stream,
            (const uint4 *)(pool + 32 * first), n, v, edges, st->keys, st->mask, st->tilehist);
  WD_LAUNCH(kKScan, "particles: scan launch", k_scan, dim3(512 / kScanCols), dim3(256), 0, ctx->stream, tiles,
            st->tilehist, st->prefix_lo, st->ctl);
  WD_LAUNCH(kKScatterLo, "particles: low-byte pass launch", (k_scatter<false, false>), dim3(supers), dim3(256), 0,
            ctx->stream, st->keys, st->mask, (const void *)nullptr, n, st->tilehist, 512u, st->prefix_lo, st->ctl,
            (void *)st->packed, 0u, (uint32_t *)nullptr);
  WD_LAUNCH(kKHist, "particles: histogram launch", k_hist<false>, dim3(tiles), dim3(256), 0, ctx->stream,
            (const void *)st->packed, st->ctl, st->tilehist3);
  WD_LAUNCH(kKScan2, "particles: second scan launch", k_scan2, dim3(256 / kScanCols), dim3(256), 0, ctx->stream,
            st->tilehist3, st->prefix_hi, st->ctl);
  WD_LAUNCH(kKScatterHi, "particles: high-byte pass launch", (k_scatter<true, false>), dim3(supers), dim3(256), 0,
            ctx->stream, (const uint16_t *)nullptr, (const uint64_t *)nullptr, (const void *)st->packed, n,
            st->tilehist3, 256u, st->prefix_hi, st->ctl, (void *)st->order, (uint32_t)first, (uint32_t *)nullptr);
#undef WD_LAUNCH
  return WD_OK;
}

}  // namespace passes4 / passes8 / passes16
}  // namespace wd

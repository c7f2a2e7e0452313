// Kernels of the learned sort: the exact order, with one histogram pass and
// one scatter pass fewer, when the predictor's depth CDF is good enough.
//
// The predictor (mlp.hip) gives, per view, 256 bin widths in key space whose
// bins should each hold 1/256 of the survivors. Then, for up to kMaxViews
// views of one cloud:
//   k_bin_table      the widths as bin edges on the 16-bit key line, each on a
//                    multiple of 16 keys; a bin wider than kBinWidth keys is
//                    split, so a key's offset in its bin fits 9 bits. Writes
//                    one u16 a 16-key block (bin_key), 8 KB a view, which the
//                    next two kernels hold in LDS.
//   k_project_bins   pass 1 of every view: the same keys, survivor bits and
//                    cov2d as k_project, and the chunk's histogram over bins.
//   (k_scan_bins)    column prefixes and totals of the bin histograms.
//   k_bin_scatter    the check: every bin's total is at most the cap, or this
//                    view falls back. Then one stable pass by bin: the
//                    survivors in index order within each bin, packed as
//                    offset-in-bin << 23 | index.
//   k_bin_sort       a workgroup a bin: one stable counting pass over the
//                    9-bit offset in LDS, so key ascending then index
//                    ascending, written to the view's order.
// Against the exact path (two passes over the key's bytes) this leaves out
// the high-byte histogram pass and its scan. A view whose check fails is sorted again by the exact path in the
// same enqueue (the gated persistent batch kernel), so the output is exact
// either way: the prediction only ever decides the speed.
#include "splat_device.h"

namespace wd {
namespace {

// Lanes of this wave whose 9-bit bin equals this lane's, among `valid`.
__device__ __forceinline__ uint64_t match_bin(uint32_t d, uint64_t valid) { return match_bits<9>(d, valid); }

// Exclusive scan of one value per thread over a workgroup of NW waves.
template <int NW>
__device__ __forceinline__ uint32_t block_scan_excl_n(uint32_t v, uint32_t *s_w, uint32_t *total) {
  const uint32_t inc = wave_scan_incl(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t off = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const uint32_t t = s_w[k];
    off += (k < w) ? t : 0u;
    sum += t;
  }
  *total = sum;
  __syncthreads();
  return off + inc - v;
}

// ---------------------------------------------------------------------------
// Bin edges from the predicted widths, and the block table. Grid (16, views):
// workgroup x fills blocks [256 x, 256 x + 256); each recomputes the edges
// (256 values) rather than wait on another.
__global__ __launch_bounds__(256) void k_bin_table(const BatchJob *__restrict__ jobs, uint32_t *__restrict__ gate) {
  const BatchJob &J = jobs[blockIdx.y];
  __shared__ uint32_t B[257], F[257], s_w[4];
  const uint32_t t = threadIdx.x;
  const float p = J.probs[t];
  // widths in fixed point, so every workgroup and every run derives the same edges
  const uint32_t q = p > 0.f ? (uint32_t)fminf(p * 1048576.f + 0.5f, 1048576.f) : 0u;
  uint32_t Q;
  const uint32_t cum = block_scan_excl(q, s_w, &Q);
  const uint32_t e = Q ? (uint32_t)(((uint64_t)cum * 65536u) / Q) : t * 256u;
  B[t] = (e + kBinBlock / 2) / kBinBlock * kBinBlock;  // nearest block edge: still non-decreasing
  if (t == 0) B[256] = 65536u;
  __syncthreads();
  const uint32_t width = B[t + 1] - B[t];
  const uint32_t parts = (width + kBinWidth - 1) / kBinWidth;  // 0 for a bin that rounded away
  uint32_t nb;
  const uint32_t first = block_scan_excl(parts, s_w, &nb);
  F[t] = first;
  if (t == 0) F[256] = nb;
  __syncthreads();
  const uint32_t key = (blockIdx.x * 256u + t) * kBinBlock;
  uint32_t lo = 0, hi = 256;  // B[lo] <= key < B[hi]: lo's bin is not empty
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (B[mid] <= key) lo = mid; else hi = mid;
  }
  const uint32_t rel = key - B[lo];
  J.lut[key / kBinBlock] = (uint16_t)((((F[lo] + rel / kBinWidth) << 9) | (rel % kBinWidth)) >> 4);
  if (blockIdx.x == 0 && blockIdx.y == 0 && t == 0) *gate = 0;  // no view has fallen back yet this run
}

// The view's block table into LDS: 8 KB, two 16-byte loads a thread. The
// caller's next barrier makes it visible.
__device__ __forceinline__ void load_blocks(uint16_t *dst, const uint16_t *__restrict__ src) {
  static_assert(kBinBlocks * 2 == 512 * 16, "two 16-byte loads a thread");
  for (uint32_t k = threadIdx.x; k < 512u; k += blockDim.x)
    reinterpret_cast<uint4 *>(dst)[k] = reinterpret_cast<const uint4 *>(src)[k];
}

// ---------------------------------------------------------------------------
// Pass 1 of each view, its histogram over the view's bins. Grid (tiles,
// views): a view's chunks are separate work items, like the batch's. (One
// workgroup projecting a chunk for every view reads the records once, but at
// the sizes this serves the records stay in the MALL between views, and the
// separate items hide latency better: it measured slower.)
template <bool STD>
__global__ __launch_bounds__(256) void k_project_bins(const BatchJob *__restrict__ jobs) {
  __shared__ ProjectSmem sm;
  __shared__ __align__(16) uint16_t blk[kBinBlocks];
  const BatchJob &J = jobs[blockIdx.y];
  load_blocks(blk, J.lut);  // project_chunk starts with a barrier
  project_chunk<STD, true>(sm, (const uint4 *)J.rec, J.n, J.view, J.keys, J.mask, (float4 *)J.cov, J.tilehist,
                           &J.ctl->recomputed_acc, blockIdx.x, blk);
}

// Column prefixes and totals of the bin histograms. Grid (512 / kScanCols, views).
__global__ __launch_bounds__(256) void k_scan_bins(const BatchJob *__restrict__ jobs) {
  __shared__ uint32_t s_w[4][kScanCols];
  const BatchJob &J = jobs[blockIdx.y];
  scan_cols<kScanCols>(xcd_item(blockIdx.x, gridDim.x) * kScanCols, J.tiles, J.tilehist, 512, J.prefix_bin, 512, 512,
            J.ctl->total, s_w);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    J.ctl->recomputed = J.ctl->recomputed_acc;
    J.ctl->recomputed_acc = 0;
  }
}

// ---------------------------------------------------------------------------
// The check and the stable pass by bin. Grid (tiles, views); thread d holds
// bins d and d + 256.
struct BinScatterSmem {
  uint32_t ent[kTile];   // bin << 23 | offset in bin << 14 | index in tile
  uint32_t wh[4][512];   // per-wave running bin counts, then wave offsets
  uint32_t shift[512];   // global start - LDS start of each bin
  uint32_t lstart[512];
  uint32_t w[4];
  __align__(16) uint16_t blk[kBinBlocks];  // the view's block table
};

__global__ __launch_bounds__(256) void k_bin_scatter(const BatchJob *__restrict__ jobs, uint32_t *__restrict__ gate) {
  __shared__ BinScatterSmem sm;
  constexpr int KPT = kKeysPerThread;
  const BatchJob &J = jobs[blockIdx.y];
  const uint32_t n = J.n, t = xcd_item(blockIdx.x, gridDim.x), tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t d0 = tid, d1 = tid + 256u, tile0 = t * (uint32_t)kTile;
  const uint32_t nwords = (n + 63u) >> 6, wbase = tile0 + w * (64u * KPT);
  uint32_t raw[KPT];
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + lane;
    raw[r] = i < n ? (uint32_t)J.keys[i] : 0u;
  }
  const uint32_t wi = (wbase >> 6) + lane;
  const uint64_t vm = (lane < (uint32_t)KPT && wi < nwords) ? J.mask[wi] : 0ull;
  load_blocks(sm.blk, J.lut);  // visible after the check's barriers
  const uint32_t tot0 = J.ctl->total[d0], tot1 = J.ctl->total[d1];
  const uint32_t cnt0 = J.tilehist[(size_t)t * 512 + d0], cnt1 = J.tilehist[(size_t)t * 512 + d1];
  const uint32_t pre0 = J.prefix_bin[(size_t)t * 512 + d0], pre1 = J.prefix_bin[(size_t)t * 512 + d1];
  // The check, made by every workgroup from the same totals.
  const int over = __syncthreads_count(tot0 > J.cap) + __syncthreads_count(tot1 > J.cap);
  uint32_t T0, T1;
  const uint32_t b0 = block_scan_excl(tot0, sm.w, &T0);
  const uint32_t b1 = T0 + block_scan_excl(tot1, sm.w, &T1);
  if (t == 0 && tid == 0) {
    J.ctl->count = T0 + T1;
    J.ctl->fallback = over ? 1u : 0u;
    J.ctl->overflow_bins = (uint32_t)over;
    if (over) atomicOr(gate, 1u);
  }
  if (over) return;
  uint32_t C0, C1;
  const uint32_t l0 = block_scan_excl(cnt0, sm.w, &C0);
  const uint32_t l1 = C0 + block_scan_excl(cnt1, sm.w, &C1);
  const uint32_t nvalid = min(C0 + C1, (uint32_t)kTile);
  sm.lstart[d0] = l0;
  sm.lstart[d1] = l1;
  sm.shift[d0] = b0 + pre0 - l0;
  sm.shift[d1] = b1 + pre1 - l1;
  uint32_t bin[KPT];
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    raw[r] = bin_key(sm.blk, raw[r]);
    bin[r] = raw[r] >> 9;
  }
  for (uint32_t k = tid; k < 2048; k += 256) (&sm.wh[0][0])[k] = 0;
  __syncthreads();
  uint32_t rank[KPT];
  uint32_t okbits = 0;
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t vlo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)vm, r);
    const uint32_t vhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(vm >> 32), r);
    const uint64_t valid = ((uint64_t)vhi << 32) | vlo;
    const bool ok = (valid >> lane) & 1ull;
    const uint32_t dd = bin[r] & 511u;
    const uint64_t peers = match_bin(dd, valid);
    const uint32_t below = count_below(peers);
    const uint32_t before = sm.wh[w][dd];  // any bin is a valid index: read unconditionally
    if (ok && below == 0) sm.wh[w][dd] = before + (uint32_t)__popcll(peers);
    rank[r] = before + below;
    okbits |= (ok ? 1u : 0u) << r;
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t d = h ? d1 : d0;
    const uint32_t c0 = sm.wh[0][d], c1 = sm.wh[1][d], c2 = sm.wh[2][d];
    const uint32_t base = sm.lstart[d];
    sm.wh[0][d] = base;
    sm.wh[1][d] = base + c0;
    sm.wh[2][d] = base + c0 + c1;
    sm.wh[3][d] = base + c0 + c1 + c2;
  }
  __syncthreads();
  uint32_t pos[KPT];
#pragma unroll
  for (int r = 0; r < KPT; ++r) pos[r] = sm.wh[w][bin[r] & 511u] + rank[r];
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + lane;
    const uint32_t dd = bin[r] & 511u;
    const uint32_t off = raw[r] & 511u;
    if (((okbits >> r) & 1u) && pos[r] < (uint32_t)kTile) sm.ent[pos[r]] = (dd << 23) | (off << 14) | (i - tile0);
  }
  __syncthreads();
  for (uint32_t p0 = 0; p0 < nvalid; p0 += 8 * 256) {
    uint32_t e[8], g[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t p = p0 + (uint32_t)k * 256u + tid;
      e[k] = p < nvalid ? sm.ent[p] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) g[k] = sm.shift[e[k] >> 23] + p0 + (uint32_t)k * 256u + tid;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t p = p0 + (uint32_t)k * 256u + tid;
      if (p < nvalid && g[k] < n) J.packed[g[k]] = (((e[k] >> 14) & 511u) << 23) | (tile0 + (e[k] & 16383u));
    }
  }
}

// The counts of k_bin_sort's 16 waves per offset -> where each wave's run of
// each offset starts in the bin (offsets ascending, waves in order).
__device__ __forceinline__ void bin_offsets(uint32_t (*cnt)[512], uint32_t *s_w) {
  __syncthreads();
  const uint32_t tid = threadIdx.x;
  uint32_t s = 0;
  if (tid < 512u)
    for (int k = 0; k < 16; ++k) s += cnt[k][tid];
  uint32_t sum;
  uint32_t run = block_scan_excl_n<16>(s, s_w, &sum);
  if (tid < 512u)
    for (int k = 0; k < 16; ++k) {
      const uint32_t x = cnt[k][tid];
      cnt[k][tid] = run;
      run += x;
    }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// A bin's survivors, in index order, into the view's order by their 9-bit
// offset in the bin: one stable counting pass. Grid (512, views), 16 waves;
// wave w ranks a contiguous run of the bin, so the ranks stay stable. A bin
// of up to kHeld rounds a wave is held in registers: ranked within each wave
// (which also counts it), placed into LDS once the waves' starts are known,
// and written out in order (whole lines, streamed: the draw order is the
// path's output). A bigger one counts, then ranks straight to memory,
// reading its run twice.
constexpr uint32_t kHeld = 24;                     // rounds a wave holds: bins up to 24,576 survivors in one go
constexpr uint32_t kStage = kHeld * 64u * 16u;     // LDS slots of a held bin

__global__ __launch_bounds__(1024) void k_bin_sort(const BatchJob *__restrict__ jobs) {
  __shared__ uint32_t cnt[16][512];
  __shared__ uint32_t stage[kStage];
  __shared__ uint32_t s_w[16], s_base, s_count;
  const BatchJob &J = jobs[blockIdx.y];
  const uint32_t b = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t tot = tid < 512u ? J.ctl->total[tid] : 0u;
  if (__syncthreads_count(tot > J.cap)) return;  // this view runs the exact path
  uint32_t all;
  const uint32_t ex = block_scan_excl_n<16>(tot, s_w, &all);
  if (tid == b) {
    s_base = ex;
    s_count = tot;
  }
  for (uint32_t k = tid; k < 16u * 512u; k += 1024) (&cnt[0][0])[k] = 0;
  __syncthreads();
  const uint32_t base = s_base, c = s_count;
  if (c == 0) return;
  const uint32_t chunk = ((c + 1023u) >> 10) << 6;  // a multiple of 64 a wave
  const uint32_t w0 = min(w * chunk, c), w1 = min(w0 + chunk, c);
  if (chunk <= 64u * kHeld) {
    // every element of the wave's run loaded at once, then both passes from registers
    uint32_t val[kHeld];
#pragma unroll
    for (uint32_t r = 0; r < kHeld; ++r) {
      const uint32_t pos = w0 + r * 64u + lane;
      val[r] = pos < w1 ? J.packed[base + pos] : 0u;
    }
    // Ranks within the wave's run first (the wave's counts per offset come out of
    // the same pass), then the waves' starts per offset, then the placement.
    uint32_t rank[kHeld];
#pragma unroll
    for (uint32_t r = 0; r < kHeld; ++r) {
      rank[r] = 0;
      if (w0 + r * 64u >= w1) continue;
      const bool ok = w0 + r * 64u + lane < w1;
      const uint64_t valid = __ballot(ok);
      const uint32_t k = val[r] >> 23;
      const uint64_t peers = match_bin(k, valid);
      const uint32_t below = count_below(peers);
      const uint32_t before = cnt[w][k];
      if (ok && below == 0) cnt[w][k] = before + (uint32_t)__popcll(peers);
      rank[r] = before + below;
    }
    bin_offsets(cnt, s_w);
#pragma unroll
    for (uint32_t r = 0; r < kHeld; ++r) {
      if (w0 + r * 64u >= w1) break;
      const bool ok = w0 + r * 64u + lane < w1;
      const uint32_t rk = cnt[w][val[r] >> 23] + rank[r];
      if (ok && rk < kStage) stage[rk] = val[r] & 0x7FFFFFu;  // rk < c <= kStage by the counts
    }
    __syncthreads();
    for (uint32_t p = tid; p < c; p += 1024) __builtin_nontemporal_store(stage[p], &J.order[base + p]);
    return;
  }
  // a bin too big to hold: the same two passes, reading the run twice
  for (uint32_t p0 = w0; p0 < w1; p0 += 64) {
    const uint32_t pos = p0 + lane;
    const bool ok = pos < w1;
    const uint64_t valid = __ballot(ok);
    const uint32_t k = ok ? J.packed[base + pos] >> 23 : 0u;
    const uint32_t k0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)k);
    if (__ballot(ok && k != k0) == 0) {
      if (lane == 0) cnt[w][k0] += (uint32_t)__popcll(valid);
    } else if (ok) {
      atomicAdd(&cnt[w][k], 1u);
    }
  }
  bin_offsets(cnt, s_w);
  for (uint32_t p0 = w0; p0 < w1; p0 += 64) {
    const uint32_t pos = p0 + lane;
    const bool ok = pos < w1;
    const uint64_t valid = __ballot(ok);
    const uint32_t v = ok ? J.packed[base + pos] : 0u;
    const uint32_t k = v >> 23;
    const uint64_t peers = match_bin(k, valid);
    const uint32_t below = count_below(peers);
    uint32_t before = 0;
    if (ok) before = cnt[w][k];
    if (ok && below == 0) cnt[w][k] = before + (uint32_t)__popcll(peers);
    const uint32_t r = before + below;
    if (ok && r < c) J.order[base + r] = v & 0x7FFFFFu;
  }
}

// Training targets: the key at ranks floor(k V / 256) of a sorted state, k =
// 1..255, with edges 0 and 65536 at the ends (V = 0: even edges).
__global__ __launch_bounds__(256) void k_quantiles(const uint32_t *__restrict__ order, const uint16_t *__restrict__ keys,
                                                   const Ctl *__restrict__ ctl, uint32_t *__restrict__ edges) {
  const uint32_t k = threadIdx.x, V = ctl->count;
  uint32_t e = k * 256u;
  if (V && k) e = keys[order[(uint32_t)(((uint64_t)k * V) / 256u)]];
  edges[k] = e;
  if (k == 0) edges[256] = 65536u;
}

}  // namespace

int launch_quantiles(wd_context *ctx, const wd_state *s, uint32_t *edges_dev) {
  hipLaunchKernelGGL(k_quantiles, dim3(1), dim3(256), 0, ctx->stream, s->order, s->keys, s->ctl, edges_dev);
  WD_HIP(hipGetLastError(), "quantiles: launch");
  return WD_OK;
}

int launch_learned(wd_context *ctx, const BatchJob *jobs, uint32_t nv, uint32_t n, bool std_form, uint32_t *gate) {
  const uint32_t tiles = (n + kTile - 1) / kTile;
  hipEvent_t ev = nullptr;
#define WD_LAUNCH(id, what, ...)                  \
  do {                                            \
    if (ctx->prof_begin(id, &ev)) return WD_E_HIP; \
    hipLaunchKernelGGL(__VA_ARGS__);              \
    WD_HIP(hipGetLastError(), what);              \
    if (ctx->prof_end(id, ev)) return WD_E_HIP;   \
  } while (0)
  WD_LAUNCH(kKBinTable, "learned: bin table launch", k_bin_table, dim3(kBinBlocks / 256, nv), dim3(256), 0, ctx->stream,
            jobs, gate);
  if (std_form)
    WD_LAUNCH(kKProject, "learned: project launch", k_project_bins<true>, dim3(tiles, nv), dim3(256), 0, ctx->stream,
              jobs);
  else
    WD_LAUNCH(kKProject, "learned: project launch", k_project_bins<false>, dim3(tiles, nv), dim3(256), 0, ctx->stream,
              jobs);
  WD_LAUNCH(kKScan, "learned: scan launch", k_scan_bins, dim3(512 / kScanCols, nv), dim3(256), 0, ctx->stream, jobs);
  WD_LAUNCH(kKBinScatter, "learned: bin scatter launch", k_bin_scatter, dim3(tiles, nv), dim3(256), 0, ctx->stream,
            jobs, gate);
  WD_LAUNCH(kKBinSort, "learned: bin sort launch", k_bin_sort, dim3(512, nv), dim3(1024), 0, ctx->stream, jobs);
#undef WD_LAUNCH
  return WD_OK;
}

}  // namespace wd

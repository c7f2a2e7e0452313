// Internal declarations shared by the host API (api.cpp) and the kernels
// (splat_sort.hip, batch.hip). Not installed; include/winding_amd.h is the ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "winding_amd.h"

namespace wd {

// Records the message for wd_last_error() and returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
// A HIP failure: "<what>: <hipGetErrorString>".
int hip_fail(hipError_t e, const char *what);

#define WD_HIP(call, what)                    \
  do {                                        \
    hipError_t wd_e_ = (call);                \
    if (wd_e_ != hipSuccess) return ::wd::hip_fail(wd_e_, what); \
  } while (0)

constexpr int kThreads = 256;        // workgroup size of every kernel here
// Keys a thread ranks in a scatter tile. Set per translation unit by the
// build: 16 by default; sort_passes.hip is built at 4, 8 and 16 (launchers in
// wd::passes4, passes8, passes16) and batch.hip at 8 and 16 (wd::kpt8,
// wd::kpt16).
#ifndef WD_KEYS_PER_THREAD
#define WD_KEYS_PER_THREAD 16
#endif
constexpr int kKeysPerThread = WD_KEYS_PER_THREAD;
static_assert(kKeysPerThread == 4 || kKeysPerThread == 8 || kKeysPerThread == 16 || kKeysPerThread == 32,
              "4, 8, 16 or 32 keys a thread");
constexpr int kTile = 256 * kKeysPerThread;  // splats per projection chunk and scatter tile
// Clouds up to this many splats carry high byte << 24 | index between the two
// passes (4 bytes); bigger ones carry 8 bytes (high byte << 32 | index).
constexpr uint32_t kNarrowMax = WD_MAX_NARROW;
// A single sort of n splats takes tiles of 1024 up to kSortSmallMax, of 2048
// up to kSortMidMax, of 4096 past it (sort_keys_per_thread); batches take
// 2048 or 4096. The per-tile buffers of a state of capacity C hold the most
// tiles any of those needs.
constexpr uint32_t kSortSmallMax = 1u << 19;
constexpr uint32_t kSortMidMax = 3u << 20;
constexpr uint64_t sort_tiles_capacity(uint64_t C) {
  return ((C + 2047) / 2048) > ((C < kSortSmallMax ? C : kSortSmallMax) + 1023) / 1024
             ? (C + 2047) / 2048
             : ((C < kSortSmallMax ? C : kSortSmallMax) + 1023) / 1024;
}
constexpr uint32_t kScanCols = 4;    // histogram columns a scan workgroup takes (a single sort)
constexpr uint32_t kBatchScanCols = 8;  // the same in a batch's scans

// The no-allocation sort's scratch (wd_sort_noalloc): the intermediates of one sort of up to
// WD_NOALLOC_MAX_SPLATS splats, each region 256-byte aligned, in one array per device.
constexpr uint64_t align256(uint64_t x) { return (x + 255u) & ~uint64_t(255); }
struct NoallocLayout {
  uint64_t ctl, keys, mask, packed, tilehist, tilehist3, prefix_lo, prefix_hi, bytes;
};
constexpr uint64_t kNoallocTiles = sort_tiles_capacity(WD_NOALLOC_MAX_SPLATS);

// Polls of a bounded wait before it gives up and flags the sort as failed.
constexpr uint32_t kSpinLimit = 1u << 20;

// Device control block of one sort state. Nothing in it needs clearing
// between sorts: every field is written before it is read in each sort.
struct Ctl {
  uint32_t total[512];    // survivors per digit: low byte (0..255), high byte (256..511)
  uint32_t total3[256];   // high-byte totals again, from pass 3's own histograms
  uint32_t count;         // V, survivors of the last sort
  uint32_t recomputed;    // FP64 area decisions in the last sort
  uint32_t recomputed_acc;
  uint32_t fallback;      // learned sort: 1 when the exact path ran
  uint32_t overflow_bins;
  uint32_t stuck;         // a bounded wait ran out (a bug, reported, never a hang)
  uint32_t pad[58];
};
static_assert(sizeof(Ctl) % 256 == 0, "Ctl keeps 256-byte alignment");

constexpr NoallocLayout noalloc_layout() {
  NoallocLayout L{};
  const uint64_t N = WD_NOALLOC_MAX_SPLATS, T = kNoallocTiles;
  uint64_t at = 0;
  L.ctl = at;
  at = align256(at + sizeof(Ctl));
  L.keys = at;
  at = align256(at + N * 2);
  L.mask = at;
  at = align256(at + (N + 63) / 64 * 8);
  L.packed = at;
  at = align256(at + N * 4);
  L.tilehist = at;
  at = align256(at + T * 512 * 2);
  L.tilehist3 = at;
  at = align256(at + T * 256 * 2);
  L.prefix_lo = at;
  at = align256(at + T * 256 * 4);
  L.prefix_hi = at;
  at = align256(at + T * 256 * 4);
  L.bytes = at;
  return L;
}
constexpr uint64_t kNoallocBytes = noalloc_layout().bytes;

// The per-view constants a kernel takes by value: the uniform plus what the
// f32 fast path and its error bound precompute once per view.
struct ViewK {
  double dmv[16];     // the uniform's matrices widened once on the host, so the
  double dP[16];      // kernels read them as scalars instead of converting per lane
  double ddmin, ddscale, dcull;
  float mv[16];
  float P[16];
  float vpw, vph;
  float dmin, dscale, cull;
  float fx, fy;       // P00 * W / 2, P11 * H / 2 (f32)
  float limx, limy;   // 1.3 / P00, 1.3 / P11 (f32)
  float n0, n1;       // row norms of T's bound (see k_project)
  int persp;          // P[15] == 0
  float eye[3];       // the eye in the capture's space, for the harmonics
};

// Builds ViewK from a wd_view (host).
ViewK make_viewk(const wd_view &v);
// The reverse-Z infinite perspective and an affine model-view: the kernels'
// fast form (zero terms of the FP64 products left out).
bool is_std(const wd_view &v);

// One (cloud, view, state) of a batch, as the batch kernel reads it.
// In device code the job table's pointers are global-memory pointers: read
// out of a table, a plain pointer could point anywhere as far as the compiler
// knows, and it would access it with flat instructions, which count against
// the LDS wait counter too (every LDS wait would then also wait for them).
// The host sees plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define WD_GLOBAL __attribute__((address_space(1)))
#else
#define WD_GLOBAL
#endif

struct BatchJob {
  ViewK view;
  const WD_GLOBAL void *rec;  // uint4 records
  WD_GLOBAL uint16_t *keys;
  WD_GLOBAL uint64_t *mask;
  WD_GLOBAL float *cov;
  WD_GLOBAL uint32_t *packed;
  WD_GLOBAL uint32_t *order;
  WD_GLOBAL Ctl *ctl;
  WD_GLOBAL uint16_t *tilehist, *tilehist3;
  WD_GLOBAL uint32_t *prefix_lo, *prefix_hi;
  // the learned sort's (unused by the exact path)
  WD_GLOBAL uint16_t *lut;          // [4096] 16-key block -> (bin << 9 | the block's offset in its bin) >> 4
  WD_GLOBAL uint32_t *prefix_bin;   // [tiles][512] where a tile's run of each bin starts
  const WD_GLOBAL float *probs;     // [256] predicted bin widths (the predictor's output row)
  // a cloud given as centres and Sigma: rec is the centres, sig Sigma (null for .splat records)
  const WD_GLOBAL uint2 *sig;
  const WD_GLOBAL uint32_t *sh;     // harmonics (degree > 0), the cloud's sh words
  uint32_t degree;
  // the draw's ellipse set-up, when the state takes it: two planes of capacity float4s, then the
  // reach per alpha byte (null when off)
  WD_GLOBAL float *ellipse;
  uint64_t capacity;
  uint32_t cap;           // most survivors a bin may hold; more and the view runs the exact path
  uint32_t n, std_form;
  uint32_t tiles;    // projection chunks = scatter tiles, of kTile splats
};
// Tiles a scatter work item takes (a super-tile) on the single-sort and
// persistent paths: 1 measured faster than 2 at 1M and 8M splats (more
// workgroups in flight); the phased batch lets the autotuner pick 1 or 2.
constexpr int kSuper = 1;
constexpr uint32_t kMaxRun = 64;        // longest run of tiles a persistent batch item takes
constexpr uint32_t kPersistentRun = 8;  // the run the tuner and the learned sort's fallback use
// The learned sort: bins a view's key line is cut into (256 predicted, the
// wide ones split), keys a bin may span (its offsets fit 9 bits), views one
// projection launch takes, and the largest cloud (indices packed in 23 bits).
constexpr uint32_t kBins = 512;
constexpr uint32_t kBinWidth = 512;
constexpr uint32_t kBinBlock = 16;   // bin edges fall on multiples of 16 keys
constexpr uint32_t kBinBlocks = 65536 / kBinBlock;
constexpr uint32_t kMaxViews = 8;
constexpr uint32_t kLearnedMaxSplats = 1u << 23;
// One GEMM of the predictor on the matrix cores: C = alpha * A B, element (m, k)
// of A at A[m * sa_m + k * sa_k] (f32 or bf16), (k, n) of B likewise, then
// + bias[n], relu, zero where mask <= 0, += C, and an optional bf16 copy.
struct MlpGemm {
  int M, N, K;
  const void *A;
  int a_bf16;
  long sa_m, sa_k;
  const void *B;
  int b_bf16;
  long sb_k, sb_n;
  float *C;
  long sc_m, sc_n;
  const float *bias;
  const float *mask;
  int relu, accumulate;
  float alpha;
  uint16_t *C_bf16;
};

// Device control block of a batch: tickets, exits, the run counter, and per
// job five counters of the persistent schedule.
struct BatchCtl {
  uint32_t ticket;
  uint32_t exited;
  uint32_t runs;
  uint32_t stuck;       // a bounded wait of this run ran out; the run's waits give up early from then on
  uint32_t stuck_seen;  // some run since the host last asked was stuck (cleared by wd_batch_stuck)
  // followed by five counters a job: projection chunks done, scan column
  // groups done, low-byte super-tiles done, histogram tiles done, second-scan
  // column groups done
};
constexpr int kBatchCounters = 5;

}  // namespace wd

struct wd_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int cus = 0;
  uint64_t mem_total = 0;
  uint64_t mem_free_at_open = 0;
  bool closed = false;
  // Clouds, states and pools alive on this context, each with its serial (one
  // per object ever made here, never reused): a batch keeps the serials of its
  // jobs' objects and checks them before a run, so a destroyed object is an
  // error even when a new one was given the same address.
  std::unordered_map<const void *, uint64_t> live;
  uint64_t next_serial = 1;
  void adopt(const void *p) { live[p] = next_serial++; }
  uint64_t serial(const void *p) const {
    const auto it = live.find(p);
    return it == live.end() ? 0 : it->second;
  }
  // Per-kernel event timing.
  bool profiling = false;
  struct Pending { int kernel; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> spare;
  double prof_ms[16] = {};
  uint64_t prof_launches[16] = {};
  double *particle_edges = nullptr;  // device: the particle sort's 4096 bucket edges, made at first use
  uint8_t *noalloc = nullptr;        // device: the no-allocation sort's scratch on this device
  int prof_begin(int kernel, hipEvent_t *a);
  int prof_end(int kernel, hipEvent_t a);
  int prof_collect();
};

struct wd_pool {
  wd_context *ctx = nullptr;
  uint8_t *particles = nullptr;  // device, n * 32 bytes
  uint64_t n = 0;
  uint32_t magic = 0x504f4f4c;  // cleared by destroy
};

struct wd_cloud {
  wd_context *ctx = nullptr;
  uint8_t *records = nullptr;  // device, n * 32 bytes (null for a cloud given as centres and Sigma)
  float *centers = nullptr;    // device, n * 4 floats: x, y, z, rgba8 bits (null for .splat records)
  float *covs = nullptr;       // device, n * 6 floats: Sigma's xx xy xz yy yz zz (null for .splat records)
  uint64_t n = 0;
  float min[3] = {0, 0, 0}, max[3] = {0, 0, 0};
  uint32_t *sh = nullptr;       // device: spherical harmonics, sh_words(degree) words a splat (+1)
  uint32_t degree = 0;          // 0: none
  uint32_t magic = 0x53504c54;  // cleared by destroy
};

namespace wd {
// Coefficients a colour channel has past the constant one, by degree, and the
// 32-bit words a splat's three channels of half floats take.
constexpr uint32_t kShCoefficients[4] = {0, 3, 8, 15};
constexpr uint32_t sh_words(uint32_t degree) { return (kShCoefficients[degree] * 3 + 1) / 2; }
}  // namespace wd

struct wd_batch {
  wd_context *ctx = nullptr;
  uint32_t max_jobs = 0, n = 0;
  wd::BatchJob *jobs = nullptr;   // device
  uint32_t *starts = nullptr;     // device: 2 x (max_jobs + 1) item starts (tiles, super-tiles)
  wd::BatchCtl *ctl = nullptr;    // device
  std::vector<wd::BatchJob> host_jobs;
  std::vector<uint32_t> host_starts;
  std::vector<std::pair<const void *, uint64_t>> shape;  // clouds and states of the set jobs, with their serials
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t grid = 0, total = 0;
  int schedule = WD_SCHEDULE_PHASED;
  uint32_t super_tiles = 1;  // phased: tiles a scatter workgroup takes at once (1 or 2); persistent: tiles an item
  uint32_t kpt = 16;         // keys a thread, 8 or 16: tiles of 256 kpt splats
  std::vector<wd_job> api_jobs;                 // the jobs as set, to rebuild the tables on a change
  std::vector<std::pair<const uint32_t *, uint32_t>> sh_set;  // each job's harmonics (words, degree) as set
  std::vector<const float *> ellipse_set;                      // each job's ellipse buffer as set
  void *stage[2] = {nullptr, nullptr};          // pinned staging for the tables
  hipEvent_t staged[2] = {nullptr, nullptr};    // the copy out of stage[i] has finished
  bool staged_used[2] = {false, false};
  int flip = 0;
  size_t stage_bytes = 0;
  uint32_t magic = 0x42415443;
};

struct wd_state {
  wd_context *ctx = nullptr;
  uint64_t capacity = 0;
  uint64_t last_n = 0;
  wd::Ctl *ctl = nullptr;      // device
  uint16_t *keys = nullptr;    // n keys, every splat
  uint64_t *mask = nullptr;    // survivor bits, ceil(n/64) words
  uint32_t *packed = nullptr;  // low-digit pass output: high byte << 24 | index (8 bytes an element past kNarrowMax)
  uint32_t *order = nullptr;   // the draw order
  float *cov = nullptr;        // n x 4 floats
  float *ellipse = nullptr;    // n x 8 floats when the state takes the draw's ellipses, else null
  uint16_t *tilehist = nullptr;   // per chunk (one tile): 256 low-byte then 256 high-byte counts
  uint16_t *tilehist3 = nullptr;  // per tile of the low-byte pass's output: 256 high-byte counts
  uint32_t *prefix_lo = nullptr, *prefix_hi = nullptr;  // [tiles][256]: where a tile's digit run starts
  uint16_t *lut = nullptr;                             // learned sort: 16-key block -> bin and offset
  uint32_t *prefix_bin = nullptr;                      // learned sort: [tiles][512]
  uint64_t tiles = 0;  // per-tile buffers hold this many tiles (sort_tiles_capacity)
  uint32_t kpt = 0;    // keys a thread its single sorts take; 0: by their size
  uint32_t magic = 0x53544154;
};

namespace wd {

enum KernelId {
  kKBounds = 0, kKProject, kKScan, kKScatterLo, kKHist, kKScan2, kKScatterHi, kKBatch, kKMlp, kKBinTable, kKBinScatter,
  kKBinSort, kKParticleKeys, kKDecode, kKernelCount
};
extern const char *const kKernelNames[kKernelCount];

// Kernel launchers (splat_sort.hip). Enqueue on ctx->stream.
// The cloud check's box is reduced without global atomics: each workgroup
// writes kBoxWords words (ordered min xyz, ordered max xyz, first zero per
// axis, first non-finite index) to a column of its own (word k of workgroup g
// of a launch of G at k * G + g), and one workgroup reduces the columns
// (launch_box_reduce, with count = G).
constexpr uint32_t kBoxWords = 10;
// scratch: 12 words of results (ordered min xyz, ordered max xyz, first bad
// index, first zero per axis), then bounds_partials(ctx) columns.
uint32_t bounds_partials(const wd_context *ctx);
int launch_bounds(wd_context *ctx, const uint8_t *records, uint64_t n, uint32_t *scratch);
int launch_box_reduce(wd_context *ctx, const uint32_t *partials, uint32_t count, uint32_t *lo3, uint32_t *hi3,
                      uint32_t *zero3, uint32_t *bad);
// The box put() reports, from the kernels' reduction: ordered bits of the least and greatest value
// per axis, made floats; a zero end takes the sign of the axis's first zero (zero_index, its value
// read by value_at), since put()'s strict compares keep the first of equal values.
void box_from_ordered(const uint32_t lo[3], const uint32_t hi[3], const uint32_t zero_index[3],
                      float (*value_at)(const void *, uint32_t, int), const void *arg, float mn[3], float mx[3]);
int launch_sort(wd_context *ctx, const wd_cloud *cloud, const wd_view &view, wd_state *st);
// What one sort reads: .splat records (first) or centres (first) with Sigma (covs), the
// harmonics (sh words, degree 0 when none), n splats.
struct SortInput {
  const void *first;
  const void *covs;
  const void *sh;
  uint32_t degree;
  uint32_t n;
};
// Where one sort writes: its intermediates and outputs. count_out, when given, gets V too;
// profile times each launch with HIP events (the no-allocation path never does).
struct SortBuffers {
  Ctl *ctl;
  uint16_t *keys;
  uint64_t *mask;
  uint32_t *packed, *order;
  float *cov;
  uint16_t *tilehist, *tilehist3;
  uint32_t *prefix_lo, *prefix_hi;
  float *ellipse;
  uint64_t capacity;
  uint32_t *count_out;
  bool profile;
  uint32_t kpt;  // keys a thread the sort takes; 0: by its size
};
// The keys a thread (4, 8 or 16) a single sort of n splats takes; forced: 0 or the caller's choice.
uint32_t sort_keys_per_thread(uint32_t n, uint32_t forced);
// The sort's passes at one tile size (sort_passes.hip).
#define WD_SORT_PASSES                                                                                          \
  int launch_sort_passes(wd_context *ctx, const SortInput &in, const wd_view &view, const SortBuffers &b);      \
  int launch_particle_passes(wd_context *ctx, const uint8_t *pool, uint64_t first, uint32_t n, const double *view6, \
                             const double *edges, wd_state *st);
namespace passes4 {
WD_SORT_PASSES
}
namespace passes8 {
WD_SORT_PASSES
}
namespace passes16 {
WD_SORT_PASSES
}
#undef WD_SORT_PASSES
int launch_sort_into(wd_context *ctx, const SortInput &in, const wd_view &view, const SortBuffers &b);
// The no-allocation sort's scratch on the current device (hipGetSymbolAddress of the array).
int noalloc_scratch(void **ptr);
// ---- capture decoders (decode.hip, decode_host.cpp) ----
// Where a decoder writes: the cloud's centres (x, y, z, rgba bits), Sigma (six
// floats), harmonics (sh_words words a splat, zeroed before), the least index
// of a non-finite position (UINT32_MAX when none), the box as ordered bits
// (min xyz, then max xyz) and each axis's first zero.
struct DecodeOut {
  float *centers;
  float *covs;
  uint32_t *sh;
  uint32_t sh_words;
  uint32_t *bad;
  uint32_t *box;
  uint32_t *zero;  // per axis, the least index whose coordinate is a zero of either sign
  uint32_t *partials;  // kBoxWords a workgroup, word-major: its own box, first zeros and first bad index
};
// .ply property types, as PLY_TYPES names them (src/scene/splats.js:392-397).
enum PlyType : uint8_t { kPlyI8, kPlyU8, kPlyI16, kPlyU16, kPlyI32, kPlyU32, kPlyF32, kPlyF64 };
// x y z f_dc_0..2 opacity scale_0..2 rot_0..3 (PLY_NEEDS), then f_rest_0.. (up to 45).
constexpr int kPlyNeeds = 14;
struct PlyLayout {
  uint32_t stride;
  uint32_t rest;
  uint32_t off[kPlyNeeds + 45];
  uint8_t type[kPlyNeeds + 45];
};
// A gunzipped .spz's sections (byte offsets) and header (src/scene/splats.js:230-256).
struct SpzLayout {
  uint32_t version, rotation_bytes, coefficients, file_coefficients;
  uint64_t positions, alphas, colours, scales, rotations, sh;
  double fixed;  // 1 / (1 << fraction), with JavaScript's 32-bit shift
};
// A .sog's decoded images (device copies, rgba8, a splat a pixel) and tables
// made on the host from its codebooks: the colour byte, exp of the scale and
// the harmonics' half bits, per codebook index (src/scene/splats.js:136-203).
struct SogLayout {
  uint32_t count, per, sh_count, centroids_width;
  uint64_t centroids_bytes;
  double mins[3], maxs[3];
  const uint8_t *means_l, *means_u, *quats, *scales, *sh0, *labels, *centroids;
  const uint8_t *colour_byte;
  const double *scale_exp;
  const uint16_t *sh_half;
};
// Each writes the workgroups it launched (at most n / 64 rounded up), whose
// boxes launch_box_reduce then folds.
int launch_decode_ply(wd_context *ctx, const uint8_t *body, uint32_t n, const PlyLayout &L, const DecodeOut &o,
                      uint32_t *groups);
int launch_decode_spz(wd_context *ctx, const uint8_t *data, uint32_t n, const SpzLayout &L, const DecodeOut &o,
                      uint32_t *groups);
int launch_decode_sog(wd_context *ctx, const SogLayout &L, const DecodeOut &o, uint32_t *groups);

// The particle sort over slots [first, first + n) of a pool (view6: eye, forward in FP64;
// edges: the 4096 FP64 bucket edges on the device).
int launch_particle_sort(wd_context *ctx, const uint8_t *pool, uint64_t first, uint32_t n, const double *view6,
                         const double *edges, wd_state *st);
// The batch kernels (batch.hip), one set a keys-a-thread build. launch_batch:
// one run of a batch on its schedule. launch_persistent: the persistent
// kernel over a job table, skipped unless *gate != 0 when gate is given.
// batch_grid: the persistent kernel's grid for `items` items.
#define WD_BATCH_LAUNCHERS                                                                                     \
  int launch_batch(wd_context *ctx, wd_batch *b);                                                              \
  int launch_persistent(wd_context *ctx, const BatchJob *jobs, const uint32_t *starts, uint32_t nj,            \
                        uint32_t max_jobs, BatchCtl *bctl, uint32_t grid, const uint32_t *gate, uint32_t run); \
  int batch_grid(wd_context *ctx, uint32_t items, uint32_t *grid);
namespace kpt8 {
WD_BATCH_LAUNCHERS
}
namespace kpt16 {
WD_BATCH_LAUNCHERS
}
#undef WD_BATCH_LAUNCHERS
// The 257 equal-mass edges of a sorted state's keys (learned.hip).
int launch_quantiles(wd_context *ctx, const wd_state *s, uint32_t *edges_dev);
// The learned sort's kernels (learned.hip) over nv views of one n-splat cloud.
int launch_learned(wd_context *ctx, const BatchJob *jobs, uint32_t nv, uint32_t n, bool std_form, uint32_t *gate);
// The predictor's forward pass over `rows` poses already in device memory;
// *probs gets the device address of the rows x 256 softmax (mlp_host.cpp).
int mlp_infer(wd_context *ctx, wd_mlp *m, const float *poses_dev, uint32_t rows, const float **probs);
int launch_copy(wd_context *ctx, const void *src, void *dst, size_t bytes);
// Predictor kernels (mlp.hip).
int launch_gemm(wd_context *ctx, const MlpGemm &g);
// The predictor's forward for `rows` poses in one launch: P[rows][256].
int launch_mlp_infer(wd_context *ctx, const float *X, const uint16_t *wb, const float *w, int rows, float *P);
int launch_softmax_xent(wd_context *ctx, const float *z, const float *target, int rows, float scale, float *p, float *dz,
                        float *loss);
int launch_colsum(wd_context *ctx, const float *x, int rows, int cols, float *out);
int launch_adam(wd_context *ctx, float *w, uint16_t *wb, const float *g, float *m, float *v, int n, float lr, float b1,
                float b2, int step, float eps);
int launch_to_bf16(wd_context *ctx, const float *w, uint16_t *wb, int n);
int launch_canonical_debug(wd_context *ctx, const wd_cloud *cloud, const wd_view &view, double *out5);

}  // namespace wd

/*
 * winding_amd.h -- C ABI of the MI355X splat projection + stable depth sort.
 *
 * One library, libwinding_amd.so, holds the gfx950 kernels and everything
 * that owns device memory. The JavaScript module (winding_amd/js) reaches it
 * through a thin N-API addon (winding_amd.node); Python reaches it through
 * ctypes. Every entry point takes plain pointers and sizes, returns WD_OK (0)
 * or a negative WD_E* status, and leaves a message for wd_last_error().
 *
 * Reference interfaces replaced (paths relative to the Winding v1.9.0 tree):
 *   new Splats(rhi, data) / destroy()        src/render/splats.js:389-416
 *   storageCapacity(rhi, bytesPerItem)       src/rhi/buffer.js:32-36
 *   SplatPass._stateFor / destroyState       src/render/splats.js:629-674,735-737
 *   SplatPass.prepare (per-cloud uniform)    src/render/splats.js:546-607
 *   _encodeSort: makeKeys, scanCounts, place src/render/splats.js:155-219,676-697
 *   Cloud uniform layout                     src/render/splats.js:80-92
 *
 * Conventions (all from the reference):
 *   - matrices are column-major, m[col*4 + row]   (src/core/math/mat4.js:1-12)
 *   - a .splat record is 32 bytes: pos f32x3, scale f32x3, rgba u8x4,
 *     quaternion u8x4 stored w,x,y,z as (q*128+128)  (src/scene/splats.js:373-388)
 *   - key = 65535 - floor(clamp((-view.z - depthMin) * depthScale, 0, 65535));
 *     ascending key is back to front            (src/render/splats.js:179-181)
 *
 * What differs, on purpose: the order within a key is stable (splat index
 * ascending), where the reference's comes out of atomics
 * (src/render/splats.js:22-25), and every cull and key decision equals the
 * FP64 evaluation of the reference's formulas on the f32 inputs.
 */
#ifndef WINDING_AMD_H
#define WINDING_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WD_OK 0
#define WD_E_ARG -1      /* a bad argument: null, wrong size, out of range */
#define WD_E_HIP -2      /* the HIP runtime failed */
#define WD_E_CAPACITY -3 /* more than a buffer or a sort holds */
#define WD_E_DATA -4     /* input refused by content (a NaN position, ...) */
#define WD_E_STATE -5    /* wrong object state: destroyed, mismatched, ... */
#define WD_E_FALLBACK -6 /* not an error: a learned sort ran the exact path */

/* Bytes in one .splat record. src/scene/splats.js:41 */
#define WD_RECORD_BYTES 32
/* Depth buckets; keys are 16 bits. src/render/splats.js:64 */
#define WD_BUCKETS 65536
/* Default cull, opacity x pixels. src/render/splats.js:78 */
#define WD_SPLAT_CULL 0.5f
/*
 * Most splats one cloud and one sort take. Up to 2^24 the stable scatter
 * carries high byte << 24 | index between its passes (4 bytes a survivor);
 * past that, 8 bytes. A batch job and a particle pool take up to 2^24.
 */
#define WD_MAX_SORT (1u << 28)
/* Most splats a batch job takes, and most particles a pool holds. */
#define WD_MAX_NARROW (1u << 24)
/* Coarse depth bins the learned predictor works in: the key's high byte. */
#define WD_COARSE_BINS 256
/* Pose features the predictor reads: the model-view's top three rows. */
#define WD_POSE_FLOATS 12
/* Most splats wd_sort_noalloc takes (its intermediates live in a fixed array of about 16 MiB). */
#define WD_NOALLOC_MAX_SPLATS (1u << 21)
/* Bytes in one particle: position, age, velocity, lifetime. src/render/particles.js:51-56 */
#define WD_PARTICLE_BYTES 32
/* The particle sort's log-depth buckets. src/render/particles.js:45 */
#define WD_PARTICLE_BUCKETS 4096

typedef struct wd_context wd_context;
typedef struct wd_cloud wd_cloud;
typedef struct wd_state wd_state;
typedef struct wd_batch wd_batch;
typedef struct wd_mlp wd_mlp;
typedef struct wd_pool wd_pool;

/*
 * The particle sort's view: the reference's 32-byte View uniform
 * (src/render/particles.js:171-174, written at :688-693): the camera's
 * position, and its forward axis, -(view[2], view[6], view[10]).
 */
typedef struct wd_particle_view {
  float eye[4];     /* xyz used */
  float forward[4]; /* xyz used */
} wd_particle_view;

/*
 * What one sort depends on: the first 37 floats of the reference's Cloud
 * uniform (src/render/splats.js:80-92, written at :574-587), all f32 as there.
 */
typedef struct wd_view {
  float model_view[16]; /* view * world, column-major */
  float projection[16]; /* column-major; [15] == 0 means perspective */
  float viewport[2];    /* width, height in pixels */
  float depth_min;      /* nearest view depth the cloud's box reaches, >= 0 */
  float depth_scale;    /* buckets per unit depth: 65535 / max(far - near, 1e-6) */
  float cull;           /* opacity x pixels a splat must add; 0 keeps all */
} wd_view;

/* Counters a sort leaves, read back with wd_state_stats. */
typedef struct wd_stats {
  uint32_t count;         /* survivors V: the reference's DrawArgs.instanceCount */
  uint32_t recomputed;    /* splats whose area cull was decided in FP64 */
  uint32_t fallback;      /* learned sorts that ran the exact path (0 or 1) */
  uint32_t overflow_bins; /* coarse bins whose prediction was too small */
} wd_stats;

/* ---- errors, devices ------------------------------------------------- */

/* The last failure's message on this thread ("" when none). */
const char *wd_last_error(void);

/* Opens device `device` (HIP ordinal) with one stream of default priority. */
int wd_open(int device, wd_context **out);
int wd_close(wd_context *ctx);
/* Waits for everything enqueued on the context's stream. */
int wd_sync(wd_context *ctx);
/* The hipStream_t the context enqueues on, for callers that time or chain. */
int wd_stream(wd_context *ctx, void **stream);

/*
 * How many items of `bytes_per_item` one allocation may hold on this device.
 * Replaces storageCapacity (src/rhi/buffer.js:32-36): the limit is the
 * device's free memory at open, capped at WD_MAX_SORT for sortable items.
 */
int wd_capacity(wd_context *ctx, uint64_t bytes_per_item, uint64_t *out);

/* ---- clouds: new Splats(rhi, data), src/render/splats.js:389-416 ------ */

/*
 * Uploads `bytes` of .splat records (host memory) and checks them on the GPU:
 * bytes % 32 == 0 and > 0 (src/scene/splats.js:374-376), every position
 * finite (src/scene/splats.js:336-338). Fills min/max of the centres
 * (src/scene/splats.js:343-348). Refuses more than WD_MAX_SORT splats by name.
 */
int wd_cloud_create(wd_context *ctx, const void *records, uint64_t bytes,
                    wd_cloud **out, float min_out[3], float max_out[3]);
/* As wd_cloud_create, from records already in device memory (copied). */
int wd_cloud_create_device(wd_context *ctx, const void *device_records,
                           uint64_t bytes, wd_cloud **out, float min_out[3],
                           float max_out[3]);
int wd_cloud_info(const wd_cloud *cloud, uint64_t *count, float min_out[3],
                  float max_out[3]);
/* Device address of the records (32 B each), for callers that read them. */
int wd_cloud_records(const wd_cloud *cloud, const void **device_records);
int wd_cloud_destroy(wd_cloud *cloud);

/*
 * Spherical harmonics for a cloud, degree 1 to 3 (0 removes them), in the
 * reference's `sh` layout (src/scene/splats.js:306-332): for each splat,
 * ceil(3 n / 2) 32-bit words, n = 3, 8 or 15 coefficients past the constant
 * one, coefficient by coefficient, red, green and blue as IEEE half floats,
 * two a word, the lower half first. `words` holds at least count of those.
 * From then on wd_sort writes, as each survivor's colour, what the reference's
 * shade pass computes (src/render/splats.js:315-381): the record's colour plus
 * the harmonics along the direction from the eye (the inverse model-view's
 * translation, as SplatPass.prepare sets it, :585-586), clamped to [0, 1] and
 * packed as unorm bytes, alpha kept; in f32, within one unit of the byte of
 * the FP64 evaluation. Batches shade it too under the reverse-Z perspective
 * (another projection is refused there); learned runs refuse it.
 */
int wd_cloud_sh(wd_context *ctx, wd_cloud *cloud, const uint32_t *words,
                uint64_t n_words, uint32_t degree);

/*
 * A cloud given as the reference's Splats data (new Splats(rhi, data),
 * src/render/splats.js:389-416, from parseSplats' output, src/scene/splats.js:
 * 305-370): `count` centres of 4 floats (x, y, z, and the rgba8 colour's bits
 * in the fourth) and 6 covariance floats each (Sigma's xx xy xz yy yz zz), host
 * memory. Refuses a non-finite position by index, as put() does
 * (src/scene/splats.js:336-338), and fills the box. Its sorts decide the
 * footprint and the cull in FP64 from the f32 Sigma. Batches take such a
 * cloud under the reverse-Z perspective; learned runs refuse it;
 * wd_cloud_records gives null for it.
 */
int wd_cloud_create_splats(wd_context *ctx, const float *centers,
                           const float *covariances, uint64_t count,
                           wd_cloud **out, float min_out[3], float max_out[3]);
/*
 * Copies a cloud's centres (4 floats a splat), covariances (6) and harmonics
 * (sh words, the reference's layout) to the host, and its degree; any pointer
 * may be null. A cloud of .splat records has no centres to read.
 */
int wd_cloud_read(wd_context *ctx, const wd_cloud *cloud, float *centers,
                  float *covariances, uint32_t *sh, uint32_t *degree);

/* ---- capture decoders on the GPU: src/scene/splats.js:58-303,390-468 -- */

/*
 * Decodes a capture into a cloud of centres and covariances, with its
 * harmonics attached (degree from the file): the header on the host, every
 * splat on the GPU, with put()'s formulas in FP64 as the reference runs them
 * in JavaScript numbers. The reference's refusals, with its messages.
 *   wd_decode_ply  a binary_little_endian .ply as training writes it
 *                  (parsePly, src/scene/splats.js:398-468)
 *   wd_decode_spz  a .spz of version 1 to 3, gzipped or already gunzipped
 *                  (readSplats + parseSpz, src/scene/splats.js:84-92,225-303)
 *   wd_decode_sog  a .sog of version 2 whose zip, meta.json and WebP images the
 *                  caller has unpacked (readSog, src/scene/splats.js:128-203):
 *                  each image as rgba8 pixels, as decodeImage hands them over
 */
int wd_decode_ply(wd_context *ctx, const void *bytes, uint64_t len,
                  wd_cloud **out, float min_out[3], float max_out[3]);
int wd_decode_spz(wd_context *ctx, const void *bytes, uint64_t len,
                  wd_cloud **out, float min_out[3], float max_out[3]);

typedef struct wd_sog_image {
  const uint8_t *rgba; /* width * height * 4 bytes, host memory */
  uint32_t width, height;
} wd_sog_image;

typedef struct wd_sog {
  uint32_t count;    /* meta.count */
  uint32_t bands;    /* meta.shN.bands: the harmonics' degree, 0 when none */
  uint32_t sh_count; /* meta.shN.count: palette entries; a label at or past it has no harmonics */
  double mins[3], maxs[3];             /* meta.means.mins, maxs */
  const double *scale_codebook;        /* meta.scales.codebook */
  uint32_t scale_codebook_size;
  const double *colour_codebook;       /* meta.sh0.codebook */
  uint32_t colour_codebook_size;
  const double *sh_codebook;           /* meta.shN.codebook (bands > 0) */
  uint32_t sh_codebook_size;
  wd_sog_image means_l, means_u, quats, scales, sh0;
  wd_sog_image centroids, labels;      /* meta.shN.files (bands > 0) */
} wd_sog;

int wd_decode_sog(wd_context *ctx, const wd_sog *sog, wd_cloud **out,
                  float min_out[3], float max_out[3]);

/* ---- the per-view uniform: SplatPass.prepare, src/render/splats.js:559-587 */

/*
 * modelView = view * world (mat4Multiply, src/core/math/mat4.js:96-126, in
 * f64 rounded to f32); depth range over the box's eight corners; depthMin =
 * max(near, 0); depthScale = 65535 / max(far - near, 1e-6). Returns 1 (and
 * leaves *out untouched) when far <= 0: the cloud is wholly behind the
 * camera and the reference skips it (src/render/splats.js:572).
 * world may be null for identity.
 */
int wd_view_prepare(const float view[16], const float world[16],
                    const float projection[16], float width, float height,
                    const float min[3], const float max[3], float cull,
                    wd_view *out);

/* ---- sort state: SplatPass._stateFor, src/render/splats.js:629-674 ---- */

/* Buffers for sorting clouds of up to `capacity` splats. */
int wd_state_create(wd_context *ctx, uint64_t capacity, wd_state **out);
int wd_state_destroy(wd_state *state);
/*
 * Device addresses of the outputs:
 *   order  u32 x capacity : splat indices, back to front, count valid
 *   count  u32            : survivors V
 *   cov2d  f32x4 x capacity indexed by splat: (xx, xy, yy) in px^2 before the
 *          draw's +0.3 (src/render/splats.js:95-96), and the rgba8 bits in .w
 *          (src/scene/splats.js:342); written for survivors only.
 */
int wd_state_buffers(wd_state *state, void **order, void **count, void **cov2d);

/*
 * Projects, culls, keys and stably sorts `cloud` for `view` into `state`
 * (makeKeys -> scanCounts -> place, src/render/splats.js:155-219). Enqueued
 * on the context's stream; returns before the GPU finishes.
 */
int wd_sort(wd_context *ctx, const wd_cloud *cloud, const wd_view *view,
            wd_state *state);

/*
 * wd_sort for callers that cannot allocate device memory (code under HIP graph
 * capture, a serving loop that must not allocate): no counterpart in the
 * reference. It allocates nothing and frees nothing: the sort's intermediates
 * live in one fixed array per device (a few MiB, held until the process ends),
 * and the caller passes device buffers for the input and the outputs:
 *   records  n .splat records (32 B each), taken as they are (the finite check
 *            of wd_cloud_create is not made here)
 *   order    u32 x n: splat indices, back to front; the first *count are valid
 *   count    one u32: survivors V
 *   cov2d    f32x4 x n, as wd_state_buffers' cov2d (zeros for culled splats)
 * 1 <= n <= WD_NOALLOC_MAX_SPLATS, or WD_E_CAPACITY. Enqueued on the context's
 * stream. Every call on one device shares that array and leaves its contents
 * behind: calls that overlap on one device, from any context, race, so issue
 * them on one stream or wait between them. Results equal wd_sort's bit for bit.
 */
int wd_sort_noalloc(wd_context *ctx, const void *records, uint64_t n, const wd_view *view, uint32_t *order,
                    uint32_t *count, float *cov2d);

/* Waits, then copies the results to host memory. Any pointer may be null. */
int wd_state_read(wd_context *ctx, wd_state *state, uint32_t *order,
                  uint64_t order_capacity, float *cov2d, uint64_t cov2d_capacity,
                  wd_stats *stats);
int wd_state_stats(wd_context *ctx, wd_state *state, wd_stats *stats);

/*
 * Copies the 16-bit keys of every splat and the survivor bitmask (bit i of
 * word i/64) of the last wd_sort: the parity tests compare them per splat.
 */
int wd_state_read_keys(wd_context *ctx, wd_state *state, uint16_t *keys,
                       uint64_t *mask, uint64_t n);

/*
 * The draw's per-splat ellipse set-up, made by the projection beside cov2d.
 * Replaces the per-instance work of the splat vertex shader `vs`
 * (src/render/splats.js:263-291): with it a rasteriser reads each splat once
 * instead of projecting it again. wd_state_ellipses(ctx, state, 1) allocates
 * two planes of capacity float4s and from then on wd_sort writes, for splat i:
 *   plane 0, [i]  its clip position (x, y, z, w), the FP64 value rounded to f32;
 *   plane 1, [i]  (along e, across e) in pixels: the eigen-axes of the footprint plus
 *          0.3 px^2 each way, each sqrt(2 lambda) long and capped at 1024,
 *          along the larger first, across = along turned a quarter clockwise;
 *          e = sqrt(min(log(255 alpha), 4)). Corner q of the quad, q in
 *          {-1, 1}^2, lies at clip.xy + (q.x along e + q.y across e) 2 /
 *          viewport clip.w; the Gaussian's offset there is q e.
 * Both are zero for a splat the shader would not draw (culled, or a minor
 * axis or reach not above zero). 0 frees them. f32, within 1e-5 of the
 * largest term of the FP64 evaluation. Batches write them too, under the
 * reverse-Z perspective only (another projection is refused there); learned
 * runs refuse such a state. wd_state_ellipse_buffer gives the device address of plane 0
 * (plane 1 follows at capacity float4s; null when off); wd_state_read_ellipses
 * copies the first n splats' 8 floats (clip position, then axes) to the host.
 */
int wd_state_ellipses(wd_context *ctx, wd_state *state, int on);
/*
 * The tile this state's single sorts (wd_sort, wd_particles_sort) take:
 * keys_per_thread 4, 8 or 16 (tiles of 1024, 2048 or 4096 splats, 256
 * threads a workgroup), or 0, the default: by the cloud's size, 4 up to
 * 512Ki splats, 8 up to 3Mi, 16 past that (the fastest of the three at each
 * size on MI355X). A sort whose tiles are more than the state's per-tile
 * buffers hold (a forced 4 past 512Ki splats) is refused with WD_E_CAPACITY.
 * No counterpart in the reference, whose sort has no tiles.
 */
int wd_state_keys_per_thread(wd_state *state, uint32_t keys_per_thread);
int wd_state_ellipse_buffer(wd_state *state, void **ellipses);
int wd_state_read_ellipses(wd_context *ctx, wd_state *state, float *out,
                           uint64_t n);

/* ---- particles: the particle sort, src/render/particles.js:169-240 ---- */

/*
 * A pool of `count` particles (WD_PARTICLE_BYTES each, host memory: position
 * xyz, age, velocity xyz, lifetime, all f32) uploaded to the device, as the
 * reference's `particles` buffer (src/render/particles.js:720-730). A slot
 * whose age is not below its lifetime is dead (a zeroed slot is). The device
 * address (wd_pool_device) is for a simulation that writes the pool in place.
 * Refuses more than WD_MAX_NARROW particles by name.
 */
int wd_pool_create(wd_context *ctx, const void *particles, uint64_t count,
                   wd_pool **out);
int wd_pool_device(const wd_pool *pool, void **device_particles);
int wd_pool_destroy(wd_pool *pool);
/*
 * Sorts one emitter's ring, slots [first, first + count) of the pool, far to
 * near (makeKeys -> scanCounts -> place, src/render/particles.js:198-240):
 * each live particle's key is 4095 - bucket, bucket = floor((log2(depth) + 10)
 * 4095 / 30) clamped to [0, 4095], depth = max(dot(position - eye, forward),
 * 1e-6), decided in FP64 against the 4095 bucket edges 2^((30 b - 40950) /
 * 4095). The state's order gets the live particles' slots (first + index),
 * key ascending and, unlike the reference's atomics, slot ascending within a
 * key; its count the live particles (the draw's instanceCount). Enqueued on
 * the context's stream. Read back with wd_state_read (cov2d is not written).
 */
int wd_particles_sort(wd_context *ctx, const wd_pool *pool, uint64_t first,
                      uint64_t count, const wd_particle_view *view,
                      wd_state *state);

/* ---- timing: HIP events around each kernel on the context's stream ---- */

/* Turns per-kernel event timing on (1) or off (0); clears the totals. */
int wd_profile(wd_context *ctx, int on);
/*
 * Per-kernel totals since wd_profile(ctx, 1): names[k] (static strings),
 * summed milliseconds and launches. Returns the number of kernels in *n.
 */
int wd_profile_read(wd_context *ctx, const char **names, double *ms,
                    uint64_t *launches, int capacity, int *n);
/*
 * Runs `reps` sorts of `cloud` for `view` into `state`, each bracketed by HIP
 * events on the context's stream, and returns each one's milliseconds in
 * ms[reps] (the per-sort spread a benchmark reports as median and p95).
 */
int wd_time_sorts(wd_context *ctx, const wd_cloud *cloud, const wd_view *view,
                  wd_state *state, uint32_t reps, double *ms);

/* ---- batches: many (cloud, view) sorts in one graph launch ------------ */

/* One sort of a batch: what wd_sort takes. No two jobs may share a state. */
typedef struct wd_job {
  const wd_cloud *cloud;
  wd_view view;
  wd_state *state;
} wd_job;

/*
 * A batch replaces the reference's per-cloud loop of dispatches in
 * _encodeSort (src/render/splats.js:676-697): every job's projection and both
 * scatter passes run from one hipGraph launch, on one of two schedules:
 *   WD_SCHEDULE_PERSISTENT  ONE persistent kernel taking every job's work
 *                           run by run (a few tiles) from a ticketed queue;
 *   WD_SCHEDULE_PHASED      six kernels, each over every job's tiles.
 * A new batch runs phased with super-tiles of 1; wd_batch_tune times the
 * configurations and keeps the fastest. On MI355X the phased schedule wins:
 * the persistent one must make each item's writes visible to the other XCDs'
 * L2s before counting it done, and that write-back costs more than the five
 * launch boundaries it replaces.
 */
#define WD_SCHEDULE_PHASED 0
#define WD_SCHEDULE_PERSISTENT 1
int wd_batch_create(wd_context *ctx, uint32_t max_jobs, wd_batch **out);
/*
 * Picks the schedule, its work item and the tile. Phased: super_tiles tiles a
 * scatter workgroup takes at once, 1 or 2. Persistent: super_tiles tiles an
 * item takes one after the other, 1 to 64 (one release and one count a run).
 * keys_per_thread: 8 or 16, tiles of 2048 or 4096 splats (256 threads a
 * workgroup). The graph is captured again on the next run.
 */
int wd_batch_configure(wd_batch *batch, int schedule, uint32_t super_tiles, uint32_t keys_per_thread);
/* wd_batch_configure with the batch's keys a thread unchanged (16 on a new batch). */
int wd_batch_schedule(wd_batch *batch, int schedule, uint32_t super_tiles);
/* One configuration the autotuner timed, and its mean ms a run. */
typedef struct wd_tuned {
  int schedule;
  uint32_t super_tiles;
  uint32_t keys_per_thread;
  double ms;
} wd_tuned;
#define WD_TUNE_CONFIGS 6
/*
 * The autotuner's report: the configuration it kept, and every one it timed,
 * in this order: phased 1 and 2 super-tiles at 16 keys a thread, the same at
 * 8, persistent runs of 8 tiles at 16 and at 8. The digits are fixed at two
 * passes of 8 bits: any two-pass split of a 16-bit key ranks 16 bits a key,
 * and 8 + 8 makes the smallest tile histograms (2^a + 2^b counters) and
 * scans; three passes would add a full read and write of the keys.
 */
typedef struct wd_tuning {
  int schedule;
  uint32_t super_tiles;
  uint32_t keys_per_thread;
  uint32_t tile;        /* splats a tile: 256 x keys_per_thread */
  uint32_t radix_bits;  /* digit width of each pass (8) */
  uint32_t passes;      /* scatter passes (2) */
  wd_tuned timed[WD_TUNE_CONFIGS];
} wd_tuning;
/* Runs the set jobs `reps` times on each configuration and keeps the fastest. */
int wd_batch_tune(wd_context *ctx, wd_batch *batch, uint32_t reps, wd_tuning *out);
/*
 * Sets the jobs (views may change between runs at no cost; other clouds or
 * states re-capture the graph on the next run). A job whose cloud has
 * harmonics or is given as centres and covariances, or whose state takes the
 * draw's ellipses, is taken under the reverse-Z perspective only (WD_E_ARG
 * otherwise; wd_sort takes any projection).
 */
int wd_batch_set(wd_batch *batch, const wd_job *jobs, uint32_t n);
/*
 * Enqueues one run of every job; results land in each job's state. Refused
 * (WD_E_STATE) when a job's harmonics or ellipses changed since the jobs were
 * set: set them again.
 */
int wd_batch_run(wd_context *ctx, wd_batch *batch);
/*
 * Waits, then reports whether a bounded wait of the persistent kernel ran out
 * in any run since the last call (a bug; each call reports a run once). Such a
 * run also marks every one of its jobs' states, so wd_state_stats and
 * wd_state_read fail on them (WD_E_STATE) until the next sort.
 */
int wd_batch_stuck(wd_context *ctx, wd_batch *batch, uint32_t *stuck);
int wd_batch_destroy(wd_batch *batch);

/* ---- the learned depth-CDF predictor (no counterpart in the reference) */

/*
 * A small MLP, 12 pose floats -> 128 -> 128 -> 256, bf16 weights on the
 * matrix cores with fp32 accumulation. Its 256 outputs (a softmax) are the
 * widths, in key space, of 256 equal-mass bins of the survivors' depth keys:
 * the predicted depth CDF, inverted. Parameters (WD_MLP_PARAMS fp32, flat):
 * W1[128][12] b1[128] W2[128][128] b2[128] W3[256][128] b3[256].
 */
#define WD_MLP_PARAMS 51200
int wd_mlp_create(wd_context *ctx, uint64_t seed, wd_mlp **out);
int wd_mlp_destroy(wd_mlp *mlp);
int wd_mlp_set(wd_mlp *mlp, const float *params);
int wd_mlp_get(wd_mlp *mlp, float *params);
/* A saved predictor: "WDMLP1", the bf16 weights, the fp32 biases. */
int wd_mlp_save(wd_mlp *mlp, void *bytes, uint64_t capacity, uint64_t *size);
int wd_mlp_load(wd_mlp *mlp, const void *bytes, uint64_t size);
/*
 * One forward and backward pass on the GPU, no update: the mean cross-entropy
 * against targets (n x 256 distributions) and every parameter's gradient.
 */
int wd_mlp_grad(wd_context *ctx, wd_mlp *mlp, const float *poses,
                const float *targets, uint32_t n, float *grads, float *loss);
/* Adam on the GPU over `epochs` shuffled passes of minibatches of `batch`. */
int wd_mlp_train(wd_context *ctx, wd_mlp *mlp, const float *poses,
                 const float *targets, uint32_t n, uint32_t epochs,
                 uint32_t batch, float lr, uint64_t seed, float *loss);
/* n x 256 predicted bin widths (probabilities). */
int wd_mlp_predict(wd_context *ctx, wd_mlp *mlp, const float *poses, uint32_t n,
                   float *probs);
/* The 12 pose floats of a view: the model-view's top rows, translation / 10. */
int wd_pose_features(const wd_view *view, float out[12]);

/* ---- the learned sort (no counterpart in the reference) --------------- */

typedef struct wd_learned wd_learned;
/* Buffers for learned runs of up to max_views (1..8) views of one cloud. */
int wd_learned_create(wd_context *ctx, uint32_t max_views, wd_learned **out);
/*
 * Sorts `cloud` for n_views views into states[0..n_views) with the predictor
 * `mlp`: its depth CDF cuts each view's key line into bins, every record is
 * read once for all the views, and each bin is sorted in LDS. The result is
 * the exact order (what wd_sort gives), bit for bit: each view's bins are
 * checked on the GPU (no bin above `cap` survivors; 0 = 8 x n / 256 + 4096),
 * and when any view fails, the exact path re-sorts the run's views in the
 * same enqueue; wd_state_stats then reports fallback = 1 and the bins over
 * the cap for that view. Enqueued on the context's stream, no host wait.
 * Clouds of up to 2^23 splats.
 */
int wd_learned_run(wd_context *ctx, wd_learned *learned, wd_mlp *mlp,
                   const wd_cloud *cloud, const wd_view *views,
                   wd_state *const *states, uint32_t n_views, uint32_t cap);
int wd_learned_destroy(wd_learned *learned);
/*
 * Training targets from an exact sort in `state`: edges[k] = the key at rank
 * floor(k V / 256) (k = 1..255), edges[0] = 0, edges[256] = 65536. The
 * predictor's target widths are (edges[k+1] - edges[k]) / 65536.
 */
int wd_state_quantiles(wd_context *ctx, wd_state *state, uint32_t edges[257]);

/* ---- synthetic inputs: the workloads' seeded generator ---------------- */

/*
 * Writes n .splat records from the seeded generator. kind 0: three-component
 * Gaussian mixture; kind 1: 80% slab |z| <= 0.25 plus 20% uniform in a 40^3
 * box. Host memory; deterministic for (kind, seed, n).
 */
int wd_synth_records(int kind, uint64_t seed, uint64_t n, void *records_out);

/* ---- checks of the exact path (used by the parity tests) ------------- */

/*
 * The FP64 definition on the GPU for every splat, no f32 fast path: per splat
 * five doubles (passes frustum and alpha 0/1, alpha*area, key, 0, 0). Needs
 * the reverse-Z perspective and an affine model-view.
 */
int wd_debug_canonical(wd_context *ctx, const wd_cloud *cloud, const wd_view *view,
                       double *out5);
/*
 * Runs the DPP wave scan on known values and f64 sqrt(a[2i]) and
 * a[2i]/a[2i+1] against host-rounded expect[2i], expect[2i+1]; counts
 * mismatches of each into mismatch3.
 */
int wd_debug_selftest(wd_context *ctx, const double *args, const double *expect,
                      uint32_t count, uint32_t mismatch3[3]);
/*
 * A plain device-to-device copy of `bytes` (read once, written once), `reps`
 * times; *ms_mean gets the mean duration from HIP events. What HBM sustains,
 * and the yardstick the FETCH_SIZE / WRITE_SIZE counters are scaled by.
 */
int wd_debug_copy(wd_context *ctx, uint64_t bytes, int reps, double *ms_mean);
/*
 * Sets the batch's live "a wait ran out" flag from the host, as a wait that
 * ran out would, so that tests can check the reporting above without a hang:
 * the next persistent run then gives up any wait longer than 1024 polls and
 * marks its jobs' states as not valid.
 */
int wd_debug_batch_force_stuck(wd_context *ctx, wd_batch *batch);

#ifdef __cplusplus
}
#endif
#endif /* WINDING_AMD_H */

/*
 * winding_amd.node -- the N-API addon the JavaScript module loads: a thin
 * shim from JS values to the C ABI of libwinding_amd.so (include/winding_amd.h).
 *
 * Handles are externals holding a small box; GPU objects also keep a strong
 * reference to their context's external, so a context outlives its clouds and
 * states. destroy() frees at once; a box the JS side dropped without
 * destroy() is freed by its finaliser. A call on a destroyed handle throws,
 * and every C failure becomes a JS Error with wd_last_error()'s message, the
 * reference's synchronous `throw new Error('splats: ...')` convention
 * (src/render/splats.js:394-400).
 */
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "winding_amd.h"

enum { K_CTX = 1, K_CLOUD = 2, K_STATE = 3, K_BATCH = 4, K_MLP = 5, K_LEARNED = 6, K_POOL = 7 };

typedef struct Box {
  int kind;
  void *ptr;         /* wd_context*, wd_cloud*, wd_state*, wd_batch*, wd_mlp*, wd_learned*; NULL once destroyed */
  napi_ref parent;   /* the context's external, for clouds and states */
  struct Box *ctx;   /* the context's box */
  int children;      /* contexts: live clouds and states */
} Box;

#define CHECK_NAPI(call)                                   \
  do {                                                     \
    if ((call) != napi_ok) {                               \
      napi_throw_error(env, NULL, "winding_amd: N-API call failed"); \
      return NULL;                                         \
    }                                                      \
  } while (0)

static napi_value throw_wd(napi_env env, int rc) {
  const char *msg = wd_last_error();
  (void)rc;
  napi_throw_error(env, NULL, msg && msg[0] ? msg : "winding_amd: failed");
  return NULL;
}

static void release_box(napi_env env, Box *b) {
  if (b->ptr) {
    if (b->kind == K_CLOUD) wd_cloud_destroy((wd_cloud *)b->ptr);
    if (b->kind == K_STATE) wd_state_destroy((wd_state *)b->ptr);
    if (b->kind == K_BATCH) wd_batch_destroy((wd_batch *)b->ptr);
    if (b->kind == K_MLP) wd_mlp_destroy((wd_mlp *)b->ptr);
    if (b->kind == K_LEARNED) wd_learned_destroy((wd_learned *)b->ptr);
    if (b->kind == K_POOL) wd_pool_destroy((wd_pool *)b->ptr);
    if (b->kind == K_CTX) wd_close((wd_context *)b->ptr);
    b->ptr = NULL;
    if (b->ctx) b->ctx->children--;
  }
  if (b->parent) {
    napi_delete_reference(env, b->parent);
    b->parent = NULL;
  }
}

static void finalize_box(napi_env env, void *data, void *hint) {
  (void)hint;
  Box *b = (Box *)data;
  /* a context is only collected once no cloud or state refers to it */
  release_box(env, b);
  free(b);
}

static napi_value make_handle(napi_env env, int kind, void *ptr, napi_value ctx_value, Box *ctx_box) {
  Box *b = (Box *)calloc(1, sizeof(Box));
  napi_value ext;
  b->kind = kind;
  b->ptr = ptr;
  b->ctx = ctx_box;
  if (napi_create_external(env, b, finalize_box, NULL, &ext) != napi_ok) {
    free(b);
    napi_throw_error(env, NULL, "winding_amd: N-API call failed");
    return NULL;
  }
  if (ctx_value) {
    napi_create_reference(env, ctx_value, 1, &b->parent);
    ctx_box->children++;
  }
  return ext;
}

static Box *get_box(napi_env env, napi_value v, int kind) {
  void *data = NULL;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data || ((Box *)data)->kind != kind) {
    napi_throw_type_error(env, NULL, "winding_amd: wrong handle");
    return NULL;
  }
  Box *b = (Box *)data;
  if (!b->ptr) {
    napi_throw_error(env, NULL,
                     kind == K_CLOUD ? "addSplats: these splats were unloaded; remove the node first"
                                     : "winding_amd: this handle was destroyed");
    return NULL;
  }
  return b;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok || argc < want) {
    napi_throw_type_error(env, NULL, "winding_amd: too few arguments");
    return 0;
  }
  return 1;
}

static int typed(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *length) {
  bool is = false;
  napi_typedarray_type t;
  napi_value buf;
  size_t off;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is ||
      napi_get_typedarray_info(env, v, &t, length, data, &buf, &off) != napi_ok || t != want) {
    napi_throw_type_error(env, NULL, "winding_amd: expected a typed array of another type");
    return 0;
  }
  return 1;
}

static napi_value f32_array(napi_env env, const float *src, size_t n) {
  napi_value ab, arr;
  void *data;
  if (napi_create_arraybuffer(env, n * 4, &data, &ab) != napi_ok) return NULL;
  memcpy(data, src, n * 4);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr);
  return arr;
}

static napi_value num(napi_env env, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  return v;
}

/* open(device) -> context */
static napi_value js_open(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  int32_t dev = 0;
  if (!get_args(env, info, 1, argv)) return NULL;
  CHECK_NAPI(napi_get_value_int32(env, argv[0], &dev));
  wd_context *ctx = NULL;
  int rc = wd_open(dev, &ctx);
  if (rc) return throw_wd(env, rc);
  return make_handle(env, K_CTX, ctx, NULL, NULL);
}

/* close(context): refuses while clouds or states are alive */
static napi_value js_close(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Box *b = get_box(env, argv[0], K_CTX);
  if (!b) return NULL;
  if (b->children > 0) {
    napi_throw_error(env, NULL, "winding_amd: close: destroy its splats and sort states first");
    return NULL;
  }
  release_box(env, b);
  return NULL;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  Box *b = get_box(env, argv[0], K_CTX);
  if (!b) return NULL;
  int rc = wd_sync((wd_context *)b->ptr);
  return rc ? throw_wd(env, rc) : NULL;
}

/* cloudCreate(context, Uint8Array records) -> {handle, count, min, max} */
static napi_value js_cloud_create(napi_env env, napi_callback_info info) {
  napi_value argv[2], out;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  void *data;
  size_t len;
  if (!typed(env, argv[1], napi_uint8_array, &data, &len)) return NULL;
  wd_cloud *cloud = NULL;
  float mn[3], mx[3];
  int rc = wd_cloud_create((wd_context *)c->ptr, data, len, &cloud, mn, mx);
  if (rc) return throw_wd(env, rc);
  napi_value h = make_handle(env, K_CLOUD, cloud, argv[0], c);
  if (!h) return NULL;
  CHECK_NAPI(napi_create_object(env, &out));
  napi_set_named_property(env, out, "handle", h);
  napi_set_named_property(env, out, "count", num(env, (double)(len / 32)));
  napi_set_named_property(env, out, "min", f32_array(env, mn, 3));
  napi_set_named_property(env, out, "max", f32_array(env, mx, 3));
  return out;
}

/* cloudSH(context, cloud, Uint32Array sh, degree): the capture's harmonics, degree 1-3 (0 removes them) */
static napi_value js_cloud_sh(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *cl = get_box(env, argv[1], K_CLOUD);
  if (!cl) return NULL;
  void *data;
  size_t len;
  if (!typed(env, argv[2], napi_uint32_array, &data, &len)) return NULL;
  uint32_t degree = 0;
  CHECK_NAPI(napi_get_value_uint32(env, argv[3], &degree));
  int rc = wd_cloud_sh((wd_context *)c->ptr, (wd_cloud *)cl->ptr, (const uint32_t *)data, len, degree);  /* len: words */
  return rc ? throw_wd(env, rc) : NULL;
}

/* A decoded or given cloud as {handle, count, degree, min, max}. */
static napi_value cloud_result(napi_env env, napi_value ctx_value, Box *c, wd_cloud *cloud, const float *mn,
                               const float *mx) {
  napi_value out;
  napi_value h = make_handle(env, K_CLOUD, cloud, ctx_value, c);
  if (!h) return NULL;
  uint64_t count = 0;
  uint32_t degree = 0;
  wd_cloud_info(cloud, &count, NULL, NULL);
  wd_cloud_read((wd_context *)c->ptr, cloud, NULL, NULL, NULL, &degree);
  CHECK_NAPI(napi_create_object(env, &out));
  napi_set_named_property(env, out, "handle", h);
  napi_set_named_property(env, out, "count", num(env, (double)count));
  napi_set_named_property(env, out, "degree", num(env, degree));
  napi_set_named_property(env, out, "min", f32_array(env, mn, 3));
  napi_set_named_property(env, out, "max", f32_array(env, mx, 3));
  return out;
}

/* cloudCreateSplats(context, Float32Array centers (4 n), Float32Array covariances (6 n)) -> cloud */
static napi_value js_cloud_create_splats(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  void *centers, *covs;
  size_t lc, lv;
  if (!typed(env, argv[1], napi_float32_array, &centers, &lc)) return NULL;
  if (!typed(env, argv[2], napi_float32_array, &covs, &lv)) return NULL;
  if (lc % 4 != 0 || lv != lc / 4 * 6) {
    napi_throw_range_error(env, NULL, "splats: 4 centre floats and 6 covariance floats a splat");
    return NULL;
  }
  wd_cloud *cloud = NULL;
  float mn[3], mx[3];
  int rc = wd_cloud_create_splats((wd_context *)c->ptr, (const float *)centers, (const float *)covs, lc / 4, &cloud,
                                  mn, mx);
  if (rc) return throw_wd(env, rc);
  return cloud_result(env, argv[0], c, cloud, mn, mx);
}

/* decodePly / decodeSpz(context, Uint8Array bytes) -> cloud, decoded on the GPU */
static napi_value decode_bytes(napi_env env, napi_callback_info info, int spz) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  void *data;
  size_t len;
  if (!typed(env, argv[1], napi_uint8_array, &data, &len)) return NULL;
  wd_cloud *cloud = NULL;
  float mn[3], mx[3];
  int rc = spz ? wd_decode_spz((wd_context *)c->ptr, data, len, &cloud, mn, mx)
               : wd_decode_ply((wd_context *)c->ptr, data, len, &cloud, mn, mx);
  if (rc) return throw_wd(env, rc);
  return cloud_result(env, argv[0], c, cloud, mn, mx);
}
static napi_value js_decode_ply(napi_env env, napi_callback_info info) { return decode_bytes(env, info, 0); }
static napi_value js_decode_spz(napi_env env, napi_callback_info info) { return decode_bytes(env, info, 1); }

static int get_named(napi_env env, napi_value obj, const char *name, napi_value *out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return 0;
  return napi_get_named_property(env, obj, name, out) == napi_ok;
}

static int sog_image(napi_env env, napi_value images, const char *name, wd_sog_image *img) {
  napi_value o, v;
  size_t len;
  void *data;
  uint32_t w = 0, h = 0;
  if (!get_named(env, images, name, &o)) return 1;  /* absent: left null, the library says which */
  if (!get_named(env, o, "rgba", &v) || !typed(env, v, napi_uint8_array, &data, &len)) return 0;
  if (get_named(env, o, "width", &v)) napi_get_value_uint32(env, v, &w);
  if (get_named(env, o, "height", &v)) napi_get_value_uint32(env, v, &h);
  if ((uint64_t)w * h * 4 > len) {
    napi_throw_range_error(env, NULL, "splats: a .sog image holds fewer bytes than width x height x 4");
    return 0;
  }
  img->rgba = (const uint8_t *)data;
  img->width = w;
  img->height = h;
  return 1;
}

/*
 * decodeSog(context, {count, bands, shCount, mins, maxs, scaleCodebook, colourCodebook, shCodebook
 * (Float64Arrays), images: {means_l, means_u, quats, scales, sh0, centroids, labels: {width, height,
 * rgba}}}) -> cloud: the numbers of meta.json and the decoded images (readSog's loop on the GPU).
 */
static napi_value js_decode_sog(napi_env env, napi_callback_info info) {
  napi_value argv[2], v, images;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  wd_sog s;
  memset(&s, 0, sizeof s);
  napi_value o = argv[1];
  if (get_named(env, o, "count", &v)) napi_get_value_uint32(env, v, &s.count);
  if (get_named(env, o, "bands", &v)) napi_get_value_uint32(env, v, &s.bands);
  s.sh_count = 0xFFFFFFFFu;
  if (get_named(env, o, "shCount", &v)) napi_get_value_uint32(env, v, &s.sh_count);
  const char *vec[2] = {"mins", "maxs"};
  double *dst[2] = {s.mins, s.maxs};
  for (int k = 0; k < 2; ++k) {
    void *d;
    size_t l;
    if (!get_named(env, o, vec[k], &v) || !typed(env, v, napi_float64_array, &d, &l) || l < 3) return NULL;
    memcpy(dst[k], d, 3 * sizeof(double));
  }
  const char *books[3] = {"scaleCodebook", "colourCodebook", "shCodebook"};
  const double **bp[3] = {&s.scale_codebook, &s.colour_codebook, &s.sh_codebook};
  uint32_t *bn[3] = {&s.scale_codebook_size, &s.colour_codebook_size, &s.sh_codebook_size};
  for (int k = 0; k < 3; ++k) {
    void *d;
    size_t l;
    if (!get_named(env, o, books[k], &v)) continue;
    if (!typed(env, v, napi_float64_array, &d, &l)) return NULL;
    *bp[k] = (const double *)d;
    *bn[k] = (uint32_t)l;
  }
  if (!get_named(env, o, "images", &images)) {
    napi_throw_type_error(env, NULL, "splats: decodeSog needs images");
    return NULL;
  }
  const char *names[7] = {"means_l", "means_u", "quats", "scales", "sh0", "centroids", "labels"};
  wd_sog_image *imgs[7] = {&s.means_l, &s.means_u, &s.quats, &s.scales, &s.sh0, &s.centroids, &s.labels};
  for (int k = 0; k < 7; ++k)
    if (!sog_image(env, images, names[k], imgs[k])) return NULL;
  wd_cloud *cloud = NULL;
  float mn[3], mx[3];
  int rc = wd_decode_sog((wd_context *)c->ptr, &s, &cloud, mn, mx);
  if (rc) return throw_wd(env, rc);
  return cloud_result(env, argv[0], c, cloud, mn, mx);
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  void *data = NULL;
  if (!get_args(env, info, 1, argv)) return NULL;
  if (napi_get_value_external(env, argv[0], &data) != napi_ok || !data) {
    napi_throw_type_error(env, NULL, "winding_amd: wrong handle");
    return NULL;
  }
  Box *b = (Box *)data;
  if (b->kind == K_CTX) {
    napi_throw_type_error(env, NULL, "winding_amd: close a context with close()");
    return NULL;
  }
  release_box(env, b);
  return NULL;
}

/*
 * viewPrepare(view16, world16|null, projection16, width, height, min3, max3, cull)
 * -> Float32Array(37) uniform, or null when the cloud is wholly behind.
 */
static napi_value js_view_prepare(napi_env env, napi_callback_info info) {
  napi_value argv[8], nul;
  if (!get_args(env, info, 8, argv)) return NULL;
  void *view, *proj, *mn, *mx, *world = NULL;
  size_t l;
  napi_valuetype t;
  double w, h, cull;
  if (!typed(env, argv[0], napi_float32_array, &view, &l) || l < 16) return NULL;
  napi_typeof(env, argv[1], &t);
  if (t != napi_null && t != napi_undefined && (!typed(env, argv[1], napi_float32_array, &world, &l) || l < 16))
    return NULL;
  if (!typed(env, argv[2], napi_float32_array, &proj, &l) || l < 16) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[3], &w));
  CHECK_NAPI(napi_get_value_double(env, argv[4], &h));
  if (!typed(env, argv[5], napi_float32_array, &mn, &l) || l < 3) return NULL;
  if (!typed(env, argv[6], napi_float32_array, &mx, &l) || l < 3) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[7], &cull));
  wd_view v;
  int rc = wd_view_prepare((const float *)view, (const float *)world, (const float *)proj, (float)w, (float)h,
                           (const float *)mn, (const float *)mx, (float)cull, &v);
  if (rc < 0) return throw_wd(env, rc);
  if (rc == 1) {
    napi_get_null(env, &nul);
    return nul;
  }
  return f32_array(env, (const float *)&v, sizeof v / 4);
}

/* stateCreate(context, capacity) -> state */
static napi_value js_state_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  double cap;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[1], &cap));
  wd_state *s = NULL;
  int rc = wd_state_create((wd_context *)c->ptr, (uint64_t)cap, &s);
  if (rc) return throw_wd(env, rc);
  return make_handle(env, K_STATE, s, argv[0], c);
}

/* sort(context, cloud, uniform Float32Array(37), state) */
static napi_value js_sort(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *cl = get_box(env, argv[1], K_CLOUD);
  if (!cl) return NULL;
  void *u;
  size_t l;
  if (!typed(env, argv[2], napi_float32_array, &u, &l) || l < sizeof(wd_view) / 4) return NULL;
  Box *s = get_box(env, argv[3], K_STATE);
  if (!s) return NULL;
  wd_view v;
  memcpy(&v, u, sizeof v);
  int rc = wd_sort((wd_context *)c->ptr, (wd_cloud *)cl->ptr, &v, (wd_state *)s->ptr);
  return rc ? throw_wd(env, rc) : NULL;
}

/* stateRead(context, state, covCount) -> {count, recomputed, order, cov2d?} (waits) */
static napi_value js_state_read(napi_env env, napi_callback_info info) {
  napi_value argv[3], out, ab, arr;
  double ncov;
  if (!get_args(env, info, 3, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *s = get_box(env, argv[1], K_STATE);
  if (!s) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[2], &ncov));
  wd_stats st;
  int rc = wd_state_stats((wd_context *)c->ptr, (wd_state *)s->ptr, &st);
  if (rc) return throw_wd(env, rc);
  void *order_data, *cov_data = NULL;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)st.count * 4 + 4, &order_data, &ab));
  napi_value cov_ab = NULL;
  if (ncov > 0) CHECK_NAPI(napi_create_arraybuffer(env, (size_t)ncov * 16, &cov_data, &cov_ab));
  rc = wd_state_read((wd_context *)c->ptr, (wd_state *)s->ptr, (uint32_t *)order_data, st.count + 1,
                     (float *)cov_data, (uint64_t)ncov, &st);
  if (rc) return throw_wd(env, rc);
  CHECK_NAPI(napi_create_object(env, &out));
  napi_create_typedarray(env, napi_uint32_array, st.count, ab, 0, &arr);
  napi_set_named_property(env, out, "order", arr);
  napi_set_named_property(env, out, "count", num(env, st.count));
  napi_set_named_property(env, out, "recomputed", num(env, st.recomputed));
  napi_set_named_property(env, out, "fallback", num(env, st.fallback));
  napi_set_named_property(env, out, "overflowBins", num(env, st.overflow_bins));
  if (cov_ab) {
    napi_create_typedarray(env, napi_float32_array, (size_t)ncov * 4, cov_ab, 0, &arr);
    napi_set_named_property(env, out, "cov2d", arr);
  }
  return out;
}

/* stateEllipses(ctx, state, on): whether each sort also writes the draw's ellipse set-up */
static napi_value js_state_ellipses(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  bool on;
  if (!get_args(env, info, 3, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *s = get_box(env, argv[1], K_STATE);
  if (!s) return NULL;
  CHECK_NAPI(napi_get_value_bool(env, argv[2], &on));
  int rc = wd_state_ellipses((wd_context *)c->ptr, (wd_state *)s->ptr, on ? 1 : 0);
  return rc ? throw_wd(env, rc) : NULL;
}

/* stateReadEllipses(ctx, state, n) -> Float32Array(8 n): clip x y z w, then (along e, across e) */
static napi_value js_state_read_ellipses(napi_env env, napi_callback_info info) {
  napi_value argv[3], ab, arr;
  double n;
  void *data;
  if (!get_args(env, info, 3, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *s = get_box(env, argv[1], K_STATE);
  if (!s) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[2], &n));
  if (!(n >= 0)) n = 0;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)n * 32 + 4, &data, &ab));
  int rc = wd_state_read_ellipses((wd_context *)c->ptr, (wd_state *)s->ptr, (float *)data, (uint64_t)n);
  if (rc) return throw_wd(env, rc);
  CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, (size_t)n * 8, ab, 0, &arr));
  return arr;
}

/* poolCreate(context, Uint8Array particles) -> handle: 32-byte particles (position, age, velocity, lifetime) */
static napi_value js_pool_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  void *data;
  size_t len;
  if (!typed(env, argv[1], napi_uint8_array, &data, &len)) return NULL;
  if (len % WD_PARTICLE_BYTES != 0) {
    napi_throw_error(env, NULL, "particles: not a whole number of 32-byte particles");
    return NULL;
  }
  wd_pool *pool = NULL;
  int rc = wd_pool_create((wd_context *)c->ptr, data, len / WD_PARTICLE_BYTES, &pool);
  if (rc) return throw_wd(env, rc);
  return make_handle(env, K_POOL, pool, argv[0], c);
}

/* particlesSort(context, pool, first, count, Float32Array(8) view: eye xyz_, forward xyz_, state) */
static napi_value js_particles_sort(napi_env env, napi_callback_info info) {
  napi_value argv[6];
  double first, count;
  void *u;
  size_t l;
  if (!get_args(env, info, 6, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *p = get_box(env, argv[1], K_POOL);
  if (!p) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[2], &first));
  CHECK_NAPI(napi_get_value_double(env, argv[3], &count));
  if (!typed(env, argv[4], napi_float32_array, &u, &l) || l < 8) return NULL;
  Box *s = get_box(env, argv[5], K_STATE);
  if (!s) return NULL;
  if (!(first >= 0) || !(count >= 0)) {
    napi_throw_range_error(env, NULL, "particles: first and count must be 0 or more");
    return NULL;
  }
  wd_particle_view v;
  memcpy(&v, u, sizeof v);
  int rc = wd_particles_sort((wd_context *)c->ptr, (wd_pool *)p->ptr, (uint64_t)first, (uint64_t)count, &v,
                             (wd_state *)s->ptr);
  return rc ? throw_wd(env, rc) : NULL;
}

/* synth(kind, seed, n) -> Uint8Array of n seeded .splat records (host only) */
static napi_value js_synth(napi_env env, napi_callback_info info) {
  napi_value argv[3], ab, arr;
  int32_t kind;
  double seed, n;
  if (!get_args(env, info, 3, argv)) return NULL;
  CHECK_NAPI(napi_get_value_int32(env, argv[0], &kind));
  CHECK_NAPI(napi_get_value_double(env, argv[1], &seed));
  CHECK_NAPI(napi_get_value_double(env, argv[2], &n));
  void *data;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)n * 32, &data, &ab));
  int rc = wd_synth_records(kind, (uint64_t)seed, (uint64_t)n, data);
  if (rc) return throw_wd(env, rc);
  napi_create_typedarray(env, napi_uint8_array, (size_t)n * 32, ab, 0, &arr);
  return arr;
}

/* profile(context, on) / profileRead(context) -> {name: [ms, launches]} */
static napi_value js_profile(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  bool on;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  CHECK_NAPI(napi_get_value_bool(env, argv[1], &on));
  int rc = wd_profile((wd_context *)c->ptr, on ? 1 : 0);
  return rc ? throw_wd(env, rc) : NULL;
}

static napi_value js_profile_read(napi_env env, napi_callback_info info) {
  napi_value argv[1], out, pair;
  if (!get_args(env, info, 1, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  const char *names[16];
  double ms[16];
  uint64_t launches[16];
  int n = 0;
  int rc = wd_profile_read((wd_context *)c->ptr, names, ms, launches, 16, &n);
  if (rc) return throw_wd(env, rc);
  CHECK_NAPI(napi_create_object(env, &out));
  for (int i = 0; i < n; ++i) {
    if (!launches[i]) continue;
    napi_create_array_with_length(env, 2, &pair);
    napi_set_element(env, pair, 0, num(env, ms[i]));
    napi_set_element(env, pair, 1, num(env, (double)launches[i]));
    napi_set_named_property(env, out, names[i], pair);
  }
  return out;
}

/* ---- batches, the predictor and the learned sort --------------------- */

static napi_value handle_of(napi_env env, int kind, void *ptr, napi_value ctx_value, Box *ctx_box) {
  return make_handle(env, kind, ptr, ctx_value, ctx_box);
}

static int u32_arg(napi_env env, napi_value v, uint32_t *out) {
  double d;
  if (napi_get_value_double(env, v, &d) != napi_ok || !(d >= 0) || d > 4294967295.0) {
    napi_throw_type_error(env, NULL, "winding_amd: expected a count");
    return 0;
  }
  *out = (uint32_t)d;
  return 1;
}

/* Reads a JS array of n handles of `kind` into ptrs. */
static int handle_list(napi_env env, napi_value arr, int kind, void **ptrs, uint32_t cap, uint32_t *n) {
  bool is = false;
  if (napi_is_array(env, arr, &is) != napi_ok || !is) {
    napi_throw_type_error(env, NULL, "winding_amd: expected an array");
    return 0;
  }
  napi_get_array_length(env, arr, n);
  if (*n > cap) {
    napi_throw_range_error(env, NULL, "winding_amd: too many items");
    return 0;
  }
  for (uint32_t i = 0; i < *n; ++i) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    Box *b = get_box(env, e, kind);
    if (!b) return 0;
    ptrs[i] = b->ptr;
  }
  return 1;
}

/* Reads a JS array of n Float32Array(37) uniforms. */
static int view_list(napi_env env, napi_value arr, wd_view *views, uint32_t cap, uint32_t *n) {
  bool is = false;
  if (napi_is_array(env, arr, &is) != napi_ok || !is) {
    napi_throw_type_error(env, NULL, "winding_amd: expected an array of uniforms");
    return 0;
  }
  napi_get_array_length(env, arr, n);
  if (*n > cap) {
    napi_throw_range_error(env, NULL, "winding_amd: too many views");
    return 0;
  }
  for (uint32_t i = 0; i < *n; ++i) {
    napi_value e;
    void *u;
    size_t l;
    napi_get_element(env, arr, i, &e);
    if (!typed(env, e, napi_float32_array, &u, &l) || l < sizeof(wd_view) / 4) return 0;
    memcpy(&views[i], u, sizeof(wd_view));
  }
  return 1;
}

#define MAX_LIST 4096

/* batchCreate(context, maxJobs) -> batch */
static napi_value js_batch_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  uint32_t m;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c || !u32_arg(env, argv[1], &m)) return NULL;
  wd_batch *b = NULL;
  int rc = wd_batch_create((wd_context *)c->ptr, m, &b);
  return rc ? throw_wd(env, rc) : handle_of(env, K_BATCH, b, argv[0], c);
}

/* batchSet(batch, [clouds], [uniforms], [states]) */
static napi_value js_batch_set(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  Box *b = get_box(env, argv[0], K_BATCH);
  if (!b) return NULL;
  uint32_t nc = 0, nv = 0, ns = 0;
  void **clouds = (void **)calloc(MAX_LIST, sizeof(void *));
  void **states = (void **)calloc(MAX_LIST, sizeof(void *));
  wd_view *views = (wd_view *)calloc(MAX_LIST, sizeof(wd_view));
  wd_job *jobs = (wd_job *)calloc(MAX_LIST, sizeof(wd_job));
  napi_value ret = NULL;
  if (clouds && states && views && jobs && handle_list(env, argv[1], K_CLOUD, clouds, MAX_LIST, &nc) &&
      view_list(env, argv[2], views, MAX_LIST, &nv) && handle_list(env, argv[3], K_STATE, states, MAX_LIST, &ns)) {
    if (nc != nv || nc != ns) {
      napi_throw_range_error(env, NULL, "batch: clouds, uniforms and states differ in number");
    } else {
      for (uint32_t i = 0; i < nc; ++i) {
        jobs[i].cloud = (const wd_cloud *)clouds[i];
        jobs[i].view = views[i];
        jobs[i].state = (wd_state *)states[i];
      }
      int rc = wd_batch_set((wd_batch *)b->ptr, jobs, nc);
      if (rc) throw_wd(env, rc);
    }
  }
  free(clouds);
  free(states);
  free(views);
  free(jobs);
  return ret;
}

/* batchRun(context, batch) */
static napi_value js_batch_run(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *b = get_box(env, argv[1], K_BATCH);
  if (!b) return NULL;
  int rc = wd_batch_run((wd_context *)c->ptr, (wd_batch *)b->ptr);
  return rc ? throw_wd(env, rc) : NULL;
}

/* batchConfigure(batch, schedule: 'phased'|'persistent', superTiles, keysPerThread) */
static napi_value js_batch_configure(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  char which[16];
  size_t len = 0;
  uint32_t super_tiles, kpt;
  if (!get_args(env, info, 4, argv)) return NULL;
  Box *b = get_box(env, argv[0], K_BATCH);
  if (!b) return NULL;
  if (napi_get_value_string_utf8(env, argv[1], which, sizeof which, &len) != napi_ok) {
    napi_throw_type_error(env, NULL, "batchConfigure: the schedule is 'phased' or 'persistent'");
    return NULL;
  }
  if (!u32_arg(env, argv[2], &super_tiles) || !u32_arg(env, argv[3], &kpt)) return NULL;
  int schedule = -1;
  if (strcmp(which, "phased") == 0) schedule = WD_SCHEDULE_PHASED;
  if (strcmp(which, "persistent") == 0) schedule = WD_SCHEDULE_PERSISTENT;
  if (schedule < 0) {
    napi_throw_type_error(env, NULL, "batchConfigure: the schedule is 'phased' or 'persistent'");
    return NULL;
  }
  int rc = wd_batch_configure((wd_batch *)b->ptr, schedule, super_tiles, kpt);
  return rc ? throw_wd(env, rc) : NULL;
}

/* stateKeysPerThread(state, keysPerThread): 0 (by the cloud's size), 4, 8 or 16 */
static napi_value js_state_keys_per_thread(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  uint32_t kpt;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *s = get_box(env, argv[0], K_STATE);
  if (!s || !u32_arg(env, argv[1], &kpt)) return NULL;
  int rc = wd_state_keys_per_thread((wd_state *)s->ptr, kpt);
  return rc ? throw_wd(env, rc) : NULL;
}

/* batchTune(context, batch, reps) -> {schedule: 'phased'|'persistent', superTiles, keysPerThread, tile, radixBits,
 * passes, timed: [{schedule, superTiles, keysPerThread, ms}, ...]} */
static napi_value js_batch_tune(napi_env env, napi_callback_info info) {
  napi_value argv[3], out, sv;
  uint32_t reps;
  if (!get_args(env, info, 3, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *b = get_box(env, argv[1], K_BATCH);
  if (!b || !u32_arg(env, argv[2], &reps)) return NULL;
  wd_tuning t;
  int rc = wd_batch_tune((wd_context *)c->ptr, (wd_batch *)b->ptr, reps, &t);
  if (rc) return throw_wd(env, rc);
  CHECK_NAPI(napi_create_object(env, &out));
  napi_create_string_utf8(env, t.schedule == WD_SCHEDULE_PERSISTENT ? "persistent" : "phased", NAPI_AUTO_LENGTH, &sv);
  napi_set_named_property(env, out, "schedule", sv);
  napi_set_named_property(env, out, "superTiles", num(env, t.super_tiles));
  napi_set_named_property(env, out, "keysPerThread", num(env, t.keys_per_thread));
  napi_set_named_property(env, out, "tile", num(env, t.tile));
  napi_set_named_property(env, out, "radixBits", num(env, t.radix_bits));
  napi_set_named_property(env, out, "passes", num(env, t.passes));
  napi_value list;
  CHECK_NAPI(napi_create_array_with_length(env, WD_TUNE_CONFIGS, &list));
  for (uint32_t i = 0; i < WD_TUNE_CONFIGS; ++i) {
    napi_value e, es;
    CHECK_NAPI(napi_create_object(env, &e));
    napi_create_string_utf8(env, t.timed[i].schedule == WD_SCHEDULE_PERSISTENT ? "persistent" : "phased",
                            NAPI_AUTO_LENGTH, &es);
    napi_set_named_property(env, e, "schedule", es);
    napi_set_named_property(env, e, "superTiles", num(env, t.timed[i].super_tiles));
    napi_set_named_property(env, e, "keysPerThread", num(env, t.timed[i].keys_per_thread));
    napi_set_named_property(env, e, "ms", num(env, t.timed[i].ms));
    CHECK_NAPI(napi_set_element(env, list, i, e));
  }
  napi_set_named_property(env, out, "timed", list);
  return out;
}

/* mlpCreate(context, seed) -> predictor */
static napi_value js_mlp_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  uint32_t seed;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c || !u32_arg(env, argv[1], &seed)) return NULL;
  wd_mlp *m = NULL;
  int rc = wd_mlp_create((wd_context *)c->ptr, seed, &m);
  return rc ? throw_wd(env, rc) : handle_of(env, K_MLP, m, argv[0], c);
}

/* mlpSave(predictor) -> Uint8Array;  mlpLoad(predictor, Uint8Array) */
static napi_value js_mlp_save(napi_env env, napi_callback_info info) {
  napi_value argv[1], ab, arr;
  if (!get_args(env, info, 1, argv)) return NULL;
  Box *m = get_box(env, argv[0], K_MLP);
  if (!m) return NULL;
  uint64_t size = 0;
  int rc = wd_mlp_save((wd_mlp *)m->ptr, NULL, 0, &size);
  if (rc) return throw_wd(env, rc);
  void *data;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)size, &data, &ab));
  rc = wd_mlp_save((wd_mlp *)m->ptr, data, size, &size);
  if (rc) return throw_wd(env, rc);
  napi_create_typedarray(env, napi_uint8_array, (size_t)size, ab, 0, &arr);
  return arr;
}

static napi_value js_mlp_load(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  void *data;
  size_t l;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *m = get_box(env, argv[0], K_MLP);
  if (!m || !typed(env, argv[1], napi_uint8_array, &data, &l)) return NULL;
  int rc = wd_mlp_load((wd_mlp *)m->ptr, data, l);
  return rc ? throw_wd(env, rc) : NULL;
}

/* mlpTrain(context, predictor, poses Float32Array(n*12), targets Float32Array(n*256), epochs, batch, lr, seed) -> loss */
static napi_value js_mlp_train(napi_env env, napi_callback_info info) {
  napi_value argv[8];
  void *x, *t;
  size_t lx, lt;
  uint32_t epochs, batch, seed;
  double lr;
  if (!get_args(env, info, 8, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *m = get_box(env, argv[1], K_MLP);
  if (!m || !typed(env, argv[2], napi_float32_array, &x, &lx) || !typed(env, argv[3], napi_float32_array, &t, &lt))
    return NULL;
  if (!u32_arg(env, argv[4], &epochs) || !u32_arg(env, argv[5], &batch) || !u32_arg(env, argv[7], &seed)) return NULL;
  CHECK_NAPI(napi_get_value_double(env, argv[6], &lr));
  const uint32_t n = (uint32_t)(lx / WD_POSE_FLOATS);
  if (lx != (size_t)n * WD_POSE_FLOATS || lt != (size_t)n * WD_COARSE_BINS) {
    napi_throw_range_error(env, NULL, "mlp: poses are n x 12 floats and targets n x 256");
    return NULL;
  }
  float loss = 0.f;
  int rc = wd_mlp_train((wd_context *)c->ptr, (wd_mlp *)m->ptr, (const float *)x, (const float *)t, n, epochs, batch,
                        (float)lr, seed, &loss);
  return rc ? throw_wd(env, rc) : num(env, loss);
}

/* mlpPredict(context, predictor, poses Float32Array(n*12)) -> Float32Array(n*256) */
static napi_value js_mlp_predict(napi_env env, napi_callback_info info) {
  napi_value argv[3], ab, arr;
  void *x, *out;
  size_t lx;
  if (!get_args(env, info, 3, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *m = get_box(env, argv[1], K_MLP);
  if (!m || !typed(env, argv[2], napi_float32_array, &x, &lx)) return NULL;
  const uint32_t n = (uint32_t)(lx / WD_POSE_FLOATS);
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)n * WD_COARSE_BINS * 4, &out, &ab));
  int rc = wd_mlp_predict((wd_context *)c->ptr, (wd_mlp *)m->ptr, (const float *)x, n, (float *)out);
  if (rc) return throw_wd(env, rc);
  napi_create_typedarray(env, napi_float32_array, (size_t)n * WD_COARSE_BINS, ab, 0, &arr);
  return arr;
}

/* poseFeatures(uniform Float32Array(37)) -> Float32Array(12) */
static napi_value js_pose_features(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  void *u;
  size_t l;
  if (!get_args(env, info, 1, argv)) return NULL;
  if (!typed(env, argv[0], napi_float32_array, &u, &l) || l < sizeof(wd_view) / 4) return NULL;
  wd_view v;
  float f[WD_POSE_FLOATS];
  memcpy(&v, u, sizeof v);
  int rc = wd_pose_features(&v, f);
  return rc ? throw_wd(env, rc) : f32_array(env, f, WD_POSE_FLOATS);
}

/* stateQuantiles(context, state) -> Uint32Array(257) */
static napi_value js_state_quantiles(napi_env env, napi_callback_info info) {
  napi_value argv[2], ab, arr;
  void *data;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *s = get_box(env, argv[1], K_STATE);
  if (!s) return NULL;
  CHECK_NAPI(napi_create_arraybuffer(env, 257 * 4, &data, &ab));
  int rc = wd_state_quantiles((wd_context *)c->ptr, (wd_state *)s->ptr, (uint32_t *)data);
  if (rc) return throw_wd(env, rc);
  napi_create_typedarray(env, napi_uint32_array, 257, ab, 0, &arr);
  return arr;
}

/* learnedCreate(context, maxViews) -> learned */
static napi_value js_learned_create(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  uint32_t m;
  if (!get_args(env, info, 2, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c || !u32_arg(env, argv[1], &m)) return NULL;
  wd_learned *l = NULL;
  int rc = wd_learned_create((wd_context *)c->ptr, m, &l);
  return rc ? throw_wd(env, rc) : handle_of(env, K_LEARNED, l, argv[0], c);
}

/* learnedRun(context, learned, predictor, cloud, [uniforms], [states], cap) */
static napi_value js_learned_run(napi_env env, napi_callback_info info) {
  napi_value argv[7];
  uint32_t nv = 0, ns = 0, cap;
  wd_view views[8];
  void *states[8];
  if (!get_args(env, info, 7, argv)) return NULL;
  Box *c = get_box(env, argv[0], K_CTX);
  if (!c) return NULL;
  Box *l = get_box(env, argv[1], K_LEARNED);
  if (!l) return NULL;
  Box *m = get_box(env, argv[2], K_MLP);
  if (!m) return NULL;
  Box *cl = get_box(env, argv[3], K_CLOUD);
  if (!cl) return NULL;
  if (!view_list(env, argv[4], views, 8, &nv) || !handle_list(env, argv[5], K_STATE, states, 8, &ns) ||
      !u32_arg(env, argv[6], &cap))
    return NULL;
  if (nv != ns) {
    napi_throw_range_error(env, NULL, "learned: uniforms and states differ in number");
    return NULL;
  }
  int rc = wd_learned_run((wd_context *)c->ptr, (wd_learned *)l->ptr, (wd_mlp *)m->ptr, (const wd_cloud *)cl->ptr,
                          views, (wd_state *const *)states, nv, cap);
  return rc ? throw_wd(env, rc) : NULL;
}

static napi_value init(napi_env env, napi_value exports) {
  const struct {
    const char *name;
    napi_callback fn;
  } fns[] = {
      {"open", js_open},           {"close", js_close},           {"sync", js_sync},
      {"cloudCreate", js_cloud_create}, {"destroy", js_destroy},  {"viewPrepare", js_view_prepare},
      {"cloudSH", js_cloud_sh},         {"cloudCreateSplats", js_cloud_create_splats},
      {"decodePly", js_decode_ply},     {"decodeSpz", js_decode_spz}, {"decodeSog", js_decode_sog},
      {"stateCreate", js_state_create}, {"sort", js_sort},        {"stateRead", js_state_read},
      {"stateEllipses", js_state_ellipses}, {"stateReadEllipses", js_state_read_ellipses},
      {"poolCreate", js_pool_create},   {"particlesSort", js_particles_sort},
      {"synth", js_synth},         {"profile", js_profile},       {"profileRead", js_profile_read},
      {"batchCreate", js_batch_create}, {"batchSet", js_batch_set}, {"batchRun", js_batch_run},
      {"batchTune", js_batch_tune},     {"batchConfigure", js_batch_configure},
      {"stateKeysPerThread", js_state_keys_per_thread}, {"mlpCreate", js_mlp_create}, {"mlpSave", js_mlp_save},
      {"mlpLoad", js_mlp_load},         {"mlpTrain", js_mlp_train}, {"mlpPredict", js_mlp_predict},
      {"poseFeatures", js_pose_features}, {"stateQuantiles", js_state_quantiles},
      {"learnedCreate", js_learned_create}, {"learnedRun", js_learned_run},
  };
  for (size_t i = 0; i < sizeof fns / sizeof fns[0]; ++i) {
    napi_value f;
    if (napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f) != napi_ok) return NULL;
    napi_set_named_property(env, exports, fns[i].name, f);
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)

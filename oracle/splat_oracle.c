/*
 * splat_oracle.c -- CPU restatement of Winding's splat projection, key and
 * stable counting sort, in FP64. TEST INFRASTRUCTURE ONLY: the parity tests,
 * the smoke check and the benchmark's CPU-baseline leg load it as the answer
 * the GPU's results are compared with.
 * The product (libwinding_amd.so) never links, calls or falls back to it.
 *
 * Canonical definition (the reference's formulas, decided in FP64):
 *   Every quantity is computed in double from the f32 inputs (record floats,
 *   record bytes, the f32 uniform) with the reference's own expression order,
 *   no fused multiply-add (built with -ffp-contract=off), IEEE division and
 *   square root. Constants are the WGSL f32 literals widened: 1.3f, 0.3f,
 *   6.2831853f. The reference runs these formulas in f32 on whatever WebGPU
 *   stack it meets (src/render/splats.js:155-186), so its exact bits are not
 *   fixed anywhere; this file fixes them.
 *
 * Functions and what each follows (Winding v1.9.0 paths):
 *   or_sigma      put(): quaternion made unit, R, M = R S, Sigma = M M^T
 *                 src/scene/splats.js:350-368; length = sqrt(sum of squares)
 *                 as hypot3 takes it (src/core/math/vec3.js:92-94), exact for
 *                 byte quaternions, where put's Math.hypot may differ by an
 *                 ulp of f64 (never visible after the f32 rounding put does).
 *   or_footprint  footprint(): J W Sigma W^T J^T, src/render/splats.js:102-125
 *   or_splat      makeKeys() cull and key, src/render/splats.js:155-186
 *   or_splat_sigma, or_run_sigma  the same for a splat given as its centre
 *                 and f32 Sigma (the reference's Splats data), and the sort
 *   or_sort       stable one-pass counting sort over 65536 buckets, in the
 *                 idiom of DrawList.sort (src/render/drawlist.js:158-203);
 *                 ties keep splat index order (the reference's place() uses
 *                 atomics and does not, src/render/splats.js:214-219)
 *   or_bounds     put()'s finite check and min/max, src/scene/splats.js:336-348
 *   or_prepare    SplatPass.prepare's per-cloud uniform,
 *                 src/render/splats.js:559-587, with mat4Multiply
 *                 src/core/math/mat4.js:96-126
 *   or_look_at, or_perspective  mat4LookAt / mat4PerspectiveReverseZInfinite
 *                 src/core/math/mat4.js:243-273,455-468
 *   or_eye        the eye in a capture's space, as SplatPass.prepare sets it
 *                 (src/render/splats.js:585-586): the translation of mat4Invert
 *                 (src/core/math/mat4.js:186-228) of the f32 model-view, f64
 *                 maths stored as f32
 *   or_ellipse, or_ellipses  the splat vertex shader's per-instance set-up
 *                 (src/render/splats.js:263-291): eigen-axes of the footprint
 *                 plus 0.3, their lengths and cap, the reach; with the clip
 *                 position of each splat
 *   or_particle_keys, or_particle_edges  the particle sort's live test and
 *                 log-depth key (src/render/particles.js:198-211), the bucket
 *                 decided against FP64 edges (definition at the function)
 *   or_shade      the shade pass's colour of one splat with harmonics,
 *                 src/render/splats.js:335-381 (coefficient, harmonics,
 *                 shade), halves decoded exactly, in FP64; the `sh` layout
 *                 of src/scene/splats.js:306-332
 *
 * Pinned by: tests/test_oracle.py (known answers from test/splats.test.js:51-71,
 * test/core.test.js:214-223,343-360, test/render.test.js:428-440 and
 * test/gpu.test.js:1804-1962 restated as orders; closed forms; brute force).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct or_view {
  float model_view[16];
  float projection[16];
  float viewport[2];
  float depth_min;
  float depth_scale;
  float cull;
} or_view;

#define OR_BUCKETS 65536

uint32_t or_sort(const uint16_t *keys, const uint8_t *keep, uint64_t n, uint32_t *order);

static const double C13 = (double)1.3f;
static const double C03 = (double)0.3f;
static const double C2PI = (double)6.2831853f;

static float rd_f32(const uint8_t *p) {
  float f;
  memcpy(&f, p, 4);
  return f;
}

/* put(): src/scene/splats.js:350-368. Sigma as xx, xy, xz, yy, yz, zz. */
void or_sigma(const uint8_t *rec, double sigma[6]) {
  double sx = rd_f32(rec + 12), sy = rd_f32(rec + 16), sz = rd_f32(rec + 20);
  double qw = (rec[28] - 128) / 128.0, qx = (rec[29] - 128) / 128.0;
  double qy = (rec[30] - 128) / 128.0, qz = (rec[31] - 128) / 128.0;
  double length = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  if (length == 0) length = 1;
  double w = qw / length, u = qx / length, v = qy / length, t = qz / length;
  double r00 = 1 - 2 * (v * v + t * t), r01 = 2 * (u * v - w * t), r02 = 2 * (u * t + w * v);
  double r10 = 2 * (u * v + w * t), r11 = 1 - 2 * (u * u + t * t), r12 = 2 * (v * t - w * u);
  double r20 = 2 * (u * t - w * v), r21 = 2 * (v * t + w * u), r22 = 1 - 2 * (u * u + v * v);
  double m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
  double m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
  double m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
  sigma[0] = m00 * m00 + m01 * m01 + m02 * m02;
  sigma[1] = m00 * m10 + m01 * m11 + m02 * m12;
  sigma[2] = m00 * m20 + m01 * m21 + m02 * m22;
  sigma[3] = m10 * m10 + m11 * m11 + m12 * m12;
  sigma[4] = m10 * m20 + m11 * m21 + m12 * m22;
  sigma[5] = m20 * m20 + m21 * m21 + m22 * m22;
}

/* view = modelView * (p, 1), columns summed left to right (WGSL mat * vec). */
static void view_of(const or_view *v, double x, double y, double z, double out[4]) {
  const float *m = v->model_view;
  for (int r = 0; r < 4; r++) out[r] = (double)m[r] * x + (double)m[4 + r] * y + (double)m[8 + r] * z + (double)m[12 + r];
}

/*
 * footprint(): src/render/splats.js:102-125. T = J W (rows 0 and 1; J's third
 * row is zero), cov = (T Sigma) T^T.
 */
void or_footprint(const or_view *v, const double view[4], const double sg[6], double cov[3]) {
  const float *p = v->projection, *m = v->model_view;
  double p00 = p[0], p11 = p[5];
  double fx = p00 * (double)v->viewport[0] * 0.5;
  double fy = p11 * (double)v->viewport[1] * 0.5;
  double j00 = fx, j11 = fy, j02 = 0, j12 = 0;
  if (p[15] == 0.0f) {
    double d = -view[2];
    double lx = C13 / p00, ly = C13 / p11;
    double cx = view[0] / d, cy = view[1] / d;
    if (cx < -lx) cx = -lx;
    if (cx > lx) cx = lx;
    if (cy < -ly) cy = -ly;
    if (cy > ly) cy = ly;
    double x = cx * d, y = cy * d;
    j00 = fx / d;
    j11 = fy / d;
    j02 = fx * x / (d * d);
    j12 = fy * y / (d * d);
  }
  /* W[k][c] = modelView[c*4 + k] */
  double t0[3], t1[3];
  for (int c = 0; c < 3; c++) {
    t0[c] = j00 * (double)m[c * 4 + 0] + j02 * (double)m[c * 4 + 2];
    t1[c] = j11 * (double)m[c * 4 + 1] + j12 * (double)m[c * 4 + 2];
  }
  double S[3][3] = {{sg[0], sg[1], sg[2]}, {sg[1], sg[3], sg[4]}, {sg[2], sg[4], sg[5]}};
  double u0[3], u1[3];
  for (int c = 0; c < 3; c++) {
    u0[c] = t0[0] * S[0][c] + t0[1] * S[1][c] + t0[2] * S[2][c];
    u1[c] = t1[0] * S[0][c] + t1[1] * S[1][c] + t1[2] * S[2][c];
  }
  cov[0] = u0[0] * t0[0] + u0[1] * t0[1] + u0[2] * t0[2];
  cov[1] = u0[0] * t1[0] + u0[1] * t1[1] + u0[2] * t1[2];
  cov[2] = u1[0] * t1[0] + u1[1] * t1[1] + u1[2] * t1[2];
}

/*
 * makeKeys() for one record: src/render/splats.js:157-185. Returns 1 for a
 * survivor and writes its key; cov (may be null) gets the FP64 footprint
 * whenever the splat passes the frustum and alpha tests; *value (may be null)
 * gets alpha * area when the area test ran, else NAN.
 */
int or_splat(const uint8_t *rec, const or_view *v, uint16_t *key, double cov[3], double *value) {
  double px = rd_f32(rec), py = rd_f32(rec + 4), pz = rd_f32(rec + 8);
  double view[4], clip[4];
  view_of(v, px, py, pz, view);
  const float *P = v->projection;
  for (int r = 0; r < 4; r++)
    clip[r] = (double)P[r] * view[0] + (double)P[4 + r] * view[1] + (double)P[8 + r] * view[2] + (double)P[12 + r] * view[3];
  double w = clip[3];
  if (value) *value = NAN;
  if (view[2] >= 0 || clip[2] > w || fabs(clip[0]) > C13 * w || fabs(clip[1]) > C13 * w) return 0;
  unsigned a = rec[27];
  if (a <= 1) return 0;
  double f[3];
  double sg[6];
  or_sigma(rec, sg);
  /* Sigma as put() stores it: each term rounded to f32 (a Float32Array, src/scene/splats.js:318,361-367) */
  for (int k = 0; k < 6; k++) sg[k] = (double)(float)sg[k];
  or_footprint(v, view, sg, f);
  if (cov) memcpy(cov, f, sizeof f);
  double cull = v->cull;
  if (cull > 0) {
    double det = (f[0] + C03) * (f[2] + C03) - f[1] * f[1];
    double area = C2PI * sqrt(det > 0 ? det : 0);
    double alpha = a / 255.0;
    if (value) *value = alpha * area;
    if (alpha * area < cull) return 0;
  }
  double t = (-view[2] - (double)v->depth_min) * (double)v->depth_scale;
  if (t < 0) t = 0;
  if (t > 65535.0) t = 65535.0;
  uint32_t bucket = (uint32_t)t;
  *key = (uint16_t)(65535u - bucket);
  return 1;
}

/*
 * makeKeys() for a splat given as its centre (4 floats, colour bits in the fourth) and its f32
 * Sigma (xx xy xz yy yz zz), as the reference's Splats data holds it: the same decisions as
 * or_splat with Sigma read instead of built from a record.
 */
int or_splat_sigma(const float *centre, const float *sigma6, const or_view *v, uint16_t *key, double cov[3]) {
  double view[4], clip[4];
  view_of(v, centre[0], centre[1], centre[2], view);
  const float *P = v->projection;
  for (int r = 0; r < 4; r++)
    clip[r] = (double)P[r] * view[0] + (double)P[4 + r] * view[1] + (double)P[8 + r] * view[2] + (double)P[12 + r] * view[3];
  double w = clip[3];
  if (view[2] >= 0 || clip[2] > w || fabs(clip[0]) > C13 * w || fabs(clip[1]) > C13 * w) return 0;
  uint32_t word;
  memcpy(&word, &centre[3], 4);
  unsigned a = word >> 24;
  if (a <= 1) return 0;
  double sg[6], f[3];
  for (int k = 0; k < 6; k++) sg[k] = sigma6[k];
  or_footprint(v, view, sg, f);
  if (cov) memcpy(cov, f, sizeof f);
  if (v->cull > 0) {
    double det = (f[0] + C03) * (f[2] + C03) - f[1] * f[1];
    double area = C2PI * sqrt(det > 0 ? det : 0);
    if (a / 255.0 * area < (double)v->cull) return 0;
  }
  double t = (-view[2] - (double)v->depth_min) * (double)v->depth_scale;
  if (t < 0) t = 0;
  if (t > 65535.0) t = 65535.0;
  *key = (uint16_t)(65535u - (uint32_t)t);
  return 1;
}

/* or_splat_sigma over a cloud, then the stable sort: returns V; keys, keep and cov (n x 3) per splat. */
uint32_t or_run_sigma(const float *centers, const float *covs, uint64_t n, const or_view *v, uint32_t *order,
                      uint16_t *keys, uint8_t *keep, double *cov) {
  for (uint64_t i = 0; i < n; i++) {
    uint16_t k = 0;
    double c[3] = {NAN, NAN, NAN};
    int s = or_splat_sigma(centers + 4 * i, covs + 6 * i, v, &k, c);
    keys[i] = s ? k : 0;
    keep[i] = (uint8_t)s;
    if (cov) memcpy(cov + 3 * i, c, sizeof c);
  }
  return or_sort(keys, keep, n, order);
}

/* Every splat of a cloud: keys (culled get 0), keep flags, FP64 cov (n x 3, may be null), value (may be null). */
void or_keys(const uint8_t *records, uint64_t n, const or_view *v, uint16_t *keys, uint8_t *keep,
             double *cov, double *value) {
  for (uint64_t i = 0; i < n; i++) {
    uint16_t k = 0;
    double c[3] = {NAN, NAN, NAN}, val;
    int s = or_splat(records + i * 32, v, &k, c, &val);
    keys[i] = s ? k : 0;
    keep[i] = (uint8_t)s;
    if (cov) memcpy(cov + 3 * i, c, sizeof c);
    if (value) value[i] = val;
  }
}

/*
 * Stable counting sort of the survivors by key, ties by index: one pass,
 * since the key fits 65536 buckets. Returns the survivor count.
 */
uint32_t or_sort(const uint16_t *keys, const uint8_t *keep, uint64_t n, uint32_t *order) {
  uint32_t *counts = (uint32_t *)calloc(OR_BUCKETS, sizeof(uint32_t));
  if (!counts) return 0xffffffffu;
  for (uint64_t i = 0; i < n; i++)
    if (keep[i]) counts[keys[i]]++;
  uint32_t total = 0;
  for (uint32_t b = 0; b < OR_BUCKETS; b++) {
    uint32_t c = counts[b];
    counts[b] = total;
    total += c;
  }
  for (uint64_t i = 0; i < n; i++)
    if (keep[i]) order[counts[keys[i]]++] = (uint32_t)i;
  free(counts);
  return total;
}

/* The whole path for one view: returns V; order must hold n. cov n x 3 (may be null). */
uint32_t or_run(const uint8_t *records, uint64_t n, const or_view *v, uint32_t *order, uint16_t *keys,
                uint8_t *keep, double *cov) {
  or_keys(records, n, v, keys, keep, cov, NULL);
  return or_sort(keys, keep, n, order);
}

/* Finite check and bounds, as put(): returns the first bad index, or -1. */
int64_t or_bounds(const uint8_t *records, uint64_t n, float mn[3], float mx[3]) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = 0; i < n; i++) {
    double p[3];
    for (int k = 0; k < 3; k++) p[k] = rd_f32(records + i * 32 + 4 * k);
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))) return (int64_t)i;
    for (int k = 0; k < 3; k++) {
      if (p[k] < lo[k]) lo[k] = p[k];
      if (p[k] > hi[k]) hi[k] = p[k];
    }
  }
  for (int k = 0; k < 3; k++) {
    mn[k] = (float)lo[k];
    mx[k] = (float)hi[k];
  }
  return -1;
}

/* mat4Multiply(out, a, b): out = a * b, each entry summed in f64 and stored as f32. */
void or_mat4_multiply(float out[16], const float a[16], const float b[16]) {
  float r[16];
  for (int c = 0; c < 4; c++) {
    double b0 = b[c * 4], b1 = b[c * 4 + 1], b2 = b[c * 4 + 2], b3 = b[c * 4 + 3];
    for (int row = 0; row < 4; row++)
      r[c * 4 + row] = (float)(b0 * a[row] + b1 * a[4 + row] + b2 * a[8 + row] + b3 * a[12 + row]);
  }
  memcpy(out, r, sizeof r);
}

/* SplatPass.prepare for one cloud. Returns 1 when the cloud is wholly behind (far <= 0). */
int or_prepare(const float view[16], const float world[16], const float proj[16], float width, float height,
               const float mn[3], const float mx[3], float cull, or_view *out) {
  static const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  float mv[16];
  or_mat4_multiply(mv, view, world ? world : identity);
  double near = INFINITY, far = -INFINITY;
  for (int k = 0; k < 8; k++) {
    double x = (k & 1) ? mx[0] : mn[0];
    double y = (k & 2) ? mx[1] : mn[1];
    double z = (k & 4) ? mx[2] : mn[2];
    double depth = -((double)mv[2] * x + (double)mv[6] * y + (double)mv[10] * z + (double)mv[14]);
    if (depth < near) near = depth;
    if (depth > far) far = depth;
  }
  if (far <= 0) return 1;
  near = near > 0 ? near : 0;
  memcpy(out->model_view, mv, sizeof mv);
  memcpy(out->projection, proj, 16 * sizeof(float));
  out->viewport[0] = width;
  out->viewport[1] = height;
  out->depth_min = (float)near;
  double span = far - near;
  out->depth_scale = (float)(65535.0 / (span > 1e-6 ? span : 1e-6));
  out->cull = cull;
  return 0;
}

/* mat4LookAt: src/core/math/mat4.js:243-273 (hypot3 = sqrt of the sum of squares). */
void or_look_at(float out[16], const double eye[3], const double center[3], const double up[3]) {
  double zx = eye[0] - center[0], zy = eye[1] - center[1], zz = eye[2] - center[2];
  double zl = sqrt(zx * zx + zy * zy + zz * zz);
  zx /= zl; zy /= zl; zz /= zl;
  double xx = up[1] * zz - up[2] * zy, xy = up[2] * zx - up[0] * zz, xz = up[0] * zy - up[1] * zx;
  double xl = sqrt(xx * xx + xy * xy + xz * xz);
  xx /= xl; xy /= xl; xz /= xl;
  double yx = zy * xz - zz * xy, yy = zz * xx - zx * xz, yz = zx * xy - zy * xx;
  out[0] = (float)xx; out[1] = (float)yx; out[2] = (float)zx; out[3] = 0;
  out[4] = (float)xy; out[5] = (float)yy; out[6] = (float)zy; out[7] = 0;
  out[8] = (float)xz; out[9] = (float)yz; out[10] = (float)zz; out[11] = 0;
  out[12] = (float)(-(xx * eye[0] + xy * eye[1] + xz * eye[2]));
  out[13] = (float)(-(yx * eye[0] + yy * eye[1] + yz * eye[2]));
  out[14] = (float)(-(zx * eye[0] + zy * eye[1] + zz * eye[2]));
  out[15] = 1;
}

/* The eye: the translation column of mat4Invert(mv) (src/core/math/mat4.js:186-228), cofactors in f64 in
 * the reference's expression order, stored as f32. Returns 1 (eye left at the origin) for a singular matrix. */
int or_eye(const float m[16], float eye[3]) {
  double a00 = m[0], a01 = m[1], a02 = m[2], a03 = m[3], a10 = m[4], a11 = m[5], a12 = m[6], a13 = m[7];
  double a20 = m[8], a21 = m[9], a22 = m[10], a23 = m[11], a30 = m[12], a31 = m[13], a32 = m[14], a33 = m[15];
  double b00 = a00 * a11 - a01 * a10, b01 = a00 * a12 - a02 * a10, b02 = a00 * a13 - a03 * a10;
  double b03 = a01 * a12 - a02 * a11, b04 = a01 * a13 - a03 * a11, b05 = a02 * a13 - a03 * a12;
  double b06 = a20 * a31 - a21 * a30, b07 = a20 * a32 - a22 * a30, b08 = a20 * a33 - a23 * a30;
  double b09 = a21 * a32 - a22 * a31, b10 = a21 * a33 - a23 * a31, b11 = a22 * a33 - a23 * a32;
  double det = b00 * b11 - b01 * b10 + b02 * b09 + b03 * b08 - b04 * b07 + b05 * b06;
  eye[0] = eye[1] = eye[2] = 0;
  if (det == 0) return 1;
  det = 1 / det;
  eye[0] = (float)((a11 * b07 - a10 * b09 - a12 * b06) * det);
  eye[1] = (float)((a00 * b09 - a01 * b07 + a02 * b06) * det);
  eye[2] = (float)((a31 * b01 - a30 * b03 - a32 * b00) * det);
  return 0;
}

/* An IEEE half's value, exactly. */
static double half_value(uint16_t h) {
  int e = (h >> 10) & 31, m = h & 1023;
  double v = e == 0 ? ldexp((double)m, -24) : e == 31 ? (m ? NAN : INFINITY) : ldexp((double)(1024 + m), e - 25);
  return (h & 0x8000) ? -v : v;
}

static const int OR_SH_COEFFICIENTS[4] = {0, 3, 8, 15};

/* Words a splat's harmonics take (shWords, src/scene/splats.js:47). */
int or_sh_words(int degree) { return (OR_SH_COEFFICIENTS[degree] * 3 + 1) / 2; }

/* coefficient(): the three halves of coefficient k of a splat whose words start at w (src/render/splats.js:338-344). */
static void coefficient(const uint32_t *w, int k, double c[3]) {
  int h = k * 3;
  uint32_t a = w[h >> 1], b = w[(h >> 1) + 1];
  if (h & 1) {
    c[0] = half_value((uint16_t)(a >> 16)); c[1] = half_value((uint16_t)b); c[2] = half_value((uint16_t)(b >> 16));
  } else {
    c[0] = half_value((uint16_t)a); c[1] = half_value((uint16_t)(a >> 16)); c[2] = half_value((uint16_t)b);
  }
}

/* The shade pass for one splat (src/render/splats.js:346-381), FP64, the f32 literals widened: colour bytes,
 * alpha kept. rec: the 32-byte record; w: the splat's sh_words(degree) words; eye as or_eye gives it. */
uint32_t or_shade(const uint8_t *rec, const uint32_t *w, int degree, const float eye[3]) {
  double dx = (double)rd_f32(rec) - eye[0], dy = (double)rd_f32(rec + 4) - eye[1], dz = (double)rd_f32(rec + 8) - eye[2];
  double len = sqrt(dx * dx + dy * dy + dz * dz);
  double x = dx / len, y = dy / len, z = dz / len;
  double c[3], k[15][3];
  for (int j = 0; j < OR_SH_COEFFICIENTS[degree]; j++) coefficient(w, j, k[j]);
  for (int ch = 0; ch < 3; ch++) c[ch] = (double)0.4886025119029199f * (-y * k[0][ch] + z * k[1][ch] - x * k[2][ch]);
  if (degree >= 2) {
    double xx = x * x, yy = y * y, zz = z * z;
    for (int ch = 0; ch < 3; ch++)
      c[ch] += (double)1.0925484305920792f * x * y * k[3][ch] - (double)1.0925484305920792f * y * z * k[4][ch] +
               (double)0.31539156525252005f * (2.0 * zz - xx - yy) * k[5][ch] -
               (double)1.0925484305920792f * x * z * k[6][ch] + (double)0.5462742152960396f * (xx - yy) * k[7][ch];
    if (degree >= 3)
      for (int ch = 0; ch < 3; ch++)
        c[ch] += -(double)0.5900435899266435f * y * (3.0 * xx - yy) * k[8][ch] +
                 (double)2.890611442640554f * x * y * z * k[9][ch] -
                 (double)0.4570457994644658f * y * (4.0 * zz - xx - yy) * k[10][ch] +
                 (double)0.3731763325901154f * z * (2.0 * zz - 3.0 * xx - 3.0 * yy) * k[11][ch] -
                 (double)0.4570457994644658f * x * (4.0 * zz - xx - yy) * k[12][ch] +
                 (double)1.445305721320277f * z * (xx - yy) * k[13][ch] -
                 (double)0.5900435899266435f * x * (xx - 3.0 * yy) * k[14][ch];
  }
  uint32_t out = (uint32_t)rec[27] << 24;
  for (int ch = 0; ch < 3; ch++) {
    double seen = rec[24 + ch] / 255.0 + c[ch];
    seen = seen < 0 ? 0 : seen > 1 ? 1 : seen;
    out |= (uint32_t)floor(0.5 + 255.0 * seen) << (8 * ch);
  }
  return out;
}

/* or_shade for splats 0..n-1 of a cloud (sh: sh_words(degree) words a splat). */
void or_shade_all(const uint8_t *records, uint64_t n, const uint32_t *sh, int degree, const float eye[3],
                  uint32_t *out) {
  int W = or_sh_words(degree);
  for (uint64_t i = 0; i < n; i++) out[i] = or_shade(records + 32 * i, sh + (uint64_t)W * i, degree, eye);
}

/*
 * The splat vertex shader's per-instance set-up (src/render/splats.js:263-291)
 * for a splat with FP64 footprint f = (xx, xy, yy) and alpha byte abyte, in
 * FP64, the f32 literals widened: out = (along e, across e), the eigen-axes
 * of [[xx + 0.3, xy], [xy, yy + 0.3]] as long as sqrt(2 lambda) capped at 1024,
 * across = along turned a quarter clockwise, e = sqrt(min(log(255 alpha), 4)).
 * Returns 0 (out zeros) when the shader draws nothing: minor axis or reach
 * not above zero.
 */
int or_ellipse(const double f[3], unsigned abyte, double out[4]) {
  double a = f[0] + C03, b = f[1], c = f[2] + C03;
  double mid = 0.5 * (a + c);
  double h = 0.5 * (a - c);
  double radius = sqrt(h * h + b * b);
  double major = mid + radius, minor = mid - radius;
  double reach = log(255.0 * (abyte / 255.0));
  out[0] = out[1] = out[2] = out[3] = 0;
  if (!(minor > 0) || !(reach > 0)) return 0;
  double ax = 1, ay = 0;
  if (fabs(b) > (double)1e-12f) {
    double uy = major - a, len = sqrt(b * b + uy * uy);
    ax = b / len;
    ay = uy / len;
  } else if (c > a) {
    ax = 0;
    ay = 1;
  }
  double e = sqrt(reach < 4 ? reach : 4);
  double la = sqrt(2 * major), lc = sqrt(2 * minor);
  la = (la < 1024 ? la : 1024) * e;
  lc = (lc < 1024 ? lc : 1024) * e;
  out[0] = la * ax;
  out[1] = la * ay;
  out[2] = lc * ay;
  out[3] = -(lc * ax);
  return 1;
}

/* The draw set-up of every splat of a cloud, 8 doubles a splat: its clip position (modelView, then projection,
 * as vs and makeKeys compute it), then or_ellipse; all zero for a splat culled or not drawn. */
void or_ellipses(const uint8_t *records, uint64_t n, const or_view *v, double *out) {
  for (uint64_t i = 0; i < n; i++) {
    const uint8_t *rec = records + 32 * i;
    double *o = out + 8 * i;
    uint16_t k;
    double f[3];
    memset(o, 0, 8 * sizeof(double));
    if (!or_splat(rec, v, &k, f, NULL)) continue;
    if (!or_ellipse(f, rec[27], o + 4)) continue;
    double view[4];
    view_of(v, rd_f32(rec), rd_f32(rec + 4), rd_f32(rec + 8), view);
    const float *P = v->projection;
    for (int r = 0; r < 4; r++)
      o[r] = (double)P[r] * view[0] + (double)P[4 + r] * view[1] + (double)P[8 + r] * view[2] + (double)P[12 + r] * view[3];
  }
}

/*
 * Particles (src/render/particles.js:169-240). A particle is 32 bytes: position (3 f32), age,
 * velocity (3 f32), lifetime. A live one (age < lifetime) gets the key 4095 - bucket, far
 * first, with bucket = floor((log2(depth) - LOG_NEAR) (BUCKETS - 1) / (LOG_FAR - LOG_NEAR))
 * clamped to [0, 4095] (LOG_NEAR = -10, LOG_FAR = 20, BUCKETS = 4096), and
 * depth = max(dot(position - eye, forward), 1e-6), makeKeys at :198-211.
 *
 * Canonical definition: depth in FP64 from the f32 inputs (the three products summed left to
 * right, the clamp at the f32 literal 1e-6f; a NaN depth takes the clamp). The bucket is the
 * number of edges E[b], b = 1..4095, at or below depth, where E[b] = exp2((30 b - 40950) / 4095)
 * in FP64: floor of the formula above with log2 taken exactly, up to the rounding of each edge
 * to a double, so that the key does not depend on how a log2 rounds.
 */
#define OR_PARTICLE_BUCKETS 4096

void or_particle_edges(double *edges) {
  edges[0] = 0;
  for (int b = 1; b < OR_PARTICLE_BUCKETS; b++) edges[b] = exp2((30.0 * b - 40950.0) / 4095.0);
}

/* Per particle of a pool of n: key (0 when dead) and keep (1 when live). view: eye (3 f32), forward (3 f32). */
void or_particle_keys(const uint8_t *pool, uint64_t n, const float eye[3], const float forward[3], uint16_t *keys,
                      uint8_t *keep) {
  double edges[OR_PARTICLE_BUCKETS];
  or_particle_edges(edges);
  for (uint64_t i = 0; i < n; i++) {
    const uint8_t *p = pool + 32 * i;
    float age = rd_f32(p + 12), lifetime = rd_f32(p + 28);
    keys[i] = 0;
    keep[i] = 0;
    if (age >= lifetime) continue; /* as the shader tests it: a NaN age or lifetime is live */
    double d = ((double)rd_f32(p) - eye[0]) * forward[0] + ((double)rd_f32(p + 4) - eye[1]) * forward[1] +
               ((double)rd_f32(p + 8) - eye[2]) * forward[2];
    if (!(d >= (double)1e-6f)) d = (double)1e-6f;
    int lo = 0, hi = OR_PARTICLE_BUCKETS - 1; /* the largest b with edges[b] <= d */
    while (lo < hi) {
      int mid = (lo + hi + 1) / 2;
      if (edges[mid] <= d) lo = mid; else hi = mid - 1;
    }
    keys[i] = (uint16_t)(OR_PARTICLE_BUCKETS - 1 - lo);
    keep[i] = 1;
  }
}

/* mat4PerspectiveReverseZInfinite: src/core/math/mat4.js:455-468. */
void or_perspective(float out[16], double fovy, double aspect, double near) {
  double f = 1 / tan(fovy * 0.5);
  memset(out, 0, 16 * sizeof(float));
  out[0] = (float)(f / aspect);
  out[5] = (float)f;
  out[11] = -1;
  out[14] = (float)near;
}

/*
 * asan_check.c -- drives every entry point of the two C oracles (splat_oracle.c,
 * decode_oracle.c) over seeded inputs sized exactly, so that a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer reports any read or write past
 * a buffer, use of freed memory or undefined arithmetic. TEST INFRASTRUCTURE
 * ONLY, like the oracles it drives. Built and run by tests/test_oracle_asan.py:
 *
 *   gcc -O1 -g -fsanitize=address,undefined -fno-omit-frame-pointer -ffp-contract=off \
 *       oracle/asan_check.c oracle/splat_oracle.c oracle/decode_oracle.c -lm -o oracle/_build/asan_check
 *
 * Prints one line per group and exits 0; a sanitizer report ends it earlier.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
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

void or_sigma(const uint8_t *rec, double sigma[6]);
void or_keys(const uint8_t *records, uint64_t n, const or_view *v, uint16_t *keys, uint8_t *keep, double *cov,
             double *value);
uint32_t or_sort(const uint16_t *keys, const uint8_t *keep, uint64_t n, uint32_t *order);
uint32_t or_run(const uint8_t *records, uint64_t n, const or_view *v, uint32_t *order, uint16_t *keys, uint8_t *keep,
                double *cov);
uint32_t or_run_sigma(const float *centers, const float *covs, uint64_t n, const or_view *v, uint32_t *order,
                      uint16_t *keys, uint8_t *keep, double *cov);
int64_t or_bounds(const uint8_t *records, uint64_t n, float mn[3], float mx[3]);
int or_prepare(const float view[16], const float world[16], const float proj[16], float width, float height,
               const float mn[3], const float mx[3], float cull, or_view *out);
void or_look_at(float out[16], const double eye[3], const double center[3], const double up[3]);
void or_perspective(float out[16], double fovy, double aspect, double near);
int or_eye(const float m[16], float eye[3]);
int or_sh_words(int degree);
void or_shade_all(const uint8_t *records, uint64_t n, const uint32_t *sh, int degree, const float eye[3],
                  uint32_t *out);
void or_ellipses(const uint8_t *records, uint64_t n, const or_view *v, double *out);
void or_particle_keys(const uint8_t *pool, uint64_t n, const float eye[3], const float forward[3], uint16_t *keys,
                      uint8_t *keep);
int64_t od_decode_ply(const uint8_t *body, uint64_t n, uint32_t stride, const uint32_t *off, const uint8_t *type,
                      uint32_t rest, int degree, float *centers, float *covs, uint16_t *halves);
int64_t od_decode_spz(const uint8_t *data, uint64_t len, uint32_t *count_out, int *degree_out, float *centers,
                      float *covs, uint16_t *halves);
int64_t od_decode_sog(uint32_t count, int degree, uint32_t sh_count, const double *mins, const double *maxs,
                      const uint8_t *means_l, const uint8_t *means_u, const uint8_t *quats, const uint8_t *scales,
                      const uint8_t *sh0, const uint8_t *labels, const uint8_t *centroids, uint32_t centroids_width,
                      uint64_t centroids_bytes, const double *scale_book, const double *colour_book,
                      const double *sh_book, float *centers, float *covs, uint16_t *halves);

static uint64_t state = 0x2545F4914F6CDD1Dull;
static uint32_t next32(void) {
  state ^= state << 13;
  state ^= state >> 7;
  state ^= state << 17;
  return (uint32_t)(state >> 16);
}
static float unit(void) { return (next32() & 0xFFFFFF) / 16777216.0f; }

static void *need(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

/* n .splat records: positions in a box, small scales, every byte of colour and quaternion random. */
static uint8_t *records(uint64_t n) {
  uint8_t *r = need(32 * n);
  for (uint64_t i = 0; i < n; i++) {
    float f[6] = {unit() * 8 - 4, unit() * 8 - 4, unit() * 8 - 4, 0.01f + unit() * 0.2f, 0.01f + unit() * 0.2f,
                  0.01f + unit() * 0.2f};
    memcpy(r + 32 * i, f, sizeof f);
    for (int k = 24; k < 32; k++) r[32 * i + k] = (uint8_t)next32();
  }
  return r;
}

static void splat_path(uint64_t n, int off_axis, float cull) {
  uint8_t *r = records(n);
  float view[16], proj[16], mn[3] = {-4, -4, -4}, mx[3] = {4, 4, 4};
  const double eye[3] = {0.5, 1.0, 9.0}, at[3] = {0, 0, 0}, up[3] = {0, 1, 0};
  or_look_at(view, eye, at, up);
  or_perspective(proj, 1.0, 16.0 / 9.0, 0.1);
  if (off_axis) proj[8] = 0.6f;
  if (n) or_bounds(r, n, mn, mx);
  or_view v;
  if (or_prepare(view, NULL, proj, 1280, 720, mn, mx, cull, &v) == 1) {
    free(r);
    return;
  }
  uint16_t *keys = need(2 * n);
  uint8_t *keep = need(n);
  double *cov = need(24 * n), *val = need(8 * n), *ell = need(64 * n);
  uint32_t *order = need(4 * n);
  or_keys(r, n, &v, keys, keep, cov, val);
  uint32_t V = or_sort(keys, keep, n, order);
  uint32_t V2 = or_run(r, n, &v, order, keys, keep, cov);
  or_ellipses(r, n, &v, ell);
  float *centers = need(16 * n), *covs = need(24 * n);
  for (uint64_t i = 0; i < n; i++) {
    double s[6];
    memcpy(centers + 4 * i, r + 32 * i, 12);
    memcpy(centers + 4 * i + 3, r + 32 * i + 24, 4);
    or_sigma(r + 32 * i, s);
    for (int k = 0; k < 6; k++) covs[6 * i + k] = (float)s[k];
  }
  uint32_t V3 = or_run_sigma(centers, covs, n, &v, order, keys, keep, cov);
  float e[3];
  or_eye(v.model_view, e);
  for (int degree = 1; degree <= 3; degree++) {
    const int W = or_sh_words(degree);
    uint32_t *sh = need(4 * (n * W + 1)), *shaded = need(4 * n);
    for (uint64_t k = 0; k < n * W + 1; k++) sh[k] = next32() & 0x3BFF3BFFu;
    or_shade_all(r, n, sh, degree, e, shaded);
    free(sh);
    free(shaded);
  }
  printf("splats n=%llu off_axis=%d cull=%g: V %u %u %u\n", (unsigned long long)n, off_axis, cull, V, V2, V3);
  free(centers);
  free(covs);
  free(keys);
  free(keep);
  free(cov);
  free(val);
  free(ell);
  free(order);
  free(r);
}

static void bounds_refusal(void) {
  uint8_t *r = records(5);
  const float nan = NAN;
  memcpy(r + 32 * 3 + 4, &nan, 4);
  float mn[3], mx[3];
  printf("bounds: first bad %lld\n", (long long)or_bounds(r, 5, mn, mx));
  free(r);
}

static void particles(uint64_t n) {
  uint8_t *pool = need(32 * n);
  for (uint64_t i = 0; i < n; i++) {
    float f[8] = {unit() * 40 - 20, unit() * 40 - 20, unit() * 200 - 100, unit() * 4,
                  unit(),           unit(),           unit(),            (i % 7) ? unit() * 4 : NAN};
    memcpy(pool + 32 * i, f, sizeof f);
  }
  const float eye[3] = {0, 0, 100}, fwd[3] = {0, 0, -1};
  uint16_t *keys = need(2 * n);
  uint8_t *keep = need(n);
  uint32_t *order = need(4 * n);
  or_particle_keys(pool, n, eye, fwd, keys, keep);
  printf("particles n=%llu: live %u\n", (unsigned long long)n, or_sort(keys, keep, n, order));
  free(pool);
  free(keys);
  free(keep);
  free(order);
}

static void ply(uint64_t n, uint32_t rest, int degree) {
  /* float properties in the order training writes them, then f_rest_, then three normals */
  const uint32_t props = 14 + rest + 3, stride = 4 * props;
  uint8_t *body = need((size_t)stride * n);
  for (uint64_t i = 0; i < n * props; i++) {
    const float x = unit() * 6 - 3;
    memcpy(body + 4 * i, &x, 4);
  }
  uint32_t off[14 + 45];
  uint8_t type[14 + 45];
  for (uint32_t k = 0; k < 14 + rest; k++) {
    off[k] = 4 * k;
    type[k] = 6;
  }
  const int W = degree ? or_sh_words(degree) : 0;
  float *centers = need(16 * n), *covs = need(24 * n);
  uint16_t *halves = need(2 * (2 * (size_t)W * n + 2));
  printf("ply n=%llu degree=%d: %lld\n", (unsigned long long)n, degree,
         (long long)od_decode_ply(body, n, stride, off, type, rest, degree, centers, covs, halves));
  free(body);
  free(centers);
  free(covs);
  free(halves);
}

static void spz(uint32_t version, uint32_t count, int file_degree, int cut) {
  static const int coefs[5] = {0, 3, 8, 15, 24};
  const uint64_t pos = version == 1 ? 6 : 9, rot = version >= 3 ? 4 : 3;
  const uint64_t body = (uint64_t)count * (pos + 1 + 3 + 3 + rot + 3 * (uint64_t)coefs[file_degree]);
  const uint64_t len = 16 + body - (cut ? 1 : 0);
  uint8_t *data = need(len);
  for (uint64_t i = 0; i < len; i++) data[i] = (uint8_t)next32();
  const uint32_t head[3] = {0x5053474eu, version, count};
  memcpy(data, head, 12);
  data[12] = (uint8_t)file_degree;
  data[13] = 12;
  uint32_t n = 0;
  int degree = 0;
  int64_t rc = od_decode_spz(data, len, &n, &degree, NULL, NULL, NULL);
  if (rc == -1) {
    const int W = degree ? or_sh_words(degree) : 0;
    float *centers = need(16 * (size_t)n), *covs = need(24 * (size_t)n);
    uint16_t *halves = need(2 * (2 * (size_t)W * n + 2));
    rc = od_decode_spz(data, len, &n, &degree, centers, covs, halves);
    free(centers);
    free(covs);
    free(halves);
  }
  printf("spz v%u count=%u degree=%d cut=%d: %lld\n", version, count, file_degree, cut, (long long)rc);
  free(data);
}

static void sog(uint32_t count, int degree) {
  static const int coefs[4] = {0, 3, 8, 15};
  const int per = coefs[degree];
  uint8_t *img[6];
  for (int k = 0; k < 6; k++) {
    img[k] = need(4 * (size_t)count);
    for (uint32_t i = 0; i < 4 * count; i++) img[k][i] = (uint8_t)next32();
  }
  for (uint32_t i = 0; i < count; i++) img[2][4 * i + 3] = (uint8_t)(248 + next32() % 8); /* the quaternion tags */
  const uint32_t palette = 100, width = degree ? 64 * per : 1;
  const uint64_t cbytes = degree ? 4ull * width * ((palette + 63) / 64) : 4;
  uint8_t *centroids = need(cbytes);
  for (uint64_t i = 0; i < cbytes; i++) centroids[i] = (uint8_t)next32();
  double books[3][256], mins[3] = {-2, -2, -2}, maxs[3] = {2, 2, 0};
  for (int b = 0; b < 3; b++)
    for (int k = 0; k < 256; k++) books[b][k] = k < 200 ? (double)k / 64 - 2 : NAN;
  const int W = degree ? or_sh_words(degree) : 0;
  float *centers = need(16 * (size_t)count), *covs = need(24 * (size_t)count);
  uint16_t *halves = need(2 * (2 * (size_t)W * count + 2));
  printf("sog count=%u degree=%d: %lld\n", count, degree,
         (long long)od_decode_sog(count, degree, palette, mins, maxs, img[0], img[1], img[2], img[3], img[4], img[5],
                                  centroids, width, cbytes, books[0], books[1], books[2], centers, covs, halves));
  for (int k = 0; k < 6; k++) free(img[k]);
  free(centroids);
  free(centers);
  free(covs);
  free(halves);
}

int main(void) {
  const uint64_t sizes[] = {0, 1, 63, 4097};
  for (int k = 0; k < 4; k++) {
    splat_path(sizes[k], 0, 0.5f);
    splat_path(sizes[k], 1, 0.0f);
  }
  bounds_refusal();
  particles(1);
  particles(5000);
  ply(1, 0, 0);
  ply(300, 9, 1);
  ply(300, 45, 3);
  for (uint32_t v = 1; v <= 3; v++) {
    spz(v, 1, 0, 0);
    spz(v, 200, 3, 0);
    spz(v, 200, 4, 0);
    spz(v, 200, 2, 1);
  }
  sog(1, 0);
  sog(500, 2);
  sog(500, 3);
  printf("asan check: done\n");
  return 0;
}

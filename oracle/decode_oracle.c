/*
 * decode_oracle.c -- CPU restatement of Winding's capture decoders, in FP64
 * as the reference runs them in JavaScript numbers. TEST INFRASTRUCTURE ONLY:
 * the parity tests compare the GPU's decodes with it; the product never links it.
 *
 * Functions and what each follows (Winding v1.9.0, src/scene/splats.js):
 *   od_put          put(): centre, colour word, covariance from the
 *                   quaternion made unit length, :335-370 (Math.hypot as the
 *                   square root of the sum of squares: parity unpinned at
 *                   V8's last f64 bit, which the f32 store hides)
 *   od_byte         byte(): clamp(Math.round(v * 255)), :372; a NaN gives 0
 *                   as the colour word's `|` makes of it
 *   od_half         signedHalf + halfBits (src/render/hdr.js:152-163)
 *   od_decode_ply   parsePly's vertex loop, :457-467 (the header is parsed by
 *                   the Python side of the oracle, oracle/__init__.py)
 *   od_decode_spz   parseSpz, :225-303, on gunzipped bytes
 *   od_decode_sog   readSog's loop, :162-203, on decoded images
 * Each decoder writes centres (4 floats a splat), covariances (6), halves
 * (2 sh_words a splat) and returns the first splat whose position is not a
 * finite number, or -1; the spz decoder returns a negative code for a refused
 * header instead (see there). Pinned by tests/test_decode.py.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static const double SH_C0 = 0.28209479177387814;
static const int COEFS[5] = {0, 3, 8, 15, 24};

static int sh_words(int degree) { return (COEFS[degree] * 3 + 1) / 2; }

static double js_round(double x) {
  double r = floor(x);
  return (x - r >= 0.5) ? r + 1 : r;
}

unsigned od_byte(double v) {
  double r = js_round(v * 255);
  if (!(r >= 0)) return 0;
  return r > 255 ? 255u : (unsigned)r;
}

static unsigned half_bits(double value) {
  if (!(value > 0)) return 0;
  if (value >= 65504) return 0x7bff;
  if (value < ldexp(1, -24)) return 0;
  if (value < ldexp(1, -14)) return (unsigned)js_round(value / ldexp(1, -24));
  int exponent = (int)floor(log2(value));
  double mantissa = js_round((value / ldexp(1, exponent) - 1) * 1024);
  if (mantissa == 1024) {
    mantissa = 0;
    exponent++;
  }
  /* where log2 rounded up to a power of two, the mantissa is -0: that power, as `|` makes of it */
  return ((unsigned)(exponent + 15) << 10) | (unsigned)(mantissa > 0 ? mantissa : 0);
}

uint16_t od_half(double v) { return (uint16_t)(v < 0 ? 0x8000u | half_bits(-v) : half_bits(v)); }

typedef struct od_out {
  float *centers;
  float *covs;
  uint16_t *halves;
  int degree;
} od_out;

static void od_put(od_out *o, uint64_t i, double x, double y, double z, unsigned r, unsigned g, unsigned b,
                   unsigned a, double sx, double sy, double sz, double qw, double qx, double qy, double qz) {
  float *c = o->centers + 4 * i;
  c[0] = (float)x;
  c[1] = (float)y;
  c[2] = (float)z;
  uint32_t word = r | (g << 8) | (b << 16) | (a << 24);
  memcpy(&c[3], &word, 4);
  double length = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  if (length == 0 || length != length) length = 1;
  double w = qw / length, u = qx / length, v = qy / length, t = qz / length;
  double r00 = 1 - 2 * (v * v + t * t), r01 = 2 * (u * v - w * t), r02 = 2 * (u * t + w * v);
  double r10 = 2 * (u * v + w * t), r11 = 1 - 2 * (u * u + t * t), r12 = 2 * (v * t - w * u);
  double r20 = 2 * (u * t - w * v), r21 = 2 * (v * t + w * u), r22 = 1 - 2 * (u * u + v * v);
  double m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
  double m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
  double m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
  float *s = o->covs + 6 * i;
  s[0] = (float)(m00 * m00 + m01 * m01 + m02 * m02);
  s[1] = (float)(m00 * m10 + m01 * m11 + m02 * m12);
  s[2] = (float)(m00 * m20 + m01 * m21 + m02 * m22);
  s[3] = (float)(m10 * m10 + m11 * m11 + m12 * m12);
  s[4] = (float)(m10 * m20 + m11 * m21 + m12 * m22);
  s[5] = (float)(m20 * m20 + m21 * m21 + m22 * m22);
}

static void od_put_sh(od_out *o, uint64_t i, int k, double r, double g, double b) {
  uint16_t *h = o->halves + i * sh_words(o->degree) * 2 + k * 3;
  h[0] = od_half(r);
  h[1] = od_half(g);
  h[2] = od_half(b);
}

static double ply_get(const uint8_t *p, int type) {
  switch (type) {
    case 0: return (double)(int8_t)p[0];
    case 1: return (double)p[0];
    case 2: { int16_t v; memcpy(&v, p, 2); return v; }
    case 3: { uint16_t v; memcpy(&v, p, 2); return v; }
    case 4: { int32_t v; memcpy(&v, p, 4); return v; }
    case 5: { uint32_t v; memcpy(&v, p, 4); return v; }
    case 6: { float v; memcpy(&v, p, 4); return v; }
    default: { double v; memcpy(&v, p, 8); return v; }
  }
}

/* off/type: the 14 needed properties (x y z f_dc_0..2 opacity scale_0..2 rot_0..3), then f_rest_0..rest-1 */
int64_t od_decode_ply(const uint8_t *body, uint64_t n, uint32_t stride, const uint32_t *off, const uint8_t *type,
                      uint32_t rest, int degree, float *centers, float *covs, uint16_t *halves) {
  od_out o = {centers, covs, halves, degree};
  uint32_t per = rest / 3;
  for (uint64_t i = 0; i < n; i++) {
    const uint8_t *b = body + i * stride;
    double v[14 + 45];
    for (uint32_t k = 0; k < 14 + rest; k++) v[k] = ply_get(b + off[k], type[k]);
    if (!(isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]))) return (int64_t)i;
    double alpha = 1 / (1 + exp(-v[6]));
    od_put(&o, i, v[0], v[1], v[2], od_byte(0.5 + SH_C0 * v[3]), od_byte(0.5 + SH_C0 * v[4]),
           od_byte(0.5 + SH_C0 * v[5]), od_byte(alpha), exp(v[7]), exp(v[8]), exp(v[9]), v[10], v[11], v[12], v[13]);
    for (uint32_t k = 0; k < per; k++) od_put_sh(&o, i, (int)k, v[14 + k], v[14 + k + per], v[14 + k + 2 * per]);
  }
  return -1;
}

static const int SPZ_SH_FLIP[15] = {-1, -1, 1, -1, 1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};

static uint32_t rd32(const uint8_t *p) { return (uint32_t)p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

static double from_half(unsigned h) {
  double sign = (h & 0x8000) ? -1 : 1;
  int exponent = (h >> 10) & 31;
  unsigned mantissa = h & 1023;
  if (exponent == 31) return mantissa ? NAN : sign * INFINITY;
  return sign * (exponent == 0 ? mantissa * ldexp(1, -24) : (1 + mantissa / 1024.0) * ldexp(1, exponent - 15));
}

/*
 * A gunzipped .spz. Writes *count and *degree first; returns -1 when decoded, the first
 * splat with a non-finite position, or -2 not a .spz, -3 a version outside 1-3, -4 a degree
 * above 4, -5 no splats, -6 cut short. Call with null outputs to read the header only.
 */
int64_t od_decode_spz(const uint8_t *data, uint64_t len, uint32_t *count_out, int *degree_out, float *centers,
                      float *covs, uint16_t *halves) {
  if (len < 16 || rd32(data) != 0x5053474eu) return -2;
  uint32_t version = rd32(data + 4);
  if (version < 1 || version > 3) return -3;
  uint32_t count = rd32(data + 8);
  int file_degree = data[12];
  int fraction = data[13];
  if (file_degree > 4) return -4;
  int degree = file_degree < 3 ? file_degree : 3;
  int file_coefficients = COEFS[file_degree];
  uint64_t position_bytes = version == 1 ? 6 : 9, rotation_bytes = version >= 3 ? 4 : 3;
  uint64_t sizes[6] = {position_bytes, 1, 3, 3, rotation_bytes, (uint64_t)file_coefficients * 3}, offs[6];
  uint64_t at = 16;
  for (int k = 0; k < 6; k++) {
    offs[k] = at;
    at += (uint64_t)count * sizes[k];
  }
  *count_out = count;
  *degree_out = degree;
  if (count == 0) return -5;
  if (at > len) return -6;
  if (!centers) return -1;
  od_out o = {centers, covs, halves, degree};
  double fixed = 1.0 / (double)(int32_t)(1u << (fraction & 31));
  for (uint64_t i = 0; i < count; i++) {
    double position[3], q[4];
    for (int a = 0; a < 3; a++) {
      if (version == 1) {
        const uint8_t *p = data + offs[0] + i * 6 + a * 2;
        position[a] = from_half(p[0] | (p[1] << 8));
      } else {
        const uint8_t *p = data + offs[0] + i * 9 + a * 3;
        position[a] = (double)((int32_t)((uint32_t)(p[0] | (p[1] << 8) | (p[2] << 16)) << 8) >> 8) * fixed;
      }
    }
    const uint8_t *r = data + offs[4] + i * rotation_bytes;
    if (rotation_bytes == 3) {
      for (int a = 0; a < 3; a++) q[a] = r[a] / 127.5 - 1;
      double s = 1 - (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      q[3] = sqrt(s > 0 ? s : 0);
    } else {
      uint32_t packed = rd32(r);
      unsigned largest = packed >> 30;
      double sum = 0;
      for (int a = 3; a >= 0; a--) {
        if ((unsigned)a == largest) continue;
        unsigned magnitude = packed & 511, negative = (packed >> 9) & 1;
        packed >>= 10;
        q[a] = 0.7071067811865476 * magnitude / 511 * (negative ? -1 : 1);
        sum += q[a] * q[a];
      }
      double s = 1 - sum;
      q[largest] = sqrt(s > 0 ? s : 0);
    }
    if (!(isfinite(position[0]) && isfinite(position[1]) && isfinite(position[2]))) return (int64_t)i;
    const uint8_t *c = data + offs[2] + i * 3, *s = data + offs[3] + i * 3;
    unsigned col[3];
    double sc[3];
    for (int k = 0; k < 3; k++) {
      col[k] = od_byte(0.5 + SH_C0 * ((c[k] / 255.0 - 0.5) / 0.15));
      sc[k] = exp(s[k] / 16.0 - 10);
    }
    od_put(&o, i, position[0], -position[1], -position[2], col[0], col[1], col[2], data[offs[1] + i], sc[0], sc[1],
           sc[2], q[3], q[0], -q[1], -q[2]);
    const uint8_t *h = data + offs[5] + i * file_coefficients * 3;
    for (int k = 0; k < COEFS[degree]; k++)
      od_put_sh(&o, i, k, (h[k * 3] - 128) / 128.0 * SPZ_SH_FLIP[k], (h[k * 3 + 1] - 128) / 128.0 * SPZ_SH_FLIP[k],
                (h[k * 3 + 2] - 128) / 128.0 * SPZ_SH_FLIP[k]);
  }
  return -1;
}

/*
 * A .sog's loop on its decoded images (rgba8, a splat a pixel). books: 256 scale, colour and
 * harmonic codebook values each (NaN past a codebook's end); centroids: the palette image,
 * width 64 per, of centroids_bytes bytes.
 */
int64_t od_decode_sog(uint32_t count, int degree, uint32_t sh_count, const double *mins, const double *maxs,
                      const uint8_t *means_l, const uint8_t *means_u, const uint8_t *quats, const uint8_t *scales,
                      const uint8_t *sh0, const uint8_t *labels, const uint8_t *centroids, uint32_t centroids_width,
                      uint64_t centroids_bytes, const double *scale_book, const double *colour_book,
                      const double *sh_book, float *centers, float *covs, uint16_t *halves) {
  od_out o = {centers, covs, halves, degree};
  int per = COEFS[degree];
  for (uint64_t i = 0; i < count; i++) {
    uint64_t px = 4 * i;
    double m[3], q[4] = {1, 0, 0, 0};
    for (int a = 0; a < 3; a++) {
      double v = means_l[px + a] | (means_u[px + a] << 8);
      double range = maxs[a] - mins[a];
      double n = mins[a] + ((range != 0 && range == range) ? range : 1) * v / 65535;
      double e = exp(fabs(n)) - 1;
      m[a] = n < 0 ? -e : e;
    }
    unsigned tag = quats[px + 3];
    if (tag >= 252) {
      unsigned largest = tag - 252, k = 0;
      double sum = 0;
      for (unsigned c = 0; c < 4; c++) {
        if (c == largest) continue;
        q[c] = ((quats[px + k++] / 255.0) * 2 - 1) / 1.4142135623730951;
        sum += q[c] * q[c];
      }
      double s = 1 - sum;
      q[largest] = sqrt(s > 0 ? s : 0);
    }
    if (!(isfinite(m[0]) && isfinite(m[1]) && isfinite(m[2]))) return (int64_t)i;
    od_put(&o, i, m[0], m[1], m[2], od_byte(0.5 + SH_C0 * colour_book[sh0[px]]),
           od_byte(0.5 + SH_C0 * colour_book[sh0[px + 1]]), od_byte(0.5 + SH_C0 * colour_book[sh0[px + 2]]),
           sh0[px + 3], exp(scale_book[scales[px]]), exp(scale_book[scales[px + 1]]), exp(scale_book[scales[px + 2]]),
           q[0], q[1], q[2], q[3]);
    if (per == 0) continue;
    unsigned label = labels[px] | (labels[px + 1] << 8);
    if (label >= sh_count) continue;
    uint64_t row = (uint64_t)(label / 64) * centroids_width + (label % 64) * per;
    for (int k = 0; k < per; k++) {
      uint64_t c = (row + k) * 4;
      if (c + 2 >= centroids_bytes) break;
      od_put_sh(&o, i, k, sh_book[centroids[c]], sh_book[centroids[c + 1]], sh_book[centroids[c + 2]]);
    }
  }
  return -1;
}

// Capture decoders on the GPU: a .ply, a gunzipped .spz or a .sog's decoded
// images in, a cloud of centres and Sigma (the reference's Splats data,
// src/scene/splats.js:305-370) out, with its harmonics. One lane a splat; the
// host has parsed the header and checked every size before a launch.
//
// Numerics. The reference decodes in JavaScript numbers, FP64, and stores f32
// (centres, covariances) and half floats (harmonics). Here the same formulas
// run in FP64 in the same order: Math.round as floor plus a half-step test,
// byte() with JavaScript's NaN behaviour, halfBits with its rounding (half up,
// not to even) and its saturation at 0x7bff. exp and the quaternion's length
// (Math.hypot there, the square root of the sum of squares here) may differ
// from V8's in the last f64 bit, which the f32 store hides but for values at
// an f32 rounding edge.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "wd_internal.h"

namespace wd {
namespace {

constexpr double kShC0 = 0.28209479177387814;
// Turning RUB to RDF flips these harmonics' signs, in 3DGS order
// (src/scene/splats.js:219): bits 0, 1, 3, 6, 8, 10, 11, 13.
constexpr uint32_t kSpzShNegative = 0x2D4Bu;

__device__ __forceinline__ double js_round(double x) {
  const double r = floor(x);
  return (x - r >= 0.5) ? r + 1.0 : r;
}

// byte(v): max(0, min(255, round(v * 255))); a NaN gives 0, as `|` makes of it
__device__ __forceinline__ uint32_t js_byte(double v) {
  const double r = js_round(v * 255.0);
  if (!(r >= 0.0)) return 0u;
  return r > 255.0 ? 255u : (uint32_t)r;
}

// halfBits (src/render/hdr.js:152-163) of a value, and signedHalf
// (src/scene/splats.js:50-52): the sign, then the magnitude's half bits.
__device__ __forceinline__ uint32_t half_bits(double v) {
  if (!(v > 0.0)) return 0u;
  if (v >= 65504.0) return 0x7bffu;
  if (v < 0x1p-24) return 0u;
  if (v < 0x1p-14) return (uint32_t)js_round(v * 0x1p24);
  int e;
  const double f = frexp(v, &e);  // v = f 2^e, f in [0.5, 1)
  int exponent = e - 1;
  double mantissa = js_round((ldexp(f, 1) - 1.0) * 1024.0);
  if (mantissa == 1024.0) {
    mantissa = 0.0;
    ++exponent;
  }
  return (uint32_t)((exponent + 15) << 10) | (uint32_t)mantissa;
}

__device__ __forceinline__ uint32_t signed_half(double v) { return v < 0 ? 0x8000u | half_bits(-v) : half_bits(v); }

// The float's bits in an order that compares as the float does.
__device__ __forceinline__ uint32_t ordered(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// put() (src/scene/splats.js:335-370) for splat i: centre, colour word, box,
// Sigma = (R S)(R S)^T from the quaternion made unit length; a non-finite
// position is recorded in bad (the least such index wins) and left out of the box.
__device__ __forceinline__ void put(DecodeOut o, uint32_t i, double x, double y, double z, uint32_t r, uint32_t g,
                                    uint32_t b, uint32_t a, double sx, double sy, double sz, double qw, double qx,
                                    double qy, double qz) {
  const bool finite = isfinite(x) && isfinite(y) && isfinite(z);
  if (!finite) atomicMin(o.bad, i);
  const float fx = (float)x, fy = (float)y, fz = (float)z;
  reinterpret_cast<float4 *>(o.centers)[i] = make_float4(fx, fy, fz, __uint_as_float(r | (g << 8) | (b << 16) | (a << 24)));
  if (finite) {
    const float p[3] = {fx, fy, fz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      atomicMin(&o.box[k], ordered(p[k]));
      atomicMax(&o.box[3 + k], ordered(p[k]));
    }
  }
  double length = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  if (length == 0.0 || length != length) length = 1.0;  // `|| 1`: a zero or NaN length is 1
  const double w = qw / length, u = qx / length, v = qy / length, t = qz / length;
  const double r00 = 1 - 2 * (v * v + t * t), r01 = 2 * (u * v - w * t), r02 = 2 * (u * t + w * v);
  const double r10 = 2 * (u * v + w * t), r11 = 1 - 2 * (u * u + t * t), r12 = 2 * (v * t - w * u);
  const double r20 = 2 * (u * t - w * v), r21 = 2 * (v * t + w * u), r22 = 1 - 2 * (u * u + v * v);
  const double m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
  const double m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
  const double m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
  float *c = o.covs + 6 * (size_t)i;
  c[0] = (float)(m00 * m00 + m01 * m01 + m02 * m02);
  c[1] = (float)(m00 * m10 + m01 * m11 + m02 * m12);
  c[2] = (float)(m00 * m20 + m01 * m21 + m02 * m22);
  c[3] = (float)(m10 * m10 + m11 * m11 + m12 * m12);
  c[4] = (float)(m10 * m20 + m11 * m21 + m12 * m22);
  c[5] = (float)(m20 * m20 + m21 * m21 + m22 * m22);
}

// putSH (src/scene/splats.js:327-332): coefficient k of splat i, red, green, blue.
__device__ __forceinline__ void put_sh(DecodeOut o, uint32_t i, uint32_t k, double r, double g, double b) {
  uint16_t *h = reinterpret_cast<uint16_t *>(o.sh) + (size_t)i * o.sh_words * 2 + k * 3;
  h[0] = (uint16_t)signed_half(r);
  h[1] = (uint16_t)signed_half(g);
  h[2] = (uint16_t)signed_half(b);
}

// One .ply property of a vertex, as a JavaScript number.
__device__ __forceinline__ double ply_get(const uint8_t *base, uint32_t off, uint32_t type) {
  const uint8_t *p = base + off;
  switch (type) {
    case kPlyI8: return (double)(int8_t)p[0];
    case kPlyU8: return (double)p[0];
    case kPlyI16: { int16_t v; __builtin_memcpy(&v, p, 2); return (double)v; }
    case kPlyU16: { uint16_t v; __builtin_memcpy(&v, p, 2); return (double)v; }
    case kPlyI32: { int32_t v; __builtin_memcpy(&v, p, 4); return (double)v; }
    case kPlyU32: { uint32_t v; __builtin_memcpy(&v, p, 4); return (double)v; }
    case kPlyF32: { float v; __builtin_memcpy(&v, p, 4); return (double)v; }
    default: { double v; __builtin_memcpy(&v, p, 8); return v; }
  }
}

// parsePly's vertex loop (src/scene/splats.js:457-467).
__global__ __launch_bounds__(256) void k_decode_ply(const uint8_t *__restrict__ body, uint32_t n, PlyLayout L,
                                                    DecodeOut o) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint8_t *b = body + (size_t)i * L.stride;
  auto get = [&](int k) { return ply_get(b, L.off[k], L.type[k]); };
  const double op = get(6);
  const double alpha = 1.0 / (1.0 + exp(-op));
  put(o, i, get(0), get(1), get(2), js_byte(0.5 + kShC0 * get(3)), js_byte(0.5 + kShC0 * get(4)),
      js_byte(0.5 + kShC0 * get(5)), js_byte(alpha), exp(get(7)), exp(get(8)), exp(get(9)), get(10), get(11),
      get(12), get(13));
  const uint32_t per = L.rest / 3;
  for (uint32_t k = 0; k < per; ++k)
    put_sh(o, i, k, get(kPlyNeeds + k), get(kPlyNeeds + k + per), get(kPlyNeeds + k + 2 * per));
}

// parseSpz's loop (src/scene/splats.js:257-301), over the sections of a gunzipped .spz.
__global__ __launch_bounds__(256) void k_decode_spz(const uint8_t *__restrict__ data, uint32_t n, SpzLayout L,
                                                    DecodeOut o) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  double pos[3];
  for (int a = 0; a < 3; ++a) {
    if (L.version == 1) {
      const uint8_t *p = data + L.positions + (size_t)i * 6 + a * 2;
      const uint32_t h = p[0] | (p[1] << 8);
      const int e = (h >> 10) & 31;
      const uint32_t m = h & 1023u;
      const double sign = (h & 0x8000u) ? -1.0 : 1.0;
      pos[a] = e == 31 ? (m ? NAN : sign * INFINITY)
                       : sign * (e == 0 ? (double)m * 0x1p-24 : (1.0 + m / 1024.0) * ldexp(1.0, e - 15));
    } else {
      const uint8_t *p = data + L.positions + (size_t)i * 9 + a * 3;
      const int32_t fixed = (int32_t)((uint32_t)(p[0] | (p[1] << 8) | (p[2] << 16)) << 8) >> 8;
      pos[a] = (double)fixed * L.fixed;
    }
  }
  double q[4];
  const uint8_t *r = data + L.rotations + (size_t)i * L.rotation_bytes;
  if (L.rotation_bytes == 3) {
    for (int a = 0; a < 3; ++a) q[a] = r[a] / 127.5 - 1.0;
    const double s = 1.0 - (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    q[3] = sqrt(s > 0.0 ? s : 0.0);
  } else {
    // smallest three: the largest component's index in the top two bits, the
    // others as a sign and nine bits of magnitude each, last first
    uint32_t packed = r[0] | (r[1] << 8) | (r[2] << 16) | ((uint32_t)r[3] << 24);
    const uint32_t largest = packed >> 30;
    double sum = 0.0;
    for (int a = 3; a >= 0; --a) {
      if ((uint32_t)a == largest) continue;
      const uint32_t magnitude = packed & 511u;
      const uint32_t negative = (packed >> 9) & 1u;
      packed >>= 10;
      q[a] = 0.7071067811865476 * magnitude / 511.0 * (negative ? -1.0 : 1.0);
      sum += q[a] * q[a];
    }
    const double s = 1.0 - sum;
    q[largest] = sqrt(s > 0.0 ? s : 0.0);
  }
  const uint8_t *c = data + L.colours + (size_t)i * 3;
  const uint8_t *s = data + L.scales + (size_t)i * 3;
  auto colour = [&](int k) { return js_byte(0.5 + kShC0 * ((c[k] / 255.0 - 0.5) / 0.15)); };
  auto scale = [&](int k) { return exp(s[k] / 16.0 - 10.0); };
  // RUB to RDF: a position's y and z negate, and so do a rotation's y and z
  put(o, i, pos[0], -pos[1], -pos[2], colour(0), colour(1), colour(2), data[L.alphas + i], scale(0), scale(1),
      scale(2), q[3], q[0], -q[1], -q[2]);
  const uint8_t *h = data + L.sh + (size_t)i * L.file_coefficients * 3;
  for (uint32_t k = 0; k < L.coefficients; ++k) {
    const double flip = ((kSpzShNegative >> k) & 1u) ? -1.0 : 1.0;
    put_sh(o, i, k, (h[k * 3] - 128) / 128.0 * flip, (h[k * 3 + 1] - 128) / 128.0 * flip,
           (h[k * 3 + 2] - 128) / 128.0 * flip);
  }
}

// readSog's loop (src/scene/splats.js:170-203), over the decoded images.
__global__ __launch_bounds__(256) void k_decode_sog(SogLayout L, DecodeOut o) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= L.count) return;
  const uint32_t px = 4 * i;
  double m[3];
  for (int a = 0; a < 3; ++a) {
    // stored as sign(x) ln(|x| + 1) in a 16-bit range
    const double v = L.means_l[px + a] | (L.means_u[px + a] << 8);
    const double range = L.maxs[a] - L.mins[a];
    const double n = L.mins[a] + ((range != 0.0 && range == range) ? range : 1.0) * v / 65535.0;
    const double e = exp(fabs(n)) - 1.0;
    m[a] = n < 0 ? -e : e;
  }
  double q[4] = {1.0, 0.0, 0.0, 0.0};
  const uint32_t tag = L.quats[px + 3];
  if (tag >= 252) {
    const uint32_t largest = tag - 252;
    uint32_t k = 0;
    double sum = 0.0;
    for (uint32_t cpt = 0; cpt < 4; ++cpt) {
      if (cpt == largest) continue;
      q[cpt] = ((L.quats[px + k++] / 255.0) * 2.0 - 1.0) / 1.4142135623730951;
      sum += q[cpt] * q[cpt];
    }
    const double s = 1.0 - sum;
    q[largest] = sqrt(s > 0.0 ? s : 0.0);
  }
  put(o, i, m[0], m[1], m[2], L.colour_byte[L.sh0[px]], L.colour_byte[L.sh0[px + 1]], L.colour_byte[L.sh0[px + 2]],
      L.sh0[px + 3], L.scale_exp[L.scales[px]], L.scale_exp[L.scales[px + 1]], L.scale_exp[L.scales[px + 2]], q[0],
      q[1], q[2], q[3]);
  if (L.per == 0) return;
  const uint32_t label = L.labels[px] | (L.labels[px + 1] << 8);
  if (label >= L.sh_count) return;
  const uint64_t row = (uint64_t)(label / 64) * L.centroids_width + (label % 64) * L.per;
  for (uint32_t k = 0; k < L.per; ++k) {
    const uint64_t c = (row + k) * 4;
    if (c + 2 >= L.centroids_bytes) break;
    uint16_t *h = reinterpret_cast<uint16_t *>(o.sh) + (size_t)i * o.sh_words * 2 + k * 3;
    h[0] = L.sh_half[L.centroids[c]];
    h[1] = L.sh_half[L.centroids[c + 1]];
    h[2] = L.sh_half[L.centroids[c + 2]];
  }
}

}  // namespace

int launch_decode_ply(wd_context *ctx, const uint8_t *body, uint32_t n, const PlyLayout &L, const DecodeOut &o) {
  hipLaunchKernelGGL(k_decode_ply, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, body, n, L, o);
  WD_HIP(hipGetLastError(), "splats: .ply decode launch");
  return WD_OK;
}

int launch_decode_spz(wd_context *ctx, const uint8_t *data, uint32_t n, const SpzLayout &L, const DecodeOut &o) {
  hipLaunchKernelGGL(k_decode_spz, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, data, n, L, o);
  WD_HIP(hipGetLastError(), "splats: .spz decode launch");
  return WD_OK;
}

int launch_decode_sog(wd_context *ctx, const SogLayout &L, const DecodeOut &o) {
  hipLaunchKernelGGL(k_decode_sog, dim3((L.count + 255) / 256), dim3(256), 0, ctx->stream, L, o);
  WD_HIP(hipGetLastError(), "splats: .sog decode launch");
  return WD_OK;
}

}  // namespace wd

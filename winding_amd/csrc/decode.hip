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

#include <algorithm>
#include <cstdint>

#include "splat_device.h"
#include "wd_internal.h"

namespace wd {
namespace {

constexpr double kShC0 = 0.28209479177387814;
// Turning RUB to RDF flips these harmonics' signs, in 3DGS order
// (src/scene/splats.js:219): bits 0, 1, 3, 6, 8, 10, 11, 13.
constexpr uint32_t kSpzShNegative = 0x2D4Bu;
constexpr uint32_t kShCoeffs[4] = {0, 3, 8, 15};
// The staged .ply form holds a workgroup's 256 vertices in LDS: up to 64 KB
// (62 floats a vertex, the layout training writes, take 62 KB).
constexpr uint32_t kPlyStagedStride = 256;

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

// signed_half of a value that is an f32, from its bits: the same answer
// (halfBits' tests and its rounding, half up, are exact on the f32's
// mantissa) in integer operations and selects instead of FP64 ones and
// branches. |v| = m 2^(e - 150) with m the 24-bit significand: a subnormal
// half is round(|v| 2^24) = (m + 2^(s - 1)) >> s, s = 126 - e in [14, 23]; a
// normal one the top 10 mantissa bits plus one when the 13 dropped are half
// or more (a carry into the exponent is halfBits' mantissa == 1024 step).
// Equal to the branch-by-branch form for all 2^32 inputs.
__device__ __forceinline__ uint32_t signed_half_f32(float v) {
  const uint32_t bits = __float_as_uint(v);
  const uint32_t a = bits & 0x7FFFFFFFu, e = a >> 23;  // |v|
  const uint32_t m = (a & 0x7FFFFFu) | 0x800000u;
  const uint32_t s = (e >= 103u && e <= 112u) ? 126u - e : 14u;
  const uint32_t sub = (m + (1u << (s - 1u))) >> s;
  const uint32_t normal = ((e - 112u) << 10) + ((a >> 13) & 1023u) + ((a >> 12) & 1u);
  const uint32_t h = a < 0x33800000u   ? 0u       // |v| < 2^-24, zero included
                     : a < 0x38800000u ? sub      // |v| < 2^-14
                     : a < 0x477FE000u ? normal   // |v| < 65504 less half a step
                     : a <= 0x7F800000u ? 0x7bffu  // to infinity
                                        : 0u;     // NaN
  const bool negative = (bits >> 31) != 0u && a != 0u && a <= 0x7F800000u;  // v < 0
  return (negative ? 0x8000u : 0u) | h;
}

// signed_half_f32 for a wave: a value that is zero or in a normal half's range
// (2^-14 <= |v| < 65504 less half a step) takes the short form, the top 10
// mantissa bits and the exponent rebased in one add, plus the first dropped bit;
// the branch-by-branch form runs only for a value that some lane of the wave
// holds outside that range (equal to it on every value the short form takes).
__device__ __forceinline__ uint32_t signed_half_f32_wave(float v) {
  const uint32_t bits = __float_as_uint(v), a = bits & 0x7FFFFFFFu;
  const bool usual = a == 0u || a - 0x38800000u < 0x477FE000u - 0x38800000u;
  uint32_t h = a == 0u ? 0u : ((a >> 13) + ((a >> 12) & 1u) - 0x1C000u) | ((bits >> 16) & 0x8000u);
  if (__builtin_expect(!__all(usual), 0)) h = usual ? h : signed_half_f32(v);
  return h;
}

// The float's bits in an order that compares as the float does.
__device__ __forceinline__ uint32_t ordered(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// One decoded splat in registers: its centre word and Sigma, as put() makes
// them (src/scene/splats.js:335-370), and its harmonics as packed halves.
template <int DEG>
struct Decoded {
  static constexpr uint32_t W = DEG ? sh_words(DEG) : 1;
  float4 centre;
  float cov[6];
  uint32_t sh[W];
  bool finite;
};

// put(): centre, colour word, Sigma = (R S)(R S)^T from the quaternion made
// unit length (`Math.hypot(...) || 1`).
template <int DEG>
__device__ __forceinline__ void put(Decoded<DEG> &d, double x, double y, double z, uint32_t r, uint32_t g, uint32_t b,
                                    uint32_t a, double sx, double sy, double sz, double qw, double qx, double qy,
                                    double qz) {
  d.finite = isfinite(x) && isfinite(y) && isfinite(z);
  d.centre = make_float4((float)x, (float)y, (float)z, __uint_as_float(r | (g << 8) | (b << 16) | (a << 24)));
  double length = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  if (length == 0.0 || length != length) length = 1.0;
  const double w = qw / length, u = qx / length, v = qy / length, t = qz / length;
  const double r00 = 1 - 2 * (v * v + t * t), r01 = 2 * (u * v - w * t), r02 = 2 * (u * t + w * v);
  const double r10 = 2 * (u * v + w * t), r11 = 1 - 2 * (u * u + t * t), r12 = 2 * (v * t - w * u);
  const double r20 = 2 * (u * t - w * v), r21 = 2 * (v * t + w * u), r22 = 1 - 2 * (u * u + v * v);
  const double m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
  const double m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
  const double m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
  d.cov[0] = (float)(m00 * m00 + m01 * m01 + m02 * m02);
  d.cov[1] = (float)(m00 * m10 + m01 * m11 + m02 * m12);
  d.cov[2] = (float)(m00 * m20 + m01 * m21 + m02 * m22);
  d.cov[3] = (float)(m10 * m10 + m11 * m11 + m12 * m12);
  d.cov[4] = (float)(m10 * m20 + m11 * m21 + m12 * m22);
  d.cov[5] = (float)(m20 * m20 + m21 * m21 + m22 * m22);
}

// Half h of a splat's harmonics (coefficient k, channel c at h = 3 k + c: the
// `sh` layout, two halves a word, the lower first). h is a constant after unrolling.
template <int DEG>
__device__ __forceinline__ void set_half(Decoded<DEG> &d, uint32_t h, uint32_t bits) {
  uint32_t &word = d.sh[h >> 1];
  word = (h & 1u) ? ((word & 0xFFFFu) | (bits << 16)) : ((word & 0xFFFF0000u) | (bits & 0xFFFFu));
}

// putSH (src/scene/splats.js:327-332): coefficient k, red, green, blue.
template <int DEG>
__device__ __forceinline__ void put_sh(Decoded<DEG> &d, uint32_t k, double r, double g, double b) {
  set_half(d, 3 * k, signed_half(r));
  set_half(d, 3 * k + 1, signed_half(g));
  set_half(d, 3 * k + 2, signed_half(b));
}
// The same for values that are f32s (a .ply's float properties, a .spz's byte fractions).
template <int DEG>
__device__ __forceinline__ void put_sh_f32(Decoded<DEG> &d, uint32_t k, float r, float g, float b) {
  set_half(d, 3 * k, signed_half_f32_wave(r));
  set_half(d, 3 * k + 1, signed_half_f32_wave(g));
  set_half(d, 3 * k + 2, signed_half_f32_wave(b));
}

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, s));
  return x;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) x = max(x, (uint32_t)__shfl_xor((int)x, s));
  return x;
}

// The workgroup's decoded splats out: the box of the finite centres and the
// least index of a non-finite position, reduced in the wave and across the
// workgroup into the workgroup's own column of o.partials (reduced over the
// workgroups afterwards by launch_box_reduce: no global atomics), and
// centres, Sigma and harmonics staged in LDS so that each goes out as one
// contiguous run of whole lines. The workgroup is G splats; lds holds
// G (10 + W) words, and the caller has finished with it.
template <int DEG, uint32_t G = 256>
__device__ __forceinline__ void emit(const DecodeOut &o, uint32_t i0, uint32_t cnt, bool valid, const Decoded<DEG> &d,
                                     uint32_t *lds) {
  constexpr uint32_t W = DEG ? sh_words(DEG) : 0;
  __shared__ uint32_t box[kBoxWords];  // ordered min xyz, ordered max xyz, first zero per axis, first bad
  const uint32_t t = threadIdx.x;
  if (t < kBoxWords) box[t] = (t < 3 || t >= 6) ? 0xFFFFFFFFu : 0u;
  __syncthreads();
  if (valid && !d.finite) atomicMin(&box[9], i0 + t);
  const bool in_box = valid && d.finite;
  const float p[3] = {d.centre.x, d.centre.y, d.centre.z};
  uint32_t lo[3], hi[3], zi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = wave_min(in_box ? ordered(p[k]) : 0xFFFFFFFFu);
    hi[k] = wave_max(in_box ? ordered(p[k]) : 0u);
    zi[k] = wave_min(in_box && p[k] == 0.0f ? i0 + t : 0xFFFFFFFFu);
  }
  uint32_t *centres = lds, *covs = lds + G * 4, *sh = lds + G * 10;
  if (valid) {
    reinterpret_cast<float4 *>(centres)[t] = d.centre;
#pragma unroll
    for (int k = 0; k < 6; ++k) covs[6 * t + k] = __float_as_uint(d.cov[k]);
    if constexpr (DEG > 0)
#pragma unroll
      for (uint32_t k = 0; k < W; ++k) sh[W * t + k] = d.sh[k];
  }
  __syncthreads();
  if ((t & 63u) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      atomicMin(&box[k], lo[k]);
      atomicMax(&box[3 + k], hi[k]);
      atomicMin(&box[6 + k], zi[k]);
    }
  }
  // Each region leaves in 16-byte pieces, then its last words: the LDS regions start at
  // multiples of 16 bytes, and so do the outputs (i0 is a multiple of 128 splats).
  auto copy_out = [&](uint32_t *dst, const uint32_t *src, uint32_t words) {
    const uint32_t n4 = words >> 2;
    for (uint32_t k = t; k < n4; k += G) reinterpret_cast<uint4 *>(dst)[k] = reinterpret_cast<const uint4 *>(src)[k];
    for (uint32_t k = 4u * n4 + t; k < words; k += G) dst[k] = src[k];
  };
  copy_out(reinterpret_cast<uint32_t *>(o.centers) + (size_t)i0 * 4, centres, cnt * 4);
  copy_out(reinterpret_cast<uint32_t *>(o.covs) + (size_t)i0 * 6, covs, cnt * 6);
  if constexpr (DEG > 0) copy_out(o.sh + (size_t)i0 * W, sh, cnt * W);
  __syncthreads();
  // word-major, and a column per workgroup in XCD-contiguous order, so that the
  // neighbouring words of a line come from workgroups that share an L2
  if (t < kBoxWords) o.partials[(size_t)t * gridDim.x + xcd_item(blockIdx.x, gridDim.x)] = box[t];
}

constexpr uint32_t emit_words(int DEG, uint32_t G = 256) { return G * (10u + (DEG ? sh_words(DEG) : 0u)); }

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

// parsePly's vertex loop (src/scene/splats.js:457-467), a workgroup kPlyGroup
// vertices. STAGED (a stride and offsets of whole words, every property a
// float, stride <= kPlyStagedStride): the workgroup's vertices are first
// copied into LDS with 16-byte loads, so the body is read once, in order.
// A workgroup is one wave: the LDS still holds ten of them a CU, and no wave
// waits at a barrier for another (88 µs for a degree-3 .ply of 1M vertices,
// against 98 at 128 vertices a workgroup).
constexpr uint32_t kPlyGroup = 64;
template <int DEG, bool STAGED>
__global__ __launch_bounds__(kPlyGroup) void k_decode_ply(const uint8_t *__restrict__ body, uint32_t n, PlyLayout L,
                                                          DecodeOut o) {
  constexpr uint32_t G = kPlyGroup;
  extern __shared__ __align__(16) uint32_t lds[];
  const uint32_t t = threadIdx.x, i0 = blockIdx.x * G, cnt = min(G, n - i0);
  const bool valid = t < cnt;
  const uint8_t *b;
  if constexpr (STAGED) {
    const uint4 *src = reinterpret_cast<const uint4 *>(body + (size_t)i0 * L.stride);
    const uint32_t n16 = (cnt * L.stride + 15u) / 16u;  // the body is padded to 16 bytes past its end
    // every load issued before the first LDS store: a thread's 16-byte pieces all in flight at once
    constexpr uint32_t kPieces = kPlyStagedStride / 16u;  // G vertices of at most kPlyStagedStride bytes
    uint4 piece[kPieces];
    // Past n16 a thread loads its last piece again and stores it past the staged vertices, in
    // the launch's G kPlyStagedStride bytes of LDS: no branch, so no load waits behind a store.
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j) piece[j] = src[min(t + j * G, n16 - 1u)];
#pragma unroll
    for (uint32_t j = 0; j < kPieces; ++j) reinterpret_cast<uint4 *>(lds)[t + j * G] = piece[j];
    __syncthreads();
    b = reinterpret_cast<const uint8_t *>(lds) + (size_t)t * L.stride;
  } else {
    b = body + (size_t)(i0 + (valid ? t : 0u)) * L.stride;
  }
  Decoded<DEG> d{};
  if (valid) {
    auto get = [&](int k) {
      if constexpr (STAGED) return (double)__uint_as_float(*reinterpret_cast<const uint32_t *>(b + L.off[k]));
      return ply_get(b, L.off[k], L.type[k]);
    };
    const double alpha = 1.0 / (1.0 + exp(-get(6)));
    put(d, get(0), get(1), get(2), js_byte(0.5 + kShC0 * get(3)), js_byte(0.5 + kShC0 * get(4)),
        js_byte(0.5 + kShC0 * get(5)), js_byte(alpha), exp(get(7)), exp(get(8)), exp(get(9)), get(10), get(11),
        get(12), get(13));
    constexpr uint32_t per = kShCoeffs[DEG];
    if constexpr (STAGED) {
      auto getf = [&](int k) { return __uint_as_float(*reinterpret_cast<const uint32_t *>(b + L.off[k])); };
#pragma unroll
      for (uint32_t k = 0; k < per; ++k)
        put_sh_f32(d, k, getf(kPlyNeeds + k), getf(kPlyNeeds + k + per), getf(kPlyNeeds + k + 2 * per));
    } else {
#pragma unroll
      for (uint32_t k = 0; k < per; ++k)
        put_sh(d, k, get(kPlyNeeds + k), get(kPlyNeeds + k + per), get(kPlyNeeds + k + 2 * per));
    }
  }
  if constexpr (STAGED) __syncthreads();  // the staged vertices are read; the LDS takes the outputs
  emit<DEG, G>(o, i0, cnt, valid, d, lds);
}

// The words of the .spz's six sections that a workgroup's 256 splats read, staged
// into LDS one section after another: all of them at most 256 (9 + 1 + 3 + 3 + 4 + 72)
// bytes, plus a word of misalignment a section.
constexpr uint32_t kSpzStageWords = 256u * 92u / 4u + 6u;

// parseSpz's loop (src/scene/splats.js:257-301), over the sections of a gunzipped .spz.
// The workgroup first copies its splats' bytes of each section into LDS with
// coalesced word loads (the data is 4-byte aligned and padded past its end),
// then each lane reads its own bytes from there.
template <int DEG>
__global__ __launch_bounds__(256) void k_decode_spz(const uint8_t *__restrict__ data, uint32_t n, SpzLayout L,
                                                    DecodeOut o) {
  extern __shared__ __align__(16) uint32_t lds[];
  const uint32_t t = threadIdx.x, i0 = blockIdx.x * 256u, cnt = min(256u, n - i0);
  const bool valid = t < cnt;
  // section k: bytes [sec + i0 w, sec + (i0 + cnt) w), from the word holding its first byte
  const uint64_t sec[6] = {L.positions, L.alphas, L.colours, L.scales, L.rotations, L.sh};
  const uint32_t width[6] = {L.version == 1 ? 6u : 9u, 1u, 3u, 3u, L.rotation_bytes, L.file_coefficients * 3u};
  uint64_t word0[6];
  uint32_t cum[7], lead[6];
  cum[0] = 0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const uint64_t b0 = sec[k] + (uint64_t)i0 * width[k], b1 = b0 + (uint64_t)cnt * width[k];
    word0[k] = b0 >> 2;
    lead[k] = (uint32_t)(b0 & 3u);
    cum[k + 1] = cum[k] + (uint32_t)(((b1 + 3u) >> 2) - (b0 >> 2));
  }
  const uint32_t *words = reinterpret_cast<const uint32_t *>(data);
  // A colour, a scale and a harmonic are each a function of one byte: the
  // workgroup makes the three tables of 256 once, while its loads are in
  // flight, from the same expressions a splat would evaluate.
  __shared__ double scale_of[256];
  __shared__ uint16_t half_of[256];
  __shared__ uint8_t colour_of[256];
  // Section by section, from a base the whole workgroup shares: a section of
  // w bytes a splat is at most 64 w + 1 words (3, 1, 1, 1, 2 and 19 loads a
  // thread for 9, 1, 3, 3, 4 and 72 bytes). Every load is issued before the
  // first store; past a section's words a thread loads its last word again and
  // does not store it.
  uint32_t vp[3], va[1], vc[1], vs[1], vr[2], vh[19];
  auto load = [&](auto &v, int k) {
    const uint32_t *src = words + word0[k];
    const uint32_t nw = cum[k + 1] - cum[k];
    constexpr uint32_t M = sizeof(v) / sizeof(v[0]);
#pragma unroll
    for (uint32_t q = 0; q < M; ++q)
      if (M < 4 || q * 256u < nw) v[q] = src[min(t + q * 256u, nw - 1u)];  // the test is the workgroup's
  };
  auto store = [&](auto &v, int k) {
    const uint32_t nw = cum[k + 1] - cum[k];
    constexpr uint32_t M = sizeof(v) / sizeof(v[0]);
#pragma unroll
    for (uint32_t q = 0; q < M; ++q)
      if (t + q * 256u < nw) lds[cum[k] + t + q * 256u] = v[q];
  };
  load(vp, 0);
  load(va, 1);
  load(vc, 2);
  load(vs, 3);
  load(vr, 4);
  load(vh, 5);
  scale_of[t] = exp(t / 16.0 - 10.0);
  colour_of[t] = (uint8_t)js_byte(0.5 + kShC0 * ((t / 255.0 - 0.5) / 0.15));
  half_of[t] = (uint16_t)signed_half_f32(((int)t - 128) / 128.0f);  // (byte - 128) / 128, exact in f32
  store(vp, 0);
  store(va, 1);
  store(vc, 2);
  store(vs, 3);
  store(vr, 4);
  store(vh, 5);
  __syncthreads();
  const uint8_t *staged = reinterpret_cast<const uint8_t *>(lds);
  auto at = [&](int k) { return staged + 4u * cum[k] + lead[k] + t * width[k]; };
  Decoded<DEG> d{};
  if (valid) {
    double pos[3];
    for (int a = 0; a < 3; ++a) {
      if (L.version == 1) {
        const uint8_t *p = at(0) + a * 2;
        const uint32_t h = p[0] | (p[1] << 8);
        const int e = (h >> 10) & 31;
        const uint32_t m = h & 1023u;
        const double sign = (h & 0x8000u) ? -1.0 : 1.0;
        pos[a] = e == 31 ? (m ? NAN : sign * INFINITY)
                         : sign * (e == 0 ? (double)m * 0x1p-24 : (1.0 + m / 1024.0) * ldexp(1.0, e - 15));
      } else {
        const uint8_t *p = at(0) + a * 3;
        const int32_t fixed = (int32_t)((uint32_t)(p[0] | (p[1] << 8) | (p[2] << 16)) << 8) >> 8;
        pos[a] = (double)fixed * L.fixed;
      }
    }
    double q[4];
    const uint8_t *r = at(4);
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
    const uint8_t *c = at(2);
    const uint8_t *s = at(3);
    auto colour = [&](int k) { return (uint32_t)colour_of[c[k]]; };
    auto scale = [&](int k) { return scale_of[s[k]]; };
    // RUB to RDF: a position's y and z negate, and so do a rotation's y and z
    put(d, pos[0], -pos[1], -pos[2], colour(0), colour(1), colour(2), *at(1), scale(0), scale(1), scale(2), q[3],
        q[0], -q[1], -q[2]);
    const uint8_t *h = at(5);
#pragma unroll
    for (uint32_t k = 0; k < kShCoeffs[DEG]; ++k) {
      // (byte - 128) / 128, flipped: the table's half with its sign turned, but for a zero
      const uint32_t flip = ((kSpzShNegative >> k) & 1u) ? 0x8000u : 0u;
#pragma unroll
      for (uint32_t c3 = 0; c3 < 3; ++c3) {
        const uint32_t bits = half_of[h[k * 3 + c3]];
        set_half(d, 3 * k + c3, bits ? bits ^ flip : 0u);
      }
    }
  }
  emit(o, i0, cnt, valid, d, lds);
}

// readSog's loop (src/scene/splats.js:170-203), over the decoded images.
template <int DEG>
__global__ __launch_bounds__(256) void k_decode_sog(SogLayout L, DecodeOut o) {
  extern __shared__ __align__(16) uint32_t lds[];
  const uint32_t t = threadIdx.x, i0 = blockIdx.x * 256u, cnt = min(256u, L.count - i0);
  const bool valid = t < cnt;
  const uint32_t px = 4 * (i0 + t);
  Decoded<DEG> d{};
  if (valid) {
    double m[3];
    for (int a = 0; a < 3; ++a) {
      // stored as sign(x) ln(|x| + 1) in a 16-bit range
      const double v = L.means_l[px + a] | (L.means_u[px + a] << 8);
      const double range = L.maxs[a] - L.mins[a];
      const double nn = L.mins[a] + ((range != 0.0 && range == range) ? range : 1.0) * v / 65535.0;
      const double e = exp(fabs(nn)) - 1.0;
      m[a] = nn < 0 ? -e : e;
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
    put(d, m[0], m[1], m[2], L.colour_byte[L.sh0[px]], L.colour_byte[L.sh0[px + 1]], L.colour_byte[L.sh0[px + 2]],
        L.sh0[px + 3], L.scale_exp[L.scales[px]], L.scale_exp[L.scales[px + 1]], L.scale_exp[L.scales[px + 2]],
        q[0], q[1], q[2], q[3]);
    if constexpr (DEG > 0) {
      const uint32_t label = L.labels[px] | (L.labels[px + 1] << 8);
      if (label < L.sh_count) {
        constexpr uint32_t per = kShCoeffs[DEG];
        const uint64_t row = (uint64_t)(label / 64) * L.centroids_width + (label % 64) * per;
#pragma unroll
        for (uint32_t k = 0; k < per; ++k) {
          const uint64_t c = (row + k) * 4;
          if (c + 2 >= L.centroids_bytes) break;
          set_half(d, 3 * k, L.sh_half[L.centroids[c]]);
          set_half(d, 3 * k + 1, L.sh_half[L.centroids[c + 1]]);
          set_half(d, 3 * k + 2, L.sh_half[L.centroids[c + 2]]);
        }
      }
    }
  }
  emit(o, i0, cnt, valid, d, lds);
}

}  // namespace

template <typename K>
static K pick(const K (&forms)[4], uint32_t degree) { return forms[degree <= 3 ? degree : 0]; }

int launch_decode_ply(wd_context *ctx, const uint8_t *body, uint32_t n, const PlyLayout &L, const DecodeOut &o,
                      uint32_t *groups) {
  using Kernel = void (*)(const uint8_t *, uint32_t, PlyLayout, DecodeOut);
  static const Kernel kDirect[4] = {k_decode_ply<0, false>, k_decode_ply<1, false>, k_decode_ply<2, false>,
                                    k_decode_ply<3, false>};
  static const Kernel kStaged[4] = {k_decode_ply<0, true>, k_decode_ply<1, true>, k_decode_ply<2, true>,
                                    k_decode_ply<3, true>};
  const uint32_t degree = L.rest / 3 == 15 ? 3 : L.rest / 3 == 8 ? 2 : L.rest / 3 == 3 ? 1 : 0;
  bool staged = L.stride % 4 == 0 && L.stride <= kPlyStagedStride;
  for (uint32_t k = 0; k < (uint32_t)kPlyNeeds + L.rest && staged; ++k)
    staged = L.type[k] == kPlyF32 && L.off[k] % 4 == 0;
  const size_t lds = std::max<size_t>(staged ? kPlyGroup * kPlyStagedStride : 0u, 4u * emit_words(3, kPlyGroup));
  *groups = (n + kPlyGroup - 1) / kPlyGroup;
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKDecode, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(pick(staged ? kStaged : kDirect, degree), dim3(*groups), dim3(kPlyGroup), lds, ctx->stream,
                     body, n, L, o);
  WD_HIP(hipGetLastError(), "splats: .ply decode launch");
  if (ctx->prof_end(kKDecode, ev)) return WD_E_HIP;
  return WD_OK;
}

int launch_decode_spz(wd_context *ctx, const uint8_t *data, uint32_t n, const SpzLayout &L, const DecodeOut &o,
                      uint32_t *groups) {
  using Kernel = void (*)(const uint8_t *, uint32_t, SpzLayout, DecodeOut);
  static const Kernel kForms[4] = {k_decode_spz<0>, k_decode_spz<1>, k_decode_spz<2>, k_decode_spz<3>};
  const uint32_t degree = L.coefficients == 15 ? 3 : L.coefficients == 8 ? 2 : L.coefficients == 3 ? 1 : 0;
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKDecode, &ev)) return WD_E_HIP;
  static_assert(kSpzStageWords <= emit_words(3), "the staged sections fit the LDS the outputs take");
  *groups = (n + 255) / 256;
  hipLaunchKernelGGL(pick(kForms, degree), dim3(*groups), dim3(256), 4u * emit_words(3), ctx->stream, data, n, L, o);
  WD_HIP(hipGetLastError(), "splats: .spz decode launch");
  if (ctx->prof_end(kKDecode, ev)) return WD_E_HIP;
  return WD_OK;
}

int launch_decode_sog(wd_context *ctx, const SogLayout &L, const DecodeOut &o, uint32_t *groups) {
  using Kernel = void (*)(SogLayout, DecodeOut);
  static const Kernel kForms[4] = {k_decode_sog<0>, k_decode_sog<1>, k_decode_sog<2>, k_decode_sog<3>};
  const uint32_t degree = L.per == 15 ? 3 : L.per == 8 ? 2 : L.per == 3 ? 1 : 0;
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKDecode, &ev)) return WD_E_HIP;
  *groups = (L.count + 255) / 256;
  hipLaunchKernelGGL(pick(kForms, degree), dim3(*groups), dim3(256), 4u * emit_words(3), ctx->stream, L, o);
  WD_HIP(hipGetLastError(), "splats: .sog decode launch");
  if (ctx->prof_end(kKDecode, ev)) return WD_E_HIP;
  return WD_OK;
}

}  // namespace wd

// Device code of the splat path shared by the single-view kernels (splat_sort.hip)
// and the batch kernel (batch.hip): numerics, wave primitives, the projection
// body and the stable scatter body.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "wd_internal.h"

namespace wd {
namespace {


constexpr double kC13 = (double)1.3f;
constexpr double kC03 = (double)0.3f;
constexpr double kC2PI = (double)6.2831853f;
constexpr float kU = 5.9604644775390625e-8f;  // 2^-24, unit roundoff of f32
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float f32_of(uint32_t bits) { return __uint_as_float(bits); }

// Lanes below this one whose bit is set in m.
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive scan over the 64 lanes of a wave on DPP: Hillis-Steele within
// each row of 16 (row_shr 1, 2, 4, 8), then row 0's total into row 1 and
// row 2's into row 3 (row_bcast:15), then the first half's into the second
// (row_bcast:31).
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
  int x = (int)v;
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, true);
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);
  return (uint32_t)x;
}

// Exclusive scan of one value per thread over a 256-thread workgroup.
// s_w holds 4 words of LDS; *total gets the workgroup's sum.
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t *s_w, uint32_t *total) {
  const uint32_t inc = wave_scan_incl(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) s_w[w] = inc;
  __syncthreads();
  uint32_t off = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t t = s_w[k];
    off += (k < w) ? t : 0u;
    sum += t;
  }
  *total = sum;
  __syncthreads();
  return off + inc - v;
}

// XCD-aware item order. Workgroups go to the 8 XCDs in turn (workgroup b to
// XCD b mod 8), and each XCD has its own L2. A scatter item's output runs
// end where the next item's runs of the same digits begin, so items that
// are neighbours in the output should share an L2: there the two partial
// lines at each seam merge into one whole line before they are written
// back. This maps workgroup b to item k-th-of-its-XCD + the XCD's first item,
// each XCD taking a contiguous range (a bijection on [0, G)).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t G) {
  const uint32_t x = b % kXcds, k = b / kXcds, q = G / kXcds, r = G % kXcds;
  return x * q + min(x, r) + k;
}

// Lanes of this wave whose BITS-bit digit equals this lane's, among `valid`:
// for each bit, the ballot of the wave XOR this lane's bit spread over 32
// bits (0 or ~0) marks the lanes that differ, and those are cleared. Kept in
// two 32-bit halves; each half's "differs |= ballot ^ spread" is one 3-input
// bitwise operation (v_bitop3, truth table 0xF6 over (acc, ballot, spread)),
// and the ballot is taken of the spread bit itself, so a bit costs four VALU
// operations.
template <int BITS>
__device__ __forceinline__ uint64_t match_bits(uint32_t d, uint64_t valid) {
  uint32_t dlo = 0, dhi = 0;  // lanes differing in some bit
#pragma unroll
  for (int b = 0; b < BITS; ++b) {
    const uint32_t s = (uint32_t)__builtin_amdgcn_sbfe((int)d, b, 1);
    // the ballot as a sign test of the spread bit: one compare on s itself
    const uint64_t bal = __builtin_amdgcn_ballot_w64((int)s < 0);
    dlo = __builtin_amdgcn_bitop3_b32(dlo, (uint32_t)bal, s, 0xF6);
    dhi = __builtin_amdgcn_bitop3_b32(dhi, (uint32_t)(bal >> 32), s, 0xF6);
  }
  return ((uint64_t)(~dhi & (uint32_t)(valid >> 32)) << 32) | (~dlo & (uint32_t)valid);
}

__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) { return match_bits<8>(d, valid); }

// ---------------------------------------------------------------------------
// The FP64 definition, for the rare splat the f32 bound cannot decide.
// put() (src/scene/splats.js:350-368), footprint() (src/render/splats.js:
// 102-125) and the area test (:174-178), in the oracle's expression order.
// Returns alpha * area.
// The FP64 footprint (footprint(), src/render/splats.js:102-125) of a splat
// whose Sigma is S, at view position (vx, vy, vz): T = J W, cov = (T S) T^T.
__device__ __forceinline__ void footprint64(const ViewK &v, const double (&S)[3][3], double vx, double vy, double vz,
                                            double &xx, double &xy, double &yy) {
  const double p00 = v.dP[0], p11 = v.dP[5];
  const double fx = p00 * (double)v.vpw * 0.5, fy = p11 * (double)v.vph * 0.5;
  double j00 = fx, j11 = fy, j02 = 0, j12 = 0;
  if (v.persp) {
    const double d = -vz;
    const double lx = kC13 / p00, ly = kC13 / p11;
    double cx = vx / d, cy = vy / d;
    if (cx < -lx) cx = -lx;
    if (cx > lx) cx = lx;
    if (cy < -ly) cy = -ly;
    if (cy > ly) cy = ly;
    const double x = cx * d, y = cy * d;
    j00 = fx / d;
    j11 = fy / d;
    j02 = fx * x / (d * d);
    j12 = fy * y / (d * d);
  }
  double t0[3], t1[3], u0[3], u1[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    t0[c] = j00 * v.dmv[c * 4 + 0] + j02 * v.dmv[c * 4 + 2];
    t1[c] = j11 * v.dmv[c * 4 + 1] + j12 * v.dmv[c * 4 + 2];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    u0[c] = t0[0] * S[0][c] + t0[1] * S[1][c] + t0[2] * S[2][c];
    u1[c] = t1[0] * S[0][c] + t1[1] * S[1][c] + t1[2] * S[2][c];
  }
  xx = u0[0] * t0[0] + u0[1] * t0[1] + u0[2] * t0[2];
  xy = u0[0] * t1[0] + u0[1] * t1[1] + u0[2] * t1[2];
  yy = u1[0] * t1[0] + u1[1] * t1[1] + u1[2] * t1[2];
}

// alpha * area, the cull's value (src/render/splats.js:169-176), in FP64.
__device__ __forceinline__ double cull_value64(double xx, double xy, double yy, uint32_t abyte) {
  const double det = (xx + kC03) * (yy + kC03) - xy * xy;
  const double area = kC2PI * sqrt(det > 0 ? det : 0);
  return (double)abyte / 255.0 * area;
}

__device__ __forceinline__ double canonical_value(const ViewK &v, uint4 a, uint4 b, double vx, double vy, double vz) {
  const double sx = f32_of(a.w), sy = f32_of(b.x), sz = f32_of(b.y);
  const uint32_t q = b.w;
  const double qw = ((int)(q & 255u) - 128) / 128.0, qx = ((int)((q >> 8) & 255u) - 128) / 128.0;
  const double qy = ((int)((q >> 16) & 255u) - 128) / 128.0, qz = ((int)(q >> 24) - 128) / 128.0;
  double length = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  if (length == 0) length = 1;
  const double w = qw / length, u = qx / length, vq = qy / length, t = qz / length;
  const double r00 = 1 - 2 * (vq * vq + t * t), r01 = 2 * (u * vq - w * t), r02 = 2 * (u * t + w * vq);
  const double r10 = 2 * (u * vq + w * t), r11 = 1 - 2 * (u * u + t * t), r12 = 2 * (vq * t - w * u);
  const double r20 = 2 * (u * t - w * vq), r21 = 2 * (vq * t + w * u), r22 = 1 - 2 * (u * u + vq * vq);
  const double m00 = r00 * sx, m01 = r01 * sy, m02 = r02 * sz;
  const double m10 = r10 * sx, m11 = r11 * sy, m12 = r12 * sz;
  const double m20 = r20 * sx, m21 = r21 * sy, m22 = r22 * sz;
  // Sigma as put() stores it: each term rounded to f32 (src/scene/splats.js:318,361-367).
  const double sxx = (float)(m00 * m00 + m01 * m01 + m02 * m02), sxy = (float)(m00 * m10 + m01 * m11 + m02 * m12);
  const double sxz = (float)(m00 * m20 + m01 * m21 + m02 * m22), syy = (float)(m10 * m10 + m11 * m11 + m12 * m12);
  const double syz = (float)(m10 * m20 + m11 * m21 + m12 * m22), szz = (float)(m20 * m20 + m21 * m21 + m22 * m22);
  const double S[3][3] = {{sxx, sxy, sxz}, {sxy, syy, syz}, {sxz, syz, szz}};
  double xx, xy, yy;
  footprint64(v, S, vx, vy, vz, xx, xy, yy);
  return cull_value64(xx, xy, yy, b.z >> 24);
}

// One splat through makeKeys. STD: the projection is the reverse-Z infinite
// perspective (nonzero only at 0, 5, 8, 9, 11 = -1, 14) and the model-view is
// affine, so the zero terms of the FP64 products are left out: adding an
// exact zero changes at most the sign of a zero, which no test here sees.
struct Projected {
  bool keep;
  bool recomputed;
  uint32_t key;
  float xx, xy, yy;
  float cx, cy, cz, cw;  // clip position, the FP64 value rounded to f32
};

// What the footprint needs of a splat whatever the view: its rotation from
// the byte quaternion (unnormalised: the integer products are exact in f32)
// and its scales. A kernel projecting one record for several views makes it
// once.
struct SplatLocal {
  float r00, r01, r02, r10, r11, r12, r20, r21, r22;
  float sx, sy, sz;
};

__device__ __forceinline__ SplatLocal splat_local(uint4 a, uint4 b) {
#pragma clang fp contract(fast)
  SplatLocal L;
  const uint32_t qb = b.w;
  const float qw = (float)(qb & 255u) - 128.f, qx = (float)((qb >> 8) & 255u) - 128.f;
  const float qy = (float)((qb >> 16) & 255u) - 128.f, qz = (float)(qb >> 24) - 128.f;
  const float xx2 = qx * qx, yy2 = qy * qy, zz2 = qz * qz;
  const float n2 = qw * qw + xx2 + yy2 + zz2;
  const float s = n2 > 0.f ? 2.f * __builtin_amdgcn_rcpf(n2) : 0.f;
  L.r00 = 1.f - s * (yy2 + zz2);
  L.r01 = s * (qx * qy - qw * qz);
  L.r02 = s * (qx * qz + qw * qy);
  L.r10 = s * (qx * qy + qw * qz);
  L.r11 = 1.f - s * (xx2 + zz2);
  L.r12 = s * (qy * qz - qw * qx);
  L.r20 = s * (qx * qz - qw * qy);
  L.r21 = s * (qy * qz + qw * qx);
  L.r22 = 1.f - s * (xx2 + yy2);
  L.sx = f32_of(a.w);
  L.sy = f32_of(b.x);
  L.sz = f32_of(b.y);
  return L;
}

// The part of makeKeys every splat form shares (src/render/splats.js:157-181):
// view and clip position, key, and the frustum and alpha tests, all FP64.
template <bool STD>
__device__ __forceinline__ Projected project_front(const ViewK &v, double x, double y, double z, uint32_t abyte,
                                                   double &vx, double &vy, double &vz) {
  // view = modelView * (p, 1); each product is exact, so fma is the sum as written.
  vx = fma(v.dmv[8], z, fma(v.dmv[4], y, v.dmv[0] * x)) + v.dmv[12];
  vy = fma(v.dmv[9], z, fma(v.dmv[5], y, v.dmv[1] * x)) + v.dmv[13];
  vz = fma(v.dmv[10], z, fma(v.dmv[6], y, v.dmv[2] * x)) + v.dmv[14];
  double cx, cy, cz, cw;
  if (STD) {
    cx = v.dP[0] * vx + v.dP[8] * vz;
    cy = v.dP[5] * vy + v.dP[9] * vz;
    cz = v.dP[14];
    cw = -vz;
  } else {
    const double vw = fma(v.dmv[11], z, fma(v.dmv[7], y, v.dmv[3] * x)) + v.dmv[15];
    const double vv[4] = {vx, vy, vz, vw};
    double c[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
      c[r] = v.dP[r] * vv[0] + v.dP[4 + r] * vv[1] + v.dP[8 + r] * vv[2] + v.dP[12 + r] * vv[3];
    cx = c[0];
    cy = c[1];
    cz = c[2];
    cw = c[3];
  }
  // Far first: the farthest key is the smallest (src/render/splats.js:179-181).
  double t = (-vz - v.ddmin) * v.ddscale;
  t = t < 0.0 ? 0.0 : t;
  t = t > 65535.0 ? 65535.0 : t;
  Projected o;
  o.key = 65535u - (uint32_t)t;
  o.recomputed = false;
  o.cx = (float)cx;
  o.cy = (float)cy;
  o.cz = (float)cz;
  o.cw = (float)cw;
  o.xx = o.xy = o.yy = 0.f;
  // bitwise on purpose: every test is evaluated, no branch
  o.keep = !((int)(vz >= 0) | (int)(cz > cw) | (int)(fabs(cx) > kC13 * cw) | (int)(fabs(cy) > kC13 * cw) |
             (int)(abyte <= 1u));
  return o;
}

// A splat given as its centre and its f32 Sigma (the reference's Splats data,
// src/render/splats.js:389-416: centres with the colour in .w, six covariance
// floats xx xy xz yy yz zz): makeKeys with the footprint and cull in FP64.
template <bool STD>
__device__ __forceinline__ Projected project_sigma(const ViewK &v, uint4 centre, const float (&c6)[6]) {
  double vx, vy, vz;
  const uint32_t abyte = centre.w >> 24;
  Projected o = project_front<STD>(v, f32_of(centre.x), f32_of(centre.y), f32_of(centre.z), abyte, vx, vy, vz);
  if (o.keep) {
    const double S[3][3] = {{c6[0], c6[1], c6[2]}, {c6[1], c6[3], c6[4]}, {c6[2], c6[4], c6[5]}};
    double xx, xy, yy;
    footprint64(v, S, vx, vy, vz, xx, xy, yy);
    o.xx = (float)xx;
    o.xy = (float)xy;
    o.yy = (float)yy;
    if (v.cull > 0.f) o.keep = !(cull_value64(xx, xy, yy, abyte) < v.dcull);
  }
  return o;
}

template <bool STD, bool HAVE_LOCAL>
__device__ __forceinline__ Projected project_with(const ViewK &v, uint4 a, uint4 b, const SplatLocal *Lp,
                                                  float bn0w0, float bn1w1, float bw2) {
  double vx, vy, vz;
  const uint32_t abyte = b.z >> 24;
  Projected o = project_front<STD>(v, f32_of(a.x), f32_of(a.y), f32_of(a.z), abyte, vx, vy, vz);
  if (o.keep) {
    // f32 from here to the decision: fusing multiply-adds only tightens the bound below.
#pragma clang fp contract(fast)
    // ---- footprint: cov = diag(j) G G^T diag(j), G = E R S ----
    float j0 = v.fx, j1 = v.fy, xd = 0.f, yd = 0.f, l0 = 0.f, l1 = 0.f;
    if (v.persp) {
      const float id = __builtin_amdgcn_rcpf((float)(-vz));
      const float rx = (float)vx * id, ry = (float)vy * id;
      xd = fminf(fmaxf(rx, -v.limx), v.limx);
      yd = fminf(fmaxf(ry, -v.limy), v.limy);
      l0 = fmaxf(fabsf(rx), v.limx);
      l1 = fmaxf(fabsf(ry), v.limy);
      j0 *= id;
      j1 *= id;
    }
    // E = rows of W (the model-view's 3x3, W[k][c] = mv[c*4+k]) with the
    // perspective's third column folded in: T = diag(j0, j1) E.
    const float e00 = xd * v.mv[2] + v.mv[0], e01 = xd * v.mv[6] + v.mv[4], e02 = xd * v.mv[10] + v.mv[8];
    const float e10 = yd * v.mv[2] + v.mv[1], e11 = yd * v.mv[6] + v.mv[5], e12 = yd * v.mv[10] + v.mv[9];
    const SplatLocal L = HAVE_LOCAL ? *Lp : splat_local(a, b);
    const float r00 = L.r00, r01 = L.r01, r02 = L.r02, r10 = L.r10, r11 = L.r11, r12 = L.r12;
    const float r20 = L.r20, r21 = L.r21, r22 = L.r22, sx = L.sx, sy = L.sy, sz = L.sz;
    const float g00 = (e00 * r00 + e01 * r10 + e02 * r20) * sx;
    const float g01 = (e00 * r01 + e01 * r11 + e02 * r21) * sy;
    const float g02 = (e00 * r02 + e01 * r12 + e02 * r22) * sz;
    const float g10 = (e10 * r00 + e11 * r10 + e12 * r20) * sx;
    const float g11 = (e10 * r01 + e11 * r11 + e12 * r21) * sy;
    const float g12 = (e10 * r02 + e11 * r12 + e12 * r22) * sz;
    o.xx = (j0 * j0) * (g00 * g00 + g01 * g01 + g02 * g02);
    o.xy = (j0 * j1) * (g00 * g10 + g01 * g11 + g02 * g12);
    o.yy = (j1 * j1) * (g10 * g10 + g11 * g11 + g12 * g12);

    if (v.cull > 0.f) {
      // The f32 value and its bound against the FP64 one. With n_r the row
      // norms of E (|W_r| + l_r |W_2|, l_r >= |x/d|), S2 = |s|^2 and u = 2^-24,
      // |xx - xx64| <= 52 u (j0 n0)^2 S2, likewise yy and xy; Sigma's rounding
      // to f32 (|dSigma_ij| <= u |Sigma_ij| <= u S2) adds at most u (j0 n0)^2
      // S2 more; 64 is used.
      const float a0 = j0 * (l0 * bw2 + bn0w0), a1 = j1 * (l1 * bw2 + bn1w1);
      const float c = (64.f * kU) * (sx * sx + sy * sy + sz * sz);
      const float exx = c * a0 * a0, eyy = c * a1 * a1, exy = c * a0 * a1;
      const float X = o.xx + 0.3f, Y = o.yy + 0.3f;
      const float xy2 = o.xy * o.xy;
      const float det = X * Y - xy2;
      const float edet = 1.01f * (Y * exx + X * eyy + exx * eyy + 2.f * fabsf(o.xy) * exy + exy * exy +
                                  4.f * kU * (X * Y + xy2));
      const float root = __builtin_amdgcn_sqrtf(fmaxf(det, 0.f));
      const float alpha2pi = (float)abyte * (6.2831853f / 255.f);
      const float val = alpha2pi * root;
      const float esq = det > 2.f * edet ? edet * __builtin_amdgcn_rsqf(det - edet) : __builtin_amdgcn_sqrtf(3.f * edet);
      const float eval = 1.02f * alpha2pi * (esq + 8.f * kU * root) + 1e-30f;
      if (!(val - v.cull > eval)) {
        if (v.cull - val > eval) {
          o.keep = false;
        } else {
          o.recomputed = true;
          o.keep = !(canonical_value(v, a, b, vx, vy, vz) < v.dcull);
        }
      }
    }
  }
  return o;
}

template <bool STD>
__device__ __forceinline__ Projected project_one(const ViewK &v, uint4 a, uint4 b, float bn0w0, float bn1w1,
                                                 float bw2) {
  return project_with<STD, false>(v, a, b, nullptr, bn0w0, bn1w1, bw2);
}

// ---------------------------------------------------------------------------
// Pass 1 body: project + key over one chunk of kTile splats (16 tiles of 256),
// the chunk the low-byte scatter later takes as one tile. Writes every splat's
// key, survivor bit and cov2d slot (culled splats get zeros, so every cache
// line is written whole, by one wave), and the chunk's two 256-bin digit
// histograms: no global atomics.
// The learned sort's bin key of a 16-bit key: bin << 9 | the key's offset in
// its bin. Bin edges fall on multiples of 16 keys, so a 4096-entry table of
// u16 (the block's bin key >> 4) in LDS gives it with one read and an or.
__device__ __forceinline__ uint32_t bin_key(const uint16_t *blk, uint32_t key) {
  return ((uint32_t)blk[key >> 4] << 4) | (key & 15u);
}

struct ProjectSmem {
  uint32_t h[512];  // survivors per low key byte, then per high key byte
  uint32_t recomp;
};

// ---------------------------------------------------------------------------
// The reference's shade pass for one survivor (src/render/splats.js:315-381),
// fused into the projection: the record's colour plus the harmonics past the
// constant term along the direction from the eye to the splat, clamped to
// [0, 1] and packed as unorm bytes, floor(0.5 + 255 x) (pack4x8unorm), the
// alpha byte kept. f32 in the reference's expression order (no contraction);
// the coefficients are IEEE halves, SHD = 1, 2, 3 (3, 8, 15 coefficients a
// channel), sh_words(SHD) words a splat, coefficient by coefficient, red,
// green and blue.
struct Rgb {
  float r, g, b;
};

// A splat's harmonics row (sh_words(SHD) words), loaded with its record so
// that the row's latency hides behind the tiles worked on meanwhile.
template <int SHD>
struct ShRow {
  uint32_t w[SHD > 0 ? sh_words(SHD) : 1];
};

template <int SHD>
__device__ __forceinline__ uint32_t shade(const ViewK &v, const ShRow<SHD> &row, float px, float py, float pz,
                                          uint32_t rgba) {
  const uint32_t(&w)[SHD > 0 ? sh_words(SHD) : 1] = row.w;
  auto half = [](uint32_t bits) { return (float)__builtin_bit_cast(_Float16, (uint16_t)bits); };
  // coefficient k: three halves from the two words they lie across
  auto coef = [&](uint32_t k) -> Rgb {
    const uint32_t h = k * 3u, a = w[h >> 1], b = w[(h >> 1) + 1];
    if (h & 1u) return Rgb{half(a >> 16), half(b & 0xFFFFu), half(b >> 16)};
    return Rgb{half(a & 0xFFFFu), half(a >> 16), half(b & 0xFFFFu)};
  };
  const float dx = px - v.eye[0], dy = py - v.eye[1], dz = pz - v.eye[2];
  const float len = sqrtf(dx * dx + dy * dy + dz * dz);
  const float x = dx / len, y = dy / len, z = dz / len;
  Rgb c;
  {
    const Rgb c0 = coef(0), c1 = coef(1), c2 = coef(2);
    const float k = 0.4886025119029199f;
    c.r = k * (-y * c0.r + z * c1.r - x * c2.r);
    c.g = k * (-y * c0.g + z * c1.g - x * c2.g);
    c.b = k * (-y * c0.b + z * c1.b - x * c2.b);
  }
  if (SHD >= 2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float a3 = 1.0925484305920792f * x * y, a4 = 1.0925484305920792f * y * z;
    const float a5 = 0.31539156525252005f * (2.0f * zz - xx - yy), a6 = 1.0925484305920792f * x * z;
    const float a7 = 0.5462742152960396f * (xx - yy);
    const Rgb c3 = coef(3), c4 = coef(4), c5 = coef(5), c6 = coef(6), c7 = coef(7);
    c.r += a3 * c3.r - a4 * c4.r + a5 * c5.r - a6 * c6.r + a7 * c7.r;
    c.g += a3 * c3.g - a4 * c4.g + a5 * c5.g - a6 * c6.g + a7 * c7.g;
    c.b += a3 * c3.b - a4 * c4.b + a5 * c5.b - a6 * c6.b + a7 * c7.b;
    if (SHD >= 3) {
      const float a8 = -0.5900435899266435f * y * (3.0f * xx - yy), a9 = 2.890611442640554f * x * y * z;
      const float a10 = 0.4570457994644658f * y * (4.0f * zz - xx - yy);
      const float a11 = 0.3731763325901154f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
      const float a12 = 0.4570457994644658f * x * (4.0f * zz - xx - yy), a13 = 1.445305721320277f * z * (xx - yy);
      const float a14 = 0.5900435899266435f * x * (xx - 3.0f * yy);
      const Rgb c8 = coef(8), c9 = coef(9), c10 = coef(10), c11 = coef(11), c12 = coef(12), c13 = coef(13),
                c14 = coef(14);
      c.r += a8 * c8.r + a9 * c9.r - a10 * c10.r + a11 * c11.r - a12 * c12.r + a13 * c13.r - a14 * c14.r;
      c.g += a8 * c8.g + a9 * c9.g - a10 * c10.g + a11 * c11.g - a12 * c12.g + a13 * c13.g - a14 * c14.g;
      c.b += a8 * c8.b + a9 * c9.b - a10 * c10.b + a11 * c11.b - a12 * c12.b + a13 * c13.b - a14 * c14.b;
    }
  }
  auto byte = [](uint32_t word, int k, float add) {
    const float base = (float)((word >> (8 * k)) & 255u) / 255.0f;
    const float seen = fminf(fmaxf(base + add, 0.0f), 1.0f);
    return (uint32_t)floorf(0.5f + 255.0f * seen);
  };
  return byte(rgba, 0, c.r) | (byte(rgba, 1, c.g) << 8) | (byte(rgba, 2, c.b) << 16) | (rgba & 0xFF000000u);
}

// The splat vertex shader's per-instance set-up (src/render/splats.js:263-291)
// for a survivor, made once here beside cov2d so that a rasteriser reads it
// instead of projecting the splat again. From the footprint plus a third of a
// pixel each way, C = [[a, b], [b, c]]: the eigen-axes, each as long as
// sqrt(2 lambda) capped at 1024 px, scaled by how far the splat reaches,
// e = sqrt(min(log(255 alpha), 4)). Writes (along e, across e) in pixels; the
// quad's corner q in {-1, 1}^2 is at clip.xy + (q.x along e + q.y across e)
// 2 / viewport clip.w, with the Gaussian's offset q e. Zeros when the shader
// would draw nothing (minor axis not > 0, or reach not > 0: alpha byte < 2).
// f32 with the hardware square root and reciprocal square root (about 1 ulp
// each); e comes from a 256-entry table of the FP64 value per alpha byte
// (reach[0] = reach[1] = 0). The shader's direction (b, major - a) loses its
// second component to cancellation when |b| << a - c, so for a >= c it is
// taken as the same direction written without the difference,
// (sign(b) (radius + h), |b|), since major - a = radius - h = b^2 / (radius + h).
__device__ __forceinline__ f32x4 ellipse_axes(float xx, float xy, float yy, uint32_t abyte,
                                              const float *__restrict__ reach) {
  const float a = xx + 0.3f, b = xy, c = yy + 0.3f;
  const float mid = 0.5f * (a + c);
  const float h = 0.5f * (a - c);
  const float radius = __builtin_amdgcn_sqrtf(h * h + b * b);
  const float major = mid + radius, minor = mid - radius;
  const float e = reach[abyte];
  if (!(minor > 0.0f) || !(e > 0.0f)) return f32x4{0.f, 0.f, 0.f, 0.f};
  float ax = 1.0f, ay = 0.0f;
  if (fabsf(b) > 1e-12f) {
    const float ux = h >= 0.0f ? copysignf(radius + h, b) : b;
    const float uy = h >= 0.0f ? fabsf(b) : radius - h;
    const float inv = __builtin_amdgcn_rsqf(ux * ux + uy * uy);
    ax = ux * inv;
    ay = uy * inv;
  } else if (c > a) {
    ax = 0.0f;
    ay = 1.0f;
  }
  const float la = fminf(__builtin_amdgcn_sqrtf(2.0f * major), 1024.0f) * e;
  const float lc = fminf(__builtin_amdgcn_sqrtf(2.0f * minor), 1024.0f) * e;
  return f32x4{la * ax, la * ay, lc * ay, -(lc * ax)};
}

// BINS (the learned sort): the chunk's histogram is over the view's 512 bins
// instead of over the two key bytes; a survivor's bin comes from the view's
// block table, already in LDS (bin_key).
// SHD > 0: the survivors' colour is shaded by the cloud's harmonics (shade).
// ELL: each splat's draw set-up goes to two planes of n_cap float4s, its clip
// position to ell[i] and its ellipse_axes to ell[n_cap + i] (each store of a
// wave covers whole lines), zeros for one not drawn.
// SHPRE (SHD > 0): each splat's harmonics row is loaded with its record, two
// tiles ahead; off, a survivor's row is loaded when it is shaded. A kernel
// whose one body serves every degree (the batch) keeps it off: its registers
// are the most any branch needs, and three tiles of rows would halve its
// occupancy for clouds without harmonics too.
template <bool STD, bool BINS = false, int SHD = 0, bool ELL = false, bool SIG = false, bool SHPRE = true>
__device__ __forceinline__ void project_chunk(ProjectSmem &sm, const uint4 *__restrict__ rec, uint32_t n,
                                              const ViewK &v, uint16_t *__restrict__ keys,
                                              uint64_t *__restrict__ mask, float4 *__restrict__ cov,
                                              uint16_t *__restrict__ tilehist, uint32_t *__restrict__ recomp_acc,
                                              uint32_t chunk, const uint16_t *blk = nullptr,
                                              const uint32_t *__restrict__ sh = nullptr,
                                              float4 *__restrict__ ell = nullptr,
                                              const float *__restrict__ reach = nullptr,
                                              const uint2 *__restrict__ sig = nullptr, uint64_t n_cap = 0) {
  const uint32_t tid = threadIdx.x;
  __syncthreads();  // the LDS may still be read by the previous work item
  for (uint32_t k = tid; k < 512; k += 256) sm.h[k] = 0;
  if (tid == 0) sm.recomp = 0;
  __syncthreads();
  // Per-view parts of the bound's row norms.
  const float bw0 = fabsf(v.mv[0]) + fabsf(v.mv[4]) + fabsf(v.mv[8]);
  const float bw1 = fabsf(v.mv[1]) + fabsf(v.mv[5]) + fabsf(v.mv[9]);
  const float bw2 = fabsf(v.mv[2]) + fabsf(v.mv[6]) + fabsf(v.mv[10]);
  uint32_t recomp = 0;
  const uint32_t first = chunk * (kTile / 256), end = min(first + kTile / 256, (n + 255u) >> 8);
  // Three register buffers: two tiles' records are in flight while a tile is
  // worked on, which hides the load latency at four waves a SIMD better than
  // one tile in flight does at five.
  // SIG: rec holds the centres (16 bytes a splat) and sig the six covariance
  // floats (24 bytes a splat, three 8-byte loads); a is the centre, b and c Sigma.
  // SHD > 0: the slot holds the splat's harmonics row too (loaded for every
  // splat of the tile: the cull is not known yet).
  struct Slot {
    uint4 a, b;
    uint2 c;
    ShRow<SHPRE ? SHD : 0> h;
  };
  auto load = [&](uint32_t tile, Slot &s) {
    const uint32_t i = (tile << 8) + tid;
    if (tile < end && i < n) {
      if constexpr (SIG) {
        s.a = rec[i];
        const uint2 s0 = sig[3 * (size_t)i], s1 = sig[3 * (size_t)i + 1];
        s.b = make_uint4(s0.x, s0.y, s1.x, s1.y);
        s.c = sig[3 * (size_t)i + 2];
      } else {
        s.a = rec[2 * (size_t)i];
        s.b = rec[2 * (size_t)i + 1];
      }
      if constexpr (SHD > 0 && SHPRE) {
        const uint32_t *row = sh + (size_t)i * sh_words(SHD);
#pragma unroll
        for (uint32_t k = 0; k < sh_words(SHD); ++k) s.h.w[k] = row[k];
      }
    }
  };
  auto work = [&](uint32_t tile, const Slot &s) {
    const uint4 &a = s.a, &b = s.b;
    const uint32_t i = (tile << 8) + tid;
    const bool in = i < n;
    Projected p;
    uint32_t rgba;
    if constexpr (SIG) {
      const float c6[6] = {__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z),
                           __uint_as_float(b.w), __uint_as_float(s.c.x), __uint_as_float(s.c.y)};
      p = project_sigma<STD>(v, a, c6);
      rgba = a.w;
    } else {
      p = project_one<STD>(v, a, b, bw0, bw1, bw2);
      rgba = b.z;
    }
    p.keep = p.keep && in;
    const uint64_t bal = __ballot(p.keep);
    if ((tid & 63u) == 0 && in) mask[i >> 6] = bal;
    if (in) {
      __builtin_nontemporal_store((uint16_t)p.key, &keys[i]);
      // Streaming stores for the keys and cov2d (the sort never reads cov2d back; the
      // keys are read once, by the next pass): lines that sit dirty in L2 at the end
      // of the launch are written back before the next launch starts.
      uint32_t colour = rgba;
      if constexpr (SHD > 0)
        if (p.keep) {
          if constexpr (SHPRE) {
            colour = shade<SHD>(v, s.h, f32_of(a.x), f32_of(a.y), f32_of(a.z), rgba);
          } else {
            ShRow<SHD> row;
            const uint32_t *src = sh + (size_t)i * sh_words(SHD);
#pragma unroll
            for (uint32_t k = 0; k < sh_words(SHD); ++k) row.w[k] = src[k];
            colour = shade<SHD>(v, row, f32_of(a.x), f32_of(a.y), f32_of(a.z), rgba);
          }
        }
      const f32x4 c = p.keep ? f32x4{p.xx, p.xy, p.yy, __uint_as_float(colour)} : f32x4{0.f, 0.f, 0.f, 0.f};
      __builtin_nontemporal_store(c, (f32x4 *)&cov[i]);
      if constexpr (ELL) {
        const f32x4 zero{0.f, 0.f, 0.f, 0.f};
        const f32x4 axes = p.keep ? ellipse_axes(p.xx, p.xy, p.yy, rgba >> 24, reach) : zero;
        const bool drawn = axes[0] != 0.f || axes[1] != 0.f;
        __builtin_nontemporal_store(drawn ? f32x4{p.cx, p.cy, p.cz, p.cw} : zero, (f32x4 *)&ell[i]);
        __builtin_nontemporal_store(axes, (f32x4 *)&ell[n_cap + i]);
      }
    }
    recomp += (p.recomputed && in) ? 1u : 0u;
    if (BINS) {
      if (p.keep) atomicAdd(&sm.h[bin_key(blk, p.key) >> 9], 1u);
    } else if (p.keep) {
      atomicAdd(&sm.h[p.key & 255u], 1u);
      atomicAdd(&sm.h[256u + (p.key >> 8)], 1u);
    }
  };
  Slot s0{make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint2(0, 0), ShRow<SHPRE ? SHD : 0>{}}, s1 = s0,
       s2 = s0;
  load(first, s0);
  load(first + 1, s1);
  for (uint32_t tile = first; tile < end; tile += 3) {
    load(tile + 2, s2);
    work(tile, s0);
    load(tile + 3, s0);
    if (tile + 1 < end) work(tile + 1, s1);
    load(tile + 4, s1);
    if (tile + 2 < end) work(tile + 2, s2);
  }
  if (recomp) atomicAdd(&sm.recomp, recomp);
  __syncthreads();
  for (uint32_t k = tid; k < 512; k += 256) tilehist[(size_t)chunk * 512 + k] = (uint16_t)sm.h[k];
  if (tid == 0 && sm.recomp) atomicAdd(recomp_acc, sm.recomp);
}

// ---------------------------------------------------------------------------
// Scans between the passes. A workgroup takes C adjacent columns (digits) of
// a [T][stride] u16 histogram, one row block a thread: thread t sums rows
// [t R, (t + 1) R), R = ceil(T / 256), with one load a row (8 bytes for C = 4,
// 16 for C = 8), the workgroup scans the C column sums (a wave scan and one
// barrier), and the thread writes its rows' exclusive prefixes,
// prefix[r * pstride + c], in 16-byte stores, from the rows it still holds
// (up to 16 rows a thread for C = 4, 8 for C = 8, so 32 or 64 registers;
// beyond that it reads them again). Only columns below pcols get prefixes;
// every column's sum goes to total[c]. No workgroup waits on another: a
// scatter workgroup turns the totals into digit bases itself (digit_bases), so
// no fence, look-back or last-block step is needed. A single sort's scans run
// in one round of workgroups and take C = 4 (kScanCols); a batch of many jobs
// takes C = 8 (kBatchScanCols), which halves its workgroups and its rounds.
template <uint32_t C>
__device__ __forceinline__ void scan_cols(uint32_t c0, uint32_t T, const uint16_t *__restrict__ hist, uint32_t stride,
                                          uint32_t *__restrict__ prefix, uint32_t pstride, uint32_t pcols,
                                          uint32_t *__restrict__ total, uint32_t (*s_w)[C]) {
  static_assert(C == 4 || C == 8, "one 8- or 16-byte load a row");
  using Row = typename std::conditional<C == 8, uint4, uint2>::type;
  constexpr uint32_t kHeld = C == 8 ? 8 : 16;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t R = (T + 255u) >> 8;
  const uint32_t r0 = min(tid * R, T), r1 = min(r0 + R, T);
  auto row = [&](uint32_t r) { return *reinterpret_cast<const Row *>(hist + (size_t)r * stride + c0); };
  auto add = [](uint32_t (&acc)[C], const Row &x) {
    const uint32_t *v = reinterpret_cast<const uint32_t *>(&x);
#pragma unroll
    for (uint32_t i = 0; i < C / 2; ++i) {
      acc[2 * i] += v[i] & 0xFFFFu;
      acc[2 * i + 1] += v[i] >> 16;
    }
  };
  uint32_t sum[C];
#pragma unroll
  for (uint32_t c = 0; c < C; ++c) sum[c] = 0;
  Row held[kHeld];
  const bool hold = R <= kHeld;
  if (hold && T > 0) {
    // every row is loaded before any is added: a thread's loads are in flight together
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) {
      const bool ok = r0 + k < r1;
      const Row x = row(ok ? r0 + k : 0u);  // row 0 exists when T > 0
      held[k] = ok ? x : Row{};
    }
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) add(sum, held[k]);
  } else {
    for (uint32_t r = r0; r < r1; ++r) add(sum, row(r));
  }
  uint32_t inc[C];
#pragma unroll
  for (uint32_t c = 0; c < C; ++c) inc[c] = wave_scan_incl(sum[c]);
  if (lane == 63u)
#pragma unroll
    for (uint32_t c = 0; c < C; ++c) s_w[w][c] = inc[c];
  __syncthreads();
  uint32_t run[C];
#pragma unroll
  for (uint32_t c = 0; c < C; ++c) {
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t t = s_w[k][c];
      off += k < w ? t : 0u;
      tot += t;
    }
    run[c] = off + inc[c] - sum[c];
    if (tid == c) total[c0 + c] = tot;
  }
  if (!prefix || c0 >= pcols || T == 0) return;
  auto put = [&](uint32_t r) {
    uint4 *q = reinterpret_cast<uint4 *>(prefix + (size_t)r * pstride + c0);
#pragma unroll
    for (uint32_t i = 0; i < C / 4; ++i) q[i] = make_uint4(run[4 * i], run[4 * i + 1], run[4 * i + 2], run[4 * i + 3]);
  };
  if (hold) {
#pragma unroll
    for (uint32_t k = 0; k < kHeld; ++k) {
      if (r0 + k < r1) put(r0 + k);
      add(run, held[k]);
    }
  } else {
    for (uint32_t r = r0; r < r1; ++r) {
      put(r);
      add(run, row(r));
    }
  }
}

// Where digit d (this thread's) starts: the exclusive scan of the 256 digit
// totals; *sum gets V.
__device__ __forceinline__ uint32_t digit_bases(uint32_t total_d, uint32_t *s_w, uint32_t *sum) {
  return block_scan_excl(total_d, s_w, sum);
}

// ---------------------------------------------------------------------------
// The high-byte histogram of tile t of the low-byte pass's output (positions
// t * kTile ...), the upsweep of the high-byte pass: counts only, so plain
// LDS adds, each wave into its own row (a slab seen edge on, whose survivors
// share a few high bytes, then contends within a wave only). Same-address
// adds of one instruction serialise in the LDS, and that still costs less
// than peeling the hot digits off with ballots first (measured on a slab).
struct HistSmem {
  uint32_t h[4][256];
};

// Pack: the low-byte pass's element, high byte << 24 | index (uint32_t), or,
// for clouds past 2^24 splats, high byte << 32 | index (uint64_t).
template <typename Pack>
__device__ __forceinline__ uint32_t pack_digit(Pack v) {
  return sizeof(Pack) == 8 ? (uint32_t)((uint64_t)v >> 32) : (uint32_t)v >> 24;
}

// RUNS: the thread's digits are sorted first and each run of equal digits is
// one add (below); measured on MI355X it saves 1.4 µs of 12.7 on slab-8m,
// whose survivors share a few high bytes, and costs 5 µs on batch-64, whose
// scenes' high bytes spread, so the single sort takes it and the batch not.
template <typename Pack, bool RUNS = false>
__device__ __forceinline__ void hist_tile(HistSmem &sm, uint32_t t, const Pack *__restrict__ packed,
                                          const Ctl *__restrict__ ctl, uint16_t *__restrict__ tilehist3) {
  constexpr int KPT = kKeysPerThread;
  const uint32_t tid = threadIdx.x, w = tid >> 6;
  const uint32_t limit = ctl->count, tile0 = t * (uint32_t)kTile;
  if (tile0 >= limit) return;
  // Counting needs no order: thread tid takes 16 consecutive elements, in 16-byte loads.
  uint32_t raw[KPT];  // the digits
  const uint32_t first = tile0 + tid * (uint32_t)KPT;
  if (first + KPT <= limit) {
    const uint4 *q = reinterpret_cast<const uint4 *>(packed + first);
    constexpr int kPer = 16 / (int)sizeof(Pack);  // elements a 16-byte load
#pragma unroll
    for (int k = 0; k < KPT / kPer; ++k) {
      const uint4 v = q[k];
      if (sizeof(Pack) == 8) {
        raw[2 * k] = v.y;
        raw[2 * k + 1] = v.w;
      } else {
        raw[4 * k] = v.x >> 24;
        raw[4 * k + 1] = v.y >> 24;
        raw[4 * k + 2] = v.z >> 24;
        raw[4 * k + 3] = v.w >> 24;
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < KPT; ++r) raw[r] = first + r < limit ? pack_digit(packed[first + r]) : 0u;
  }
  __syncthreads();  // the LDS may still be read by the previous work item
  for (uint32_t k = tid; k < 1024; k += 256) (&sm.h[0][0])[k] = 0;
  __syncthreads();
  if (!RUNS) {
#pragma unroll
    for (int r = 0; r < KPT; ++r)
      if (first + r < limit) atomicAdd(&sm.h[w][raw[r]], 1u);
    __syncthreads();
    tilehist3[(size_t)t * 256 + tid] = (uint16_t)(sm.h[0][tid] + sm.h[1][tid] + sm.h[2][tid] + sm.h[3][tid]);
    return;
  }
  // The thread's digits sorted in registers (a bitonic network, min and max a
  // comparator), so each run of equal digits costs one LDS add of its length:
  // a slab's few high bytes then collide in the LDS a few times a thread, not
  // KPT times. Elements past V sort last as 256 and are not counted.
#pragma unroll
  for (int r = 0; r < KPT; ++r)
    if (first + r >= limit) raw[r] = 256u;
#pragma unroll
  for (int k = 2; k <= KPT; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < KPT; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const uint32_t a = raw[i], b = raw[l];
          const bool up = (i & k) == 0;
          raw[i] = up ? min(a, b) : max(a, b);
          raw[l] = up ? max(a, b) : min(a, b);
        }
      }
  uint32_t run = 1;
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    if (r == KPT - 1 || raw[r] != raw[r + 1 < KPT ? r + 1 : r]) {
      if (raw[r] < 256u) atomicAdd(&sm.h[w][raw[r]], run);
      run = 1;
    } else {
      ++run;
    }
  }
  __syncthreads();
  tilehist3[(size_t)t * 256 + tid] = (uint16_t)(sm.h[0][tid] + sm.h[1][tid] + sm.h[2][tid] + sm.h[3][tid]);
}

// ---------------------------------------------------------------------------
// Pass 2 and 3 body: one stable counting pass over an 8-bit digit on super-
// tile st, S consecutive tiles of kTile. HI = false: input is every splat in
// index order (keys + survivor bits), digit = key & 255, output high byte
// << 24 | index. HI = true: input is that output (V from the control block),
// digit = value >> 24, output the index. Each tile is ranked within its waves
// by a ballot match; the super-tile's counts per digit (known beforehand from
// the tile histograms) place every element in LDS, and the super-tile is then
// written out in digit runs S times longer than one tile's.
//
// One LDS word an element: digit << 24 | payload. HI: the payload is the
// index. LO: the high byte << kLocalBits | the index within the super-tile.
// 40 KB at S = 2, four workgroups a CU.
template <int S, bool WIDE = false>
struct ScatterSmem {
  uint32_t ent[S * kTile];
  uint8_t dig[WIDE ? S * kTile : 4];  // WIDE, high-byte pass: each slot's digit (ent holds the index)
  uint32_t wh[4][256];   // per-wave running digit counts, then wave offsets
  uint32_t sub[S][256];  // each tile's start of digit d in the super-tile's LDS order
  uint32_t shift[256];   // global start - LDS start of digit d
  uint32_t w[4];
  uint32_t nvalid;
};

// WIDE: the elements between the passes are 64-bit (high byte << 32 |
// index), so indices may pass 2^24; the high-byte pass then keeps each slot's
// digit in its own LDS byte. out: the low-byte pass's elements (Pack) or the
// draw order (uint32_t).
template <bool HI, int S, bool WIDE = false>
__device__ __forceinline__ void scatter_super(ScatterSmem<S, WIDE> &sm, uint32_t st,
                                              const uint16_t *__restrict__ keys, const uint64_t *__restrict__ mask,
                                              const typename std::conditional<WIDE, uint64_t, uint32_t>::type
                                                  *__restrict__ in_packed,
                                              uint32_t n, const uint16_t *__restrict__ hist, uint32_t hstride,
                                              const uint32_t *__restrict__ prefix, const uint32_t *__restrict__ total,
                                              Ctl *__restrict__ ctl, void *__restrict__ out_v, uint32_t out_base = 0) {
  using Pack = typename std::conditional<WIDE, uint64_t, uint32_t>::type;
  constexpr int KPT = kKeysPerThread;
  constexpr uint32_t kLocalBits = S * kTile <= 8192 ? 13u : 14u;
  static_assert(S * kTile <= 16384, "index within the super-tile in 14 bits");
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  const uint32_t limit = HI ? ctl->count : n;
  const uint32_t T = (limit + kTile - 1) / kTile;
  const uint32_t t0 = st * (uint32_t)S;
  if (t0 >= T) return;  // past the survivors (HI)
  const uint32_t ns = min((uint32_t)S, T - t0);
  const uint32_t d = tid;
  const uint32_t nwords = (n + 63u) >> 6;
  // Every global read of the item is issued here, before any is waited on:
  // the tiles' elements (and survivor words), the digit counts, the starts.
  uint32_t raw[S][KPT];  // HI: the element (WIDE: its index); LO: the key
  uint32_t dgt[S][KPT];  // HI and WIDE: the element's digit
#pragma unroll
  for (int s = 0; s < S; ++s)
#pragma unroll
    for (int r = 0; r < KPT; ++r) dgt[s][r] = 0u;
  uint64_t vm[S];  // LO: lane k < 16 holds the survivor word of round k of this wave
  uint32_t cnt[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const uint32_t wbase = (t0 + s) * (uint32_t)kTile + w * (64u * KPT);
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
      const uint32_t i = wbase + (uint32_t)r * 64u + lane;
      if (HI) {
        const Pack v = i < limit ? in_packed[i] : (Pack)0;
        raw[s][r] = (uint32_t)v;
        if (WIDE) dgt[s][r] = pack_digit(v);
      } else {
        raw[s][r] = i < n ? (uint32_t)keys[i] : 0u;
      }
    }
    if (!HI) {
      const uint32_t wi = (wbase >> 6) + lane;
      vm[s] = (lane < (uint32_t)KPT && wi < nwords && (uint32_t)s < ns) ? mask[wi] : 0ull;
    }
    cnt[s] = (uint32_t)s < ns ? (uint32_t)hist[(size_t)(t0 + s) * hstride + d] : 0u;
  }
  const uint32_t dtotal = total[d], dprefix = prefix[(size_t)t0 * 256 + d];
  __syncthreads();  // the LDS may still be read by the previous work item
  uint32_t V;
  const uint32_t gbase = digit_bases(dtotal, sm.w, &V) + dprefix;
  if (!HI && st == 0 && tid == 0) ctl->count = V;  // read by the passes after this one
  uint32_t tot = 0;
  uint32_t subs[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    subs[s] = tot;
    tot += cnt[s];
  }
  uint32_t nvalid;
  const uint32_t lstart = block_scan_excl(tot, sm.w, &nvalid);
#pragma unroll
  for (int s = 0; s < S; ++s) sm.sub[s][d] = lstart + subs[s];
  sm.shift[d] = gbase - lstart;
  if (tid == 0) sm.nvalid = nvalid;
  const uint32_t sbase = t0 * (uint32_t)kTile;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if ((uint32_t)s >= ns) break;
    const uint32_t wbase = (t0 + s) * (uint32_t)kTile + w * (64u * KPT);
    for (uint32_t k = tid; k < 1024; k += 256) (&sm.wh[0][0])[k] = 0;
    __syncthreads();
    uint32_t rank[KPT];
    uint32_t okbits = 0;
    // FULL: every element of the tile is valid (the high-byte pass away from the
    // end of V), so the round needs no ballot of the bounds and no mask.
    auto rank_rounds = [&](auto full_c) {
      constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
      for (int r = 0; r < KPT; ++r) {
        const uint32_t i = wbase + (uint32_t)r * 64u + lane;
        uint64_t valid;
        uint32_t dd;
        if (HI) {
          valid = FULL ? ~0ull : __ballot(i < limit);
          dd = WIDE ? dgt[s][r] : raw[s][r] >> 24;
        } else {
          const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)vm[s], r);
          const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(vm[s] >> 32), r);
          valid = ((uint64_t)hi << 32) | lo;
          dd = raw[s][r] & 255u;
        }
        const bool ok = FULL ? true : ((valid >> lane) & 1ull) != 0;
        const uint64_t peers = match_digit(dd, valid);
        const uint32_t below = count_below(peers);
        const uint32_t before = sm.wh[w][dd];  // any digit is a valid index: read unconditionally
        if (ok && below == 0) sm.wh[w][dd] = before + (uint32_t)__popcll(peers);
        rank[r] = before + below;
        okbits |= (ok ? 1u : 0u) << r;
      }
    };
    if (HI && (t0 + s + 1) * (uint32_t)kTile <= limit)
      rank_rounds(std::integral_constant<bool, true>{});
    else
      rank_rounds(std::integral_constant<bool, false>{});
    __syncthreads();
    {
      // wave offsets of digit d within this tile, plus the tile's start
      const uint32_t c0 = sm.wh[0][d], c1 = sm.wh[1][d], c2 = sm.wh[2][d];
      const uint32_t base = sm.sub[s][d];
      sm.wh[0][d] = base;
      sm.wh[1][d] = base + c0;
      sm.wh[2][d] = base + c0 + c1;
      sm.wh[3][d] = base + c0 + c1 + c2;
    }
    __syncthreads();
    // every slot read first, then every write: no read waits behind a write
    uint32_t pos[KPT];
#pragma unroll
    for (int r = 0; r < KPT; ++r)
      pos[r] = sm.wh[w][HI ? (WIDE ? dgt[s][r] : raw[s][r] >> 24) : raw[s][r] & 255u] + rank[r];
#pragma unroll
    for (int r = 0; r < KPT; ++r) {
      const uint32_t i = wbase + (uint32_t)r * 64u + lane;
      const uint32_t dd = HI ? (WIDE ? dgt[s][r] : raw[s][r] >> 24) : raw[s][r] & 255u;
      const uint32_t e = HI ? raw[s][r] : (dd << 24) | ((raw[s][r] >> 8) << kLocalBits) | (i - sbase);
      // pos is always in range when the counts agree; the test is a guard
      if (((okbits >> r) & 1u) && pos[r] < (uint32_t)(S * kTile)) {
        sm.ent[pos[r]] = e;
        if (HI && WIDE) sm.dig[pos[r]] = (uint8_t)dd;
      }
    }
    __syncthreads();
  }
  const uint32_t nv = min(sm.nvalid, (uint32_t)(S * kTile));
  // eight slots a thread at a time, so the two LDS reads of each overlap the others'
  for (uint32_t p0 = 0; p0 < nv; p0 += 8 * 256) {
    uint32_t e[8], g[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t p = p0 + (uint32_t)k * 256u + tid;
      e[k] = p < nv ? sm.ent[p] : 0u;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t p = p0 + (uint32_t)k * 256u + tid;
      const uint32_t dd = HI && WIDE ? (p < nv ? (uint32_t)sm.dig[p] : 0u) : e[k] >> 24;
      g[k] = sm.shift[dd] + p;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t p = p0 + (uint32_t)k * 256u + tid;
      // n bounds both outputs; a guard against a miscount. The draw order is the
      // path's output: streamed out, so that the next launch does not start behind
      // its write-back.
      if (p < nv && g[k] < n) {
        if (HI) {
          const uint32_t index = (WIDE ? e[k] : e[k] & 0xFFFFFFu) + out_base;
          __builtin_nontemporal_store(index, &static_cast<uint32_t *>(out_v)[g[k]]);
        } else {
          const uint32_t hb = (e[k] >> kLocalBits) & 255u, index = sbase + (e[k] & ((1u << kLocalBits) - 1u));
          if (WIDE)
            static_cast<uint64_t *>(out_v)[g[k]] = ((uint64_t)hb << 32) | index;
          else
            static_cast<uint32_t *>(out_v)[g[k]] = (hb << 24) | index;
        }
      }
    }
  }
}

}  // namespace
}  // namespace wd

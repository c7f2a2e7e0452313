// gfx950 kernels of the splat hot path: project + key, then a two-pass
// stable LSD scatter over the 16-bit key with decoupled look-back.
//
//   k_project   one thread a splat, reading its 32-byte record once: the
//               reference's makeKeys (src/render/splats.js:155-186) with the
//               footprint (:102-125) fused in. Writes the 16-bit key of every
//               splat, a survivor bit per splat (a wave ballot, one u64 per 64
//               splats), (xx, xy, yy, rgba) for survivors, and counts both key
//               bytes in LDS. The last workgroup to finish scans the two
//               256-bin histograms into digit bases: the reference's
//               scanCounts (:188-212), which also zeroes the counts for the
//               next sort, as this does.
//   k_scatter   the reference's place (:214-219), made stable: tiles of 4096
//               ranked within each wave by a ballot match on the 8-bit digit,
//               tile offsets by decoupled look-back, staged through LDS and
//               written out in digit runs. Run once on the low byte (splats in
//               index order in, high byte << 24 | index out) and once on the
//               high byte (order out). LSD passes are stable, so the result
//               is key ascending, then index ascending.
//
// Numerics. Every decision equals the FP64 evaluation of the reference's
// formulas on the f32 inputs (the definition the CPU oracle restates):
//   - view and clip coordinates and the key are computed in FP64 here too. The
//     view's products are of two f32 values, exact in f64, so fma and mul+add
//     agree; everything else is built with -ffp-contract=off.
//   - the area cull runs in f32 with a proven bound on its error against the
//     FP64 value; only a splat whose value lies inside the bound is decided
//     again in FP64 (canonical_value, rare), so the f32 fast path never
//     changes a decision.
#include <hip/hip_runtime.h>

#include "wd_internal.h"

namespace wd {
namespace {

constexpr double kC13 = (double)1.3f;
constexpr double kC03 = (double)0.3f;
constexpr double kC2PI = (double)6.2831853f;
constexpr float kU = 5.9604644775390625e-8f;  // 2^-24, unit roundoff of f32

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

// Lanes of this wave whose 8-bit digit equals this lane's, among `valid`.
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t valid) {
  uint64_t m = valid;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

__device__ __forceinline__ uint64_t load_status(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint64_t *p, uint32_t epoch, uint32_t flag, uint32_t value) {
  const uint64_t w = ((uint64_t)((epoch << 2) | flag) << 32) | value;
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// The FP64 definition, for the rare splat the f32 bound cannot decide.
// put() (src/scene/splats.js:350-368), footprint() (src/render/splats.js:
// 102-125) and the area test (:174-178), in the oracle's expression order.
// Returns alpha * area.
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
  const double S[3][3] = {
      {m00 * m00 + m01 * m01 + m02 * m02, m00 * m10 + m01 * m11 + m02 * m12, m00 * m20 + m01 * m21 + m02 * m22},
      {m00 * m10 + m01 * m11 + m02 * m12, m10 * m10 + m11 * m11 + m12 * m12, m10 * m20 + m11 * m21 + m12 * m22},
      {m00 * m20 + m01 * m21 + m02 * m22, m10 * m20 + m11 * m21 + m12 * m22, m20 * m20 + m21 * m21 + m22 * m22}};
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
  const double xx = u0[0] * t0[0] + u0[1] * t0[1] + u0[2] * t0[2];
  const double xy = u0[0] * t1[0] + u0[1] * t1[1] + u0[2] * t1[2];
  const double yy = u1[0] * t1[0] + u1[1] * t1[1] + u1[2] * t1[2];
  const double det = (xx + kC03) * (yy + kC03) - xy * xy;
  const double area = kC2PI * sqrt(det > 0 ? det : 0);
  const double alpha = (double)(b.z >> 24) / 255.0;
  return alpha * area;
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
};

template <bool STD>
__device__ __forceinline__ Projected project_one(const ViewK &v, uint4 a, uint4 b, float bn0w0, float bn1w1,
                                                 float bw2) {
  Projected o;
  o.keep = false;
  o.recomputed = false;
  o.key = 0;
  o.xx = o.xy = o.yy = 0.f;
  const double x = f32_of(a.x), y = f32_of(a.y), z = f32_of(a.z);
  // view = modelView * (p, 1); each product is exact, so fma is the sum as written.
  const double vx = fma(v.dmv[8], z, fma(v.dmv[4], y, v.dmv[0] * x)) + v.dmv[12];
  const double vy = fma(v.dmv[9], z, fma(v.dmv[5], y, v.dmv[1] * x)) + v.dmv[13];
  const double vz = fma(v.dmv[10], z, fma(v.dmv[6], y, v.dmv[2] * x)) + v.dmv[14];
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
  const uint32_t abyte = b.z >> 24;
  if (vz >= 0 || cz > cw || fabs(cx) > kC13 * cw || fabs(cy) > kC13 * cw || abyte <= 1u) return o;

  // ---- footprint in f32: cov = diag(j) G G^T diag(j), G = E R S ----
  float j0, j1, xd = 0.f, yd = 0.f, l0 = 0.f, l1 = 0.f;
  if (v.persp) {
    const float d = (float)(-vz);
    const float id = __builtin_amdgcn_rcpf(d);
    const float rx = (float)vx * id, ry = (float)vy * id;
    xd = fminf(fmaxf(rx, -v.limx), v.limx);
    yd = fminf(fmaxf(ry, -v.limy), v.limy);
    l0 = fmaxf(fabsf(rx), v.limx);
    l1 = fmaxf(fabsf(ry), v.limy);
    j0 = v.fx * id;
    j1 = v.fy * id;
  } else {
    j0 = v.fx;
    j1 = v.fy;
  }
  // E = rows of W (the model-view's 3x3, W[k][c] = mv[c*4+k]) with the
  // perspective's third column folded in: T = diag(j0, j1) E.
  const float e00 = fmaf(xd, v.mv[2], v.mv[0]), e01 = fmaf(xd, v.mv[6], v.mv[4]), e02 = fmaf(xd, v.mv[10], v.mv[8]);
  const float e10 = fmaf(yd, v.mv[2], v.mv[1]), e11 = fmaf(yd, v.mv[6], v.mv[5]), e12 = fmaf(yd, v.mv[10], v.mv[9]);
  // R from the byte quaternion, unnormalised: integer products are exact in f32.
  const uint32_t qb = b.w;
  const float qw = (float)((int)(qb & 255u) - 128), qx = (float)((int)((qb >> 8) & 255u) - 128);
  const float qy = (float)((int)((qb >> 16) & 255u) - 128), qz = (float)((int)(qb >> 24) - 128);
  const float n2 = qw * qw + qx * qx + qy * qy + qz * qz;
  const float s = n2 > 0.f ? 2.f * __builtin_amdgcn_rcpf(n2) : 0.f;
  const float r00 = 1.f - s * (qy * qy + qz * qz), r01 = s * (qx * qy - qw * qz), r02 = s * (qx * qz + qw * qy);
  const float r10 = s * (qx * qy + qw * qz), r11 = 1.f - s * (qx * qx + qz * qz), r12 = s * (qy * qz - qw * qx);
  const float r20 = s * (qx * qz - qw * qy), r21 = s * (qy * qz + qw * qx), r22 = 1.f - s * (qx * qx + qy * qy);
  const float sx = f32_of(a.w), sy = f32_of(b.x), sz = f32_of(b.y);
  const float g00 = fmaf(e02, r20, fmaf(e01, r10, e00 * r00)) * sx;
  const float g01 = fmaf(e02, r21, fmaf(e01, r11, e00 * r01)) * sy;
  const float g02 = fmaf(e02, r22, fmaf(e01, r12, e00 * r02)) * sz;
  const float g10 = fmaf(e12, r20, fmaf(e11, r10, e10 * r00)) * sx;
  const float g11 = fmaf(e12, r21, fmaf(e11, r11, e10 * r01)) * sy;
  const float g12 = fmaf(e12, r22, fmaf(e11, r12, e10 * r02)) * sz;
  const float q0 = fmaf(g02, g02, fmaf(g01, g01, g00 * g00));
  const float q1 = fmaf(g12, g12, fmaf(g11, g11, g10 * g10));
  const float p01 = fmaf(g02, g12, fmaf(g01, g11, g00 * g10));
  o.xx = (j0 * j0) * q0;
  o.xy = (j0 * j1) * p01;
  o.yy = (j1 * j1) * q1;

  if (v.cull > 0.f) {
    // The f32 value and its bound against the FP64 one. With n_r the row
    // norms of E (|W_r| + l_r |W_2|, l_r >= |x/d|), S2 = |s|^2 and u = 2^-24,
    // |xx - xx64| <= 52 u (j0 n0)^2 S2, likewise yy and xy; 64 is used.
    const float n0 = fmaf(l0, bw2, bn0w0), n1 = fmaf(l1, bw2, bn1w1);
    const float a0 = j0 * n0, a1 = j1 * n1;
    const float c = (64.f * kU) * fmaf(sz, sz, fmaf(sy, sy, sx * sx));
    const float exx = c * a0 * a0, eyy = c * a1 * a1, exy = c * a0 * a1;
    const float X = o.xx + 0.3f, Y = o.yy + 0.3f;
    const float det = fmaf(X, Y, -(o.xy * o.xy));
    const float edet = 1.01f * (Y * exx + X * eyy + exx * eyy + 2.f * fabsf(o.xy) * exy + exy * exy +
                                4.f * kU * (X * Y + o.xy * o.xy));
    const float root = __builtin_amdgcn_sqrtf(fmaxf(det, 0.f));
    const float alpha = (float)abyte * (1.f / 255.f);
    const float val = alpha * (6.2831853f * root);
    const float esq = det > 2.f * edet ? edet * __builtin_amdgcn_rsqf(det - edet) : __builtin_amdgcn_sqrtf(3.f * edet);
    const float eval = 1.01f * (alpha * 6.2831853f * (1.01f * esq + 4.f * kU * root) + 8.f * kU * val) + 1e-30f;
    if (!(val - v.cull > eval)) {
      if (v.cull - val > eval) return o;
      o.recomputed = true;
      if (canonical_value(v, a, b, vx, vy, vz) < v.dcull) return o;
    }
  }
  // Far first: the farthest key is the smallest (src/render/splats.js:179-181).
  double t = (-vz - v.ddmin) * v.ddscale;
  t = t < 0.0 ? 0.0 : t;
  t = t > 65535.0 ? 65535.0 : t;
  o.key = 65535u - (uint32_t)t;
  o.keep = true;
  return o;
}

// ---------------------------------------------------------------------------
// K1: project + key. A grid of at most a few workgroups per CU strides over
// 256-splat tiles, so the histogram flush is paid once per workgroup.
template <bool STD>
__global__ __launch_bounds__(256) void k_project(const uint4 *__restrict__ rec, uint32_t n, ViewK v,
                                                 uint16_t *__restrict__ keys, uint64_t *__restrict__ mask,
                                                 float4 *__restrict__ cov, Ctl *__restrict__ ctl) {
  __shared__ uint32_t h[512];
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_recomp;
  __shared__ uint32_t s_last;
  const uint32_t tid = threadIdx.x;
  for (uint32_t k = tid; k < 512; k += 256) h[k] = 0;
  if (tid == 0) s_recomp = 0;
  __syncthreads();
  // Per-view parts of the bound's row norms.
  const float bw0 = fabsf(v.mv[0]) + fabsf(v.mv[4]) + fabsf(v.mv[8]);
  const float bw1 = fabsf(v.mv[1]) + fabsf(v.mv[5]) + fabsf(v.mv[9]);
  const float bw2 = fabsf(v.mv[2]) + fabsf(v.mv[6]) + fabsf(v.mv[10]);
  uint32_t recomp = 0;
  const uint32_t ntiles = (n + 255u) >> 8;
  // Records of the next tile are loaded before this one is worked on, so a
  // wave always has a tile's loads in flight behind its arithmetic.
  uint4 na = make_uint4(0, 0, 0, 0), nb = make_uint4(0, 0, 0, 0);
  {
    const uint32_t i0 = (blockIdx.x << 8) + tid;
    if (i0 < n) {
      na = rec[2 * (size_t)i0];
      nb = rec[2 * (size_t)i0 + 1];
    }
  }
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t i = (tile << 8) + tid;
    const bool in = i < n;
    const uint4 a = na, b = nb;
    const uint32_t inext = ((tile + gridDim.x) << 8) + tid;
    if (tile + gridDim.x < ntiles && inext < n) {
      na = rec[2 * (size_t)inext];
      nb = rec[2 * (size_t)inext + 1];
    }
    Projected p;
    p.keep = false;
    p.recomputed = false;
    p.key = 0;
    if (in) p = project_one<STD>(v, a, b, bw0, bw1, bw2);
    const uint64_t bal = __ballot(p.keep);
    if ((tid & 63u) == 0 && in) mask[i >> 6] = bal;
    if (in) keys[i] = (uint16_t)p.key;
    recomp += p.recomputed ? 1u : 0u;
    if (p.keep) {
      cov[i] = make_float4(p.xx, p.xy, p.yy, __uint_as_float(b.z));
      atomicAdd(&h[p.key & 255u], 1u);
      atomicAdd(&h[256u + (p.key >> 8)], 1u);
    }
  }
  if (recomp) atomicAdd(&s_recomp, recomp);
  __syncthreads();
  for (uint32_t k = tid; k < 512; k += 256)
    if (h[k]) atomicAdd(&ctl->hist[k], h[k]);
  if (tid == 0 && s_recomp) atomicAdd(&ctl->recomputed_acc, s_recomp);
  __threadfence();
  __syncthreads();
  if (tid == 0) s_last = (atomicAdd(&ctl->done, 1u) == gridDim.x - 1) ? 1u : 0u;
  __syncthreads();
  if (!s_last) return;
  // The last workgroup: every other one's counts are in. Scan them into the
  // digit bases and zero them for the next sort.
  __threadfence();
  const uint32_t c0 = __hip_atomic_load(&ctl->hist[tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t c1 = __hip_atomic_load(&ctl->hist[256 + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  ctl->hist[tid] = 0;
  ctl->hist[256 + tid] = 0;
  uint32_t total0, total1;
  const uint32_t e0 = block_scan_excl(c0, s_w, &total0);
  const uint32_t e1 = block_scan_excl(c1, s_w, &total1);
  ctl->base[tid] = e0;
  ctl->base[256 + tid] = e1;
  if (tid == 0) {
    ctl->count = total0;
    ctl->done = 0;
    ctl->ticket_lo = 0;
    ctl->ticket_hi = 0;
    ctl->recomputed = __hip_atomic_load(&ctl->recomputed_acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ctl->recomputed_acc = 0;
  }
}

// ---------------------------------------------------------------------------
// K2/K3: one stable counting pass over an 8-bit digit, a 4096-element tile per
// workgroup. HI = false: input is every splat in index order (keys + survivor
// bits), digit = key & 255, output high byte << 24 | index. HI = true: input
// is that output (count from the control block), digit = value >> 24, output
// the index.
template <bool HI>
__global__ __launch_bounds__(256) void k_scatter(const uint16_t *__restrict__ keys, const uint64_t *__restrict__ mask,
                                                 const uint32_t *__restrict__ in_packed, uint32_t n,
                                                 Ctl *__restrict__ ctl, uint64_t *__restrict__ status,
                                                 uint32_t epoch, uint32_t *__restrict__ out) {
  constexpr int KPT = kKeysPerThread;
  constexpr int TILE = kTile;
  __shared__ uint32_t s_val[TILE];
  __shared__ uint8_t s_dig[TILE];
  __shared__ uint32_t s_wh[4][256];  // per-wave running digit counts, then wave offsets
  __shared__ uint32_t s_lstart[256];
  __shared__ uint32_t s_gstart[256];
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_nvalid;

  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  if (tid == 0) s_tile = atomicAdd(HI ? &ctl->ticket_hi : &ctl->ticket_lo, 1u);
  for (uint32_t k = tid; k < 1024; k += 256) (&s_wh[0][0])[k] = 0;
  __syncthreads();
  const uint32_t t = s_tile;
  const uint32_t limit = HI ? ctl->count : n;
  const uint32_t tile0 = t * (uint32_t)TILE;
  if (tile0 >= limit) return;  // past the survivors (HI); never looked back to
  const uint32_t wbase = tile0 + w * (64u * KPT);
  const uint32_t nwords = (n + 63u) >> 6;

  // Every load of the wave's 16 rounds is issued before any is used.
  uint32_t raw[KPT];
  uint64_t vmask[KPT];
  const uint32_t wave0 = tile0 + (uint32_t)__builtin_amdgcn_readfirstlane(w) * (64u * KPT);
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + lane;
    if (HI) {
      raw[r] = i < limit ? in_packed[i] : 0u;
    } else {
      raw[r] = i < n ? (uint32_t)keys[i] : 0u;
      const uint32_t wi = (wave0 + (uint32_t)r * 64u) >> 6;
      vmask[r] = wi < nwords ? mask[wi] : 0ull;
    }
  }
  uint32_t dig[KPT], rank[KPT], val[KPT];
  uint32_t okbits = 0;
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    const uint32_t i = wbase + (uint32_t)r * 64u + lane;
    uint64_t valid;
    uint32_t d, value;
    if (HI) {
      valid = __ballot(i < limit);
      d = raw[r] >> 24;
      value = raw[r] & 0xFFFFFFu;
    } else {
      valid = vmask[r];
      d = raw[r] & 255u;
      value = ((raw[r] >> 8) << 24) | i;
    }
    const bool ok = (valid >> lane) & 1ull;
    const uint64_t peers = match_digit(d, valid);
    const uint32_t below = count_below(peers);
    uint32_t before = 0;
    if (ok) before = s_wh[w][d];
    if (ok && below == 0) s_wh[w][d] = before + (uint32_t)__popcll(peers);
    dig[r] = d;
    rank[r] = before + below;
    val[r] = value;
    okbits |= (ok ? 1u : 0u) << r;
  }
  __syncthreads();
  // Thread d owns digit d: wave offsets, the tile's count, the look-back.
  const uint32_t d = tid;
  const uint32_t c0 = s_wh[0][d], c1 = s_wh[1][d], c2 = s_wh[2][d], c3 = s_wh[3][d];
  const uint32_t tot = c0 + c1 + c2 + c3;
  uint64_t *st = status + (size_t)t * 256u + d;
  uint32_t prefix = 0;
  if (t == 0) {
    store_status(st, epoch, kFlagInclusive, tot);
  } else {
    store_status(st, epoch, kFlagAggregate, tot);
    int p = (int)t - 1;
    while (p >= 0) {
      const uint64_t s = load_status(status + (size_t)p * 256u + d);
      const uint32_t tag = (uint32_t)(s >> 32);
      if ((tag >> 2) != epoch || (tag & 3u) == 0u) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      prefix += (uint32_t)s;
      if ((tag & 3u) == kFlagInclusive) break;
      --p;
    }
    store_status(st, epoch, kFlagInclusive, prefix + tot);
  }
  s_wh[0][d] = 0;
  s_wh[1][d] = c0;
  s_wh[2][d] = c0 + c1;
  s_wh[3][d] = c0 + c1 + c2;
  uint32_t nvalid;
  const uint32_t lstart = block_scan_excl(tot, s_w, &nvalid);
  s_lstart[d] = lstart;
  s_gstart[d] = ctl->base[(HI ? 256u : 0u) + d] + prefix;
  if (tid == 0) s_nvalid = nvalid;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < KPT; ++r) {
    if ((okbits >> r) & 1u) {
      const uint32_t pos = s_lstart[dig[r]] + s_wh[w][dig[r]] + rank[r];
      if (pos < (uint32_t)TILE) {  // always, when the counts agree; a guard, not a branch taken
        s_val[pos] = val[r];
        s_dig[pos] = (uint8_t)dig[r];
      }
    }
  }
  __syncthreads();
  const uint32_t nv = min(s_nvalid, (uint32_t)TILE);
  for (uint32_t p = tid; p < nv; p += 256) {
    const uint32_t dd = s_dig[p];
    const uint32_t g = s_gstart[dd] + (p - s_lstart[dd]);
    if (g < n) out[g] = s_val[p];  // n bounds both outputs; a guard against a miscount
  }
}

// ---------------------------------------------------------------------------
// Cloud check at load: first non-finite position and the centres' box.
__device__ __forceinline__ uint32_t ordered_bits(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ __launch_bounds__(256) void k_bounds(const uint4 *__restrict__ rec, uint32_t n, uint32_t *out8) {
  // out8: [0..2] ordered min, [3..5] ordered max, [6] first bad index
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0, 0, 0}, bad = 0xFFFFFFFFu;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint4 a = rec[2 * (size_t)i];
    const float p[3] = {f32_of(a.x), f32_of(a.y), f32_of(a.z)};
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) finite = finite && isfinite(p[k]);
    if (!finite) {
      bad = min(bad, i);
      continue;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const uint32_t o = ordered_bits(p[k]);
      lo[k] = min(lo[k], o);
      hi[k] = max(hi[k], o);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    atomicMin(&out8[k], lo[k]);
    atomicMax(&out8[3 + k], hi[k]);
  }
  if (bad != 0xFFFFFFFFu) atomicMin(&out8[6], bad);
}

// ---------------------------------------------------------------------------
// Debug: the FP64 definition for every splat, for the exactness tests.
// out5 per splat: frustum+alpha pass (0/1), alpha*area, key (as computed for
// any splat passing frustum+alpha), and spare.
template <bool STD>
__global__ __launch_bounds__(256) void k_canonical(const uint4 *__restrict__ rec, uint32_t n, ViewK v, double *out5) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint4 a = rec[2 * (size_t)i], b = rec[2 * (size_t)i + 1];
  const double x = f32_of(a.x), y = f32_of(a.y), z = f32_of(a.z);
  const double vx = fma(v.dmv[8], z, fma(v.dmv[4], y, v.dmv[0] * x)) + v.dmv[12];
  const double vy = fma(v.dmv[9], z, fma(v.dmv[5], y, v.dmv[1] * x)) + v.dmv[13];
  const double vz = fma(v.dmv[10], z, fma(v.dmv[6], y, v.dmv[2] * x)) + v.dmv[14];
  double *o = out5 + 5 * (size_t)i;
  o[0] = 0;
  o[1] = __longlong_as_double(0x7FF8000000000000ll);
  o[2] = 0;
  o[3] = 0;
  o[4] = 0;
  const double cw = -vz;
  const double cx = v.dP[0] * vx + v.dP[8] * vz;
  const double cy = v.dP[5] * vy + v.dP[9] * vz;
  const double cz = v.dP[14];
  if (vz >= 0 || cz > cw || fabs(cx) > kC13 * cw || fabs(cy) > kC13 * cw || (b.z >> 24) <= 1u) return;
  o[0] = 1;
  o[1] = canonical_value(v, a, b, vx, vy, vz);
  double t = (-vz - v.ddmin) * v.ddscale;
  t = t < 0.0 ? 0.0 : t;
  t = t > 65535.0 ? 65535.0 : t;
  o[2] = (double)(65535u - (uint32_t)t);
}

// Self-test of the primitives the exact path leans on: DPP wave scan, f64
// sqrt and division rounding (against values the host computed).
__global__ __launch_bounds__(256) void k_selftest(const double *args, const double *expect, uint32_t count,
                                                  uint32_t *mismatch3) {
  // wave scan of lane+1 over each wave: inclusive sum (l+1)(l+2)/2
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t s = wave_scan_incl(lane + 1u);
  if (s != (lane + 1u) * (lane + 2u) / 2u) atomicAdd(&mismatch3[0], 1u);
  uint32_t tot;
  __shared__ uint32_t s_w[4];
  const uint32_t e = block_scan_excl(threadIdx.x + 1u, s_w, &tot);
  if (e != threadIdx.x * (threadIdx.x + 1u) / 2u || tot != 256u * 257u / 2u) atomicAdd(&mismatch3[0], 1u);
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < count; i += gridDim.x * 256u) {
    const double x = args[2 * i], y = args[2 * i + 1];
    if (__double_as_longlong(sqrt(x)) != __double_as_longlong(expect[2 * i])) atomicAdd(&mismatch3[1], 1u);
    if (__double_as_longlong(x / y) != __double_as_longlong(expect[2 * i + 1])) atomicAdd(&mismatch3[2], 1u);
  }
}

bool is_std(const wd_view &v) {
  const float *P = v.projection, *m = v.model_view;
  return P[1] == 0 && P[2] == 0 && P[3] == 0 && P[4] == 0 && P[6] == 0 && P[7] == 0 && P[10] == 0 &&
         P[11] == -1.f && P[12] == 0 && P[13] == 0 && P[15] == 0 && m[3] == 0 && m[7] == 0 && m[11] == 0 &&
         m[15] == 1.f;
}

}  // namespace

ViewK make_viewk(const wd_view &v) {
  ViewK k;
  for (int i = 0; i < 16; ++i) {
    k.mv[i] = v.model_view[i];
    k.P[i] = v.projection[i];
  }
  k.vpw = v.viewport[0];
  k.vph = v.viewport[1];
  k.dmin = v.depth_min;
  k.dscale = v.depth_scale;
  k.cull = v.cull;
  k.fx = (float)((double)v.projection[0] * (double)v.viewport[0] * 0.5);
  k.fy = (float)((double)v.projection[5] * (double)v.viewport[1] * 0.5);
  k.limx = (float)((double)1.3f / (double)v.projection[0]);
  k.limy = (float)((double)1.3f / (double)v.projection[5]);
  k.n0 = 0;
  k.n1 = 0;
  k.persp = v.projection[15] == 0.0f ? 1 : 0;
  for (int i = 0; i < 16; ++i) {
    k.dmv[i] = (double)v.model_view[i];
    k.dP[i] = (double)v.projection[i];
  }
  k.ddmin = (double)v.depth_min;
  k.ddscale = (double)v.depth_scale;
  k.dcull = (double)v.cull;
  return k;
}

const char *const kKernelNames[kKernelCount] = {"k_bounds", "k_project", "k_scatter_lo", "k_scatter_hi",
                                                "k_batch",  "k_mlp_fwd", "k_mlp_bwd",    "k_learned"};

int launch_bounds(wd_context *ctx, const uint8_t *records, uint64_t n, uint32_t *scratch8) {
  static const uint32_t init[8] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 0, 0, 0xFFFFFFFFu, 0};
  WD_HIP(hipMemcpyAsync(scratch8, init, sizeof init, hipMemcpyHostToDevice, ctx->stream), "cloud check: upload");
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, (uint64_t)ctx->cus * 4);
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKBounds, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, ctx->stream, (const uint4 *)records, (uint32_t)n, scratch8);
  WD_HIP(hipGetLastError(), "cloud check: launch");
  return ctx->prof_end(kKBounds, ev);
}

int launch_sort(wd_context *ctx, const wd_cloud *cloud, const wd_view &view, wd_state *st) {
  const uint32_t n = (uint32_t)cloud->n;
  const ViewK vk = make_viewk(view);
  const bool std_form = is_std(view);
  if (++st->epoch >= (1u << 29)) {
    WD_HIP(hipMemsetAsync(st->status_lo, 0, st->tiles * 256 * sizeof(uint64_t), ctx->stream), "sort: status reset");
    WD_HIP(hipMemsetAsync(st->status_hi, 0, st->tiles * 256 * sizeof(uint64_t), ctx->stream), "sort: status reset");
    st->epoch = 1;
  }
  st->last_n = n;
  const uint32_t tiles256 = (n + 255u) / 256u;
  const uint32_t g1 = std::min<uint32_t>(tiles256, (uint32_t)ctx->cus * 4u);
  const uint32_t tiles = (n + kTile - 1) / kTile;
  const uint4 *rec = (const uint4 *)cloud->records;
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKProject, &ev)) return WD_E_HIP;
  if (std_form)
    hipLaunchKernelGGL(k_project<true>, dim3(g1), dim3(256), 0, ctx->stream, rec, n, vk, st->keys, st->mask,
                       (float4 *)st->cov, st->ctl);
  else
    hipLaunchKernelGGL(k_project<false>, dim3(g1), dim3(256), 0, ctx->stream, rec, n, vk, st->keys, st->mask,
                       (float4 *)st->cov, st->ctl);
  WD_HIP(hipGetLastError(), "sort: project launch");
  if (ctx->prof_end(kKProject, ev)) return WD_E_HIP;
  if (ctx->prof_begin(kKScatterLo, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(k_scatter<false>, dim3(tiles), dim3(256), 0, ctx->stream, st->keys, st->mask,
                     (const uint32_t *)nullptr, n, st->ctl, st->status_lo, st->epoch, st->packed);
  WD_HIP(hipGetLastError(), "sort: low-byte scatter launch");
  if (ctx->prof_end(kKScatterLo, ev)) return WD_E_HIP;
  if (ctx->prof_begin(kKScatterHi, &ev)) return WD_E_HIP;
  hipLaunchKernelGGL(k_scatter<true>, dim3(tiles), dim3(256), 0, ctx->stream, (const uint16_t *)nullptr,
                     (const uint64_t *)nullptr, st->packed, n, st->ctl, st->status_hi, st->epoch, st->order);
  WD_HIP(hipGetLastError(), "sort: high-byte scatter launch");
  return ctx->prof_end(kKScatterHi, ev);
}

int launch_canonical_debug(wd_context *ctx, const wd_cloud *cloud, const wd_view &view, double *out5) {
  if (!is_std(view)) return fail(WD_E_ARG, "canonical debug: needs the reverse-Z perspective and an affine model-view");
  const uint32_t n = (uint32_t)cloud->n;
  hipLaunchKernelGGL(k_canonical<true>, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                     (const uint4 *)cloud->records, n, make_viewk(view), out5);
  WD_HIP(hipGetLastError(), "canonical debug: launch");
  return WD_OK;
}

}  // namespace wd

// Launch used by the self-test entry point (args/expect are device arrays of 2*count doubles).
namespace wd {
int launch_selftest_kernel(wd_context *ctx, const double *args, const double *expect, uint32_t count,
                           uint32_t *mismatch3) {
  hipLaunchKernelGGL(k_selftest, dim3(64), dim3(256), 0, ctx->stream, args, expect, count, mismatch3);
  WD_HIP(hipGetLastError(), "selftest: launch");
  return WD_OK;
}
}  // namespace wd

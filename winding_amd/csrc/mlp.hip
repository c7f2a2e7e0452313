// Kernels of the learned depth-CDF predictor: 12 pose floats -> 128 -> 128 -> 256.
//
//   k_gemm          C = alpha * A B (+ bias, relu, relu' mask, += C) on the bf16
//                   matrix cores: one wave a 16x16 tile, v_mfma_f32_16x16x16_bf16
//                   over K in steps of 16, fp32 accumulate. Operands are read
//                   through strides, so one kernel serves the forward GEMMs and
//                   the transposed backward ones; f32 operands are rounded to
//                   bf16 (nearest even) as they are loaded.
//   k_softmax_xent  the head: softmax over 256 logits, cross-entropy against the
//                   target distribution, and dZ = (p - t) / batch.
//   k_colsum        bias gradients: column sums.
//   k_adam          Adam on the fp32 master parameters, and their bf16 copy.
//   k_mlp_infer     inference, one pose a workgroup in one launch (the
//                   learned sort's, and wd_mlp_predict's): the same bf16
//                   weights and bf16 activations as the matrix-core forward,
//                   products exact in fp32, summed in fp32 in another order.
#include <hip/hip_runtime.h>

#include "wd_internal.h"

namespace wd {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t load_bf16(const void *p, int is_bf16, long idx) {
  return is_bf16 ? ((const uint16_t *)p)[idx] : to_bf16(((const float *)p)[idx]);
}

__global__ __launch_bounds__(64) void k_gemm(MlpGemm g) {
  const int tn = blockIdx.x * 16, tm = blockIdx.y * 16;
  const int lane = threadIdx.x;
  const int r = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = k0 + kq + i, m = tm + r, n = tn + r;
      a[i] = (short)((m < g.M && k < g.K) ? load_bf16(g.A, g.a_bf16, (long)m * g.sa_m + (long)k * g.sa_k) : 0);
      b[i] = (short)((k < g.K && n < g.N) ? load_bf16(g.B, g.b_bf16, (long)k * g.sb_k + (long)n * g.sb_n) : 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
  // D rows kq..kq+3, column r
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = tm + kq + i, n = tn + r;
    if (m >= g.M || n >= g.N) continue;
    const long o = (long)m * g.sc_m + (long)n * g.sc_n;
    float v = g.alpha * acc[i];
    if (g.bias) v += g.bias[n];
    if (g.relu) v = fmaxf(v, 0.f);
    if (g.mask && !(g.mask[o] > 0.f)) v = 0.f;
    if (g.accumulate) v += g.C[o];
    g.C[o] = v;
    if (g.C_bf16) g.C_bf16[o] = to_bf16(v);
  }
}

// One wave a row of 256 logits.
__global__ __launch_bounds__(256) void k_softmax_xent(const float *__restrict__ z, const float *__restrict__ target,
                                                      int rows, float scale, float *__restrict__ p,
                                                      float *__restrict__ dz, float *__restrict__ loss) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const float *zr = z + (long)row * 256;
  float v[4], mx = -3.4e38f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = zr[lane + 64 * i];
    mx = fmaxf(mx, v[i]);
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = expf(v[i] - mx);
    s += v[i];
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  const float inv = 1.f / s;
  float l = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const long idx = (long)row * 256 + lane + 64 * i;
    const float pi = v[i] * inv;
    if (p) p[idx] = pi;
    if (target) {
      const float t = target[idx];
      if (t > 0.f) l -= t * logf(fmaxf(pi, 1e-30f));
      if (dz) dz[idx] = (pi - t) * scale;
    }
  }
  if (loss && target) {
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (lane == 0) atomicAdd(loss, l * scale);
  }
}

// The whole forward for one pose: layers 1 and 2 by threads 0..127 (one
// output each), layer 3 by all 256, the softmax across the workgroup.
// Parameters: W1 [128][12] | b1 | W2 [128][128] | b2 | W3 [256][128] | b3,
// bf16 weights (wb) and fp32 biases (w).
constexpr int kMIn = WD_POSE_FLOATS, kMH = 128;
constexpr int kOW1 = 0, kOB1 = kOW1 + kMH * kMIn, kOW2 = kOB1 + kMH, kOB2 = kOW2 + kMH * kMH, kOW3 = kOB2 + kMH,
              kOB3 = kOW3 + WD_COARSE_BINS * kMH;
static_assert(WD_COARSE_BINS == 256, "one output a thread");
static_assert(kOW2 % 8 == 0 && kOW3 % 8 == 0, "16-byte rows");

__device__ __forceinline__ float dot_row_bf16(const uint16_t *__restrict__ row, const float *h) {
  // 128 bf16 weights, 16 bytes at a time, against 128 activations in LDS
  const uint4 *q = reinterpret_cast<const uint4 *>(row);
  float acc = 0.f;
#pragma unroll 4
  for (int j = 0; j < kMH / 8; ++j) {
    const uint4 v = q[j];
    const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      acc += __uint_as_float(u[e] << 16) * h[8 * j + 2 * e];
      acc += __uint_as_float(u[e] & 0xFFFF0000u) * h[8 * j + 2 * e + 1];
    }
  }
  return acc;
}

__global__ __launch_bounds__(256) void k_mlp_infer(const float *__restrict__ X, const uint16_t *__restrict__ wb,
                                                   const float *__restrict__ w, float *__restrict__ P) {
  __shared__ float h1[kMH], h2[kMH], red[4];
  const int row = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  if (t < kMH) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < kMIn; ++k) acc += from_bf16(wb[kOW1 + t * kMIn + k]) * from_bf16(to_bf16(X[row * kMIn + k]));
    h1[t] = from_bf16(to_bf16(fmaxf(acc + w[kOB1 + t], 0.f)));
  }
  __syncthreads();
  if (t < kMH) h2[t] = from_bf16(to_bf16(fmaxf(dot_row_bf16(wb + kOW2 + t * kMH, h1) + w[kOB2 + t], 0.f)));
  __syncthreads();
  const float z = dot_row_bf16(wb + kOW3 + t * kMH, h2) + w[kOB3 + t];
  float mx = z;
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  if (lane == 0) red[wv] = mx;
  __syncthreads();
  mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  __syncthreads();
  const float e = expf(z - mx);
  float s = e;
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) red[wv] = s;
  __syncthreads();
  s = red[0] + red[1] + red[2] + red[3];
  P[row * WD_COARSE_BINS + t] = e / s;
}

__global__ __launch_bounds__(256) void k_colsum(const float *__restrict__ x, int rows, int cols, float *__restrict__ out) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; ++r) s += x[(long)r * cols + c];
  out[c] = s;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, uint16_t *__restrict__ wb, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int n, float lr, float b1,
                                              float b2, float c1, float c2, float eps) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wb[i] = to_bf16(wi);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ w, uint16_t *__restrict__ wb, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) wb[i] = to_bf16(w[i]);
}

}  // namespace

int launch_gemm(wd_context *ctx, const MlpGemm &g) {
  hipLaunchKernelGGL(k_gemm, dim3((g.N + 15) / 16, (g.M + 15) / 16), dim3(64), 0, ctx->stream, g);
  WD_HIP(hipGetLastError(), "mlp: gemm launch");
  return WD_OK;
}

int launch_mlp_infer(wd_context *ctx, const float *X, const uint16_t *wb, const float *w, int rows, float *P) {
  hipLaunchKernelGGL(k_mlp_infer, dim3(rows), dim3(256), 0, ctx->stream, X, wb, w, P);
  WD_HIP(hipGetLastError(), "mlp: inference launch");
  return WD_OK;
}

int launch_softmax_xent(wd_context *ctx, const float *z, const float *target, int rows, float scale, float *p, float *dz,
                        float *loss) {
  hipLaunchKernelGGL(k_softmax_xent, dim3((rows + 3) / 4), dim3(256), 0, ctx->stream, z, target, rows, scale, p, dz, loss);
  WD_HIP(hipGetLastError(), "mlp: softmax launch");
  return WD_OK;
}

int launch_colsum(wd_context *ctx, const float *x, int rows, int cols, float *out) {
  hipLaunchKernelGGL(k_colsum, dim3((cols + 255) / 256), dim3(256), 0, ctx->stream, x, rows, cols, out);
  WD_HIP(hipGetLastError(), "mlp: colsum launch");
  return WD_OK;
}

int launch_adam(wd_context *ctx, float *w, uint16_t *wb, const float *g, float *m, float *v, int n, float lr, float b1,
                float b2, int step, float eps) {
  const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, w, wb, g, m, v, n, lr, b1, b2, c1, c2, eps);
  WD_HIP(hipGetLastError(), "mlp: adam launch");
  return WD_OK;
}

int launch_to_bf16(wd_context *ctx, const float *w, uint16_t *wb, int n) {
  hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, w, wb, n);
  WD_HIP(hipGetLastError(), "mlp: cast launch");
  return WD_OK;
}

}  // namespace wd

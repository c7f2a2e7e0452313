// Host side of the learned depth-CDF predictor (wd_mlp_*): parameters, the
// forward and backward passes as sequences of the matrix-core GEMM, and Adam.
//
// Network: x (12 pose floats) -> relu(W1 x + b1) (128) -> relu(W2 h1 + b2)
// (128) -> z = W3 h2 + b3 (256) -> p = softmax(z). p is the predicted mass of
// 256 equal-probability depth bins read as widths: bin k covers keys
// [65536 * sum_{j<k} p_j, 65536 * sum_{j<=k} p_j), so p is the derivative of
// the inverse CDF of the survivors' keys, sampled at 256 points. Trained by
// cross-entropy against the exact widths the sort itself measures.
//
// Parameters, flat, fp32 master (bf16 copy for the matrix cores):
//   W1 [128][12] | b1 [128] | W2 [128][128] | b2 [128] | W3 [256][128] | b3 [256]
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "wd_internal.h"

using wd::fail;

namespace {

constexpr int kIn = WD_POSE_FLOATS, kH = 128, kOut = WD_COARSE_BINS;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kOut * kH;
constexpr int kParams = oB3 + kOut;
static_assert(kParams == 51200, "parameter count");

}  // namespace

struct wd_mlp {
  wd_context *ctx = nullptr;
  float *w = nullptr;      // fp32 master
  uint16_t *wb = nullptr;  // bf16 copy
  float *g = nullptr, *m = nullptr, *v = nullptr;
  int step = 0;
  // activations for up to `cap` rows
  uint32_t cap = 0;
  float *X = nullptr, *H1 = nullptr, *H2 = nullptr, *Z = nullptr, *P = nullptr, *T = nullptr, *dZ = nullptr;
  float *dH1 = nullptr, *dH2 = nullptr, *loss = nullptr;
  uint16_t *H1b = nullptr, *H2b = nullptr;
  uint32_t magic = 0x4d4c5031;
};

namespace {

int ensure_rows(wd_mlp *m, uint32_t rows) {
  if (rows <= m->cap) return WD_OK;
  void *ptrs[] = {m->X, m->H1, m->H2, m->Z, m->P, m->T, m->dZ, m->dH1, m->dH2, m->H1b, m->H2b};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  const size_t r = rows;
  hipError_t e = hipMalloc(&m->X, r * kIn * 4);
  if (e == hipSuccess) e = hipMalloc(&m->H1, r * kH * 4);
  if (e == hipSuccess) e = hipMalloc(&m->H2, r * kH * 4);
  if (e == hipSuccess) e = hipMalloc(&m->Z, r * kOut * 4);
  if (e == hipSuccess) e = hipMalloc(&m->P, r * kOut * 4);
  if (e == hipSuccess) e = hipMalloc(&m->T, r * kOut * 4);
  if (e == hipSuccess) e = hipMalloc(&m->dZ, r * kOut * 4);
  if (e == hipSuccess) e = hipMalloc(&m->dH1, r * kH * 4);
  if (e == hipSuccess) e = hipMalloc(&m->dH2, r * kH * 4);
  if (e == hipSuccess) e = hipMalloc(&m->H1b, r * kH * 2);
  if (e == hipSuccess) e = hipMalloc(&m->H2b, r * kH * 2);
  if (e != hipSuccess) {
    m->cap = 0;
    return wd::hip_fail(e, "mlp: activations");
  }
  m->cap = rows;
  return WD_OK;
}

wd::MlpGemm gemm(int M, int N, int K, const void *A, int abf, long sam, long sak, const void *B, int bbf, long sbk,
                 long sbn, float *C, long scm, long scn) {
  wd::MlpGemm g;
  std::memset(&g, 0, sizeof g);
  g.M = M; g.N = N; g.K = K;
  g.A = A; g.a_bf16 = abf; g.sa_m = sam; g.sa_k = sak;
  g.B = B; g.b_bf16 = bbf; g.sb_k = sbk; g.sb_n = sbn;
  g.C = C; g.sc_m = scm; g.sc_n = scn;
  g.alpha = 1.f;
  return g;
}

// Forward over `rows` rows already in m->X; logits into Z, probabilities into P,
// and with targets in T: dZ = (P - T) / rows and the mean loss added to *m->loss.
int forward(wd_context *ctx, wd_mlp *m, int rows, bool with_targets, const float *X = nullptr) {
  wd::MlpGemm g1 = gemm(rows, kH, kIn, X ? X : m->X, 0, kIn, 1, m->wb + oW1, 1, 1, kIn, m->H1, kH, 1);
  g1.bias = m->w + oB1;
  g1.relu = 1;
  g1.C_bf16 = m->H1b;
  int rc = wd::launch_gemm(ctx, g1);
  wd::MlpGemm g2 = gemm(rows, kH, kH, m->H1b, 1, kH, 1, m->wb + oW2, 1, 1, kH, m->H2, kH, 1);
  g2.bias = m->w + oB2;
  g2.relu = 1;
  g2.C_bf16 = m->H2b;
  if (!rc) rc = wd::launch_gemm(ctx, g2);
  wd::MlpGemm g3 = gemm(rows, kOut, kH, m->H2b, 1, kH, 1, m->wb + oW3, 1, 1, kH, m->Z, kOut, 1);
  g3.bias = m->w + oB3;
  if (!rc) rc = wd::launch_gemm(ctx, g3);
  if (!rc)
    rc = wd::launch_softmax_xent(ctx, m->Z, with_targets ? m->T : nullptr, rows, 1.f / rows, m->P,
                                 with_targets ? m->dZ : nullptr, with_targets ? m->loss : nullptr);
  return rc;
}

// Backward from dZ into m->g (every parameter's gradient).
int backward(wd_context *ctx, wd_mlp *m, int rows) {
  int rc = wd::launch_gemm(ctx, gemm(kOut, kH, rows, m->dZ, 0, 1, kOut, m->H2b, 1, kH, 1, m->g + oW3, kH, 1));
  if (!rc) rc = wd::launch_colsum(ctx, m->dZ, rows, kOut, m->g + oB3);
  wd::MlpGemm d2 = gemm(rows, kH, kOut, m->dZ, 0, kOut, 1, m->wb + oW3, 1, kH, 1, m->dH2, kH, 1);
  d2.mask = m->H2;
  if (!rc) rc = wd::launch_gemm(ctx, d2);
  if (!rc) rc = wd::launch_gemm(ctx, gemm(kH, kH, rows, m->dH2, 0, 1, kH, m->H1b, 1, kH, 1, m->g + oW2, kH, 1));
  if (!rc) rc = wd::launch_colsum(ctx, m->dH2, rows, kH, m->g + oB2);
  wd::MlpGemm d1 = gemm(rows, kH, kH, m->dH2, 0, kH, 1, m->wb + oW2, 1, kH, 1, m->dH1, kH, 1);
  d1.mask = m->H1;
  if (!rc) rc = wd::launch_gemm(ctx, d1);
  if (!rc) rc = wd::launch_gemm(ctx, gemm(kH, kIn, rows, m->dH1, 0, 1, kH, m->X, 0, kIn, 1, m->g + oW1, kIn, 1));
  if (!rc) rc = wd::launch_colsum(ctx, m->dH1, rows, kH, m->g + oB1);
  return rc;
}

bool alive(const wd_mlp *m) { return m && m->magic == 0x4d4c5031; }

}  // namespace

namespace wd {
int mlp_infer(wd_context *ctx, wd_mlp *m, const float *poses_dev, uint32_t rows, const float **probs) {
  if (!ctx || !alive(m) || !poses_dev || rows == 0) return fail(WD_E_ARG, "mlp: bad argument");
  if (m->ctx != ctx) return fail(WD_E_STATE, "mlp: of another device");
  int rc = ensure_rows(m, rows);
  if (rc) return rc;
  hipEvent_t ev = nullptr;
  if (ctx->prof_begin(kKMlp, &ev)) return WD_E_HIP;
  if ((rc = launch_mlp_infer(ctx, poses_dev, m->wb, m->w, (int)rows, m->P))) return rc;
  if (ctx->prof_end(kKMlp, ev)) return WD_E_HIP;
  *probs = m->P;
  return WD_OK;
}
}  // namespace wd

extern "C" {

int wd_mlp_create(wd_context *ctx, uint64_t seed, wd_mlp **out) {
  if (!ctx || !out) return fail(WD_E_ARG, "mlp: null argument");
  *out = nullptr;
  WD_HIP(hipSetDevice(ctx->device), "mlp: set device");
  wd_mlp *m = new wd_mlp();
  m->ctx = ctx;
  hipError_t e = hipMalloc(&m->w, kParams * 4);
  if (e == hipSuccess) e = hipMalloc(&m->wb, kParams * 2);
  if (e == hipSuccess) e = hipMalloc(&m->g, kParams * 4);
  if (e == hipSuccess) e = hipMalloc(&m->m, kParams * 4);
  if (e == hipSuccess) e = hipMalloc(&m->v, kParams * 4);
  if (e == hipSuccess) e = hipMalloc(&m->loss, 4);
  if (e != hipSuccess) {
    wd_mlp_destroy(m);
    return wd::hip_fail(e, "mlp: allocation");
  }
  // He initialisation from a seeded generator; biases zero.
  std::vector<float> w(kParams, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_out, int fan_in) {
    std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / fan_in));
    for (int i = 0; i < fan_out * fan_in; ++i) w[off + i] = (float)nd(rng);
  };
  fill(oW1, kH, kIn);
  fill(oW2, kH, kH);
  fill(oW3, kOut, kH);
  int rc = wd_mlp_set(m, w.data());
  if (rc) {
    wd_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return WD_OK;
}

int wd_mlp_destroy(wd_mlp *m) {
  if (!alive(m)) return fail(WD_E_STATE, "mlp: destroyed twice");
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  void *ptrs[] = {m->w, m->wb, m->g, m->m, m->v, m->loss, m->X, m->H1, m->H2, m->Z, m->P,
                  m->T, m->dZ, m->dH1, m->dH2, m->H1b, m->H2b};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  m->magic = 0;
  delete m;
  return WD_OK;
}

int wd_mlp_set(wd_mlp *m, const float *params) {
  if (!alive(m) || !params) return fail(WD_E_ARG, "mlp: null argument or destroyed");
  wd_context *ctx = m->ctx;
  WD_HIP(hipMemcpyAsync(m->w, params, kParams * 4, hipMemcpyHostToDevice, ctx->stream), "mlp: set");
  WD_HIP(hipMemsetAsync(m->m, 0, kParams * 4, ctx->stream), "mlp: set");
  WD_HIP(hipMemsetAsync(m->v, 0, kParams * 4, ctx->stream), "mlp: set");
  m->step = 0;
  int rc = wd::launch_to_bf16(ctx, m->w, m->wb, kParams);
  if (rc) return rc;
  WD_HIP(hipStreamSynchronize(ctx->stream), "mlp: set");
  return WD_OK;
}

int wd_mlp_get(wd_mlp *m, float *params) {
  if (!alive(m) || !params) return fail(WD_E_ARG, "mlp: null argument or destroyed");
  WD_HIP(hipStreamSynchronize(m->ctx->stream), "mlp: get");
  WD_HIP(hipMemcpy(params, m->w, kParams * 4, hipMemcpyDeviceToHost), "mlp: get");
  return WD_OK;
}

int wd_mlp_save(wd_mlp *m, void *bytes, uint64_t capacity, uint64_t *size) {
  // "WDMLP1", the bf16 weight copy (what the matrix cores read), then the fp32 biases.
  if (!alive(m) || !size) return fail(WD_E_ARG, "mlp: null argument or destroyed");
  const uint64_t need = 8 + (uint64_t)kParams * 2 + (uint64_t)(kH + kH + kOut) * 4;
  *size = need;
  if (!bytes) return WD_OK;
  if (capacity < need) return fail(WD_E_CAPACITY, "mlp: save needs %llu bytes", (unsigned long long)need);
  std::vector<uint16_t> wb(kParams);
  std::vector<float> w(kParams);
  WD_HIP(hipStreamSynchronize(m->ctx->stream), "mlp: save");
  WD_HIP(hipMemcpy(wb.data(), m->wb, kParams * 2, hipMemcpyDeviceToHost), "mlp: save");
  WD_HIP(hipMemcpy(w.data(), m->w, kParams * 4, hipMemcpyDeviceToHost), "mlp: save");
  uint8_t *p = (uint8_t *)bytes;
  std::memcpy(p, "WDMLP1\0\0", 8);
  std::memcpy(p + 8, wb.data(), kParams * 2);
  float *b = (float *)(p + 8 + kParams * 2);
  std::memcpy(b, w.data() + oB1, kH * 4);
  std::memcpy(b + kH, w.data() + oB2, kH * 4);
  std::memcpy(b + 2 * kH, w.data() + oB3, kOut * 4);
  return WD_OK;
}

int wd_mlp_load(wd_mlp *m, const void *bytes, uint64_t size) {
  if (!alive(m) || !bytes) return fail(WD_E_ARG, "mlp: null argument or destroyed");
  const uint64_t need = 8 + (uint64_t)kParams * 2 + (uint64_t)(kH + kH + kOut) * 4;
  const uint8_t *p = (const uint8_t *)bytes;
  if (size != need || std::memcmp(p, "WDMLP1", 6) != 0)
    return fail(WD_E_DATA, "mlp: not a saved predictor (%llu bytes; one is %llu)", (unsigned long long)size,
                (unsigned long long)need);
  std::vector<float> w(kParams);
  const uint16_t *wb = (const uint16_t *)(p + 8);
  for (int i = 0; i < kParams; ++i) {
    const uint32_t u = (uint32_t)wb[i] << 16;
    std::memcpy(&w[i], &u, 4);
  }
  const float *b = (const float *)(p + 8 + kParams * 2);
  std::memcpy(w.data() + oB1, b, kH * 4);
  std::memcpy(w.data() + oB2, b + kH, kH * 4);
  std::memcpy(w.data() + oB3, b + 2 * kH, kOut * 4);
  return wd_mlp_set(m, w.data());
}

int wd_mlp_grad(wd_context *ctx, wd_mlp *m, const float *poses, const float *targets, uint32_t n, float *grads,
                float *loss) {
  if (!ctx || !alive(m) || !poses || !targets || n == 0) return fail(WD_E_ARG, "mlp: bad argument");
  int rc = ensure_rows(m, n);
  if (rc) return rc;
  WD_HIP(hipMemcpyAsync(m->X, poses, (size_t)n * kIn * 4, hipMemcpyHostToDevice, ctx->stream), "mlp: upload");
  WD_HIP(hipMemcpyAsync(m->T, targets, (size_t)n * kOut * 4, hipMemcpyHostToDevice, ctx->stream), "mlp: upload");
  WD_HIP(hipMemsetAsync(m->loss, 0, 4, ctx->stream), "mlp: loss");
  if ((rc = forward(ctx, m, (int)n, true)) || (rc = backward(ctx, m, (int)n))) return rc;
  WD_HIP(hipStreamSynchronize(ctx->stream), "mlp: grad");
  if (grads) WD_HIP(hipMemcpy(grads, m->g, kParams * 4, hipMemcpyDeviceToHost), "mlp: grads");
  if (loss) WD_HIP(hipMemcpy(loss, m->loss, 4, hipMemcpyDeviceToHost), "mlp: loss");
  return WD_OK;
}

int wd_mlp_train(wd_context *ctx, wd_mlp *m, const float *poses, const float *targets, uint32_t n, uint32_t epochs,
                 uint32_t batch, float lr, uint64_t seed, float *loss) {
  if (!ctx || !alive(m) || !poses || !targets || n == 0 || batch == 0) return fail(WD_E_ARG, "mlp: bad argument");
  batch = std::min(batch, n);
  int rc = ensure_rows(m, batch);
  if (rc) return rc;
  std::vector<uint32_t> idx(n);
  for (uint32_t i = 0; i < n; ++i) idx[i] = i;
  std::mt19937_64 rng(seed);
  std::vector<float> xb((size_t)batch * kIn), tb((size_t)batch * kOut);
  float last = 0.f;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    std::shuffle(idx.begin(), idx.end(), rng);
    WD_HIP(hipMemsetAsync(m->loss, 0, 4, ctx->stream), "mlp: loss");
    const uint32_t steps = n / batch;
    for (uint32_t s = 0; s < steps; ++s) {
      for (uint32_t r = 0; r < batch; ++r) {
        const uint32_t k = idx[s * batch + r];
        std::memcpy(&xb[(size_t)r * kIn], poses + (size_t)k * kIn, kIn * 4);
        std::memcpy(&tb[(size_t)r * kOut], targets + (size_t)k * kOut, kOut * 4);
      }
      // Staging copies are synchronous for pageable memory, so the buffers are free again on return.
      WD_HIP(hipMemcpyAsync(m->X, xb.data(), xb.size() * 4, hipMemcpyHostToDevice, ctx->stream), "mlp: upload");
      WD_HIP(hipMemcpyAsync(m->T, tb.data(), tb.size() * 4, hipMemcpyHostToDevice, ctx->stream), "mlp: upload");
      if ((rc = forward(ctx, m, (int)batch, true)) || (rc = backward(ctx, m, (int)batch))) return rc;
      m->step++;
      if ((rc = wd::launch_adam(ctx, m->w, m->wb, m->g, m->m, m->v, kParams, lr, 0.9f, 0.999f, m->step, 1e-8f))) return rc;
      WD_HIP(hipStreamSynchronize(ctx->stream), "mlp: step");
    }
    WD_HIP(hipMemcpy(&last, m->loss, 4, hipMemcpyDeviceToHost), "mlp: loss");
    last /= (float)std::max(1u, n / batch);
  }
  if (loss) *loss = last;
  return WD_OK;
}

int wd_mlp_predict(wd_context *ctx, wd_mlp *m, const float *poses, uint32_t n, float *probs) {
  if (!ctx || !alive(m) || !poses || !probs || n == 0) return fail(WD_E_ARG, "mlp: bad argument");
  int rc = ensure_rows(m, n);
  if (rc) return rc;
  WD_HIP(hipMemcpyAsync(m->X, poses, (size_t)n * kIn * 4, hipMemcpyHostToDevice, ctx->stream), "mlp: upload");
  if ((rc = wd::launch_mlp_infer(ctx, m->X, m->wb, m->w, (int)n, m->P))) return rc;
  WD_HIP(hipStreamSynchronize(ctx->stream), "mlp: predict");
  WD_HIP(hipMemcpy(probs, m->P, (size_t)n * kOut * 4, hipMemcpyDeviceToHost), "mlp: predict");
  return WD_OK;
}

int wd_pose_features(const wd_view *view, float out[12]) {
  // The model-view's top three rows: rotation and translation, translation / 10.
  if (!view || !out) return fail(WD_E_ARG, "pose: null argument");
  const float *m = view->model_view;
  for (int r = 0; r < 3; ++r) {
    out[4 * r + 0] = m[r];
    out[4 * r + 1] = m[4 + r];
    out[4 * r + 2] = m[8 + r];
    out[4 * r + 3] = m[12 + r] * 0.1f;
  }
  return WD_OK;
}

}  // extern "C"

// Seeded synthetic .splat records for the benchmark workloads.
//
// Each splat draws from its own splitmix64 stream (seeded by the cloud seed
// and its index), so records are identical however many threads write them.
// Uniforms take the top 53 bits; normals are Box-Muller (cosine branch).
//
// kind 0, Gaussian mixture: three components, means U[-3,3]^3, isotropic
//   sigma U[1,3] (drawn from the cloud's own stream); scale exp(N(-4,0.5))
//   per axis; rgba U{0..255}; quaternion a normalised N(0,1)^4 stored as
//   clamp(round(q*128+128), 0, 255) in w,x,y,z order, the .splat layout of
//   src/scene/splats.js:373-388.
// kind 1, slab: 80% in |z| <= 0.25, x in [-10.5,10.5], y in [-6,6] (the
//   frustum's cross-section at distance 10 for fovY 60deg, 16:9), 20% uniform
//   in [-20,20]^3; other attributes as kind 0.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "winding_amd.h"
#include "wd_internal.h"

namespace {

struct Stream {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double normal() {
    double u1 = 1.0 - uniform();
    double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

uint64_t mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void put_f32(uint8_t *p, double v) {
  float f = (float)v;
  std::memcpy(p, &f, 4);
}

uint8_t quat_byte(double q) {
  double b = std::round(q * 128.0 + 128.0);
  if (b < 0) b = 0;
  if (b > 255) b = 255;
  return (uint8_t)b;
}

}  // namespace

extern "C" int wd_synth_records(int kind, uint64_t seed, uint64_t n, void *records_out) {
  if (!records_out || (kind != 0 && kind != 1)) return wd::fail(WD_E_ARG, "synth: kind must be 0 or 1 and the output not null");
  Stream cloud{mix(seed ^ 0xC2B2AE3D27D4EB4Full)};
  double mean[3][3], sigma[3];
  for (int c = 0; c < 3; c++) {
    for (int k = 0; k < 3; k++) mean[c][k] = -3.0 + 6.0 * cloud.uniform();
    sigma[c] = 1.0 + 2.0 * cloud.uniform();
  }
  uint8_t *out = (uint8_t *)records_out;
  const uint64_t base = mix(seed + 0x165667B19E3779F9ull);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    Stream s{mix(base ^ (uint64_t)i)};
    uint8_t *r = out + (uint64_t)i * 32;
    double p[3];
    if (kind == 0) {
      int c = (int)(s.uniform() * 3.0);
      for (int k = 0; k < 3; k++) p[k] = mean[c][k] + sigma[c] * s.normal();
    } else if (s.uniform() < 0.8) {
      p[0] = -10.5 + 21.0 * s.uniform();
      p[1] = -6.0 + 12.0 * s.uniform();
      p[2] = -0.25 + 0.5 * s.uniform();
    } else {
      for (int k = 0; k < 3; k++) p[k] = -20.0 + 40.0 * s.uniform();
    }
    for (int k = 0; k < 3; k++) put_f32(r + 4 * k, p[k]);
    for (int k = 0; k < 3; k++) put_f32(r + 12 + 4 * k, std::exp(-4.0 + 0.5 * s.normal()));
    for (int k = 0; k < 4; k++) r[24 + k] = (uint8_t)(s.uniform() * 256.0);
    double q[4], len2 = 0;
    for (int k = 0; k < 4; k++) {
      q[k] = s.normal();
      len2 += q[k] * q[k];
    }
    double len = std::sqrt(len2);
    if (len == 0) len = 1;
    for (int k = 0; k < 4; k++) r[28 + k] = quat_byte(q[k] / len);
  }
  return WD_OK;
}

// Host side of the capture decoders (include/winding_amd.h, wd_decode_*): the
// .ply header, the .spz header and gunzip, the .sog tables, and the checks and
// messages of the reference's parsers (src/scene/splats.js:58-303,390-468).
// The per-splat work runs on the GPU (decode.hip).
#include <hip/hip_runtime_api.h>
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "wd_internal.h"

using wd::fail;

namespace {

const int kShCoefficients[5] = {0, 3, 8, 15, 24};

// Math.round: to the nearest, a half up.
double js_round(double x) {
  const double r = std::floor(x);
  return (x - r >= 0.5) ? r + 1.0 : r;
}

// halfBits (src/render/hdr.js:152-163) and signedHalf (src/scene/splats.js:50-52).
uint16_t signed_half_host(double v) {
  const bool neg = v < 0;
  if (neg) v = -v;
  uint32_t h;
  if (!(v > 0.0)) {
    h = 0;
  } else if (v >= 65504.0) {
    h = 0x7bff;
  } else if (v < 0x1p-24) {
    h = 0;
  } else if (v < 0x1p-14) {
    h = (uint32_t)js_round(v * 0x1p24);
  } else {
    int e;
    const double f = std::frexp(v, &e);
    int exponent = e - 1;
    double mantissa = js_round((std::ldexp(f, 1) - 1.0) * 1024.0);
    if (mantissa == 1024.0) {
      mantissa = 0.0;
      ++exponent;
    }
    h = (uint32_t)((exponent + 15) << 10) | (uint32_t)mantissa;
  }
  return (uint16_t)(neg ? 0x8000u | h : h);
}

// The decoded splats' buffers and the launch that fills them; frees all on failure.
int decode_into(wd_context *ctx, uint64_t n, uint32_t degree,
                const std::function<int(const wd::DecodeOut &, uint32_t *)> &launch,
                wd_cloud **out, float *min_out, float *max_out) {
  if (n > WD_MAX_SORT)
    return fail(WD_E_CAPACITY, "splats: %llu splats are more than this device holds in one buffer, %u",
                (unsigned long long)n, WD_MAX_SORT);
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  wd_cloud *c = new wd_cloud();
  c->ctx = ctx;
  c->n = n;
  const uint32_t W = degree ? wd::sh_words(degree) : 0;
  uint32_t *scratch = nullptr;
  auto undo = [&]() {
    (void)hipStreamSynchronize(ctx->stream);
    for (void *p : {(void *)c->centers, (void *)c->covs, (void *)c->sh, (void *)scratch})
      if (p) (void)hipFree(p);
    delete c;
  };
  hipError_t e = hipMalloc(&c->centers, n * 16);
  if (e == hipSuccess) e = hipMalloc(&c->covs, n * 24);
  if (e == hipSuccess && W) e = hipMalloc(&c->sh, (n * W + 1) * 4);
  if (e == hipSuccess && W) e = hipMemsetAsync(c->sh, 0, (n * W + 1) * 4, ctx->stream);
  // bad index, ordered min xyz, ordered max xyz, first zero per axis, 2 spare; then a box column a
  // workgroup (every decoder's workgroups are 64 splats or more)
  const uint64_t most_groups = (n + 63) / 64;
  if (e == hipSuccess) e = hipMalloc(&scratch, (12 + (size_t)wd::kBoxWords * most_groups) * sizeof(uint32_t));
  if (e != hipSuccess) {
    undo();
    return wd::hip_fail(e, "splats: decode buffers");
  }
  wd::DecodeOut o{c->centers, c->covs, c->sh, W, scratch, scratch + 1, scratch + 7, scratch + 12};
  uint32_t groups = 0;
  int rc = launch(o, &groups);
  if (rc == WD_OK && (groups == 0 || groups > most_groups)) rc = fail(WD_E_STATE, "splats: a decode of %u workgroups", groups);
  if (rc == WD_OK) rc = wd::launch_box_reduce(ctx, o.partials, groups, o.box, o.box + 3, o.zero, o.bad);
  uint32_t res[12];
  if (rc == WD_OK) {
    e = hipMemcpyAsync(res, scratch, sizeof res, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = wd::hip_fail(e, "splats: decode");
  }
  if (rc == WD_OK && res[0] != 0xFFFFFFFFu)
    rc = fail(WD_E_DATA, "splats: splat %u has a position that is not a number", res[0]);
  if (rc != WD_OK) {
    undo();
    return rc;
  }
  (void)hipFree(scratch);
  scratch = nullptr;
  // a zero end of the box takes its sign from the first zero (put()'s strict compares), read back
  auto at = [](const void *centers, uint32_t i, int k) -> float {
    float v = 0.f;
    if (hipMemcpy(&v, (const float *)centers + (size_t)i * 4 + k, 4, hipMemcpyDeviceToHost) != hipSuccess) return 0.f;
    return v;
  };
  wd::box_from_ordered(res + 1, res + 4, res + 7, at, c->centers, c->min, c->max);
  c->degree = degree;
  if (min_out) std::memcpy(min_out, c->min, sizeof c->min);
  if (max_out) std::memcpy(max_out, c->max, sizeof c->max);
  ctx->adopt(c);
  *out = c;
  return WD_OK;
}

// A host buffer copied to the device for the length of one decode.
// 16 bytes over the end, zeroed, so that 16-byte loads of the last whole words stay inside it.
struct DeviceCopy {
  void *p = nullptr;
  ~DeviceCopy() {
    if (p) (void)hipFree(p);
  }
  hipError_t from(const void *src, size_t bytes) {
    hipError_t e = hipMalloc(&p, bytes + 16);
    if (e == hipSuccess) e = hipMemset((uint8_t *)p + bytes, 0, 16);
    return e == hipSuccess && bytes ? hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) : e;
  }
};

std::vector<std::string> words_of(const std::string &line) {
  std::vector<std::string> w;
  std::istringstream in(line);
  std::string t;
  while (in >> t) w.push_back(t);
  return w;
}

// gunzip, as the browser's DecompressionStream('gzip') does for readSplats.
int gunzip(const uint8_t *src, uint64_t len, std::vector<uint8_t> &out) {
  z_stream z;
  std::memset(&z, 0, sizeof z);
  if (inflateInit2(&z, 16 + MAX_WBITS) != Z_OK) return fail(WD_E_DATA, "splats: gunzip could not start");
  z.next_in = const_cast<Bytef *>(src);
  z.avail_in = (uInt)len;
  out.clear();
  uint8_t buf[1 << 16];
  int r;
  do {
    z.next_out = buf;
    z.avail_out = sizeof buf;
    r = inflate(&z, Z_NO_FLUSH);
    if (r != Z_OK && r != Z_STREAM_END) {
      inflateEnd(&z);
      return fail(WD_E_DATA, "splats: the gzip stream is damaged");
    }
    out.insert(out.end(), buf, buf + (sizeof buf - z.avail_out));
  } while (r != Z_STREAM_END);
  inflateEnd(&z);
  return WD_OK;
}

}  // namespace

extern "C" {

int wd_decode_ply(wd_context *ctx, const void *bytes, uint64_t len, wd_cloud **out, float min_out[3],
                  float max_out[3]) {
  if (!ctx || !bytes || !out) return fail(WD_E_ARG, "splats: null argument");
  *out = nullptr;
  const uint8_t *data = (const uint8_t *)bytes;
  // The header ends at the first "end_header" line; the vertex bytes start after its line break
  // (src/scene/splats.js:400-410).
  const char *marker = "end_header";
  const int64_t scan = (int64_t)std::min<uint64_t>(len, 65536) - 10;
  int64_t end = -1;
  for (int64_t i = 0; i < scan; ++i)
    if (std::memcmp(data + i, marker, 10) == 0) {
      end = i + 10;
      break;
    }
  if (end < 0) return fail(WD_E_DATA, "splats: a .ply with no end_header");
  if ((uint64_t)end < len && data[end] == 0x0d) ++end;
  ++end;  // the newline
  std::string header((const char *)data, (size_t)std::min<uint64_t>((uint64_t)end, len));
  std::string format;
  bool have_vertices = false;
  double count = 0;
  std::string count_text;
  uint32_t stride = 0;
  std::map<std::string, std::pair<uint32_t, uint8_t>> at;  // name -> offset, type (the last of a name wins)
  static const std::map<std::string, std::pair<uint8_t, uint32_t>> kTypes = {
      {"char", {wd::kPlyI8, 1}},   {"int8", {wd::kPlyI8, 1}},     {"uchar", {wd::kPlyU8, 1}},
      {"uint8", {wd::kPlyU8, 1}},  {"short", {wd::kPlyI16, 2}},   {"int16", {wd::kPlyI16, 2}},
      {"ushort", {wd::kPlyU16, 2}}, {"uint16", {wd::kPlyU16, 2}}, {"int", {wd::kPlyI32, 4}},
      {"int32", {wd::kPlyI32, 4}}, {"uint", {wd::kPlyU32, 4}},    {"uint32", {wd::kPlyU32, 4}},
      {"float", {wd::kPlyF32, 4}}, {"float32", {wd::kPlyF32, 4}}, {"double", {wd::kPlyF64, 8}},
      {"float64", {wd::kPlyF64, 8}}};
  std::istringstream lines(header);
  std::string line;
  while (std::getline(lines, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    const std::vector<std::string> w = words_of(line);
    if (w.empty()) continue;
    const std::string w1 = w.size() > 1 ? w[1] : "undefined";
    if (w[0] == "format") {
      format = w1;
    } else if (w[0] == "element") {
      if (!have_vertices && w1 == "chunk")
        return fail(WD_E_DATA, "splats: a compressed .ply (from SuperSplat) is not read; export an uncompressed .ply or a .splat");
      if (!have_vertices && w1 != "vertex")
        return fail(WD_E_DATA, "splats: a .ply whose first element is '%s', not 'vertex'", w1.c_str());
      if (have_vertices) break;  // elements after the splats are not needed
      have_vertices = true;
      count_text = w.size() > 2 ? w[2] : "NaN";
      char *endp = nullptr;
      count = std::strtod(count_text.c_str(), &endp);
      if (!endp || *endp) count = NAN;
    } else if (w[0] == "property" && have_vertices) {
      if (w1 == "list") return fail(WD_E_DATA, "splats: a .ply with list properties on its vertices is not a splat capture");
      auto t = kTypes.find(w1);
      if (t == kTypes.end()) return fail(WD_E_DATA, "splats: unknown .ply property type '%s'", w1.c_str());
      at[w.size() > 2 ? w[2] : "undefined"] = {stride, t->second.first};
      stride += t->second.second;
    }
  }
  if (!have_vertices) return fail(WD_E_DATA, "splats: a .ply with no vertices");
  if (format != "binary_little_endian")
    return fail(WD_E_DATA, "splats: a .ply in '%s' format; only binary_little_endian is read",
                format.empty() ? "null" : format.c_str());
  if (!(count > 0 && std::floor(count) == count)) return fail(WD_E_DATA, "splats: a .ply with %s vertices", count_text.c_str());
  static const char *kNeeds[wd::kPlyNeeds] = {"x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                              "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
  std::string missing;
  for (const char *name : kNeeds)
    if (!at.count(name)) missing += (missing.empty() ? "" : ", ") + std::string(name);
  if (!missing.empty()) return fail(WD_E_DATA, "splats: this .ply is not a splat capture: no %s", missing.c_str());
  if (count > (double)WD_MAX_SORT)
    return fail(WD_E_CAPACITY, "splats: %.0f splats are more than this device holds in one buffer, %u", count, WD_MAX_SORT);
  const uint64_t n = (uint64_t)count;
  if ((uint64_t)end + n * stride > len)
    return fail(WD_E_DATA, "splats: the .ply is cut short: %llu vertices need %llu bytes, it has %lld",
                (unsigned long long)n, (unsigned long long)(n * stride), (long long)len - (long long)end);
  // Count the consecutive f_rest_k properties: 9, 24 or 45 of them give degree 1, 2 or 3.
  uint32_t rest = 0;
  while (rest < 46 && at.count("f_rest_" + std::to_string(rest))) ++rest;
  int degree = -1;
  for (int d = 0; d < 4; ++d)
    if (rest == 3u * kShCoefficients[d]) degree = d;
  if (rest > 0 && degree < 1)
    return fail(WD_E_DATA, "splats: a .ply with %u f_rest_ properties; harmonics of degree 1, 2 or 3 have 9, 24 or 45", rest);
  wd::PlyLayout L;
  std::memset(&L, 0, sizeof L);
  L.stride = stride;
  L.rest = rest;
  for (int k = 0; k < wd::kPlyNeeds; ++k) {
    L.off[k] = at[kNeeds[k]].first;
    L.type[k] = at[kNeeds[k]].second;
  }
  for (uint32_t k = 0; k < rest; ++k) {
    const auto &p = at["f_rest_" + std::to_string(k)];
    L.off[wd::kPlyNeeds + k] = p.first;
    L.type[wd::kPlyNeeds + k] = p.second;
  }
  DeviceCopy body;
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  WD_HIP(body.from(data + end, n * stride), "splats: .ply upload");
  return decode_into(
      ctx, n, (uint32_t)degree,
      [&](const wd::DecodeOut &o, uint32_t *groups) {
        return wd::launch_decode_ply(ctx, (const uint8_t *)body.p, (uint32_t)n, L, o, groups);
      },
      out, min_out, max_out);
}

int wd_decode_spz(wd_context *ctx, const void *bytes, uint64_t len, wd_cloud **out, float min_out[3],
                  float max_out[3]) {
  if (!ctx || !bytes || !out) return fail(WD_E_ARG, "splats: null argument");
  *out = nullptr;
  const uint8_t *src = (const uint8_t *)bytes;
  std::vector<uint8_t> unpacked;
  const uint8_t *data = src;
  uint64_t size = len;
  auto u32 = [](const uint8_t *p) { return (uint32_t)p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); };
  if (len >= 2 && src[0] == 0x1f && src[1] == 0x8b) {
    int rc = gunzip(src, len, unpacked);
    if (rc) return rc;
    data = unpacked.data();
    size = unpacked.size();
  } else if (len >= 8 && u32(src) == 0x5053474eu && u32(src + 4) >= 4) {
    return fail(WD_E_DATA, "splats: a .spz of version 4, compressed with zstd, which browsers cannot unpack; "
                           "save it as version 3, or as a .ply or .splat");
  }
  // parseSpz's header and sections (src/scene/splats.js:230-256)
  if (size < 16 || u32(data) != 0x5053474eu) return fail(WD_E_DATA, "splats: gzipped, but not a .spz");
  const uint32_t version = u32(data + 4);
  if (version < 1 || version > 3) return fail(WD_E_DATA, "splats: a .spz of version %u; versions 1 to 3 are read", version);
  const uint32_t count = u32(data + 8);
  const uint32_t file_degree = data[12], fraction = data[13];
  if (file_degree > 4) return fail(WD_E_DATA, "splats: a .spz with harmonics of degree %u", file_degree);
  const uint32_t degree = std::min(file_degree, 3u);  // degree 4's band, which nothing here evaluates, is left out
  wd::SpzLayout L;
  std::memset(&L, 0, sizeof L);
  L.version = version;
  L.rotation_bytes = version >= 3 ? 4 : 3;
  L.file_coefficients = kShCoefficients[file_degree];
  L.coefficients = kShCoefficients[degree];
  const uint64_t position_bytes = version == 1 ? 6 : 9;
  uint64_t at = 16;
  const uint64_t sizes[6] = {position_bytes, 1, 3, 3, L.rotation_bytes, (uint64_t)L.file_coefficients * 3};
  uint64_t *offs[6] = {&L.positions, &L.alphas, &L.colours, &L.scales, &L.rotations, &L.sh};
  for (int k = 0; k < 6; ++k) {
    *offs[k] = at;
    at += (uint64_t)count * sizes[k];
  }
  if (count == 0) return fail(WD_E_DATA, "splats: a .spz with no splats");
  if (at > size)
    return fail(WD_E_DATA, "splats: the .spz is cut short: %u splats need %llu bytes, it has %llu", count,
                (unsigned long long)at, (unsigned long long)size);
  L.fixed = 1.0 / (double)(int32_t)(1u << (fraction & 31u));  // JavaScript's 1 << fraction
  DeviceCopy dev;
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  WD_HIP(dev.from(data, at), "splats: .spz upload");
  return decode_into(
      ctx, count, degree,
      [&](const wd::DecodeOut &o, uint32_t *groups) {
        return wd::launch_decode_spz(ctx, (const uint8_t *)dev.p, count, L, o, groups);
      },
      out,
      min_out, max_out);
}

int wd_decode_sog(wd_context *ctx, const wd_sog *sog, wd_cloud **out, float min_out[3], float max_out[3]) {
  if (!ctx || !sog || !out) return fail(WD_E_ARG, "splats: null argument");
  *out = nullptr;
  const uint32_t count = sog->count;
  if (count == 0) return fail(WD_E_DATA, "splats: a .sog of 0 splats");
  const uint32_t degree = sog->bands >= 1 && sog->bands <= 3 ? sog->bands : 0;
  const uint32_t per = kShCoefficients[degree];
  const wd_sog_image *need[6] = {&sog->means_l, &sog->means_u, &sog->quats, &sog->scales, &sog->sh0, &sog->labels};
  static const char *kNames[6] = {"means_l", "means_u", "quats", "scales", "sh0", "shN_labels"};
  for (int k = 0; k < (degree ? 6 : 5); ++k) {
    if (!need[k]->rgba) return fail(WD_E_ARG, "splats: the .sog's %s image is missing", kNames[k]);
    const uint64_t px = (uint64_t)need[k]->width * need[k]->height;
    if (px < count) return fail(WD_E_DATA, "splats: %s has %llu pixels, for %u splats", kNames[k], (unsigned long long)px, count);
  }
  if (!sog->scale_codebook || !sog->colour_codebook || (degree && !sog->sh_codebook))
    return fail(WD_E_ARG, "splats: the .sog's codebooks are missing");
  if (degree) {
    if (!sog->centroids.rgba) return fail(WD_E_ARG, "splats: the .sog's shN_centroids image is missing");
    if (sog->centroids.width != 64 * per)
      return fail(WD_E_DATA, "splats: the .sog's harmonic palette is %u wide; %u for degree %u", sog->centroids.width,
                  64 * per, degree);
  }
  // Per codebook index: the colour byte, exp of the scale, the harmonic's half bits. An index
  // past a codebook's end reads undefined there, a NaN here.
  auto book = [](const double *b, uint32_t n, int k) { return (uint32_t)k < n ? b[k] : NAN; };
  uint8_t colour_byte[256];
  double scale_exp[256];
  uint16_t sh_half[256];
  for (int k = 0; k < 256; ++k) {
    const double r = js_round((0.5 + 0.28209479177387814 * book(sog->colour_codebook, sog->colour_codebook_size, k)) * 255.0);
    colour_byte[k] = !(r >= 0.0) ? 0 : r > 255.0 ? 255 : (uint8_t)r;
    scale_exp[k] = std::exp(book(sog->scale_codebook, sog->scale_codebook_size, k));
    sh_half[k] = degree ? signed_half_host(book(sog->sh_codebook, sog->sh_codebook_size, k)) : 0;
  }
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  DeviceCopy img[7], tabs[3];
  const wd_sog_image *all[7] = {&sog->means_l, &sog->means_u, &sog->quats, &sog->scales, &sog->sh0, &sog->labels,
                                &sog->centroids};
  for (int k = 0; k < 7; ++k) {
    if (k >= 5 && !degree) continue;
    WD_HIP(img[k].from(all[k]->rgba, (size_t)4 * all[k]->width * all[k]->height), "splats: .sog upload");
  }
  WD_HIP(tabs[0].from(colour_byte, sizeof colour_byte), "splats: .sog tables");
  WD_HIP(tabs[1].from(scale_exp, sizeof scale_exp), "splats: .sog tables");
  WD_HIP(tabs[2].from(sh_half, sizeof sh_half), "splats: .sog tables");
  wd::SogLayout L;
  std::memset(&L, 0, sizeof L);
  L.count = count;
  L.per = per;
  L.sh_count = sog->sh_count;
  L.centroids_width = sog->centroids.width;
  L.centroids_bytes = degree ? (uint64_t)4 * sog->centroids.width * sog->centroids.height : 0;
  for (int a = 0; a < 3; ++a) {
    L.mins[a] = sog->mins[a];
    L.maxs[a] = sog->maxs[a];
  }
  const uint8_t **ptrs[7] = {&L.means_l, &L.means_u, &L.quats, &L.scales, &L.sh0, &L.labels, &L.centroids};
  for (int k = 0; k < 7; ++k) *ptrs[k] = (const uint8_t *)img[k].p;
  L.colour_byte = (const uint8_t *)tabs[0].p;
  L.scale_exp = (const double *)tabs[1].p;
  L.sh_half = (const uint16_t *)tabs[2].p;
  return decode_into(
      ctx, count, degree, [&](const wd::DecodeOut &o, uint32_t *groups) { return wd::launch_decode_sog(ctx, L, o, groups); },
      out, min_out,
      max_out);
}

int wd_cloud_read(wd_context *ctx, const wd_cloud *cloud, float *centers, float *covariances, uint32_t *sh,
                  uint32_t *degree) {
  if (!ctx || !cloud || !ctx->live.count(cloud)) return fail(WD_E_STATE, "splats: these splats were unloaded");
  WD_HIP(hipSetDevice(ctx->device), "splats: set device");
  WD_HIP(hipStreamSynchronize(ctx->stream), "splats: read");
  if ((centers || covariances) && !cloud->centers) return fail(WD_E_ARG, "splats: a cloud of .splat records has no centres to read");
  if (centers) WD_HIP(hipMemcpy(centers, cloud->centers, cloud->n * 16, hipMemcpyDeviceToHost), "splats: read");
  if (covariances) WD_HIP(hipMemcpy(covariances, cloud->covs, cloud->n * 24, hipMemcpyDeviceToHost), "splats: read");
  if (sh && cloud->degree)
    WD_HIP(hipMemcpy(sh, cloud->sh, cloud->n * wd::sh_words(cloud->degree) * 4, hipMemcpyDeviceToHost), "splats: read");
  if (degree) *degree = cloud->degree;
  return WD_OK;
}

}  // extern "C"

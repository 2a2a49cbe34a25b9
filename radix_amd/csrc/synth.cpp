// Seeded synthetic splat clouds (host, C++): the benchmark and test inputs.
//
// Random numbers are mulberry32, the reference's test generator
// (test/gpu/renderUtils.ts:23-30); normals are Box-Muller pairs in double.
// mulberry32's state only advances by a constant, so draw k of a stream is a
// pure function of seed + (k + 1) * 0x6d2b79f5: splat i starts at draw i * D
// (D draws per splat, fixed per cloud), which lets the splats be generated in
// parallel with the same numbers as a sequential run.
//
// Also the packed encoder (splats.ts:271-309) and packed SH encoder
// (sh.ts:137-174) used to build quantized tables.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr uint32_t GOLDEN = 0x6d2b79f5u;

struct Mulberry {
  uint32_t s;
  explicit Mulberry(uint32_t seed) : s(seed) {}
  double next() {
    s += GOLDEN;
    uint32_t t = (s ^ (s >> 15)) * (1u | s);
    t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
    return (double)(t ^ (t >> 14)) / 4294967296.0;
  }
};

struct Normals {
  Mulberry& r;
  bool have = false;
  double spare = 0.0;
  explicit Normals(Mulberry& m) : r(m) {}
  double next() {
    if (have) {
      have = false;
      return spare;
    }
    double u1 = r.next(), u2 = r.next();
    if (u1 < 1e-300) u1 = 1e-300;
    const double rad = sqrt(-2.0 * log(u1));
    spare = rad * sin(2.0 * M_PI * u2);
    have = true;
    return rad * cos(2.0 * M_PI * u2);
  }
};

const int SH_COEFFS[4] = {0, 3, 8, 15};

}  // namespace

extern "C" {

// A cloud of `count` splats into f32 planes of `stride` floats:
// centers: 0 = uniform in [-1, 1]^3 + offset; 1 = a mixture of `clusters`
// anisotropic Gaussians (means uniform in [-extent, extent]^3, per-axis std
// log-uniform in [std_lo, std_hi], drawn from seed ^ 0x9e3779b9).
// log-scales N(log_scale_mean, log_scale_std) per axis; rotation a
// normalised 4D standard normal; opacity sigmoid(N(0, 1)); base colour
// uniform [0, 1]; SH bands (degree sh_degree) N(0, sh_std).
typedef struct rdxs_cloud {
  uint32_t seed;
  uint32_t count;
  int sh_degree;
  int centers;
  double offset[3];
  double log_scale_mean, log_scale_std;
  double sh_std;
  int clusters;
  double extent, std_lo, std_hi;
} rdxs_cloud;

int rdxs_draws_per_splat(const rdxs_cloud* c) {
  const int K = SH_COEFFS[std::max(0, std::min(c->sh_degree, 3))];
  // uniforms: cluster pick (mixture) or 3 (cube), 3 colour; normals rounded up to pairs
  const int normals = (c->centers == 1 ? 3 : 0) + 3 + 4 + 1 + 3 * K;
  const int uniforms = (c->centers == 1 ? 1 : 3) + 3;
  return uniforms + 2 * ((normals + 1) / 2);
}

int rdxs_generate(const rdxs_cloud* c, float* center, float* log_scale, float* quat, float* opacity, float* rgb,
                  float* sh, size_t stride) {
  const uint32_t n = c->count;
  if (!stride) stride = n;
  const int K = SH_COEFFS[std::max(0, std::min(c->sh_degree, 3))];
  const int D = rdxs_draws_per_splat(c);
  std::vector<double> cm, cs;
  if (c->centers == 1) {
    Mulberry r(c->seed ^ 0x9e3779b9u);
    cm.resize(3 * c->clusters);
    cs.resize(3 * c->clusters);
    for (int k = 0; k < c->clusters; k++) {
      for (int a = 0; a < 3; a++) cm[3 * k + a] = (r.next() * 2.0 - 1.0) * c->extent;
      for (int a = 0; a < 3; a++) cs[3 * k + a] = exp(log(c->std_lo) + r.next() * (log(c->std_hi) - log(c->std_lo)));
    }
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    Mulberry r(c->seed + (uint32_t)((uint64_t)i * (uint64_t)D * GOLDEN));
    Normals g(r);
    double p[3];
    if (c->centers == 1) {
      int k = (int)(r.next() * c->clusters);
      if (k >= c->clusters) k = c->clusters - 1;
      for (int a = 0; a < 3; a++) p[a] = cm[3 * k + a] + cs[3 * k + a] * g.next();
    } else {
      for (int a = 0; a < 3; a++) p[a] = r.next() * 2.0 - 1.0 + c->offset[a];
    }
    double ls[3], q[4];
    for (int a = 0; a < 3; a++) ls[a] = c->log_scale_mean + c->log_scale_std * g.next();
    for (int a = 0; a < 4; a++) q[a] = g.next();
    const double ql = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double op = 1.0 / (1.0 + exp(-g.next()));
    double col[3];
    for (int a = 0; a < 3; a++) col[a] = r.next();
    for (int a = 0; a < 3; a++) {
      center[a * stride + i] = (float)p[a];
      log_scale[a * stride + i] = (float)ls[a];
      rgb[a * stride + i] = (float)col[a];
    }
    for (int a = 0; a < 4; a++) quat[a * stride + i] = (float)(ql > 0 ? q[a] / ql : (a == 3 ? 1.0 : 0.0));
    opacity[i] = (float)op;
    if (sh)
      for (int v = 0; v < 3 * K; v++) sh[(size_t)v * stride + i] = (float)(c->sh_std * g.next());
  }
  return 0;
}

// ---- packed encoder (splats.ts:271-309, encoding.ts) ----------------------

static double js_round(double x) { return floor(x + 0.5); }

static uint32_t half_bits(float f32) {
  uint32_t bits;
  memcpy(&bits, &f32, 4);
  const uint32_t sign = (bits >> 16) & 0x8000u;
  const uint32_t ex = (bits >> 23) & 0xffu;
  const uint32_t man = bits & 0x7fffffu;
  if (ex == 0xffu) return sign | 0x7c00u | (man ? 0x200u : 0u);
  const int e = (int)ex - 127 + 15;
  if (e >= 0x1f) return sign | 0x7c00u;
  if (e <= 0) {
    const double m = fabs((double)f32) * 16777216.0;
    double fl = floor(m);
    const double fr = m - fl;
    uint32_t r = (uint32_t)fl;
    if (fr > 0.5 || (fr == 0.5 && (r & 1u))) r++;
    return sign | r;
  }
  const uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13);
  const uint32_t rest = man & 0x1fffu;
  return (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) ? h + 1u : h;
}

static uint32_t oct_encode(const double* v, int bits) {
  const double mx = ldexp(1.0, bits) - 1.0;
  const uint32_t ones = (uint32_t)mx | ((uint32_t)mx << bits);
  const double l1 = fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
  if (!(l1 > 0)) return 0;
  double px = v[0] / l1, py = v[1] / l1;
  if (v[2] < 0) {
    const double ax = fabs(px);
    px = (1.0 - fabs(py)) * (px >= 0 ? 1.0 : -1.0);
    py = (1.0 - ax) * (py >= 0 ? 1.0 : -1.0);
  }
  const uint32_t code = (uint32_t)js_round((px * 0.5 + 0.5) * mx) | ((uint32_t)js_round((py * 0.5 + 0.5) * mx) << bits);
  return code == 0 ? ones : code;
}

static uint32_t quat_oct_encode(const double* q, int axis_bits, int angle_bits) {
  const double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (!(len > 0)) return 0;
  const double amax = ldexp(1.0, angle_bits) - 1.0;
  const double s = (q[3] < 0 ? -1.0 : 1.0) / len;
  const double axis[3] = {q[0] * s, q[1] * s, q[2] * s};
  const double angle = 2.0 * acos(std::min(1.0, q[3] * s));
  const uint32_t ac = (uint32_t)js_round(angle / M_PI * amax);
  return oct_encode(axis, axis_bits) | (ac << (2 * axis_bits));
}

static uint32_t unorm8(double v) { return (uint32_t)std::max(0.0, std::min(255.0, js_round(v * 255.0))); }

// f32 planes to packedSplats words (uint32x4 per splat) with encoding enc
// (rgbMin, rgbMax, lnScaleMin, lnScaleMax).
int rdxs_pack_splats(uint32_t n, const float* center, const float* log_scale, const float* quat, const float* opacity,
                     const float* rgb, size_t stride, const double* enc, uint32_t* words) {
  if (!stride) stride = n;
  const double range = enc[1] - enc[0];
  const double step = (enc[3] - enc[2]) / 254.0;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    double q[4];
    for (int a = 0; a < 4; a++) q[a] = quat[a * stride + i];
    const uint32_t qc = quat_oct_encode(q, 8, 8);
    uint32_t sc[3];
    for (int a = 0; a < 3; a++) {
      const double s = exp((double)log_scale[a * stride + i]);
      sc[a] = s <= 0 ? 0u : (uint32_t)std::max(1.0, std::min(255.0, js_round((log(s) - enc[2]) / step) + 1.0));
    }
    uint32_t* w = words + 4 * (size_t)i;
    w[0] = unorm8((rgb[i] - enc[0]) / range) | (unorm8((rgb[stride + i] - enc[0]) / range) << 8) |
           (unorm8((rgb[2 * stride + i] - enc[0]) / range) << 16) | (unorm8(opacity[i]) << 24);
    w[1] = half_bits(center[i]) | (half_bits(center[stride + i]) << 16);
    w[2] = half_bits(center[2 * stride + i]) | ((qc & 0xffffu) << 16);
    w[3] = sc[0] | (sc[1] << 8) | (sc[2] << 16) | ((qc >> 16) << 24);
  }
  return 0;
}

static double sh_scale_of(uint32_t code) { return code == 0 ? 0.0 : exp2(((double)code - 192.0) / 16.0); }

static uint32_t sh_scale_code(double mx) {
  if (!(mx > 0)) return 0;
  uint32_t code = (uint32_t)std::max(1.0, std::min(255.0, ceil(log2(mx) * 16.0) + 192.0));
  while (code < 255 && sh_scale_of(code) < mx) code++;
  return code;
}

// f32 SH planes (3K planes, plane 3k + c) to packedSh channels: `channels`
// uint32x4 arrays (degree 3: 3 of them), channel t of splat i at
// channels[t] + 4 i.
int rdxs_pack_sh(uint32_t n, int degree, const float* sh, size_t stride, uint32_t* const* channels) {
  if (!stride) stride = n;
  static const int band_bytes[3] = {10, 16, 22};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t w[12] = {0};
    int byte = 0;
    for (int band = 0; band < degree; band++) {
      const int first = SH_COEFFS[band] * 3, end = SH_COEFFS[band + 1] * 3;
      double mx = 0;
      for (int v = first; v < end; v++) mx = std::max(mx, fabs((double)sh[(size_t)v * stride + i]));
      const uint32_t code = sh_scale_code(mx);
      const double scale = sh_scale_of(code);
      w[byte >> 2] |= code << ((byte & 3) * 8);
      int b = byte + 1;
      for (int v = first; v < end; v++, b++) {
        const int qv = scale > 0 ? (int)std::max(-127.0, std::min(127.0, js_round(sh[(size_t)v * stride + i] / scale * 127.0))) : 0;
        w[b >> 2] |= ((uint32_t)qv & 0xffu) << ((b & 3) * 8);
      }
      byte += band_bytes[band];
    }
    const int texels = degree;
    for (int t = 0; t < texels; t++)
      for (int k = 0; k < 4; k++) channels[t][4 * (size_t)i + k] = w[4 * t + k];
  }
  return 0;
}

// Uniform draws of a seeded stream (for cameras and small picks).
int rdxs_uniforms(uint32_t seed, uint32_t n, double* out) {
  Mulberry r(seed);
  for (uint32_t i = 0; i < n; i++) out[i] = r.next();
  return 0;
}

// Standard normals of a seeded stream (Box-Muller pairs).
int rdxs_normals(uint32_t seed, uint32_t n, double* out) {
  Mulberry r(seed);
  Normals g(r);
  for (uint32_t i = 0; i < n; i++) out[i] = g.next();
  return 0;
}

}  // extern "C"

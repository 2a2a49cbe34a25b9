/*
 * Oracle: scalar/vector codecs, splat layouts, SH codecs and basis, and the
 * covariance of a scale + rotation, restated in double from src/core
 * (encoding.ts, splats.ts, sh.ts, math.ts). Test infrastructure only.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

/* Math.round (ties toward +infinity) */
static double js_round(double x) { return floor(x + 0.5); }
static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* --- encoding.ts:156-189: half bits of f via its float32 bits, ties to even --- */
uint32_t orc_encode_half(double value) {
  float f32 = (float)value;
  uint32_t bits;
  memcpy(&bits, &f32, 4);
  uint32_t sign = (bits >> 16) & 0x8000u;
  uint32_t exponent = (bits >> 23) & 0xffu;
  uint32_t mantissa = bits & 0x7fffffu;
  if (exponent == 0xffu) return sign | 0x7c00u | (mantissa ? 0x200u : 0u);
  int e = (int)exponent - 127 + 15;
  if (e >= 0x1f) return sign | 0x7c00u;
  if (e <= 0) {
    double m = fabs((double)f32) * 16777216.0; /* multiples of 2^-24 */
    double r = floor(m);
    double frac = m - r;
    uint32_t rounded = (uint32_t)r;
    if (frac > 0.5 || (frac == 0.5 && (rounded & 1u))) rounded += 1u;
    return sign | rounded;
  }
  uint32_t half = sign | ((uint32_t)e << 10) | (mantissa >> 13);
  uint32_t rest = mantissa & 0x1fffu;
  return (rest > 0x1000u || (rest == 0x1000u && (half & 1u))) ? half + 1u : half;
}

/* encoding.ts:191-202 */
double orc_decode_half(uint32_t h) {
  double sign = (h & 0x8000u) ? -1.0 : 1.0;
  uint32_t exponent = (h >> 10) & 0x1fu;
  uint32_t mantissa = h & 0x3ffu;
  if (exponent == 0) return sign * (double)mantissa * ldexp(1.0, -24);
  if (exponent == 0x1f) return mantissa ? NAN : sign * INFINITY;
  return sign * (1.0 + mantissa / 1024.0) * ldexp(1.0, (int)exponent - 15);
}

/* ceil(log2(m)) for m > 0, exactly (frexp: m = f * 2^e, f in [0.5, 1)) */
static int ceil_log2(double m) {
  int e;
  double f = frexp(m, &e);
  return f == 0.5 ? e - 1 : e;
}

/* encoding.ts:219-231: 8-bit magnitudes, 3 sign bits, 5-bit shared exponent */
uint32_t orc_encode_ext_rgb(double r, double g, double b) {
  double v[3] = {r, g, b};
  double max_abs = fmax(fabs(r), fmax(fabs(g), fabs(b)));
  int base = (int)clampd((max_abs > 0 ? ceil_log2(max_abs) : 0) + 15, 0, 31);
  double scale = 255.0 * ldexp(1.0, 15 - base);
  uint32_t word = 0;
  for (int c = 0; c < 3; c++) {
    double m = fmin(255.0, js_round(fabs(v[c]) * scale));
    word |= (uint32_t)m << (8 * c);
    if (v[c] < 0) word |= 1u << (24 + c);
  }
  return word | ((uint32_t)base << 27);
}

/* encoding.ts:233-239 */
void orc_decode_ext_rgb(uint32_t e, double* rgb) {
  double unit = ldexp(1.0, (int)((e >> 27) & 0x1fu) - 15) / 255.0;
  for (int c = 0; c < 3; c++) {
    double m = (double)((e >> (8 * c)) & 0xffu) * unit;
    rgb[c] = (e & (1u << (24 + c))) ? -m : m;
  }
}

/* encoding.ts:25-47: octahedral direction, `bits` per axis; code 0 is none */
uint32_t orc_encode_oct(const double* v, int bits) {
  const double max = ldexp(1.0, bits) - 1.0;
  const uint32_t all_ones = (uint32_t)max | ((uint32_t)max << bits);
  double l1 = fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
  if (!(l1 > 0)) return 0;
  double px = v[0] / l1, py = v[1] / l1;
  if (v[2] < 0) {
    double ax = fabs(px);
    px = (1.0 - fabs(py)) * (px >= 0 ? 1.0 : -1.0);
    py = (1.0 - ax) * (py >= 0 ? 1.0 : -1.0);
  }
  uint32_t code = (uint32_t)js_round((px * 0.5 + 0.5) * max) | ((uint32_t)js_round((py * 0.5 + 0.5) * max) << bits);
  return code == 0 ? all_ones : code;
}

/* encoding.ts:49-66 */
void orc_decode_oct(uint32_t code, int bits, double* out) {
  if (code == 0) {
    out[0] = out[1] = out[2] = 0;
    return;
  }
  const uint32_t mask = (1u << bits) - 1u;
  const double max = (double)mask;
  double x = ((code & mask) / max) * 2.0 - 1.0;
  double y = (((code >> bits) & mask) / max) * 2.0 - 1.0;
  double z = 1.0 - fabs(x) - fabs(y);
  double t = fmax(-z, 0.0);
  x += x >= 0 ? -t : t;
  y += y >= 0 ? -t : t;
  double len = sqrt(x * x + y * y + z * z);
  out[0] = x / len;
  out[1] = y / len;
  out[2] = z / len;
}

/* encoding.ts:85-98: axis (octahedral) low, angle 0..pi above, w >= 0 */
uint32_t orc_encode_quat_oct(const double* q, int axis_bits, int angle_bits) {
  double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (!(len > 0)) return 0;
  const double angle_max = ldexp(1.0, angle_bits) - 1.0;
  double s = (q[3] < 0 ? -1.0 : 1.0) / len;
  double axis[3] = {q[0] * s, q[1] * s, q[2] * s};
  double angle = 2.0 * acos(fmin(1.0, q[3] * s));
  uint32_t angle_code = (uint32_t)js_round((angle / M_PI) * angle_max);
  return orc_encode_oct(axis, axis_bits) | (angle_code << (2 * axis_bits));
}

/* encoding.ts:100-110 */
void orc_decode_quat_oct(uint32_t code, int axis_bits, int angle_bits, double* q) {
  const uint32_t axis_mask = (uint32_t)((1ull << (2 * axis_bits)) - 1u);
  const uint32_t angle_max = (1u << angle_bits) - 1u;
  double axis[3];
  orc_decode_oct(code & axis_mask, axis_bits, axis);
  double angle = ((double)((code >> (2 * axis_bits)) & angle_max) / angle_max) * M_PI;
  double s = sin(angle / 2);
  q[0] = axis[0] * s;
  q[1] = axis[1] * s;
  q[2] = axis[2] * s;
  q[3] = cos(angle / 2);
}

/* encoding.ts:136-138 (Math.max(0, Math.min(255, Math.round(v * 255)))) */
static uint32_t unorm8(double v) { return (uint32_t)clampd(js_round(v * 255.0), 0, 255); }

/* --- splats.ts:271-309: the packed codec --- */
void orc_packed_encode(const double* center, const double* scales, const double* quat, const double* rgb,
                       double opacity, const double* enc, uint32_t* w) {
  const double rgb_range = enc[1] - enc[0];
  const double ln_step = (enc[3] - enc[2]) / 254.0;
  uint32_t qc = orc_encode_quat_oct(quat, 8, 8);
  uint32_t sc[3];
  for (int k = 0; k < 3; k++) {
    double v = scales[k];
    sc[k] = v <= 0 ? 0u : (uint32_t)clampd(js_round((log(v) - enc[2]) / ln_step) + 1.0, 1.0, 255.0);
  }
  w[0] = unorm8((rgb[0] - enc[0]) / rgb_range) | (unorm8((rgb[1] - enc[0]) / rgb_range) << 8) |
         (unorm8((rgb[2] - enc[0]) / rgb_range) << 16) | (unorm8(opacity) << 24);
  w[1] = orc_encode_half(center[0]) | (orc_encode_half(center[1]) << 16);
  w[2] = orc_encode_half(center[2]) | ((qc & 0xffffu) << 16);
  w[3] = sc[0] | (sc[1] << 8) | (sc[2] << 16) | ((qc >> 16) << 24);
}

void orc_packed_decode(const uint32_t* w, const double* enc, double* center, double* scales, double* quat,
                       double* rgb, double* opacity) {
  const double rgb_range = enc[1] - enc[0];
  const double ln_step = (enc[3] - enc[2]) / 254.0;
  for (int c = 0; c < 3; c++) rgb[c] = enc[0] + ((w[0] >> (8 * c)) & 0xffu) / 255.0 * rgb_range;
  *opacity = (w[0] >> 24) / 255.0;
  center[0] = orc_decode_half(w[1] & 0xffffu);
  center[1] = orc_decode_half(w[1] >> 16);
  center[2] = orc_decode_half(w[2] & 0xffffu);
  for (int k = 0; k < 3; k++) {
    uint32_t code = (w[3] >> (8 * k)) & 0xffu;
    scales[k] = code == 0 ? 0.0 : exp(enc[2] + (code - 1.0) * ln_step);
  }
  orc_decode_quat_oct((w[2] >> 16) | ((w[3] >> 24) << 16), 8, 8, quat);
}

/* --- splats.ts:173-197: the ext codec (no normal, roughness, metallic) --- */
void orc_ext_encode(const double* center, const double* scales, const double* quat, const double* rgb,
                    double opacity, uint32_t* a, uint32_t* b) {
  for (int k = 0; k < 3; k++) {
    float c = (float)center[k];
    memcpy(&a[k], &c, 4);
  }
  a[3] = (uint32_t)clampd(js_round(opacity * 128.0), 0, 255) << 24; /* encoding.ts:140-143 */
  b[0] = orc_encode_ext_rgb(rgb[0], rgb[1], rgb[2]);
  b[1] = orc_encode_half(log(scales[0])) | (orc_encode_half(log(scales[1])) << 16);
  b[2] = orc_encode_half(log(scales[2]));
  b[3] = orc_encode_quat_oct(quat, 10, 12);
}

void orc_ext_decode(const uint32_t* a, const uint32_t* b, double* center, double* scales, double* quat, double* rgb,
                    double* opacity) {
  for (int k = 0; k < 3; k++) {
    float c;
    memcpy(&c, &a[k], 4);
    center[k] = c;
  }
  *opacity = (a[3] >> 24) / 128.0;
  orc_decode_ext_rgb(b[0], rgb);
  scales[0] = exp(orc_decode_half(b[1] & 0xffffu));
  scales[1] = exp(orc_decode_half(b[1] >> 16));
  scales[2] = exp(orc_decode_half(b[2] & 0xffffu));
  orc_decode_quat_oct(b[3], 10, 12, quat);
}

/* --- splats.ts:219-266: the cov words (A as ext, no normal; B.x ext rgb,
 * B.y half(ln xx) | half(ln yy) << 16, B.z half(ln zz) | half(c_xy) << 16,
 * B.w half(c_xz) | half(c_yz) << 16, c = sign * ln|correlation|) --- */

/* splats.ts:221-228: the correlation value / sqrt(a b) as sign(c) ln|c|
 * (Math.max(-1, Math.min(1, NaN)) is NaN, which encodes as 0) */
double orc_cov_encode_correlation(double value, double a, double b) {
  double c = value / sqrt(a * b);
  if (isnan(c)) return 0.0;
  c = clampd(c, -1.0, 1.0);
  if (c == 0.0) return 0.0;
  return (c > 0 ? 1.0 : -1.0) * clampd(log(fabs(c)), -100.0, -1e-7);
}

/* splats.ts:230-232 */
double orc_cov_decode_correlation(double e) { return e == 0.0 ? 0.0 : (e > 0 ? -1.0 : 1.0) * exp(-fabs(e)); }

void orc_cov_encode(const double* center, const double* cov6, const double* rgb, double opacity, uint32_t* a,
                    uint32_t* b) {
  const double xx = cov6[0], yy = cov6[1], zz = cov6[2], xy = cov6[3], xz = cov6[4], yz = cov6[5];
  for (int k = 0; k < 3; k++) {
    float c = (float)center[k];
    memcpy(&a[k], &c, 4);
  }
  a[3] = (uint32_t)clampd(js_round(opacity * 128.0), 0, 255) << 24; /* encoding.ts:140-143; normal none */
  b[0] = orc_encode_ext_rgb(rgb[0], rgb[1], rgb[2]);
  b[1] = orc_encode_half(log(xx)) | (orc_encode_half(log(yy)) << 16);
  b[2] = orc_encode_half(log(zz)) | (orc_encode_half(orc_cov_encode_correlation(xy, xx, yy)) << 16);
  b[3] = orc_encode_half(orc_cov_encode_correlation(xz, xx, zz)) |
         (orc_encode_half(orc_cov_encode_correlation(yz, yy, zz)) << 16);
}

void orc_cov_decode(const uint32_t* a, const uint32_t* b, double* center, double* cov6, double* rgb,
                    double* opacity) {
  for (int k = 0; k < 3; k++) {
    float c;
    memcpy(&c, &a[k], 4);
    center[k] = c;
  }
  *opacity = (a[3] >> 24) / 128.0;
  orc_decode_ext_rgb(b[0], rgb);
  const double xx = exp(orc_decode_half(b[1] & 0xffffu)), yy = exp(orc_decode_half(b[1] >> 16));
  const double zz = exp(orc_decode_half(b[2] & 0xffffu));
  cov6[0] = xx;
  cov6[1] = yy;
  cov6[2] = zz;
  cov6[3] = orc_cov_decode_correlation(orc_decode_half(b[2] >> 16)) * sqrt(xx * yy);
  cov6[4] = orc_cov_decode_correlation(orc_decode_half(b[3] & 0xffffu)) * sqrt(xx * zz);
  cov6[5] = orc_cov_decode_correlation(orc_decode_half(b[3] >> 16)) * sqrt(yy * zz);
}

/* --- sh.ts: coefficients past DC per degree, packed band sizes --- */
static const int SH_COEFFS[5] = {0, 3, 8, 15, 24};
static const int PACKED_BAND_BYTES[4] = {10, 16, 22, 28};
static const int PACKED_BYTES[5] = {0, 10, 26, 48, 76};

/* sh.ts:82-84 */
double orc_sh_scale_decode(uint32_t code) { return code == 0 ? 0.0 : exp2(((double)code - 192.0) / 16.0); }

/* sh.ts:71-80: the smallest code whose scale covers max */
uint32_t orc_sh_scale_encode(double max) {
  if (!(max > 0)) return 0;
  double c = clampd(ceil(log2(max) * 16.0) + 192.0, 1.0, 255.0);
  uint32_t code = (uint32_t)c;
  while (code < 255 && orc_sh_scale_decode(code) < max) code++;
  return code;
}

int orc_sh_words(int codec, int degree) {
  if (codec == 2) return SH_COEFFS[degree];
  return (PACKED_BYTES[degree] + 3) / 4;
}

/* sh.ts:114-174 */
void orc_sh_encode(int codec, const double* coeffs, int degree, uint32_t* words) {
  if (codec == 2) {
    for (int k = 0; k < SH_COEFFS[degree]; k++)
      words[k] = orc_encode_ext_rgb(coeffs[3 * k], coeffs[3 * k + 1], coeffs[3 * k + 2]);
    return;
  }
  int nwords = orc_sh_words(codec, degree);
  memset(words, 0, (size_t)nwords * 4);
  int byte = 0;
  for (int band = 0; band < degree; band++) {
    int first = SH_COEFFS[band] * 3, end = SH_COEFFS[band + 1] * 3;
    double max = 0;
    for (int v = first; v < end; v++) max = fmax(max, fabs(coeffs[v]));
    uint32_t code = orc_sh_scale_encode(max);
    double scale = orc_sh_scale_decode(code);
    words[byte >> 2] |= code << ((byte & 3) * 8);
    byte++;
    for (int v = first; v < end; v++) {
      int q = scale > 0 ? (int)clampd(js_round(coeffs[v] / scale * 127.0), -127, 127) : 0;
      words[byte >> 2] |= ((uint32_t)q & 0xffu) << ((byte & 3) * 8);
      byte++;
    }
  }
}

static uint32_t get_byte(const uint32_t* words, int b) { return (words[b >> 2] >> ((b & 3) * 8)) & 0xffu; }

void orc_sh_decode(int codec, const uint32_t* words, int degree, double* coeffs) {
  if (codec == 2) {
    for (int k = 0; k < SH_COEFFS[degree]; k++) orc_decode_ext_rgb(words[k], &coeffs[3 * k]);
    return;
  }
  int byte = 0;
  for (int band = 0; band < degree; band++) {
    int first = SH_COEFFS[band] * 3, end = SH_COEFFS[band + 1] * 3;
    double scale = orc_sh_scale_decode(get_byte(words, byte)) / 127.0;
    for (int v = first; v < end; v++) {
      int8_t q = (int8_t)get_byte(words, byte + 1 + v - first);
      coeffs[v] = q * scale;
    }
    byte += PACKED_BAND_BYTES[band];
  }
}

/* sh.ts:376-407: the 24 basis functions of bands 1-4 (3DGS signs) */
void orc_sh_basis(const double* d, double* o) {
  double x = d[0], y = d[1], z = d[2];
  double xx = x * x, yy = y * y, zz = z * z;
  o[0] = -0.4886025 * y;
  o[1] = 0.4886025 * z;
  o[2] = -0.4886025 * x;
  o[3] = 1.0925484 * x * y;
  o[4] = -1.0925484 * y * z;
  o[5] = 0.3153916 * (2 * zz - xx - yy);
  o[6] = -1.0925484 * x * z;
  o[7] = 0.5462742 * (xx - yy);
  o[8] = -0.5900436 * y * (3 * xx - yy);
  o[9] = 2.8906114 * x * y * z;
  o[10] = -0.4570458 * y * (4 * zz - xx - yy);
  o[11] = 0.3731763 * z * (2 * zz - 3 * xx - 3 * yy);
  o[12] = -0.4570458 * x * (4 * zz - xx - yy);
  o[13] = 1.4453057 * z * (xx - yy);
  o[14] = -0.5900436 * x * (xx - 3 * yy);
  o[15] = 2.5033429 * x * y * (xx - yy);
  o[16] = -1.7701308 * y * z * (3 * xx - yy);
  o[17] = 0.9461747 * x * y * (7 * zz - 1);
  o[18] = -0.6690465 * y * z * (7 * zz - 3);
  o[19] = 0.1057855 * (zz * (35 * zz - 30) + 3);
  o[20] = -0.6690465 * x * z * (7 * zz - 3);
  o[21] = 0.4730873 * (xx - yy) * (7 * zz - 1);
  o[22] = -1.7701308 * x * z * (xx - 3 * yy);
  o[23] = 0.6258357 * (xx * (xx - 3 * yy) - yy * (3 * xx - yy));
}

/* sh.ts:362-373 */
void orc_sh_evaluate(const double* c, const double* dir, int degree, double* rgb) {
  double basis[24];
  orc_sh_basis(dir, basis);
  int deg = degree < 0 ? 0 : (degree > 4 ? 4 : degree);
  rgb[0] = rgb[1] = rgb[2] = 0;
  for (int k = 0; k < SH_COEFFS[deg]; k++)
    for (int ch = 0; ch < 3; ch++) rgb[ch] += c[3 * k + ch] * basis[k];
}

/* Sh.evaluate over a table (the CPU baseline's "+SH" column): for splat i,
 * rgb_i += evaluate(coefficients of i, normalize(centre_i - origin)); f32
 * planes of `stride` (centres 3, SH coefficient-major 3K), rgb 3 planes. */
void orc_sh_apply_table(size_t n, size_t stride, const float* center, const float* sh, int degree, const double* origin,
                        float* rgb) {
  const int K = SH_COEFFS[degree < 0 ? 0 : (degree > 3 ? 3 : degree)];
  double c[45], out[3], d[3];
  for (size_t i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) d[k] = (double)center[k * stride + i] - origin[k];
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (len > 0)
      for (int k = 0; k < 3; k++) d[k] /= len;
    for (int v = 0; v < 3 * K; v++) c[v] = sh[v * stride + i];
    orc_sh_evaluate(c, d, degree, out);
    for (int k = 0; k < 3; k++) rgb[k * stride + i] += (float)out[k];
  }
}

/* math.ts:29-45: Sigma = M M^T, M = R(q) diag(scales) (q normalised; zero: identity) */
void orc_cov_from_scale_quat(const double* s, const double* q, double* cov) {
  double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double x = 0, y = 0, z = 0, w = 1;
  if (len > 0) {
    x = q[0] / len;
    y = q[1] / len;
    z = q[2] / len;
    w = q[3] / len;
  }
  /* column k of R times scale k; rows r */
  double m[3][3] = {
      {s[0] * (1 - 2 * (y * y + z * z)), s[1] * 2 * (x * y - w * z), s[2] * 2 * (x * z + w * y)},
      {s[0] * 2 * (x * y + w * z), s[1] * (1 - 2 * (x * x + z * z)), s[2] * 2 * (y * z - w * x)},
      {s[0] * 2 * (x * z - w * y), s[1] * 2 * (y * z + w * x), s[2] * (1 - 2 * (x * x + y * y))},
  };
  double c[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c[i][j] = m[i][0] * m[j][0] + m[i][1] * m[j][1] + m[i][2] * m[j][2];
  cov[0] = c[0][0];
  cov[1] = c[1][1];
  cov[2] = c[2][2];
  cov[3] = c[0][1];
  cov[4] = c[0][2];
  cov[5] = c[1][2];
}

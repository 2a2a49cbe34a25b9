// Fused projection kernel for gfx950: decode a splat (f32 planes, ext or
// packed words, with their SH), add the SH colour, project it for one or
// two eyes with EWA, and write the reference's ewa words and its sort key,
// in one pass over the table.
//
// Reads are structure-of-arrays: lane i of a wave reads splat i of each
// plane (f32, as buffer loads with one 32-bit lane offset shared by every
// plane) or its uint32x4 words (ext, packed), so every load is one
// contiguous run per wave, and all of a splat's loads are issued before any
// is used. The table is read once, so its loads are nontemporal: over 59
// f32 planes that streams at ~6.0 TB/s against ~5.4 TB/s for default-policy
// loads (measured on 4M-splat planes). The eye-space centre and the key distance are computed in double
// (a dozen FMAs); the rest in float. The 2D determinant uses the closed form
//   det(J S' J^T) = (fx fy / z^3)^2 det(L)^2 (sy^2 sz^2 lx^2 + sx^2 sz^2 ly^2 + sx^2 sy^2 lz^2),
// l = R^T (c - eye) (Cauchy-Binet with J0 x J1 = fx fy c / z^3), a sum of
// squares that stays accurate for thin splats seen edge-on, where a*d - b^2
// cancels.
#pragma once

#include "common.hpp"
#include "sort_kernels.hpp"

namespace rdx {

constexpr int LAYOUT_F32 = 0, LAYOUT_EXT = 1, LAYOUT_PACKED = 2, LAYOUT_COV = 3;
constexpr int SHC_NONE = 0, SHC_F32 = 1, SHC_EXT = 2, SHC_PACKED = 3;
// f32 SH planes of a table too large for one 4 GiB window over them (read
// in groups of SH_GROUP coefficients, f32_rsrc)
constexpr int SHC_F32_WIDE = 4;
constexpr uint32_t BUFFER_DWORD3 = 0x00020000u;  // raw buffer resource, gfx9 family

struct EyeU {
  double L[9];      // object -> eye linear part, row-major
  double e[3];      // the eye in object space (eye-space centre = L (c - e))
  float Lf[9];
  float Pf[16];     // projection, row-major
  float fx, fy;     // focal length in pixels: size / 2 * (P00, P11)
  float det_l2;     // det(L)^2
  float ortho_l[3]; // L^-1 (0, 0, 1) (ortho only)
  int ortho;
};

struct ProjU {
  EyeU eye[2];
  int eyes;
  int head_is_eye0;   // mono: the head is eye 0 (the key distance is eye 0's)
  double h[3];        // the head in object space: SH origin
  double H[9];        // object -> head linear part (head-space centre = H (c - h))
  float size[2];
  float ndc_z_min;
  float max_std_dev, min_alpha, blur, pre_blur, clip_xy, min_pr, max_pr, min_contrib;
  float focal_distance, aperture_tan, sort_guard;
  int lod_inflate, sh_degree, sort_bits, cull_keys, key_per_eye;
  float enc[4];       // packed encoding: rgbMin, rgbMax, lnScaleMin, lnScaleMax
  // object -> head is a rotation times a uniform scale (isUniformScale,
  // splatTransform.tsl.ts:46-62), and that scale (cbrt |det|)
  int head_uniform;
  float head_scale;
  int shared_axes;  // eye 1's Lf equals eye 0's (a rig whose eyes differ by a translation)
};

struct TableP {
  uint32_t count;
  uint32_t stride;         // floats per plane
  const float* center;     // 3 planes
  const float* log_scale;  // 3 planes
  const float* quat;       // 4 planes
  const float* opacity;
  const float* rgb;        // 3 planes
  const float* sh;         // 3K planes
  const uint4* ext_a;
  const uint4* ext_b;
  const uint4* packed;
  const uint4* shc[4];
};

struct OutP {
  uint4* words[2];
  uint32_t* keys;
  size_t key_stride;
  uint32_t at;
  // side outputs (float planes of `sstride`), any may be NULL
  size_t sstride;
  float* conic;
  float* cov2d;
  float* opac;
  float* ndc;
  float* vz;
  float* bound;
  float* rgb;
  float* dist;
};

// ---- codecs (src/gpu/encoding.tsl.ts, splats.tsl.ts, sh.tsl.ts) --------------

// 2^(e - 15) / 255 from the exponent field: exact (encoding.tsl.ts:36-44)
__device__ __forceinline__ float3 decode_ext_rgb(uint32_t e) {
  const float unit = bitsf(((e >> 27) + 112u) << 23) * (1.0f / 255.0f);
  float r = (float)(e & 0xffu) * unit, g = (float)((e >> 8) & 0xffu) * unit, b = (float)((e >> 16) & 0xffu) * unit;
  return make_float3((e & 0x1000000u) ? -r : r, (e & 0x2000000u) ? -g : g, (e & 0x4000000u) ? -b : b);
}

// 8-bit magnitudes, 3 signs, 5-bit shared exponent (encoding.ts:219-231);
// ceil(log2(max)) from the float's bits, exactly.
__device__ __forceinline__ uint32_t encode_ext_rgb(float r, float g, float b) {
  const float ar = fabsf(r), ag = fabsf(g), ab = fabsf(b);
  const float mx = fmaxf(ar, fmaxf(ag, ab));
  int cl = 0;
  if (mx > 0.0f) {
    const uint32_t u = fbits(mx);
    const int ex = (int)(u >> 23);
    cl = ex == 0 ? -126 : ex - 127 + ((u & 0x7fffffu) ? 1 : 0);  // subnormal: the base clamps to 0 anyway
  }
  const int base = min(max(cl + 15, 0), 31);
  // 255 * 2^(15 - base) from the exponent field (142 - base): exact
  const float scale = bitsf((uint32_t)(142 - base) << 23) * 255.0f;
  // Math.round(|v| * scale) of the exact product: the float product p is off
  // by e = fma(|v|, scale, -p), |e| <= ulp(p) / 2; its fraction f = p - floor(p)
  // is exact and on p's grid, as is 0.5, so only f == 0.5 can be decided by e.
  auto mag = [&](float v) {
    const float p = v * scale;
    const float e = fmaf(v, scale, -p);
    float rr = floorf(p);
    const float f = p - rr;
    if (f > 0.5f || (f == 0.5f && e >= 0.0f)) rr += 1.0f;
    return (uint32_t)fminf(rr, 255.0f);
  };
  const uint32_t mr = mag(ar), mg = mag(ag), mb = mag(ab);
  const uint32_t signs = (r < 0.0f ? 1u : 0u) | (g < 0.0f ? 2u : 0u) | (b < 0.0f ? 4u : 0u);
  return mr | (mg << 8) | (mb << 16) | ((((uint32_t)base << 3) | signs) << 24);
}

__device__ __forceinline__ float3 oct_decode(uint32_t c, int bits) {
  if (c == 0u) return make_float3(0.0f, 0.0f, 0.0f);
  const uint32_t mx = (1u << bits) - 1u;
  const float inv = 1.0f / (float)mx;
  float x = (float)(c & mx) * inv * 2.0f - 1.0f;
  float y = (float)((c >> bits) & mx) * inv * 2.0f - 1.0f;
  const float z = 1.0f - fabsf(x) - fabsf(y);
  const float t = fmaxf(-z, 0.0f);
  x += x >= 0.0f ? -t : t;
  y += y >= 0.0f ? -t : t;
  const float rl = rsqrtf(x * x + y * y + z * z);
  return make_float3(x * rl, y * rl, z * rl);
}

// axis octahedral low, angle above (encoding.tsl.ts:113-119)
__device__ __forceinline__ float4 quat_oct_decode(uint32_t code, int axis_bits, int angle_bits) {
  const float3 axis = oct_decode(code & ((1u << (2 * axis_bits)) - 1u), axis_bits);
  const uint32_t amax = (1u << angle_bits) - 1u;
  const float half = (float)((code >> (2 * axis_bits)) & amax) / (float)amax * 1.5707963267948966f;
  float s, c;
  __sincosf(half, &s, &c);
  return make_float4(axis.x * s, axis.y * s, axis.z * s, c);
}

// ---- word codecs (src/render/projection.tsl.ts:277-310) ---------------------

// atan2 for the 12-bit angle code (steps of pi / 4096 = 7.7e-4 rad): the
// octant reduction a = min / max in [0, 1], atan(a) = a P(a^2) with P of
// degree 7 fitted here to atan(a) / a on [0, 1] (weighted least squares
// toward the minimax error; 1.3e-7 rad evaluated in float), then the
// octant's reflections. Arguments not both zero.
__device__ __forceinline__ float atan2_code(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float a = fminf(ax, ay) * __builtin_amdgcn_rcpf(fmaxf(ax, ay));
  const float s = a * a;
  float p = -0.004054663702845573f;
  p = fmaf(p, s, 0.02186330407857895f);
  p = fmaf(p, s, -0.055912818759679794f);
  p = fmaf(p, s, 0.09642232209444046f);
  p = fmaf(p, s, -0.13908642530441284f);
  p = fmaf(p, s, 0.19946567714214325f);
  p = fmaf(p, s, -0.33329862356185913f);
  p = fmaf(p, s, 0.9999993443489075f);
  float r = a * p;
  if (ay > ax) r = 1.57079632679489662f - r;
  if (x < 0.0f) r = 3.14159265358979324f - r;
  return y < 0.0f ? -r : r;
}

__device__ __forceinline__ uint32_t encode_fixed16(float v) {
  return (uint32_t)fminf(fmaxf(floorf((v + 2.0f) * 0.25f * 65535.0f + 0.5f), 0.0f), 65535.0f);
}
__device__ __forceinline__ uint32_t encode_depth20(float view_z) {
  const uint32_t lo = 113u << 23, hi = (145u << 23) - 1u;
  const uint32_t f = min(max(fbits(-view_z), lo), hi);
  return min((f - lo + 128u) >> 8, 0xfffffu);
}
__device__ __forceinline__ uint32_t encode_angle12(float angle) {
  return (uint32_t)floorf(angle * (float)(1.0 / 3.14159265358979323846) * 4096.0f + 0.5f) & 0xfffu;
}
__device__ __forceinline__ uint32_t encode_log_scale12(float scale) {
  const float top = 4094.0f;
  // v_log_f32 (1 ulp) for the 12-bit code's 1/256-octave steps
  const float l = fminf(fmaxf(__builtin_amdgcn_logf(fmaxf(scale, 1e-20f)), -4.0f), 12.0f);
  const float n = (l + 4.0f) * (1.0f / 16.0f) * top;
  return (uint32_t)fminf(fmaxf(floorf(n + 0.5f), 0.0f), top) + 1u;
}
__device__ __forceinline__ uint32_t encode_opacity8(float o) {
  const uint32_t c = (uint32_t)floorf(fminf(fmaxf(o * 128.0f, 0.0f), 255.0f) + 0.5f);
  return max(c, 1u);
}

// ---- SH (src/gpu/sh.tsl.ts:60-139; basis constants of src/core/sh.ts:381-405) ----

// The real SH basis of bands 1..deg along unit d, w[k] for coefficient k.
__device__ __forceinline__ void sh_basis(float3 d, int deg, float w[15]) {
  const float x = d.x, y = d.y, z = d.z;
  if (deg >= 1) {
    w[0] = -0.4886025f * y;
    w[1] = 0.4886025f * z;
    w[2] = -0.4886025f * x;
    if (deg >= 2) {
      const float xx = x * x, yy = y * y, zz = z * z;
      w[3] = 1.0925484f * x * y;
      w[4] = -1.0925484f * y * z;
      w[5] = 0.3153916f * (2.0f * zz - xx - yy);
      w[6] = -1.0925484f * x * z;
      w[7] = 0.5462742f * (xx - yy);
      if (deg >= 3) {
        w[8] = -0.5900436f * y * (3.0f * xx - yy);
        w[9] = 2.8906114f * x * y * z;
        w[10] = -0.4570458f * y * (4.0f * zz - xx - yy);
        w[11] = 0.3731763f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
        w[12] = -0.4570458f * x * (4.0f * zz - xx - yy);
        w[13] = 1.4453057f * z * (xx - yy);
        w[14] = -0.5900436f * x * (xx - 3.0f * yy);
      }
    }
  }
}

// Bands 1..deg of coefficients coef[k] (float3) along unit d. With `band`
// (packed SH: coef[k] holds the band's integer codes), each band's sum is
// scaled by its band[b] once instead of every coefficient.
__device__ __forceinline__ float3 sh_eval(float3 d, int deg, const float3* coef, const float* band = nullptr) {
  float w[15];
  sh_basis(d, deg, w);
  const int K = deg >= 3 ? 15 : (deg == 2 ? 8 : (deg == 1 ? 3 : 0));
  float3 s = make_float3(0.0f, 0.0f, 0.0f);
  if (!band) {
#pragma unroll
    for (int k = 0; k < 15; k++) {
      if (k < K) {
        s.x = fmaf(coef[k].x, w[k], s.x);
        s.y = fmaf(coef[k].y, w[k], s.y);
        s.z = fmaf(coef[k].z, w[k], s.z);
      }
    }
    return s;
  }
#pragma unroll
  for (int b = 0; b < 3; b++) {
    const int k0 = b == 0 ? 0 : (b == 1 ? 3 : 8), k1 = b == 0 ? 3 : (b == 1 ? 8 : 15);
    if (k0 < K) {
      float3 t = make_float3(0.0f, 0.0f, 0.0f);
#pragma unroll
      for (int k = k0; k < k1; k++) {
        t.x = fmaf(coef[k].x, w[k], t.x);
        t.y = fmaf(coef[k].y, w[k], t.y);
        t.z = fmaf(coef[k].z, w[k], t.z);
      }
      s.x = fmaf(t.x, band[b], s.x);
      s.y = fmaf(t.y, band[b], s.y);
      s.z = fmaf(t.z, band[b], s.z);
    }
  }
  return s;
}

// Packed SH bytes: scale code at the band's first byte, then sint8 codes
// (src/core/sh.ts:137-174; sh.tsl.ts:40-54)
__device__ __forceinline__ uint32_t sh_byte(const uint32_t* w, int b) { return (w[b >> 2] >> ((b & 3) * 8)) & 0xffu; }
__device__ __forceinline__ float sh_scale(uint32_t code) {
  return code == 0u ? 0.0f : exp2f(((float)code - 192.0f) * (1.0f / 16.0f)) * (1.0f / 127.0f);
}

struct Splat {
  float3 c;      // object-space centre (as stored, float)
  float3 s;      // scales
  float4 q;      // quaternion (any length)
  float3 rgb;    // base colour
  float op;
  // covSplats: the object-space covariance xx, yy, zz, xy, xz, yz (the cov
  // form), decoded in double: the 2D covariance's determinant of a thin
  // splat cancels, and float entries would leave it off by the 2D
  // condition number times 2^-24
  double cv[6];
};

// splats.ts:230-232: a correlation stored as sign(c) ln|c|
__device__ __forceinline__ double decode_correlation(double e) { return e == 0.0 ? 0.0 : (e > 0.0 ? -exp(-e) : exp(e)); }

// Buffer resources of an f32 table's plane groups (one 32-bit lane offset
// i * 4 serves every plane; a plane's start is a scalar offset). A resource
// spans at most 4 GiB: one resource serves all SH planes of a table up to
// f32_sh_window splats per plane; larger tables (SHC_F32_WIDE) read them
// through three resources of SH_GROUP coefficients (15 planes) each, up to
// F32_MAX_STRIDE splats per plane.
constexpr int SH_GROUP = 5;
constexpr uint64_t F32_MAX_STRIDE = 0xffffffffull / (4ull * 3 * SH_GROUP);
__host__ __device__ constexpr int sh_coefs(int deg) { return deg >= 3 ? 15 : (deg == 2 ? 8 : (deg == 1 ? 3 : 0)); }
__host__ __device__ constexpr bool f32_sh_wide(uint64_t stride, int deg) {
  return stride * 4ull * 3ull * (uint64_t)sh_coefs(deg) > 0xffffffffull;
}
struct F32Rsrc {
  __amdgpu_buffer_rsrc_t c, l, q, o, g, h[3];
  uint32_t st4;  // bytes per plane
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* p, uint32_t planes, uint32_t st4) {
  const uint64_t bytes = (uint64_t)planes * st4;
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(bytes > 0xffffffffull ? 0xffffffffu : (uint32_t)bytes),
                                           (int)BUFFER_DWORD3);
}

template <int SHC>
__device__ __forceinline__ F32Rsrc f32_rsrc(const TableP& t, int deg) {
  constexpr int SHG = SHC == SHC_F32_WIDE ? SH_GROUP : 15;
  F32Rsrc r;
  r.st4 = t.stride * 4u;
  r.c = plane_rsrc(t.center, 3, r.st4);
  r.l = plane_rsrc(t.log_scale, 3, r.st4);
  r.q = plane_rsrc(t.quat, 4, r.st4);
  r.o = plane_rsrc(t.opacity, 1, r.st4);
  r.g = plane_rsrc(t.rgb, 3, r.st4);
  const int K = deg >= 3 ? 15 : (deg == 2 ? 8 : (deg == 1 ? 3 : 0));
  const float* sh = t.sh ? t.sh : t.rgb;
#pragma unroll
  for (int g = 0; g < 3; g++) {
    const int k0 = g * SHG, k1 = min(K, k0 + SHG);
    r.h[g] = plane_rsrc(sh + (size_t)(3 * k0) * t.stride, k1 > k0 ? (uint32_t)(3 * (k1 - k0)) : 0u, r.st4);
  }
  return r;
}

// Cache-policy operand of a buffer load: nt (nontemporal, streaming).
constexpr int CPOL_NT = 2;

__device__ __forceinline__ float ldp(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, CPOL_NT));
}

// A table word read once: a nontemporal 16-byte load.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 ld_stream(const uint4* p) {
  const u32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const u32x4_t*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
// Output words nothing on the path reads back: a nontemporal 16-byte store.
__device__ __forceinline__ void st_stream(uint4* p, uint4 w) {
  u32x4_t v;
  v.x = w.x;
  v.y = w.y;
  v.z = w.z;
  v.w = w.w;
  __builtin_nontemporal_store(v, reinterpret_cast<u32x4_t*>(p));
}

// All of splat i: geometry, base colour and SH bands 1..deg into coef.
// DEG >= 0: a compile-time degree (f32 planes); DEG < 0: the runtime `deg`.
template <int LAYOUT, int SHC, int DEG>
__device__ __forceinline__ void load_splat(const TableP& t, const F32Rsrc& rs, const float* enc, uint32_t i, int deg,
                                           Splat& sp, float3* coef, float* band) {
  if constexpr (LAYOUT == LAYOUT_F32) {
    constexpr int K = DEG >= 3 ? 15 : (DEG == 2 ? 8 : (DEG == 1 ? 3 : 0));
    const uint32_t v = i * 4u, s = rs.st4;
    sp.c = make_float3(ldp(rs.c, v, 0), ldp(rs.c, v, s), ldp(rs.c, v, 2 * s));
    const float l0 = ldp(rs.l, v, 0), l1 = ldp(rs.l, v, s), l2 = ldp(rs.l, v, 2 * s);
    sp.q = make_float4(ldp(rs.q, v, 0), ldp(rs.q, v, s), ldp(rs.q, v, 2 * s), ldp(rs.q, v, 3 * s));
    sp.op = ldp(rs.o, v, 0);
    sp.rgb = make_float3(ldp(rs.g, v, 0), ldp(rs.g, v, s), ldp(rs.g, v, 2 * s));
#pragma unroll
    for (int k = 0; k < K; k++) {
      constexpr int SHG = SHC == SHC_F32_WIDE ? SH_GROUP : 15;
      const __amdgpu_buffer_rsrc_t h = rs.h[k / SHG];
      const uint32_t p = (uint32_t)(3 * (k % SHG));
      coef[k] = make_float3(ldp(h, v, p * s), ldp(h, v, (p + 1) * s), ldp(h, v, (p + 2) * s));
    }
    sp.s = make_float3(__expf(l0), __expf(l1), __expf(l2));
  } else {
    // uint32x4 words: issue the SH texels' loads with the splat's
    uint4 tex[4];
    if constexpr (SHC == SHC_EXT) {
      const int texels = deg <= 0 ? 0 : (deg == 1 ? 1 : (deg == 2 ? 2 : 4));
#pragma unroll
      for (int x = 0; x < 4; x++)
        if (x < texels) tex[x] = ld_stream(t.shc[x] + i);
    } else if constexpr (SHC == SHC_PACKED) {
      const int texels = deg <= 0 ? 0 : deg;
#pragma unroll
      for (int x = 0; x < 3; x++)
        if (x < texels) tex[x] = ld_stream(t.shc[x] + i);
    }
    if constexpr (LAYOUT == LAYOUT_EXT) {
      const uint4 A = ld_stream(t.ext_a + i), B = ld_stream(t.ext_b + i);
      sp.c = make_float3(bitsf(A.x), bitsf(A.y), bitsf(A.z));
      sp.op = (float)(A.w >> 24) * (1.0f / 128.0f);
      sp.rgb = decode_ext_rgb(B.x);
      sp.s = make_float3(__expf(half_to_float(B.y)), __expf(half_to_float(B.y >> 16)), __expf(half_to_float(B.z)));
      sp.q = quat_oct_decode(B.w, 10, 12);
    } else if constexpr (LAYOUT == LAYOUT_COV) {
      // covSplats (splats.tsl.ts:105-121): A as ext, B the covariance
      const uint4 A = ld_stream(t.ext_a + i), B = ld_stream(t.ext_b + i);
      sp.c = make_float3(bitsf(A.x), bitsf(A.y), bitsf(A.z));
      sp.op = (float)(A.w >> 24) * (1.0f / 128.0f);
      sp.rgb = decode_ext_rgb(B.x);
      const double xx = exp((double)half_to_float(B.y)), yy = exp((double)half_to_float(B.y >> 16)),
                   zz = exp((double)half_to_float(B.z));
      sp.cv[0] = xx;
      sp.cv[1] = yy;
      sp.cv[2] = zz;
      sp.cv[3] = decode_correlation(half_to_float(B.z >> 16)) * sqrt(xx * yy);
      sp.cv[4] = decode_correlation(half_to_float(B.w)) * sqrt(xx * zz);
      sp.cv[5] = decode_correlation(half_to_float(B.w >> 16)) * sqrt(yy * zz);
      sp.s = make_float3(0.0f, 0.0f, 0.0f);
      sp.q = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    } else {
      const uint4 W = ld_stream(t.packed + i);
      const float lo = enc[0], range = enc[1] - enc[0];
      sp.rgb = make_float3(lo + (float)(W.x & 0xffu) * (1.0f / 255.0f) * range,
                           lo + (float)((W.x >> 8) & 0xffu) * (1.0f / 255.0f) * range,
                           lo + (float)((W.x >> 16) & 0xffu) * (1.0f / 255.0f) * range);
      sp.op = (float)(W.x >> 24) * (1.0f / 255.0f);
      sp.c = make_float3(half_to_float(W.y), half_to_float(W.y >> 16), half_to_float(W.z));
      const float step = (enc[3] - enc[2]) * (1.0f / 254.0f);
      auto sc = [&](uint32_t code) { return code == 0u ? 0.0f : __expf(enc[2] + ((float)code - 1.0f) * step); };
      sp.s = make_float3(sc(W.w & 0xffu), sc((W.w >> 8) & 0xffu), sc((W.w >> 16) & 0xffu));
      sp.q = quat_oct_decode((W.z >> 16) | ((W.w >> 24) << 16), 8, 8);
    }
    if constexpr (SHC == SHC_EXT) {
      const int K = deg >= 3 ? 15 : (deg == 2 ? 8 : (deg == 1 ? 3 : 0));
#pragma unroll
      for (int k = 0; k < 15; k++) {
        if (k < K) {
          const uint4 v = tex[k >> 2];
          const uint32_t w = (k & 3) == 0 ? v.x : ((k & 3) == 1 ? v.y : ((k & 3) == 2 ? v.z : v.w));
          coef[k] = decode_ext_rgb(w);
        }
      }
    } else if constexpr (SHC == SHC_PACKED) {
      if (deg > 0) {
        uint32_t w[12];
#pragma unroll
        for (int x = 0; x < 3; x++) {
          const bool have = x < deg;
          w[4 * x] = have ? tex[x].x : 0u;
          w[4 * x + 1] = have ? tex[x].y : 0u;
          w[4 * x + 2] = have ? tex[x].z : 0u;
          w[4 * x + 3] = have ? tex[x].w : 0u;
        }
        // the codes here, each band's scale applied to its sum (sh_eval)
        band[0] = sh_scale(sh_byte(w, 0));
        band[1] = deg >= 2 ? sh_scale(sh_byte(w, 10)) : 0.0f;
        band[2] = deg >= 3 ? sh_scale(sh_byte(w, 26)) : 0.0f;
#pragma unroll
        for (int k = 0; k < 15; k++) {
          const int b = k < 3 ? 0 : (k < 8 ? 1 : 2);
          const int start = b == 0 ? 0 : (b == 1 ? 10 : 26);
          const int first = b == 0 ? 0 : (b == 1 ? 3 : 8);
          const int at = start + 1 + (k - first) * 3;
          coef[k] = make_float3((float)(int8_t)sh_byte(w, at), (float)(int8_t)sh_byte(w, at + 1),
                                (float)(int8_t)sh_byte(w, at + 2));
        }
      }
    }
  }
}

// base + SH(normalize(c - origin)), clamped at 0 (sh.tsl.ts:255-266); band:
// the packed bands' scales (NULL for the other codecs)
__device__ __forceinline__ float3 sh_colour(const Splat& sp, const double* origin, int deg, const float3* coef,
                                            const float* band = nullptr) {
  if (deg <= 0) return sp.rgb;
  const float dx = (float)((double)sp.c.x - origin[0]);
  const float dy = (float)((double)sp.c.y - origin[1]);
  const float dz = (float)((double)sp.c.z - origin[2]);
  const float rl = rsqrtf(dx * dx + dy * dy + dz * dz);
  const float3 s = sh_eval(make_float3(dx * rl, dy * rl, dz * rl), deg, coef, band);
  return make_float3(fmaxf(sp.rgb.x + s.x, 0.0f), fmaxf(sp.rgb.y + s.y, 0.0f), fmaxf(sp.rgb.z + s.z, 0.0f));
}

// Rotation of a normalised quaternion, row-major (math.ts:29-45, math.tsl.ts:21-40)
__device__ __forceinline__ void quat_matrix(float4 q, float R[9]) {
  const float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  float x = 0.0f, y = 0.0f, z = 0.0f, w = 1.0f;
  if (n2 > 0.0f) {
    const float r = rsqrtf(n2);
    x = q.x * r;
    y = q.y * r;
    z = q.z * r;
    w = q.w * r;
  }
  R[0] = 1.0f - 2.0f * (y * y + z * z);
  R[1] = 2.0f * (x * y - w * z);
  R[2] = 2.0f * (x * z + w * y);
  R[3] = 2.0f * (x * y + w * z);
  R[4] = 1.0f - 2.0f * (x * x + z * z);
  R[5] = 2.0f * (y * z - w * x);
  R[6] = 2.0f * (x * z - w * y);
  R[7] = 2.0f * (y * z + w * x);
  R[8] = 1.0f - 2.0f * (x * x + y * y);
}

// A splat's centre seen from one eye: object-space offset from the eye, the
// eye-space centre (double, and its distance), clip coordinates, and the
// centre culls (cullDepth, cullClipXY, projection.tsl.ts:132-140).
struct EyeCentre {
  double ce[3];
  double dist;
  float3 d;      // c - eye (object space)
  float clip[4];
  bool pass;
};

__device__ __forceinline__ EyeCentre eye_centre(const EyeU& E, float3 c, float ndc_z_min, float clip_xy) {
  EyeCentre r;
  const double dx = (double)c.x - E.e[0], dy = (double)c.y - E.e[1], dz = (double)c.z - E.e[2];
#pragma unroll
  for (int i = 0; i < 3; i++) r.ce[i] = fma(E.L[3 * i], dx, fma(E.L[3 * i + 1], dy, E.L[3 * i + 2] * dz));
  r.dist = sqrt(fma(r.ce[0], r.ce[0], fma(r.ce[1], r.ce[1], r.ce[2] * r.ce[2])));
  r.d = make_float3((float)dx, (float)dy, (float)dz);
  const float x = (float)r.ce[0], y = (float)r.ce[1], z = (float)r.ce[2];
#pragma unroll
  for (int i = 0; i < 4; i++) r.clip[i] = fmaf(E.Pf[4 * i], x, fmaf(E.Pf[4 * i + 1], y, fmaf(E.Pf[4 * i + 2], z, E.Pf[4 * i + 3])));
  const bool cull_depth = z >= 0.0f || r.clip[2] < ndc_z_min * r.clip[3] || r.clip[2] >= r.clip[3];
  const float limit = fminf(clip_xy, 2.0f) * r.clip[3];
  const bool cull_xy = fabsf(r.clip[0]) > limit || fabsf(r.clip[1]) > limit;
  r.pass = !(cull_depth || cull_xy);
  return r;
}

struct EwaResult {
  bool ok;
  float a, b, d, det, opc, sc1, sc2, angle, z;
  float2 ndc;
  uint4 words;
};

// Q = L R, the splat's axes in eye space (projections.tsl.ts:104-110)
__device__ __forceinline__ void eye_axes(const EyeU& E, const float* R, float Q[9]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      Q[3 * i + j] = fmaf(E.Lf[3 * i], R[j], fmaf(E.Lf[3 * i + 1], R[3 + j], E.Lf[3 * i + 2] * R[6 + j]));
}

// An eye's 2D covariance before the blur (JΣJ^T, projections.tsl.ts:111-118)
// and its determinant.
struct Cov2 {
  float a, b, d, det;
};

// The splat form: J rows (projections.tsl.ts:111-117), T = (J Q) diag(s),
// cov2D = T T^T, and the determinant by Cauchy-Binet (no cancellation): R is
// the rotation, Q = eye_axes(E, R), s the (inflated) scales.
__device__ __forceinline__ Cov2 cov2d_splat(const EyeU& E, const EyeCentre& C, const float* R, const float* Q,
                                            float3 s) {
  const float x = (float)C.ce[0], y = (float)C.ce[1], z = (float)C.ce[2];
  float j00, j02, j11, j12, kdet, l0, l1, l2;
  if (E.ortho) {
    j00 = E.fx;
    j02 = 0.0f;
    j11 = E.fy;
    j12 = 0.0f;
    kdet = E.fx * E.fy;
    l0 = R[0] * E.ortho_l[0] + R[3] * E.ortho_l[1] + R[6] * E.ortho_l[2];
    l1 = R[1] * E.ortho_l[0] + R[4] * E.ortho_l[1] + R[7] * E.ortho_l[2];
    l2 = R[2] * E.ortho_l[0] + R[5] * E.ortho_l[1] + R[8] * E.ortho_l[2];
  } else {
    const float iz = __builtin_amdgcn_rcpf(z);
    j00 = E.fx * iz;
    j11 = E.fy * iz;
    j02 = -j00 * x * iz;
    j12 = -j11 * y * iz;
    kdet = j00 * j11 * iz;  // fx fy / z^3
    l0 = R[0] * C.d.x + R[3] * C.d.y + R[6] * C.d.z;
    l1 = R[1] * C.d.x + R[4] * C.d.y + R[7] * C.d.z;
    l2 = R[2] * C.d.x + R[5] * C.d.y + R[8] * C.d.z;
  }
  const float t00 = (j00 * Q[0] + j02 * Q[6]) * s.x, t01 = (j00 * Q[1] + j02 * Q[7]) * s.y,
              t02 = (j00 * Q[2] + j02 * Q[8]) * s.z;
  const float t10 = (j11 * Q[3] + j12 * Q[6]) * s.x, t11 = (j11 * Q[4] + j12 * Q[7]) * s.y,
              t12 = (j11 * Q[5] + j12 * Q[8]) * s.z;
  Cov2 c;
  c.a = t00 * t00 + t01 * t01 + t02 * t02;
  c.d = t10 * t10 + t11 * t11 + t12 * t12;
  c.b = t00 * t10 + t01 * t11 + t02 * t12;
  const float sxx = s.x * s.x, syy = s.y * s.y, szz = s.z * s.z;
  c.det = kdet * kdet * E.det_l2 * (syy * szz * l0 * l0 + sxx * szz * l1 * l1 + sxx * syy * l2 * l2);
  return c;
}

// The cov form (covSplats): K = J L (L: object to eye), cov2D = K Σ K^T for
// the object-space covariance cv (xx, yy, zz, xy, xz, yz, already
// inflated) and its determinant a d - b^2, in double (the determinant of a
// thin splat's 2D covariance cancels).
__device__ __forceinline__ Cov2 cov2d_cov(const EyeU& E, const EyeCentre& C, const double* cv) {
  const float x = (float)C.ce[0], y = (float)C.ce[1], z = (float)C.ce[2];
  float j00, j02, j11, j12;
  if (E.ortho) {
    j00 = E.fx;
    j02 = 0.0f;
    j11 = E.fy;
    j12 = 0.0f;
  } else {
    const float iz = __builtin_amdgcn_rcpf(z);
    j00 = E.fx * iz;
    j11 = E.fy * iz;
    j02 = -j00 * x * iz;
    j12 = -j11 * y * iz;
  }
  const float* L = E.Lf;
  double k0[3], k1[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    k0[j] = (double)j00 * L[j] + (double)j02 * L[6 + j];
    k1[j] = (double)j11 * L[3 + j] + (double)j12 * L[6 + j];
  }
  // Σ k0 and Σ k1
  const double u0 = cv[0] * k0[0] + cv[3] * k0[1] + cv[4] * k0[2];
  const double u1 = cv[3] * k0[0] + cv[1] * k0[1] + cv[5] * k0[2];
  const double u2 = cv[4] * k0[0] + cv[5] * k0[1] + cv[2] * k0[2];
  const double w0 = cv[0] * k1[0] + cv[3] * k1[1] + cv[4] * k1[2];
  const double w1 = cv[3] * k1[0] + cv[1] * k1[1] + cv[5] * k1[2];
  const double w2 = cv[4] * k1[0] + cv[5] * k1[1] + cv[2] * k1[2];
  const double a = k0[0] * u0 + k0[1] * u1 + k0[2] * u2;
  const double b = k1[0] * u0 + k1[1] * u1 + k1[2] * u2;
  const double d = k1[0] * w0 + k1[1] * w1 + k1[2] * w2;
  Cov2 c;
  c.a = (float)a;
  c.b = (float)b;
  c.d = (float)d;
  c.det = (float)(a * d - b * b);
  return c;
}

// The ewa method for one eye after its 2D covariance (projections.tsl.ts:118-148
// with boundFromCov, projection.tsl.ts:215-240, and the word codecs): `pass`
// is the eye's centre culls, rgb_word the colour's word (encode_ext_rgb, the
// same for every eye of one SH origin). U supplies the params (min_alpha,
// blur, ...).
template <typename U>
__device__ __forceinline__ EwaResult ewa_words(const EyeU& E, const EyeCentre& C, bool pass, Cov2 c0, float op,
                                               float std_dev, uint32_t rgb_word, const U& u) {
  EwaResult r;
  const float z = (float)C.ce[2];
  r.z = z;
  const float a0 = c0.a, b0 = c0.b, d0 = c0.d, det0 = c0.det;
  // blur and opacity compensation (projections.tsl.ts:119-129), depth of field (projection.tsl.ts:118-126)
  float blur = u.blur;
  if (u.focal_distance > 0.0f && u.aperture_tan > 0.0f && z < 0.0f) {
    const float depth = -z;
    const float coc_view = fabsf(depth - u.focal_distance) * u.aperture_tan;
    const float focal_pixels = E.fy;  // renderSize.y / 2 * P11
    const float coc = E.ortho ? coc_view * focal_pixels : coc_view * focal_pixels * __builtin_amdgcn_rcpf(depth);
    blur = fminf(fmaxf(coc * coc, blur), u.max_pr * u.max_pr);
  }
  const float pre = u.pre_blur;
  const float ap = a0 + pre, dp = d0 + pre;
  const float det_orig = det0 + pre * (a0 + d0) + pre * pre;
  r.a = ap + blur;
  r.d = dp + blur;
  r.b = b0;
  r.det = det_orig + blur * (ap + dp) + blur * blur;
  // v_sqrt_f32 (1 ulp) throughout: the outputs are 8- to 16-bit codes
  r.opc = op * __builtin_amdgcn_sqrtf(fmaxf(0.0f, det_orig * __builtin_amdgcn_rcpf(r.det)));
  const float contribution = r.opc * 6.283185307179586f * __builtin_amdgcn_sqrtf(fmaxf(0.0f, r.det));
  // boundFromCov (projection.tsl.ts:223-240)
  const float half = (r.a - r.d) * 0.5f;
  const float eig1 = (r.a + r.d) * 0.5f + __builtin_amdgcn_sqrtf(half * half + r.b * r.b);
  const float eig2 = eig1 > 0.0f ? fmaxf(r.det * __builtin_amdgcn_rcpf(eig1), 0.0f) : 0.0f;
  r.sc1 = fminf(u.max_pr, std_dev * __builtin_amdgcn_sqrtf(fmaxf(eig1, 0.0f)));
  r.sc2 = fminf(u.max_pr, std_dev * __builtin_amdgcn_sqrtf(eig2));
  const bool bound_ok = r.sc1 >= u.min_pr || r.sc2 >= u.min_pr;
  {
    float vx, vy;
    if (fabsf(r.a - eig1) >= fabsf(r.d - eig1)) {
      vx = r.b;
      vy = eig1 - r.a;
    } else {
      vx = eig1 - r.d;
      vy = r.b;
    }
    const float len2 = vx * vx + vy * vy;
    if (len2 > 0.0f) {
      const float rl = rsqrtf(len2);
      vx *= rl;
      vy *= rl;
    } else {
      vx = 1.0f;
      vy = 0.0f;
    }
    r.angle = atan2_code(vy, vx);
    if (r.angle < 0.0f) r.angle += 3.14159265358979323846f;
  }
  r.ok = pass && r.opc >= u.min_alpha && contribution >= u.min_contrib && bound_ok;
  r.ndc = make_float2(0.0f, 0.0f);
  if (C.clip[3] > 0.0f) {
    const float iw = __builtin_amdgcn_rcpf(C.clip[3]);
    r.ndc = make_float2(C.clip[0] * iw, C.clip[1] * iw);
  }
  r.words = make_uint4(0u, 0u, 0u, 0u);
  if (r.ok) {
    const float stored = r.opc > 1.0f ? (r.opc + 3.0f) * 0.25f : r.opc;
    r.words = make_uint4(encode_fixed16(r.ndc.x) | (encode_fixed16(r.ndc.y) << 16),
                         encode_depth20(z) | (encode_angle12(r.angle) << 20), rgb_word,
                         encode_opacity8(stored) | (encode_log_scale12(r.sc1) << 8) | (encode_log_scale12(r.sc2) << 20));
  }
  return r;
}

// The ewa method for one eye of a splat (scales s, rotation R, Q = eye_axes(E, R))
template <typename U>
__device__ __forceinline__ EwaResult ewa_eye(const EyeU& E, const EyeCentre& C, bool pass, const float* R,
                                             const float* Q, float3 s, float op, float std_dev, uint32_t rgb_word,
                                             const U& u) {
  return ewa_words(E, C, pass, cov2d_splat(E, C, R, Q, s), op, std_dev, rgb_word, u);
}

// The ewa method for one eye of a covariance (the cov form, cv inflated)
template <typename U>
__device__ __forceinline__ EwaResult ewa_eye_cov(const EyeU& E, const EyeCentre& C, bool pass, const double* cv, float op,
                                                 float std_dev, uint32_t rgb_word, const U& u) {
  return ewa_words(E, C, pass, cov2d_cov(E, C, cv), op, std_dev, rgb_word, u);
}

// prepareSplat (projection.tsl.ts:42-64) and cullInvisible (66-69). COVF:
// the cov form (sp.cv grows; s is the inflation), else the splat form (s the
// inflated scales).
template <bool COVF, typename U>
__device__ __forceinline__ void prepare(Splat& sp, const U& u, float& op, float3& s, float& std_dev, bool& invisible,
                                        float& inflate) {
  op = sp.op;
  inflate = 1.0f;
  if (op > 1.0f) {
    op = fminf(op * 4.0f - 3.0f, 5.0f);
    if (u.lod_inflate) {
      inflate = cbrtf(__expf((op * op - 1.0f) * 0.36787944117144233f));
      op = 1.0f;
    }
  }
  std_dev = op > 1.0f ? u.max_std_dev + (op - 1.0f) * 0.7f : u.max_std_dev;
  if constexpr (COVF) {
    // the cov form grows by inflate^2 (projection.tsl.ts:54-56); s carries the factor
    const double i2 = (double)inflate * inflate;
#pragma unroll
    for (int k = 0; k < 6; k++) sp.cv[k] *= i2;
    s = make_float3(inflate, inflate, inflate);
    invisible = op <= 0.0f || op < u.min_alpha || (sp.cv[0] == 0.0 && sp.cv[1] == 0.0 && sp.cv[2] == 0.0);
  } else {
    s = make_float3(sp.s.x * inflate, sp.s.y * inflate, sp.s.z * inflate);
    invisible = op <= 0.0f || op < u.min_alpha || (s.x == 0.0f && s.y == 0.0f && s.z == 0.0f);
  }
}
template <bool COVF, typename U>
__device__ __forceinline__ void prepare(Splat& sp, const U& u, float& op, float3& s, float& std_dev, bool& invisible) {
  float inflate;
  prepare<COVF>(sp, u, op, s, std_dev, invisible, inflate);
}

// inSortGuard (projection.tsl.ts:144-155): a CPU-sorted splat keeps its key
// only if its head-space bound (`radius` = extent x maxStdDev) reaches the
// guard-widened screen of eye 0, or the head is inside it.
__device__ __forceinline__ bool in_sort_guard(float radius, const double* hc, const ProjU& u) {
  const float* P0 = u.eye[0].Pf;
  const float hx0 = (float)hc[0], hy0 = (float)hc[1], hz0 = (float)hc[2];
  const float hw = P0[12] * hx0 + P0[13] * hy0 + P0[14] * hz0 + P0[15];
  const float hx = P0[0] * hx0 + P0[1] * hy0 + P0[2] * hz0 + P0[3];
  const float hy = P0[4] * hx0 + P0[5] * hy0 + P0[6] * hz0 + P0[7];
  const float w = fmaxf(hw, 1e-6f);
  const float rx = P0[0] * radius / w, ry = P0[5] * radius / w;
  const bool inside = fabsf(hx / w) - rx <= u.sort_guard + 1.0f && fabsf(hy / w) - ry <= u.sort_guard + 1.0f;
  const bool near = -hz0 <= radius;
  return hz0 < radius && (inside || near);
}

// The head-space covariance H Σ H^T of an object-space covariance cv (xx,
// yy, zz, xy, xz, yz), in double: A as xx, yy, zz, xy, xz, yz.
__device__ __forceinline__ void head_cov(const double* cv, const ProjU& u, double* A) {
  const double S[9] = {cv[0], cv[3], cv[4], cv[3], cv[1], cv[5], cv[4], cv[5], cv[2]};
  double HS[9];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) HS[3 * r + k] = u.H[3 * r] * S[k] + u.H[3 * r + 1] * S[3 + k] + u.H[3 * r + 2] * S[6 + k];
  const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int e = 0; e < 6; e++) {
    const int a = ij[e][0], b = ij[e][1];
    A[e] = HS[3 * a] * u.H[3 * b] + HS[3 * a + 1] * u.H[3 * b + 1] + HS[3 * a + 2] * u.H[3 * b + 2];
  }
}

// The covariance form's extent (the ewa method's input): the square root of
// the largest diagonal entry of the head-space covariance (H R S)(H R S)^T,
// or H Σ H^T for a covSplats splat (COVF: sp.cv, inflated).
template <bool COVF = false>
__device__ __forceinline__ float cov_extent(const Splat& sp, float3 s, const ProjU& u) {
  if constexpr (COVF) {
    double A[6];
    head_cov(sp.cv, u, A);
    return (float)sqrt(fmax(fmax(fmax(A[0], A[1]), A[2]), 0.0));
  }
  float Rg[9];
  quat_matrix(sp.q, Rg);
  float diag[3];
  for (int r = 0; r < 3; r++) {
    float acc = 0.0f;
    for (int k = 0; k < 3; k++) {
      const float hr = (float)u.H[3 * r] * Rg[k] + (float)u.H[3 * r + 1] * Rg[3 + k] + (float)u.H[3 * r + 2] * Rg[6 + k];
      const float sk = k == 0 ? s.x : (k == 1 ? s.y : s.z);
      acc += hr * hr * sk * sk;
    }
    diag[r] = acc;
  }
  return sqrtf(fmaxf(fmaxf(diag[0], diag[1]), diag[2]));
}

// The splat form's extent (the keys pass: a stored splat stays a splat,
// SplatProjector.ts:152-194): its largest head-space scale. A uniform
// transform scales the scales (transformSplat, splatTransform.tsl.ts:154-165);
// any other goes through the covariance (fromCov, splats.tsl.ts:189-198),
// whose scales are the square roots of the eigenvalues of M M^T, M = H R S:
// the largest by the trigonometric solution of the characteristic cubic
// (Smith, "Eigenvalues of a symmetric 3x3 matrix", CACM 1961), in double.
__device__ __forceinline__ double sym3_max_eigen(const double* A);

template <bool COVF = false>
__device__ __forceinline__ float splat_extent(const Splat& sp, float3 s, const ProjU& u) {
  if constexpr (COVF) {
    // a covSplats splat (sp.cv, inflated) in the splat form is fromCov of
    // its head-space covariance: the largest scale is sqrt(lambda_max)
    double A[6];
    head_cov(sp.cv, u, A);
    return (float)sqrt(fmax(sym3_max_eigen(A), 0.0));
  }
  if (u.head_uniform) return fmaxf(fmaxf(s.x, s.y), s.z) * u.head_scale;
  float Rg[9];
  quat_matrix(sp.q, Rg);
  double M[9];
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) {
      const double sk = k == 0 ? s.x : (k == 1 ? s.y : s.z);
      M[3 * r + k] = (u.H[3 * r] * Rg[k] + u.H[3 * r + 1] * Rg[3 + k] + u.H[3 * r + 2] * Rg[6 + k]) * sk;
    }
  double A[6];  // xx, yy, zz, xy, xz, yz of M M^T
  const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int e = 0; e < 6; e++) {
    const int a = ij[e][0], b = ij[e][1];
    A[e] = M[3 * a] * M[3 * b] + M[3 * a + 1] * M[3 * b + 1] + M[3 * a + 2] * M[3 * b + 2];
  }
  return (float)sqrt(fmax(sym3_max_eigen(A), 0.0));
}

// The largest eigenvalue of a symmetric 3x3 (xx, yy, zz, xy, xz, yz) by the
// trigonometric solution of its characteristic cubic (Smith, CACM 1961).
__device__ __forceinline__ double sym3_max_eigen(const double* A) {
  const double p1 = A[3] * A[3] + A[4] * A[4] + A[5] * A[5];
  const double q = (A[0] + A[1] + A[2]) / 3.0;
  const double d0 = A[0] - q, d1 = A[1] - q, d2 = A[2] - q;
  const double p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * p1;
  double lmax = q;
  if (p2 > 0.0) {
    const double p = sqrt(p2 / 6.0), ip = 1.0 / p;
    const double b0 = d0 * ip, b1 = d1 * ip, b2 = d2 * ip, b3 = A[3] * ip, b4 = A[4] * ip, b5 = A[5] * ip;
    const double det = b0 * (b1 * b2 - b5 * b5) - b3 * (b3 * b2 - b5 * b4) + b4 * (b3 * b5 - b1 * b4);
    const double r = fmin(fmax(det * 0.5, -1.0), 1.0);
    lmax = q + 2.0 * p * cos(acos(r) / 3.0);
  }
  return lmax;
}

// The projection kernel. DEG: the f32 table's SH degree (compile time), or
// -1 for word tables (u.sh_degree at run time).
template <int LAYOUT, int SHC, int DEG, int EYES, bool SIDE>
__global__ __launch_bounds__(256) void k_project(TableP t, ProjU u, OutP o) {
  constexpr int BLOCK = 256;
  const uint32_t n = t.count;
  const uint32_t invalid = key_invalid(u.sort_bits);
  const bool per_eye = EYES == 2 && u.key_per_eye;
  const int deg = DEG >= 0 ? DEG : u.sh_degree;
  F32Rsrc rs;
  if constexpr (LAYOUT == LAYOUT_F32) rs = f32_rsrc<SHC>(t, deg);

  // One splat per thread, no grid-stride loop: a loop lets the compiler hoist
  // the per-eye uniforms out of it and keep them all live in SGPRs, which
  // spills (165 SGPRs for two eyes) and costs occupancy.
  const uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) {
    Splat sp;
    float3 coef[15];
    float band[3];
    load_splat<LAYOUT, SHC, DEG>(t, rs, u.enc, i, deg, sp, coef, band);
    float op, std_dev;
    float3 s;
    bool invisible;
    prepare<LAYOUT == LAYOUT_COV>(sp, u, op, s, std_dev, invisible);
    // Two eyes: the colour first (from the head, for both eyes), so the SH
    // coefficients are dead before the eyes' centres and projections are
    // live (stereo packed SH3: 128 -> 86 VGPRs, 3 % faster); a splat both
    // eyes cull has its colour computed and not used. One eye: after the
    // cull (the mono f32 SH3 kernel is bound by its reads at 4 waves per
    // SIMD, and the earlier colour measured 0.5 % slower there).
    constexpr bool SH_FIRST = EYES == 2;
    float3 rgb = sp.rgb;
    uint32_t rgb_word = 0u;
    if (SH_FIRST) {
      if (SIDE || !invisible) rgb = sh_colour(sp, u.h, deg, coef, SHC == SHC_PACKED ? band : nullptr);
      rgb_word = encode_ext_rgb(rgb.x, rgb.y, rgb.z);
    }

    EyeCentre C[EYES];
    bool any = false;
#pragma unroll
    for (int e = 0; e < EYES; e++) {
      C[e] = eye_centre(u.eye[e], sp.c, u.ndc_z_min, u.clip_xy);
      C[e].pass = C[e].pass && !invisible;
      any |= C[e].pass;
    }
    // The key distance: |head-space centre| (per eye with key_per_eye); the
    // head-space centre also places the sortGuard test
    double head_dist = C[0].dist;
    double hc[3] = {C[0].ce[0], C[0].ce[1], C[0].ce[2]};
    const bool need_head = !per_eye || SIDE || (!u.cull_keys && u.sort_guard > 0.0f);
    if (!u.head_is_eye0 && need_head) {
      const double dx = (double)sp.c.x - u.h[0], dy = (double)sp.c.y - u.h[1], dz = (double)sp.c.z - u.h[2];
      for (int r = 0; r < 3; r++) hc[r] = fma(u.H[3 * r], dx, fma(u.H[3 * r + 1], dy, u.H[3 * r + 2] * dz));
      head_dist = sqrt(fma(hc[0], hc[0], fma(hc[1], hc[1], hc[2] * hc[2])));
    }

    uint4 words[EYES];
    bool vis[EYES];
    bool visible = false;
#pragma unroll
    for (int e = 0; e < EYES; e++) {
      words[e] = make_uint4(0u, 0u, 0u, 0u);
      vis[e] = false;
    }
    float R[9];
    if (any || SIDE) {
      if constexpr (LAYOUT != LAYOUT_COV) quat_matrix(sp.q, R);
      if (!SH_FIRST) {
        rgb = sh_colour(sp, u.h, deg, coef, SHC == SHC_PACKED ? band : nullptr);
        rgb_word = encode_ext_rgb(rgb.x, rgb.y, rgb.z);
      }
      // the eyes of a rig usually share their rotation (shared_axes): then
      // one Q serves both
      float Q0[9];
      if constexpr (LAYOUT != LAYOUT_COV) eye_axes(u.eye[0], R, Q0);
#pragma unroll
      for (int e = 0; e < EYES; e++) {
        if (!C[e].pass && !SIDE) continue;
        EwaResult r;
        if constexpr (LAYOUT == LAYOUT_COV) {
          r = ewa_eye_cov(u.eye[e], C[e], C[e].pass, sp.cv, op, std_dev, rgb_word, u);
        } else {
          float Qe[9];
          const float* Q = Q0;
          if (e > 0 && !u.shared_axes) {
            eye_axes(u.eye[e], R, Qe);
            Q = Qe;
          }
          r = ewa_eye(u.eye[e], C[e], C[e].pass, R, Q, s, op, std_dev, rgb_word, u);
        }
        if (r.ok) {
          words[e] = r.words;
          vis[e] = true;
          visible = true;
        }
        if constexpr (SIDE) {
          const size_t ss = o.sstride;
          if (o.conic) {
            float* cp = o.conic + (size_t)e * 3 * ss;
            cp[i] = r.d / r.det;
            cp[ss + i] = -r.b / r.det;
            cp[2 * ss + i] = r.a / r.det;
          }
          if (o.cov2d) {
            float* cp = o.cov2d + (size_t)e * 3 * ss;
            cp[i] = r.a;
            cp[ss + i] = r.b;
            cp[2 * ss + i] = r.d;
          }
          if (o.opac) o.opac[(size_t)e * ss + i] = r.opc;
          if (o.ndc) {
            o.ndc[(size_t)e * 2 * ss + i] = r.ndc.x;
            o.ndc[(size_t)e * 2 * ss + ss + i] = r.ndc.y;
          }
          if (o.vz) o.vz[(size_t)e * ss + i] = r.z;
          if (o.bound) {
            float* bp = o.bound + (size_t)e * 3 * ss;
            bp[i] = r.sc1;
            bp[ss + i] = r.sc2;
            bp[2 * ss + i] = r.angle;
          }
        }
      }
    }

    // sortGuard for the CPU sort's keys (projection.tsl.ts:144-155)
    const bool guarded = u.cull_keys || !(u.sort_guard > 0.0f) || in_sort_guard(cov_extent<LAYOUT == LAYOUT_COV>(sp, s, u) * u.max_std_dev, hc, u);

    // Keys (SplatProjection.ts:143-151)
    const bool keyed = u.cull_keys ? visible : (!invisible && guarded);
    uint32_t key[EYES];
    if (per_eye) {
#pragma unroll
      for (int e = 0; e < EYES; e++) {
        const bool ke = u.cull_keys ? vis[e] : keyed;
        key[e] = ke ? sort_key(fbits((float)C[e].dist), u.sort_bits) : invalid;
      }
    } else {
      key[0] = keyed ? sort_key(fbits((float)head_dist), u.sort_bits) : invalid;
    }

    // Stores
    const uint32_t at = o.at + i;
#pragma unroll
    for (int e = 0; e < EYES; e++)
      if (o.words[e]) st_stream(o.words[e] + at, words[e]);
    if (o.keys) {
      o.keys[at] = key[0];
      if (per_eye) o.keys[o.key_stride + at] = key[EYES - 1];
    }
    if constexpr (SIDE) {
      const size_t ss = o.sstride;
      if (o.rgb) {
        o.rgb[i] = rgb.x;
        o.rgb[ss + i] = rgb.y;
        o.rgb[2 * ss + i] = rgb.z;
      }
      if (o.dist) {
        if (per_eye) {
          for (int e = 0; e < EYES; e++) o.dist[(size_t)e * ss + i] = (float)C[e].dist;
        } else {
          o.dist[i] = (float)head_dist;
        }
      }
    }
  }
}

// The keys-only pass (SplatProjection.writeKeys, SplatProjection.ts:165-172,
// the "keys" pass of SplatProjector.ts:184-189): the CPU sort's keys, from
// the splat's centre, scales, rotation and opacity alone (no SH, no EWA):
// key = cullInvisible || !guarded ? invalid : compute(|head centre|), with
// the splat form's sortGuard extent (splat_extent). The head-space centre
// and distance are computed as k_project computes them.
template <int LAYOUT>
__global__ __launch_bounds__(256) void k_project_keys(TableP t, ProjU u, uint32_t* keys, uint32_t at) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.count) return;
  F32Rsrc rs;
  if constexpr (LAYOUT == LAYOUT_F32) rs = f32_rsrc<SHC_NONE>(t, 0);
  Splat sp;
  float3 coef[1];
  float band[3];
  load_splat<LAYOUT, SHC_NONE, LAYOUT == LAYOUT_F32 ? 0 : -1>(t, rs, u.enc, i, 0, sp, coef, band);
  float op, std_dev;
  float3 s;
  bool invisible;
  prepare<LAYOUT == LAYOUT_COV>(sp, u, op, s, std_dev, invisible);
  double hc[3], head_dist;
  if (u.head_is_eye0) {
    const EyeCentre C = eye_centre(u.eye[0], sp.c, u.ndc_z_min, u.clip_xy);
    for (int r = 0; r < 3; r++) hc[r] = C.ce[r];
    head_dist = C.dist;
  } else {
    const double dx = (double)sp.c.x - u.h[0], dy = (double)sp.c.y - u.h[1], dz = (double)sp.c.z - u.h[2];
    for (int r = 0; r < 3; r++) hc[r] = fma(u.H[3 * r], dx, fma(u.H[3 * r + 1], dy, u.H[3 * r + 2] * dz));
    head_dist = sqrt(fma(hc[0], hc[0], fma(hc[1], hc[1], hc[2] * hc[2])));
  }
  const bool guarded = !(u.sort_guard > 0.0f) || in_sort_guard(splat_extent<LAYOUT == LAYOUT_COV>(sp, s, u) * u.max_std_dev, hc, u);
  keys[at + i] = !invisible && guarded ? sort_key(fbits((float)head_dist), u.sort_bits) : key_invalid(u.sort_bits);
}

}  // namespace rdx

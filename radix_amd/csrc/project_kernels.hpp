// Fused projection kernel for gfx950: decode a splat (f32 planes, ext or
// packed words, with their SH), add the SH colour, project it for one or
// two eyes with EWA, write the reference's ewa words and its sort key, and
// count the key's digits for the radix sort, in one pass over the table.
//
// Reads are structure-of-arrays: lane i of a wave reads splat i of each
// plane (f32) or its uint32x4 words (ext, packed), so every load is one
// contiguous run per wave. Positions, the eye-space centre, clip
// coordinates and the key distance are computed in double (a few dozen
// FMAs); the rest in float. The 2D determinant uses the closed form
//   det(J S' J^T) = (fx fy / z^3)^2 det(L)^2 (sy^2 sz^2 lx^2 + sx^2 sz^2 ly^2 + sx^2 sy^2 lz^2),
// l = R^T (c - eye) (Cauchy-Binet with J0 x J1 = fx fy c / z^3), a sum of
// squares that stays accurate for thin splats seen edge-on, where a*d - b^2
// cancels.
#pragma once

#include "common.hpp"
#include "sort_kernels.hpp"

namespace rdx {

constexpr int LAYOUT_F32 = 0, LAYOUT_EXT = 1, LAYOUT_PACKED = 2;
constexpr int SHC_NONE = 0, SHC_F32 = 1, SHC_EXT = 2, SHC_PACKED = 3;

struct EyeU {
  double L[9];  // object -> eye linear part, row-major
  double t[3];  // object -> eye translation
  double e[3];  // the eye in object space
  double P[16]; // projection, row-major
  float Lf[9];
  float fx, fy;    // focal length in pixels: size / 2 * (P00, P11)
  float det_l2;    // det(L)^2
  float ortho_l[3];// L^-1 (0, 0, 1) (ortho only)
  int ortho;
};

struct ProjU {
  EyeU eye[2];
  int eyes;
  int head_is_eye0;   // mono: the head is eye 0 (the key distance is eye 0's)
  double h[3];        // the head in object space: SH origin
  double H[9];        // object -> head linear part (key distance when not eye 0)
  double Ht[3];
  float size[2];
  float ndc_z_min;
  float max_std_dev, min_alpha, blur, pre_blur, clip_xy, min_pr, max_pr, min_contrib;
  float focal_distance, aperture_tan, sort_guard;
  int lod_inflate, sh_degree, sort_bits, cull_keys, key_per_eye;
  int passes;         // digit passes the histograms count (0: none)
  int rb;             // their digit width
  float enc[4];       // packed encoding: rgbMin, rgbMax, lnScaleMin, lnScaleMax
};

struct TableP {
  uint32_t count;
  size_t stride;
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
  uint32_t* hist;        // digit histograms (HIST_WORDS words per segment), or NULL
  size_t hist_stride;    // eye e's histograms at hist + e * hist_stride (key_per_eye)
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
    if (ex == 0) {
      cl = -126;  // subnormal: below 2^-126, the base clamps to 0 anyway
    } else {
      cl = ex - 127 + ((u & 0x7fffffu) ? 1 : 0);
    }
  }
  const int base = min(max(cl + 15, 0), 31);
  // 255 * 2^(15 - base) from the exponent field (142 - base): exact
  const float scale = bitsf((uint32_t)(142 - base) << 23) * 255.0f;
  const uint32_t mr = (uint32_t)fminf(floorf(ar * scale + 0.5f), 255.0f);
  const uint32_t mg = (uint32_t)fminf(floorf(ag * scale + 0.5f), 255.0f);
  const uint32_t mb = (uint32_t)fminf(floorf(ab * scale + 0.5f), 255.0f);
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
  const float l = fminf(fmaxf(log2f(fmaxf(scale, 1e-20f)), -4.0f), 12.0f);
  const float n = (l + 4.0f) * (1.0f / 16.0f) * top;
  return (uint32_t)fminf(fmaxf(floorf(n + 0.5f), 0.0f), top) + 1u;
}
__device__ __forceinline__ uint32_t encode_opacity8(float o) {
  const uint32_t c = (uint32_t)floorf(fminf(fmaxf(o * 128.0f, 0.0f), 255.0f) + 0.5f);
  return max(c, 1u);
}

// ---- SH (src/gpu/sh.tsl.ts:60-139; basis constants of src/core/sh.ts:381-405) ----

// Adds bands 1..deg of coefficients coef(k) (float3) along unit d.
template <typename Coef>
__device__ __forceinline__ float3 sh_eval(float3 d, int deg, Coef coef) {
  const float x = d.x, y = d.y, z = d.z;
  float3 s = make_float3(0.0f, 0.0f, 0.0f);
  auto acc = [&](int k, float w) {
    const float3 c = coef(k);
    s.x = fmaf(c.x, w, s.x);
    s.y = fmaf(c.y, w, s.y);
    s.z = fmaf(c.z, w, s.z);
  };
  if (deg >= 1) {
    acc(0, -0.4886025f * y);
    acc(1, 0.4886025f * z);
    acc(2, -0.4886025f * x);
    if (deg >= 2) {
      const float xx = x * x, yy = y * y, zz = z * z;
      acc(3, 1.0925484f * x * y);
      acc(4, -1.0925484f * y * z);
      acc(5, 0.3153916f * (2.0f * zz - xx - yy));
      acc(6, -1.0925484f * x * z);
      acc(7, 0.5462742f * (xx - yy));
      if (deg >= 3) {
        acc(8, -0.5900436f * y * (3.0f * xx - yy));
        acc(9, 2.8906114f * x * y * z);
        acc(10, -0.4570458f * y * (4.0f * zz - xx - yy));
        acc(11, 0.3731763f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy));
        acc(12, -0.4570458f * x * (4.0f * zz - xx - yy));
        acc(13, 1.4453057f * z * (xx - yy));
        acc(14, -0.5900436f * x * (xx - 3.0f * yy));
      }
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
};

template <int LAYOUT>
__device__ __forceinline__ void load_geometry(const TableP& t, const ProjU& u, uint32_t i, Splat& sp) {
  if constexpr (LAYOUT == LAYOUT_F32) {
    const size_t st = t.stride;
    sp.c = make_float3(t.center[i], t.center[st + i], t.center[2 * st + i]);
    sp.s = make_float3(__expf(t.log_scale[i]), __expf(t.log_scale[st + i]), __expf(t.log_scale[2 * st + i]));
    sp.q = make_float4(t.quat[i], t.quat[st + i], t.quat[2 * st + i], t.quat[3 * st + i]);
    sp.op = t.opacity[i];
  } else if constexpr (LAYOUT == LAYOUT_EXT) {
    const uint4 A = t.ext_a[i], B = t.ext_b[i];
    sp.c = make_float3(bitsf(A.x), bitsf(A.y), bitsf(A.z));
    sp.op = (float)(A.w >> 24) * (1.0f / 128.0f);
    sp.rgb = decode_ext_rgb(B.x);
    sp.s = make_float3(__expf(half_to_float(B.y)), __expf(half_to_float(B.y >> 16)), __expf(half_to_float(B.z)));
    sp.q = quat_oct_decode(B.w, 10, 12);
  } else {
    const uint4 W = t.packed[i];
    const float lo = u.enc[0], range = u.enc[1] - u.enc[0];
    sp.rgb = make_float3(lo + (float)(W.x & 0xffu) * (1.0f / 255.0f) * range,
                         lo + (float)((W.x >> 8) & 0xffu) * (1.0f / 255.0f) * range,
                         lo + (float)((W.x >> 16) & 0xffu) * (1.0f / 255.0f) * range);
    sp.op = (float)(W.x >> 24) * (1.0f / 255.0f);
    sp.c = make_float3(half_to_float(W.y), half_to_float(W.y >> 16), half_to_float(W.z));
    const float step = (u.enc[3] - u.enc[2]) * (1.0f / 254.0f);
    auto sc = [&](uint32_t code) { return code == 0u ? 0.0f : __expf(u.enc[2] + ((float)code - 1.0f) * step); };
    sp.s = make_float3(sc(W.w & 0xffu), sc((W.w >> 8) & 0xffu), sc((W.w >> 16) & 0xffu));
    sp.q = quat_oct_decode((W.z >> 16) | ((W.w >> 24) << 16), 8, 8);
  }
}

// SH colour of splat i (base + bands), clamped at 0 (sh.tsl.ts:255-266)
template <int LAYOUT, int SHC>
__device__ __forceinline__ float3 shade(const TableP& t, const ProjU& u, uint32_t i, const Splat& sp) {
  float3 rgb = sp.rgb;
  if constexpr (LAYOUT == LAYOUT_F32) {
    const size_t st = t.stride;
    rgb = make_float3(t.rgb[i], t.rgb[st + i], t.rgb[2 * st + i]);
  }
  const int deg = u.sh_degree;
  if constexpr (SHC == SHC_NONE) {
    return rgb;
  } else {
    if (deg <= 0) return rgb;
    // direction from the head (object space) to the centre
    const float dx = (float)((double)sp.c.x - u.h[0]);
    const float dy = (float)((double)sp.c.y - u.h[1]);
    const float dz = (float)((double)sp.c.z - u.h[2]);
    const float rl = rsqrtf(dx * dx + dy * dy + dz * dz);
    const float3 d = make_float3(dx * rl, dy * rl, dz * rl);
    float3 s;
    if constexpr (SHC == SHC_F32) {
      const size_t st = t.stride;
      const float* p = t.sh + i;
      s = sh_eval(d, deg, [&](int k) { return make_float3(p[(3 * k) * st], p[(3 * k + 1) * st], p[(3 * k + 2) * st]); });
    } else if constexpr (SHC == SHC_EXT) {
      uint32_t w[16];
      const int texels = deg == 1 ? 1 : (deg == 2 ? 2 : 4);
#pragma unroll
      for (int x = 0; x < 4; x++) {
        if (x < texels) {
          const uint4 v = t.shc[x][i];
          w[4 * x] = v.x;
          w[4 * x + 1] = v.y;
          w[4 * x + 2] = v.z;
          w[4 * x + 3] = v.w;
        }
      }
      s = sh_eval(d, deg, [&](int k) { return decode_ext_rgb(w[k]); });
    } else {
      uint32_t w[12];
      const int texels = deg == 1 ? 1 : (deg == 2 ? 2 : 3);
#pragma unroll
      for (int x = 0; x < 3; x++) {
        if (x < texels) {
          const uint4 v = t.shc[x][i];
          w[4 * x] = v.x;
          w[4 * x + 1] = v.y;
          w[4 * x + 2] = v.z;
          w[4 * x + 3] = v.w;
        }
      }
      const float sc1 = sh_scale(sh_byte(w, 0));
      const float sc2 = deg >= 2 ? sh_scale(sh_byte(w, 10)) : 0.0f;
      const float sc3 = deg >= 3 ? sh_scale(sh_byte(w, 26)) : 0.0f;
      s = sh_eval(d, deg, [&](int k) {
        const int band_start = k < 3 ? 0 : (k < 8 ? 10 : 26);
        const int first = k < 3 ? 0 : (k < 8 ? 3 : 8);
        const float sc = k < 3 ? sc1 : (k < 8 ? sc2 : sc3);
        const int at = band_start + 1 + (k - first) * 3;
        return make_float3((float)(int8_t)sh_byte(w, at) * sc, (float)(int8_t)sh_byte(w, at + 1) * sc,
                           (float)(int8_t)sh_byte(w, at + 2) * sc);
      });
    }
    return make_float3(fmaxf(rgb.x + s.x, 0.0f), fmaxf(rgb.y + s.y, 0.0f), fmaxf(rgb.z + s.z, 0.0f));
  }
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

struct EyeResult {
  bool visible;
  uint4 words;
};

template <int LAYOUT, int SHC, int EYES, bool SIDE, bool HIST>
__global__ __launch_bounds__(256) void k_project(TableP t, ProjU u, OutP o) {
  constexpr int BLOCK = 256;
  constexpr int HEYES = HIST ? EYES : 1;
  __shared__ uint32_t s_hist[HIST ? HEYES * MAX_PASSES * HIST_BINS : 1];
  if constexpr (HIST) {
    for (int x = threadIdx.x; x < HEYES * u.passes * HIST_BINS; x += BLOCK) s_hist[x] = 0u;
    __syncthreads();
  }
  const uint32_t n = t.count;
  const uint32_t invalid = key_invalid(u.sort_bits);
  const bool per_eye = EYES == 2 && u.key_per_eye;

  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    Splat sp;
    load_geometry<LAYOUT>(t, u, i, sp);

    // prepareSplat (projection.tsl.ts:42-64)
    float op = sp.op, inflate = 1.0f;
    if (op > 1.0f) {
      op = fminf(op * 4.0f - 3.0f, 5.0f);
      if (u.lod_inflate) {
        inflate = cbrtf(__expf((op * op - 1.0f) * 0.36787944117144233f));
        op = 1.0f;
      }
    }
    const float3 s = make_float3(sp.s.x * inflate, sp.s.y * inflate, sp.s.z * inflate);
    const float std_dev = op > 1.0f ? u.max_std_dev + (op - 1.0f) * 0.7f : u.max_std_dev;
    // cullInvisible (projection.tsl.ts:66-69)
    const bool invisible = op <= 0.0f || op < u.min_alpha || (s.x == 0.0f && s.y == 0.0f && s.z == 0.0f);

    // Eye-space centres and clip coordinates, in double.
    double ce[EYES][3], clip[EYES][4];
    bool pass[EYES];
    bool any = false;
#pragma unroll
    for (int e = 0; e < EYES; e++) {
      const EyeU& E = u.eye[e];
      const double px = (double)sp.c.x, py = (double)sp.c.y, pz = (double)sp.c.z;
      for (int r = 0; r < 3; r++)
        ce[e][r] = fma(E.L[3 * r], px, fma(E.L[3 * r + 1], py, fma(E.L[3 * r + 2], pz, E.t[r])));
      for (int r = 0; r < 4; r++)
        clip[e][r] = fma(E.P[4 * r], ce[e][0], fma(E.P[4 * r + 1], ce[e][1], fma(E.P[4 * r + 2], ce[e][2], E.P[4 * r + 3])));
      // cullDepth, cullClipXY (projection.tsl.ts:132-140)
      const bool cull_depth = ce[e][2] >= 0.0 || clip[e][2] < (double)u.ndc_z_min * clip[e][3] || clip[e][2] >= clip[e][3];
      const double limit = (double)fminf(u.clip_xy, 2.0f) * clip[e][3];
      const bool cull_xy = fabs(clip[e][0]) > limit || fabs(clip[e][1]) > limit;
      pass[e] = !(invisible || cull_depth || cull_xy);
      any |= pass[e];
    }

    // The key distance: |head-space centre| (per eye with key_per_eye)
    double dist[EYES];
#pragma unroll
    for (int e = 0; e < EYES; e++)
      dist[e] = sqrt(fma(ce[e][0], ce[e][0], fma(ce[e][1], ce[e][1], ce[e][2] * ce[e][2])));
    double head_dist = dist[0];
    double hc[3] = {ce[0][0], ce[0][1], ce[0][2]};
    if (!u.head_is_eye0) {
      const double px = (double)sp.c.x, py = (double)sp.c.y, pz = (double)sp.c.z;
      for (int r = 0; r < 3; r++) hc[r] = fma(u.H[3 * r], px, fma(u.H[3 * r + 1], py, fma(u.H[3 * r + 2], pz, u.Ht[r])));
      head_dist = sqrt(fma(hc[0], hc[0], fma(hc[1], hc[1], hc[2] * hc[2])));
    }

    EyeResult res[EYES];
    float R[9];
    float3 rgb = make_float3(0.0f, 0.0f, 0.0f);
    bool visible = false;
#pragma unroll
    for (int e = 0; e < EYES; e++) {
      res[e].visible = false;
      res[e].words = make_uint4(0u, 0u, 0u, 0u);
    }

    if (any || SIDE) {
      quat_matrix(sp.q, R);
      rgb = shade<LAYOUT, SHC>(t, u, i, sp);
    }
    if (any || SIDE) {
#pragma unroll
      for (int e = 0; e < EYES; e++) {
        if (!pass[e] && !SIDE) continue;
        const EyeU& E = u.eye[e];
        const float x = (float)ce[e][0], y = (float)ce[e][1], z = (float)ce[e][2];
        // Q = L R, J rows (projections.tsl.ts:104-117)
        float Q[9];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int c = 0; c < 3; c++)
            Q[3 * r + c] = fmaf(E.Lf[3 * r], R[c], fmaf(E.Lf[3 * r + 1], R[3 + c], E.Lf[3 * r + 2] * R[6 + c]));
        float j00, j02, j11, j12, kdet;
        float l0, l1, l2;
        if (E.ortho) {
          j00 = E.fx; j02 = 0.0f; j11 = E.fy; j12 = 0.0f;
          kdet = E.fx * E.fy;
          l0 = R[0] * E.ortho_l[0] + R[3] * E.ortho_l[1] + R[6] * E.ortho_l[2];
          l1 = R[1] * E.ortho_l[0] + R[4] * E.ortho_l[1] + R[7] * E.ortho_l[2];
          l2 = R[2] * E.ortho_l[0] + R[5] * E.ortho_l[1] + R[8] * E.ortho_l[2];
        } else {
          const float iz = 1.0f / z;
          j00 = E.fx * iz;
          j11 = E.fy * iz;
          j02 = -j00 * x * iz;
          j12 = -j11 * y * iz;
          kdet = j00 * j11 * iz;  // fx fy / z^3
          const float dx = (float)((double)sp.c.x - E.e[0]);
          const float dy = (float)((double)sp.c.y - E.e[1]);
          const float dz = (float)((double)sp.c.z - E.e[2]);
          l0 = R[0] * dx + R[3] * dy + R[6] * dz;
          l1 = R[1] * dx + R[4] * dy + R[7] * dz;
          l2 = R[2] * dx + R[5] * dy + R[8] * dz;
        }
        // T rows = (J Q) diag(s)
        const float t00 = (j00 * Q[0] + j02 * Q[6]) * s.x, t01 = (j00 * Q[1] + j02 * Q[7]) * s.y,
                    t02 = (j00 * Q[2] + j02 * Q[8]) * s.z;
        const float t10 = (j11 * Q[3] + j12 * Q[6]) * s.x, t11 = (j11 * Q[4] + j12 * Q[7]) * s.y,
                    t12 = (j11 * Q[5] + j12 * Q[8]) * s.z;
        const float a0 = t00 * t00 + t01 * t01 + t02 * t02;
        const float d0 = t10 * t10 + t11 * t11 + t12 * t12;
        const float b0 = t00 * t10 + t01 * t11 + t02 * t12;
        const float sxx = s.x * s.x, syy = s.y * s.y, szz = s.z * s.z;
        const float det0 = kdet * kdet * E.det_l2 * (syy * szz * l0 * l0 + sxx * szz * l1 * l1 + sxx * syy * l2 * l2);
        // blur and opacity compensation (projections.tsl.ts:119-129)
        float blur = u.blur;
        if (u.focal_distance > 0.0f && u.aperture_tan > 0.0f && z < 0.0f) {
          const float depth = -z;
          const float coc_view = fabsf(depth - u.focal_distance) * u.aperture_tan;
          const float focal_pixels = u.size[1] * 0.5f * (float)E.P[5];
          const float coc = E.ortho ? coc_view * focal_pixels : coc_view * focal_pixels / depth;
          blur = fminf(fmaxf(coc * coc, blur), u.max_pr * u.max_pr);
        }
        const float pre = u.pre_blur;
        const float ap = a0 + pre, dp = d0 + pre;
        const float det_orig = det0 + pre * (a0 + d0) + pre * pre;
        const float a = ap + blur, d = dp + blur, b = b0;
        const float det = det_orig + blur * (ap + dp) + blur * blur;
        const float opc = op * sqrtf(fmaxf(0.0f, det_orig / det));
        const float contribution = opc * 6.283185307179586f * sqrtf(fmaxf(0.0f, det));
        // boundFromCov (projection.tsl.ts:223-240)
        const float half = (a - d) * 0.5f;
        const float eig1 = (a + d) * 0.5f + sqrtf(half * half + b * b);
        const float eig2 = eig1 > 0.0f ? fmaxf(det / eig1, 0.0f) : 0.0f;
        const float sc1 = fminf(u.max_pr, std_dev * sqrtf(fmaxf(eig1, 0.0f)));
        const float sc2 = fminf(u.max_pr, std_dev * sqrtf(eig2));
        const bool bound_ok = sc1 >= u.min_pr || sc2 >= u.min_pr;
        float angle = 0.0f;
        {
          float vx, vy;
          if (fabsf(a - eig1) >= fabsf(d - eig1)) {
            vx = b;
            vy = eig1 - a;
          } else {
            vx = eig1 - d;
            vy = b;
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
          angle = atan2f(vy, vx);
          if (angle < 0.0f) angle += 3.14159265358979323846f;
        }
        const bool ok = pass[e] && opc >= u.min_alpha && contribution >= u.min_contrib && bound_ok;
        float2 ndc = make_float2(0.0f, 0.0f);
        if (clip[e][3] > 0.0) ndc = make_float2((float)(clip[e][0] / clip[e][3]), (float)(clip[e][1] / clip[e][3]));
        if (ok) {
          const float stored = opc > 1.0f ? (opc + 3.0f) * 0.25f : opc;
          res[e].words = make_uint4(encode_fixed16(ndc.x) | (encode_fixed16(ndc.y) << 16),
                                    encode_depth20(z) | (encode_angle12(angle) << 20),
                                    encode_ext_rgb(rgb.x, rgb.y, rgb.z),
                                    encode_opacity8(stored) | (encode_log_scale12(sc1) << 8) |
                                        (encode_log_scale12(sc2) << 20));
          res[e].visible = true;
          visible = true;
        }
        if constexpr (SIDE) {
          const size_t ss = o.sstride;
          if (o.conic) {
            float* cp = o.conic + (size_t)e * 3 * ss;
            cp[i] = d / det;
            cp[ss + i] = -b / det;
            cp[2 * ss + i] = a / det;
          }
          if (o.cov2d) {
            float* cp = o.cov2d + (size_t)e * 3 * ss;
            cp[i] = a;
            cp[ss + i] = b;
            cp[2 * ss + i] = d;
          }
          if (o.opac) o.opac[(size_t)e * ss + i] = opc;
          if (o.ndc) {
            o.ndc[(size_t)e * 2 * ss + i] = ndc.x;
            o.ndc[(size_t)e * 2 * ss + ss + i] = ndc.y;
          }
          if (o.vz) o.vz[(size_t)e * ss + i] = z;
          if (o.bound) {
            float* bp = o.bound + (size_t)e * 3 * ss;
            bp[i] = sc1;
            bp[ss + i] = sc2;
            bp[2 * ss + i] = angle;
          }
        }
      }
    }

    // sortGuard for the CPU sort's keys (projection.tsl.ts:144-155)
    bool guarded = true;
    if (!u.cull_keys && u.sort_guard > 0.0f) {
      float Rg[9];
      quat_matrix(sp.q, Rg);
      // head-space cov diagonal: (H3 R S)(H3 R S)^T
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
      const float radius = sqrtf(fmaxf(fmaxf(diag[0], diag[1]), diag[2])) * u.max_std_dev;
      const double* P0 = u.eye[0].P;
      const double hw = fma(P0[12], hc[0], fma(P0[13], hc[1], fma(P0[14], hc[2], P0[15])));
      const double hx = fma(P0[0], hc[0], fma(P0[1], hc[1], fma(P0[2], hc[2], P0[3])));
      const double hy = fma(P0[4], hc[0], fma(P0[5], hc[1], fma(P0[6], hc[2], P0[7])));
      const float w = fmaxf((float)hw, 1e-6f);
      const float rx = (float)P0[0] * radius / w, ry = (float)P0[5] * radius / w;
      const bool inside = fabsf((float)hx / w) - rx <= u.sort_guard + 1.0f && fabsf((float)hy / w) - ry <= u.sort_guard + 1.0f;
      const bool near = -(float)hc[2] <= radius;
      guarded = (float)hc[2] < radius && (inside || near);
    }

    // Keys (SplatProjection.ts:143-151)
    const bool keyed = u.cull_keys ? visible : (!invisible && guarded);
    uint32_t key[EYES];
    if (per_eye) {
#pragma unroll
      for (int e = 0; e < EYES; e++) {
        const bool ke = u.cull_keys ? res[e].visible : keyed;
        key[e] = ke ? sort_key(fbits((float)dist[e]), u.sort_bits) : invalid;
      }
    } else {
      key[0] = keyed ? sort_key(fbits((float)head_dist), u.sort_bits) : invalid;
    }

    // Stores
    const uint32_t at = o.at + i;
#pragma unroll
    for (int e = 0; e < EYES; e++)
      if (o.words[e]) o.words[e][at] = res[e].words;
    if (o.keys) {
      o.keys[at] = key[0];
      if (per_eye) o.keys[o.key_stride + at] = key[EYES - 1];
    }
    if constexpr (HIST) {
      if (key[0] != invalid) hist_add_key(s_hist, key[0], u.passes, u.rb);
      if (EYES == 2 && per_eye && key[EYES - 1] != invalid)
        hist_add_key(s_hist + MAX_PASSES * HIST_BINS, key[EYES - 1], u.passes, u.rb);
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
          for (int e = 0; e < EYES; e++) o.dist[(size_t)e * ss + i] = (float)dist[e];
        } else {
          o.dist[i] = (float)head_dist;
        }
      }
    }
  }
  if constexpr (HIST) {
    __syncthreads();
    hist_flush<BLOCK>(s_hist, o.hist, u.passes);
    if (EYES == 2 && per_eye) hist_flush<BLOCK>(s_hist + MAX_PASSES * HIST_BINS, o.hist + o.hist_stride, u.passes);
  }
}

}  // namespace rdx

// The ray-evaluated projection methods (the projection words of
// src/render/projections.tsl.ts:170-429 with the helpers of
// projection.tsl.ts:169-275, 312-360). 3dgut and htgs: per splat, six head
// words (the head-space centre as an octahedral direction and a float
// distance, the colour, the head-space rotation, the MIP-inflated scales as
// half floats, the opacity) and two words per eye (a screen bound: ndc
// 2x16, a 10-bit angle and two 11-bit log2 pixel scales rounded outward).
// raygs: four head words (centre, colour, opacity) and three per eye (the
// frame of a quad facing the eye, c^2, its two axis lengths).
#pragma once

#include <hip/hip_fp16.h>

#include "project_kernels.hpp"

namespace rdx {

constexpr int METHOD_EWA = 0, METHOD_3DGUT = 1, METHOD_HTGS = 2, METHOD_RAYGS = 3;

// Uniforms of a ray projection: the ewa uniforms plus the rotations of the
// object transform (uniform scale times a rotation; splatTransform.tsl.ts:142-152).
struct RayU {
  ProjU u;
  float hq[4];     // object -> head rotation (x, y, z, w)
  float eq[2][4];  // object -> eye rotations
  float oscale;    // the object transform's scale
  // An object transform that is not a rotation times a uniform scale: the
  // splat's head-space form is fromCov of H Σ H^T (u.H), and each eye turns
  // it by its head -> eye rotation ehq.
  int nonuniform;
  float ehq[2][4];
};

// Word channels: ch[0], ch[1] uint32x4; with two eyes ch[2] uint32x2
// (SplatProjection.ts:46-53).
struct RayOut {
  uint32_t* ch[3];
  uint32_t* keys;
  uint32_t at;
};

__device__ __forceinline__ float4 quat_mul(float4 a, float4 b) {
  return make_float4(a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
                     a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
__device__ __forceinline__ float4 quat_normalize(float4 q) {
  const float n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (!(n2 > 0.0f)) return q;
  const float r = rsqrtf(n2);
  return make_float4(q.x * r, q.y * r, q.z * r, q.w * r);
}
// v rotated by unit q
__device__ __forceinline__ float3 quat_rotate(float4 q, float3 v) {
  const float tx = 2.0f * (q.y * v.z - q.z * v.y), ty = 2.0f * (q.z * v.x - q.x * v.z), tz = 2.0f * (q.x * v.y - q.y * v.x);
  return make_float3(v.x + q.w * tx + (q.y * tz - q.z * ty), v.y + q.w * ty + (q.z * tx - q.x * tz),
                     v.z + q.w * tz + (q.x * ty - q.y * tx));
}

// A bound: centre in ndc, angle of the first axis in [0, pi), half extents in pixels
struct RayBound {
  float nx, ny, angle, s1, s2;
};

// boundFromCov (projection.tsl.ts:223-240)
template <typename U>
__device__ __forceinline__ bool bound_from_cov(float a, float b, float d, float std_dev, const U& u, RayBound& bd) {
  const float det = a * d - b * b;
  const float half = (a - d) * 0.5f;
  const float eig1 = (a + d) * 0.5f + sqrtf(half * half + b * b);
  const float eig2 = eig1 > 0.0f ? fmaxf(det / eig1, 0.0f) : 0.0f;
  const float sc1 = fminf(u.max_pr, std_dev * sqrtf(fmaxf(eig1, 0.0f)));
  const float sc2 = fminf(u.max_pr, std_dev * sqrtf(eig2));
  const bool ok = sc1 >= u.min_pr || sc2 >= u.min_pr;
  bd.angle = 0.0f;
  bd.s1 = sc1;
  bd.s2 = sc2;
  if (ok) {
    float vx, vy;
    if (fabsf(a - eig1) >= fabsf(d - eig1)) {
      vx = b;
      vy = eig1 - a;
    } else {
      vx = eig1 - d;
      vy = b;
    }
    const float len = sqrtf(vx * vx + vy * vy);
    if (len > 0.0f) {
      vx /= len;
      vy /= len;
    } else {
      vx = 1.0f;
      vy = 0.0f;
    }
    float ang = atan2f(vy, vx);
    bd.angle = ang < 0.0f ? ang + 3.14159265358979323846f : ang;
  }
  return ok;
}

// screenBound (projection.tsl.ts:243-248)
template <typename U>
__device__ __forceinline__ RayBound screen_bound(const U& u) {
  const float s = fminf(u.max_pr, sqrtf(u.size[0] * u.size[0] + u.size[1] * u.size[1]) * 0.5f);
  return RayBound{0.0f, 0.0f, 0.0f, s, s};
}

// clipBoundToScreen (projection.tsl.ts:250-275)
__device__ __forceinline__ bool clip_bound_to_screen(const float* size, RayBound& bd) {
  const float hx = size[0] * 0.5f, hy = size[1] * 0.5f;
  float sn, cs;
  sincosf(bd.angle, &sn, &cs);
  const float cx = cs * bd.s1, cy = cs * bd.s2, sx = sn * bd.s1, sy = sn * bd.s2;
  const float ex = sqrtf(cx * cx + sy * sy), ey = sqrtf(sx * sx + cy * cy);
  const float px = bd.nx * hx, py = bd.ny * hy;
  float lox = px - ex, loy = py - ey, hix = px + ex, hiy = py + ey;
  if (!(lox >= -hx && loy >= -hy && hix <= hx && hiy <= hy)) {
    lox = fmaxf(lox, -hx);
    loy = fmaxf(loy, -hy);
    hix = fminf(hix, hx);
    hiy = fminf(hiy, hy);
    if (lox >= hix || loy >= hiy) return false;
    bd.nx = (lox + hix) * 0.5f / hx;
    bd.ny = (loy + hiy) * 0.5f / hy;
    bd.angle = 0.0f;
    bd.s1 = (hix - lox) * 0.5f;
    bd.s2 = (hiy - loy) * 0.5f;
  }
  return true;
}

// The columns of R(q) diag(s), q unit (scaleQuatMatrix, math.tsl.ts)
__device__ __forceinline__ void scaled_axes(float4 q, float3 s, float3 col[3]) {
  col[0] = quat_rotate(q, make_float3(s.x, 0.0f, 0.0f));
  col[1] = quat_rotate(q, make_float3(0.0f, s.y, 0.0f));
  col[2] = quat_rotate(q, make_float3(0.0f, 0.0f, s.z));
}

__device__ __forceinline__ float4 clip_point(const float* P, float3 p, float w) {
  return make_float4(fmaf(P[0], p.x, fmaf(P[1], p.y, fmaf(P[2], p.z, P[3] * w))),
                     fmaf(P[4], p.x, fmaf(P[5], p.y, fmaf(P[6], p.z, P[7] * w))),
                     fmaf(P[8], p.x, fmaf(P[9], p.y, fmaf(P[10], p.z, P[11] * w))),
                     fmaf(P[12], p.x, fmaf(P[13], p.y, fmaf(P[14], p.z, P[15] * w))));
}

// gutBound: the unscented transform of the 7 sigma points c, c +- sqrt(3) s_i R e_i
// (projections.tsl.ts:183-222)
template <typename U>
__device__ __forceinline__ bool gut_bound(const float* P, const float* size, float3 c, float4 q, float3 s, float std_dev,
                                          const U& u, RayBound& bd) {
  const float k = 1.7320508075688772f;
  float3 col[3];
  scaled_axes(q, make_float3(s.x * k, s.y * k, s.z * k), col);
  const float hx = size[0] * 0.5f, hy = size[1] * 0.5f;
  bool front = true;
  float px[7], py[7];
#pragma unroll
  for (int i = 0; i < 7; i++) {
    float3 p = c;
    if (i > 0) {
      const float3 a = col[(i - 1) >> 1];
      const float sg = (i & 1) ? 1.0f : -1.0f;
      p = make_float3(c.x + sg * a.x, c.y + sg * a.y, c.z + sg * a.z);
    }
    const float4 cl = clip_point(P, p, 1.0f);
    front = front && cl.w > 0.0f;
    px[i] = cl.x / cl.w * hx;
    py[i] = cl.y / cl.w * hy;
  }
  if (!front) {
    bd = screen_bound(u);
    return true;
  }
  float mx = 0.0f, my = 0.0f;
#pragma unroll
  for (int i = 1; i < 7; i++) {
    mx += px[i] * (1.0f / 6.0f);
    my += py[i] * (1.0f / 6.0f);
  }
  const float d0x = px[0] - mx, d0y = py[0] - my;
  float a = d0x * d0x * 2.0f, b = d0x * d0y * 2.0f, d = d0y * d0y * 2.0f;
#pragma unroll
  for (int i = 1; i < 7; i++) {
    const float dx = px[i] - mx, dy = py[i] - my;
    a += dx * dx * (1.0f / 6.0f);
    b += dx * dy * (1.0f / 6.0f);
    d += dy * dy * (1.0f / 6.0f);
  }
  RayBound r;
  r.nx = mx / hx;
  r.ny = my / hy;
  const bool ok = bound_from_cov(a, b, d, std_dev, u, r);
  bd = r;
  return ok;
}

// htgsBound: the exact outline of the std-dev ellipsoid from its dual
// quadric (projections.tsl.ts:224-264)
template <typename U>
__device__ __forceinline__ bool htgs_bound(const float* P, const float* size, float3 c, float4 q, float3 s, float std_dev,
                                           float2 eye_ndc, const U& u, RayBound& bd) {
  float3 col[3];
  scaled_axes(q, make_float3(s.x * std_dev, s.y * std_dev, s.z * std_dev), col);
  float4 T[3];
#pragma unroll
  for (int i = 0; i < 3; i++) T[i] = clip_point(P, col[i], 0.0f);
  const float4 c3 = clip_point(P, c, 1.0f);
  const float xs[3] = {T[0].x, T[1].x, T[2].x}, ys[3] = {T[0].y, T[1].y, T[2].y}, ws[3] = {T[0].w, T[1].w, T[2].w};
  const float sww = ws[0] * ws[0] + ws[1] * ws[1] + ws[2] * ws[2];
  const float qww = sww - c3.w * c3.w;
  bool valid = qww < 0.0f && c3.w > 0.0f;
  float a = 0.0f, b = 0.0f, d = 0.0f, cx = 0.0f, cy = 0.0f;
  if (valid) {
    float uu[3], vv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      uu[i] = xs[i] * c3.w - ws[i] * c3.x;
      vv[i] = ys[i] * c3.w - ws[i] * c3.y;
    }
    const float xw = xs[0] * ws[0] + xs[1] * ws[1] + xs[2] * ws[2];
    const float yw = ys[0] * ws[0] + ys[1] * ws[1] + ys[2] * ws[2];
    const float xx = xs[0] * xs[0] + xs[1] * xs[1] + xs[2] * xs[2];
    const float yy = ys[0] * ys[0] + ys[1] * ys[1] + ys[2] * ys[2];
    const float xy = xs[0] * ys[0] + xs[1] * ys[1] + xs[2] * ys[2];
    const float uu2 = uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2];
    const float vv2 = vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2];
    const float uv = uu[0] * vv[0] + uu[1] * vv[1] + uu[2] * vv[2];
    const float inv_c2 = 1.0f / (qww * qww);
    const float hx = size[0] * 0.5f, hy = size[1] * 0.5f;
    a = (uu2 - (sww * xx - xw * xw)) * inv_c2 * hx * hx;
    b = (uv - (sww * xy - xw * yw)) * inv_c2 * hx * hy;
    d = (vv2 - (sww * yy - yw * yw)) * inv_c2 * hy * hy;
    const float wu = ws[0] * uu[0] + ws[1] * uu[1] + ws[2] * uu[2];
    const float wv = ws[0] * vv[0] + ws[1] * vv[1] + ws[2] * vv[2];
    const float den = c3.w * qww;
    cx = eye_ndc.x + wu / den;
    cy = eye_ndc.y + wv / den;
    valid = a > 0.0f && d > 0.0f && a * d - b * b > 0.0f;
  }
  if (!valid) {
    bd = screen_bound(u);
    return true;
  }
  RayBound r;
  r.nx = cx;
  r.ny = cy;
  const bool ok = bound_from_cov(a, b, d, 1.0f, u, r);
  bd = r;
  return ok;
}

// ---- word codecs (projection.tsl.ts:277-360, encoding.tsl.ts:50-134) ------

// log2 pixel scale in `bits`, rounded outward (encodeLogScale)
__device__ __forceinline__ uint32_t encode_log_scale_out(float scale, int bits) {
  const float top = (float)((1u << bits) - 2u);
  const float nrm = (fminf(fmaxf(log2f(fmaxf(scale, 1e-20f)), -4.0f), 12.0f) + 4.0f) * (1.0f / 16.0f) * top;
  return (uint32_t)fminf(fmaxf(ceilf(nrm), 0.0f), top) + 1u;
}

// encodeBound: [ndc 2x16, angle10 | log2 s1 (11, outward) << 10 | log2 s2 << 21]
__device__ __forceinline__ uint2 encode_bound(const RayBound& bd) {
  const uint32_t angle = (uint32_t)floorf(bd.angle * (1024.0f / 3.14159265358979323846f) + 0.5f) & 1023u;
  // sin(pi / 1024 / 2) * 1.001: the angle's half step folded into the scales
  const float slack = fmaxf(bd.s1, bd.s2) * (0.0015339801862847655f * 1.001f);
  return make_uint2(encode_fixed16(bd.nx) | (encode_fixed16(bd.ny) << 16),
                    angle | (encode_log_scale_out(bd.s1 + slack, 11) << 10) | (encode_log_scale_out(bd.s2 + slack, 11) << 21));
}

// encodeCenter: octahedral direction (20 + 20 bits) and the float distance;
// low holds the direction's low 4 + 4 bits in bits 24-31
__device__ __forceinline__ void encode_center(float3 c, uint32_t& dir, uint32_t& dist, uint32_t& low) {
  dist = fbits(sqrtf(c.x * c.x + c.y * c.y + c.z * c.z));
  const float l1 = fmaxf(fabsf(c.x) + fabsf(c.y) + fabsf(c.z), 1e-30f);
  float px = c.x / l1, py = c.y / l1;
  const float sx = px >= 0.0f ? 1.0f : -1.0f, sy = py >= 0.0f ? 1.0f : -1.0f;
  if (c.z < 0.0f) {
    const float fx = (1.0f - fabsf(py)) * sx, fy = (1.0f - fabsf(px)) * sy;
    px = fx;
    py = fy;
  }
  const uint32_t qx = (uint32_t)floorf((px * 0.5f + 0.5f) * 1048575.0f + 0.5f);
  const uint32_t qy = (uint32_t)floorf((py * 0.5f + 0.5f) * 1048575.0f + 0.5f);
  dir = (qx >> 4) | ((qy >> 4) << 16);
  low = ((qx & 15u) << 24) | (qy << 28);
}

// Half-float bits (round to nearest even)
__device__ __forceinline__ uint32_t half_bits(float f) { return (uint32_t)__half_as_ushort(__float2half_rn(f)); }

// encodeQuatOct32: 10 + 10 bit octahedral axis, 12-bit angle, for w >= 0
__device__ __forceinline__ uint32_t encode_quat_oct32(float4 q) {
  const float len = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
  if (!(len > 0.0f)) return 0u;
  const float s = (q.w < 0.0f ? -1.0f : 1.0f) / len;
  const float angle = acosf(fminf(q.w * s, 1.0f)) * 2.0f;
  const uint32_t acode = (uint32_t)floorf(angle * (4095.0f / 3.14159265358979323846f) + 0.5f);
  const float ax = q.x * s, ay = q.y * s, az = q.z * s;
  uint32_t oct = 0u;
  const float l1 = fabsf(ax) + fabsf(ay) + fabsf(az);
  if (l1 > 0.0f) {
    float px = ax / l1, py = ay / l1;
    if (az < 0.0f) {
      const float ox = px;
      px = (1.0f - fabsf(py)) * (ox >= 0.0f ? 1.0f : -1.0f);
      py = (1.0f - fabsf(ox)) * (py >= 0.0f ? 1.0f : -1.0f);
    }
    const uint32_t cu = (uint32_t)floorf((px * 0.5f + 0.5f) * 1023.0f + 0.5f);
    const uint32_t cv = (uint32_t)floorf((py * 0.5f + 0.5f) * 1023.0f + 0.5f);
    oct = cu | (cv << 10);
    if (oct == 0u) oct = 0xfffffu;
  }
  return oct | (acode << 20);
}

// ---- raygs: a tilted 3D quad facing the eye (projections.tsl.ts:333-429) ----------

// basisToQuat (math.tsl.ts:116-148): the rotation with columns x, y, z
__device__ __forceinline__ float4 basis_to_quat(float3 x, float3 y, float3 z) {
  float4 q;
  if (z.z <= 0.0f) {
    const float dif10 = y.y - x.x, omm22 = 1.0f - z.z;
    if (dif10 <= 0.0f) {
      const float four = omm22 - dif10, inv = rsqrtf(fmaxf(four, 1e-30f)) * 0.5f;
      q = make_float4(four * inv, (x.y + y.x) * inv, (x.z + z.x) * inv, (y.z - z.y) * inv);
    } else {
      const float four = omm22 + dif10, inv = rsqrtf(fmaxf(four, 1e-30f)) * 0.5f;
      q = make_float4((x.y + y.x) * inv, four * inv, (y.z + z.y) * inv, (z.x - x.z) * inv);
    }
  } else {
    const float sum10 = y.y + x.x, opm22 = 1.0f + z.z;
    if (sum10 <= 0.0f) {
      const float four = opm22 - sum10, inv = rsqrtf(fmaxf(four, 1e-30f)) * 0.5f;
      q = make_float4((x.z + z.x) * inv, (y.z + z.y) * inv, four * inv, (x.y - y.x) * inv);
    } else {
      const float four = opm22 + sum10, inv = rsqrtf(fmaxf(four, 1e-30f)) * 0.5f;
      q = make_float4((y.z - z.y) * inv, (z.x - x.z) * inv, (x.y - y.x) * inv, four * inv);
    }
  }
  return q;
}

// One Jacobi rotation of symEigens (math.tsl.ts:57-86) in the plane (P, Q):
// rows then columns of A, then the rows of E (E's row j is eigenvector j).
template <int P, int Q>
__device__ __forceinline__ void jacobi_rotate(double (&A)[3][3], double (&E)[3][3]) {
  const double phi = 0.5 * atan2(2.0 * A[P][Q], A[Q][Q] - A[P][P]);
  double s, c;
  sincos(phi, &s, &c);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double x = A[P][k], y = A[Q][k];
    A[P][k] = c * x - s * y;
    A[Q][k] = s * x + c * y;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double x = A[k][P], y = A[k][Q];
    A[k][P] = c * x - s * y;
    A[k][Q] = s * x + c * y;
  }
  A[P][Q] = 0.0;
  A[Q][P] = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double x = E[P][k], y = E[Q][k];
    E[P][k] = c * x - s * y;
    E[Q][k] = s * x + c * y;
  }
}

// fromCov (splats.tsl.ts:189-198): the scales (square roots of the
// eigenvalues, descending) and the rotation (basisToQuat of the right-handed
// eigenbasis, w >= 0) of a covariance c6 = xx, yy, zz, xy, xz, yz, by
// symEigens (math.tsl.ts:46-113): at most 16 Jacobi rotations, each on the
// largest off-diagonal entry, until it is within 1e-6 of the trace. In
// double: a thin splat's small eigenvalue from float sweeps carries the
// large one's rounding (its relative error times the aspect ratio), and the
// raygs frame's c^2 scales with its inverse.
__device__ __forceinline__ void from_cov(const double* c6, float3& scales, float4& quat) {
  double A[3][3] = {{c6[0], c6[3], c6[4]}, {c6[3], c6[1], c6[5]}, {c6[4], c6[5], c6[2]}};
  double E[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  const double eps = 1e-6 * (fabs(c6[0]) + fabs(c6[1]) + fabs(c6[2]));
#pragma unroll 1
  for (int it = 0; it < 16; it++) {
    int pq = 0;
    double mx = fabs(A[0][1]);
    if (fabs(A[0][2]) > mx) {
      mx = fabs(A[0][2]);
      pq = 1;
    }
    if (fabs(A[1][2]) > mx) {
      mx = fabs(A[1][2]);
      pq = 2;
    }
    if (mx <= eps) break;
    if (pq == 0) jacobi_rotate<0, 1>(A, E);
    else if (pq == 1) jacobi_rotate<0, 2>(A, E);
    else jacobi_rotate<1, 2>(A, E);
  }
  double l[3] = {A[0][0], A[1][1], A[2][2]};
  double v[3][3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const double r = 1.0 / sqrt(E[j][0] * E[j][0] + E[j][1] * E[j][1] + E[j][2] * E[j][2]);
#pragma unroll
    for (int k = 0; k < 3; k++) v[j][k] = E[j][k] * r;
  }
  const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int a = pairs[k][0], b = pairs[k][1];
    if (l[a] < l[b]) {
      const double t = l[a];
      l[a] = l[b];
      l[b] = t;
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const double w = v[a][m];
        v[a][m] = v[b][m];
        v[b][m] = w;
      }
    }
  }
  const double x0 = v[0][1] * v[1][2] - v[0][2] * v[1][1], x1 = v[0][2] * v[1][0] - v[0][0] * v[1][2],
               x2 = v[0][0] * v[1][1] - v[0][1] * v[1][0];
  if (x0 * v[2][0] + x1 * v[2][1] + x2 * v[2][2] < 0.0)
#pragma unroll
    for (int m = 0; m < 3; m++) v[2][m] = -v[2][m];
  scales = make_float3((float)sqrt(fmax(l[0], 0.0)), (float)sqrt(fmax(l[1], 0.0)), (float)sqrt(fmax(l[2], 0.0)));
  // basisToQuat (math.tsl.ts:117-150) of the columns x, y, z
  const double* x = v[0];
  const double* y = v[1];
  const double* z = v[2];
  double q[4];
  if (z[2] <= 0.0) {
    const double dif10 = y[1] - x[0], omm22 = 1.0 - z[2];
    if (dif10 <= 0.0) {
      const double four = omm22 - dif10, inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = four * inv, q[1] = (x[1] + y[0]) * inv, q[2] = (x[2] + z[0]) * inv, q[3] = (y[2] - z[1]) * inv;
    } else {
      const double four = omm22 + dif10, inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = (x[1] + y[0]) * inv, q[1] = four * inv, q[2] = (y[2] + z[1]) * inv, q[3] = (z[0] - x[2]) * inv;
    }
  } else {
    const double sum10 = y[1] + x[0], opm22 = 1.0 + z[2];
    if (sum10 <= 0.0) {
      const double four = opm22 - sum10, inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = (x[2] + z[0]) * inv, q[1] = (y[2] + z[1]) * inv, q[2] = four * inv, q[3] = (x[1] - y[0]) * inv;
    } else {
      const double four = opm22 + sum10, inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = (y[2] - z[1]) * inv, q[1] = (z[0] - x[2]) * inv, q[2] = (x[1] - y[0]) * inv, q[3] = four * inv;
    }
  }
  const double n = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  const double sg = q[3] < 0.0 ? -n : n;
  quat = make_float4((float)(q[0] * sg), (float)(q[1] * sg), (float)(q[2] * sg), (float)(q[3] * sg));
}

// quadOffScreen: every corner mu +- a0 +- a1 in front of the camera and
// all four beyond the same side of the frustum
__device__ __forceinline__ bool quad_off_screen(const float* P, float3 mu, float3 a0, float3 a1) {
  const float4 c = clip_point(P, mu, 1.0f), A0 = clip_point(P, a0, 0.0f), A1 = clip_point(P, a1, 0.0f);
  bool front = true, o0 = true, o1 = true, o2 = true, o3 = true;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float s0 = (i & 1) ? 1.0f : -1.0f, s1 = (i & 2) ? 1.0f : -1.0f;
    const float x = c.x + s0 * A0.x + s1 * A1.x, y = c.y + s0 * A0.y + s1 * A1.y, w = c.w + s0 * A0.w + s1 * A1.w;
    front = front && w > 0.0f;
    o0 = o0 && x < -w;
    o1 = o1 && x > w;
    o2 = o2 && y < -w;
    o3 = o3 && y > w;
  }
  return front && (o0 || o1 || o2 || o3);
}

// raygsFrame and the eye words [frame QuatOct32, float c^2, half2 ln(axis
// lengths)]; false when the camera is inside the splat's support or the
// quad is off screen
__device__ __forceinline__ bool raygs_eye(const EyeU& E, float3 c, float4 q, float3 s, float std_dev, uint32_t* out) {
  const float kappa = std_dev * std_dev;
  const float4 qc = make_float4(-q.x, -q.y, -q.z, q.w);
  float3 lc = quat_rotate(qc, c);
  lc = make_float3(lc.x / s.x, lc.y / s.y, lc.z / s.z);
  float c2 = lc.x * lc.x + lc.y * lc.y + lc.z * lc.z;
  bool ok = true;
  float3 mu;
  if (E.ortho) {
    float3 d = quat_rotate(qc, make_float3(0.0f, 0.0f, -1.0f));
    d = make_float3(d.x / s.x, d.y / s.y, d.z / s.z);
    const float r = rsqrtf(d.x * d.x + d.y * d.y + d.z * d.z);
    mu = make_float3(d.x * r, d.y * r, d.z * r);
    c2 = 1e18f;
  } else {
    ok = c2 > kappa * 1.0001f;
    const float r = rsqrtf(c2);
    mu = make_float3(lc.x * r, lc.y * r, lc.z * r);
  }
  // orthonormal basis of the plane orthogonal to mu (Householder)
  const float3 t = make_float3(mu.x, mu.y, mu.z + (mu.z >= 0.0f ? 1.0f : -1.0f));
  const float inv_den = 2.0f / (t.x * t.x + t.y * t.y + t.z * t.z);
  const float fx = t.x * inv_den, fy = t.y * inv_den;
  const float3 q0 = quat_rotate(q, make_float3(s.x * (t.x * fx - 1.0f), s.y * (t.y * fx), s.z * (t.z * fx)));
  const float3 q1 = quat_rotate(q, make_float3(s.x * (t.x * fy), s.y * (t.y * fy - 1.0f), s.z * (t.z * fy)));
  // major eigenvector of Q^T Q; u0 is it rotated 90 degrees
  const float bxx = q0.x * q0.x + q0.y * q0.y + q0.z * q0.z;
  const float bxy = q0.x * q1.x + q0.y * q1.y + q0.z * q1.z;
  const float byy = q1.x * q1.x + q1.y * q1.y + q1.z * q1.z;
  const float diff = bxx - byy;
  const float eig1 = (bxx + byy) * 0.5f + sqrtf(fmaxf(0.0f, diff * diff * 0.25f + bxy * bxy));
  float vx, vy;
  if (fabsf(bxx - eig1) >= fabsf(byy - eig1)) {
    vx = bxy;
    vy = eig1 - bxx;
  } else {
    vx = eig1 - byy;
    vy = bxy;
  }
  const float vl = sqrtf(vx * vx + vy * vy);
  if (vl > 0.0f) {
    vx /= vl;
    vy /= vl;
  } else {
    vx = 1.0f;
    vy = 0.0f;
  }
  const float u0x = -vy, u0y = vx;
  const float3 a0 = make_float3(q0.x * u0x + q1.x * u0y, q0.y * u0x + q1.y * u0y, q0.z * u0x + q1.z * u0y);
  const float3 a1 = make_float3(q0.x * vx + q1.x * vy, q0.y * vx + q1.y * vy, q0.z * vx + q1.z * vy);
  const float lx = sqrtf(a0.x * a0.x + a0.y * a0.y + a0.z * a0.z), ly = sqrtf(a1.x * a1.x + a1.y * a1.y + a1.z * a1.z);
  ok = ok && lx > 0.0f && ly > 0.0f;
  const float zs = std_dev * rsqrtf(fmaxf(1.0f - kappa / c2, 1e-4f));
  ok = ok && !quad_off_screen(E.Pf, c, make_float3(a0.x * zs, a0.y * zs, a0.z * zs),
                              make_float3(a1.x * zs, a1.y * zs, a1.z * zs));
  if (!ok) return false;
  const float3 n0 = make_float3(a0.x / lx, a0.y / lx, a0.z / lx), n1 = make_float3(a1.x / ly, a1.y / ly, a1.z / ly);
  const float3 n2 = make_float3(n0.y * n1.z - n0.z * n1.y, n0.z * n1.x - n0.x * n1.z, n0.x * n1.y - n0.y * n1.x);
  out[0] = encode_quat_oct32(basis_to_quat(n0, n1, n2));
  out[1] = fbits(c2);
  out[2] = half_bits(logf(lx)) | (half_bits(logf(ly)) << 16);
  return true;
}

// ---- the kernel ----------------------------------------------------------------

// NU: the object transform is not a rotation times a uniform scale (the
// head-space splat is fromCov of H Σ H^T); a compile-time choice, so the
// common kernels carry none of its registers.
// One splat per thread: prepareSplat, the head basis with MIP inflation
// (projections.tsl.ts:172-180, projection.tsl.ts:169-197), per eye the
// bound (cullDepth only), the words and the key (SplatProjection.ts:131-151;
// a splat-form sortGuard, projection.tsl.ts:145-155).
template <int METHOD, int LAYOUT, int SHC, int DEG, int EYES, bool NU>
__global__ __launch_bounds__(256) void k_project_ray(TableP t, RayU r, RayOut o) {
  const ProjU& u = r.u;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= t.count) return;
  const int deg = DEG >= 0 ? DEG : u.sh_degree;
  F32Rsrc rs;
  if constexpr (LAYOUT == LAYOUT_F32) rs = f32_rsrc<SHC>(t, deg);
  Splat sp;
  float3 coef[15];
  float band[3];
  load_splat<LAYOUT, SHC, DEG>(t, rs, u.enc, i, deg, sp, coef, band);
  if constexpr (LAYOUT == LAYOUT_COV) {
    // the splat form of a covSplats splat (GpuCovSplatReader.loadSplat, splatAccess.ts:103-105)
    from_cov(sp.cv, sp.s, sp.q);
  }
  float op, std_dev, inflate;
  float3 s;
  bool invisible;
  prepare<false>(sp, u, op, s, std_dev, invisible, inflate);
  const float3 rgb = sh_colour(sp, u.h, deg, coef, SHC == SHC_PACKED ? band : nullptr);

  // head-space centre (double) and its distance, as the ewa kernel takes them
  double hc[3], head_dist;
  if (u.head_is_eye0) {
    const EyeCentre C0 = eye_centre(u.eye[0], sp.c, u.ndc_z_min, u.clip_xy);
    for (int k = 0; k < 3; k++) hc[k] = C0.ce[k];
    head_dist = C0.dist;
  } else {
    const double dx = (double)sp.c.x - u.h[0], dy = (double)sp.c.y - u.h[1], dz = (double)sp.c.z - u.h[2];
    for (int k = 0; k < 3; k++) hc[k] = fma(u.H[3 * k], dx, fma(u.H[3 * k + 1], dy, u.H[3 * k + 2] * dz));
    head_dist = sqrt(fma(hc[0], hc[0], fma(hc[1], hc[1], hc[2] * hc[2])));
  }
  const float3 hcf = make_float3((float)hc[0], (float)hc[1], (float)hc[2]);

  // viewBasis in the head frame: the head-space rotation and scales, flat
  // splats kept a sliver thick
  const float4 qn = quat_normalize(sp.q);
  float4 hq;
  float3 S;
  if constexpr (NU) {
    // transformSplat's general case (splatTransform.tsl.ts:154-165): fromCov
    // of H Σ H^T, Σ the object-space covariance, then grown by the LoD inflation
    float R[9];
    quat_matrix(qn, R);
    double M[9];  // H R diag(s)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double sk = k == 0 ? sp.s.x : (k == 1 ? sp.s.y : sp.s.z);
        M[3 * a + k] = (u.H[3 * a] * R[k] + u.H[3 * a + 1] * R[3 + k] + u.H[3 * a + 2] * R[6 + k]) * sk;
      }
    double c6[6];
    const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
#pragma unroll
    for (int e = 0; e < 6; e++)
      c6[e] = M[3 * ij[e][0]] * M[3 * ij[e][1]] + M[3 * ij[e][0] + 1] * M[3 * ij[e][1] + 1] +
              M[3 * ij[e][0] + 2] * M[3 * ij[e][1] + 2];
    float3 sh;
    from_cov(c6, sh, hq);
    S = make_float3(sh.x * inflate, sh.y * inflate, sh.z * inflate);
  } else {
    hq = quat_normalize(quat_mul(make_float4(r.hq[0], r.hq[1], r.hq[2], r.hq[3]), qn));
    S = make_float3(s.x * r.oscale, s.y * r.oscale, s.z * r.oscale);
  }
  const float3 s_guard = S;
  {
    const float floor_s = fmaxf(S.x, fmaxf(S.y, S.z)) * 1e-4f;
    S = make_float3(fmaxf(S.x, floor_s), fmaxf(S.y, floor_s), fmaxf(S.z, floor_s));
  }

  // mipInflate with the head view (eye 0's lens)
  float oph = op;
  {
    const EyeU& E0 = u.eye[0];
    const float3 local = quat_rotate(make_float4(-hq.x, -hq.y, -hq.z, hq.w), hcf);
    const float l2x = local.x * local.x, l2y = local.y * local.y, l2z = local.z * local.z;
    const float pixel = 2.0f / (u.size[1] * E0.Pf[5]);
    const float footprint = pixel * pixel * (E0.ortho ? 1.0f : hcf.x * hcf.x + hcf.y * hcf.y + hcf.z * hcf.z);
    float blur = u.blur;
    if (u.focal_distance > 0.0f && u.aperture_tan > 0.0f && hcf.z < 0.0f) {
      const float depth = -hcf.z;
      const float coc_view = fabsf(depth - u.focal_distance) * u.aperture_tan;
      const float coc = E0.ortho ? coc_view * E0.fy : coc_view * E0.fy / depth;
      blur = fminf(fmaxf(coc * coc, blur), u.max_pr * u.max_pr);
    }
    const float s2x = S.x * S.x + u.pre_blur * footprint, s2y = S.y * S.y + u.pre_blur * footprint,
                s2z = S.z * S.z + u.pre_blur * footprint;
    const float bx = s2x + blur * footprint, by = s2y + blur * footprint, bz = s2z + blur * footprint;
    const float num = l2x * s2y * s2z + l2y * s2x * s2z + l2z * s2x * s2y;
    const float den = l2x * by * bz + l2y * bx * bz + l2z * bx * by;
    oph *= den > 0.0f ? sqrtf(fmaxf(num / den, 0.0f)) : 1.0f;
    S = make_float3(sqrtf(bx), sqrtf(by), sqrtf(bz));
  }
  const bool ok = !invisible && oph >= u.min_alpha;

  // words: 6 head + 2 per eye (3dgut, htgs) or 4 head + 3 per eye (raygs)
  constexpr int HEAD = METHOD == METHOD_RAYGS ? 4 : 6;
  constexpr int EW = METHOD == METHOD_RAYGS ? 3 : 2;
  uint32_t w[10];
#pragma unroll
  for (int k = 0; k < 10; k++) w[k] = 0u;
  bool visible = false;
#pragma unroll
  for (int e = 0; e < EYES; e++) {
    const EyeU& E = u.eye[e];
    const EyeCentre C = eye_centre(E, sp.c, u.ndc_z_min, u.clip_xy);
    const float z = (float)C.ce[2];
    const bool cull_depth = z >= 0.0f || C.clip[2] < u.ndc_z_min * C.clip[3] || C.clip[2] >= C.clip[3];
    if (ok && !cull_depth) {
      const float3 c = make_float3((float)C.ce[0], (float)C.ce[1], z);
      const float4 eq = NU ? quat_normalize(quat_mul(make_float4(r.ehq[e][0], r.ehq[e][1], r.ehq[e][2], r.ehq[e][3]), hq))
                           : quat_normalize(quat_mul(make_float4(r.eq[e][0], r.eq[e][1], r.eq[e][2], r.eq[e][3]), qn));
      if constexpr (METHOD == METHOD_RAYGS) {
        if (raygs_eye(E, c, eq, S, std_dev, w + HEAD + EW * e)) visible = true;
      } else {
        RayBound bd = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        bool bok;
        if constexpr (METHOD == METHOD_3DGUT) {
          bok = gut_bound(E.Pf, u.size, c, eq, S, std_dev, u, bd);
        } else {
          const float2 ndc =
              C.clip[3] > 0.0f ? make_float2(C.clip[0] / C.clip[3], C.clip[1] / C.clip[3]) : make_float2(0.0f, 0.0f);
          bok = htgs_bound(E.Pf, u.size, c, eq, S, std_dev, ndc, u, bd);
        }
        if (bok && clip_bound_to_screen(u.size, bd)) {
          const uint2 bw = encode_bound(bd);
          w[HEAD + EW * e] = bw.x;
          w[HEAD + EW * e + 1] = bw.y;
          visible = true;
        }
      }
    }
  }
  if (ok) {
    uint32_t dir, dist, low;
    encode_center(hcf, dir, dist, low);
    const float stored = oph > 1.0f ? (oph + 3.0f) * 0.25f : oph;
    w[0] = dir;
    w[1] = dist;
    w[2] = encode_ext_rgb(rgb.x, rgb.y, rgb.z);
    if constexpr (METHOD == METHOD_RAYGS) {
      w[3] = encode_opacity8(stored) | low;
    } else {
      w[3] = encode_quat_oct32(hq);
      w[4] = half_bits(logf(S.x)) | (half_bits(logf(S.y)) << 16);
      w[5] = half_bits(logf(S.z)) | (encode_opacity8(stored) << 16) | low;
    }
  }
  const size_t at = (size_t)o.at + i;
  reinterpret_cast<uint4*>(o.ch[0])[at] = make_uint4(w[0], w[1], w[2], w[3]);
  reinterpret_cast<uint4*>(o.ch[1])[at] = make_uint4(w[4], w[5], w[6], w[7]);
  if constexpr (EYES == 2) reinterpret_cast<uint2*>(o.ch[2])[at] = make_uint2(w[8], w[9]);

  if (o.keys) {
    bool guarded = true;
    if (!u.cull_keys && u.sort_guard > 0.0f) {
      // inSortGuard for the splat form: the reach is the largest scale
      const float radius = fmaxf(s_guard.x, fmaxf(s_guard.y, s_guard.z)) * u.max_std_dev;
      const float* P0 = u.eye[0].Pf;
      const float4 hcl = clip_point(P0, hcf, 1.0f);
      const float wv = fmaxf(hcl.w, 1e-6f);
      const float rx = P0[0] * radius / wv, ry = P0[5] * radius / wv;
      const bool inside = fabsf(hcl.x / wv) - rx <= u.sort_guard + 1.0f && fabsf(hcl.y / wv) - ry <= u.sort_guard + 1.0f;
      const bool near = -hcf.z <= radius;
      guarded = hcf.z < radius && (inside || near);
    }
    const bool keyed = u.cull_keys ? visible : (!invisible && guarded);
    o.keys[o.at + i] = keyed ? sort_key(fbits((float)head_dist), u.sort_bits) : key_invalid(u.sort_bits);
  }
}

}  // namespace rdx

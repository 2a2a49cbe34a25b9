// The draw of the ray-evaluated projection methods (3dgut, htgs, raygs): the
// reference draws a quad per splat whose vertex stage turns the method's
// words into varyings and whose fragment stage evaluates the method's
// density (src/render/splatDraw.ts:98-181 with projections.tsl.ts:310-330,
// 432-466 and projection.tsl.ts:396-442). The varyings are affine in the
// quad's screen position (3dgut, htgs: localDir and localCross; raygs: the
// homogeneous quad coordinates, perspective-correct), so each is written
// here as an affine function of the pixel and evaluated at the pixel centre
// by the compositor's tile renderer (ray_z2 in raster_kernels.hpp).
#pragma once

#include "draw_kernels.hpp"
#include "ray_kernels.hpp"

namespace rdx {

// The eye being drawn: head -> eye (SplatCamera's fromHead, row-major 3x4)
// and its rotation, the eye's projection (row-major), and its words.
struct RayDrawU {
  float fh[12];
  float fq[4];  // (x, y, z, w)
  float P[16];
  int ortho;
  int eye;
  int eyes;
};

// decodeLogScale for `bits` (projection.tsl.ts:305-307): log2 pixel scales on [-4, 12]
__device__ __forceinline__ float decode_log_scale(uint32_t code, int bits) {
  return exp2f((float)(code - 1u) / (float)((1u << bits) - 2u) * 16.0f - 4.0f);
}

// decodeCenter (projection.tsl.ts:335-349): the octahedral direction of 20 +
// 20 bits (16 + 16 in `dir`, the low 4 + 4 in bits 24-31 of `last`) times
// the float distance
__device__ __forceinline__ float3 decode_center(uint32_t dir, uint32_t dist, uint32_t last) {
  const uint32_t qx = ((dir & 0xffffu) << 4) | ((last >> 24) & 15u), qy = ((dir >> 16) << 4) | (last >> 28);
  float x = (float)qx * (2.0f / 1048575.0f) - 1.0f, y = (float)qy * (2.0f / 1048575.0f) - 1.0f;
  const float z = 1.0f - fabsf(x) - fabsf(y);
  const float t = fmaxf(-z, 0.0f);
  x += x >= 0.0f ? -t : t;
  y += y >= 0.0f ? -t : t;
  const float d = __uint_as_float(dist) * rsqrtf(x * x + y * y + z * z);
  return make_float3(x * d, y * d, z * d);
}

__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 div3(float3 a, float3 s) { return make_float3(a.x / s.x, a.y / s.y, a.z / s.z); }

// Draw position i (back to front) of m: splat order[i]'s words of the eye,
// as the record the renderer reads (7 float4, RAY_REC), its touched tiles
// and rectangle. Positions past n, and splats this eye culls (a zero second
// eye word: look.visible), are not drawn.
template <int METHOD>
__global__ __launch_bounds__(256) void k_draw_ray_pre(const uint4* __restrict__ ch0, const uint4* __restrict__ ch1,
                                                      const uint2* __restrict__ ch2, uint32_t n,
                                                      const uint32_t* __restrict__ order, const uint32_t* m_dev,
                                                      DiffU u, RayDrawU e, int linear, float4* __restrict__ pre,
                                                      uint32_t* __restrict__ touch, uint4* __restrict__ rect) {
  constexpr bool RAYGS = METHOD == METHOD_RAYGS;
  constexpr int HEAD = RAYGS ? 4 : 6, EW = RAYGS ? 3 : 2;
  const uint32_t m = *m_dev;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  touch[i] = 0u;
  const uint32_t sp = order[i];
  if (sp >= n) return;
  uint32_t w[10];
  {
    const uint4 a = ch0[sp], b = ch1[sp];
    w[0] = a.x, w[1] = a.y, w[2] = a.z, w[3] = a.w, w[4] = b.x, w[5] = b.y, w[6] = b.z, w[7] = b.w;
    w[8] = w[9] = 0u;
    if (e.eyes == 2) {
      const uint2 c = ch2[sp];
      w[8] = c.x, w[9] = c.y;
    }
  }
  uint32_t ew[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    // the eye's words: one eye or the other, by a select over both (the
    // indices stay constant, so w stays in registers)
    const uint32_t a0 = k < EW ? w[HEAD + k] : 0u, a1 = k < EW ? w[HEAD + EW + k < 10 ? HEAD + EW + k : 0] : 0u;
    ew[k] = e.eye == 0 ? a0 : a1;
  }
  if (ew[1] == 0u) return;  // look.visible
  const uint32_t last = RAYGS ? w[3] : w[5];
  const uint32_t byte = RAYGS ? (w[3] & 0xffu) : ((w[5] >> 16) & 0xffu);
  // the LoD / cutoff rule (splatDraw.ts:127-140)
  float alpha = (float)byte * (1.0f / 128.0f);
  float bound = u.max_std_dev, cut = u.max_std_dev;
  float op;
  if (alpha > 1.0f) {
    alpha = fminf(alpha * 4.0f - 3.0f, 5.0f);
    bound = u.max_std_dev + (alpha - 1.0f) * 0.7f;
    cut = bound;
    op = -expf((alpha * alpha - 1.0f) * 0.36787944117144233f);  // -a, a = exp((alpha^2 - 1) / e)
  } else {
    if (u.falloff >= 1.0f) cut = fminf(cut, sqrtf(logf(fmaxf(alpha, u.min_alpha) / u.min_alpha) * 2.0f));
    op = alpha;
  }
  float3 col = decode_ext_rgb(w[2]);
  col = make_float3(fmaxf(col.x, 0.0f), fmaxf(col.y, 0.0f), fmaxf(col.z, 0.0f));
  if (linear) col = make_float3(srgb_to_linear(col.x), srgb_to_linear(col.y), srgb_to_linear(col.z));
  // the centre in the eye (fromHead)
  const float3 hc = decode_center(w[0], w[1], last);
  const float3 c = make_float3(fmaf(e.fh[0], hc.x, fmaf(e.fh[1], hc.y, fmaf(e.fh[2], hc.z, e.fh[3]))),
                               fmaf(e.fh[4], hc.x, fmaf(e.fh[5], hc.y, fmaf(e.fh[6], hc.z, e.fh[7]))),
                               fmaf(e.fh[8], hc.x, fmaf(e.fh[9], hc.y, fmaf(e.fh[10], hc.z, e.fh[11]))));
  const float W = (float)u.width, H = (float)u.height;
  // view position under the pixel (px, py) at the centre's depth (unproject,
  // projection.tsl.ts:386-393): affine in the pixel, d/dpx = (kx, 0, 0),
  // d/dpy = (0, ky, 0) (framebuffer y runs down)
  const float wz = fmaf(e.P[14], c.z, e.P[15]);
  const float kx = 2.0f * wz / (W * e.P[0]), ky = -2.0f * wz / (H * e.P[5]);
  auto view_at = [&](float px, float py) {
    const float nx = px * (2.0f / W) - 1.0f, ny = 1.0f - py * (2.0f / H);
    return make_float3((nx * wz - e.P[2] * c.z - e.P[3]) / e.P[0], (ny * wz - e.P[6] * c.z - e.P[7]) / e.P[5], c.z);
  };
  float4 r[RAY_REC];
#pragma unroll
  for (int k = 0; k < RAY_REC; k++) r[k] = make_float4(0.f, 0.f, 0.f, 0.f);
  float cx, cy, ex, ey;
  if constexpr (!RAYGS) {
    const float4 q = quat_normalize(quat_mul(make_float4(e.fq[0], e.fq[1], e.fq[2], e.fq[3]), quat_oct_decode(w[3], 10, 12)));
    const float3 sc = make_float3(expf(half_to_float(w[4])), expf(half_to_float(w[4] >> 16)), expf(half_to_float(w[5])));
    // decodeBound (projection.tsl.ts:362-368) and the quad of rayVertex (:412-435)
    const float bnx = decode_fixed16(ew[0] & 0xffffu), bny = decode_fixed16(ew[0] >> 16);
    const float ang = (float)(ew[1] & 1023u) * (3.14159265358979323846f / 1024.0f);
    const float bsx = decode_log_scale((ew[1] >> 10) & 0x7ffu, 11), bsy = decode_log_scale(ew[1] >> 21, 11);
    float sn, cs;
    sincosf(ang, &sn, &cs);
    const float hx = W * 0.5f, hy = H * 0.5f;
    // axis1 = (cos, sin), axis2 = (sin, -cos) in ndc (y up)
    const float ext_x = sqrtf(cs * bsx * cs * bsx + sn * bsy * sn * bsy);
    const float ext_y = sqrtf(sn * bsx * sn * bsx + cs * bsy * cs * bsy);
    const bool on_screen = fabsf(bnx * hx) + ext_x < hx && fabsf(bny * hy) + ext_y < hy;
    const float corner = on_screen ? (METHOD == METHOD_3DGUT ? cut / bound : 1.0f) : 1.0f;
    cx = (bnx + 1.0f) * hx;
    cy = (1.0f - bny) * hy;
    // in framebuffer pixels (y down): u bsx = dx cos - dy sin, v bsy = dx sin + dy cos
    r[1] = make_float4(cs, -sn, bsx * corner, 0.0f);
    const float qy = bsy * corner;
    // the box of the quad (the rectangle +-qx axis1 +-qy axis2), a pixel wider
    ex = fabsf(cs) * bsx * corner + fabsf(sn) * qy + 1.0f;
    ey = fabsf(sn) * bsx * corner + fabsf(cs) * qy + 1.0f;
    // the ray in the splat's whitened frame, affine in the pixel
    const float4 qc = make_float4(-q.x, -q.y, -q.z, q.w);
    const float3 v0 = view_at(cx, cy);
    float3 LC0, LCx, LCy, LD0, LDx, LDy;
    if (!e.ortho) {
      // origin 0, direction the view position
      const float3 lo = div3(quat_rotate(qc, make_float3(-c.x, -c.y, -c.z)), sc);
      LD0 = div3(quat_rotate(qc, v0), sc);
      LDx = div3(quat_rotate(qc, make_float3(kx, 0.0f, 0.0f)), sc);
      LDy = div3(quat_rotate(qc, make_float3(0.0f, ky, 0.0f)), sc);
      LC0 = cross3(lo, LD0);
      LCx = cross3(lo, LDx);
      LCy = cross3(lo, LDy);
    } else {
      // origin (view x, y, 0), direction -z
      const float3 ld = div3(quat_rotate(qc, make_float3(0.0f, 0.0f, -1.0f)), sc);
      const float3 lo0 = div3(quat_rotate(qc, make_float3(v0.x - c.x, v0.y - c.y, -c.z)), sc);
      const float3 lox = div3(quat_rotate(qc, make_float3(kx, 0.0f, 0.0f)), sc);
      const float3 loy = div3(quat_rotate(qc, make_float3(0.0f, ky, 0.0f)), sc);
      LD0 = ld;
      LDx = LDy = make_float3(0.0f, 0.0f, 0.0f);
      LC0 = cross3(lo0, ld);
      LCx = cross3(lox, ld);
      LCy = cross3(loy, ld);
    }
    // |L0 + dx Lx + dy Ly|^2 = |L0|^2 + 2 dx L0.Lx + 2 dy L0.Ly + dx^2 |Lx|^2
    // + 2 dx dy Lx.Ly + dy^2 |Ly|^2 for localCross (N) and localDir (D)
    r[3] = make_float4(qy, dot3(LC0, LC0), 2.0f * dot3(LC0, LCx), 2.0f * dot3(LC0, LCy));
    r[4] = make_float4(dot3(LCx, LCx), 2.0f * dot3(LCx, LCy), dot3(LCy, LCy), dot3(LD0, LD0));
    r[5] = make_float4(2.0f * dot3(LD0, LDx), 2.0f * dot3(LD0, LDy), dot3(LDx, LDx), 2.0f * dot3(LDx, LDy));
    r[6] = make_float4(dot3(LDy, LDy), 0.0f, 0.0f, 0.0f);
  } else {
    // the facing quad (raygs vertex, projections.tsl.ts:442-460): frame, c^2, axis lengths
    const float4 q = quat_oct_decode(ew[0], 10, 12);
    const float c2 = __uint_as_float(ew[1]);
    const float lx = expf(half_to_float(ew[2])), ly = expf(half_to_float(ew[2] >> 16));
    const float3 n0 = quat_rotate(q, make_float3(1.0f, 0.0f, 0.0f)), n1 = quat_rotate(q, make_float3(0.0f, 1.0f, 0.0f)),
                 nr = quat_rotate(q, make_float3(0.0f, 0.0f, 1.0f));
    const float zmax = cut * rsqrtf(fmaxf(1.0f - cut * cut / c2, 1e-4f));
    // the box of the quad's corners c + (+-zmax lx) n0 + (+-zmax ly) n1
    // around the centre's pixel; a corner behind the eye: the whole screen
    const float4 cc = clip_point(e.P, c, 1.0f);
    cx = (cc.x / cc.w + 1.0f) * 0.5f * W;
    cy = (1.0f - cc.y / cc.w) * 0.5f * H;
    {
      float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
      bool front = cc.w > 0.0f;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float a = (k & 1) ? zmax * lx : -zmax * lx, b = (k & 2) ? zmax * ly : -zmax * ly;
        const float3 p = make_float3(c.x + a * n0.x + b * n1.x, c.y + a * n0.y + b * n1.y, c.z + a * n0.z + b * n1.z);
        const float4 k4 = clip_point(e.P, p, 1.0f);
        front = front && k4.w > 0.0f;
        const float px = (k4.x / k4.w + 1.0f) * 0.5f * W, py = (1.0f - k4.y / k4.w) * 0.5f * H;
        x0 = fminf(x0, px), x1 = fmaxf(x1, px), y0 = fminf(y0, py), y1 = fmaxf(y1, py);
      }
      if (front) {
        ex = fmaxf(x1 - cx, cx - x0) + 1.0f;
        ey = fmaxf(y1 - cy, cy - y0) + 1.0f;
      } else {
        cx = 0.5f * W;
        cy = 0.5f * H;
        ex = 0.5f * W + 1.0f;
        ey = 0.5f * H + 1.0f;
      }
    }
    // z = (U, V) / W, the hit of the pixel's ray on the quad's plane, affine
    // about (cx, cy): perspective (origin 0, direction D(p)):
    // U lx = (c.n) (D.n0) - (c.n0) (D.n), W = D.n; orthographic (origin
    // O(p) = (view x, y, 0), D = -z): U lx = (O - c).n0 (D.n) + ((c - O).n) (D.n0), W = D.n
    const float3 v0 = view_at(cx, cy);
    float U0, Ux, Uy, V0, Vx, Vy, W0, Wx, Wy;
    if (!e.ortho) {
      const float cn = dot3(c, nr), c0 = dot3(c, n0), c1 = dot3(c, n1);
      const float3 dx = make_float3(kx, 0.0f, 0.0f), dy = make_float3(0.0f, ky, 0.0f);
      const float sg = cn > 0.0f ? 1.0f : -1.0f;  // the hit lies in front: t = (c.n) / W > 0
      U0 = sg * (cn * dot3(v0, n0) - c0 * dot3(v0, nr)) / lx;
      Ux = sg * (cn * dot3(dx, n0) - c0 * dot3(dx, nr)) / lx;
      Uy = sg * (cn * dot3(dy, n0) - c0 * dot3(dy, nr)) / lx;
      V0 = sg * (cn * dot3(v0, n1) - c1 * dot3(v0, nr)) / ly;
      Vx = sg * (cn * dot3(dx, n1) - c1 * dot3(dx, nr)) / ly;
      Vy = sg * (cn * dot3(dy, n1) - c1 * dot3(dy, nr)) / ly;
      W0 = sg * dot3(v0, nr);
      Wx = sg * dot3(dx, nr);
      Wy = sg * dot3(dy, nr);
    } else {
      const float3 D = make_float3(0.0f, 0.0f, -1.0f);
      const float dn = dot3(D, nr), d0 = dot3(D, n0), d1 = dot3(D, n1);
      const float sg = dn >= 0.0f ? 1.0f : -1.0f;
      const float3 o0 = make_float3(v0.x - c.x, v0.y - c.y, -c.z);  // O - c at (cx, cy)
      const float3 ox = make_float3(kx, 0.0f, 0.0f), oy = make_float3(0.0f, ky, 0.0f);
      U0 = sg * (dot3(o0, n0) * dn - dot3(o0, nr) * d0) / lx;
      Ux = sg * (dot3(ox, n0) * dn - dot3(ox, nr) * d0) / lx;
      Uy = sg * (dot3(oy, n0) * dn - dot3(oy, nr) * d0) / lx;
      V0 = sg * (dot3(o0, n1) * dn - dot3(o0, nr) * d1) / ly;
      Vx = sg * (dot3(ox, n1) * dn - dot3(ox, nr) * d1) / ly;
      Vy = sg * (dot3(oy, n1) * dn - dot3(oy, nr) * d1) / ly;
      W0 = sg * dn;
      Wx = Wy = 0.0f;
    }
    r[1] = make_float4(zmax, c2, 0.0f, 0.0f);
    r[3] = make_float4(U0, Ux, Uy, V0);
    r[4] = make_float4(Vx, Vy, W0, Wx);
    r[5] = make_float4(Wy, 0.0f, 0.0f, 0.0f);
  }
  r[0] = make_float4(cx, cy, op, cut * cut);
  r[1].w = ex;
  r[2] = make_float4(col.x, col.y, col.z, ey);
#pragma unroll
  for (int k = 0; k < RAY_REC; k++) pre[RAY_REC * (size_t)i + k] = r[k];
  // tiles of pixel centres within the box
  const float bx0 = cx - ex, bx1 = cx + ex, by0 = cy - ey, by1 = cy + ey;
  if (!(bx1 >= 0.0f && by1 >= 0.0f && bx0 < W && by0 < H)) return;
  const int tx0 = max(0, (int)floorf(bx0 / DIFF_TILE)), tx1 = min(u.tiles_x - 1, (int)floorf(bx1 / DIFF_TILE));
  const int ty0 = max(0, (int)floorf(by0 / DIFF_TILE)), ty1 = min(u.tiles_y - 1, (int)floorf(by1 / DIFF_TILE));
  rect[i] = make_uint4((uint32_t)tx0, (uint32_t)ty0, (uint32_t)tx1, (uint32_t)ty1);
  touch[i] = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
}

}  // namespace rdx

// The splat draw in production form: the reference draws a quad per splat in
// the sorted order and blends the ewa method's fragment premultiplied "over"
// the target (src/render/splatDraw.ts:98-181, projections.tsl.ts:150-167,
// projection.tsl.ts:396-405, 444-455). Here the same fragments are
// composited per 16x16 tile by the compositor's renderer: one record per
// draw position, decoded from the eye's ewa words, then binning in draw
// order, a stable sort of the entries by tile and front-to-back compositing
// (equal to back-to-front "over" up to rounding).
#pragma once

#include "project_kernels.hpp"
#include "raster_kernels.hpp"

namespace rdx {

// The word decoders (projection.tsl.ts:277-311).
__device__ __forceinline__ float decode_fixed16(uint32_t code) { return (float)code * (4.0f / 65535.0f) - 2.0f; }
__device__ __forceinline__ float decode_log_scale12(uint32_t code) {
  return exp2f((float)(code - 1u) * (16.0f / 4094.0f) - 4.0f);
}

// sRGBTransferEOTF: the colour of an output that stores linear values.
__device__ __forceinline__ float srgb_to_linear(float c) {
  return c <= 0.04045f ? c * (1.0f / 12.92f) : powf((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}

// counts[C_VISIBLE] = min(*count_dev, capacity); the other words cleared.
static __global__ void k_draw_counts(const uint32_t* count_dev, uint32_t capacity, uint32_t* counts) {
  if (threadIdx.x < 4) counts[threadIdx.x] = threadIdx.x == 0 ? min(*count_dev, capacity) : 0u;
}

// Draw position i (back to front) of m: splat order[i]'s words. Indices at
// or past n are not drawn (the vertex stage's guard, splatDraw.ts:121-122),
// nor are words with a zero opacity byte (culled in this eye). The record,
// in the renderer's layout, is in framebuffer pixels (row 0 at the top,
// centres at + 0.5):
//   (cx, cy, opacity or -a, cut^2), (Q0, 2 Q1, Q2, ex), (r, g, b, ey)
// with z^2 = d^T Q d = bound^2 ((d.a1 / sx)^2 + (d.a2 / sy)^2) the squared
// radius of the quad's uv at the pixel (uv = corner * cut, the quad spanning
// sx * cut / bound along a1 and sy * cut / bound along a2), and (ex, ey) the
// widened half extents of the ellipse z^2 <= cut^2.
static __global__ __launch_bounds__(256) void k_draw_pre(const uint4* __restrict__ words, uint32_t n,
                                                         const uint32_t* __restrict__ order, const uint32_t* m_dev,
                                                         DiffU u, int linear, float4* __restrict__ pre,
                                                         uint32_t* __restrict__ touch, uint4* __restrict__ rect) {
  const uint32_t m = *m_dev;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  touch[i] = 0u;
  const uint32_t sp = order[i];
  if (sp >= n) return;
  const uint4 W = words[sp];
  const uint32_t byte = W.w & 0xffu;
  if (byte == 0u) return;
  // look (projections.tsl.ts:150-153) and the LoD / cutoff rule (splatDraw.ts:127-140)
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
  float3 c = decode_ext_rgb(W.z);
  c = make_float3(fmaxf(c.x, 0.0f), fmaxf(c.y, 0.0f), fmaxf(c.z, 0.0f));
  if (linear) c = make_float3(srgb_to_linear(c.x), srgb_to_linear(c.y), srgb_to_linear(c.z));
  // the bound (projections.tsl.ts:155-165); framebuffer y runs down, ndc y up
  const float nx = decode_fixed16(W.x & 0xffffu), ny = decode_fixed16(W.x >> 16);
  const float ang = (float)(W.y >> 20) * (3.14159265358979323846f / 4096.0f);
  const float sx = decode_log_scale12((W.w >> 8) & 0xfffu), sy = decode_log_scale12(W.w >> 20);
  float sn, cs;
  sincosf(ang, &sn, &cs);
  const float a1x = cs, a1y = -sn;  // axis1 = (cos, sin), y flipped
  const float a2x = sn, a2y = cs;   // axis2 = (sin, -cos), y flipped
  const float cx = (nx + 1.0f) * 0.5f * (float)u.width, cy = (1.0f - ny) * 0.5f * (float)u.height;
  const float b2 = bound * bound, ix = 1.0f / (sx * sx), iy = 1.0f / (sy * sy);
  const float Q0 = b2 * (a1x * a1x * ix + a2x * a2x * iy);
  const float Q1 = b2 * (a1x * a1y * ix + a2x * a2y * iy);
  const float Q2 = b2 * (a1y * a1y * ix + a2y * a2y * iy);
  // the ellipse z^2 <= cut^2: semi-axes sx * cut / bound along a1, sy * cut / bound along a2
  const float k = cut / bound, px = sx * k, py = sy * k;
  const float hx = sqrtf(a1x * a1x * px * px + a2x * a2x * py * py), hy = sqrtf(a1y * a1y * px * px + a2y * a2y * py * py);
  const float ex = hx * 1.02f + 0.5f, ey = hy * 1.02f + 0.5f;
  pre[3 * (size_t)i] = make_float4(cx, cy, op, cut * cut);
  pre[3 * (size_t)i + 1] = make_float4(Q0, 2.0f * Q1, Q2, ex);
  pre[3 * (size_t)i + 2] = make_float4(c.x, c.y, c.z, ey);
  // tiles of pixel centres within the widened box
  const float x0 = cx - ex, x1 = cx + ex, y0 = cy - ey, y1 = cy + ey;
  if (x1 < 0.0f || y1 < 0.0f || x0 >= (float)u.width || y0 >= (float)u.height) return;
  const int tx0 = max(0, (int)floorf(x0 / DIFF_TILE)), tx1 = min(u.tiles_x - 1, (int)floorf(x1 / DIFF_TILE));
  const int ty0 = max(0, (int)floorf(y0 / DIFF_TILE)), ty1 = min(u.tiles_y - 1, (int)floorf(y1 / DIFF_TILE));
  rect[i] = make_uint4((uint32_t)tx0, (uint32_t)ty0, (uint32_t)tx1, (uint32_t)ty1);
  touch[i] = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
}

// One block: each tile's number of chunks (at least one: an empty tile's
// block writes its zero pixels), their exclusive scan -> item_start, and
// the work items (tile, chunk) in tile order; the item count to *items.
static __global__ __launch_bounds__(1024) void k_draw_items(const uint2* range, uint32_t tiles, uint32_t* item_start,
                                                             uint32_t* item_tile, uint32_t* item_chunk, uint32_t* items) {
  __shared__ uint32_t s_scan[1024 / RDX_WAVE];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < tiles; b0 += 1024) {
    const uint32_t t = b0 + threadIdx.x;
    uint32_t k = 0;
    if (t < tiles) {
      const uint2 r = range[t];
      k = max(1u, (r.y - r.x + DRAW_CHUNK - 1) / DRAW_CHUNK);
    }
    uint32_t sum;
    const uint32_t at = carry + block_exclusive_sum<1024>(k, s_scan, &sum);
    if (t < tiles) {
      item_start[t] = at;
      for (uint32_t c = 0; c < k; c++) {
        item_tile[at + c] = t;
        item_chunk[at + c] = c;
      }
    }
    carry += sum;
  }
  if (threadIdx.x == 0) *items = carry;
}

// A tile of several chunks: its pixels are the chunks' partials folded front
// to back, C = sum_k (prod_{j < k} T_j) C_k, T = prod_k T_k.
static __global__ __launch_bounds__(DIFF_BLOCK) void k_draw_combine(const uint2* range, const uint32_t* item_start,
                                                                     const float4* partial, DiffU u, float* image,
                                                                     float* alpha_out) {
  const uint32_t tile = blockIdx.x;
  const uint2 r = range[tile];
  const uint32_t k = max(1u, (r.y - r.x + DRAW_CHUNK - 1) / DRAW_CHUNK);
  if (k <= 1) return;
  const DiffQuad q = diff_quad(tile, u);
  if (q.x >= u.width || q.y >= u.height) return;
  float C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, T = 1.0f;
  const float4* part = partial + (size_t)item_start[tile] * DIFF_BLOCK + threadIdx.x;
  for (uint32_t c = 0; c < k; c++) {
    const float4 v = part[(size_t)c * DIFF_BLOCK];
    C0 = fmaf(T, v.x, C0);
    C1 = fmaf(T, v.y, C1);
    C2 = fmaf(T, v.z, C2);
    T *= v.w;
  }
  const size_t p = (size_t)q.y * u.width + q.x, plane_px = (size_t)u.width * u.height;
  image[p] = C0;
  image[plane_px + p] = C1;
  image[2 * plane_px + p] = C2;
  alpha_out[p] = 1.0f - T;
}

}  // namespace rdx

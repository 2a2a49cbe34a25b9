// The differentiable compositor for gfx950: front-to-back alpha compositing
// of the ewa splats over 16 x 16 pixel tiles, and its adjoint.
//
// No reference counterpart as a whole: the reference draws the sorted splats
// with rasterised quads and premultiplied "over" blending and has no
// backward (src/render/splatDraw.ts). The compositor follows its fragment:
// alpha = opacity * exp(-z^2 / 2) at falloff 1 (projection.tsl.ts:444-455),
// z^2 the squared Mahalanobis distance under the blurred 2D covariance
// (projections.tsl.ts:155-167), discarded where z^2 > cutoff^2 or alpha <
// minAlpha, cutoff = min(maxStdDev, sqrt(2 ln(opacity / minAlpha)))
// (splatDraw.ts:136-145, 165-174), splats in the sort's order (keys
// ascending = back to front, composited here front to back). Compositing
// stops before the splat that would bring transmittance below t_stop.
//
// Parameters: 59 planes of n floats (the f32 table's planes, with the
// opacity plane holding the LOGIT): centre 3, log-scale 3, quaternion 4
// (x, y, z, w), opacity logit 1, base colour 3, SH bands 1-3 45
// (coefficient-major). Pixels (x + 0.5, y + 0.5), y up; images planar
// [channel][y][x].
//
// Pipeline (raster.hip): k_diff_pre (per splat: centre, conic, opacity,
// cutoff, colour, key, tile rectangle) -> radix sort of the keys (back to
// front) -> scan of the tile counts in front-to-back order -> k_diff_emit
// (one entry per (splat, tile), in front-to-back order) -> stable radix sort
// of the entries by tile -> k_diff_ranges -> k_diff_render (a block per
// tile). Backward: k_diff_render_bwd (a block per tile, back to front from
// each pixel's last contributor, transmittance recomputed by division from
// the final one) and k_diff_pre_bwd (per splat, the sum of its entries'
// partials and the chain rule to the 59 parameters).
#pragma once

#include "common.hpp"

namespace rdx {

constexpr int DIFF_PARAMS = 59;
constexpr int DIFF_TILE = 16;
constexpr int DIFF_BLOCK = DIFF_TILE * DIFF_TILE;
constexpr int DIFF_PIX_GRADS = 9;  // mu2 2, conic 3, opacity 1, colour 3
constexpr int DIFF_BWD_BATCH = 64;  // entries staged per round of the adjoint

struct DiffU {
  double Vl[9], Vt[3], cam[3];  // world -> view (row-major linear part), camera position
  double P[16];                 // projection, row-major
  float fx, fy;                 // renderSize / 2 * (P00, P11)
  int width, height, tiles_x, tiles_y;
  float ndc_z_min, max_std_dev, min_alpha, blur, clip_xy, t_stop;
  int sh_degree;
  float falloff;  // the draw's fragment shaping (SplatParams.falloff); the differentiable form uses 1
};

// ---- SH basis (sh.ts:376-407, the reference's constants) and its gradient ----

__device__ __forceinline__ void sh_basis15(float x, float y, float z, float* o) {
  const float xx = x * x, yy = y * y, zz = z * z;
  o[0] = -0.4886025f * y;
  o[1] = 0.4886025f * z;
  o[2] = -0.4886025f * x;
  o[3] = 1.0925484f * x * y;
  o[4] = -1.0925484f * y * z;
  o[5] = 0.3153916f * (2.0f * zz - xx - yy);
  o[6] = -1.0925484f * x * z;
  o[7] = 0.5462742f * (xx - yy);
  o[8] = -0.5900436f * y * (3.0f * xx - yy);
  o[9] = 2.8906114f * x * y * z;
  o[10] = -0.4570458f * y * (4.0f * zz - xx - yy);
  o[11] = 0.3731763f * z * (2.0f * zz - 3.0f * xx - 3.0f * yy);
  o[12] = -0.4570458f * x * (4.0f * zz - xx - yy);
  o[13] = 1.4453057f * z * (xx - yy);
  o[14] = -0.5900436f * x * (xx - 3.0f * yy);
}

// d(basis)/d(direction): g[k] = (d/dx, d/dy, d/dz) of o[k].
__device__ __forceinline__ void sh_basis15_grad(float x, float y, float z, float3* g) {
  const float xx = x * x, yy = y * y, zz = z * z;
  g[0] = make_float3(0.0f, -0.4886025f, 0.0f);
  g[1] = make_float3(0.0f, 0.0f, 0.4886025f);
  g[2] = make_float3(-0.4886025f, 0.0f, 0.0f);
  g[3] = make_float3(1.0925484f * y, 1.0925484f * x, 0.0f);
  g[4] = make_float3(0.0f, -1.0925484f * z, -1.0925484f * y);
  g[5] = make_float3(-0.6307832f * x, -0.6307832f * y, 1.2615664f * z);
  g[6] = make_float3(-1.0925484f * z, 0.0f, -1.0925484f * x);
  g[7] = make_float3(1.0925484f * x, -1.0925484f * y, 0.0f);
  g[8] = make_float3(-0.5900436f * 6.0f * x * y, -0.5900436f * (3.0f * xx - 3.0f * yy), 0.0f);
  g[9] = make_float3(2.8906114f * y * z, 2.8906114f * x * z, 2.8906114f * x * y);
  g[10] = make_float3(0.9140916f * x * y, -0.4570458f * (4.0f * zz - xx - 3.0f * yy), -3.6563664f * y * z);
  g[11] = make_float3(-2.2390578f * x * z, -2.2390578f * y * z, 0.3731763f * (6.0f * zz - 3.0f * xx - 3.0f * yy));
  g[12] = make_float3(-0.4570458f * (4.0f * zz - 3.0f * xx - yy), 0.9140916f * x * y, -3.6563664f * x * z);
  g[13] = make_float3(2.8906114f * x * z, -2.8906114f * y * z, 1.4453057f * (xx - yy));
  g[14] = make_float3(-0.5900436f * (3.0f * xx - 3.0f * yy), 3.5402616f * x * y, 0.0f);
}

// ---- one splat's forward -------------------------------------------------------

struct DiffSplat {
  bool ok;
  double t[3], dist, clip[4];
  float mu2x, mu2y;
  float qv[4], qlen, R[9], s[3], M[9], Sig[9];
  float J[6], T[6];
  float a, b, d, ap, dp, det_o, det_p, ratio, op, opc, cut;
  float Q0, Q1, Q2;
  float dir[3], dlen, Y[15], col_raw[3], col[3];
};

__device__ __forceinline__ float plane(const float* th, uint32_t n, int k, uint32_t i) { return th[(size_t)k * n + i]; }

__device__ inline void diff_splat_forward(const float* th, uint32_t n, uint32_t i, const DiffU& u, DiffSplat& s) {
  const double mu[3] = {plane(th, n, 0, i), plane(th, n, 1, i), plane(th, n, 2, i)};
#pragma unroll
  for (int r = 0; r < 3; r++) s.t[r] = u.Vl[3 * r] * mu[0] + u.Vl[3 * r + 1] * mu[1] + u.Vl[3 * r + 2] * mu[2] + u.Vt[r];
  s.dist = sqrt(s.t[0] * s.t[0] + s.t[1] * s.t[1] + s.t[2] * s.t[2]);
#pragma unroll
  for (int r = 0; r < 4; r++) s.clip[r] = u.P[4 * r] * s.t[0] + u.P[4 * r + 1] * s.t[1] + u.P[4 * r + 2] * s.t[2] + u.P[4 * r + 3];
  const double w = s.clip[3];
  const bool cull_depth = s.t[2] >= 0.0 || s.clip[2] < (double)u.ndc_z_min * w || s.clip[2] >= w;
  const double limit = fmin((double)u.clip_xy, 2.0) * w;
  const bool cull_xy = fabs(s.clip[0]) > limit || fabs(s.clip[1]) > limit;
  s.mu2x = (float)((s.clip[0] / w + 1.0) * 0.5 * u.width);
  s.mu2y = (float)((s.clip[1] / w + 1.0) * 0.5 * u.height);
  // covariance
  float q[4] = {plane(th, n, 6, i), plane(th, n, 7, i), plane(th, n, 8, i), plane(th, n, 9, i)};
  s.qlen = sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
#pragma unroll
  for (int k = 0; k < 4; k++) s.qv[k] = s.qlen > 0.0f ? q[k] / s.qlen : (k == 3 ? 1.0f : 0.0f);
  const float qx = s.qv[0], qy = s.qv[1], qz = s.qv[2], qw = s.qv[3];
  s.R[0] = 1.0f - 2.0f * (qy * qy + qz * qz);
  s.R[1] = 2.0f * (qx * qy - qw * qz);
  s.R[2] = 2.0f * (qx * qz + qw * qy);
  s.R[3] = 2.0f * (qx * qy + qw * qz);
  s.R[4] = 1.0f - 2.0f * (qx * qx + qz * qz);
  s.R[5] = 2.0f * (qy * qz - qw * qx);
  s.R[6] = 2.0f * (qx * qz - qw * qy);
  s.R[7] = 2.0f * (qy * qz + qw * qx);
  s.R[8] = 1.0f - 2.0f * (qx * qx + qy * qy);
#pragma unroll
  for (int k = 0; k < 3; k++) s.s[k] = __expf(plane(th, n, 3 + k, i));
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) s.M[3 * r + c] = s.R[3 * r + c] * s.s[c];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      s.Sig[3 * r + c] = s.M[3 * r] * s.M[3 * c] + s.M[3 * r + 1] * s.M[3 * c + 1] + s.M[3 * r + 2] * s.M[3 * c + 2];
  const float x = (float)s.t[0], y = (float)s.t[1], z = (float)s.t[2];
  const float iz = 1.0f / z;
  s.J[0] = u.fx * iz;
  s.J[1] = 0.0f;
  s.J[2] = -u.fx * x * iz * iz;
  s.J[3] = 0.0f;
  s.J[4] = u.fy * iz;
  s.J[5] = -u.fy * y * iz * iz;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      s.T[3 * r + c] = s.J[3 * r] * (float)u.Vl[c] + s.J[3 * r + 1] * (float)u.Vl[3 + c] + s.J[3 * r + 2] * (float)u.Vl[6 + c];
  float TS[6];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) TS[3 * r + c] = s.T[3 * r] * s.Sig[c] + s.T[3 * r + 1] * s.Sig[3 + c] + s.T[3 * r + 2] * s.Sig[6 + c];
  s.a = TS[0] * s.T[0] + TS[1] * s.T[1] + TS[2] * s.T[2];
  s.b = TS[0] * s.T[3] + TS[1] * s.T[4] + TS[2] * s.T[5];
  s.d = TS[3] * s.T[3] + TS[4] * s.T[4] + TS[5] * s.T[5];
  s.ap = s.a + u.blur;
  s.dp = s.d + u.blur;
  s.det_o = s.a * s.d - s.b * s.b;
  s.det_p = s.ap * s.dp - s.b * s.b;
  s.ratio = s.det_o > 0.0f ? s.det_o / s.det_p : 0.0f;
  s.op = 1.0f / (1.0f + __expf(-plane(th, n, 10, i)));
  s.opc = s.op * sqrtf(s.ratio);
  const float idet = 1.0f / s.det_p;
  s.Q0 = s.dp * idet;
  s.Q1 = -s.b * idet;
  s.Q2 = s.ap * idet;
  s.cut = fminf(u.max_std_dev, sqrtf(2.0f * __logf(fmaxf(s.opc, u.min_alpha) / u.min_alpha)));
  // colour: base + SH along the direction from the camera, clamped at 0
  const float dv[3] = {(float)(mu[0] - u.cam[0]), (float)(mu[1] - u.cam[1]), (float)(mu[2] - u.cam[2])};
  s.dlen = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
#pragma unroll
  for (int k = 0; k < 3; k++) s.dir[k] = s.dlen > 0.0f ? dv[k] / s.dlen : 0.0f;
  sh_basis15(s.dir[0], s.dir[1], s.dir[2], s.Y);
  const int nk = u.sh_degree >= 3 ? 15 : (u.sh_degree == 2 ? 8 : (u.sh_degree == 1 ? 3 : 0));
#pragma unroll
  for (int c = 0; c < 3; c++) {
    float v = plane(th, n, 11 + c, i);
#pragma unroll
    for (int k = 0; k < 15; k++)  // constant indices: Y stays in registers
      if (k < nk) v = fmaf(plane(th, n, 14 + 3 * k + c, i), s.Y[k], v);
    s.col_raw[c] = v;
    s.col[c] = fmaxf(v, 0.0f);
  }
  s.ok = !cull_depth && !cull_xy && s.op >= u.min_alpha && s.opc >= u.min_alpha && s.det_p > 0.0f;
}

// ---- per splat: what the pixels need, the key and the tiles -------------------

// pre[3 i .. 3 i + 2] = (mu2x, mu2y, opc, cut^2), (Q0, 2 Q1, Q2, ex), (r, g, b, ey): (ex, ey)
// the half extents of the cutoff ellipse's bounding box, widened by 2 % and
// half a pixel (the renderer's waves skip entries whose widened box misses
// their pixels; z2's float error is far below that margin unless the
// ellipse is extremely elongated, which gets an unbounded box).
static __global__ __launch_bounds__(256) void k_diff_pre(const float* __restrict__ th, uint32_t n, DiffU u,
                                                         float4* __restrict__ pre, uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ touch, uint4* __restrict__ rect) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DiffSplat s;
  diff_splat_forward(th, n, i, u, s);
  float ex = s.cut * sqrtf(s.ap) * 1.02f + 0.5f, ey = s.cut * sqrtf(s.dp) * 1.02f + 0.5f;
  // condition proxy (trace^2 / det of the conic's inverse): beyond 1e6 the box is not trusted
  if (!((s.ap + s.dp) * (s.ap + s.dp) <= 1e6f * s.det_p)) ex = ey = INFINITY;
  pre[3 * (size_t)i] = make_float4(s.mu2x, s.mu2y, s.opc, s.cut * s.cut);
  pre[3 * (size_t)i + 1] = make_float4(s.Q0, 2.0f * s.Q1, s.Q2, ex);
  pre[3 * (size_t)i + 2] = make_float4(s.col[0], s.col[1], s.col[2], ey);
  uint32_t cnt = 0;
  uint4 rc = make_uint4(0u, 0u, 0u, 0u);
  if (s.ok) {
    // the cutoff ellipse's bounding box: cut * sqrt of the covariance diagonal
    const float rx = s.cut * sqrtf(s.ap), ry = s.cut * sqrtf(s.dp);
    const float x0 = s.mu2x - rx, x1 = s.mu2x + rx, y0 = s.mu2y - ry, y1 = s.mu2y + ry;
    if (x1 >= 0.0f && y1 >= 0.0f && x0 < (float)u.width && y0 < (float)u.height) {
      const int tx0 = max(0, (int)floorf(x0 / DIFF_TILE)), tx1 = min(u.tiles_x - 1, (int)floorf(x1 / DIFF_TILE));
      const int ty0 = max(0, (int)floorf(y0 / DIFF_TILE)), ty1 = min(u.tiles_y - 1, (int)floorf(y1 / DIFF_TILE));
      rc = make_uint4((uint32_t)tx0, (uint32_t)ty0, (uint32_t)tx1, (uint32_t)ty1);
      cnt = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
    }
  }
  keys[i] = s.ok ? sort_key(fbits((float)s.dist), 32) : 0xffffffffu;
  touch[i] = cnt;
  rect[i] = rc;
}

// ---- exclusive scan of the tile counts in front-to-back order -------------------
// Position j (front to back) is splat order[m - 1 - j] of the back-to-front
// order. Blocks of SCAN_TILE positions: partial sums, then the block sums in
// one block, then the offsets.
constexpr int DSCAN_BLOCK = 256;
constexpr int DSCAN_ITEMS = 4;
constexpr int DSCAN_TILE = DSCAN_BLOCK * DSCAN_ITEMS;

// Front-to-back position j of m is back-to-front position m - 1 - j: the
// splat order[m - 1 - j], or that position itself when order is NULL (the
// draw, whose per-entry records are stored by position).
__device__ __forceinline__ uint32_t f2b_item(const uint32_t* order, uint32_t m, uint32_t j) {
  return order ? order[m - 1 - j] : m - 1 - j;
}
__device__ __forceinline__ uint32_t f2b_touch(const uint32_t* order, const uint32_t* touch, uint32_t m, uint32_t j) {
  return j < m ? touch[f2b_item(order, m, j)] : 0u;
}

static __global__ __launch_bounds__(DSCAN_BLOCK) void k_diff_scan_partial(const uint32_t* order, const uint32_t* touch,
                                                                          const uint32_t* m_dev, uint32_t* bsum) {
  __shared__ uint32_t s_scan[DSCAN_BLOCK / RDX_WAVE];
  const uint32_t m = *m_dev;
  const uint32_t j0 = blockIdx.x * DSCAN_TILE + threadIdx.x * DSCAN_ITEMS;
  uint32_t v = 0;
#pragma unroll
  for (int k = 0; k < DSCAN_ITEMS; k++) v += f2b_touch(order, touch, m, j0 + k);
  uint32_t total;
  (void)block_exclusive_sum<DSCAN_BLOCK>(v, s_scan, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// One block: exclusive scan of the block sums in place; the total to *total,
// or 0 with *overflow set when the entries would not fit in `cap` (nothing
// is then drawn: no kernel reads an entry that was not written).
static __global__ __launch_bounds__(1024) void k_diff_scan_blocks(uint32_t* bsum, uint32_t blocks, uint32_t* total,
                                                                   uint32_t cap, uint32_t* overflow) {
  __shared__ uint32_t s_scan[1024 / RDX_WAVE];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < blocks; b0 += 1024) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t v = b < blocks ? bsum[b] : 0u;
    uint32_t sum;
    const uint32_t ex = block_exclusive_sum<1024>(v, s_scan, &sum);
    if (b < blocks) bsum[b] = carry + ex;
    carry += sum;
  }
  if (threadIdx.x == 0) {
    *total = carry <= cap ? carry : 0u;
    if (carry > cap) *overflow = 1u;
  }
}

static __global__ __launch_bounds__(DSCAN_BLOCK) void k_diff_scan_final(const uint32_t* order, const uint32_t* touch,
                                                                        const uint32_t* m_dev, const uint32_t* bsum,
                                                                        uint32_t* offsets) {
  __shared__ uint32_t s_scan[DSCAN_BLOCK / RDX_WAVE];
  const uint32_t m = *m_dev;
  const uint32_t j0 = blockIdx.x * DSCAN_TILE + threadIdx.x * DSCAN_ITEMS;
  uint32_t v[DSCAN_ITEMS], sum = 0;
#pragma unroll
  for (int k = 0; k < DSCAN_ITEMS; k++) {
    v[k] = f2b_touch(order, touch, m, j0 + k);
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = bsum[blockIdx.x] + block_exclusive_sum<DSCAN_BLOCK>(sum, s_scan, &total);
#pragma unroll
  for (int k = 0; k < DSCAN_ITEMS; k++) {
    if (j0 + k < m) offsets[j0 + k] = run;
    run += v[k];
  }
}

// ---- entries: one per (splat, tile), front to back -------------------------------

static __global__ __launch_bounds__(256) void k_diff_emit(const uint32_t* order, const uint32_t* m_dev, const uint32_t* touch,
                                                          const uint4* rect, const uint32_t* offsets, uint32_t tiles_x,
                                                          uint32_t cap, uint32_t* tile_key, uint32_t* emit_idx,
                                                          uint32_t* emit_splat, uint32_t* first_entry, uint32_t* overflow) {
  const uint32_t m = *m_dev;
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t sp = f2b_item(order, m, j);
  const uint32_t cnt = touch[sp];
  if (cnt == 0u) return;
  const uint4 r = rect[sp];
  uint32_t q = offsets[j];
  if (q + cnt > cap) {
    store_agent(overflow, 1u);
    return;
  }
  if (first_entry) first_entry[sp] = q;  // the adjoint sums the splat's entries [q, q + cnt)
  for (uint32_t ty = r.y; ty <= r.w; ty++)
    for (uint32_t tx = r.x; tx <= r.z; tx++, q++) {
      tile_key[q] = ty * tiles_x + tx;
      emit_idx[q] = q;
      emit_splat[q] = sp;
    }
}

// Each tile's [start, end) in the tile-sorted entries (tiles with none keep 0, 0).
static __global__ void k_diff_ranges(const uint32_t* tile_key, const uint32_t* e_dev, uint32_t cap, uint32_t tiles,
                                     uint2* range) {
  const uint32_t e = min(*e_dev, cap);
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x) {
    const uint32_t t = tile_key[k];
    if (t >= tiles) continue;  // never: every entry below e was written with its tile
    if (k == 0 || tile_key[k - 1] != t) range[t].x = k;
    if (k == e - 1 || tile_key[k + 1] != t) range[t].y = k + 1;
  }
}

// counts: visible splats, entries, overflow, a zero word
static __global__ void k_fill_counts(uint32_t* counts) {
  if (threadIdx.x < 4) counts[threadIdx.x] = 0u;
}

static __global__ void k_diff_clear_ranges(uint2* range, uint32_t tiles) {
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < tiles; t += gridDim.x * blockDim.x)
    range[t] = make_uint2(0u, 0u);
}

// ---- the tile renderer -----------------------------------------------------------

// exp(-z2 / 2) as one multiply and v_exp_f32 (2^x): the same value as
// __expf(-0.5f * z2), whose halving is exact.
__device__ __forceinline__ float gauss(float z2) { return __builtin_amdgcn_exp2f(z2 * (-0.5f * 1.44269504088896340736f)); }

// The draw's fragment alpha (fragmentAlpha, projection.tsl.ts:444-455): op
// > 0 is the opacity, mixed toward op * g by the falloff; op <= 0 marks an
// LoD opacity above 1 whose exponent a = exp((alpha^2 - 1) / e) the entry
// carries as -op: 1 - (1 - g)^a, mixed toward 1.
__device__ __forceinline__ float draw_alpha(float op, float g, float falloff) {
  if (op > 0.0f) return op * fmaf(falloff, g, 1.0f - falloff);
  const float lod = 1.0f - __builtin_amdgcn_exp2f(-op * __builtin_amdgcn_logf(1.0f - g));
  return fmaf(falloff, lod, 1.0f - falloff);
}

// The tile's pixels: wave w takes the 8x8 quadrant (w & 1, w >> 1) of the
// 16x16 tile, lane l its pixel (l & 7, l >> 3). Entries are staged in LDS
// 256 at a time (the next batch's indices and pre words are loaded while
// this one is composited); each entry's widened box is tested against the
// four quadrants once, and a wave walks only the entries that can reach its
// pixels, with predicated arithmetic per lane.
struct DiffQuad {
  int w, lane, x, y;
  float px, py;
  float cx0[4], cy0[4];  // pixel-centre ranges of the four quadrants: [c0, c0 + 7]
};

__device__ __forceinline__ DiffQuad diff_quad(uint32_t tile, const DiffU& u) {
  DiffQuad q;
  q.lane = threadIdx.x & (RDX_WAVE - 1);
  q.w = threadIdx.x / RDX_WAVE;
  const int tx = (int)(tile % (uint32_t)u.tiles_x), ty = (int)(tile / (uint32_t)u.tiles_x);
  q.x = tx * DIFF_TILE + (q.w & 1) * 8 + (q.lane & 7);
  q.y = ty * DIFF_TILE + (q.w >> 1) * 8 + (q.lane >> 3);
  q.px = (float)q.x + 0.5f;
  q.py = (float)q.y + 0.5f;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    q.cx0[v] = (float)(tx * DIFF_TILE + (v & 1) * 8) + 0.5f;
    q.cy0[v] = (float)(ty * DIFF_TILE + (v >> 1) * 8) + 0.5f;
  }
  return q;
}

// Bit v: the entry's widened box holds a pixel centre of quadrant v.
__device__ __forceinline__ uint32_t quad_mask(const DiffQuad& q, float4 p0, float4 p1, float4 p2) {
  uint32_t m = 0;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const bool hx = p0.x + p1.w >= q.cx0[v] && p0.x - p1.w <= q.cx0[v] + 7.0f;
    const bool hy = p0.y + p2.w >= q.cy0[v] && p0.y - p2.w <= q.cy0[v] + 7.0f;
    m |= (hx && hy) ? (1u << v) : 0u;
  }
  return m;
}

__device__ __forceinline__ void load_entry(const uint32_t* emit_idx, const uint32_t* emit_splat, const float4* pre,
                                           uint32_t e, float4& p0, float4& p1, float4& p2) {
  const uint32_t sp = emit_splat[emit_idx[e]];
  p0 = pre[3 * (size_t)sp];
  p1 = pre[3 * (size_t)sp + 1];
  p2 = pre[3 * (size_t)sp + 2];
}

// Records of the draw's ray-evaluated methods (draw_ray_kernels.hpp): 7
// float4 per draw position, the first three laid out as the ewa record's
// (centre, opacity, cut^2; ..., ex; colour, ey), the rest the method's.
constexpr int REC_EWA = 0, REC_BOUND = 1, REC_RAYGS = 2;
constexpr int RAY_REC = 7;

__device__ __forceinline__ void load_entry_ray(const uint32_t* emit_idx, const uint32_t* emit_splat, const float4* pre,
                                               uint32_t e, float4* r) {
  const uint32_t sp = emit_splat[emit_idx[e]];
#pragma unroll
  for (int j = 0; j < RAY_REC; j++) r[j] = pre[RAY_REC * (size_t)sp + j];
}

// A ray method's density at the pixel (dx, dy) from the record's centre, or
// +inf where the pixel is outside the method's quad (so it never blends).
//  REC_BOUND (3dgut, htgs): the quad is the screen bound, |u| <= qx and
//  |v| <= qy along its axes; localCross and localDir are affine in the
//  pixel, so their squared lengths are quadratic forms in (dx, dy), held as
//  6 coefficients each: z^2 = |localCross|^2 / |localDir|^2 (rayDensity,
//  projection.tsl.ts:437-442), 1e30 for a null direction. (The record's
//  centre is the projected splat centre, where localCross is about 0, so
//  the expanded squares lose nothing to cancellation.)
//  REC_RAYGS: z = (U, V) / W, affine numerators and denominator of the
//  pixel ray's hit on the facing quad; inside where W > 0 and |z| <= zmax;
//  z^2 = |z|^2 c^2 / (c^2 + |z|^2) = (U^2 + V^2) c^2 / (c^2 W^2 + U^2 + V^2)
//  (projections.tsl.ts:462-466).
template <int REC>
__device__ __forceinline__ float ray_z2(float4 r1, float4 r3, float4 r4, float4 r5, float4 r6, float dx, float dy) {
  if constexpr (REC == REC_BOUND) {
    const float u = dx * r1.x + dy * r1.y, v = dy * r1.x - dx * r1.y;
    const bool in = fabsf(u) <= r1.z && fabsf(v) <= r3.x;
    const float xx = dx * dx, xy = dx * dy, yy = dy * dy;
    const float num = r3.y + dx * r3.z + dy * r3.w + xx * r4.x + xy * r4.y + yy * r4.z;
    const float den = r4.w + dx * r5.x + dy * r5.y + xx * r5.z + xy * r5.w + yy * r6.x;
    const float z2 = den <= 0.0f ? 1e30f : fmaxf(num, 0.0f) * __builtin_amdgcn_rcpf(den);
    return in ? z2 : __builtin_inff();
  } else {
    const float U = r3.x + dx * r3.y + dy * r3.z, V = r3.w + dx * r4.x + dy * r4.y, W = r4.z + dx * r4.w + dy * r5.x;
    const float lim = r1.x * W;
    const bool in = W > 0.0f && fabsf(U) <= lim && fabsf(V) <= lim;
    const float c2 = r1.y, ss = U * U + V * V;
    return in ? ss * c2 * __builtin_amdgcn_rcpf(fmaf(c2 * W, W, ss)) : __builtin_inff();
  }
}

// The draw splits a tile's entries into chunks of DRAW_CHUNK: one block
// composites each (a few tiles of a clustered frame hold tens of thousands
// of entries) into a partial (colour, transmittance), and k_draw_combine
// folds a tile's partials front to back ("over" is associative).
constexpr uint32_t DRAW_CHUNK = 1024;
struct DrawItems {
  const uint32_t* items;       // [1]: the number of work items
  const uint32_t* item_tile;   // per item: its tile
  const uint32_t* item_chunk;  // per item: its chunk of the tile's entries
  float4* partial;             // per item and pixel slot: (C, T) of a tile of several chunks
};

// The compositor's adjoint runs per (tile, chunk of DIFF_CHUNK entries of
// the tile's list), so a tile that holds thousands of entries is spread over
// as many blocks (one block per tile left the few densest tiles running
// three times as long as the average one). The forward, which composites a
// tile's list in order in one block (its early stop is sequential), records
// at the end of each chunk, per pixel, the colour that chunk's contributors
// added and the transmittance after them; a chunk's adjoint starts from
// those: its pixels' transmittance behind the chunk, and the colour of the
// later chunks over it. Chunk c of the tile whose list starts at entry x has
// record row tile + x / DIFF_CHUNK + c (distinct for all (tile, chunk): a
// tile of e entries has at most max(1, ceil(e / DIFF_CHUNK)) chunks), so the
// rows need no scan: at most tiles + entries / DIFF_CHUNK + 1.
constexpr uint32_t DIFF_CHUNK = DIFF_BLOCK;  // the forward's batches of entries
struct DiffChunks {
  float4* rec;           // [row][DIFF_BLOCK]: (colour of the chunk's contributors, transmittance after them)
  uint32_t* tile_reach;  // per tile: its pixels' furthest contributor (the largest n_contrib)
};
__host__ __device__ constexpr uint32_t diff_chunk_rows(uint32_t tiles, uint32_t entries) {
  return tiles + entries / DIFF_CHUNK + 1u;
}

// REC: the record kind (REC_EWA for the compositor and the ewa draw; the
// ray methods' records are drawn only, DRAW true).
template <bool DRAW, int REC = REC_EWA>
__global__ __launch_bounds__(DIFF_BLOCK) void k_diff_render(const uint2* range, const uint32_t* emit_idx,
                                                                   const uint32_t* emit_splat, const float4* pre, DiffU u,
                                                                   float* image, float* alpha_out, float* t_final,
                                                                   uint32_t* n_contrib, DrawItems di, DiffChunks dc) {
  static_assert(DIFF_CHUNK == DIFF_BLOCK, "a chunk is one batch of the forward");
  static_assert(DRAW || REC == REC_EWA, "the ray records are drawn only");
  constexpr int RX = REC == REC_EWA ? 1 : DIFF_BLOCK;  // the ray records' other four float4
  __shared__ float4 s_p0[DIFF_BLOCK], s_p1[DIFF_BLOCK], s_p2[DIFF_BLOCK];
  __shared__ float4 s_p3[RX], s_p4[RX], s_p5[RX], s_p6[RX];
  static_assert(REC == REC_EWA || RAY_REC == 7, "the ray records' float4 staged here");
  __shared__ uint32_t s_ov[DIFF_BLOCK];
  uint32_t tile = blockIdx.x;
  uint2 rg;
  uint32_t nchunks = 1;
  if constexpr (DRAW) {
    if (blockIdx.x >= *di.items) return;  // uniform per block
    tile = di.item_tile[blockIdx.x];
    const uint2 all = range[tile];
    nchunks = max(1u, (all.y - all.x + DRAW_CHUNK - 1) / DRAW_CHUNK);
    const uint32_t b0 = all.x + di.item_chunk[blockIdx.x] * DRAW_CHUNK;
    rg = make_uint2(b0, min(b0 + DRAW_CHUNK, all.y));
  } else {
    rg = range[tile];
  }
  const DiffQuad q = diff_quad(tile, u);
  const bool inside = q.x < u.width && q.y < u.height;
  const float cut_alpha = u.min_alpha;
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
  float Cc0 = 0.0f, Cc1 = 0.0f, Cc2 = 0.0f;  // this batch's contributors
  uint32_t last = 0;
  bool done = !inside;
  float4 nr[REC == REC_EWA ? 3 : RAY_REC];
#pragma unroll
  for (int j = 0; j < (REC == REC_EWA ? 3 : RAY_REC); j++) nr[j] = make_float4(0.f, 0.f, 0.f, 0.f);
  auto load = [&](uint32_t e) {
    if constexpr (REC == REC_EWA)
      load_entry(emit_idx, emit_splat, pre, e, nr[0], nr[1], nr[2]);
    else
      load_entry_ray(emit_idx, emit_splat, pre, e, nr);
  };
  if (rg.x + threadIdx.x < rg.y) load(rg.x + threadIdx.x);
  for (uint32_t b = rg.x; b < rg.y; b += DIFF_BLOCK) {
    if (__syncthreads_count(!done) == 0) break;
    const uint32_t e = b + threadIdx.x;
    s_p0[threadIdx.x] = nr[0];
    s_p1[threadIdx.x] = nr[1];
    s_p2[threadIdx.x] = nr[2];
    if constexpr (REC != REC_EWA) {
      s_p3[threadIdx.x] = nr[3];
      s_p4[threadIdx.x] = nr[4];
      s_p5[threadIdx.x] = nr[5];
      s_p6[threadIdx.x] = nr[6];
    }
    s_ov[threadIdx.x] = e < rg.y ? quad_mask(q, nr[0], nr[1], nr[2]) : 0u;
    if (e + DIFF_BLOCK < rg.y) load(e + DIFF_BLOCK);
    __syncthreads();
    const uint32_t cnt = min((uint32_t)DIFF_BLOCK, rg.y - b);
    for (uint32_t c0 = 0; c0 < cnt; c0 += RDX_WAVE) {
      if (!ballot(!done)) break;
      uint64_t m = ballot(c0 + q.lane < cnt && ((s_ov[c0 + q.lane] >> q.w) & 1u));
      while (m) {
        const uint32_t k = c0 + (uint32_t)__builtin_ctzll(m);
        m &= m - 1;
        const float4 p0 = s_p0[k], p1 = s_p1[k];
        const float dx = q.px - p0.x, dy = q.py - p0.y;
        float z2;
        if constexpr (REC == REC_EWA)
          z2 = p1.x * dx * dx + p1.y * dx * dy + p1.z * dy * dy;
        else
          z2 = ray_z2<REC>(p1, s_p3[k], s_p4[k], s_p5[k], s_p6[k], dx, dy);
        const float a = DRAW ? draw_alpha(p0.z, gauss(z2), u.falloff) : p0.z * gauss(z2);
        const bool hit = !done && z2 <= p0.w && a >= cut_alpha;
        const float Tn = T * (1.0f - a);
        // the compositor stops before the splat that would take the
        // transmittance below t_stop; the draw blends it and stops after
        // (what it leaves out is then below t_stop times a colour)
        const bool stop = hit && Tn < u.t_stop;
        const bool use = DRAW ? hit : hit && !stop;
        done = done || stop;
        if (ballot(use)) {
          const float4 c = s_p2[k];
          const float wgt = use ? a * T : 0.0f;
          Cc0 = fmaf(c.x, wgt, Cc0);
          Cc1 = fmaf(c.y, wgt, Cc1);
          Cc2 = fmaf(c.z, wgt, Cc2);
          T = use ? Tn : T;
          last = use ? b - rg.x + k + 1 : last;
        }
      }
    }
    if constexpr (!DRAW)
      dc.rec[((size_t)tile + rg.x / DIFF_CHUNK + (b - rg.x) / DIFF_CHUNK) * DIFF_BLOCK + threadIdx.x] =
          make_float4(Cc0, Cc1, Cc2, T);
    C0 += Cc0;
    C1 += Cc1;
    C2 += Cc2;
    Cc0 = Cc1 = Cc2 = 0.0f;
    __syncthreads();
  }
  if constexpr (!DRAW) {
    // the tile's reach: how many of its entries the adjoint walks
    uint32_t r = last;
    for (int off = 32; off >= 1; off >>= 1) r = max(r, (uint32_t)__shfl_xor((int)r, off));
    if (q.lane == 0) s_ov[q.w] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t m = 0;
      for (int v = 0; v < DIFF_BLOCK / RDX_WAVE; v++) m = max(m, s_ov[v]);
      dc.tile_reach[tile] = m;
    }
  }
  if constexpr (DRAW) {
    if (nchunks > 1) {
      di.partial[(size_t)blockIdx.x * DIFF_BLOCK + threadIdx.x] = make_float4(C0, C1, C2, T);
      return;
    }
  }
  if (!inside) return;
  const size_t p = (size_t)q.y * u.width + q.x, plane_px = (size_t)u.width * u.height;
  image[p] = C0;
  image[plane_px + p] = C1;
  image[2 * plane_px + p] = C2;
  if (alpha_out) alpha_out[p] = 1.0f - T;
  if constexpr (!DRAW) {
    t_final[p] = T;
    n_contrib[p] = last;
  }
}

// Sums over each row of 16 lanes to the row's last lane (DPP row_shr 1, 2, 4, 8
// with zeros shifted in), for N values at once: the steps of different values
// are independent, so they fill each other's DPP wait states.
template <int N>
__device__ __forceinline__ void row_sums_to_last(float* v) {
#pragma unroll
  for (int j = 0; j < N; j++)
    v[j] += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v[j]), 0x111, 0xf, 0xf, true));
#pragma unroll
  for (int j = 0; j < N; j++)
    v[j] += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v[j]), 0x112, 0xf, 0xf, true));
#pragma unroll
  for (int j = 0; j < N; j++)
    v[j] += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v[j]), 0x114, 0xf, 0xf, true));
#pragma unroll
  for (int j = 0; j < N; j++)
    v[j] += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v[j]), 0x118, 0xf, 0xf, true));
}

// One block: the adjoint's work items, (tile, chunk) for each tile's chunks
// up to its reach (at least one: a tile's last item writes the zero
// partials of the entries no pixel reached); their count to *items. The
// full chunks come first and each tile's partial last chunk after them, so
// the blocks dispatched last are the short ones.
static __global__ __launch_bounds__(1024) void k_diff_bwd_items(const uint32_t* tile_reach, uint32_t tiles,
                                                                uint32_t* item_tile, uint32_t* item_chunk, uint32_t* items) {
  __shared__ uint32_t s_scan[1024 / RDX_WAVE];
  uint32_t full = 0;
  for (int part = 0; part < 2; part++) {
    uint32_t carry = part ? full : 0u;
    for (uint32_t b0 = 0; b0 < tiles; b0 += 1024) {
      const uint32_t t = b0 + threadIdx.x;
      const uint32_t reach = t < tiles ? tile_reach[t] : 0u;
      const uint32_t nfull = reach / DIFF_CHUNK;
      // part 0: the full chunks; part 1: the rest (an empty tile's one item)
      const uint32_t k = t >= tiles ? 0u : part == 0 ? nfull : (reach % DIFF_CHUNK != 0u || reach == 0u ? 1u : 0u);
      uint32_t sum;
      const uint32_t at = carry + block_exclusive_sum<1024>(k, s_scan, &sum);
      for (uint32_t c = 0; c < k; c++) {
        item_tile[at + c] = t;
        item_chunk[at + c] = part == 0 ? c : nfull;
      }
      carry += sum;
    }
    if (part == 0) full = carry;
    else if (threadIdx.x == 0) *items = carry;
  }
}

struct BwdItems {
  const uint32_t* items;       // [1]: the number of work items
  const uint32_t* item_tile;   // per item: its tile
  const uint32_t* item_chunk;  // per item: its chunk of the tile's entries
  const float4* rec;           // the forward's chunk records (DiffChunks::rec)
};

// The adjoint of the renderer: per (tile, chunk) item, back to front over
// each pixel's contributors within the chunk (the pixel layout and the entry
// culling of k_diff_render), starting from the forward's record of the
// chunk: the pixel's transmittance after the chunk's contributors, and B,
// the colour composited behind them (the later chunks' colours over that
// transmittance); every entry's pixel-side partials summed over the tile's
// pixels (per row of 16 lanes with DPP, then the 16 rows of the tile in a
// fixed order: no float atomics, reproducible) into gentry[emission
// index][9]. Each entry is written by the item of its chunk alone.
static __global__ __launch_bounds__(DIFF_BLOCK) void k_diff_render_bwd(const uint2* range, const uint32_t* emit_idx,
                                                                       const uint32_t* emit_splat, const float4* pre, DiffU u,
                                                                       const uint32_t* n_contrib, const float* grad_image,
                                                                       float* gentry, BwdItems bi) {
  constexpr int WAVES = DIFF_BLOCK / RDX_WAVE;
  constexpr int BATCH = DIFF_BWD_BATCH;
  constexpr int SLOTS = WAVES * 4;  // a row of 16 lanes per slot
  // a slot's partials, padded so that the four rows of a wave, which write
  // together, fall in different LDS banks (BATCH * 9 words is a multiple of
  // the bank count)
  constexpr int GSTRIDE = BATCH * DIFF_PIX_GRADS + 4;
  constexpr int GWORDS = SLOTS * GSTRIDE;
  __shared__ float4 s_p0[BATCH], s_p1[BATCH], s_p2[BATCH];
  __shared__ uint32_t s_q[BATCH];
  __shared__ uint32_t s_ov[BATCH];
  __shared__ __align__(16) float s_g[GWORDS];
  __shared__ uint32_t s_max[WAVES];
  static_assert(GWORDS % 4 == 0, "partials cleared as float4");
  static_assert(BATCH % RDX_WAVE == 0 && BATCH <= DIFF_BLOCK, "whole waves of entries");
  if (blockIdx.x >= *bi.items) return;  // uniform per block
  const uint32_t tile = bi.item_tile[blockIdx.x], chunk = bi.item_chunk[blockIdx.x];
  const DiffQuad q = diff_quad(tile, u);
  const bool inside = q.x < u.width && q.y < u.height;
  const uint2 rg = range[tile];
  const uint32_t c0e = chunk * DIFF_CHUNK;  // the chunk's first entry, in the tile's list
  const uint32_t cs = rg.x + c0e, ce = min(cs + DIFF_CHUNK, rg.y);
  const size_t p = (size_t)q.y * u.width + q.x, plane_px = (size_t)u.width * u.height;
  const uint32_t nc = inside ? n_contrib[p] : 0u;
  float dC0 = 0.0f, dC1 = 0.0f, dC2 = 0.0f;
  if (inside) {
    dC0 = grad_image[p];
    dC1 = grad_image[plane_px + p];
    dC2 = grad_image[2 * plane_px + p];
  }
  // The pixel behind this chunk: a pixel whose contributors reach into or
  // past it starts from its transmittance after the chunk (>= t_stop > 0)
  // and B = the later chunks' colours (summed back to front) over it.
  float T = 1.0f, B0 = 0.0f, B1 = 0.0f, B2 = 0.0f;
  if (nc > c0e) {
    const float4* rp = bi.rec + ((size_t)tile + rg.x / DIFF_CHUNK) * DIFF_BLOCK + threadIdx.x;
    T = rp[(size_t)chunk * DIFF_BLOCK].w;
    for (uint32_t m = (nc - 1u) / DIFF_CHUNK; m > chunk; m--) {
      const float4 v = rp[(size_t)m * DIFF_BLOCK];
      B0 += v.x;
      B1 += v.y;
      B2 += v.z;
    }
    const float iT = 1.0f / T;
    B0 *= iT;
    B1 *= iT;
    B2 *= iT;
  }
  // the furthest any pixel of the wave, and of the tile, went
  uint32_t wreach = nc;
  for (int off = 32; off >= 1; off >>= 1) wreach = max(wreach, (uint32_t)__shfl_xor((int)wreach, off));
  if (q.lane == 0) s_max[q.w] = wreach;
  __syncthreads();
  uint32_t reach = 0;
  for (int v = 0; v < WAVES; v++) reach = max(reach, s_max[v]);
  const uint32_t end = rg.x + reach;
  // entries no pixel reached get zero partials: this chunk's, and after the
  // tile's last chunk item everything to the end of its list
  {
    const bool last_item = chunk + 1u == max(1u, (reach + DIFF_CHUNK - 1) / DIFF_CHUNK);
    const uint32_t zhi = last_item ? rg.y : ce;
    for (uint32_t e = max(cs, end) + threadIdx.x; e < zhi; e += DIFF_BLOCK)
      for (int k = 0; k < DIFF_PIX_GRADS; k++) gentry[(size_t)emit_idx[e] * DIFF_PIX_GRADS + k] = 0.0f;
  }
  const uint32_t top = min(ce, max(end, cs));  // this chunk's entries below the reach: [cs, top)
  float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
  uint32_t nq = 0;
  {
    const uint32_t lo0 = top - min((uint32_t)BATCH, top - cs);
    if (threadIdx.x < BATCH && lo0 + threadIdx.x < top) {
      nq = emit_idx[lo0 + threadIdx.x];
      const uint32_t sp = emit_splat[nq];
      n0 = pre[3 * (size_t)sp];
      n1 = pre[3 * (size_t)sp + 1];
      n2 = pre[3 * (size_t)sp + 2];
    }
  }
  uint32_t hi = top;
  while (hi > cs) {
    const uint32_t lo = hi - min((uint32_t)BATCH, hi - cs);
    const uint32_t cnt = hi - lo;
    __syncthreads();  // the previous batch's write-out is done
    if (threadIdx.x < BATCH) {
      s_q[threadIdx.x] = nq;
      s_p0[threadIdx.x] = n0;
      s_p1[threadIdx.x] = n1;
      s_p2[threadIdx.x] = n2;
      s_ov[threadIdx.x] = threadIdx.x < cnt ? quad_mask(q, n0, n1, n2) : 0u;
    }
    {
      float4* g4 = reinterpret_cast<float4*>(s_g);
#pragma unroll
      for (int k = threadIdx.x; k < GWORDS / 4; k += DIFF_BLOCK) g4[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (lo > cs) {
      const uint32_t lo2 = lo - min((uint32_t)BATCH, lo - cs);
      if (threadIdx.x < BATCH && lo2 + threadIdx.x < lo) {
        nq = emit_idx[lo2 + threadIdx.x];
        const uint32_t sp = emit_splat[nq];
        n0 = pre[3 * (size_t)sp];
        n1 = pre[3 * (size_t)sp + 1];
        n2 = pre[3 * (size_t)sp + 2];
      }
    }
    __syncthreads();
    // this wave's entries of the batch below its reach, back to front
    const uint32_t wend = wreach > lo - rg.x ? min(cnt, wreach - (lo - rg.x)) : 0u;
    for (int c0 = (int)((wend + RDX_WAVE - 1) / RDX_WAVE) * RDX_WAVE - RDX_WAVE; c0 >= 0; c0 -= RDX_WAVE) {
      uint64_t m = ballot((uint32_t)c0 + q.lane < wend && ((s_ov[c0 + q.lane] >> q.w) & 1u));
      while (m) {
        const int hb = 63 - __builtin_clzll(m);
        m &= ~(1ull << hb);
        const int k = c0 + hb;
        const uint32_t idx = lo - rg.x + (uint32_t)k;  // position in the tile's list
        const float4 p0 = s_p0[k], p1 = s_p1[k];
        const float dx = q.px - p0.x, dy = q.py - p0.y;
        const float z2 = p1.x * dx * dx + p1.y * dx * dy + p1.z * dy * dy;
        const float g = gauss(z2);
        const float a = p0.z * g;
        const bool contrib = idx < nc && z2 <= p0.w && a >= u.min_alpha;
        if (!ballot(contrib)) continue;
        const float4 c = s_p2[k];
        float gv[DIFF_PIX_GRADS];
        // Branch-free: a lane that does not composite this splat takes alpha
        // 0, which leaves its transmittance (divided by 1) and B exactly as
        // they were and makes every partial below vanish.
        const float ae = contrib ? a : 0.0f;
        const float Tb = T * __builtin_amdgcn_rcpf(1.0f - ae);  // the transmittance in front of this splat
        const float aT = ae * Tb;
        const float galpha = Tb * ((c.x - B0) * dC0 + (c.y - B1) * dC1 + (c.z - B2) * dC2);
        const float gz2 = -0.5f * ae * galpha;
        const float q1 = 0.5f * p1.y;
        gv[0] = gz2 * -2.0f * (p1.x * dx + q1 * dy);
        gv[1] = gz2 * -2.0f * (q1 * dx + p1.z * dy);
        gv[2] = gz2 * dx * dx;
        gv[3] = gz2 * 2.0f * dx * dy;
        gv[4] = gz2 * dy * dy;
        gv[5] = (contrib ? g : 0.0f) * galpha;
        gv[6] = aT * dC0;
        gv[7] = aT * dC1;
        gv[8] = aT * dC2;
        T = Tb;
        B0 = fmaf(c.x, ae, (1.0f - ae) * B0);
        B1 = fmaf(c.y, ae, (1.0f - ae) * B1);
        B2 = fmaf(c.z, ae, (1.0f - ae) * B2);
        row_sums_to_last<DIFF_PIX_GRADS>(gv);
        if ((q.lane & 15) == 15) {
          float* gs = s_g + (q.w * 4 + (q.lane >> 4)) * GSTRIDE + k * DIFF_PIX_GRADS;
#pragma unroll
          for (int j = 0; j < DIFF_PIX_GRADS; j++) gs[j] = gv[j];
        }
      }
    }
    __syncthreads();
    // entry k's value j: the sum of the slots in a fixed order
    for (uint32_t o = threadIdx.x; o < cnt * DIFF_PIX_GRADS; o += DIFF_BLOCK) {
      float v = 0.0f;
#pragma unroll
      for (int sl = 0; sl < SLOTS; sl++) v += s_g[sl * GSTRIDE + o];
      const uint32_t k = o / DIFF_PIX_GRADS, j = o - k * DIFF_PIX_GRADS;
      gentry[(size_t)s_q[k] * DIFF_PIX_GRADS + j] = v;
    }
    hi = lo;
  }
}

// The chain rule from the pixel-side partials to the 59 parameters. A
// splat's pixel-side partials are the sum of its entries' (emitted
// consecutively from first_entry, summed in emission order).
static __global__ __launch_bounds__(256) void k_diff_pre_bwd(const float* __restrict__ th, uint32_t n, DiffU u,
                                                             const uint32_t* __restrict__ touch,
                                                             const uint32_t* __restrict__ first_entry,
                                                             const float* __restrict__ gentry, const uint32_t* e_dev,
                                                             uint32_t cap, float* __restrict__ grad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // each parameter's gradient is stored as soon as it is known, so no
  // array of all 59 stays live (157 -> fewer VGPRs: more waves resident)
  float* gp = grad + i;
  auto put = [&](int k, float v) { gp[(size_t)k * n] = v; };
  // The splat's pixel-side partials first, while few registers are live:
  // two entries' loads in flight at a time, added in entry order.
  float G[DIFF_PIX_GRADS];
#pragma unroll
  for (int k = 0; k < DIFF_PIX_GRADS; k++) G[k] = 0.0f;
  {
    const uint32_t cnt = touch[i];
    const uint32_t q0 = cnt ? first_entry[i] : 0u;
    // always, unless the emission overflowed (an error of the forward): the
    // entries written, within the buffer
    if (cnt && q0 + cnt <= min(*e_dev, cap)) {
      const float* ge = gentry + (size_t)q0 * DIFF_PIX_GRADS;
      uint32_t q = 0;
      for (; q + 2 <= cnt; q += 2) {
        float v[2][DIFF_PIX_GRADS];
#pragma unroll
        for (int e = 0; e < 2; e++)
#pragma unroll
          for (int k = 0; k < DIFF_PIX_GRADS; k++) v[e][k] = ge[(size_t)(q + e) * DIFF_PIX_GRADS + k];
#pragma unroll
        for (int e = 0; e < 2; e++)
#pragma unroll
          for (int k = 0; k < DIFF_PIX_GRADS; k++) G[k] += v[e][k];
      }
      if (q < cnt)
#pragma unroll
        for (int k = 0; k < DIFF_PIX_GRADS; k++) G[k] += ge[(size_t)q * DIFF_PIX_GRADS + k];
    }
  }
  DiffSplat s;
  diff_splat_forward(th, n, i, u, s);
  if (!s.ok) {
#pragma unroll
    for (int k = 0; k < DIFF_PARAMS; k++) put(k, 0.0f);
    return;
  }
  {
    const float gmu2x = G[0], gmu2y = G[1], gQ0 = G[2], gQ1 = G[3], gQ2 = G[4], gopc = G[5];
    // colour
    const int nk = u.sh_degree >= 3 ? 15 : (u.sh_degree == 2 ? 8 : (u.sh_degree == 1 ? 3 : 0));
    float graw[3];
    for (int c = 0; c < 3; c++) graw[c] = s.col_raw[c] > 0.0f ? G[6 + c] : 0.0f;
    for (int c = 0; c < 3; c++) put(11 + c, graw[c]);
    float3 Yg[15];
    sh_basis15_grad(s.dir[0], s.dir[1], s.dir[2], Yg);
    float gdir[3] = {0.0f, 0.0f, 0.0f};
    // constant indices throughout (the degree guards each band): the arrays
    // stay in registers
#pragma unroll
    for (int k = 0; k < 15; k++) {
      if (k < nk) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          put(14 + 3 * k + c, s.Y[k] * graw[c]);
          const float cg = plane(th, n, 14 + 3 * k + c, i) * graw[c];
          gdir[0] = fmaf(cg, Yg[k].x, gdir[0]);
          gdir[1] = fmaf(cg, Yg[k].y, gdir[1]);
          gdir[2] = fmaf(cg, Yg[k].z, gdir[2]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) put(14 + 3 * k + c, 0.0f);
      }
    }
    float gmu[3] = {0.0f, 0.0f, 0.0f};
    if (s.dlen > 0.0f) {
      const float dd = gdir[0] * s.dir[0] + gdir[1] * s.dir[1] + gdir[2] * s.dir[2];
      const float il = __builtin_amdgcn_rcpf(s.dlen);
      for (int e = 0; e < 3; e++) gmu[e] += (gdir[e] - s.dir[e] * dd) * il;
    }
    // conic = inverse(cov2D'): dcov' = -Q G Q
    const float G00 = gQ0, G01 = 0.5f * gQ1, G11 = gQ2;
    const float QG00 = s.Q0 * G00 + s.Q1 * G01, QG01 = s.Q0 * G01 + s.Q1 * G11;
    const float QG10 = s.Q1 * G00 + s.Q2 * G01, QG11 = s.Q1 * G01 + s.Q2 * G11;
    float ga = -(QG00 * s.Q0 + QG01 * s.Q1);
    float gd = -(QG10 * s.Q1 + QG11 * s.Q2);
    float gb = -2.0f * (QG00 * s.Q1 + QG01 * s.Q2);
    // opacity compensation
    float gop = 0.0f;
    if (s.ratio > 0.0f) {
      const float h = 0.5f * s.opc;
      const float io = __builtin_amdgcn_rcpf(s.det_o), ip = __builtin_amdgcn_rcpf(s.det_p);
      ga += gopc * h * (s.d * io - s.dp * ip);
      gd += gopc * h * (s.a * io - s.ap * ip);
      gb += gopc * h * (-2.0f * s.b * io + 2.0f * s.b * ip);
      gop = gopc * sqrtf(s.ratio);
    }
    put(10, gop * s.op * (1.0f - s.op));
    // cov2D = T Sigma T^T
    const float G2[4] = {ga, 0.5f * gb, 0.5f * gb, gd};
    float gSig[9], gT[6];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        float v = 0.0f;
        for (int k = 0; k < 2; k++)
          for (int l = 0; l < 2; l++) v += s.T[3 * k + r] * G2[2 * k + l] * s.T[3 * l + c];
        gSig[3 * r + c] = v;
      }
    for (int r = 0; r < 2; r++)
      for (int c = 0; c < 3; c++) {
        float v = 0.0f;
        for (int l = 0; l < 2; l++)
          for (int k = 0; k < 3; k++) v += G2[2 * r + l] * s.T[3 * l + k] * s.Sig[3 * k + c];
        gT[3 * r + c] = 2.0f * v;
      }
    float gJ[6];
    for (int r = 0; r < 2; r++)
      for (int c = 0; c < 3; c++)
        gJ[3 * r + c] = gT[3 * r] * (float)u.Vl[3 * c] + gT[3 * r + 1] * (float)u.Vl[3 * c + 1] + gT[3 * r + 2] * (float)u.Vl[3 * c + 2];
    const float x = (float)s.t[0], y = (float)s.t[1], z = (float)s.t[2];
    const float iz = __builtin_amdgcn_rcpf(z), iz2 = iz * iz, iz3 = iz2 * iz;
    float gt[3];
    gt[0] = gJ[2] * (-u.fx * iz2);
    gt[1] = gJ[5] * (-u.fy * iz2);
    gt[2] = gJ[0] * (-u.fx * iz2) + gJ[2] * (2.0f * u.fx * x * iz3) + gJ[4] * (-u.fy * iz2) + gJ[5] * (2.0f * u.fy * y * iz3);
    // the centre in pixels from clip
    const float iw = __builtin_amdgcn_rcpf((float)s.clip[3]);
    const float gcx = gmu2x * 0.5f * u.width * iw, gcy = gmu2y * 0.5f * u.height * iw;
    const float gcw = -(gmu2x * 0.5f * u.width * (float)s.clip[0] + gmu2y * 0.5f * u.height * (float)s.clip[1]) * (iw * iw);
    for (int k = 0; k < 3; k++) gt[k] += gcx * (float)u.P[k] + gcy * (float)u.P[4 + k] + gcw * (float)u.P[12 + k];
    for (int k = 0; k < 3; k++) gmu[k] += (float)u.Vl[k] * gt[0] + (float)u.Vl[3 + k] * gt[1] + (float)u.Vl[6 + k] * gt[2];
    for (int k = 0; k < 3; k++) put(k, gmu[k]);
    // Sigma = M M^T, M = R diag(s)
    float gM[9];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        float v = 0.0f;
        for (int k = 0; k < 3; k++) v += (gSig[3 * r + k] + gSig[3 * k + r]) * s.M[3 * k + c];
        gM[3 * r + c] = v;
      }
    float gR[9];
    for (int c = 0; c < 3; c++) {
      float gs = 0.0f;
      for (int r = 0; r < 3; r++) {
        gs += gM[3 * r + c] * s.R[3 * r + c];
        gR[3 * r + c] = gM[3 * r + c] * s.s[c];
      }
      put(3 + c, gs * s.s[c]);
    }
    const float qx = s.qv[0], qy = s.qv[1], qz = s.qv[2], qw = s.qv[3];
    float gq[4];
    gq[0] = 2.0f * (gR[1] * qy + gR[2] * qz + gR[3] * qy - gR[5] * qw + gR[6] * qz + gR[7] * qw) - 4.0f * qx * (gR[4] + gR[8]);
    gq[1] = 2.0f * (gR[1] * qx + gR[2] * qw + gR[3] * qx + gR[5] * qz - gR[6] * qw + gR[7] * qz) - 4.0f * qy * (gR[0] + gR[8]);
    gq[2] = 2.0f * (-gR[1] * qw + gR[2] * qx + gR[3] * qw + gR[5] * qy + gR[6] * qx + gR[7] * qy) - 4.0f * qz * (gR[0] + gR[4]);
    gq[3] = 2.0f * (-gR[1] * qz + gR[2] * qy + gR[3] * qz - gR[5] * qx - gR[6] * qy + gR[7] * qx);
    if (s.qlen > 0.0f) {
      const float dq = gq[0] * qx + gq[1] * qy + gq[2] * qz + gq[3] * qw;
      const float iq = __builtin_amdgcn_rcpf(s.qlen);
      for (int k = 0; k < 4; k++) put(6 + k, (gq[k] - s.qv[k] * dq) * iq);
    } else {
      for (int k = 0; k < 4; k++) put(6 + k, 0.0f);
    }
  }
}

}  // namespace rdx

// Host side of the production draw (draw_kernels.hpp): the sorted splats'
// ewa words composited per 16x16 tile into a premultiplied RGBA image, on
// one stream, with buffers sized on demand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "draw_kernels.hpp"
#include "draw_ray_kernels.hpp"
#include "project_host.hpp"
#include "sort_kernels.hpp"

namespace rdx {

struct Draw {
  Context* ctx = nullptr;
  rdx_draw_settings set;
  SortScratch sort_entries;
  Buffer pre, touch, rect, offsets, bsum, tile_key, emit_idx, emit_splat, range, counts;
  Buffer item_start, item_tile, item_chunk, partial;  // the render's work items (tile chunks)
  uint32_t capacity = 0, cap = 0, max_items = 0;     // draw positions, entries, work items
  void release() {
    sort_entries.release();
    Buffer* b[] = {&pre,     &touch,      &rect,       &offsets,    &bsum,      &tile_key, &emit_idx,
                   &emit_splat, &range, &counts, &item_start, &item_tile, &item_chunk, &partial};
    for (Buffer* x : b) x->release();
  }
};

// counts words: [0] drawn positions (the clamped count), [1] entries, [2] overflow, [3] work items
enum { D_DRAWN = 0, D_ENTRIES = 1, D_OVERFLOW = 2, D_ITEMS = 3, D_WORDS = 4 };

// Entries of 16-bit tile keys: a frame's millions fill the machine with
// 8,192-key tiles.
static SortPlan draw_entry_plan() { return make_plan(rdx_sort_tuning{8, 16, 512}, 16); }

// rec_words: float4 per draw position (3 for the ewa records, RAY_REC for the ray methods')
static int reserve(Draw* d, uint32_t capacity, uint32_t tiles, int rec_words) {
  const size_t nn = std::max<uint32_t>(capacity, 1);
  int rc;
  if ((rc = d->pre.ensure(nn * rec_words * sizeof(float4))) || (rc = d->touch.ensure(nn * 4)) ||
      (rc = d->rect.ensure(nn * sizeof(uint4))) || (rc = d->offsets.ensure(nn * 4)) ||
      (rc = d->bsum.ensure((nn / DSCAN_TILE + 2) * 4)) || (rc = d->counts.ensure(D_WORDS * 4)) ||
      (rc = d->range.ensure((size_t)tiles * sizeof(uint2))) || (rc = d->item_start.ensure((size_t)tiles * 4)))
    return rc;
  d->capacity = (uint32_t)nn;
  return RDX_OK;
}

// The entry buffers and the render's work items for `need` entries over
// `tiles` tiles: grown (an eighth to spare), never shrunk.
static int reserve_entries(Draw* d, uint32_t need, uint32_t tiles) {
  // a tile of e entries is ceil(e / DRAW_CHUNK) >= 1 items: at most tiles + cap / DRAW_CHUNK
  if (d->cap && need <= d->cap && tiles + (d->cap + DRAW_CHUNK - 1) / DRAW_CHUNK <= d->max_items) return RDX_OK;
  const uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)std::max(need, d->cap) + need / 8 + 1024, 0x7fffffffull);
  int rc;
  if ((rc = d->tile_key.ensure((size_t)cap * 4)) || (rc = d->emit_idx.ensure((size_t)cap * 4)) ||
      (rc = d->emit_splat.ensure((size_t)cap * 4)))
    return rc;
  const uint32_t max_items = tiles + (cap + DRAW_CHUNK - 1) / DRAW_CHUNK;
  if ((rc = d->item_tile.ensure((size_t)max_items * 4)) || (rc = d->item_chunk.ensure((size_t)max_items * 4)) ||
      (rc = d->partial.ensure((size_t)max_items * DIFF_BLOCK * sizeof(float4))))
    return rc;
  const SortPlan pe = draw_entry_plan();
  if ((rc = d->sort_entries.reserve(cap, 1, tiles_of(pe, cap)))) return rc;
  d->max_items = max_items;
  d->cap = cap;
  return RDX_OK;
}

static uint32_t blocks_for(size_t items, uint32_t block = 256) {
  return (uint32_t)std::max<size_t>(1, (items + block - 1) / block);
}

static int draw_uniforms(const char* who, Draw* d, uint32_t width, uint32_t height, const rdx_params* params,
                         DiffU* u) {
  if (width < 1 || height < 1 || width > 16384 || height > 16384)
    return fail(RDX_E_INVALID, "%s: render size %ux%u (1..16384)", who, width, height);
  memset(u, 0, sizeof *u);
  u->width = (int)width;
  u->height = (int)height;
  u->tiles_x = (int)((width + DIFF_TILE - 1) / DIFF_TILE);
  u->tiles_y = (int)((height + DIFF_TILE - 1) / DIFF_TILE);
  u->max_std_dev = (float)params->max_std_dev;
  u->min_alpha = (float)params->min_alpha;
  u->falloff = (float)params->falloff;
  u->t_stop = d->set.t_stop;
  if (!(u->min_alpha > 0.0f) || !(u->max_std_dev > 0.0f))
    return fail(RDX_E_INVALID, "%s: min_alpha and max_std_dev must be positive", who);
  if ((uint32_t)(u->tiles_x * u->tiles_y) > 65536) return fail(RDX_E_INVALID, "%s: more than 65536 tiles", who);
  return RDX_OK;
}

// The draw after its records: binning, the entries' stable sort by tile, the
// tile ranges and the renderer of the record kind REC (rec_words float4 a
// record); pre_launch(counts) writes the records, touches and rectangles.
template <int REC, typename PreLaunch>
static int draw_run(Draw* d, hipStream_t s, const DiffU& u, const uint32_t* count_dev, uint32_t capacity,
                    float* rgba_dev, PreLaunch pre_launch) {
  const uint32_t width = (uint32_t)u.width, height = (uint32_t)u.height;
  const uint32_t tiles = (uint32_t)(u.tiles_x * u.tiles_y);
  if (int rc = reserve(d, capacity, tiles, REC == REC_EWA ? 3 : RAY_REC)) return rc;
  Context* c = d->ctx;
  uint32_t* counts = (uint32_t*)d->counts.ptr;
  const uint32_t cap_pos = capacity;  // positions the order holds
  hipLaunchKernelGGL(k_draw_counts, dim3(1), dim3(64), 0, s, count_dev, cap_pos, counts);
  {
    Stage st(c, s, "draw_pre");
    if (cap_pos) pre_launch(counts);
  }
  const uint32_t blocks = (uint32_t)((d->capacity + DSCAN_TILE - 1) / DSCAN_TILE);
  {
    Stage st(c, s, "draw_bin");
    hipLaunchKernelGGL(k_diff_scan_partial, dim3(blocks), dim3(DSCAN_BLOCK), 0, s, nullptr,
                       (const uint32_t*)d->touch.ptr, counts + D_DRAWN, (uint32_t*)d->bsum.ptr);
    hipLaunchKernelGGL(k_diff_scan_blocks, dim3(1), dim3(1024), 0, s, (uint32_t*)d->bsum.ptr, blocks, counts + D_ENTRIES,
                       0xffffffffu, counts + D_OVERFLOW);
  }
  {
    // The frame's entry count, read back once (the stream waits for the
    // scan): the entry buffers are sized to it, so no frame is cut short; a
    // per-position limit in the settings turns a larger count into an error.
    uint32_t need = 0;
    hipError_t e = hipMemcpyAsync(&need, counts + D_ENTRIES, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "rdx_draw_splats: entry count");
    const uint64_t limit = std::min<uint64_t>(
        d->set.max_entries_per_splat ? (uint64_t)d->set.max_entries_per_splat * capacity : 0x7fffffffull, 0x7fffffffull);
    if (need > limit)
      return fail(RDX_E_CAPACITY, "rdx_draw_splats: %u (splat, tile) entries exceed the limit of %llu", need,
                  (unsigned long long)limit);
    if (int rc = reserve_entries(d, need, tiles)) return rc;
  }
  {
    Stage st(c, s, "draw_bin");
    hipLaunchKernelGGL(k_diff_scan_final, dim3(blocks), dim3(DSCAN_BLOCK), 0, s, nullptr, (const uint32_t*)d->touch.ptr,
                       counts + D_DRAWN, (const uint32_t*)d->bsum.ptr, (uint32_t*)d->offsets.ptr);
    hipLaunchKernelGGL(k_diff_emit, dim3(blocks_for(d->capacity)), dim3(256), 0, s, nullptr, counts + D_DRAWN,
                       (const uint32_t*)d->touch.ptr, (const uint4*)d->rect.ptr, (const uint32_t*)d->offsets.ptr,
                       (uint32_t)u.tiles_x, d->cap, (uint32_t*)d->tile_key.ptr, (uint32_t*)d->emit_idx.ptr,
                       (uint32_t*)d->emit_splat.ptr, nullptr, counts + D_OVERFLOW);
  }
  // stable sort of the entries by tile: front to back within each tile
  const SortPlan pe = draw_entry_plan();
  SegLayout Le;
  Le.tiles = tiles_of(pe, d->cap);
  if (int rc = run_sort(c, d->sort_entries, s, pe, Le, (const uint32_t*)d->tile_key.ptr, (const uint32_t*)d->emit_idx.ptr,
                        (uint32_t*)d->tile_key.ptr, (uint32_t*)d->emit_idx.ptr, d->cap, counts + D_ENTRIES, false, 0u,
                        false, nullptr, nullptr))
    return rc;
  {
    Stage st(c, s, "draw_ranges");
    hipLaunchKernelGGL(k_diff_clear_ranges, dim3(blocks_for(tiles)), dim3(256), 0, s, (uint2*)d->range.ptr, tiles);
    hipLaunchKernelGGL(k_diff_ranges, dim3(std::min<uint32_t>(blocks_for(d->cap), 4096u)), dim3(256), 0, s,
                       (const uint32_t*)d->tile_key.ptr, counts + D_ENTRIES, d->cap, tiles, (uint2*)d->range.ptr);
  }
  {
    Stage st(c, s, "draw_render");
    const size_t plane = (size_t)width * height;
    hipLaunchKernelGGL(k_draw_items, dim3(1), dim3(1024), 0, s, (const uint2*)d->range.ptr, tiles,
                       (uint32_t*)d->item_start.ptr, (uint32_t*)d->item_tile.ptr, (uint32_t*)d->item_chunk.ptr,
                       counts + D_ITEMS);
    const DrawItems di{counts + D_ITEMS, (const uint32_t*)d->item_tile.ptr, (const uint32_t*)d->item_chunk.ptr,
                       (float4*)d->partial.ptr};
    hipLaunchKernelGGL((k_diff_render<true, REC>), dim3(d->max_items), dim3(DIFF_BLOCK), 0, s, (const uint2*)d->range.ptr,
                       (const uint32_t*)d->emit_idx.ptr, (const uint32_t*)d->emit_splat.ptr, (const float4*)d->pre.ptr, u,
                       rgba_dev, rgba_dev + 3 * plane, nullptr, nullptr, di, DiffChunks{});
    hipLaunchKernelGGL(k_draw_combine, dim3(tiles), dim3(DIFF_BLOCK), 0, s, (const uint2*)d->range.ptr,
                       (const uint32_t*)d->item_start.ptr, (const float4*)d->partial.ptr, u, rgba_dev,
                       rgba_dev + 3 * plane);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_draw");
}

}  // namespace rdx

using namespace rdx;

extern "C" {

int rdx_draw_create(rdx_context* ctx, const rdx_draw_settings* s, rdx_draw** out) {
  if (!out) return fail(RDX_E_INVALID, "rdx_draw_create: out is NULL");
  *out = nullptr;
  if (!ctx) return fail(RDX_E_INVALID, "rdx_draw_create: no context");
  if (s && !(s->t_stop >= 0.0f && s->t_stop < 1.0f))
    return fail(RDX_E_INVALID, "rdx_draw_create: t_stop must be in [0, 1) (got %g)", (double)s->t_stop);
  Draw* d = new Draw();
  d->ctx = reinterpret_cast<Context*>(ctx);
  if (s) {
    d->set = *s;
  } else {
    memset(&d->set, 0, sizeof d->set);
  }
  *out = reinterpret_cast<rdx_draw*>(d);
  return RDX_OK;
}

int rdx_draw_destroy(rdx_draw* draw) {
  if (!draw) return RDX_OK;
  Draw* d = reinterpret_cast<Draw*>(draw);
  d->release();
  delete d;
  return RDX_OK;
}

int rdx_draw_splats(rdx_draw* draw, void* stream, const uint32_t* words_dev, uint32_t n, const uint32_t* order_dev,
                    const uint32_t* count_dev, uint32_t capacity, uint32_t width, uint32_t height,
                    const rdx_params* params, float* rgba_dev) {
  if (!draw || !count_dev || !params || !rgba_dev || (n && !words_dev) || (capacity && !order_dev))
    return fail(RDX_E_INVALID, "rdx_draw_splats: NULL argument");
  Draw* d = reinterpret_cast<Draw*>(draw);
  hipStream_t s = (hipStream_t)stream;
  DiffU u;
  if (int rc = draw_uniforms("rdx_draw_splats", d, width, height, params, &u)) return rc;
  return draw_run<REC_EWA>(d, s, u, count_dev, capacity, rgba_dev, [&](uint32_t* counts) {
    hipLaunchKernelGGL(k_draw_pre, dim3(blocks_for(capacity)), dim3(256), 0, s, (const uint4*)words_dev, n, order_dev,
                       counts + D_DRAWN, u, d->set.linear_output, (float4*)d->pre.ptr, (uint32_t*)d->touch.ptr,
                       (uint4*)d->rect.ptr);
  });
}

int rdx_draw_ray(rdx_draw* draw, void* stream, int method, uint32_t* const* channels_dev, uint32_t n,
                 const uint32_t* order_dev, const uint32_t* count_dev, uint32_t capacity, const rdx_camera* camera,
                 int eye, const rdx_params* params, float* rgba_dev) {
  if (!draw || !count_dev || !params || !rgba_dev || !camera || !channels_dev || (capacity && !order_dev))
    return fail(RDX_E_INVALID, "rdx_draw_ray: NULL argument");
  if (method != RDX_METHOD_3DGUT && method != RDX_METHOD_HTGS && method != RDX_METHOD_RAYGS)
    return fail(RDX_E_INVALID, "rdx_draw_ray: method %d is not 3dgut (1), htgs (2) or raygs (3) (ewa: rdx_draw_splats)",
                method);
  if (camera->eyes < 1 || camera->eyes > 2 || eye < 0 || eye >= camera->eyes)
    return fail(RDX_E_INVALID, "rdx_draw_ray: eye %d of a camera of %d eyes", eye, camera->eyes);
  if (n && (!channels_dev[0] || !channels_dev[1] || (camera->eyes == 2 && !channels_dev[2])))
    return fail(RDX_E_INVALID, "rdx_draw_ray: a word channel is NULL");
  Draw* d = reinterpret_cast<Draw*>(draw);
  hipStream_t s = (hipStream_t)stream;
  const double sw = camera->size[0], sh = camera->size[1];
  if (!(sw >= 1.0 && sh >= 1.0 && sw <= 16384.0 && sh <= 16384.0 && sw == std::floor(sw) && sh == std::floor(sh)))
    return fail(RDX_E_INVALID, "rdx_draw_ray: render size %gx%g (whole pixels, 1..16384)", sw, sh);
  DiffU u;
  if (int rc = draw_uniforms("rdx_draw_ray", d, (uint32_t)sw, (uint32_t)sh, params, &u)) return rc;
  // the eye (SplatCamera.setViews, splatCamera.ts:119-138): fromHead = view * head^-1 and its rotation
  RayDrawU e;
  memset(&e, 0, sizeof e);
  double hinv[16], fh[16];
  if (!invert(camera->head, hinv)) return fail(RDX_E_INVALID, "rdx_draw_ray: head matrix is singular");
  mul(camera->eye[eye].view, hinv, fh);
  double R[9];
  for (int r = 0; r < 3; r++) {
    for (int k = 0; k < 3; k++) e.fh[4 * r + k] = (float)cm(fh, r, k);
    e.fh[4 * r + 3] = (float)cm(fh, r, 3);
  }
  // the rotation of the decomposition: columns over their lengths (the first negated when det < 0)
  double len[3];
  for (int k = 0; k < 3; k++) len[k] = std::sqrt(cm(fh, 0, k) * cm(fh, 0, k) + cm(fh, 1, k) * cm(fh, 1, k) + cm(fh, 2, k) * cm(fh, 2, k));
  for (int r = 0; r < 3; r++)
    for (int k = 0; k < 3; k++) R[3 * r + k] = cm(fh, r, k);
  if (det3(R) < 0.0) len[0] = -len[0];
  for (int k = 0; k < 3; k++) {
    if (!(len[k] != 0.0)) return fail(RDX_E_INVALID, "rdx_draw_ray: the eye's fromHead is singular");
    for (int r = 0; r < 3; r++) R[3 * r + k] /= len[k];
  }
  quat_of_rotation(R, e.fq);
  for (int r = 0; r < 4; r++)
    for (int k = 0; k < 4; k++) e.P[4 * r + k] = (float)cm(camera->eye[eye].projection, r, k);
  e.ortho = camera->eye[eye].ortho ? 1 : 0;
  e.eye = eye;
  e.eyes = camera->eyes;
  const uint4* c0 = (const uint4*)channels_dev[0];
  const uint4* c1 = (const uint4*)channels_dev[1];
  const uint2* c2 = camera->eyes == 2 ? (const uint2*)channels_dev[2] : nullptr;
  auto pre = [&](uint32_t* counts) {
    const dim3 g(blocks_for(capacity)), b(256);
    float4* p = (float4*)d->pre.ptr;
    uint32_t* t = (uint32_t*)d->touch.ptr;
    uint4* rc = (uint4*)d->rect.ptr;
    if (method == RDX_METHOD_3DGUT)
      hipLaunchKernelGGL(k_draw_ray_pre<METHOD_3DGUT>, g, b, 0, s, c0, c1, c2, n, order_dev, counts + D_DRAWN, u, e,
                         d->set.linear_output, p, t, rc);
    else if (method == RDX_METHOD_HTGS)
      hipLaunchKernelGGL(k_draw_ray_pre<METHOD_HTGS>, g, b, 0, s, c0, c1, c2, n, order_dev, counts + D_DRAWN, u, e,
                         d->set.linear_output, p, t, rc);
    else
      hipLaunchKernelGGL(k_draw_ray_pre<METHOD_RAYGS>, g, b, 0, s, c0, c1, c2, n, order_dev, counts + D_DRAWN, u, e,
                         d->set.linear_output, p, t, rc);
  };
  if (method == RDX_METHOD_RAYGS) return draw_run<REC_RAYGS>(d, s, u, count_dev, capacity, rgba_dev, pre);
  return draw_run<REC_BOUND>(d, s, u, count_dev, capacity, rgba_dev, pre);
}

int rdx_draw_stats(rdx_draw* draw, void* stream, uint32_t* drawn, uint32_t* entries, uint32_t* capacity,
                   int* overflowed) {
  if (!draw) return fail(RDX_E_INVALID, "rdx_draw_stats: no draw");
  Draw* d = reinterpret_cast<Draw*>(draw);
  uint32_t h[D_WORDS] = {0, 0, 0, 0};
  if (d->counts.ptr) {
    hipError_t e = hipMemcpyAsync(h, d->counts.ptr, sizeof h, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rdx_draw_stats");
  }
  if (drawn) *drawn = h[D_DRAWN];
  if (entries) *entries = h[D_ENTRIES];
  if (capacity) *capacity = d->cap;
  if (overflowed) *overflowed = h[D_OVERFLOW] != 0;
  return RDX_OK;
}

}  // extern "C"

// Host side of the differentiable compositor (raster_kernels.hpp): buffers
// sized on demand, the forward sequence and the backward sequence on one
// stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "raster_kernels.hpp"
#include "sort_kernels.hpp"

namespace rdx {

struct Raster {
  Context* ctx = nullptr;
  rdx_raster_settings set;
  SortScratch sort_splats, sort_entries;
  Buffer pre, keys, touch, rect, order, counts, offsets, bsum, tile_key, emit_idx, emit_splat, range, t_final,
      n_contrib, gentry, first_entry;
  // the forward's chunk records and the adjoint's (tile, chunk) items
  Buffer chunk_rec, tile_reach, item_tile, item_chunk, items;
  uint32_t cap = 0;   // entries
  uint32_t rows = 0;  // chunk record rows (and the adjoint's largest item count) of the last forward
  // the last forward
  DiffU u;
  const float* params = nullptr;
  uint32_t n = 0;
  bool have_forward = false;
  void release() {
    sort_splats.release();
    sort_entries.release();
    Buffer* b[] = {&pre,   &keys,      &touch,     &rect,       &order,     &counts,     &offsets,
                   &bsum,  &tile_key,  &emit_idx,  &emit_splat, &range,     &t_final,    &n_contrib,
                   &gentry, &first_entry, &chunk_rec, &tile_reach, &item_tile, &item_chunk, &items};
    for (Buffer* x : b) x->release();
  }
};

// counts buffer words: [0] visible splats, [1] entries, [2] overflow flag
enum { C_VISIBLE = 0, C_ENTRIES = 1, C_OVERFLOW = 2, C_WORDS = 4 };

static int make_diff_uniforms(const rdx_camera* cam, const rdx_params* prm, int sh_degree, float t_stop, DiffU* u) {
  if (cam->eyes != 1) return fail(RDX_E_INVALID, "rdx_raster: one eye (got %d)", cam->eyes);
  const int W = (int)cam->size[0], H = (int)cam->size[1];
  if (W < 1 || H < 1 || W > 4096 || H > 4096) return fail(RDX_E_INVALID, "rdx_raster: render size %dx%d (1..4096)", W, H);
  memset(u, 0, sizeof *u);
  // world -> view = the eye's view (column-major) times the head
  double V[16];
  const double* A = cam->eye[0].view;
  const double* B = cam->head;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      double v = 0;
      for (int k = 0; k < 4; k++) v += A[4 * k + r] * B[4 * c + k];
      V[4 * c + r] = v;
    }
  for (int r = 0; r < 3; r++) {
    for (int c = 0; c < 3; c++) u->Vl[3 * r + c] = V[4 * c + r];
    u->Vt[r] = V[12 + r];
  }
  // camera position: solve Vl c = -Vt
  const double* L = u->Vl;
  const double det = L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) + L[2] * (L[3] * L[7] - L[4] * L[6]);
  if (!(std::fabs(det) > 0.0)) return fail(RDX_E_INVALID, "rdx_raster: singular view");
  const double inv[9] = {(L[4] * L[8] - L[5] * L[7]) / det, (L[2] * L[7] - L[1] * L[8]) / det, (L[1] * L[5] - L[2] * L[4]) / det,
                         (L[5] * L[6] - L[3] * L[8]) / det, (L[0] * L[8] - L[2] * L[6]) / det, (L[2] * L[3] - L[0] * L[5]) / det,
                         (L[3] * L[7] - L[4] * L[6]) / det, (L[1] * L[6] - L[0] * L[7]) / det, (L[0] * L[4] - L[1] * L[3]) / det};
  for (int r = 0; r < 3; r++) u->cam[r] = -(inv[3 * r] * u->Vt[0] + inv[3 * r + 1] * u->Vt[1] + inv[3 * r + 2] * u->Vt[2]);
  const double* P = cam->eye[0].projection;
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) u->P[4 * r + c] = P[4 * c + r];
  u->fx = (float)(0.5 * W * u->P[0]);
  u->fy = (float)(0.5 * H * u->P[5]);
  u->width = W;
  u->height = H;
  u->tiles_x = (W + DIFF_TILE - 1) / DIFF_TILE;
  u->tiles_y = (H + DIFF_TILE - 1) / DIFF_TILE;
  u->ndc_z_min = (float)cam->ndc_z_min;
  u->max_std_dev = (float)prm->max_std_dev;
  u->min_alpha = (float)prm->min_alpha;
  u->blur = (float)prm->blur_amount;
  u->clip_xy = (float)prm->clip_xy;
  u->t_stop = t_stop;
  u->sh_degree = std::max(0, std::min(3, sh_degree));
  u->falloff = 1.0f;
  return RDX_OK;
}

// The compositor's two sorts have their own tile shapes: a few hundred
// thousand depth keys and about a million entries fill the machine only with
// smaller tiles than the context's (8,192 keys by default). Depth: ranged
// compacting sort, 512 threads x 8 keys; entries (16-bit tile keys, sized by
// the capacity, the blocks past the device count return at once): 256 x 8.
static SortPlan depth_plan() { return compact_plan(rdx_sort_tuning{8, 8, 512}, 32); }
static SortPlan entry_plan() { return make_plan(rdx_sort_tuning{8, 8, 256}, 16); }

static int reserve(Raster* r, uint32_t n, const DiffU& u) {
  const SortPlan p = depth_plan();
  const size_t px = (size_t)u.width * u.height, tiles = (size_t)u.tiles_x * u.tiles_y;
  const size_t nn = std::max<uint32_t>(n, 1);
  int rc;
  if ((rc = r->pre.ensure(nn * 3 * sizeof(float4)))) return rc;
  if ((rc = r->keys.ensure(nn * 4)) || (rc = r->touch.ensure(nn * 4)) || (rc = r->rect.ensure(nn * sizeof(uint4))) ||
      (rc = r->order.ensure(nn * 4)) || (rc = r->offsets.ensure(nn * 4)) || (rc = r->counts.ensure(C_WORDS * 4)) ||
      (rc = r->bsum.ensure((nn / DSCAN_TILE + 2) * 4)))
    return rc;
  if ((rc = r->range.ensure(tiles * sizeof(uint2))) || (rc = r->t_final.ensure(px * 4)) || (rc = r->n_contrib.ensure(px * 4)) ||
      (rc = r->first_entry.ensure(nn * 4)) || (rc = r->tile_reach.ensure(tiles * 4)) ||
      (rc = r->items.ensure(4)))
    return rc;
  if ((rc = r->sort_splats.reserve(n, 1, tiles_of(p, n)))) return rc;
  return RDX_OK;
}

// The (splat, tile) entry buffers for `need` entries: grown (with an eighth
// to spare, so a frame's count wandering up does not reallocate every time),
// never shrunk.
static int reserve_entries(Raster* r, uint32_t need) {
  if (need <= r->cap && r->cap) return RDX_OK;
  const uint32_t cap = (uint32_t)std::min<uint64_t>((uint64_t)need + need / 8 + 1024, 0x7fffffffull);
  int rc;
  if ((rc = r->tile_key.ensure((size_t)cap * 4)) || (rc = r->emit_idx.ensure((size_t)cap * 4)) ||
      (rc = r->emit_splat.ensure((size_t)cap * 4)) || (rc = r->gentry.ensure((size_t)cap * DIFF_PIX_GRADS * 4)))
    return rc;
  const SortPlan pe = entry_plan();
  if ((rc = r->sort_entries.reserve(cap, 1, tiles_of(pe, cap)))) return rc;
  r->cap = cap;
  return RDX_OK;
}

// The frame's entry count, read back once (the stream waits for the scan):
// the entry buffers are sized to it, so no frame is ever cut short. A
// per-splat limit in the settings turns a larger count into an error.
static int size_entries(Raster* r, hipStream_t s, const uint32_t* total_dev, uint32_t n) {
  uint32_t need = 0;
  hipError_t e = hipMemcpyAsync(&need, total_dev, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "rdx_raster_forward: entry count");
  const uint64_t limit = r->set.max_entries_per_splat ? (uint64_t)r->set.max_entries_per_splat * n : 0x7fffffffull;
  if (need > std::min<uint64_t>(limit, 0x7fffffffull))
    return fail(RDX_E_CAPACITY, "rdx_raster_forward: %u (splat, tile) entries exceed the limit of %llu", need,
                (unsigned long long)std::min<uint64_t>(limit, 0x7fffffffull));
  return reserve_entries(r, need);
}

static uint32_t grid_for(size_t items, uint32_t block = 256, uint32_t max_grid = 1u << 20) {
  return (uint32_t)std::max<size_t>(1, std::min<size_t>((items + block - 1) / block, max_grid));
}

}  // namespace rdx

using namespace rdx;

extern "C" {

int rdx_raster_create(rdx_context* ctx, const rdx_raster_settings* s, rdx_raster** out) {
  if (!out) return fail(RDX_E_INVALID, "rdx_raster_create: out is NULL");
  *out = nullptr;
  if (!ctx) return fail(RDX_E_INVALID, "rdx_raster_create: no context");
  Raster* r = new Raster();
  r->ctx = reinterpret_cast<Context*>(ctx);
  if (s) {
    r->set = *s;
  } else {
    memset(&r->set, 0, sizeof r->set);
    r->set.sh_degree = 3;
    r->set.t_stop = 1e-4f;
  }
  if (!(r->set.t_stop > 0.0f)) r->set.t_stop = 1e-4f;
  *out = reinterpret_cast<rdx_raster*>(r);
  return RDX_OK;
}

int rdx_raster_destroy(rdx_raster* raster) {
  if (!raster) return RDX_OK;
  Raster* r = reinterpret_cast<Raster*>(raster);
  r->release();
  delete r;
  return RDX_OK;
}

int rdx_raster_forward(rdx_raster* raster, void* stream, const float* params_dev, uint32_t n, const rdx_camera* camera,
                       const rdx_params* params, float* image_dev, float* alpha_dev) {
  if (!raster || !camera || !params || !image_dev || (n && !params_dev))
    return fail(RDX_E_INVALID, "rdx_raster_forward: NULL argument");
  Raster* r = reinterpret_cast<Raster*>(raster);
  r->have_forward = false;  // a forward that fails leaves nothing to differentiate
  hipStream_t s = (hipStream_t)stream;
  DiffU u;
  if (int rc = make_diff_uniforms(camera, params, r->set.sh_degree, r->set.t_stop, &u)) return rc;
  if ((size_t)u.tiles_x * u.tiles_y > 65536) return fail(RDX_E_INVALID, "rdx_raster: more than 65536 tiles");
  if (int rc = reserve(r, n, u)) return rc;
  Context* c = r->ctx;
  uint32_t* counts = (uint32_t*)r->counts.ptr;
  const uint32_t tiles = (uint32_t)(u.tiles_x * u.tiles_y);
  hipLaunchKernelGGL(k_fill_counts, dim3(1), dim3(64), 0, s, counts);
  {
    Stage st(c, s, "diff_pre");
    if (n)
      hipLaunchKernelGGL(k_diff_pre, dim3(grid_for(n)), dim3(256), 0, s, params_dev, n, u, (float4*)r->pre.ptr,
                         (uint32_t*)r->keys.ptr, (uint32_t*)r->touch.ptr, (uint4*)r->rect.ptr);
  }
  // back to front, invalid keys dropped; the visible count on the device
  const SortPlan p = depth_plan();
  SegLayout L;
  L.tiles = tiles_of(p, n);
  if (int rc = run_sort(c, r->sort_splats, s, p, L, (const uint32_t*)r->keys.ptr, nullptr, nullptr, (uint32_t*)r->order.ptr,
                        std::max<uint32_t>(n, 1), n ? nullptr : counts + C_WORDS - 1, true, 0xffffffffu, false,
                        counts + C_VISIBLE, nullptr))
    return rc;
  const uint32_t blocks = (uint32_t)((std::max<uint32_t>(n, 1) + DSCAN_TILE - 1) / DSCAN_TILE);
  {
    Stage st(c, s, "diff_scan");
    hipLaunchKernelGGL(k_diff_scan_partial, dim3(blocks), dim3(DSCAN_BLOCK), 0, s, (const uint32_t*)r->order.ptr,
                       (const uint32_t*)r->touch.ptr, counts + C_VISIBLE, (uint32_t*)r->bsum.ptr);
    hipLaunchKernelGGL(k_diff_scan_blocks, dim3(1), dim3(1024), 0, s, (uint32_t*)r->bsum.ptr, blocks, counts + C_ENTRIES,
                       0xffffffffu, counts + C_OVERFLOW);
  }
  if (int rc = size_entries(r, s, counts + C_ENTRIES, n)) return rc;
  {
    // chunk records of the forward, work items of the adjoint: one per
    // (tile, chunk) at most
    const uint32_t rows = diff_chunk_rows(tiles, r->cap);
    if (int rc = r->chunk_rec.ensure((size_t)rows * DIFF_BLOCK * sizeof(float4))) return rc;
    if (int rc = r->item_tile.ensure((size_t)rows * 4)) return rc;
    if (int rc = r->item_chunk.ensure((size_t)rows * 4)) return rc;
    r->rows = rows;
  }
  {
    Stage st(c, s, "diff_scan");
    hipLaunchKernelGGL(k_diff_scan_final, dim3(blocks), dim3(DSCAN_BLOCK), 0, s, (const uint32_t*)r->order.ptr,
                       (const uint32_t*)r->touch.ptr, counts + C_VISIBLE, (const uint32_t*)r->bsum.ptr,
                       (uint32_t*)r->offsets.ptr);
  }
  {
    Stage st(c, s, "diff_emit");
    hipLaunchKernelGGL(k_diff_emit, dim3(grid_for(n)), dim3(256), 0, s, (const uint32_t*)r->order.ptr, counts + C_VISIBLE,
                       (const uint32_t*)r->touch.ptr, (const uint4*)r->rect.ptr, (const uint32_t*)r->offsets.ptr,
                       (uint32_t)u.tiles_x, r->cap, (uint32_t*)r->tile_key.ptr, (uint32_t*)r->emit_idx.ptr,
                       (uint32_t*)r->emit_splat.ptr, (uint32_t*)r->first_entry.ptr, counts + C_OVERFLOW);
  }
  // stable sort of the entries by tile: front to back within each tile
  const SortPlan pe = entry_plan();
  SegLayout Le;
  Le.tiles = tiles_of(pe, r->cap);
  if (int rc = run_sort(c, r->sort_entries, s, pe, Le, (const uint32_t*)r->tile_key.ptr, (const uint32_t*)r->emit_idx.ptr,
                        (uint32_t*)r->tile_key.ptr, (uint32_t*)r->emit_idx.ptr, r->cap, counts + C_ENTRIES, false, 0u,
                        false, nullptr, nullptr))
    return rc;
  {
    Stage st(c, s, "diff_ranges");
    hipLaunchKernelGGL(k_diff_clear_ranges, dim3(grid_for(tiles)), dim3(256), 0, s, (uint2*)r->range.ptr, tiles);
    hipLaunchKernelGGL(k_diff_ranges, dim3(grid_for(r->cap, 256, 4096)), dim3(256), 0, s, (const uint32_t*)r->tile_key.ptr,
                       counts + C_ENTRIES, r->cap, tiles, (uint2*)r->range.ptr);
  }
  {
    Stage st(c, s, "diff_render");
    hipLaunchKernelGGL(k_diff_render<false>, dim3(tiles), dim3(DIFF_BLOCK), 0, s, (const uint2*)r->range.ptr,
                       (const uint32_t*)r->emit_idx.ptr, (const uint32_t*)r->emit_splat.ptr, (const float4*)r->pre.ptr, u,
                       image_dev, alpha_dev, (float*)r->t_final.ptr, (uint32_t*)r->n_contrib.ptr, DrawItems{},
                       DiffChunks{(float4*)r->chunk_rec.ptr, (uint32_t*)r->tile_reach.ptr});
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "rdx_raster_forward");
  r->u = u;
  r->params = params_dev;
  r->n = n;
  r->have_forward = true;
  return RDX_OK;
}

int rdx_raster_backward(rdx_raster* raster, void* stream, const float* grad_image_dev, float* grad_params_dev) {
  if (!raster || !grad_image_dev || !grad_params_dev) return fail(RDX_E_INVALID, "rdx_raster_backward: NULL argument");
  Raster* r = reinterpret_cast<Raster*>(raster);
  if (!r->have_forward) return fail(RDX_E_INVALID, "rdx_raster_backward: no forward to differentiate");
  hipStream_t s = (hipStream_t)stream;
  Context* c = r->ctx;
  const DiffU& u = r->u;
  const uint32_t n = r->n;
  const uint32_t tiles = (uint32_t)(u.tiles_x * u.tiles_y);
  uint32_t* counts = (uint32_t*)r->counts.ptr;
  {
    Stage st(c, s, "diff_render_bwd");
    hipLaunchKernelGGL(k_diff_bwd_items, dim3(1), dim3(1024), 0, s, (const uint32_t*)r->tile_reach.ptr, tiles,
                       (uint32_t*)r->item_tile.ptr, (uint32_t*)r->item_chunk.ptr, (uint32_t*)r->items.ptr);
    const BwdItems bi{(const uint32_t*)r->items.ptr, (const uint32_t*)r->item_tile.ptr,
                      (const uint32_t*)r->item_chunk.ptr, (const float4*)r->chunk_rec.ptr};
    hipLaunchKernelGGL(k_diff_render_bwd, dim3(r->rows), dim3(DIFF_BLOCK), 0, s, (const uint2*)r->range.ptr,
                       (const uint32_t*)r->emit_idx.ptr, (const uint32_t*)r->emit_splat.ptr, (const float4*)r->pre.ptr, u,
                       (const uint32_t*)r->n_contrib.ptr, grad_image_dev, (float*)r->gentry.ptr, bi);
  }
  if (n) {
    Stage st(c, s, "diff_pre_bwd");
    hipLaunchKernelGGL(k_diff_pre_bwd, dim3(grid_for(n)), dim3(256), 0, s, r->params, n, u, (const uint32_t*)r->touch.ptr,
                       (const uint32_t*)r->first_entry.ptr, (const float*)r->gentry.ptr, counts + C_ENTRIES, r->cap,
                       grad_params_dev);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_raster_backward");
}

int rdx_raster_stats(rdx_raster* raster, void* stream, uint32_t* visible, uint32_t* entries, uint32_t* capacity,
                     int* overflowed) {
  if (!raster) return fail(RDX_E_INVALID, "rdx_raster_stats: no raster");
  Raster* r = reinterpret_cast<Raster*>(raster);
  uint32_t h[C_WORDS] = {0, 0, 0, 0};
  if (r->counts.ptr) {
    hipError_t e = hipMemcpyAsync(h, r->counts.ptr, sizeof h, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rdx_raster_stats");
  }
  if (visible) *visible = h[C_VISIBLE];
  if (entries) *entries = h[C_ENTRIES];
  if (capacity) *capacity = r->cap;
  if (overflowed) *overflowed = h[C_OVERFLOW] != 0;
  return RDX_OK;
}

}  // extern "C"

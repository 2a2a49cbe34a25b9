// Batched projection: many views over many tables in one launch. A block
// takes a chunk of one table's splats and projects each of them for every
// view of that table in its group (up to BATCH_MAX_VIEWS), so a table is
// read once for all the views that draw it. Each view is a segment of the
// batch's keys, with its own digit histograms for the segmented sort.
#pragma once

#include "project_kernels.hpp"

namespace rdx {

constexpr int BATCH_MAX_VIEWS = 8;   // views per group
constexpr int BATCH_HIST_WORDS = 32 * HIST_BINS;  // LDS histograms: views x passes <= 32
constexpr int BATCH_CHUNK = 256 * 16;  // splats per block

struct BatchView {
  EyeU eye;          // object (its table) -> eye
  double h[3];       // the head in object space: SH origin
  uint32_t seg;      // segment index (the view's)
  uint32_t seg_off;  // its first key in the batch
  uint32_t pad;
  uint4* words;      // its ewa words (NULL: not written)
};

struct BatchGroup {
  uint32_t table;       // table index
  uint32_t first_view;  // into the view list
  uint32_t num_views;   // <= BATCH_MAX_VIEWS and views * passes <= 32
  uint32_t block0;      // first block of the group
};

struct BatchU {
  int lod_inflate;
  float ndc_z_min, clip_xy;
  float max_std_dev, min_alpha, blur, pre_blur, min_pr, max_pr, min_contrib;
  float focal_distance, aperture_tan;
  float size[2];
  int sort_bits, cull_keys, passes, rb;
};

struct BatchDesc {
  const BatchView* views;
  const BatchGroup* groups;
  const TableP* tables;
  const int* table_sh_degree;
  const float* table_enc;  // 4 per table
  int num_groups;
};

template <int LAYOUT, int SHC, int DEG, bool HIST>
__global__ __launch_bounds__(256) void k_project_batch(BatchDesc D, BatchU u, uint32_t* keys, uint32_t* hist) {
  constexpr int BLOCK = 256;
  __shared__ uint32_t s_hist[HIST ? BATCH_HIST_WORDS : 1];
  __shared__ uint32_t s_group;

  // The block's group: the last group whose first block is <= blockIdx.x.
  if (threadIdx.x == 0) {
    uint32_t g = 0;
    for (int k = 1; k < D.num_groups; k++)
      if (D.groups[k].block0 <= blockIdx.x) g = (uint32_t)k;
    s_group = g;
  }
  __syncthreads();
  const BatchGroup G = D.groups[s_group];
  const TableP& t = D.tables[G.table];
  const int deg = DEG >= 0 ? DEG : D.table_sh_degree[G.table];
  const uint32_t nv = G.num_views;
  const uint32_t chunk = blockIdx.x - G.block0;
  if constexpr (HIST) {
    for (int x = threadIdx.x; x < (int)nv * u.passes * HIST_BINS; x += BLOCK) s_hist[x] = 0u;
    __syncthreads();
  }
  const float* enc = D.table_enc + 4 * G.table;
  const uint32_t invalid = key_invalid(u.sort_bits);
  const uint32_t lo = chunk * (uint32_t)BATCH_CHUNK;
  const uint32_t hi = min(lo + (uint32_t)BATCH_CHUNK, t.count);

  F32Rsrc rs;
  if constexpr (LAYOUT == LAYOUT_F32) rs = f32_rsrc(t, deg);

  for (uint32_t i = lo + threadIdx.x; i < hi; i += BLOCK) {
    Splat sp;
    float3 coef[15];
    load_splat<LAYOUT, SHC, DEG>(t, rs, enc, i, deg, sp, coef);
    float op, std_dev;
    float3 s;
    bool invisible;
    prepare(sp, u, op, s, std_dev, invisible);

    // Culls of every view first: the rotation only when some view draws it.
    uint32_t pass_mask = 0;
    for (uint32_t v = 0; v < nv; v++)
      if (!invisible && eye_centre(D.views[G.first_view + v].eye, sp.c, u.ndc_z_min, u.clip_xy).pass) pass_mask |= 1u << v;
    float R[9];
    if (pass_mask) quat_matrix(sp.q, R);
    for (uint32_t v = 0; v < nv; v++) {
      const BatchView& V = D.views[G.first_view + v];
      uint4 words = make_uint4(0u, 0u, 0u, 0u);
      bool visible = false;
      const EyeCentre C = eye_centre(V.eye, sp.c, u.ndc_z_min, u.clip_xy);
      if ((pass_mask >> v) & 1u) {
        const float3 rgb = sh_colour(sp, V.h, deg, coef);
        const EwaResult r = ewa_eye(V.eye, C, true, R, s, op, std_dev, rgb, u);
        if (r.ok) {
          words = r.words;
          visible = true;
        }
      }
      if (V.words) V.words[i] = words;
      const bool keyed = u.cull_keys ? visible : !invisible;
      const uint32_t key = keyed ? sort_key(fbits((float)C.dist), u.sort_bits) : invalid;
      keys[V.seg_off + i] = key;
      if constexpr (HIST) {
        if (key != invalid) hist_add_key(s_hist + v * u.passes * HIST_BINS, key, u.passes, u.rb);
      }
    }
  }
  if constexpr (HIST) {
    __syncthreads();
    for (uint32_t v = 0; v < nv; v++)
      hist_flush<BLOCK>(s_hist + v * u.passes * HIST_BINS, hist + (size_t)D.views[G.first_view + v].seg * HIST_WORDS,
                        u.passes);
  }
}

}  // namespace rdx

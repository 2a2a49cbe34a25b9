// Batched projection: many views over many tables in one launch. A block
// takes 256 splats of one table and projects each of them for every view of
// that table in its group (up to BATCH_MAX_VIEWS), so a table is read once
// for all the views that draw it. Each view is a segment of the batch's keys.
#pragma once

#include "project_kernels.hpp"

namespace rdx {

constexpr int BATCH_MAX_VIEWS = 8;  // views per group

struct BatchView {
  EyeU eye;          // object (its table) -> eye
  double h[3];       // the eye in object space: the SH origin
  uint32_t seg_off;  // the view's first key in the batch
  uint32_t pad;
  uint4* words;      // its ewa words (NULL: not written)
};

struct BatchGroup {
  uint32_t table;       // table index
  uint32_t first_view;  // into the view list
  uint32_t num_views;   // <= BATCH_MAX_VIEWS
  uint32_t block0;      // first block of the group in its launch
};

// Uniforms shared by every view of a batch (one rdx_params).
struct BatchU {
  int lod_inflate;
  float ndc_z_min, clip_xy;
  float max_std_dev, min_alpha, blur, pre_blur, min_pr, max_pr, min_contrib;
  float focal_distance, aperture_tan;
  int sort_bits, cull_keys;
};

struct BatchDesc {
  const BatchView* views;
  const BatchGroup* groups;
  const TableP* tables;
  const int* table_sh_degree;
  const float* table_enc;  // 4 per table
  int num_groups;
};

template <int LAYOUT, int SHC, int DEG>
__global__ __launch_bounds__(256) void k_project_batch(BatchDesc D, BatchU u, uint32_t* keys) {
  constexpr int BLOCK = 256;
  __shared__ uint32_t s_group;

  // The block's group: the last group whose first block is <= blockIdx.x.
  if (threadIdx.x == 0) {
    uint32_t g = 0;
    for (int k = 1; k < D.num_groups; k++)
      if (D.groups[k].block0 <= blockIdx.x) g = (uint32_t)k;
    s_group = g;
  }
  __syncthreads();
  // uniform across the block: the group's and its views' fields load through
  // the scalar cache into SGPRs instead of once per lane
  const BatchGroup G = D.groups[__builtin_amdgcn_readfirstlane(s_group)];
  const TableP& t = D.tables[G.table];
  const int deg = DEG >= 0 ? DEG : D.table_sh_degree[G.table];
  const uint32_t nv = G.num_views;
  const uint32_t chunk = blockIdx.x - G.block0;
  const float* enc = D.table_enc + 4 * G.table;
  const uint32_t invalid = key_invalid(u.sort_bits);

  F32Rsrc rs;
  if constexpr (LAYOUT == LAYOUT_F32) rs = f32_rsrc<SHC>(t, deg);

  const uint32_t i = chunk * BLOCK + threadIdx.x;
  if (i < t.count) {
    Splat sp;
    float3 coef[15];
    float band[3];
    load_splat<LAYOUT, SHC, DEG>(t, rs, enc, i, deg, sp, coef, band);
    float op, std_dev;
    float3 s;
    bool invisible;
    prepare<LAYOUT == LAYOUT_COV>(sp, u, op, s, std_dev, invisible);
    float R[9];
    bool have_r = false;
    for (uint32_t v = 0; v < nv; v++) {
      const BatchView& V = D.views[G.first_view + v];
      uint4 words = make_uint4(0u, 0u, 0u, 0u);
      bool visible = false;
      EyeCentre C = eye_centre(V.eye, sp.c, u.ndc_z_min, u.clip_xy);
      C.pass = C.pass && !invisible;
      if (C.pass) {
        const float3 rgb = sh_colour(sp, V.h, deg, coef, SHC == SHC_PACKED ? band : nullptr);
        EwaResult r;
        if constexpr (LAYOUT == LAYOUT_COV) {
          r = ewa_eye_cov(V.eye, C, true, sp.cv, op, std_dev, encode_ext_rgb(rgb.x, rgb.y, rgb.z), u);
        } else {
          if (!have_r) {
            quat_matrix(sp.q, R);
            have_r = true;
          }
          float Q[9];
          eye_axes(V.eye, R, Q);
          r = ewa_eye(V.eye, C, true, R, Q, s, op, std_dev, encode_ext_rgb(rgb.x, rgb.y, rgb.z), u);
        }
        if (r.ok) {
          words = r.words;
          visible = true;
        }
      }
      if (V.words) st_stream(V.words + i, words);
      // Mono views: the head is the eye, the key distance the eye's.
      const bool keyed = u.cull_keys ? visible : !invisible;
      const uint32_t key = keyed ? sort_key(fbits((float)C.dist), u.sort_bits) : invalid;
      keys[V.seg_off + i] = key;
    }
  }
}

}  // namespace rdx

// Host side of the projection: uniforms from the reference's camera and
// object inputs, table pointers, rdx_project_ewa and the fused
// rdx_project_sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "project_host.hpp"

namespace rdx {

// Solve L x = b for a row-major 3x3 (Cramer).
static void solve3(const double* L, const double* b, double* x) {
  const double d = det3(L);
  for (int k = 0; k < 3; k++) {
    double M[9];
    memcpy(M, L, sizeof M);
    for (int r = 0; r < 3; r++) M[3 * r + k] = b[r];
    x[k] = det3(M) / d;
  }
}

int make_uniforms(const rdx_splat_table* t, const rdx_camera* cam, const rdx_params* prm, ProjU* u) {
  memset(u, 0, sizeof *u);
  if (cam->eyes < 1 || cam->eyes > 2) return fail(RDX_E_INVALID, "SplatCamera: 1 or 2 eyes, not %d", cam->eyes);
  double A[16], Ainv[16], headInv[16];
  mul(cam->head, t->object_to_world, A);  // object to head
  if (!invert(A, Ainv)) return fail(RDX_E_INVALID, "object to head transform is singular");
  if (!invert(cam->head, headInv)) return fail(RDX_E_INVALID, "head matrix is singular");
  for (int k = 0; k < 3; k++) u->h[k] = cm(Ainv, k, 3);
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) u->H[3 * r + c] = cm(A, r, c);
  u->eyes = cam->eyes;
  u->head_is_eye0 = 0;
  for (int e = 0; e < cam->eyes; e++) {
    EyeU& E = u->eye[e];
    double F[16], G[16];
    const bool same = memcmp(cam->eye[e].view, cam->head, sizeof(double) * 16) == 0;
    if (same) {
      memcpy(G, A, sizeof G);  // fromHead is exactly the identity (splatCamera.ts:128-129)
      if (e == 0) u->head_is_eye0 = 1;
    } else {
      mul(cam->eye[e].view, headInv, F);
      mul(F, A, G);
    }
    double tr[3];
    for (int r = 0; r < 3; r++) {
      for (int c = 0; c < 3; c++) E.L[3 * r + c] = cm(G, r, c);
      tr[r] = cm(G, r, 3);
    }
    const double neg_t[3] = {-tr[0], -tr[1], -tr[2]};
    solve3(E.L, neg_t, E.e);  // the eye in object space: L e + t = 0
    const double dl = det3(E.L);
    if (!(dl != 0.0)) return fail(RDX_E_INVALID, "object to eye transform is singular");
    E.det_l2 = (float)(dl * dl);
    const double z[3] = {0.0, 0.0, 1.0};
    double ol[3];
    solve3(E.L, z, ol);
    for (int k = 0; k < 3; k++) E.ortho_l[k] = (float)ol[k];
    for (int k = 0; k < 9; k++) E.Lf[k] = (float)E.L[k];
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) E.Pf[4 * r + c] = (float)cm(cam->eye[e].projection, r, c);
    E.fx = (float)(cam->size[0] * 0.5 * cm(cam->eye[e].projection, 0, 0));
    E.fy = (float)(cam->size[1] * 0.5 * cm(cam->eye[e].projection, 1, 1));
    E.ortho = cam->eye[e].ortho ? 1 : 0;
  }
  // eyes that differ by a translation have the same float L: one Q = L R serves both
  u->shared_axes = cam->eyes == 2 && memcmp(u->eye[0].Lf, u->eye[1].Lf, sizeof u->eye[0].Lf) == 0 ? 1 : 0;
  u->size[0] = (float)cam->size[0];
  u->size[1] = (float)cam->size[1];
  u->ndc_z_min = (float)cam->ndc_z_min;
  u->max_std_dev = (float)prm->max_std_dev;
  u->min_alpha = (float)prm->min_alpha;
  u->blur = (float)prm->blur_amount;
  u->pre_blur = (float)prm->pre_blur_amount;
  u->clip_xy = (float)prm->clip_xy;
  u->min_pr = (float)prm->min_pixel_radius;
  u->max_pr = (float)prm->max_pixel_radius;
  u->min_contrib = (float)prm->min_contribution;
  u->focal_distance = (float)prm->focal_distance;
  u->aperture_tan = (float)prm->aperture_tan;
  u->sort_guard = (float)prm->sort_guard;
  u->lod_inflate = prm->lod_inflate ? 1 : 0;
  const int table_deg = t->sh_codec == RDX_SH_NONE ? 0 : t->sh_degree;
  u->sh_degree = std::max(0, std::min(std::min(table_deg, prm->max_sh_degree), 3));
  u->sort_bits = prm->sort_bits;
  u->cull_keys = prm->cull_keys ? 1 : 0;
  u->key_per_eye = prm->key_per_eye ? 1 : 0;
  for (int k = 0; k < 4; k++) u->enc[k] = t->packed_encoding[k];
  // isUniformScale (splatTransform.tsl.ts:46-62): G = M^T M within 1e-6 s2 of
  // s2 I (s2 = trace(G) / 3) and det M > 0; the scale is cbrt |det M|
  {
    double g[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) g[i][j] = cm(A, 0, i) * cm(A, 0, j) + cm(A, 1, i) * cm(A, 1, j) + cm(A, 2, i) * cm(A, 2, j);
    const double s2 = (g[0][0] + g[1][1] + g[2][2]) / 3.0;
    bool uniform = s2 > 0.0;
    for (int i = 0; i < 3 && uniform; i++)
      for (int j = i; j < 3; j++)
        if (fabs(g[i][j] - (i == j ? s2 : 0.0)) > 1e-6 * s2) uniform = false;
    const double dA = det3(u->H);
    u->head_uniform = uniform && dA > 0.0 ? 1 : 0;
    u->head_scale = (float)cbrt(fabs(dA));
  }
  return RDX_OK;
}

TableP make_table(const rdx_splat_table* t) {
  TableP p;
  memset(&p, 0, sizeof p);
  p.count = t->count;
  p.stride = (uint32_t)(t->plane_stride ? t->plane_stride : t->count);
  p.center = t->center;
  p.log_scale = t->log_scale;
  p.quat = t->quat;
  p.opacity = t->opacity;
  p.rgb = t->rgb;
  p.sh = t->sh;
  p.ext_a = (const uint4*)t->ext_a;
  p.ext_b = (const uint4*)t->ext_b;
  p.packed = (const uint4*)t->packed;
  for (int k = 0; k < 4; k++) p.shc[k] = (const uint4*)t->sh_channels[k];
  return p;
}

int check_table(const rdx_splat_table* t) {
  if (!t) return fail(RDX_E_INVALID, "no splat table");
  const int deg = t->sh_codec == RDX_SH_NONE ? 0 : t->sh_degree;
  if (deg < 0 || deg > 3) return fail(RDX_E_UNSUPPORTED, "SH degree %d: degrees 0 to 3 are supported", t->sh_degree);
  switch (t->layout) {
    case RDX_LAYOUT_F32_SOA:
      if (t->count && (!t->center || !t->log_scale || !t->quat || !t->opacity || !t->rgb))
        return fail(RDX_E_INVALID, "f32 table: missing planes");
      if (deg > 0 && (t->sh_codec != RDX_SH_F32 || !t->sh))
        return fail(RDX_E_INVALID, "f32 table: SH must be f32 planes");
      if (t->plane_stride && t->plane_stride < t->count) return fail(RDX_E_INVALID, "f32 table: plane stride below count");
      // planes are read with 32-bit offsets within groups of at most 15 planes (4 GiB each)
      if ((uint64_t)(t->plane_stride ? t->plane_stride : t->count) > F32_MAX_STRIDE)
        return fail(RDX_E_UNSUPPORTED, "f32 table: %zu splats per plane exceed the limit of %llu",
                    (size_t)(t->plane_stride ? t->plane_stride : t->count), (unsigned long long)F32_MAX_STRIDE);
      break;
    case RDX_LAYOUT_EXT:
      if (t->count && (!t->ext_a || !t->ext_b)) return fail(RDX_E_INVALID, "extSplats table: missing words");
      break;
    case RDX_LAYOUT_COV:
      if (t->count && (!t->ext_a || !t->ext_b)) return fail(RDX_E_INVALID, "covSplats table: missing words");
      break;
    case RDX_LAYOUT_PACKED:
      if (t->count && !t->packed) return fail(RDX_E_INVALID, "packedSplats table: missing words");
      if (!(t->packed_encoding[1] != t->packed_encoding[0]) || !(t->packed_encoding[3] != t->packed_encoding[2]))
        return fail(RDX_E_INVALID, "packedSplats table: empty encoding range");
      break;
    default:
      return fail(RDX_E_INVALID, "unknown splat layout %d", t->layout);
  }
  if (t->layout != RDX_LAYOUT_F32_SOA && deg > 0) {
    if (t->sh_codec != RDX_SH_EXT && t->sh_codec != RDX_SH_PACKED)
      return fail(RDX_E_INVALID, "word tables take ext or packed SH");
    const int texels = t->sh_codec == RDX_SH_EXT ? (deg == 1 ? 1 : deg == 2 ? 2 : 4) : deg;
    for (int k = 0; k < texels; k++)
      if (!t->sh_channels[k]) return fail(RDX_E_INVALID, "SH channel %d missing", k);
  }
  return RDX_OK;
}

int launch_project(hipStream_t s, const rdx_splat_table* t, const ProjU& u, const OutP& o, bool side) {
  const uint32_t n = t->count;
  if (n == 0) return RDX_OK;
  // One thread per splat, grid x 256 >= n (up to 2^24 blocks): many short
  // blocks keep the machine full to the end.
  const uint32_t grid = (uint32_t)(((uint64_t)n + 255) / 256);
  const TableP tp = make_table(t);
  const int shc = t->sh_codec == RDX_SH_PACKED ? SHC_PACKED : SHC_EXT;
  switch (t->layout) {
    case RDX_LAYOUT_F32_SOA: launch_project_f32(s, dim3(grid), tp, u, o, u.eyes, side); break;
    case RDX_LAYOUT_EXT: launch_project_ext(s, dim3(grid), shc, tp, u, o, u.eyes, side); break;
    case RDX_LAYOUT_COV: launch_project_cov(s, dim3(grid), shc, tp, u, o, u.eyes, side); break;
    case RDX_LAYOUT_PACKED: launch_project_packed(s, dim3(grid), shc, tp, u, o, u.eyes, side); break;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "projection launch");
}

static void fill_side(OutP& o, const rdx_float_out* side, uint32_t n) {
  if (!side) return;
  o.sstride = side->stride ? side->stride : n;
  o.conic = side->conic;
  o.cov2d = side->cov2d;
  o.opac = side->opacity;
  o.ndc = side->ndc;
  o.vz = side->view_z;
  o.bound = side->bound;
  o.rgb = side->rgb;
  o.dist = side->distance;
}

}  // namespace rdx

using namespace rdx;

extern "C" {

void rdx_params_default(rdx_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  // src/render/splatParams.ts:68-72
  p->max_std_dev = sqrt(8.0);
  p->min_alpha = 0.5 / 255.0;
  p->falloff = 1.0;
  p->lod_inflate = 1;
  p->blur_amount = 0.3;
  p->pre_blur_amount = 0.0;
  p->clip_xy = 1.4;
  p->min_pixel_radius = 0.0;
  p->max_pixel_radius = 512.0;
  p->min_contribution = 0.0;
  p->focal_distance = 0.0;
  p->aperture_tan = 0.0;
  p->sort_guard = 0.0;
  p->max_sh_degree = 4;
  p->sort_bits = 24;  // SplatRenderer.ts:163
  p->cull_keys = 1;   // SplatRenderer.ts:432
  p->key_per_eye = 0;
}

int rdx_project_ewa(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                    const rdx_params* params, uint32_t at, uint32_t* const* words, uint32_t* keys, size_t key_stride,
                    const rdx_float_out* side) {
  if (!camera || !params) return fail(RDX_E_INVALID, "rdx_project_ewa: NULL camera or params");
  if (int rc = check_table(table)) return rc;
  if (int rc = check_bits(params->sort_bits, "SplatProjection")) return rc;
  ProjU u;
  if (int rc = make_uniforms(table, camera, params, &u)) return rc;
  OutP o;
  memset(&o, 0, sizeof o);
  for (int e = 0; e < camera->eyes; e++) o.words[e] = words ? (uint4*)words[e] : nullptr;
  o.keys = keys;
  o.key_stride = key_stride;
  o.at = at;
  fill_side(o, side, table->count);
  if (params->key_per_eye && camera->eyes == 2 && keys && key_stride < (size_t)at + table->count)
    return fail(RDX_E_INVALID, "rdx_project_ewa: key_stride below the keys written");
  Stage st(reinterpret_cast<Context*>(ctx), (hipStream_t)stream, "project");
  return launch_project((hipStream_t)stream, table, u, o, side != nullptr);
}

int rdx_project_keys(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                     const rdx_params* params, uint32_t at, uint32_t* keys) {
  if (!camera || !params || !keys) return fail(RDX_E_INVALID, "rdx_project_keys: NULL camera, params or keys");
  if (int rc = check_table(table)) return rc;
  if (int rc = check_bits(params->sort_bits, "SplatProjection")) return rc;
  ProjU u;
  if (int rc = make_uniforms(table, camera, params, &u)) return rc;
  const uint32_t n = table->count;
  if (n == 0) return RDX_OK;
  const TableP tp = make_table(table);
  const dim3 grid((uint32_t)(((uint64_t)n + 255) / 256));
  hipStream_t s = (hipStream_t)stream;
  Stage st(reinterpret_cast<Context*>(ctx), s, "project_keys");
  switch (table->layout) {
    case RDX_LAYOUT_F32_SOA: hipLaunchKernelGGL(k_project_keys<LAYOUT_F32>, grid, dim3(256), 0, s, tp, u, keys, at); break;
    case RDX_LAYOUT_EXT: hipLaunchKernelGGL(k_project_keys<LAYOUT_EXT>, grid, dim3(256), 0, s, tp, u, keys, at); break;
    case RDX_LAYOUT_COV: hipLaunchKernelGGL(k_project_keys<LAYOUT_COV>, grid, dim3(256), 0, s, tp, u, keys, at); break;
    case RDX_LAYOUT_PACKED: hipLaunchKernelGGL(k_project_keys<LAYOUT_PACKED>, grid, dim3(256), 0, s, tp, u, keys, at); break;
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_project_keys");
}

int rdx_project_sort(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                     const rdx_params* params, uint32_t* const* words, uint32_t* keys, uint32_t* out_indices,
                     uint32_t* out_count, uint32_t* out_args, const rdx_float_out* side) {
  if (!ctx) return fail(RDX_E_INVALID, "rdx_project_sort: no context");
  if (!camera || !params) return fail(RDX_E_INVALID, "rdx_project_sort: NULL camera or params");
  if (int rc = check_table(table)) return rc;
  if (int rc = check_bits(params->sort_bits, "WgpuRadixSort")) return rc;
  if (!out_indices) return fail(RDX_E_INVALID, "rdx_project_sort: no output indices");
  Context* c = reinterpret_cast<Context*>(ctx);
  hipStream_t s = (hipStream_t)stream;
  const uint32_t n = table->count;
  // Per-eye keys (stereo): two segments, eye e's keys, indices, count and
  // draw arguments at keys + e n, out_indices + e n, out_count + e and
  // out_args + 5 e.
  const int segs = params->key_per_eye && camera->eyes == 2 ? 2 : 1;
  if ((uint64_t)n * segs > 0xffffffffull) return fail(RDX_E_CAPACITY, "rdx_project_sort: %u splats per eye", n);
  ProjU u;
  if (int rc = make_uniforms(table, camera, params, &u)) return rc;
  const SortPlan p = compact_plan(c->tuning, params->sort_bits);
  SegLayout L;
  const uint32_t nt = tiles_of(p, n);
  if (int rc = c->sort.reserve(n * segs, segs, nt * segs)) return rc;
  if (segs == 2) {
    const uint32_t off[2] = {0u, n}, cnt[2] = {n, n};
    if (int rc = small_segments(c->sort, s, p, off, cnt, 2, &L)) return rc;
  } else {
    L.tiles = nt;
  }
  if (!keys) {
    if (int rc = c->proj.ensure((size_t)n * segs * 4)) return rc;
    keys = (uint32_t*)c->proj.ptr;
  }
  OutP o;
  memset(&o, 0, sizeof o);
  for (int e = 0; e < camera->eyes; e++) o.words[e] = words ? (uint4*)words[e] : nullptr;
  o.keys = keys;
  o.key_stride = n;
  o.at = 0;
  fill_side(o, side, n);
  {
    Stage st(c, s, "project");
    if (int rc = launch_project(s, table, u, o, side != nullptr)) return rc;
  }
  return run_sort(c, c->sort, s, p, L, keys, nullptr, nullptr, out_indices, n * segs, nullptr, true,
                  key_invalid(params->sort_bits), false, out_count, out_args);
}

}  // extern "C"

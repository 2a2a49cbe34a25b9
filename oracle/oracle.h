/*
 * The oracle: a CPU restatement, in plain C (FP64 for the float math, exact
 * integers for keys and orders), of the reference's per-frame splat path.
 *
 * TEST INFRASTRUCTURE ONLY. It is the checker of the HIP path: only the
 * tests, the smoke check and the bench's CPU-baseline leg load it. The
 * product (radix_amd, libradix_hip.so) never calls it.
 *
 * Each function cites the reference file:line it restates. What pins each
 * function (golden vectors, known answers, closed forms, brute force) is in
 * the tests (tests/test_oracle_*.py); every function here is pinned.
 */
#ifndef RADIX_ORACLE_H
#define RADIX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- src/core/sort.ts ---------------------------------------------------- */
uint32_t orc_key_invalid(int bits);                 /* sort.ts:37-39 */
uint32_t orc_key_valid_max(int bits);               /* sort.ts:41-43 */
uint32_t orc_key_compute(double distance, int bits); /* sort.ts:46-58 */
void orc_keys_compute(const double* distance, size_t n, int bits, uint32_t* keys);
/* SortKeys.sort (sort.ts:64-163): stable LSD (1x16, 11+13, 2x16); returns
 * visible; ordering must hold count. Scratch is per call (malloc). */
int64_t orc_sort(const uint32_t* keys, uint32_t count, int bits, uint32_t* ordering);

/* ---- src/core/encoding.ts ------------------------------------------------ */
uint32_t orc_encode_half(double f);           /* encoding.ts:156-189 (via float32) */
double orc_decode_half(uint32_t h);           /* encoding.ts:191-202 */
uint32_t orc_encode_ext_rgb(double r, double g, double b); /* encoding.ts:219-231 */
void orc_decode_ext_rgb(uint32_t e, double* rgb);          /* encoding.ts:233-239 */
uint32_t orc_encode_oct(const double* v, int bits);        /* encoding.ts:25-47 */
void orc_decode_oct(uint32_t code, int bits, double* out); /* encoding.ts:49-66 */
uint32_t orc_encode_quat_oct(const double* q, int axis_bits, int angle_bits); /* encoding.ts:85-98 */
void orc_decode_quat_oct(uint32_t code, int axis_bits, int angle_bits, double* q); /* encoding.ts:100-110 */

/* ---- src/core/splats.ts packed and ext codecs ---------------------------- */
/* splats.ts:271-309; enc = rgbMin, rgbMax, lnScaleMin, lnScaleMax */
void orc_packed_encode(const double* center, const double* scales, const double* quat, const double* rgb,
                       double opacity, const double* enc, uint32_t* words4);
void orc_packed_decode(const uint32_t* words4, const double* enc, double* center, double* scales, double* quat,
                       double* rgb, double* opacity);
/* splats.ts:173-197 (no PBR fields: normal none, roughness and metallic 0) */
void orc_ext_encode(const double* center, const double* scales, const double* quat, const double* rgb,
                    double opacity, uint32_t* a4, uint32_t* b4);
void orc_ext_decode(const uint32_t* a4, const uint32_t* b4, double* center, double* scales, double* quat,
                    double* rgb, double* opacity);

/* splats.ts:219-266: cov words from the object-space covariance xx, yy, zz,
 * xy, xz, yz (normal none) and back */
void orc_cov_encode(const double* center, const double* cov6, const double* rgb, double opacity, uint32_t* a4,
                    uint32_t* b4);
void orc_cov_decode(const uint32_t* a4, const uint32_t* b4, double* center, double* cov6, double* rgb,
                    double* opacity);
double orc_cov_encode_correlation(double value, double a, double b); /* splats.ts:221-228 */
double orc_cov_decode_correlation(double encoded);                   /* splats.ts:230-232 */

/* ---- src/core/sh.ts ------------------------------------------------------ */
void orc_sh_basis(const double* dir, double* out24);                       /* sh.ts:376-407 */
void orc_sh_evaluate(const double* coeffs, const double* dir, int degree, double* rgb); /* sh.ts:362-373 */
void orc_sh_apply_table(size_t n, size_t stride, const float* center, const float* sh, int degree, const double* origin,
                        float* rgb);
uint32_t orc_sh_scale_encode(double max);                                  /* sh.ts:71-80 */
double orc_sh_scale_decode(uint32_t code);                                 /* sh.ts:82-84 */
int orc_sh_words(int codec, int degree); /* codec 2 ext, 3 packed (sh.ts:114-174) */
void orc_sh_encode(int codec, const double* coeffs, int degree, uint32_t* words);
void orc_sh_decode(int codec, const uint32_t* words, int degree, double* coeffs);

/* ---- src/core/math.ts ---------------------------------------------------- */
void orc_cov_from_scale_quat(const double* scales, const double* quat, double* cov6); /* math.ts:29-45 */

/* ---- The ewa projection (the render TSL restated in FP64) ----------------- */

typedef struct orc_splat {
  double center[3];
  double cov[6];     /* xx, yy, zz, xy, xz, yz (object space) */
  double rgb[3];     /* base colour plus SH, clamped at 0 (object space view) */
  double opacity;
} orc_splat;

typedef struct orc_params {
  double max_std_dev, min_alpha, falloff;
  int lod_inflate;
  double blur_amount, pre_blur_amount, clip_xy, min_pixel_radius, max_pixel_radius, min_contribution;
  double focal_distance, aperture_tan, sort_guard;
  int max_sh_degree, sort_bits, cull_keys, key_per_eye;
} orc_params;

/* One eye, prepared on the host: head to eye (4x4, column-major) and the
 * projection (column-major), ortho flag. */
typedef struct orc_eye {
  double from_head[16];
  double projection[16];
  int ortho;
} orc_eye;

typedef struct orc_view {
  double object_to_head[16]; /* column-major */
  double view_origin[3];     /* head position in object space (SH origin) */
  int eyes;
  orc_eye eye[2];
  double size[2];
  double ndc_z_min;
} orc_view;

typedef struct orc_eye_out {
  int visible;        /* drawn in this eye */
  uint32_t words[4];  /* zero when culled */
  double ndc[2], view_z;
  double a, b, d;     /* blurred 2D covariance */
  double conic[3];    /* (d, -b, a) / det */
  double opacity;     /* blur-compensated */
  double scale1, scale2, angle;
  double lambda1, lambda2;
  /* Culling margins (relative distance to a threshold; small: near a flip) */
  double margin;
  double distance;    /* |eye-space centre| */
  uint32_t key;       /* with key_per_eye */
} orc_eye_out;

typedef struct orc_out {
  orc_eye_out eye[2];
  double rgb[3];
  double distance;    /* |head-space centre| */
  int invisible, guarded, visible;
  uint32_t key;
  double margin;      /* smallest relative margin over the culls deciding the key */
} orc_out;

void orc_params_default(orc_params* p);
/* SplatProjection.project + write for ewa (SplatProjection.ts:131-153,
 * projections.tsl.ts:95-148, projection.tsl.ts:42-140, 215-240, 277-310),
 * for a splat already in object space with its SH colour applied. */
void orc_project_ewa(const orc_splat* s, const orc_view* v, const orc_params* p, orc_out* out);

/* A whole f32-SoA table (planes of `stride` floats) through the path:
 * SH colour (splatGenerator.ts:152-158, sh.tsl.ts:255-266), cov from
 * log-scales and quat, projection and key. Outputs per splat (arrays of n;
 * per eye arrays of 2n with eye e at [e*n]). */
typedef struct orc_table_f32 {
  uint32_t count;
  size_t stride;
  const float *center, *log_scale, *quat, *opacity, *rgb, *sh;
  int sh_degree;
} orc_table_f32;

typedef struct orc_batch_out {
  uint32_t* words;    /* per eye 4n */
  uint32_t* keys;     /* n (2n with key_per_eye) */
  double* conic;      /* per eye 3n (planes) */
  double* opacity;    /* per eye n */
  double* ndc;        /* per eye 2n */
  double* view_z;     /* per eye n */
  double* rgb;        /* 3n */
  double* distance;   /* n (2n with key_per_eye) */
  double* margin;     /* n */
  uint8_t* visible;   /* per eye n */
} orc_batch_out;

void orc_project_table_f32(const orc_table_f32* t, const orc_view* v, const orc_params* p, orc_batch_out* out);
/* The keys-only pass (SplatProjection.writeKeys) over an f32 table: the
 * splat-form sortGuard; margin[i]: relative distance to the nearest cull
 * threshold (may be NULL). */
void orc_project_keys_table_f32(const orc_table_f32* t, const orc_view* v, const orc_params* p, uint32_t* keys,
                                double* margin);

/* Decode then project a packed + packedSh table or an ext + extSh table
 * (uint32x4 arrays; sh_channels: one uint32x4 array per texel). */
void orc_project_table_packed(uint32_t count, const uint32_t* packed, const double* enc, int sh_codec, int sh_degree,
                              const uint32_t* const* sh_channels, const orc_view* v, const orc_params* p,
                              orc_batch_out* out);
void orc_project_table_ext(uint32_t count, const uint32_t* ext_a, const uint32_t* ext_b, int sh_codec, int sh_degree,
                           const uint32_t* const* sh_channels, const orc_view* v, const orc_params* p,
                           orc_batch_out* out);
/* A covSplats table (covA, covB words) with ext or packed SH: the cov form
 * the ewa method reads (GpuCovSplatReader.loadCovSplat, splatAccess.ts:107-109) */
void orc_project_table_cov(uint32_t count, const uint32_t* cov_a, const uint32_t* cov_b, int sh_codec, int sh_degree,
                           const uint32_t* const* sh_channels, const orc_view* v, const orc_params* p,
                           orc_batch_out* out);

/* Prepare an orc_view from the reference's camera inputs (setViews,
 * splatCamera.ts:119-138; setTransforms, splatTransform.tsl.ts:78-107):
 * world-to-head and per-eye world-to-eye views (column-major). */
void orc_view_from_camera(const double* object_to_world, const double* head, int eyes, const double* const* eye_views,
                          const double* const* eye_projections, const int* ortho, double width, double height,
                          double ndc_z_min, orc_view* out);

/* ---- Differentiable compositor (FP64): forward, loss, analytic gradient -- */

typedef struct orc_raster_cfg {
  int width, height;
  double background[3];
  double t_min;
} orc_raster_cfg;

/* Parameters in double, SoA planes of n. */
typedef struct orc_diff_params {
  uint32_t count;
  int sh_degree;
  const double *means, *log_scales, *quats, *opacity_logits, *sh;
} orc_diff_params;

/* Render (front to back, per pixel sorted by view depth, ties by index).
 * image: 3 planes of w*h. */
void orc_raster_forward(const orc_diff_params* P, const orc_view* v, const orc_params* p, const orc_raster_cfg* c,
                        double* image);
/* L = sum (image - target)^2 and its gradient wrt every parameter (planes as
 * in P; grads arrays sized like the params). Returns L. */
double orc_raster_loss_grad(const orc_diff_params* P, const orc_view* v, const orc_params* p, const orc_raster_cfg* c,
                            const double* target, double* g_means, double* g_log_scales, double* g_quats,
                            double* g_opacity_logits, double* g_sh);
/* The loss only (for finite differences), recomputing the pixels in the
 * footprint of splat `only` (or all pixels if only < 0) from `base` (the
 * unperturbed image, may be NULL when only < 0). */
double orc_raster_loss(const orc_diff_params* P, const orc_view* v, const orc_params* p, const orc_raster_cfg* c,
                       const double* target, int64_t only, const double* base, double base_loss);
/* Pixel set of splat i's footprint (bbox) in the current params: x0,y0,x1,y1. */
void orc_raster_footprint(const orc_diff_params* P, const orc_view* v, const orc_params* p, const orc_raster_cfg* c,
                          uint32_t i, int* box);

/* ---- The differentiable compositor and its adjoint (raster.c) ------------- */
#define ORC_DIFF_PARAMS 59 /* centre 3, log-scale 3, quat 4, opacity logit 1, rgb 3, SH 45 */
typedef struct OrcRaster {
  int width, height;
  double view[16]; /* world -> view, column-major */
  double proj[16]; /* column-major */
  double ndc_z_min, max_std_dev, min_alpha, blur, clip_xy, t_stop;
  int sh_degree;
} OrcRaster;
typedef struct OrcDiffFrozen OrcDiffFrozen;
int orc_diff_render(const double* theta, int n, const OrcRaster* r, double* image, double* alpha);
double orc_diff_loss(const double* theta, int n, const OrcRaster* r, const double* target);
/* 1 (the default): pixels walk 16x16 tile lists of the splats that can reach them; 0: every splat (the same result). */
void orc_raster_binning(int on);
double orc_diff_grad(const double* theta, int n, const OrcRaster* r, const double* target, double* grad);
/* the loss with every pixel's contributor list fixed at theta (smooth in theta) */
OrcDiffFrozen* orc_diff_freeze(const double* theta, int n, const OrcRaster* r, const double* target);
double orc_diff_loss_frozen(const OrcDiffFrozen* f, const double* theta, int n, const OrcRaster* r,
                            const double* target);
void orc_diff_free(OrcDiffFrozen* f);
void orc_sh_basis_grad(const double* dir, double (*grad)[3]);

/* The ewa method's splat draw (src/render/splatDraw.ts:98-181): words (n x
 * 4) of one eye, order[0..count) back to front, premultiplied "over" into
 * rgba (4 planes of width x height doubles, row 0 at the top). draw.c */
void orc_draw(const uint32_t* words, uint32_t n, const uint32_t* order, uint32_t count, int width, int height,
              double max_std_dev, double min_alpha, double falloff, int linear, double* rgba);
/* The draw of the 3dgut (1), htgs (2) and raygs (3) words of one eye
 * (splatDraw.ts:98-181, projections.tsl.ts:310-330, 432-466,
 * projection.tsl.ts:396-442), FP64; ch2 NULL with one eye. */
void orc_draw_ray(int method, const uint32_t* ch0, const uint32_t* ch1, const uint32_t* ch2, uint32_t n, int eye,
                  const uint32_t* order, uint32_t count, const double* from_head, const double* from_head_quat,
                  const double* proj, int ortho, int width, int height, double max_std_dev, double min_alpha,
                  double falloff, int linear, double* rgba);

/* ---- The 3dgut and htgs projection words (ray.c) ------------------------- */
/* SplatProjection.project / write for "3dgut" (method 1), "htgs" (2) and "raygs" (3)
 * (SplatProjection.ts:130-151, projections.tsl.ts:170-300, with
 * projection.tsl.ts:42-360): an object-space splat (scales and rotation,
 * SH colour applied) through the view's object-to-head transform, which
 * must be a rotation times a uniform scale (splatTransform.tsl.ts:142-152). */
typedef struct orc_ray_splat {
  double center[3], scales[3], quat[4], rgb[3], opacity;
} orc_ray_splat;
typedef struct orc_ray_out {
  uint32_t words[10];    /* 6 head words, then 2 per eye (raygs: 4, then 3 per eye) */
  int visible, invisible, guarded, eye_visible[2];
  uint32_t key;
  double distance;       /* |head-space centre| */
  double head_scales[3]; /* MIP-inflated */
  double head_opacity;   /* MIP-compensated */
  double bound[2][5];    /* per eye after the screen clip: ndc x, y, angle, scale1, scale2 */
  double frame_axes[2][2][3]; /* raygs: per eye the quad's axes (view space) */
  double frame_c2[2];         /* raygs: c^2 per eye */
  double frame_lengths[2][2]; /* raygs: the axes' lengths */
  double eye_margin[2];  /* relative distance to an eye's cull or clip decisions */
  double margin;         /* the smallest over the head and eyes */
} orc_ray_out;
void orc_project_ray(int method, const orc_ray_splat* s, const orc_view* v, const orc_params* p, orc_ray_out* out);
/* fromCov (splats.tsl.ts:189-198 with symEigens, math.tsl.ts:46-113, in
 * double): scales and rotation (x, y, z, w; w >= 0) of a covariance
 * xx, yy, zz, xy, xz, yz */
void orc_from_cov(const double* cov6, double* scales, double* quat);
/* A whole f32 table: words n x (6 + 2 eyes; raygs 4 + 3 eyes), keys, margins (each may be NULL but words) */
void orc_project_ray_table_f32(int method, const orc_table_f32* t, const orc_view* v, const orc_params* p, uint32_t* words,
                               uint32_t* keys, double* margin, double* eye_margin);
void orc_encode_center(const double* c, uint32_t* dir, uint32_t* dist, uint32_t* low); /* projection.tsl.ts:317-333 */
void orc_decode_center(uint32_t dir, uint32_t dist, uint32_t last, double* c);        /* projection.tsl.ts:335-349 */
uint32_t orc_encode_quat_oct32_gpu(const double* q);                                   /* encoding.tsl.ts:97-110 */

#ifdef __cplusplus
}
#endif
#endif

/*
 * Oracle: the ewa projection, SH colour and sort key of one splat, restated
 * in double from the reference's TSL (src/render/projection.tsl.ts,
 * projections.tsl.ts, SplatProjection.ts, splatTransform.tsl.ts,
 * src/gpu/splatGenerator.ts, sh.tsl.ts). Test infrastructure only.
 *
 * Matrices are column-major 4x4 (three.js Matrix4.elements): element (row
 * r, column c) at m[4c + r].
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define AT(m, r, c) ((m)[4 * (c) + (r)])
#define NDC_RANGE 2.0
#define LOG_SCALE_MIN (-4.0)
#define LOG_SCALE_MAX 12.0

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

void orc_params_default(orc_params* p) {
  /* src/render/splatParams.ts:68-72 */
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
  p->sort_bits = 24; /* SplatRenderer.ts:163 */
  p->cull_keys = 1;  /* SplatRenderer.ts:432 */
  p->key_per_eye = 0;
}

static void mat4_mul(const double* a, const double* b, double* out) {
  double r[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0;
      for (int k = 0; k < 4; k++) s += AT(a, i, k) * AT(b, k, j);
      AT(r, i, j) = s;
    }
  memcpy(out, r, sizeof r);
}

/* General 4x4 inverse (cofactors), column-major in and out */
static void mat4_invert(const double* m, double* out) {
  double inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] +
           m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] -
           m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] +
           m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] -
            m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] -
           m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] +
           m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] -
           m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] +
            m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] +
           m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] -
           m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] +
            m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] -
            m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] -
           m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] +
           m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] -
            m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] +
            m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  for (int i = 0; i < 16; i++) out[i] = inv[i] / det;
}

static void xform_point(const double* m, const double* p, double* out) {
  double r[4];
  for (int i = 0; i < 4; i++) r[i] = AT(m, i, 0) * p[0] + AT(m, i, 1) * p[1] + AT(m, i, 2) * p[2] + AT(m, i, 3);
  memcpy(out, r, 3 * sizeof(double));
}

static void clip_of(const double* P, const double* c, double* clip) {
  for (int i = 0; i < 4; i++) clip[i] = AT(P, i, 0) * c[0] + AT(P, i, 1) * c[1] + AT(P, i, 2) * c[2] + AT(P, i, 3);
}

/* M Sigma M^T for the upper-left 3x3 of column-major m; cov6 in and out
 * (xx, yy, zz, xy, xz, yz) (splatTransform.tsl.ts:120-125, projection.tsl.ts:158-165) */
static void cov_xform(const double* m, const double* c, double* out) {
  double S[3][3] = {{c[0], c[3], c[4]}, {c[3], c[1], c[5]}, {c[4], c[5], c[2]}};
  double T[3][3], R[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = AT(m, i, 0) * S[0][j] + AT(m, i, 1) * S[1][j] + AT(m, i, 2) * S[2][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = T[i][0] * AT(m, j, 0) + T[i][1] * AT(m, j, 1) + T[i][2] * AT(m, j, 2);
  out[0] = R[0][0];
  out[1] = R[1][1];
  out[2] = R[2][2];
  out[3] = R[0][1];
  out[4] = R[0][2];
  out[5] = R[1][2];
}

void orc_view_from_camera(const double* object_to_world, const double* head, int eyes, const double* const* eye_views,
                          const double* const* eye_projections, const int* ortho, double width, double height,
                          double ndc_z_min, orc_view* out) {
  memset(out, 0, sizeof *out);
  /* setTransforms with no modifier: a = head * object (splatTransform.tsl.ts:83-85) */
  mat4_mul(head, object_to_world, out->object_to_head);
  /* the generator context: viewOrigin = position of inverse(head * object)
   * (splatGenerator.ts:96-103, 211-212) */
  double inv[16];
  mat4_invert(out->object_to_head, inv);
  out->view_origin[0] = inv[12];
  out->view_origin[1] = inv[13];
  out->view_origin[2] = inv[14];
  out->eyes = eyes;
  double head_world[16];
  mat4_invert(head, head_world);
  for (int e = 0; e < eyes; e++) {
    /* fromHead: exactly the identity when the eye is the head (splatCamera.ts:128-129) */
    if (memcmp(eye_views[e], head, 16 * sizeof(double)) == 0) {
      memset(out->eye[e].from_head, 0, 16 * sizeof(double));
      for (int i = 0; i < 4; i++) AT(out->eye[e].from_head, i, i) = 1.0;
    } else {
      mat4_mul(eye_views[e], head_world, out->eye[e].from_head);
    }
    memcpy(out->eye[e].projection, eye_projections[e], 16 * sizeof(double));
    out->eye[e].ortho = ortho ? ortho[e] : 0;
  }
  out->size[0] = width;
  out->size[1] = height;
  out->ndc_z_min = ndc_z_min;
}

/* --- word codecs (projection.tsl.ts:277-310) --- */
static uint32_t encode_fixed16(double v) {
  return (uint32_t)clampd(floor((v + NDC_RANGE) / (2 * NDC_RANGE) * 65535.0 + 0.5), 0.0, 65535.0);
}
static uint32_t float_bits(double v) {
  float f = (float)v;
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static uint32_t encode_depth20(double view_z) {
  const uint32_t lo = 113u << 23, hi = (145u << 23) - 1u;
  uint32_t f = float_bits(-view_z);
  f = f < lo ? lo : (f > hi ? hi : f);
  uint32_t code = (f - lo + 128u) >> 8;
  return code < 0xfffffu ? code : 0xfffffu;
}
static uint32_t encode_angle12(double angle) { return (uint32_t)floor(angle / M_PI * 4096.0 + 0.5) & 0xfffu; }
static uint32_t encode_log_scale12(double scale) {
  const double top = 4094.0;
  double n = (clampd(log2(fmax(scale, 1e-20)), LOG_SCALE_MIN, LOG_SCALE_MAX) - LOG_SCALE_MIN) /
             (LOG_SCALE_MAX - LOG_SCALE_MIN) * top;
  return (uint32_t)clampd(floor(n + 0.5), 0.0, top) + 1u;
}
static uint32_t encode_opacity8(double o) {
  uint32_t c = (uint32_t)floor(clampd(o * 128.0, 0.0, 255.0) + 0.5);
  return c > 1u ? c : 1u;
}
static double stored_opacity(double o) { return o > 1.0 ? (o + 3.0) * 0.25 : o; }

static double rel_margin(double x, double threshold) {
  double s = fmax(fabs(threshold), fabs(x));
  return s > 0 ? fabs(x - threshold) / s : INFINITY;
}

void orc_project_ewa(const orc_splat* s, const orc_view* v, const orc_params* p, orc_out* out) {
  memset(out, 0, sizeof *out);
  double margin = INFINITY;

  /* transformCov with a = object to head (splatTransform.tsl.ts:120-140) */
  double center[3], cov[6];
  xform_point(v->object_to_head, s->center, center);
  cov_xform(v->object_to_head, s->cov, cov);

  /* prepareSplat (projection.tsl.ts:42-64) */
  double opacity = s->opacity, inflate = 1.0;
  if (opacity > 1.0) {
    opacity = fmin(opacity * 4.0 - 3.0, 5.0);
    if (p->lod_inflate) {
      inflate = pow(exp((opacity * opacity - 1.0) / M_E), 1.0 / 3.0);
      opacity = 1.0;
    }
  }
  for (int k = 0; k < 6; k++) cov[k] *= inflate * inflate;
  double std_dev = opacity > 1.0 ? p->max_std_dev + (opacity - 1.0) * 0.7 : p->max_std_dev;

  /* cullInvisible (projection.tsl.ts:66-69) */
  int degenerate = cov[0] == 0 && cov[1] == 0 && cov[2] == 0;
  out->invisible = opacity <= 0.0 || opacity < p->min_alpha || degenerate;
  margin = fmin(margin, rel_margin(opacity, p->min_alpha));

  /* the head view: the head with the first eye's lens (projection.tsl.ts:91-94, 112-130) */
  double hclip[4];
  clip_of(v->eye[0].projection, center, hclip);
  out->distance = sqrt(center[0] * center[0] + center[1] * center[1] + center[2] * center[2]);
  for (int c = 0; c < 3; c++) out->rgb[c] = s->rgb[c];

  int visible = 0;
  for (int e = 0; e < v->eyes; e++) {
    const orc_eye* eye = &v->eye[e];
    orc_eye_out* eo = &out->eye[e];
    /* splatInView with the eye frame (projection.tsl.ts:112-130) */
    double c[3], clip[4];
    xform_point(eye->from_head, center, c);
    clip_of(eye->projection, c, clip);
    eo->distance = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double blur = p->blur_amount;
    if (p->focal_distance > 0 && p->aperture_tan > 0 && c[2] < 0) {
      double depth = -c[2];
      double coc_view = fabs(depth - p->focal_distance) * p->aperture_tan;
      double focal_pixels = v->size[1] * 0.5 * AT(eye->projection, 1, 1);
      double coc = eye->ortho ? coc_view * focal_pixels : coc_view * focal_pixels / depth;
      blur = clampd(coc * coc, blur, p->max_pixel_radius * p->max_pixel_radius);
    }
    double ndc[2] = {0, 0};
    if (clip[3] > 0) {
      ndc[0] = clip[0] / clip[3];
      ndc[1] = clip[1] / clip[3];
    }
    eo->ndc[0] = ndc[0];
    eo->ndc[1] = ndc[1];
    eo->view_z = c[2];
    double em = INFINITY;

    /* cullDepth, cullClipXY (projection.tsl.ts:132-140) */
    int cull_depth = c[2] >= 0.0 || clip[2] < v->ndc_z_min * clip[3] || clip[2] >= clip[3];
    double limit = fmin(p->clip_xy, NDC_RANGE) * clip[3];
    int cull_xy = fabs(clip[0]) > limit || fabs(clip[1]) > limit;
    em = fmin(em, fabs(c[2]) / fmax(eo->distance, 1e-300));
    em = fmin(em, rel_margin(clip[2], v->ndc_z_min * clip[3]));
    em = fmin(em, rel_margin(clip[2], clip[3]));
    em = fmin(em, rel_margin(fabs(clip[0]), limit));
    em = fmin(em, rel_margin(fabs(clip[1]), limit));
    int ok = !(out->invisible || cull_depth || cull_xy);
    if (ok) {
      /* ewa.project (projections.tsl.ts:95-148) */
      double cov3[6];
      cov_xform(eye->from_head, cov, cov3);
      double fx = v->size[0] * 0.5 * AT(eye->projection, 0, 0);
      double fy = v->size[1] * 0.5 * AT(eye->projection, 1, 1);
      double J[2][3];
      if (eye->ortho) {
        double Jo[2][3] = {{fx, 0, 0}, {0, fy, 0}};
        memcpy(J, Jo, sizeof J);
      } else {
        double iz = 1.0 / c[2];
        double j1x = fx * iz, j1y = fy * iz;
        double Jp[2][3] = {{j1x, 0, -j1x * c[0] * iz}, {0, j1y, -j1y * c[1] * iz}};
        memcpy(J, Jp, sizeof J);
      }
      double S[3][3] = {{cov3[0], cov3[3], cov3[4]}, {cov3[3], cov3[1], cov3[5]}, {cov3[4], cov3[5], cov3[2]}};
      double JS[2][3];
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 3; j++) JS[i][j] = J[i][0] * S[0][j] + J[i][1] * S[1][j] + J[i][2] * S[2][j];
      double a0 = JS[0][0] * J[0][0] + JS[0][1] * J[0][1] + JS[0][2] * J[0][2];
      double b0 = JS[0][0] * J[1][0] + JS[0][1] * J[1][1] + JS[0][2] * J[1][2];
      double d0 = JS[1][0] * J[1][0] + JS[1][1] * J[1][1] + JS[1][2] * J[1][2];
      double a = a0 + p->pre_blur_amount, d = d0 + p->pre_blur_amount, b = b0;
      double det_orig = a * d - b * b;
      a += blur;
      d += blur;
      double det = a * d - b * b;
      double op = opacity * sqrt(fmax(0.0, det_orig / det));
      double contribution = op * 2 * M_PI * sqrt(fmax(0.0, det));
      /* boundFromCov (projection.tsl.ts:223-240) */
      double half = (a - d) * 0.5;
      double eig1 = (a + d) * 0.5 + sqrt(half * half + b * b);
      double eig2 = eig1 > 0.0 ? fmax(det / eig1, 0.0) : 0.0;
      double s1 = fmin(p->max_pixel_radius, std_dev * sqrt(fmax(eig1, 0.0)));
      double s2 = fmin(p->max_pixel_radius, std_dev * sqrt(eig2));
      int bound_ok = s1 >= p->min_pixel_radius || s2 >= p->min_pixel_radius;
      double angle = 0;
      if (bound_ok) {
        /* eigenvector2 (projection.tsl.ts:215-220) */
        double vx, vy;
        if (fabs(a - eig1) >= fabs(d - eig1)) {
          vx = b;
          vy = eig1 - a;
        } else {
          vx = eig1 - d;
          vy = b;
        }
        double len = sqrt(vx * vx + vy * vy);
        if (len > 0) {
          vx /= len;
          vy /= len;
        } else {
          vx = 1;
          vy = 0;
        }
        angle = atan2(vy, vx);
        if (angle < 0) angle += M_PI;
      }
      em = fmin(em, rel_margin(op, p->min_alpha));
      if (p->min_contribution > 0) em = fmin(em, rel_margin(contribution, p->min_contribution));
      ok = op >= p->min_alpha && contribution >= p->min_contribution && bound_ok;
      eo->a = a;
      eo->b = b;
      eo->d = d;
      eo->conic[0] = d / det;
      eo->conic[1] = -b / det;
      eo->conic[2] = a / det;
      eo->opacity = op;
      eo->scale1 = s1;
      eo->scale2 = s2;
      eo->angle = angle;
      eo->lambda1 = eig1;
      eo->lambda2 = eig2;
      if (ok) {
        eo->words[0] = encode_fixed16(ndc[0]) | (encode_fixed16(ndc[1]) << 16);
        eo->words[1] = encode_depth20(c[2]) | (encode_angle12(angle) << 20);
        eo->words[2] = orc_encode_ext_rgb(out->rgb[0], out->rgb[1], out->rgb[2]);
        eo->words[3] = encode_opacity8(stored_opacity(op)) | (encode_log_scale12(s1) << 8) |
                       (encode_log_scale12(s2) << 20);
        eo->visible = 1;
        visible = 1;
      }
    }
    eo->margin = em;
    eo->key = orc_key_compute(eo->distance, p->sort_bits);
  }
  out->visible = visible;

  /* inSortGuard (projection.tsl.ts:144-155) */
  int guarded = 1;
  if (p->sort_guard > 0) {
    double extent = sqrt(fmax(fmax(cov[0], cov[1]), cov[2]));
    double radius = extent * p->max_std_dev;
    double w = fmax(hclip[3], 1e-6);
    double rx = AT(v->eye[0].projection, 0, 0) * radius / w, ry = AT(v->eye[0].projection, 1, 1) * radius / w;
    int inside = fabs(hclip[0] / w) - rx <= p->sort_guard + 1.0 && fabs(hclip[1] / w) - ry <= p->sort_guard + 1.0;
    int near = -center[2] <= radius;
    guarded = center[2] < radius && (inside || near);
  }
  out->guarded = guarded;

  /* SplatProjection.write's key (SplatProjection.ts:143-151) */
  int keyed = p->cull_keys ? visible : (!out->invisible && guarded);
  if (p->cull_keys)
    for (int e = 0; e < v->eyes; e++) margin = fmin(margin, out->eye[e].margin);
  out->margin = margin;
  out->key = keyed ? orc_key_compute(out->distance, p->sort_bits) : orc_key_invalid(p->sort_bits);
  if (p->key_per_eye)
    for (int e = 0; e < v->eyes; e++)
      if (!(p->cull_keys ? out->eye[e].visible : keyed)) out->eye[e].key = orc_key_invalid(p->sort_bits);
}

/* --- whole tables --- */

static void emit(const orc_out* o, const orc_view* v, const orc_params* p, uint32_t n, uint32_t i, orc_batch_out* b) {
  for (int e = 0; e < v->eyes; e++) {
    const orc_eye_out* eo = &o->eye[e];
    size_t at = (size_t)e * n + i;
    if (b->words)
      for (int k = 0; k < 4; k++) b->words[4 * at + k] = eo->words[k];
    if (b->conic)
      for (int k = 0; k < 3; k++) b->conic[(size_t)e * 3 * n + (size_t)k * n + i] = eo->conic[k];
    if (b->opacity) b->opacity[at] = eo->opacity;
    if (b->ndc)
      for (int k = 0; k < 2; k++) b->ndc[(size_t)e * 2 * n + (size_t)k * n + i] = eo->ndc[k];
    if (b->view_z) b->view_z[at] = eo->view_z;
    if (b->visible) b->visible[at] = (uint8_t)eo->visible;
    if (p->key_per_eye) {
      if (b->keys) b->keys[at] = eo->key;
      if (b->distance) b->distance[at] = eo->distance;
    }
  }
  if (!p->key_per_eye) {
    if (b->keys) b->keys[i] = o->key;
    if (b->distance) b->distance[i] = o->distance;
  }
  if (b->rgb)
    for (int k = 0; k < 3; k++) b->rgb[(size_t)k * n + i] = o->rgb[k];
  if (b->margin) b->margin[i] = o->margin;
}

/* SH colour as the generator applies it before the transform
 * (splatGenerator.ts:152-158; GpuShReader.apply, sh.tsl.ts:261-266):
 * rgb = max(rgb + SH(normalize(center - viewOrigin)), 0), bands capped at
 * min(maxShDegree, the table's degree). */
static void apply_sh(const double* center, const double* coeffs, int degree, const orc_view* v, const orc_params* p,
                     double* rgb) {
  int deg = degree < p->max_sh_degree ? degree : p->max_sh_degree;
  if (deg <= 0) return;
  double dir[3] = {center[0] - v->view_origin[0], center[1] - v->view_origin[1], center[2] - v->view_origin[2]};
  double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  for (int k = 0; k < 3; k++) dir[k] /= len;
  double sh[3];
  orc_sh_evaluate(coeffs, dir, deg, sh);
  for (int k = 0; k < 3; k++) rgb[k] = fmax(rgb[k] + sh[k], 0.0);
}

static const int SH_COEFFS[4] = {0, 3, 8, 15};

void orc_project_table_f32(const orc_table_f32* t, const orc_view* v, const orc_params* p, orc_batch_out* out) {
  const uint32_t n = t->count;
  const size_t st = t->stride ? t->stride : n;
  for (uint32_t i = 0; i < n; i++) {
    orc_splat s;
    double scales[3], quat[4], coeffs[45];
    for (int k = 0; k < 3; k++) {
      s.center[k] = t->center[k * st + i];
      scales[k] = exp((double)t->log_scale[k * st + i]);
      s.rgb[k] = t->rgb[k * st + i];
    }
    for (int k = 0; k < 4; k++) quat[k] = t->quat[k * st + i];
    s.opacity = t->opacity[i];
    orc_cov_from_scale_quat(scales, quat, s.cov);
    if (t->sh_degree > 0 && t->sh) {
      for (int k = 0; k < 3 * SH_COEFFS[t->sh_degree]; k++) coeffs[k] = t->sh[k * st + i];
      apply_sh(s.center, coeffs, t->sh_degree, v, p, s.rgb);
    }
    orc_out o;
    orc_project_ewa(&s, v, p, &o);
    emit(&o, v, p, n, i, out);
  }
}

void orc_project_table_packed(uint32_t count, const uint32_t* packed, const double* enc, int sh_codec, int sh_degree,
                              const uint32_t* const* sh_channels, const orc_view* v, const orc_params* p,
                              orc_batch_out* out) {
  for (uint32_t i = 0; i < count; i++) {
    orc_splat s;
    double scales[3], quat[4];
    orc_packed_decode(&packed[4 * (size_t)i], enc, s.center, scales, quat, s.rgb, &s.opacity);
    orc_cov_from_scale_quat(scales, quat, s.cov);
    if (sh_codec && sh_degree > 0) {
      uint32_t words[16];
      int texels = (orc_sh_words(sh_codec, sh_degree) + 3) / 4;
      for (int t = 0; t < texels; t++)
        for (int k = 0; k < 4; k++) words[4 * t + k] = sh_channels[t][4 * (size_t)i + k];
      double coeffs[45];
      orc_sh_decode(sh_codec, words, sh_degree, coeffs);
      apply_sh(s.center, coeffs, sh_degree, v, p, s.rgb);
    }
    orc_out o;
    orc_project_ewa(&s, v, p, &o);
    emit(&o, v, p, count, i, out);
  }
}

void orc_project_table_ext(uint32_t count, const uint32_t* ext_a, const uint32_t* ext_b, int sh_codec, int sh_degree,
                           const uint32_t* const* sh_channels, const orc_view* v, const orc_params* p,
                           orc_batch_out* out) {
  for (uint32_t i = 0; i < count; i++) {
    orc_splat s;
    double scales[3], quat[4];
    orc_ext_decode(&ext_a[4 * (size_t)i], &ext_b[4 * (size_t)i], s.center, scales, quat, s.rgb, &s.opacity);
    orc_cov_from_scale_quat(scales, quat, s.cov);
    if (sh_codec && sh_degree > 0) {
      uint32_t words[16];
      int texels = (orc_sh_words(sh_codec, sh_degree) + 3) / 4;
      for (int t = 0; t < texels; t++)
        for (int k = 0; k < 4; k++) words[4 * t + k] = sh_channels[t][4 * (size_t)i + k];
      double coeffs[45];
      orc_sh_decode(sh_codec, words, sh_degree, coeffs);
      apply_sh(s.center, coeffs, sh_degree, v, p, s.rgb);
    }
    orc_out o;
    orc_project_ewa(&s, v, p, &o);
    emit(&o, v, p, count, i, out);
  }
}

void orc_project_table_cov(uint32_t count, const uint32_t* cov_a, const uint32_t* cov_b, int sh_codec, int sh_degree,
                           const uint32_t* const* sh_channels, const orc_view* v, const orc_params* p,
                           orc_batch_out* out) {
  for (uint32_t i = 0; i < count; i++) {
    orc_splat s;
    orc_cov_decode(&cov_a[4 * (size_t)i], &cov_b[4 * (size_t)i], s.center, s.cov, s.rgb, &s.opacity);
    if (sh_codec && sh_degree > 0) {
      uint32_t words[16];
      int texels = (orc_sh_words(sh_codec, sh_degree) + 3) / 4;
      for (int t = 0; t < texels; t++)
        for (int k = 0; k < 4; k++) words[4 * t + k] = sh_channels[t][4 * (size_t)i + k];
      double coeffs[45];
      orc_sh_decode(sh_codec, words, sh_degree, coeffs);
      apply_sh(s.center, coeffs, sh_degree, v, p, s.rgb);
    }
    orc_out o;
    orc_project_ewa(&s, v, p, &o);
    emit(&o, v, p, count, i, out);
  }
}

/* --- The keys-only pass (SplatProjection.writeKeys, SplatProjection.ts:165-172) ---
 * The "keys" program (SplatProjector.ts:152-194) has no wanted form, so a
 * stored-splat generator stays in the splat form: transformSplat
 * (splatTransform.tsl.ts:154-165) keeps the scales, times cbrt(|det M|), for
 * a rotation times a uniform scale (isUniformScale, :46-62: G = M^T M, every
 * entry of G - s2 I within 1e-6 s2, s2 = trace(G) / 3, and det M > 0) and
 * otherwise takes fromCov of M Sigma M^T (splats.tsl.ts:189-198), whose
 * scales are the square roots of its eigenvalues. inSortGuard's extent is
 * then the largest head-space scale (projection.tsl.ts:145). */
static double det3_cm(const double* m) {
  return AT(m, 0, 0) * (AT(m, 1, 1) * AT(m, 2, 2) - AT(m, 1, 2) * AT(m, 2, 1)) -
         AT(m, 0, 1) * (AT(m, 1, 0) * AT(m, 2, 2) - AT(m, 1, 2) * AT(m, 2, 0)) +
         AT(m, 0, 2) * (AT(m, 1, 0) * AT(m, 2, 1) - AT(m, 1, 1) * AT(m, 2, 0));
}

static int is_uniform_scale(const double* m) {
  double g[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) g[i][j] = AT(m, 0, i) * AT(m, 0, j) + AT(m, 1, i) * AT(m, 1, j) + AT(m, 2, i) * AT(m, 2, j);
  double s2 = (g[0][0] + g[1][1] + g[2][2]) / 3.0;
  if (!(s2 > 0)) return 0;
  for (int i = 0; i < 3; i++)
    for (int j = i; j < 3; j++)
      if (fabs(g[i][j] - (i == j ? s2 : 0.0)) > 1e-6 * s2) return 0;
  return det3_cm(m) > 0;
}

/* Largest eigenvalue of a symmetric 3x3 (xx, yy, zz, xy, xz, yz) by cyclic Jacobi rotations */
static double sym3_max_eigen(const double* c) {
  double a[3][3] = {{c[0], c[3], c[4]}, {c[3], c[1], c[5]}, {c[4], c[5], c[2]}};
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    double scale = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-30 * scale || off == 0.0) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < 3; k++) {
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = cs * akp - sn * akq;
          a[k][q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < 3; k++) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = cs * apk - sn * aqk;
          a[q][k] = sn * apk + cs * aqk;
        }
      }
  }
  return fmax(a[0][0], fmax(a[1][1], a[2][2]));
}

void orc_project_keys_table_f32(const orc_table_f32* t, const orc_view* v, const orc_params* p, uint32_t* keys,
                                double* margin) {
  const uint32_t n = t->count;
  const size_t st = t->stride ? t->stride : n;
  const double* M = v->object_to_head;
  const int uniform = is_uniform_scale(M);
  const double uscale = cbrt(fabs(det3_cm(M)));
  for (uint32_t i = 0; i < n; i++) {
    double c0[3], center[3], scales[3], quat[4], cov[6];
    for (int k = 0; k < 3; k++) {
      c0[k] = t->center[k * st + i];
      scales[k] = exp((double)t->log_scale[k * st + i]);
    }
    for (int k = 0; k < 4; k++) quat[k] = t->quat[k * st + i];
    xform_point(M, c0, center);
    /* prepareSplat (projection.tsl.ts:42-64), splat form: the scales grow by inflate */
    double opacity = t->opacity[i], inflate = 1.0;
    if (opacity > 1.0) {
      opacity = fmin(opacity * 4.0 - 3.0, 5.0);
      if (p->lod_inflate) {
        inflate = pow(exp((opacity * opacity - 1.0) / M_E), 1.0 / 3.0);
        opacity = 1.0;
      }
    }
    double extent;
    if (uniform) {
      extent = fmax(scales[0], fmax(scales[1], scales[2])) * uscale;
    } else {
      double c6[6];
      orc_cov_from_scale_quat(scales, quat, c6);
      cov_xform(M, c6, cov);
      extent = sqrt(fmax(sym3_max_eigen(cov), 0.0));
    }
    extent *= inflate;
    int degenerate = scales[0] == 0 && scales[1] == 0 && scales[2] == 0;
    int invisible = opacity <= 0.0 || opacity < p->min_alpha || degenerate;
    double m = rel_margin(opacity, p->min_alpha);
    int guarded = 1;
    if (p->sort_guard > 0) {
      /* inSortGuard (projection.tsl.ts:144-155) with the head frame: eye 0's lens */
      const double* P = v->eye[0].projection;
      double hclip[4];
      clip_of(P, center, hclip);
      double radius = extent * p->max_std_dev;
      double w = fmax(hclip[3], 1e-6);
      double rx = AT(P, 0, 0) * radius / w, ry = AT(P, 1, 1) * radius / w;
      int inside = fabs(hclip[0] / w) - rx <= p->sort_guard + 1.0 && fabs(hclip[1] / w) - ry <= p->sort_guard + 1.0;
      int near = -center[2] <= radius;
      guarded = center[2] < radius && (inside || near);
      m = fmin(m, rel_margin(fabs(hclip[0] / w) - rx, p->sort_guard + 1.0));
      m = fmin(m, rel_margin(fabs(hclip[1] / w) - ry, p->sort_guard + 1.0));
      m = fmin(m, rel_margin(-center[2], radius));
      m = fmin(m, rel_margin(center[2], radius));
    }
    double dist = sqrt(center[0] * center[0] + center[1] * center[1] + center[2] * center[2]);
    keys[i] = !invisible && guarded ? orc_key_compute(dist, p->sort_bits) : orc_key_invalid(p->sort_bits);
    if (margin) margin[i] = m;
  }
}

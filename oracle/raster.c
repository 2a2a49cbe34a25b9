/*
 * Oracle (FP64): the differentiable compositor and its adjoint.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). No reference counterpart as a
 * whole: the reference draws splats with rasterised quads (splatDraw.ts) and
 * has no backward pass. What it fixes, and this follows:
 *   - the ewa projection of each splat (projections.tsl.ts:95-148) with its
 *     opacity compensation, the culls (projection.tsl.ts:42-140) and the SH
 *     colour clamped at 0 (splatDraw.ts:149-151, sh.ts:362-407);
 *   - the fragment: alpha = opacity * exp(-z^2 / 2) at falloff 1
 *     (projection.tsl.ts:444-455), z^2 the squared Mahalanobis distance of
 *     the pixel from the splat centre under the blurred 2D covariance
 *     (projections.tsl.ts:155-167), discarded where z^2 > cutoff^2 and where
 *     alpha < minAlpha, cutoff = min(maxStdDev, sqrt(2 ln(opacity / minAlpha)))
 *     (splatDraw.ts:136-145, 165-174);
 *   - premultiplied "over" in the stored (sRGB) colour (splatDraw.ts:62-78),
 *     in the sort's order: keys ascending, back to front (sort.ts), which is
 *     front to back over the reversed list.
 * Added for training: compositing stops before the splat that would bring
 * the transmittance below t_stop (1e-4); the image is the premultiplied
 * colour; the loss is the mean squared error against a target image.
 *
 * Parameters per splat (59 doubles, the f32 table's plane order): centre 3,
 * log-scale 3, quaternion 4 (x, y, z, w; normalised inside), opacity LOGIT
 * 1 (opacity = sigmoid), base colour 3, SH bands 1-3 coefficient-major 45.
 * Pixels: (x + 0.5, y + 0.5) with y up (row 0 at the bottom), image planar
 * [channel][y][x].
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NP ORC_DIFF_PARAMS

typedef struct {
  int ok;
  double t[3], dist;
  double clip[4], mu2[2];
  double qv[4], qlen, R[9], s[3], M[9], Sigma[9];
  double J[6], T[6];
  double a, b, d, ap, dp, det_o, det_p, ratio;
  double op, opc, cut;
  double Q[3];
  double dvec[3], dlen, dir[3], basis[24], col_raw[3], col[3];
} DSplat;

static void view_parts(const OrcRaster* r, double Vl[9], double Vt[3], double cam[3]) {
  /* column-major world -> view */
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) Vl[3 * i + j] = r->view[4 * j + i];
    Vt[i] = r->view[12 + i];
  }
  /* camera position: -Vl^T Vt (Vl a rotation times a positive scale is not
     assumed: solve Vl c = -Vt) */
  double A[9];
  memcpy(A, Vl, sizeof A);
  double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
  double inv[9] = {(A[4] * A[8] - A[5] * A[7]) / det, (A[2] * A[7] - A[1] * A[8]) / det, (A[1] * A[5] - A[2] * A[4]) / det,
                   (A[5] * A[6] - A[3] * A[8]) / det, (A[0] * A[8] - A[2] * A[6]) / det, (A[2] * A[3] - A[0] * A[5]) / det,
                   (A[3] * A[7] - A[4] * A[6]) / det, (A[1] * A[6] - A[0] * A[7]) / det, (A[0] * A[4] - A[1] * A[3]) / det};
  for (int i = 0; i < 3; i++) cam[i] = -(inv[3 * i] * Vt[0] + inv[3 * i + 1] * Vt[1] + inv[3 * i + 2] * Vt[2]);
}

/* The forward of one splat: everything the pixels and the adjoint need. */
static void splat_forward(const double* th, const OrcRaster* r, const double Vl[9], const double Vt[3],
                          const double cam[3], DSplat* s) {
  memset(s, 0, sizeof *s);
  const double* mu = th;
  for (int i = 0; i < 3; i++) s->t[i] = Vl[3 * i] * mu[0] + Vl[3 * i + 1] * mu[1] + Vl[3 * i + 2] * mu[2] + Vt[i];
  s->dist = sqrt(s->t[0] * s->t[0] + s->t[1] * s->t[1] + s->t[2] * s->t[2]);
  const double* P = r->proj; /* column-major */
  for (int i = 0; i < 4; i++) s->clip[i] = P[i] * s->t[0] + P[4 + i] * s->t[1] + P[8 + i] * s->t[2] + P[12 + i];
  const double x = s->t[0], y = s->t[1], z = s->t[2], w = s->clip[3];
  /* cullDepth, cullClipXY (projection.tsl.ts:132-140) */
  const int cull_depth = z >= 0.0 || s->clip[2] < r->ndc_z_min * w || s->clip[2] >= w;
  const double limit = fmin(r->clip_xy, 2.0) * w;
  const int cull_xy = fabs(s->clip[0]) > limit || fabs(s->clip[1]) > limit;
  s->mu2[0] = (s->clip[0] / w + 1.0) * 0.5 * r->width;
  s->mu2[1] = (s->clip[1] / w + 1.0) * 0.5 * r->height;
  /* covariance */
  double len = sqrt(th[6] * th[6] + th[7] * th[7] + th[8] * th[8] + th[9] * th[9]);
  s->qlen = len;
  for (int k = 0; k < 4; k++) s->qv[k] = len > 0 ? th[6 + k] / len : (k == 3 ? 1.0 : 0.0);
  const double qx = s->qv[0], qy = s->qv[1], qz = s->qv[2], qw = s->qv[3];
  double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy),
                 2 * (qx * qy + qw * qz), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                 2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), 1 - 2 * (qx * qx + qy * qy)};
  memcpy(s->R, R, sizeof R);
  for (int k = 0; k < 3; k++) s->s[k] = exp(th[3 + k]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) s->M[3 * i + j] = R[3 * i + j] * s->s[j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      s->Sigma[3 * i + j] = s->M[3 * i] * s->M[3 * j] + s->M[3 * i + 1] * s->M[3 * j + 1] + s->M[3 * i + 2] * s->M[3 * j + 2];
  /* J (projections.tsl.ts:104-117): focal = renderSize / 2 * (P00, P11) */
  const double fx = 0.5 * r->width * P[0], fy = 0.5 * r->height * P[5];
  s->J[0] = fx / z;
  s->J[1] = 0.0;
  s->J[2] = -fx * x / (z * z);
  s->J[3] = 0.0;
  s->J[4] = fy / z;
  s->J[5] = -fy * y / (z * z);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++)
      s->T[3 * i + j] = s->J[3 * i] * Vl[j] + s->J[3 * i + 1] * Vl[3 + j] + s->J[3 * i + 2] * Vl[6 + j];
  double TS[6];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++)
      TS[3 * i + j] = s->T[3 * i] * s->Sigma[j] + s->T[3 * i + 1] * s->Sigma[3 + j] + s->T[3 * i + 2] * s->Sigma[6 + j];
  s->a = TS[0] * s->T[0] + TS[1] * s->T[1] + TS[2] * s->T[2];
  s->b = TS[0] * s->T[3] + TS[1] * s->T[4] + TS[2] * s->T[5];
  s->d = TS[3] * s->T[3] + TS[4] * s->T[4] + TS[5] * s->T[5];
  s->ap = s->a + r->blur;
  s->dp = s->d + r->blur;
  s->det_o = s->a * s->d - s->b * s->b;
  s->det_p = s->ap * s->dp - s->b * s->b;
  s->ratio = s->det_o > 0.0 ? s->det_o / s->det_p : 0.0;
  s->op = 1.0 / (1.0 + exp(-th[10]));
  s->opc = s->op * sqrt(s->ratio);
  s->Q[0] = s->dp / s->det_p;
  s->Q[1] = -s->b / s->det_p;
  s->Q[2] = s->ap / s->det_p;
  s->cut = fmin(r->max_std_dev, sqrt(2.0 * log(fmax(s->opc, r->min_alpha) / r->min_alpha)));
  /* colour: base + SH in the direction from the camera, clamped at 0 */
  for (int k = 0; k < 3; k++) s->dvec[k] = mu[k] - cam[k];
  s->dlen = sqrt(s->dvec[0] * s->dvec[0] + s->dvec[1] * s->dvec[1] + s->dvec[2] * s->dvec[2]);
  for (int k = 0; k < 3; k++) s->dir[k] = s->dlen > 0 ? s->dvec[k] / s->dlen : 0.0;
  orc_sh_basis(s->dir, s->basis);
  static const int coeffs[4] = {0, 3, 8, 15};
  const int nk = coeffs[r->sh_degree < 0 ? 0 : (r->sh_degree > 3 ? 3 : r->sh_degree)];
  for (int c = 0; c < 3; c++) {
    double v = th[11 + c];
    for (int k = 0; k < nk; k++) v += th[14 + 3 * k + c] * s->basis[k];
    s->col_raw[c] = v;
    s->col[c] = v > 0.0 ? v : 0.0;
  }
  /* prepare's invisible (op < minAlpha) and the ewa visibility (opc >= minAlpha) */
  s->ok = !cull_depth && !cull_xy && s->op >= r->min_alpha && s->opc >= r->min_alpha && s->det_p > 0.0;
}

typedef struct {
  uint32_t key;
  int index;
} KeyIdx;

static int cmp_key(const void* pa, const void* pb) {
  const KeyIdx* a = (const KeyIdx*)pa;
  const KeyIdx* b = (const KeyIdx*)pb;
  if (a->key != b->key) return a->key < b->key ? -1 : 1;
  return a->index < b->index ? -1 : (a->index > b->index ? 1 : 0);
}

/* Draw order front to back: the drawn splats by key ascending (back to
   front, ties by index), reversed. Returns the count. */
static int draw_order(const DSplat* sp, int n, int* order) {
  KeyIdx* k = (KeyIdx*)malloc(sizeof(KeyIdx) * (size_t)(n > 0 ? n : 1));
  int m = 0;
  for (int i = 0; i < n; i++)
    if (sp[i].ok) {
      k[m].key = orc_key_compute(sp[i].dist, 32);
      k[m].index = i;
      m++;
    }
  qsort(k, (size_t)m, sizeof(KeyIdx), cmp_key);
  for (int j = 0; j < m; j++) order[j] = k[m - 1 - j].index;
  free(k);
  return m;
}

static DSplat* forward_all(const double* theta, int n, const OrcRaster* r, int* order, int* m) {
  double Vl[9], Vt[3], cam[3];
  view_parts(r, Vl, Vt, cam);
  DSplat* sp = (DSplat*)malloc(sizeof(DSplat) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; i++) splat_forward(theta + (size_t)NP * i, r, Vl, Vt, cam, &sp[i]);
  *m = draw_order(sp, n, order);
  return sp;
}

/* Binning (an evaluation shortcut, not a change of the maths): a splat's
   fragments are discarded where z^2 > cut^2, and z^2 = d^T Q d with Q the
   inverse of the blurred 2D covariance S = [[ap, b], [b, dp]], so they lie
   within |dx| <= cut sqrt(ap), |dy| <= cut sqrt(dp) of its centre. Each
   16x16 tile lists, in draw order, the splats whose rectangle (one pixel
   wider) meets it; a pixel walks its tile's list, which holds every splat
   that can reach it in the same order as the full list. */
#define BIN_TILE 16
static int g_binning = 1;

void orc_raster_binning(int on) { g_binning = on != 0; }

typedef struct {
  int tiles_x, tiles_y;
  int* start; /* tiles + 1 */
  int* list;  /* splat indices, per tile in draw order */
} Bins;

static int splat_rect(const DSplat* s, int W, int H, int* x0, int* x1, int* y0, int* y1) {
  const double rx = s->cut * sqrt(fmax(s->ap, 0.0)) + 1.0, ry = s->cut * sqrt(fmax(s->dp, 0.0)) + 1.0;
  if (!(rx < 1e9) || !(ry < 1e9)) {
    *x0 = 0, *x1 = W - 1, *y0 = 0, *y1 = H - 1;
    return 1;
  }
  /* pixel x's centre x + 0.5 within [mu - rx, mu + rx] */
  const double fx0 = ceil(s->mu2[0] - rx - 0.5), fx1 = floor(s->mu2[0] + rx - 0.5);
  const double fy0 = ceil(s->mu2[1] - ry - 0.5), fy1 = floor(s->mu2[1] + ry - 0.5);
  if (fx1 < 0 || fy1 < 0 || fx0 > W - 1 || fy0 > H - 1 || fx0 > fx1 || fy0 > fy1) return 0;
  *x0 = fx0 < 0 ? 0 : (int)fx0;
  *x1 = fx1 > W - 1 ? W - 1 : (int)fx1;
  *y0 = fy0 < 0 ? 0 : (int)fy0;
  *y1 = fy1 > H - 1 ? H - 1 : (int)fy1;
  return 1;
}

static void bin_splats(const DSplat* sp, const int* order, int m, int W, int H, Bins* b) {
  b->tiles_x = (W + BIN_TILE - 1) / BIN_TILE;
  b->tiles_y = (H + BIN_TILE - 1) / BIN_TILE;
  const int tiles = b->tiles_x * b->tiles_y;
  b->start = (int*)calloc((size_t)tiles + 1, sizeof(int));
  for (int pass = 0; pass < 2; pass++) {
    int* fill = NULL;
    if (pass == 1) {
      for (int t = 0; t < tiles; t++) b->start[t + 1] += b->start[t];
      b->list = (int*)malloc(sizeof(int) * (size_t)(b->start[tiles] > 0 ? b->start[tiles] : 1));
      fill = (int*)malloc(sizeof(int) * (size_t)tiles);
      memcpy(fill, b->start, sizeof(int) * (size_t)tiles);
    }
    for (int j = 0; j < m; j++) {
      int x0, x1, y0, y1;
      const int i = order[j];
      if (g_binning) {
        if (!splat_rect(&sp[i], W, H, &x0, &x1, &y0, &y1)) continue;
      } else {
        x0 = 0, x1 = W - 1, y0 = 0, y1 = H - 1;
      }
      for (int ty = y0 / BIN_TILE; ty <= y1 / BIN_TILE; ty++)
        for (int tx = x0 / BIN_TILE; tx <= x1 / BIN_TILE; tx++) {
          const int t = ty * b->tiles_x + tx;
          if (pass == 0)
            b->start[t + 1]++;
          else
            b->list[fill[t]++] = i;
        }
    }
    free(fill);
  }
}

static void free_bins(Bins* b) {
  free(b->start);
  free(b->list);
}

/* alpha of splat s at pixel p, or -1 where discarded */
static double pixel_alpha(const DSplat* s, const OrcRaster* r, double px, double py, double* dx, double* dy, double* g) {
  *dx = px - s->mu2[0];
  *dy = py - s->mu2[1];
  const double z2 = s->Q[0] * *dx * *dx + 2.0 * s->Q[1] * *dx * *dy + s->Q[2] * *dy * *dy;
  if (z2 > s->cut * s->cut) return -1.0;
  *g = exp(-0.5 * z2);
  const double alpha = s->opc * *g;
  if (alpha < r->min_alpha) return -1.0;
  return alpha;
}

int orc_diff_render(const double* theta, int n, const OrcRaster* r, double* image, double* alpha_out) {
  int* order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int m;
  DSplat* sp = forward_all(theta, n, r, order, &m);
  const int W = r->width, H = r->height;
  Bins bins;
  bin_splats(sp, order, m, W, H, &bins);
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      double T = 1.0, C[3] = {0, 0, 0};
      const int t = (y / BIN_TILE) * bins.tiles_x + x / BIN_TILE;
      for (int q = bins.start[t]; q < bins.start[t + 1]; q++) {
        const DSplat* s = &sp[bins.list[q]];
        double dx, dy, g;
        const double a = pixel_alpha(s, r, x + 0.5, y + 0.5, &dx, &dy, &g);
        if (a < 0.0) continue;
        const double Tn = T * (1.0 - a);
        if (Tn < r->t_stop) break;
        for (int c = 0; c < 3; c++) C[c] += s->col[c] * a * T;
        T = Tn;
      }
      for (int c = 0; c < 3; c++) image[(size_t)c * W * H + (size_t)y * W + x] = C[c];
      if (alpha_out) alpha_out[(size_t)y * W + x] = 1.0 - T;
    }
  free_bins(&bins);
  free(sp);
  free(order);
  return m;
}

/* The adjoint of one splat's forward: from the gradients of its pixel-side
   quantities (mu2, Q, opc, col) to its 59 parameters. */
static void splat_backward(const double* th, const OrcRaster* r, const double Vl[9], const double cam[3], const DSplat* s,
                           const double gmu2[2], const double gQ[3], double gopc, const double gcol[3], double* g) {
  memset(g, 0, sizeof(double) * NP);
  (void)cam;
  /* colour */
  static const int coeffs[4] = {0, 3, 8, 15};
  const int nk = coeffs[r->sh_degree < 0 ? 0 : (r->sh_degree > 3 ? 3 : r->sh_degree)];
  double graw[3], gdir[3] = {0, 0, 0};
  for (int c = 0; c < 3; c++) graw[c] = s->col_raw[c] > 0.0 ? gcol[c] : 0.0;
  for (int c = 0; c < 3; c++) g[11 + c] = graw[c];
  double bgrad[24][3];
  orc_sh_basis_grad(s->dir, bgrad);
  for (int k = 0; k < nk; k++)
    for (int c = 0; c < 3; c++) {
      g[14 + 3 * k + c] = s->basis[k] * graw[c];
      for (int e = 0; e < 3; e++) gdir[e] += th[14 + 3 * k + c] * graw[c] * bgrad[k][e];
    }
  double gmu[3] = {0, 0, 0};
  if (s->dlen > 0) {
    const double dd = gdir[0] * s->dir[0] + gdir[1] * s->dir[1] + gdir[2] * s->dir[2];
    for (int e = 0; e < 3; e++) gmu[e] += (gdir[e] - s->dir[e] * dd) / s->dlen;
  }
  /* Q = inverse(Sigma2'): dSigma2' = -Q G Q, G = [[gQ0, gQ1/2], [gQ1/2, gQ2]] */
  const double Q00 = s->Q[0], Q01 = s->Q[1], Q11 = s->Q[2];
  const double G00 = gQ[0], G01 = 0.5 * gQ[1], G11 = gQ[2];
  const double QG00 = Q00 * G00 + Q01 * G01, QG01 = Q00 * G01 + Q01 * G11;
  const double QG10 = Q01 * G00 + Q11 * G01, QG11 = Q01 * G01 + Q11 * G11;
  const double S00 = -(QG00 * Q00 + QG01 * Q01), S01 = -(QG00 * Q01 + QG01 * Q11), S11 = -(QG10 * Q01 + QG11 * Q11);
  double ga = S00, gd = S11, gb = 2.0 * S01;
  /* opacity compensation: opc = op sqrt(det_o / det') */
  double gop = 0.0;
  if (s->ratio > 0.0) {
    const double h = 0.5 * s->opc;
    ga += gopc * h * (s->d / s->det_o - s->dp / s->det_p);
    gd += gopc * h * (s->a / s->det_o - s->ap / s->det_p);
    gb += gopc * h * (-2.0 * s->b / s->det_o + 2.0 * s->b / s->det_p);
    gop = gopc * sqrt(s->ratio);
  }
  g[10] = gop * s->op * (1.0 - s->op);
  /* cov2D = T Sigma T^T: dSigma = T^T G2 T, dT = 2 G2 T Sigma */
  const double G2[4] = {ga, 0.5 * gb, 0.5 * gb, gd};
  double gSigma[9], gT[6];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int k = 0; k < 2; k++)
        for (int l = 0; l < 2; l++) v += s->T[3 * k + i] * G2[2 * k + l] * s->T[3 * l + j];
      gSigma[3 * i + j] = v;
    }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int l = 0; l < 2; l++)
        for (int k = 0; k < 3; k++) v += G2[2 * i + l] * s->T[3 * l + k] * s->Sigma[3 * k + j];
      gT[3 * i + j] = 2.0 * v;
    }
  /* T = J Vl: dJ = dT Vl^T */
  double gJ[6];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 3; j++) gJ[3 * i + j] = gT[3 * i] * Vl[3 * j] + gT[3 * i + 1] * Vl[3 * j + 1] + gT[3 * i + 2] * Vl[3 * j + 2];
  const double* P = r->proj;
  const double fx = 0.5 * r->width * P[0], fy = 0.5 * r->height * P[5];
  const double x = s->t[0], y = s->t[1], z = s->t[2];
  double gt[3];
  gt[0] = gJ[2] * (-fx / (z * z));
  gt[1] = gJ[5] * (-fy / (z * z));
  gt[2] = gJ[0] * (-fx / (z * z)) + gJ[2] * (2.0 * fx * x / (z * z * z)) + gJ[4] * (-fy / (z * z)) +
          gJ[5] * (2.0 * fy * y / (z * z * z));
  /* mu2 = ((clip.xy / w + 1) / 2) size */
  const double w = s->clip[3];
  const double gcx = gmu2[0] * 0.5 * r->width / w, gcy = gmu2[1] * 0.5 * r->height / w;
  const double gcw = -(gmu2[0] * 0.5 * r->width * s->clip[0] + gmu2[1] * 0.5 * r->height * s->clip[1]) / (w * w);
  for (int k = 0; k < 3; k++) gt[k] += gcx * P[4 * k] + gcy * P[4 * k + 1] + gcw * P[4 * k + 3];
  /* t = Vl mu + Vt */
  for (int k = 0; k < 3; k++) gmu[k] += Vl[k] * gt[0] + Vl[3 + k] * gt[1] + Vl[6 + k] * gt[2];
  for (int k = 0; k < 3; k++) g[k] = gmu[k];
  /* Sigma = M M^T, M = R diag(s): dM = 2 dSigma M (dSigma symmetric) */
  double gM[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int k = 0; k < 3; k++) v += (gSigma[3 * i + k] + gSigma[3 * k + i]) * s->M[3 * k + j];
      gM[3 * i + j] = v;
    }
  double gR[9];
  for (int j = 0; j < 3; j++) {
    double gs = 0;
    for (int i = 0; i < 3; i++) {
      gs += gM[3 * i + j] * s->R[3 * i + j];
      gR[3 * i + j] = gM[3 * i + j] * s->s[j];
    }
    g[3 + j] = gs * s->s[j];
  }
  /* R(q) for the unit quaternion (x, y, z, w) */
  const double qx = s->qv[0], qy = s->qv[1], qz = s->qv[2], qw = s->qv[3];
  double gq[4];
  gq[0] = 2.0 * (gR[1] * qy + gR[2] * qz + gR[3] * qy + gR[5] * (-qw) + gR[6] * qz + gR[7] * qw) +
          -4.0 * qx * (gR[4] + gR[8]);
  gq[1] = 2.0 * (gR[1] * qx + gR[2] * qw + gR[3] * qx + gR[5] * qz + gR[6] * (-qw) + gR[7] * qz) +
          -4.0 * qy * (gR[0] + gR[8]);
  gq[2] = 2.0 * (gR[1] * (-qw) + gR[2] * qx + gR[3] * qw + gR[5] * qy + gR[6] * qx + gR[7] * qy) +
          -4.0 * qz * (gR[0] + gR[4]);
  gq[3] = 2.0 * (gR[1] * (-qz) + gR[2] * qy + gR[3] * qz + gR[5] * (-qx) + gR[6] * (-qy) + gR[7] * qx);
  if (s->qlen > 0) {
    const double dq = gq[0] * qx + gq[1] * qy + gq[2] * qz + gq[3] * qw;
    for (int k = 0; k < 4; k++) g[6 + k] = (gq[k] - s->qv[k] * dq) / s->qlen;
  }
}

/* Pixel pass of the loss and its adjoint. grad NULL: loss only. lists:
   when set, the per-pixel contributor lists to use instead of the cutoff,
   minAlpha and transmittance decisions (the frozen evaluation). */
typedef struct OrcDiffFrozen {
  int width, height;
  int* start;  /* H*W + 1 */
  int* list;
} OrcDiffFrozen;

static double loss_pass(const double* theta, int n, const OrcRaster* r, const double* target, double* grad,
                        const OrcDiffFrozen* frozen, OrcDiffFrozen* record) {
  int* order = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int m;
  DSplat* sp = forward_all(theta, n, r, order, &m);
  const int W = r->width, H = r->height;
  const double norm = 1.0 / (3.0 * W * H);
  double* acc = NULL; /* per splat: gmu2 2, gQ 3, gopc 1, gcol 3 */
  if (grad) acc = (double*)calloc((size_t)n * 9 + 1, sizeof(double));
  int* pl = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
  double* pa = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* pT = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* pg = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* pdx = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* pdy = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  size_t rec_cap = 0, rec_len = 0;
  if (record) {
    record->width = W;
    record->height = H;
    record->start = (int*)malloc(sizeof(int) * ((size_t)W * H + 1));
    rec_cap = (size_t)W * H * 4 + 16;
    record->list = (int*)malloc(sizeof(int) * rec_cap);
  }
  Bins bins;
  bin_splats(sp, order, frozen ? 0 : m, W, H, &bins);
  double loss = 0.0;
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const size_t p = (size_t)y * W + x;
      const int tile = (y / BIN_TILE) * bins.tiles_x + x / BIN_TILE;
      double T = 1.0, C[3] = {0, 0, 0};
      int cnt = 0;
      if (frozen) {
        for (int q = frozen->start[p]; q < frozen->start[p + 1]; q++) {
          const int i = frozen->list[q];
          const DSplat* s = &sp[i];
          double dx = x + 0.5 - s->mu2[0], dy = y + 0.5 - s->mu2[1];
          const double z2 = s->Q[0] * dx * dx + 2.0 * s->Q[1] * dx * dy + s->Q[2] * dy * dy;
          const double g = exp(-0.5 * z2);
          const double a = s->opc * g;
          pl[cnt] = i;
          pa[cnt] = a;
          pT[cnt] = T;
          pg[cnt] = g;
          pdx[cnt] = dx;
          pdy[cnt] = dy;
          cnt++;
          for (int c = 0; c < 3; c++) C[c] += s->col[c] * a * T;
          T *= 1.0 - a;
        }
      } else {
        for (int q = bins.start[tile]; q < bins.start[tile + 1]; q++) {
          const int i = bins.list[q];
          const DSplat* s = &sp[i];
          double dx, dy, g;
          const double a = pixel_alpha(s, r, x + 0.5, y + 0.5, &dx, &dy, &g);
          if (a < 0.0) continue;
          const double Tn = T * (1.0 - a);
          if (Tn < r->t_stop) break;
          pl[cnt] = i;
          pa[cnt] = a;
          pT[cnt] = T;
          pg[cnt] = g;
          pdx[cnt] = dx;
          pdy[cnt] = dy;
          cnt++;
          for (int c = 0; c < 3; c++) C[c] += s->col[c] * a * T;
          T = Tn;
        }
      }
      if (record) {
        record->start[p] = (int)rec_len;
        if (rec_len + (size_t)cnt > rec_cap) {
          rec_cap = 2 * (rec_len + (size_t)cnt);
          record->list = (int*)realloc(record->list, sizeof(int) * rec_cap);
        }
        for (int q = 0; q < cnt; q++) record->list[rec_len++] = pl[q];
      }
      double dC[3];
      for (int c = 0; c < 3; c++) {
        const double e = C[c] - target[(size_t)c * W * H + p];
        loss += e * e;
        dC[c] = 2.0 * e * norm;
      }
      if (!grad) continue;
      /* back to front: B = colour composited behind the current splat */
      double B[3] = {0, 0, 0};
      for (int q = cnt - 1; q >= 0; q--) {
        const int i = pl[q];
        const DSplat* s = &sp[i];
        double* A = acc + (size_t)9 * i;
        const double a = pa[q], Ti = pT[q];
        double galpha = 0.0;
        for (int c = 0; c < 3; c++) {
          A[6 + c] += a * Ti * dC[c];
          galpha += Ti * (s->col[c] - B[c]) * dC[c];
          B[c] = s->col[c] * a + (1.0 - a) * B[c];
        }
        A[5] += pg[q] * galpha;
        const double gz2 = -0.5 * a * galpha;
        const double dx = pdx[q], dy = pdy[q];
        A[0] += gz2 * -2.0 * (s->Q[0] * dx + s->Q[1] * dy);
        A[1] += gz2 * -2.0 * (s->Q[1] * dx + s->Q[2] * dy);
        A[2] += gz2 * dx * dx;
        A[3] += gz2 * 2.0 * dx * dy;
        A[4] += gz2 * dy * dy;
      }
    }
  if (record) record->start[(size_t)W * H] = (int)rec_len;
  loss *= norm;
  if (grad) {
    double Vl[9], Vt[3], cam[3];
    view_parts(r, Vl, Vt, cam);
    for (int i = 0; i < n; i++) {
      double* g = grad + (size_t)NP * i;
      if (!sp[i].ok) {
        memset(g, 0, sizeof(double) * NP);
        continue;
      }
      const double* A = acc + (size_t)9 * i;
      splat_backward(theta + (size_t)NP * i, r, Vl, cam, &sp[i], A, A + 2, A[5], A + 6, g);
    }
    free(acc);
  }
  free_bins(&bins);
  free(pl);
  free(pa);
  free(pT);
  free(pg);
  free(pdx);
  free(pdy);
  free(sp);
  free(order);
  return loss;
}

double orc_diff_loss(const double* theta, int n, const OrcRaster* r, const double* target) {
  return loss_pass(theta, n, r, target, NULL, NULL, NULL);
}

double orc_diff_grad(const double* theta, int n, const OrcRaster* r, const double* target, double* grad) {
  return loss_pass(theta, n, r, target, grad, NULL, NULL);
}

OrcDiffFrozen* orc_diff_freeze(const double* theta, int n, const OrcRaster* r, const double* target) {
  OrcDiffFrozen* f = (OrcDiffFrozen*)calloc(1, sizeof(OrcDiffFrozen));
  loss_pass(theta, n, r, target, NULL, NULL, f);
  return f;
}

double orc_diff_loss_frozen(const OrcDiffFrozen* f, const double* theta, int n, const OrcRaster* r,
                            const double* target) {
  return loss_pass(theta, n, r, target, NULL, f, NULL);
}

void orc_diff_free(OrcDiffFrozen* f) {
  if (!f) return;
  free(f->start);
  free(f->list);
  free(f);
}

/* Gradient of the SH basis (orc_sh_basis) with respect to the direction. */
void orc_sh_basis_grad(const double* d, double (*g)[3]) {
  const double x = d[0], y = d[1], z = d[2];
  const double xx = x * x, yy = y * y, zz = z * z;
#define G(k, a, b, c) (g[k][0] = (a), g[k][1] = (b), g[k][2] = (c))
  G(0, 0, -0.4886025, 0);
  G(1, 0, 0, 0.4886025);
  G(2, -0.4886025, 0, 0);
  G(3, 1.0925484 * y, 1.0925484 * x, 0);
  G(4, 0, -1.0925484 * z, -1.0925484 * y);
  G(5, 0.3153916 * -2 * x, 0.3153916 * -2 * y, 0.3153916 * 4 * z);
  G(6, -1.0925484 * z, 0, -1.0925484 * x);
  G(7, 0.5462742 * 2 * x, 0.5462742 * -2 * y, 0);
  G(8, -0.5900436 * y * 6 * x, -0.5900436 * (3 * xx - 3 * yy), 0);
  G(9, 2.8906114 * y * z, 2.8906114 * x * z, 2.8906114 * x * y);
  G(10, -0.4570458 * y * -2 * x, -0.4570458 * (4 * zz - xx - 3 * yy), -0.4570458 * y * 8 * z);
  G(11, 0.3731763 * z * -6 * x, 0.3731763 * z * -6 * y, 0.3731763 * (6 * zz - 3 * xx - 3 * yy));
  G(12, -0.4570458 * (4 * zz - 3 * xx - yy), -0.4570458 * x * -2 * y, -0.4570458 * x * 8 * z);
  G(13, 1.4453057 * z * 2 * x, 1.4453057 * z * -2 * y, 1.4453057 * (xx - yy));
  G(14, -0.5900436 * (3 * xx - 3 * yy), -0.5900436 * x * -6 * y, 0);
  G(15, 2.5033429 * y * (3 * xx - yy), 2.5033429 * x * (xx - 3 * yy), 0);
  G(16, -1.7701308 * y * z * 6 * x, -1.7701308 * z * (3 * xx - 3 * yy), -1.7701308 * y * (3 * xx - yy));
  G(17, 0.9461747 * y * (7 * zz - 1), 0.9461747 * x * (7 * zz - 1), 0.9461747 * x * y * 14 * z);
  G(18, 0, -0.6690465 * z * (7 * zz - 3), -0.6690465 * y * (21 * zz - 3));
  G(19, 0, 0, 0.1057855 * (140 * zz * z - 60 * z));
  G(20, -0.6690465 * z * (7 * zz - 3), 0, -0.6690465 * x * (21 * zz - 3));
  G(21, 0.4730873 * 2 * x * (7 * zz - 1), 0.4730873 * -2 * y * (7 * zz - 1), 0.4730873 * (xx - yy) * 14 * z);
  G(22, -1.7701308 * z * (3 * xx - 3 * yy), -1.7701308 * x * z * -6 * y, -1.7701308 * x * (xx - 3 * yy));
  G(23, 0.6258357 * (4 * xx * x - 12 * x * yy), 0.6258357 * (-12 * xx * y + 4 * yy * y), 0);
#undef G
}

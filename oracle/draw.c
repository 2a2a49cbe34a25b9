/* Test infrastructure: the CPU oracle of the splat draw (FP64), checked by
 * the tests, never part of the product.
 *
 * The ewa method's draw as the reference specifies it, one splat at a time
 * in draw order, each pixel blended as the GPU's blend unit would:
 *  - src/render/splatDraw.ts:98-181: splat order[i] for i < count (indices
 *    >= n skipped, :121-122); the LoD and cutoff rule (:127-140); the
 *    fragment (:165-178): discard z^2 > cut^2, alpha = fragmentAlpha,
 *    discard alpha < minAlpha, colour max(rgb, 0), sRGB-decoded for linear
 *    outputs; premultiplied "over" blending (:62-78).
 *  - src/render/projections.tsl.ts:150-167: look (visible iff the opacity
 *    byte is nonzero, rgb = decodeExtRgb(w2), opacity = byte / 128), the
 *    quad (boundCorner with the corner scaled by cut / bound), uv = corner *
 *    cut, density |uv|^2.
 *  - src/render/projection.tsl.ts:277-311 (decoders), 396-405 (boundCorner),
 *    444-455 (fragmentAlpha).
 *  - three.js r186 sRGBTransferEOTF (the IEC 61966-2-1 curve).
 * The uv of a pixel centre follows from the quad's affine map: its offset d
 * from the centre in pixels (y up, as ndc) gives corner = (d.axis1 / (sx k),
 * d.axis2 / (sy k)) with k = cut / bound. Rows of the image run top to
 * bottom (framebuffer rows): pixel (x, y) has its centre at ndc
 * ((x + 0.5) 2 / W - 1, 1 - (y + 0.5) 2 / H). */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static double decode_fixed16(uint32_t code) { return (double)code / 65535.0 * 4.0 - 2.0; }
static double decode_log_scale12(uint32_t code) { return exp2((double)(code - 1u) / 4094.0 * 16.0 - 4.0); }
static double srgb_eotf(double c) {
  return c <= 0.04045 ? c * 0.0773993808 : pow(c * 0.9478672986 + 0.0521327014, 2.4);
}

void orc_draw(const uint32_t* words, uint32_t n, const uint32_t* order, uint32_t count, int width, int height,
              double max_std_dev, double min_alpha, double falloff, int linear, double* rgba) {
  const size_t plane = (size_t)width * height;
  memset(rgba, 0, 4 * plane * sizeof(double));
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t sp = order[i];
    if (sp >= n) continue;
    const uint32_t* w = words + 4 * (size_t)sp;
    const uint32_t byte = w[3] & 0xffu;
    if (byte == 0u) continue;
    double alpha = (double)byte / 128.0;
    double bound = max_std_dev, cut = max_std_dev;
    if (alpha > 1.0) {
      alpha = fmin(alpha * 4.0 - 3.0, 5.0);
      bound = max_std_dev + (alpha - 1.0) * 0.7;
      cut = bound;
    } else if (falloff >= 1.0) {
      cut = fmin(cut, sqrt(log(fmax(alpha, min_alpha) / min_alpha) * 2.0));
    }
    double rgb[3];
    orc_decode_ext_rgb(w[2], rgb);
    for (int c = 0; c < 3; c++) {
      rgb[c] = fmax(rgb[c], 0.0);
      if (linear) rgb[c] = srgb_eotf(rgb[c]);
    }
    const double nx = decode_fixed16(w[0] & 0xffffu), ny = decode_fixed16(w[0] >> 16);
    const double ang = (double)(w[1] >> 20) / 4096.0 * M_PI;
    const double sx = decode_log_scale12((w[3] >> 8) & 0xfffu), sy = decode_log_scale12(w[3] >> 20);
    const double a1x = cos(ang), a1y = sin(ang), a2x = a1y, a2y = -a1x;
    const double k = cut / bound;
    // the quad's pixel extent (y up): corners at +-sx k axis1 +- sy k axis2
    const double ex = fabs(a1x) * sx * k + fabs(a2x) * sy * k, ey = fabs(a1y) * sx * k + fabs(a2y) * sy * k;
    const double cxp = (nx + 1.0) * 0.5 * width, cyp = (ny + 1.0) * 0.5 * height;  // centre in pixels, y up
    int x0 = (int)floor(cxp - ex - 1.0), x1 = (int)ceil(cxp + ex + 1.0);
    int yu0 = (int)floor(cyp - ey - 1.0), yu1 = (int)ceil(cyp + ey + 1.0);
    if (x0 < 0) x0 = 0;
    if (x1 > width - 1) x1 = width - 1;
    if (yu0 < 0) yu0 = 0;
    if (yu1 > height - 1) yu1 = height - 1;
    const double lod_a = alpha > 1.0 ? exp((alpha * alpha - 1.0) / M_E) : 0.0;
    for (int yu = yu0; yu <= yu1; yu++) {
      const int y = height - 1 - yu;  // framebuffer row
      for (int x = x0; x <= x1; x++) {
        // the pixel centre in ndc, then its offset from the centre in pixels
        const double px = ((x + 0.5) * 2.0 / width - 1.0 - nx) * 0.5 * width;
        const double py = ((1.0 - (y + 0.5) * 2.0 / height) - ny) * 0.5 * height;
        const double cx = (px * a1x + py * a1y) / (sx * k), cy = (px * a2x + py * a2y) / (sy * k);
        if (fabs(cx) > 1.0 || fabs(cy) > 1.0) continue;  // outside the quad
        const double u = cx * cut, v = cy * cut;
        const double z2 = u * u + v * v;
        if (z2 > cut * cut) continue;
        const double g = exp(-0.5 * z2);
        double a;
        if (alpha <= 1.0) {
          a = alpha + (alpha * g - alpha) * falloff;
        } else {
          a = 1.0 + ((1.0 - pow(1.0 - g, lod_a)) - 1.0) * falloff;
        }
        if (a < min_alpha) continue;
        const size_t p = (size_t)y * width + x;
        for (int c = 0; c < 3; c++) rgba[c * plane + p] = rgb[c] * a + (1.0 - a) * rgba[c * plane + p];
        rgba[3 * plane + p] = a + (1.0 - a) * rgba[3 * plane + p];
      }
    }
  }
}

/* --- the draw of the ray-evaluated methods (FP64) ------------------------------
 * src/render/splatDraw.ts:98-181 with the methods' vertex and density stages:
 *  - 3dgut and htgs (projections.tsl.ts:310-330; projection.tsl.ts:396-442):
 *    the head words give the head-space centre (decodeCenter), rotation
 *    (QuatOct32) and scales (half-float logs); the eye's fromHead turns them
 *    into the eye (centre = fromHead (c, 1), quat = fromHeadQuat * q). The
 *    quad is the eye's screen bound (decodeBound) with its corner scaled by
 *    cut / bound (3dgut; htgs keeps 1) when the bound lies wholly on screen.
 *    At a pixel the ray (perspective: from the eye through the pixel's view
 *    position at the centre's depth; orthographic: from that position's xy
 *    along -z) in the splat's whitened frame gives z^2 = |o x d|^2 / |d|^2
 *    (rayDensity), 1e30 for a null direction.
 *  - raygs (projections.tsl.ts:432-466): the eye words hold the facing quad
 *    (QuatOct32 frame in the eye, c^2, half-float log axis lengths); the quad
 *    spans z in [-zmax, zmax]^2, zmax = cut / sqrt(max(1 - cut^2 / c^2,
 *    1e-4)), around the centre. A pixel's ray meets the quad's plane at
 *    centre + lengths.x z.x n0 + lengths.y z.y n1 (in front of the eye under
 *    perspective: the rasteriser draws the part with w > 0); the density is
 *    |z|^2 c^2 / (c^2 + |z|^2).
 * Words per splat (SplatProjection.ts:46-53): ch0 [4 i..4 i+3], ch1
 * [4 i..4 i+3], ch2 [2 i..2 i+1] with two eyes; 3dgut / htgs: 6 head words
 * then 2 per eye, raygs: 4 head words then 3 per eye.
 * from_head: the eye's head -> eye matrix, column-major 4x4; from_head_quat
 * its rotation (x, y, z, w); proj: the eye's projection, column-major. */
static void qrot(const double* q, const double* v, double* o) {
  const double tx = 2.0 * (q[1] * v[2] - q[2] * v[1]), ty = 2.0 * (q[2] * v[0] - q[0] * v[2]),
               tz = 2.0 * (q[0] * v[1] - q[1] * v[0]);
  o[0] = v[0] + q[3] * tx + (q[1] * tz - q[2] * ty);
  o[1] = v[1] + q[3] * ty + (q[2] * tx - q[0] * tz);
  o[2] = v[2] + q[3] * tz + (q[0] * ty - q[1] * tx);
}
static void qmul(const double* a, const double* b, double* o) {
  o[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
  o[1] = a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0];
  o[2] = a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3];
  o[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
}
static double decode_log_scale11(uint32_t code) { return exp2((double)(code - 1u) / 2046.0 * 16.0 - 4.0); }

/* The view position at depth vz under ndc (unproject, projection.tsl.ts:386-393) */
static void unproject_ndc(const double* P, double nx, double ny, double vz, double* v) {
  const double w = P[11] * vz + P[15];  /* clipW: element(2).w * z + element(3).w */
  v[0] = (nx * w - P[8] * vz - P[12]) / P[0];
  v[1] = (ny * w - P[9] * vz - P[13]) / P[5];
  v[2] = vz;
}

void orc_draw_ray(int method, const uint32_t* ch0, const uint32_t* ch1, const uint32_t* ch2, uint32_t n, int eye,
                  const uint32_t* order, uint32_t count, const double* from_head, const double* from_head_quat,
                  const double* proj, int ortho, int width, int height, double max_std_dev, double min_alpha,
                  double falloff, int linear, double* rgba) {
  const size_t plane = (size_t)width * height;
  memset(rgba, 0, 4 * plane * sizeof(double));
  const int raygs = method == 3;
  const int head_words = raygs ? 4 : 6, eye_words = raygs ? 3 : 2;
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t sp = order[i];
    if (sp >= n) continue;
    uint32_t w[10];
    for (int k = 0; k < 4; k++) w[k] = ch0[4 * (size_t)sp + k];
    for (int k = 0; k < 4; k++) w[4 + k] = ch1[4 * (size_t)sp + k];
    w[8] = ch2 ? ch2[2 * (size_t)sp] : 0u;
    w[9] = ch2 ? ch2[2 * (size_t)sp + 1] : 0u;
    const uint32_t* e = w + head_words + eye_words * eye;
    if (e[1] == 0u) continue;  /* look.visible */
    const uint32_t last = raygs ? w[3] : w[5];
    double alpha = (double)(raygs ? (w[3] & 0xffu) : ((w[5] >> 16) & 0xffu)) / 128.0;
    double bound = max_std_dev, cut = max_std_dev;
    if (alpha > 1.0) {
      alpha = fmin(alpha * 4.0 - 3.0, 5.0);
      bound = max_std_dev + (alpha - 1.0) * 0.7;
      cut = bound;
    } else if (falloff >= 1.0) {
      cut = fmin(cut, sqrt(log(fmax(alpha, min_alpha) / min_alpha) * 2.0));
    }
    double rgb[3];
    orc_decode_ext_rgb(w[2], rgb);
    for (int c = 0; c < 3; c++) {
      rgb[c] = fmax(rgb[c], 0.0);
      if (linear) rgb[c] = srgb_eotf(rgb[c]);
    }
    double hc[3], c[3];
    orc_decode_center(w[0], w[1], last, hc);
    for (int r = 0; r < 3; r++)
      c[r] = from_head[r] * hc[0] + from_head[4 + r] * hc[1] + from_head[8 + r] * hc[2] + from_head[12 + r];
    const double lod_a = alpha > 1.0 ? exp((alpha * alpha - 1.0) / M_E) : 0.0;
    /* method state */
    double q[4], sc[3] = {1, 1, 1}, a1[2] = {1, 0}, bnx = 0, bny = 0, bsx = 1, bsy = 1, corner = 1;
    double n0[3], n1[3], nrm[3], len[2] = {1, 1}, c2 = 1, zmax = 0;
    int x0 = 0, x1 = width - 1, y0 = 0, y1 = height - 1;  /* the pixels to visit */
    if (!raygs) {
      double hq[4];
      orc_decode_quat_oct(w[3], 10, 12, hq);
      qmul(from_head_quat, hq, q);
      sc[0] = exp(orc_decode_half(w[4] & 0xffffu));
      sc[1] = exp(orc_decode_half(w[4] >> 16));
      sc[2] = exp(orc_decode_half(w[5] & 0xffffu));
      bnx = decode_fixed16(e[0] & 0xffffu);
      bny = decode_fixed16(e[0] >> 16);
      const double ang = (double)(e[1] & 1023u) / 1024.0 * M_PI;
      bsx = decode_log_scale11((e[1] >> 10) & 0x7ffu);
      bsy = decode_log_scale11(e[1] >> 21);
      a1[0] = cos(ang);
      a1[1] = sin(ang);
      const double a2[2] = {a1[1], -a1[0]};
      const double hx = width * 0.5, hy = height * 0.5;
      const double ext_x = sqrt(a1[0] * bsx * a1[0] * bsx + a2[0] * bsy * a2[0] * bsy);
      const double ext_y = sqrt(a1[1] * bsx * a1[1] * bsx + a2[1] * bsy * a2[1] * bsy);
      const int on_screen = fabs(bnx * hx) + ext_x < hx && fabs(bny * hy) + ext_y < hy;
      const double shrink = method == 1 ? cut / bound : 1.0;
      corner = on_screen ? shrink : 1.0;
      /* the quad's box (the rectangle +-bsx axis1 +-bsy axis2, times corner), a pixel wider */
      const double cxp = (bnx + 1.0) * hx, cyp = (1.0 - bny) * hy;  /* framebuffer pixels */
      const double bx = (fabs(a1[0]) * bsx + fabs(a1[1]) * bsy) * corner, by = (fabs(a1[1]) * bsx + fabs(a1[0]) * bsy) * corner;
      const int bx0 = (int)floor(cxp - bx - 1.0), bx1 = (int)ceil(cxp + bx + 1.0);
      const int by0 = (int)floor(cyp - by - 1.0), by1 = (int)ceil(cyp + by + 1.0);
      x0 = bx0 > 0 ? bx0 : 0;
      x1 = bx1 < width - 1 ? bx1 : width - 1;
      y0 = by0 > 0 ? by0 : 0;
      y1 = by1 < height - 1 ? by1 : height - 1;
    } else {
      orc_decode_quat_oct(e[0], 10, 12, q);
      float f;
      memcpy(&f, &e[1], 4);
      c2 = f;
      len[0] = exp(orc_decode_half(e[2] & 0xffffu));
      len[1] = exp(orc_decode_half(e[2] >> 16));
      const double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0}, ez[3] = {0, 0, 1};
      qrot(q, ex, n0);
      qrot(q, ey, n1);
      qrot(q, ez, nrm);
      zmax = cut / sqrt(fmax(1.0 - cut * cut / c2, 1e-4));
    }
    const double qc[4] = {-q[0], -q[1], -q[2], q[3]};
    for (int y = y0; y <= y1; y++) {
      for (int x = x0; x <= x1; x++) {
        const double nx = (x + 0.5) * 2.0 / width - 1.0, ny = 1.0 - (y + 0.5) * 2.0 / height;
        double vp[3], o[3], d[3];
        unproject_ndc(proj, nx, ny, c[2], vp);
        if (ortho) {
          o[0] = vp[0], o[1] = vp[1], o[2] = 0.0;
          d[0] = 0.0, d[1] = 0.0, d[2] = -1.0;
        } else {
          o[0] = o[1] = o[2] = 0.0;
          d[0] = vp[0], d[1] = vp[1], d[2] = vp[2];
        }
        double z2;
        if (!raygs) {
          /* inside the quad: the pixel's offset from the bound's centre in pixels (y up) */
          const double px = (nx - bnx) * 0.5 * width, py = (ny - bny) * 0.5 * height;
          const double u = (px * a1[0] + py * a1[1]) / bsx, v = (px * a1[1] - py * a1[0]) / bsy;
          if (fabs(u) > corner || fabs(v) > corner) continue;
          double oc[3] = {o[0] - c[0], o[1] - c[1], o[2] - c[2]}, lo[3], ld[3];
          qrot(qc, oc, lo);
          qrot(qc, d, ld);
          for (int k = 0; k < 3; k++) {
            lo[k] /= sc[k];
            ld[k] /= sc[k];
          }
          const double cr[3] = {lo[1] * ld[2] - lo[2] * ld[1], lo[2] * ld[0] - lo[0] * ld[2], lo[0] * ld[1] - lo[1] * ld[0]};
          const double dd = ld[0] * ld[0] + ld[1] * ld[1] + ld[2] * ld[2];
          z2 = dd <= 0.0 ? 1e30 : (cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]) / dd;
        } else {
          const double dn = d[0] * nrm[0] + d[1] * nrm[1] + d[2] * nrm[2];
          if (dn == 0.0) continue;
          const double t = ((c[0] - o[0]) * nrm[0] + (c[1] - o[1]) * nrm[1] + (c[2] - o[2]) * nrm[2]) / dn;
          if (!ortho && !(t > 0.0)) continue;
          const double X[3] = {o[0] + t * d[0] - c[0], o[1] + t * d[1] - c[1], o[2] + t * d[2] - c[2]};
          const double zx = (X[0] * n0[0] + X[1] * n0[1] + X[2] * n0[2]) / len[0];
          const double zy = (X[0] * n1[0] + X[1] * n1[1] + X[2] * n1[2]) / len[1];
          if (fabs(zx) > zmax || fabs(zy) > zmax) continue;
          const double zz = zx * zx + zy * zy;
          z2 = zz * c2 / (c2 + zz);
        }
        if (z2 > cut * cut) continue;
        const double g = exp(-0.5 * z2);
        double a;
        if (alpha <= 1.0) {
          a = alpha + (alpha * g - alpha) * falloff;
        } else {
          a = 1.0 + ((1.0 - pow(1.0 - g, lod_a)) - 1.0) * falloff;
        }
        if (a < min_alpha) continue;
        const size_t p = (size_t)y * width + x;
        for (int k = 0; k < 3; k++) rgba[k * plane + p] = rgb[k] * a + (1.0 - a) * rgba[k * plane + p];
        rgba[3 * plane + p] = a + (1.0 - a) * rgba[3 * plane + p];
      }
    }
  }
}

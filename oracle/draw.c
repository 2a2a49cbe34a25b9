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

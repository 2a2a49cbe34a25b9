/*
 * Oracle: the 3dgut, htgs and raygs projection words of one splat, restated
 * in double from the reference's TSL: SplatProjection.project / write
 * (src/render/SplatProjection.ts:130-151), headBasis, gutBound, htgsBound,
 * rayProjection.project, quadOffScreen, raygsFrame and raygs.project
 * (src/render/projections.tsl.ts:170-429), basisToQuat
 * (src/gpu/math.tsl.ts:116-148), and
 * the helpers of src/render/projection.tsl.ts (prepareSplat 42-64,
 * cullInvisible 66-69, splatInView 112-130, cullDepth 132-135,
 * inSortGuard 145-155, viewBasis 169-174, mipInflate 177-197,
 * boundFromCov 223-240, screenBound 243-248, clipBoundToScreen 250-275,
 * encodeLogScale 299-303, encodeOpacity 310, encodeCenter 317-333,
 * encodeBound 352-360), with the object transform of
 * src/render/splatTransform.tsl.ts:46-62 and 142-165 (a rotation times a
 * uniform scale turns the splat; any other goes through fromCov,
 * src/gpu/splats.tsl.ts:189-198, with symEigens, src/gpu/math.tsl.ts:46-113)
 * and the codecs of src/gpu/encoding.tsl.ts:50-134. Test infrastructure only.
 *
 * Matrices are column-major 4x4: element (row r, column c) at m[4c + r].
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

#define AT(m, r, c) ((m)[4 * (c) + (r)])

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double rel_margin(double x, double t) { return fabs(x - t) / fmax(fmax(fabs(x), fabs(t)), 1e-30); }

static void point_of(const double* m, const double* p, double w, double* out) {
  double r[4];
  for (int i = 0; i < 4; i++) r[i] = AT(m, i, 0) * p[0] + AT(m, i, 1) * p[1] + AT(m, i, 2) * p[2] + AT(m, i, 3) * w;
  memcpy(out, r, 4 * sizeof(double));
}

/* --- quaternions (x, y, z, w), three.js conventions --- */
static void qmul(const double* a, const double* b, double* o) {
  double r[4] = {a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1], a[3] * b[1] - a[0] * b[2] + a[1] * b[3] + a[2] * b[0],
                 a[3] * b[2] + a[0] * b[1] - a[1] * b[0] + a[2] * b[3], a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2]};
  memcpy(o, r, sizeof r);
}
static void qnorm(double* q) {
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n > 0)
    for (int k = 0; k < 4; k++) q[k] /= n;
}
static void qrot(const double* q, const double* v, double* o) {
  double t[3] = {2 * (q[1] * v[2] - q[2] * v[1]), 2 * (q[2] * v[0] - q[0] * v[2]), 2 * (q[0] * v[1] - q[1] * v[0])};
  double r[3] = {v[0] + q[3] * t[0] + (q[1] * t[2] - q[2] * t[1]), v[1] + q[3] * t[1] + (q[2] * t[0] - q[0] * t[2]),
                 v[2] + q[3] * t[2] + (q[0] * t[1] - q[1] * t[0])};
  memcpy(o, r, sizeof r);
}
/* The rotation quaternion of the upper-left 3x3 of a column-major matrix
 * divided by `scale` (Matrix4.decompose / Quaternion.setFromRotationMatrix) */
static void quat_of_matrix(const double* m, double scale, double* q) {
  double R[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) R[r][c] = AT(m, r, c) / scale;
  double tr = R[0][0] + R[1][1] + R[2][2];
  if (tr > 0) {
    double s = 0.5 / sqrt(tr + 1.0);
    q[3] = 0.25 / s;
    q[0] = (R[2][1] - R[1][2]) * s;
    q[1] = (R[0][2] - R[2][0]) * s;
    q[2] = (R[1][0] - R[0][1]) * s;
  } else if (R[0][0] > R[1][1] && R[0][0] > R[2][2]) {
    double s = 2.0 * sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]);
    q[3] = (R[2][1] - R[1][2]) / s;
    q[0] = 0.25 * s;
    q[1] = (R[0][1] + R[1][0]) / s;
    q[2] = (R[0][2] + R[2][0]) / s;
  } else if (R[1][1] > R[2][2]) {
    double s = 2.0 * sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]);
    q[3] = (R[0][2] - R[2][0]) / s;
    q[0] = (R[0][1] + R[1][0]) / s;
    q[1] = 0.25 * s;
    q[2] = (R[1][2] + R[2][1]) / s;
  } else {
    double s = 2.0 * sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]);
    q[3] = (R[1][0] - R[0][1]) / s;
    q[0] = (R[0][2] + R[2][0]) / s;
    q[1] = (R[1][2] + R[2][1]) / s;
    q[2] = 0.25 * s;
  }
  qnorm(q);
}
static double det3_of(const double* m) {
  return AT(m, 0, 0) * (AT(m, 1, 1) * AT(m, 2, 2) - AT(m, 1, 2) * AT(m, 2, 1)) -
         AT(m, 0, 1) * (AT(m, 1, 0) * AT(m, 2, 2) - AT(m, 1, 2) * AT(m, 2, 0)) +
         AT(m, 0, 2) * (AT(m, 1, 0) * AT(m, 2, 1) - AT(m, 1, 1) * AT(m, 2, 0));
}

/* --- codecs --- */
static uint32_t fixed16(double v) { return (uint32_t)clampd(floor((v + 2.0) / 4.0 * 65535.0 + 0.5), 0.0, 65535.0); }
static uint32_t log_scale(double scale, int bits, int outward) {
  double top = pow(2.0, bits) - 2.0;
  double x = (clampd(log2(fmax(scale, 1e-20)), -4.0, 12.0) + 4.0) / 16.0 * top;
  return (uint32_t)clampd(outward ? ceil(x) : floor(x + 0.5), 0.0, top) + 1u;
}
static uint32_t opacity8(double o) {
  double c = floor(clampd(o * 128.0, 0.0, 255.0) + 0.5);
  return c < 1.0 ? 1u : (uint32_t)c;
}
static uint32_t float_bits(double v) {
  float f = (float)v;
  uint32_t b;
  memcpy(&b, &f, 4);
  return b;
}
/* encodeQuatOct (encoding.tsl.ts:97-110) with octEncoder(10) (71-95) */
uint32_t orc_encode_quat_oct32_gpu(const double* q) {
  double len = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (!(len > 0)) return 0u;
  double s = (q[3] < 0 ? -1.0 : 1.0) / len;
  double angle = acos(fmin(q[3] * s, 1.0)) * 2.0;
  uint32_t acode = (uint32_t)floor(angle / M_PI * 4095.0 + 0.5);
  double ax = q[0] * s, ay = q[1] * s, az = q[2] * s;
  uint32_t oct = 0;
  double l1 = fabs(ax) + fabs(ay) + fabs(az);
  if (l1 > 0) {
    double px = ax / l1, py = ay / l1;
    if (az < 0) {
      double ox = px;
      px = (1.0 - fabs(py)) * (ox >= 0 ? 1.0 : -1.0);
      py = (1.0 - fabs(ox)) * (py >= 0 ? 1.0 : -1.0);
    }
    uint32_t cu = (uint32_t)floor((px * 0.5 + 0.5) * 1023.0 + 0.5), cv = (uint32_t)floor((py * 0.5 + 0.5) * 1023.0 + 0.5);
    oct = cu | (cv << 10);
    if (oct == 0) oct = 0xfffffu;
  }
  return oct | (acode << 20);
}
/* encodeCenter (projection.tsl.ts:317-333): dir, distance bits, low */
void orc_encode_center(const double* c, uint32_t* dir, uint32_t* dist, uint32_t* low) {
  *dist = float_bits(sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
  double l1 = fmax(fabs(c[0]) + fabs(c[1]) + fabs(c[2]), 1e-30);
  double px = c[0] / l1, py = c[1] / l1;
  double sx = px >= 0 ? 1.0 : -1.0, sy = py >= 0 ? 1.0 : -1.0;
  if (c[2] < 0) {
    double fx = (1.0 - fabs(py)) * sx, fy = (1.0 - fabs(px)) * sy;
    px = fx;
    py = fy;
  }
  uint32_t qx = (uint32_t)floor((px * 0.5 + 0.5) * 1048575.0 + 0.5), qy = (uint32_t)floor((py * 0.5 + 0.5) * 1048575.0 + 0.5);
  *dir = (qx >> 4) | ((qy >> 4) << 16);
  *low = ((qx & 15u) << 24) | (qy << 28);
}
/* decodeCenter (projection.tsl.ts:335-349) */
void orc_decode_center(uint32_t dir, uint32_t dist, uint32_t last, double* c) {
  uint32_t qx = ((dir & 0xffffu) << 4) | ((last >> 24) & 15u), qy = ((dir >> 16) << 4) | (last >> 28);
  double x = qx / 1048575.0 * 2.0 - 1.0, y = qy / 1048575.0 * 2.0 - 1.0;
  double z = 1.0 - fabs(x) - fabs(y);
  double t = fmax(-z, 0.0);
  x += x >= 0 ? -t : t;
  y += y >= 0 ? -t : t;
  double n = sqrt(x * x + y * y + z * z);
  float d;
  memcpy(&d, &dist, 4);
  c[0] = x / n * d;
  c[1] = y / n * d;
  c[2] = z / n * d;
}

typedef struct bound_t {
  double nx, ny, angle, s1, s2;
} bound_t;

/* boundFromCov; *margin: relative distance of the larger scale to minPixelRadius */
static int bound_from_cov(double a, double b, double d, double std_dev, const orc_params* p, bound_t* bd, double* margin) {
  double det = a * d - b * b;
  double half = (a - d) * 0.5;
  double e1 = (a + d) * 0.5 + sqrt(half * half + b * b);
  double e2 = e1 > 0 ? fmax(det / e1, 0.0) : 0.0;
  double s1 = fmin(p->max_pixel_radius, std_dev * sqrt(fmax(e1, 0.0)));
  double s2 = fmin(p->max_pixel_radius, std_dev * sqrt(e2));
  int ok = s1 >= p->min_pixel_radius || s2 >= p->min_pixel_radius;
  *margin = fmin(*margin, rel_margin(fmax(s1, s2), p->min_pixel_radius));
  *margin = fmin(*margin, rel_margin(std_dev * sqrt(fmax(e1, 0.0)), p->max_pixel_radius));
  if (ok) {
    double vx, vy;
    if (fabs(a - e1) >= fabs(d - e1)) {
      vx = b;
      vy = e1 - a;
    } else {
      vx = e1 - d;
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
    double ang = atan2(vy, vx);
    bd->angle = ang < 0 ? ang + M_PI : ang;
    bd->s1 = s1;
    bd->s2 = s2;
  }
  return ok;
}
static void screen_bound(const orc_view* v, const orc_params* p, bound_t* bd) {
  bd->nx = bd->ny = bd->angle = 0;
  bd->s1 = bd->s2 = fmin(p->max_pixel_radius, sqrt(v->size[0] * v->size[0] + v->size[1] * v->size[1]) * 0.5);
}
static int clip_to_screen(const orc_view* v, bound_t* bd, double* margin) {
  double hx = v->size[0] * 0.5, hy = v->size[1] * 0.5;
  double cx = cos(bd->angle) * bd->s1, cy = cos(bd->angle) * bd->s2, sx = sin(bd->angle) * bd->s1, sy = sin(bd->angle) * bd->s2;
  double ex = sqrt(cx * cx + sy * sy), ey = sqrt(sx * sx + cy * cy);
  double px = bd->nx * hx, py = bd->ny * hy;
  double lox = px - ex, loy = py - ey, hix = px + ex, hiy = py + ey;
  /* near a screen edge the box may be clipped or not: an unstable case */
  double ref = fmax(hx, hy);
  *margin = fmin(*margin, fmin(fmin(fabs(lox + hx), fabs(hix - hx)), fmin(fabs(loy + hy), fabs(hiy - hy))) / ref);
  if (!(lox >= -hx && loy >= -hy && hix <= hx && hiy <= hy)) {
    lox = fmax(lox, -hx);
    loy = fmax(loy, -hy);
    hix = fmin(hix, hx);
    hiy = fmin(hiy, hy);
    if (lox >= hix || loy >= hiy) return 0;
    bd->nx = (lox + hix) * 0.5 / hx;
    bd->ny = (loy + hiy) * 0.5 / hy;
    bd->angle = 0;
    bd->s1 = (hix - lox) * 0.5;
    bd->s2 = (hiy - loy) * 0.5;
  }
  return 1;
}
static void scaled_axes(const double* q, const double* s, double col[3][3]) {
  for (int i = 0; i < 3; i++) {
    double e[3] = {0, 0, 0};
    e[i] = s[i];
    qrot(q, e, col[i]);
  }
}

static int gut_bound(const orc_eye* E, const orc_view* v, const double* c, const double* q, const double* s, double std_dev,
                     const orc_params* p, bound_t* bd, double* margin) {
  double s3[3] = {s[0] * sqrt(3.0), s[1] * sqrt(3.0), s[2] * sqrt(3.0)};
  double col[3][3];
  scaled_axes(q, s3, col);
  double hx = v->size[0] * 0.5, hy = v->size[1] * 0.5;
  int front = 1;
  double px[7], py[7];
  for (int i = 0; i < 7; i++) {
    double pt[3] = {c[0], c[1], c[2]};
    if (i > 0) {
      double sg = (i & 1) ? 1.0 : -1.0;
      for (int k = 0; k < 3; k++) pt[k] += sg * col[(i - 1) / 2][k];
    }
    double cl[4];
    point_of(E->projection, pt, 1.0, cl);
    front = front && cl[3] > 0;
    *margin = fmin(*margin, fabs(cl[3]) / fmax(fabs(cl[3]) + fabs(cl[2]), 1e-30));
    px[i] = cl[0] / cl[3] * hx;
    py[i] = cl[1] / cl[3] * hy;
  }
  if (!front) {
    screen_bound(v, p, bd);
    return 1;
  }
  double mx = 0, my = 0;
  for (int i = 1; i < 7; i++) {
    mx += px[i] / 6.0;
    my += py[i] / 6.0;
  }
  double d0x = px[0] - mx, d0y = py[0] - my;
  double a = d0x * d0x * 2.0, b = d0x * d0y * 2.0, d = d0y * d0y * 2.0;
  for (int i = 1; i < 7; i++) {
    double dx = px[i] - mx, dy = py[i] - my;
    a += dx * dx / 6.0;
    b += dx * dy / 6.0;
    d += dy * dy / 6.0;
  }
  int ok = bound_from_cov(a, b, d, std_dev, p, bd, margin);
  bd->nx = mx / hx;
  bd->ny = my / hy;
  return ok;
}

static int htgs_bound(const orc_eye* E, const orc_view* v, const double* c, const double* q, const double* s, double std_dev,
                      const double* eye_ndc, const orc_params* p, bound_t* bd, double* margin) {
  double ss[3] = {s[0] * std_dev, s[1] * std_dev, s[2] * std_dev};
  double col[3][3];
  scaled_axes(q, ss, col);
  double T[3][4], c3[4];
  for (int i = 0; i < 3; i++) point_of(E->projection, col[i], 0.0, T[i]);
  point_of(E->projection, c, 1.0, c3);
  double xs[3] = {T[0][0], T[1][0], T[2][0]}, ys[3] = {T[0][1], T[1][1], T[2][1]}, ws[3] = {T[0][3], T[1][3], T[2][3]};
  double sww = ws[0] * ws[0] + ws[1] * ws[1] + ws[2] * ws[2];
  double qww = sww - c3[3] * c3[3];
  *margin = fmin(*margin, rel_margin(sww, c3[3] * c3[3]));
  int valid = qww < 0 && c3[3] > 0;
  double a = 0, b = 0, d = 0, cx = 0, cy = 0;
  if (valid) {
    double u[3], w[3];
    for (int i = 0; i < 3; i++) {
      u[i] = xs[i] * c3[3] - ws[i] * c3[0];
      w[i] = ys[i] * c3[3] - ws[i] * c3[1];
    }
#define DOT3(p_, q_) ((p_)[0] * (q_)[0] + (p_)[1] * (q_)[1] + (p_)[2] * (q_)[2])
    double xw = DOT3(xs, ws), yw = DOT3(ys, ws);
    double inv_c2 = 1.0 / (qww * qww);
    double hx = v->size[0] * 0.5, hy = v->size[1] * 0.5;
    a = (DOT3(u, u) - (sww * DOT3(xs, xs) - xw * xw)) * inv_c2 * hx * hx;
    b = (DOT3(u, w) - (sww * DOT3(xs, ys) - xw * yw)) * inv_c2 * hx * hy;
    d = (DOT3(w, w) - (sww * DOT3(ys, ys) - yw * yw)) * inv_c2 * hy * hy;
    cx = eye_ndc[0] + DOT3(ws, u) / (c3[3] * qww);
    cy = eye_ndc[1] + DOT3(ws, w) / (c3[3] * qww);
#undef DOT3
    valid = a > 0 && d > 0 && a * d - b * b > 0;
  }
  if (!valid) {
    screen_bound(v, p, bd);
    return 1;
  }
  int ok = bound_from_cov(a, b, d, 1.0, p, bd, margin);
  bd->nx = cx;
  bd->ny = cy;
  return ok;
}

/* basisToQuat (math.tsl.ts:116-148): the rotation with columns x, y, z */
static void basis_to_quat(const double* x, const double* y, const double* z, double* q) {
  double four, inv;
  if (z[2] <= 0) {
    double dif10 = y[1] - x[0], omm22 = 1.0 - z[2];
    if (dif10 <= 0) {
      four = omm22 - dif10;
      inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = four * inv, q[1] = (x[1] + y[0]) * inv, q[2] = (x[2] + z[0]) * inv, q[3] = (y[2] - z[1]) * inv;
    } else {
      four = omm22 + dif10;
      inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = (x[1] + y[0]) * inv, q[1] = four * inv, q[2] = (y[2] + z[1]) * inv, q[3] = (z[0] - x[2]) * inv;
    }
  } else {
    double sum10 = y[1] + x[0], opm22 = 1.0 + z[2];
    if (sum10 <= 0) {
      four = opm22 - sum10;
      inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = (x[2] + z[0]) * inv, q[1] = (y[2] + z[1]) * inv, q[2] = four * inv, q[3] = (x[1] - y[0]) * inv;
    } else {
      four = opm22 + sum10;
      inv = 0.5 / sqrt(fmax(four, 1e-30));
      q[0] = (y[2] - z[1]) * inv, q[1] = (z[0] - x[2]) * inv, q[2] = (x[1] - y[0]) * inv, q[3] = four * inv;
    }
  }
}

/* symEigens (math.tsl.ts:46-113) in double: Jacobi rotations on the
 * largest off-diagonal entry (the first of (0,1), (0,2), (1,2) at a tie), at
 * most 16, until it is within 1e-6 of the trace; eigenvalues descending by
 * three conditional swaps, eigenvectors normalized, the third flipped to a
 * right-handed basis. E's row j is eigenvector j. */
static void sym_eigens(const double* c6, double* values, double vec[3][3]) {
  double A[3][3] = {{c6[0], c6[3], c6[4]}, {c6[3], c6[1], c6[5]}, {c6[4], c6[5], c6[2]}};
  double E[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const double eps = 1e-6 * (fabs(c6[0]) + fabs(c6[1]) + fabs(c6[2]));
  for (int it = 0; it < 16; it++) {
    int p = 0, q = 1;
    double mx = fabs(A[0][1]);
    if (fabs(A[0][2]) > mx) {
      mx = fabs(A[0][2]);
      q = 2;
    }
    if (fabs(A[1][2]) > mx) {
      mx = fabs(A[1][2]);
      p = 1;
      q = 2;
    }
    if (mx <= eps) break;
    const double phi = 0.5 * atan2(2.0 * A[p][q], A[q][q] - A[p][p]);
    const double c = cos(phi), s = sin(phi);
    for (int k = 0; k < 3; k++) {
      const double x = A[p][k], y = A[q][k];
      A[p][k] = c * x - s * y;
      A[q][k] = s * x + c * y;
    }
    for (int k = 0; k < 3; k++) {
      const double x = A[k][p], y = A[k][q];
      A[k][p] = c * x - s * y;
      A[k][q] = s * x + c * y;
    }
    A[p][q] = A[q][p] = 0.0;
    for (int k = 0; k < 3; k++) {
      const double x = E[p][k], y = E[q][k];
      E[p][k] = c * x - s * y;
      E[q][k] = s * x + c * y;
    }
  }
  double l[3] = {A[0][0], A[1][1], A[2][2]};
  for (int j = 0; j < 3; j++) {
    const double n = sqrt(E[j][0] * E[j][0] + E[j][1] * E[j][1] + E[j][2] * E[j][2]);
    for (int k = 0; k < 3; k++) vec[j][k] = E[j][k] / n;
  }
  const int pairs[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int k = 0; k < 3; k++) {
    const int a = pairs[k][0], b = pairs[k][1];
    if (l[a] < l[b]) {
      double t = l[a];
      l[a] = l[b];
      l[b] = t;
      for (int m = 0; m < 3; m++) {
        t = vec[a][m];
        vec[a][m] = vec[b][m];
        vec[b][m] = t;
      }
    }
  }
  const double x[3] = {vec[0][1] * vec[1][2] - vec[0][2] * vec[1][1], vec[0][2] * vec[1][0] - vec[0][0] * vec[1][2],
                       vec[0][0] * vec[1][1] - vec[0][1] * vec[1][0]};
  if (x[0] * vec[2][0] + x[1] * vec[2][1] + x[2] * vec[2][2] < 0)
    for (int m = 0; m < 3; m++) vec[2][m] = -vec[2][m];
  for (int j = 0; j < 3; j++) values[j] = l[j];
}

/* fromCov (splats.tsl.ts:189-198): scales and rotation (w >= 0) of a covariance */
void orc_from_cov(const double* c6, double* scales, double* quat) {
  double values[3], vec[3][3];
  sym_eigens(c6, values, vec);
  for (int k = 0; k < 3; k++) scales[k] = sqrt(fmax(values[k], 0.0));
  basis_to_quat(vec[0], vec[1], vec[2], quat);
  qnorm(quat);
  if (quat[3] < 0)
    for (int k = 0; k < 4; k++) quat[k] = -quat[k];
}

/* isUniformScale (splatTransform.tsl.ts:46-62) of a column-major 4x4's
 * upper-left 3x3: G = M^T M within 1e-6 s2 of s2 I (s2 = trace(G) / 3) and
 * det M > 0 */
static int is_uniform_scale(const double* M) {
  double G[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) G[a][b] = AT(M, 0, a) * AT(M, 0, b) + AT(M, 1, a) * AT(M, 1, b) + AT(M, 2, a) * AT(M, 2, b);
  const double s2 = (G[0][0] + G[1][1] + G[2][2]) / 3.0;
  if (!(s2 > 0)) return 0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++)
      if (fabs(G[a][b] - (a == b ? s2 : 0.0)) > 1e-6 * s2) return 0;
  return det3_of(M) > 0;
}

/* quadOffScreen (projections.tsl.ts:335-352) */
static int quad_off_screen(const double* P, const double* mu, const double* a0, const double* a1, double* margin) {
  double c[4], A0[4], A1[4];
  point_of(P, mu, 1.0, c);
  point_of(P, a0, 0.0, A0);
  point_of(P, a1, 0.0, A1);
  int front = 1, o[4] = {1, 1, 1, 1};
  for (int i = 0; i < 4; i++) {
    double s0 = (i & 1) ? 1.0 : -1.0, s1 = (i & 2) ? 1.0 : -1.0;
    double x = c[0] + s0 * A0[0] + s1 * A1[0], y = c[1] + s0 * A0[1] + s1 * A1[1], w = c[3] + s0 * A0[3] + s1 * A1[3];
    front = front && w > 0;
    o[0] = o[0] && x < -w;
    o[1] = o[1] && x > w;
    o[2] = o[2] && y < -w;
    o[3] = o[3] && y > w;
    *margin = fmin(*margin, fmin(fabs(fabs(x) - fabs(w)), fabs(fabs(y) - fabs(w))) / fmax(fabs(w), 1e-30));
  }
  return front && (o[0] || o[1] || o[2] || o[3]);
}

/* raygsFrame (projections.tsl.ts:354-396) and the eye words; fills the frame axes, c^2, lengths */
static int raygs_eye(const orc_eye* E, const double* c, const double* q, const double* s, double std_dev, uint32_t* w,
                     double axes[2][3], double* c2_out, double* lengths, double* margin) {
  double kappa = std_dev * std_dev;
  double qc[4] = {-q[0], -q[1], -q[2], q[3]}, lc[3];
  qrot(qc, c, lc);
  for (int k = 0; k < 3; k++) lc[k] /= s[k];
  double c2 = lc[0] * lc[0] + lc[1] * lc[1] + lc[2] * lc[2];
  int ok = 1;
  double mu[3];
  if (E->ortho) {
    double dz[3] = {0, 0, -1}, d[3];
    qrot(qc, dz, d);
    for (int k = 0; k < 3; k++) d[k] /= s[k];
    double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int k = 0; k < 3; k++) mu[k] = d[k] / n;
    c2 = 1e18;
  } else {
    ok = c2 > kappa * 1.0001;
    *margin = fmin(*margin, rel_margin(c2, kappa * 1.0001));
    for (int k = 0; k < 3; k++) mu[k] = lc[k] / sqrt(c2);
  }
  double t[3] = {mu[0], mu[1], mu[2] + (mu[2] >= 0 ? 1.0 : -1.0)};
  double inv_den = 2.0 / (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  double v0[3] = {s[0] * (t[0] * t[0] * inv_den - 1.0), s[1] * (t[1] * t[0] * inv_den), s[2] * (t[2] * t[0] * inv_den)};
  double v1[3] = {s[0] * (t[0] * t[1] * inv_den), s[1] * (t[1] * t[1] * inv_den - 1.0), s[2] * (t[2] * t[1] * inv_den)};
  double q0[3], q1[3];
  qrot(q, v0, q0);
  qrot(q, v1, q1);
  double bxx = q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2];
  double bxy = q0[0] * q1[0] + q0[1] * q1[1] + q0[2] * q1[2];
  double byy = q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2];
  double diff = bxx - byy;
  double eig1 = (bxx + byy) * 0.5 + sqrt(fmax(0.0, diff * diff * 0.25 + bxy * bxy));
  double vx, vy;
  if (fabs(bxx - eig1) >= fabs(byy - eig1)) {
    vx = bxy;
    vy = eig1 - bxx;
  } else {
    vx = eig1 - byy;
    vy = bxy;
  }
  double vl = sqrt(vx * vx + vy * vy);
  if (vl > 0) {
    vx /= vl;
    vy /= vl;
  } else {
    vx = 1;
    vy = 0;
  }
  double u0x = -vy, u0y = vx;
  double a0[3], a1[3];
  for (int k = 0; k < 3; k++) {
    a0[k] = q0[k] * u0x + q1[k] * u0y;
    a1[k] = q0[k] * vx + q1[k] * vy;
  }
  double lx = sqrt(a0[0] * a0[0] + a0[1] * a0[1] + a0[2] * a0[2]), ly = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
  ok = ok && lx > 0 && ly > 0;
  double zs = std_dev / sqrt(fmax(1.0 - kappa / c2, 1e-4));
  double za0[3] = {a0[0] * zs, a0[1] * zs, a0[2] * zs}, za1[3] = {a1[0] * zs, a1[1] * zs, a1[2] * zs};
  ok = ok && !quad_off_screen(E->projection, c, za0, za1, margin);
  memcpy(axes[0], a0, sizeof a0);
  memcpy(axes[1], a1, sizeof a1);
  *c2_out = c2;
  lengths[0] = lx;
  lengths[1] = ly;
  if (!ok) return 0;
  double n0[3] = {a0[0] / lx, a0[1] / lx, a0[2] / lx}, n1[3] = {a1[0] / ly, a1[1] / ly, a1[2] / ly};
  double n2[3] = {n0[1] * n1[2] - n0[2] * n1[1], n0[2] * n1[0] - n0[0] * n1[2], n0[0] * n1[1] - n0[1] * n1[0]};
  double fq[4];
  basis_to_quat(n0, n1, n2, fq);
  w[0] = orc_encode_quat_oct32_gpu(fq);
  w[1] = float_bits(c2);
  w[2] = orc_encode_half(log(lx)) | (orc_encode_half(log(ly)) << 16);
  return 1;
}

void orc_project_ray(int method, const orc_ray_splat* src, const orc_view* v, const orc_params* p, orc_ray_out* out) {
  memset(out, 0, sizeof *out);
  double margin = INFINITY;
  /* transformSplat with a = object to head (splatTransform.tsl.ts:154-165):
   * transformUniform (142-152) for a rotation times a uniform scale, else
   * fromCov of M Sigma M^T */
  const double* M = v->object_to_head;
  double center[4], qn[4] = {src->quat[0], src->quat[1], src->quat[2], src->quat[3]};
  point_of(M, src->center, 1.0, center);
  qnorm(qn);
  double quat[4], scales[3];
  if (is_uniform_scale(M)) {
    double oscale = cbrt(fabs(det3_of(M)));
    double tq[4];
    quat_of_matrix(M, oscale, tq);
    qmul(tq, qn, quat);
    for (int k = 0; k < 3; k++) scales[k] = src->scales[k] * oscale;
  } else {
    double cov[6], h[6];
    orc_cov_from_scale_quat(src->scales, qn, cov);
    const double S[3][3] = {{cov[0], cov[3], cov[4]}, {cov[3], cov[1], cov[5]}, {cov[4], cov[5], cov[2]}};
    double MS[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) MS[a][b] = AT(M, a, 0) * S[0][b] + AT(M, a, 1) * S[1][b] + AT(M, a, 2) * S[2][b];
    const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int e = 0; e < 6; e++) {
      const int a = ij[e][0], b = ij[e][1];
      h[e] = MS[a][0] * AT(M, b, 0) + MS[a][1] * AT(M, b, 1) + MS[a][2] * AT(M, b, 2);
    }
    orc_from_cov(h, scales, quat);
  }

  /* prepareSplat */
  double opacity = src->opacity, inflate = 1.0;
  if (opacity > 1.0) {
    opacity = fmin(opacity * 4.0 - 3.0, 5.0);
    if (p->lod_inflate) {
      inflate = pow(exp((opacity * opacity - 1.0) / M_E), 1.0 / 3.0);
      opacity = 1.0;
    }
  }
  for (int k = 0; k < 3; k++) scales[k] *= inflate;
  double std_dev = opacity > 1.0 ? p->max_std_dev + (opacity - 1.0) * 0.7 : p->max_std_dev;
  out->invisible = opacity <= 0.0 || opacity < p->min_alpha || (scales[0] == 0 && scales[1] == 0 && scales[2] == 0);
  margin = fmin(margin, rel_margin(opacity, p->min_alpha));
  out->distance = sqrt(center[0] * center[0] + center[1] * center[1] + center[2] * center[2]);

  /* headBasis: viewBasis in the head frame, then mipInflate with the head view */
  double hq[4] = {quat[0], quat[1], quat[2], quat[3]};
  qnorm(hq);
  double S[3] = {scales[0], scales[1], scales[2]};
  double smax = fmax(S[0], fmax(S[1], S[2])) * 1e-4;
  for (int k = 0; k < 3; k++) S[k] = fmax(S[k], smax);
  const orc_eye* E0 = &v->eye[0];
  {
    double conj[4] = {-hq[0], -hq[1], -hq[2], hq[3]}, local[3];
    qrot(conj, center, local);
    double l2[3] = {local[0] * local[0], local[1] * local[1], local[2] * local[2]};
    double pixel = 2.0 / (v->size[1] * AT(E0->projection, 1, 1));
    double fp = pixel * pixel * (E0->ortho ? 1.0 : out->distance * out->distance);
    /* the head's blur, widened by depth of field (splatInView with the head frame) */
    double blur = p->blur_amount;
    if (p->focal_distance > 0 && p->aperture_tan > 0 && center[2] < 0) {
      double depth = -center[2];
      double coc_view = fabs(depth - p->focal_distance) * p->aperture_tan;
      double focal_pixels = v->size[1] * 0.5 * AT(E0->projection, 1, 1);
      double coc = E0->ortho ? coc_view * focal_pixels : coc_view * focal_pixels / depth;
      blur = clampd(coc * coc, blur, p->max_pixel_radius * p->max_pixel_radius);
    }
    double s2[3], sb[3];
    for (int k = 0; k < 3; k++) {
      s2[k] = S[k] * S[k] + p->pre_blur_amount * fp;
      sb[k] = s2[k] + blur * fp;
    }
    double num = l2[0] * s2[1] * s2[2] + l2[1] * s2[0] * s2[2] + l2[2] * s2[0] * s2[1];
    double den = l2[0] * sb[1] * sb[2] + l2[1] * sb[0] * sb[2] + l2[2] * sb[0] * sb[1];
    out->head_opacity = opacity * (den > 0 ? sqrt(fmax(num / den, 0.0)) : 1.0);
    for (int k = 0; k < 3; k++) S[k] = sqrt(sb[k]);
  }
  for (int k = 0; k < 3; k++) out->head_scales[k] = S[k];
  int ok = !out->invisible && out->head_opacity >= p->min_alpha;
  margin = fmin(margin, rel_margin(out->head_opacity, p->min_alpha));

  const int W = 6 + 2 * v->eyes;
  for (int e = 0; e < v->eyes; e++) {
    const orc_eye* E = &v->eye[e];
    double c[4], clip[4];
    point_of(E->from_head, center, 1.0, c);
    point_of(E->projection, c, 1.0, clip);
    double em = INFINITY;
    int cull_depth = c[2] >= 0.0 || clip[2] < v->ndc_z_min * clip[3] || clip[2] >= clip[3];
    em = fmin(em, fabs(c[2]) / fmax(sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]), 1e-300));
    em = fmin(em, rel_margin(clip[2], v->ndc_z_min * clip[3]));
    em = fmin(em, rel_margin(clip[2], clip[3]));
    if (ok && !cull_depth) {
      /* frameQuat: the eye's rotation of the head-space quaternion */
      double fq[4], eq[4];
      quat_of_matrix(E->from_head, 1.0, fq);
      qmul(fq, hq, eq);
      qnorm(eq);
      if (method == 3) {
        if (raygs_eye(E, c, eq, S, std_dev, out->words + 4 + 3 * e, out->frame_axes[e], &out->frame_c2[e],
                      out->frame_lengths[e], &em)) {
          out->eye_visible[e] = 1;
          out->visible = 1;
        }
        out->eye_margin[e] = em;
        margin = fmin(margin, em);
        continue;
      }
      bound_t bd;
      int bok;
      if (method == 1) {
        bok = gut_bound(E, v, c, eq, S, std_dev, p, &bd, &em);
      } else {
        double ndc[2] = {0, 0};
        if (clip[3] > 0) {
          ndc[0] = clip[0] / clip[3];
          ndc[1] = clip[1] / clip[3];
        }
        bok = htgs_bound(E, v, c, eq, S, std_dev, ndc, p, &bd, &em);
      }
      if (bok && clip_to_screen(v, &bd, &em)) {
        uint32_t angle = (uint32_t)floor(bd.angle / M_PI * 1024.0 + 0.5) & 1023u;
        double slack = fmax(bd.s1, bd.s2) * sin(M_PI / 1024.0 / 2.0) * 1.001;
        out->words[6 + 2 * e] = fixed16(bd.nx) | (fixed16(bd.ny) << 16);
        out->words[7 + 2 * e] = angle | (log_scale(bd.s1 + slack, 11, 1) << 10) | (log_scale(bd.s2 + slack, 11, 1) << 21);
        out->bound[e][0] = bd.nx;
        out->bound[e][1] = bd.ny;
        out->bound[e][2] = bd.angle;
        out->bound[e][3] = bd.s1;
        out->bound[e][4] = bd.s2;
        out->eye_visible[e] = 1;
        out->visible = 1;
      }
    }
    out->eye_margin[e] = em;
    margin = fmin(margin, em);
  }
  if (ok) {
    uint32_t dir, dist, low;
    orc_encode_center(center, &dir, &dist, &low);
    double stored = out->head_opacity > 1.0 ? (out->head_opacity + 3.0) * 0.25 : out->head_opacity;
    out->words[0] = dir;
    out->words[1] = dist;
    out->words[2] = orc_encode_ext_rgb(src->rgb[0], src->rgb[1], src->rgb[2]);
    if (method == 3) {
      out->words[3] = opacity8(stored) | low;
    } else {
      out->words[3] = orc_encode_quat_oct32_gpu(hq);
      out->words[4] = orc_encode_half(log(S[0])) | (orc_encode_half(log(S[1])) << 16);
      out->words[5] = orc_encode_half(log(S[2])) | (opacity8(stored) << 16) | low;
    }
  }
  (void)W;

  /* the key (SplatProjection.ts:141-151) with the splat-form sortGuard */
  out->guarded = 1;
  if (p->sort_guard > 0) {
    double radius = fmax(scales[0], fmax(scales[1], scales[2])) * p->max_std_dev;
    double hclip[4];
    point_of(E0->projection, center, 1.0, hclip);
    double w = fmax(hclip[3], 1e-6);
    double rx = AT(E0->projection, 0, 0) * radius / w, ry = AT(E0->projection, 1, 1) * radius / w;
    int inside = fabs(hclip[0] / w) - rx <= p->sort_guard + 1.0 && fabs(hclip[1] / w) - ry <= p->sort_guard + 1.0;
    int near = -center[2] <= radius;
    out->guarded = center[2] < radius && (inside || near);
    if (!p->cull_keys) {
      margin = fmin(margin, rel_margin(fabs(hclip[0] / w) - rx, p->sort_guard + 1.0));
      margin = fmin(margin, rel_margin(fabs(hclip[1] / w) - ry, p->sort_guard + 1.0));
      margin = fmin(margin, rel_margin(-center[2], radius));
    }
  }
  int keyed = p->cull_keys ? out->visible : (!out->invisible && out->guarded);
  out->key = keyed ? orc_key_compute(out->distance, p->sort_bits) : orc_key_invalid(p->sort_bits);
  out->margin = margin;
}

static const int RAY_SH_COEFFS[4] = {0, 3, 8, 15};

void orc_project_ray_table_f32(int method, const orc_table_f32* t, const orc_view* v, const orc_params* p, uint32_t* words,
                               uint32_t* keys, double* margin, double* eye_margin) {
  const uint32_t n = t->count;
  const size_t st = t->stride ? t->stride : n;
  const int W = method == 3 ? 4 + 3 * v->eyes : 6 + 2 * v->eyes;
  for (uint32_t i = 0; i < n; i++) {
    orc_ray_splat s;
    double coeffs[45];
    for (int k = 0; k < 3; k++) {
      s.center[k] = t->center[k * st + i];
      s.scales[k] = exp((double)t->log_scale[k * st + i]);
      s.rgb[k] = t->rgb[k * st + i];
    }
    for (int k = 0; k < 4; k++) s.quat[k] = t->quat[k * st + i];
    s.opacity = t->opacity[i];
    if (t->sh_degree > 0 && t->sh) {
      int deg = t->sh_degree < p->max_sh_degree ? t->sh_degree : p->max_sh_degree;
      if (deg > 0) {
        for (int k = 0; k < 3 * RAY_SH_COEFFS[t->sh_degree]; k++) coeffs[k] = t->sh[k * st + i];
        double dir[3] = {s.center[0] - v->view_origin[0], s.center[1] - v->view_origin[1], s.center[2] - v->view_origin[2]};
        double len = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int k = 0; k < 3; k++) dir[k] /= len;
        double sh[3];
        orc_sh_evaluate(coeffs, dir, deg, sh);
        for (int k = 0; k < 3; k++) s.rgb[k] = fmax(s.rgb[k] + sh[k], 0.0);
      }
    }
    orc_ray_out o;
    orc_project_ray(method, &s, v, p, &o);
    memcpy(words + (size_t)W * i, o.words, W * sizeof(uint32_t));
    if (keys) keys[i] = o.key;
    if (margin) margin[i] = o.margin;
    if (eye_margin)
      for (int e = 0; e < v->eyes; e++) eye_margin[(size_t)e * n + i] = o.eye_margin[e];
  }
}

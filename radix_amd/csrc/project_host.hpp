// Host side of the projection: the camera and object transforms turned into
// the kernel's uniforms (in double), and the launch of one table.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>

#include "api_internal.hpp"
#include "project_kernels.hpp"

namespace rdx {

// Column-major 4x4 (three.js Matrix4.elements): element (r, c) at m[4c + r].
inline double& cm(double* m, int r, int c) { return m[4 * c + r]; }
inline double cm(const double* m, int r, int c) { return m[4 * c + r]; }

inline void mul(const double* a, const double* b, double* out) {
  double r[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += cm(a, i, k) * cm(b, k, j);
      cm(r, i, j) = s;
    }
  memcpy(out, r, sizeof r);
}

// Inverse by Gauss-Jordan elimination with partial pivoting.
inline bool invert(const double* m, double* out) {
  double a[4][8];
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) {
      a[r][c] = cm(m, r, c);
      a[r][4 + c] = r == c ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; c++) {
    int piv = c;
    for (int r = c + 1; r < 4; r++)
      if (fabs(a[r][c]) > fabs(a[piv][c])) piv = r;
    if (a[piv][c] == 0.0) return false;
    if (piv != c)
      for (int k = 0; k < 8; k++) std::swap(a[c][k], a[piv][k]);
    const double inv = 1.0 / a[c][c];
    for (int k = 0; k < 8; k++) a[c][k] *= inv;
    for (int r = 0; r < 4; r++) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f == 0.0) continue;
      for (int k = 0; k < 8; k++) a[r][k] -= f * a[c][k];
    }
  }
  for (int r = 0; r < 4; r++)
    for (int c = 0; c < 4; c++) cm(out, r, c) = a[r][4 + c];
  return true;
}

inline double det3(const double* L) {
  return L[0] * (L[4] * L[8] - L[5] * L[7]) - L[1] * (L[3] * L[8] - L[5] * L[6]) + L[2] * (L[3] * L[7] - L[4] * L[6]);
}

// The quaternion (x, y, z, w) of a row-major rotation (Shepperd: the
// largest of the four squared components first).
inline void quat_of_rotation(const double* R, float* q) {
  const double tr = R[0] + R[4] + R[8];
  double x, y, z, w;
  if (tr > 0.0) {
    const double s = 0.5 / std::sqrt(tr + 1.0);
    w = 0.25 / s;
    x = (R[7] - R[5]) * s;
    y = (R[2] - R[6]) * s;
    z = (R[3] - R[1]) * s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    const double s = 2.0 * std::sqrt(1.0 + R[0] - R[4] - R[8]);
    w = (R[7] - R[5]) / s;
    x = 0.25 * s;
    y = (R[1] + R[3]) / s;
    z = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    const double s = 2.0 * std::sqrt(1.0 + R[4] - R[0] - R[8]);
    w = (R[2] - R[6]) / s;
    x = (R[1] + R[3]) / s;
    y = 0.25 * s;
    z = (R[5] + R[7]) / s;
  } else {
    const double s = 2.0 * std::sqrt(1.0 + R[8] - R[0] - R[4]);
    w = (R[3] - R[1]) / s;
    x = (R[2] + R[6]) / s;
    y = (R[5] + R[7]) / s;
    z = 0.25 * s;
  }
  const double n = std::sqrt(x * x + y * y + z * z + w * w);
  q[0] = (float)(x / n);
  q[1] = (float)(y / n);
  q[2] = (float)(z / n);
  q[3] = (float)(w / n);
}

// Uniforms of one table seen by one camera (setViews, splatCamera.ts:119-138;
// setTransforms, splatTransform.tsl.ts:78-107; the generator context,
// splatGenerator.ts:96-103).
int make_uniforms(const rdx_splat_table* t, const rdx_camera* cam, const rdx_params* prm, ProjU* u);

// Kernel pointers of a table.
TableP make_table(const rdx_splat_table* t);

// A table descriptor's checks (layout, pointers, SH codec, sizes).
int check_table(const rdx_splat_table* t);

// Launch the projection of a table (one thread per splat, grid-stride
// beyond 2^20 blocks).
int launch_project(hipStream_t s, const rdx_splat_table* t, const ProjU& u, const OutP& o, bool side);

// The kernels, instantiated per layout in their own translation units.
void launch_project_f32(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int eyes, bool side);
void launch_project_ext(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                        bool side);
void launch_project_packed(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                           bool side);
void launch_project_cov(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                        bool side);

}  // namespace rdx

// The ray projection methods' entry point: uniforms (the object transform
// split into a uniform scale and rotations, splatTransform.tsl.ts:46-74,
// 142-152) and the kernels per layout, SH degree and eye count.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "project_host.hpp"
#include "ray_launch.hpp"

namespace rdx {

static double det3r(const double* M) {
  return M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) + M[2] * (M[3] * M[7] - M[4] * M[6]);
}

// isUniformScale (splatTransform.tsl.ts:46-62) of a row-major 3x3: M^T M
// within 1e-6 of s^2 I, and det > 0. Returns s, or 0.
static double uniform_scale(const double* M) {
  double G[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) G[3 * i + j] = M[i] * M[j] + M[3 + i] * M[3 + j] + M[6 + i] * M[6 + j];
  const double s2 = (G[0] + G[4] + G[8]) / 3.0;
  if (!(s2 > 0.0)) return 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = i; j < 3; j++)
      if (std::fabs(G[3 * i + j] - (i == j ? s2 : 0.0)) > 1e-6 * s2) return 0.0;
  const double d = det3r(M);
  return d > 0.0 ? std::cbrt(d) : 0.0;
}

}  // namespace rdx

using namespace rdx;

extern "C" {

int rdx_project_ray(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                    const rdx_params* params, int method, uint32_t at, uint32_t* const* channels_dev, uint32_t* keys_dev) {
  if (!camera || !params || !channels_dev) return fail(RDX_E_INVALID, "rdx_project_ray: NULL camera, params or channels");
  if (method != RDX_METHOD_3DGUT && method != RDX_METHOD_HTGS && method != RDX_METHOD_RAYGS)
    return fail(method == RDX_METHOD_EWA ? RDX_E_UNSUPPORTED : RDX_E_INVALID,
                "rdx_project_ray: method %d is not 3dgut (1), htgs (2) or raygs (3) (ewa: rdx_project_ewa)", method);
  if (int rc = check_table(table)) return rc;
  if (int rc = check_bits(params->sort_bits, "SplatProjection")) return rc;
  RayU r;
  memset(&r, 0, sizeof r);
  if (int rc = make_uniforms(table, camera, params, &r.u)) return rc;
  const int channels = camera->eyes == 2 ? 3 : 2;
  for (int c = 0; c < channels; c++)
    if (!channels_dev[c]) return fail(RDX_E_INVALID, "rdx_project_ray: channel w%d is NULL", c);
  // The object transform: a uniform scale times a rotation turns and scales
  // the splat (transformUniform, splatTransform.tsl.ts:142-152); any other
  // goes through the covariance (fromCov of H Σ H^T, :154-165), after which
  // each eye turns the head-space splat by its head -> eye rotation.
  const double s = uniform_scale(r.u.H);
  double R[9];
  if (s > 0.0) {
    r.oscale = (float)s;
    for (int k = 0; k < 9; k++) R[k] = r.u.H[k] / s;
    quat_of_rotation(R, r.hq);
    for (int e = 0; e < camera->eyes; e++) {
      for (int k = 0; k < 9; k++) R[k] = r.u.eye[e].L[k] / s;
      quat_of_rotation(R, r.eq[e]);
    }
  } else {
    const double d = det3r(r.u.H);
    if (!(std::fabs(d) > 0.0)) return fail(RDX_E_INVALID, "rdx_project_ray: the object transform is singular");
    r.nonuniform = 1;
    r.oscale = 1.0f;
    // H^-1 (row-major) by cofactors; each eye's head -> eye rotation L_e H^-1
    const double* H = r.u.H;
    const double inv[9] = {(H[4] * H[8] - H[5] * H[7]) / d, (H[2] * H[7] - H[1] * H[8]) / d,
                           (H[1] * H[5] - H[2] * H[4]) / d, (H[5] * H[6] - H[3] * H[8]) / d,
                           (H[0] * H[8] - H[2] * H[6]) / d, (H[2] * H[3] - H[0] * H[5]) / d,
                           (H[3] * H[7] - H[4] * H[6]) / d, (H[1] * H[6] - H[0] * H[7]) / d,
                           (H[0] * H[4] - H[1] * H[3]) / d};
    for (int e = 0; e < camera->eyes; e++) {
      const double* L = r.u.eye[e].L;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) R[3 * a + b] = L[3 * a] * inv[b] + L[3 * a + 1] * inv[3 + b] + L[3 * a + 2] * inv[6 + b];
      quat_of_rotation(R, r.ehq[e]);
    }
  }
  const uint32_t n = table->count;
  if (n == 0) return RDX_OK;
  RayOut o;
  memset(&o, 0, sizeof o);
  for (int c = 0; c < channels; c++) o.ch[c] = channels_dev[c];
  o.keys = keys_dev;
  o.at = at;
  const TableP tp = make_table(table);
  const dim3 grid((uint32_t)(((uint64_t)n + 255) / 256));
  hipStream_t st = (hipStream_t)stream;
  Stage stage(reinterpret_cast<Context*>(ctx), st,
              method == RDX_METHOD_3DGUT ? "project_3dgut" : (method == RDX_METHOD_HTGS ? "project_htgs" : "project_raygs"));
  if (r.nonuniform)
    launch_ray_nonuniform(method, camera->eyes, st, grid, table, tp, r, o);
  else
    launch_ray_uniform(method, camera->eyes, st, grid, table, tp, r, o);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_project_ray");
}

}  // extern "C"

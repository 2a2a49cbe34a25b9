// The ray projection kernels' launches per layout, SH degree, eye count and
// object transform kind (uniform or not); the two kinds are instantiated in
// their own translation units (ray_uniform.hip, ray_nonuniform.hip).
#pragma once

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "project_host.hpp"
#include "ray_kernels.hpp"

namespace rdx {

template <int METHOD, int EYES, bool NU>
static void launch_layout(hipStream_t st, dim3 grid, const rdx_splat_table* t, const TableP& tp, const RayU& r,
                          const RayOut& o) {
  const int shc = t->sh_codec == RDX_SH_PACKED ? SHC_PACKED : SHC_EXT;
  if (t->layout == RDX_LAYOUT_F32_SOA) {
    switch (r.u.sh_degree) {
      case 0: hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_F32, SHC_F32, 0, EYES, NU>), grid, dim3(256), 0, st, tp, r, o); break;
      case 1: hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_F32, SHC_F32, 1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o); break;
      case 2:
        if (f32_sh_wide(tp.stride, 2))
          hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_F32, SHC_F32_WIDE, 2, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
        else
          hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_F32, SHC_F32, 2, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
        break;
      default:
        if (f32_sh_wide(tp.stride, 3))
          hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_F32, SHC_F32_WIDE, 3, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
        else
          hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_F32, SHC_F32, 3, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
        break;
    }
  } else if (t->layout == RDX_LAYOUT_EXT) {
    if (shc == SHC_PACKED)
      hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_EXT, SHC_PACKED, -1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
    else
      hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_EXT, SHC_EXT, -1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
  } else if (t->layout == RDX_LAYOUT_COV) {
    if (shc == SHC_PACKED)
      hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_COV, SHC_PACKED, -1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
    else
      hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_COV, SHC_EXT, -1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
  } else {
    if (shc == SHC_PACKED)
      hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_PACKED, SHC_PACKED, -1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
    else
      hipLaunchKernelGGL((k_project_ray<METHOD, LAYOUT_PACKED, SHC_EXT, -1, EYES, NU>), grid, dim3(256), 0, st, tp, r, o);
  }
}

template <bool NU>
static void launch_ray(int method, int eyes, hipStream_t st, dim3 grid, const rdx_splat_table* t, const TableP& tp,
                       const RayU& r, const RayOut& o) {
  if (method == RDX_METHOD_3DGUT) {
    if (eyes == 2) launch_layout<METHOD_3DGUT, 2, NU>(st, grid, t, tp, r, o);
    else launch_layout<METHOD_3DGUT, 1, NU>(st, grid, t, tp, r, o);
  } else if (method == RDX_METHOD_HTGS) {
    if (eyes == 2) launch_layout<METHOD_HTGS, 2, NU>(st, grid, t, tp, r, o);
    else launch_layout<METHOD_HTGS, 1, NU>(st, grid, t, tp, r, o);
  } else {
    if (eyes == 2) launch_layout<METHOD_RAYGS, 2, NU>(st, grid, t, tp, r, o);
    else launch_layout<METHOD_RAYGS, 1, NU>(st, grid, t, tp, r, o);
  }
}

void launch_ray_uniform(int method, int eyes, hipStream_t st, dim3 grid, const rdx_splat_table* t, const TableP& tp,
                        const RayU& r, const RayOut& o);
void launch_ray_nonuniform(int method, int eyes, hipStream_t st, dim3 grid, const rdx_splat_table* t, const TableP& tp,
                           const RayU& r, const RayOut& o);

}  // namespace rdx

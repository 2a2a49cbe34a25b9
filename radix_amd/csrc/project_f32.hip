// Projection kernels for f32 structure-of-arrays tables, one per SH degree.
#include "project_host.hpp"

namespace rdx {

template <int DEG, int EYES>
static void pick(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, bool side, bool count) {
  if (side)
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC_F32, DEG, EYES, true, false>), grid, dim3(256), 0, s, t, u, o);
  else if (count)
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC_F32, DEG, EYES, false, true>), grid, dim3(256), 0, s, t, u, o);
  else
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC_F32, DEG, EYES, false, false>), grid, dim3(256), 0, s, t, u, o);
}

template <int DEG>
static void eyes(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int n, bool side, bool count) {
  if (n == 1) pick<DEG, 1>(s, grid, t, u, o, side, count);
  else pick<DEG, 2>(s, grid, t, u, o, side, count);
}

void launch_project_f32(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int n_eyes, bool side,
                        bool count) {
  switch (u.sh_degree) {
    case 0: eyes<0>(s, grid, t, u, o, n_eyes, side, count); break;
    case 1: eyes<1>(s, grid, t, u, o, n_eyes, side, count); break;
    case 2: eyes<2>(s, grid, t, u, o, n_eyes, side, count); break;
    default: eyes<3>(s, grid, t, u, o, n_eyes, side, count); break;
  }
}

}  // namespace rdx

// Projection kernels for f32 structure-of-arrays tables, one per SH degree.
#include "project_host.hpp"

namespace rdx {

template <int DEG, int EYES>
static void pick(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, bool side) {
  if (side)
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC_F32, DEG, EYES, true>), grid, dim3(256), 0, s, t, u, o);
  else
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC_F32, DEG, EYES, false>), grid, dim3(256), 0, s, t, u, o);
}

template <int DEG>
static void eyes(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int n, bool side) {
  if (n == 1) pick<DEG, 1>(s, grid, t, u, o, side);
  else pick<DEG, 2>(s, grid, t, u, o, side);
}

void launch_project_f32(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int n_eyes, bool side) {
  switch (u.sh_degree) {
    case 0: eyes<0>(s, grid, t, u, o, n_eyes, side); break;
    case 1: eyes<1>(s, grid, t, u, o, n_eyes, side); break;
    case 2: eyes<2>(s, grid, t, u, o, n_eyes, side); break;
    default: eyes<3>(s, grid, t, u, o, n_eyes, side); break;
  }
}

}  // namespace rdx

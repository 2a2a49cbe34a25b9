// Projection kernels for f32 structure-of-arrays tables, one per SH degree.
#include "project_host.hpp"

namespace rdx {

template <int SHC, int DEG, int EYES>
static void pick_shc(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, bool side) {
  if (side)
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC, DEG, EYES, true>), grid, dim3(256), 0, s, t, u, o);
  else
    hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC, DEG, EYES, false>), grid, dim3(256), 0, s, t, u, o);
}

// Tables whose SH planes exceed one 4 GiB window (degrees 2 and 3 only, at
// the plane stride limit) take the grouped SH reads.
template <int DEG, int EYES>
static void pick(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, bool side) {
  if constexpr (DEG >= 2) {
    if (f32_sh_wide(t.stride, DEG)) {
      pick_shc<SHC_F32_WIDE, DEG, EYES>(s, grid, t, u, o, side);
      return;
    }
  }
  pick_shc<SHC_F32, DEG, EYES>(s, grid, t, u, o, side);
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

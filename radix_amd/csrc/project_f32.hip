// Projection kernels for f32 structure-of-arrays tables.
#include "project_host.hpp"

namespace rdx {

template <int EYES, bool SIDE, bool HIST>
static void go(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o) {
  hipLaunchKernelGGL((k_project<LAYOUT_F32, SHC_F32, EYES, SIDE, HIST>), grid, dim3(256), 0, s, t, u, o);
}

void launch_project_f32(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int eyes, bool side,
                        bool hist) {
  if (eyes == 1) {
    if (side) hist ? go<1, true, true>(s, grid, t, u, o) : go<1, true, false>(s, grid, t, u, o);
    else hist ? go<1, false, true>(s, grid, t, u, o) : go<1, false, false>(s, grid, t, u, o);
  } else {
    if (side) hist ? go<2, true, true>(s, grid, t, u, o) : go<2, true, false>(s, grid, t, u, o);
    else hist ? go<2, false, true>(s, grid, t, u, o) : go<2, false, false>(s, grid, t, u, o);
  }
}

}  // namespace rdx

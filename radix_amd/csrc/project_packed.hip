// Projection kernels for packedSplats tables (with ext or packed SH bundles).
#include "project_host.hpp"

namespace rdx {

template <int SHC, int EYES, bool SIDE, bool HIST>
static void go(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o) {
  hipLaunchKernelGGL((k_project<LAYOUT_PACKED, SHC, EYES, SIDE, HIST>), grid, dim3(256), 0, s, t, u, o);
}

template <int SHC>
static void pick(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int eyes, bool side,
                 bool hist) {
  if (eyes == 1) {
    if (side) hist ? go<SHC, 1, true, true>(s, grid, t, u, o) : go<SHC, 1, true, false>(s, grid, t, u, o);
    else hist ? go<SHC, 1, false, true>(s, grid, t, u, o) : go<SHC, 1, false, false>(s, grid, t, u, o);
  } else {
    if (side) hist ? go<SHC, 2, true, true>(s, grid, t, u, o) : go<SHC, 2, true, false>(s, grid, t, u, o);
    else hist ? go<SHC, 2, false, true>(s, grid, t, u, o) : go<SHC, 2, false, false>(s, grid, t, u, o);
  }
}

void launch_project_packed(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                         bool side, bool hist) {
  if (shc == SHC_PACKED) pick<SHC_PACKED>(s, grid, t, u, o, eyes, side, hist);
  else pick<SHC_EXT>(s, grid, t, u, o, eyes, side, hist);
}

}  // namespace rdx

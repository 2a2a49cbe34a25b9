// Projection kernels for extSplats and covSplats tables (with ext or packed
// SH bundles): two uint32x4 words per splat.
#include "project_host.hpp"

namespace rdx {

template <int LAYOUT, int SHC, int EYES>
static void pick(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, bool side) {
  if (side)
    hipLaunchKernelGGL((k_project<LAYOUT, SHC, -1, EYES, true>), grid, dim3(256), 0, s, t, u, o);
  else if (u.sh_degree == 3)  // the common degree at a compile-time constant: no degree branches
    hipLaunchKernelGGL((k_project<LAYOUT, SHC, 3, EYES, false>), grid, dim3(256), 0, s, t, u, o);
  else
    hipLaunchKernelGGL((k_project<LAYOUT, SHC, -1, EYES, false>), grid, dim3(256), 0, s, t, u, o);
}

template <int LAYOUT>
static void launch_words2(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                          bool side) {
  if (shc == SHC_PACKED) {
    if (eyes == 1) pick<LAYOUT, SHC_PACKED, 1>(s, grid, t, u, o, side);
    else pick<LAYOUT, SHC_PACKED, 2>(s, grid, t, u, o, side);
  } else {
    if (eyes == 1) pick<LAYOUT, SHC_EXT, 1>(s, grid, t, u, o, side);
    else pick<LAYOUT, SHC_EXT, 2>(s, grid, t, u, o, side);
  }
}

void launch_project_ext(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                         bool side) {
  launch_words2<LAYOUT_EXT>(s, grid, shc, t, u, o, eyes, side);
}

void launch_project_cov(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                        bool side) {
  launch_words2<LAYOUT_COV>(s, grid, shc, t, u, o, eyes, side);
}

}  // namespace rdx

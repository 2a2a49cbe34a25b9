// Ray projection kernels for any other object transform (fromCov of H Σ H^T).
#include "ray_launch.hpp"

namespace rdx {

void launch_ray_nonuniform(int method, int eyes, hipStream_t st, dim3 grid, const rdx_splat_table* t, const TableP& tp,
                           const RayU& r, const RayOut& o) {
  launch_ray<true>(method, eyes, st, grid, t, tp, r, o);
}

}  // namespace rdx

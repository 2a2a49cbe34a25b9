// Ray projection kernels for a rotation times a uniform scale.
#include "ray_launch.hpp"

namespace rdx {

void launch_ray_uniform(int method, int eyes, hipStream_t st, dim3 grid, const rdx_splat_table* t, const TableP& tp,
                        const RayU& r, const RayOut& o) {
  launch_ray<false>(method, eyes, st, grid, t, tp, r, o);
}

}  // namespace rdx

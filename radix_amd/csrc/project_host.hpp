// Host side of the projection: the camera and object transforms turned into
// the kernel's uniforms (in double), and the launch of one table.
#pragma once

#include "api_internal.hpp"
#include "project_kernels.hpp"

namespace rdx {

// Uniforms of one table seen by one camera (setViews, splatCamera.ts:119-138;
// setTransforms, splatTransform.tsl.ts:78-107; the generator context,
// splatGenerator.ts:96-103).
int make_uniforms(const rdx_splat_table* t, const rdx_camera* cam, const rdx_params* prm, ProjU* u);

// Kernel pointers of a table.
TableP make_table(const rdx_splat_table* t);

// A table descriptor's checks (layout, pointers, SH codec, sizes).
int check_table(const rdx_splat_table* t);

// Launch the projection of a table (one thread per splat, grid-stride
// beyond 2^20 blocks).
int launch_project(hipStream_t s, const rdx_splat_table* t, const ProjU& u, const OutP& o, bool side);

// The kernels, instantiated per layout in their own translation units.
void launch_project_f32(hipStream_t s, dim3 grid, const TableP& t, const ProjU& u, const OutP& o, int eyes, bool side);
void launch_project_ext(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                        bool side);
void launch_project_packed(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                           bool side);
void launch_project_cov(hipStream_t s, dim3 grid, int shc, const TableP& t, const ProjU& u, const OutP& o, int eyes,
                        bool side);

}  // namespace rdx

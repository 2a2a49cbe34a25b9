// File decode onto the device: the byte transpose of RDX payloads (an RDX
// chunk stores each channel's elements byte-transposed then gzipped,
// src/core/rdx.ts:8-17; Compression.decompressTransposed undoes the
// transpose after the gunzip, src/core/compress.ts:6-19, 197-204). The host
// gunzips; this puts the planes back into elements in device memory, the
// channel buffers a splat table reads.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "common.hpp"

namespace rdx {

// Four elements of EB bytes per thread (EB a multiple of 4, planes 4-byte
// aligned: count % 4 == 0): word j of plane k holds byte k of elements 4j ..
// 4j + 3; element 4j + e's word w gathers byte e of planes 4w .. 4w + 3.
template <int EB>
static __global__ __launch_bounds__(256) void k_untranspose4(const uint32_t* __restrict__ planes, uint32_t quads,
                                                             uint32_t* __restrict__ out) {
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= quads) return;
  uint32_t p[EB];
#pragma unroll
  for (int k = 0; k < EB; k++) p[k] = __builtin_nontemporal_load(planes + (size_t)k * quads + j);
#pragma unroll
  for (int e = 0; e < 4; e++) {
    uint32_t v[EB / 4];
#pragma unroll
    for (int w = 0; w < EB / 4; w++) {
      const uint32_t b0 = (p[4 * w] >> (8 * e)) & 0xffu, b1 = (p[4 * w + 1] >> (8 * e)) & 0xffu;
      const uint32_t b2 = (p[4 * w + 2] >> (8 * e)) & 0xffu, b3 = (p[4 * w + 3] >> (8 * e)) & 0xffu;
      v[w] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
    }
    uint32_t* dst = out + ((size_t)4 * j + e) * (EB / 4);
    if constexpr (EB == 16) {
      *reinterpret_cast<uint4*>(dst) = make_uint4(v[0], v[1], v[2], v[3]);
    } else if constexpr (EB == 8) {
      *reinterpret_cast<uint2*>(dst) = make_uint2(v[0], v[1]);
    } else {
#pragma unroll
      for (int w = 0; w < EB / 4; w++) dst[w] = v[w];
    }
  }
}

// Any element size and count: a thread per element, byte by byte.
static __global__ __launch_bounds__(256) void k_untranspose_bytes(const uint8_t* __restrict__ planes, uint32_t count,
                                                                  uint32_t eb, uint8_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  for (uint32_t k = 0; k < eb; k++) out[(size_t)i * eb + k] = planes[(size_t)k * count + i];
}

}  // namespace rdx

using namespace rdx;

extern "C" {

int rdx_untranspose_bytes(void* stream, const uint8_t* planes_dev, uint32_t count, uint32_t element_bytes, void* out_dev) {
  if (element_bytes == 0 || element_bytes > 64) return fail(RDX_E_INVALID, "rdx_untranspose_bytes: element size %u", element_bytes);
  if (count == 0) return RDX_OK;
  if (!planes_dev || !out_dev) return fail(RDX_E_INVALID, "rdx_untranspose_bytes: NULL planes or output");
  hipStream_t s = (hipStream_t)stream;
  // the vector path: whole 32-bit plane words, and outputs aligned to an element
  const bool aligned = count % 4 == 0 && ((uintptr_t)planes_dev & 3) == 0 &&
                       ((uintptr_t)out_dev & (element_bytes >= 16 ? 15 : element_bytes - 1)) == 0;
  const uint32_t quads = count / 4;
  const dim3 g4((quads + 255) / 256), g1((count + 255) / 256);
  if (aligned && element_bytes == 16) {
    hipLaunchKernelGGL(k_untranspose4<16>, g4, dim3(256), 0, s, (const uint32_t*)planes_dev, quads, (uint32_t*)out_dev);
  } else if (aligned && element_bytes == 8) {
    hipLaunchKernelGGL(k_untranspose4<8>, g4, dim3(256), 0, s, (const uint32_t*)planes_dev, quads, (uint32_t*)out_dev);
  } else if (aligned && element_bytes == 4) {
    hipLaunchKernelGGL(k_untranspose4<4>, g4, dim3(256), 0, s, (const uint32_t*)planes_dev, quads, (uint32_t*)out_dev);
  } else {
    hipLaunchKernelGGL(k_untranspose_bytes, g1, dim3(256), 0, s, planes_dev, count, element_bytes, (uint8_t*)out_dev);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_untranspose_bytes");
}

}  // extern "C"

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

// 16-byte elements (uint32x4 channels: extSplats, extSh, packedSplats),
// staged: a thread's four elements are 64 contiguous output bytes, so direct
// stores from neighbouring lanes land 64 bytes apart. The block's 1,024
// elements (16 KB) go through LDS and leave as consecutive 16-byte stores.
static __global__ __launch_bounds__(256) void k_untranspose16_staged(const uint32_t* __restrict__ planes, uint32_t quads,
                                                                     uint4* __restrict__ out) {
  __shared__ uint4 s[1024];
  const uint32_t q0 = blockIdx.x * 256u;
  const uint32_t nq = min(256u, quads - q0);  // the block's quads (the last block: fewer)
  const uint32_t t = threadIdx.x;
  if (t < nq) {
    uint32_t p[16];
#pragma unroll
    for (int k = 0; k < 16; k++) p[k] = __builtin_nontemporal_load(planes + (size_t)k * quads + q0 + t);
#pragma unroll
    for (int e = 0; e < 4; e++) {
      uint32_t v[4];
#pragma unroll
      for (int w = 0; w < 4; w++)
        v[w] = ((p[4 * w] >> (8 * e)) & 0xffu) | (((p[4 * w + 1] >> (8 * e)) & 0xffu) << 8) |
               (((p[4 * w + 2] >> (8 * e)) & 0xffu) << 16) | (((p[4 * w + 3] >> (8 * e)) & 0xffu) << 24);
      s[4 * t + e] = make_uint4(v[0], v[1], v[2], v[3]);
    }
  }
  __syncthreads();
  uint4* dst = out + (size_t)q0 * 4u;
  for (uint32_t v = t; v < 4u * nq; v += 256u) dst[v] = s[v];
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
  const dim3 g4((uint32_t)(((uint64_t)quads + 255) / 256)), g1((uint32_t)(((uint64_t)count + 255) / 256));
  if (aligned && element_bytes == 16) {
    hipLaunchKernelGGL(k_untranspose16_staged, g4, dim3(256), 0, s, (const uint32_t*)planes_dev, quads, (uint4*)out_dev);
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

// ---- SPZ (Niantic's gzip container, src/core/spz.ts) ------------------------------

namespace rdx {

constexpr double SPZ_COLOR_SCALE = 0.28209479177387814 / 0.15;  // SH_C0 / 0.15 (spz.ts:50, ply.ts:15)

// Column offsets of the unzipped body after the 16-byte header (spz.ts:152-160)
struct SpzCols {
  size_t centers, alphas, colors, scales, rotations, sh;
};

// The values of one splat into the f32 planes (centre 3, log-scale 3, quat 4,
// opacity, rgb 3, SH 3K; plane stride `st`), as readSpz computes them
// (spz.ts:297-347) before encoding them into a layout: centres as int24 x
// 2^-fractionalBits (v1: half floats), opacity a / 255, colour
// (b / 255 - 0.5) SH_C0 / 0.15 + 0.5, log-scale b / 16 - 10 (the scale
// exp(b / 16 - 10)), rotations as smallest-three (v3+) or x, y, z in
// [-1, 1] with w = sqrt(1 - |xyz|^2) (v2, v1), SH (b - 128) / 128. The
// byte pointers are the splat's own in each column, in global memory or in
// an LDS stage.
__device__ __forceinline__ void spz_splat(const uint8_t* cen, const uint8_t* alpha, const uint8_t* col,
                                          const uint8_t* sc, const uint8_t* rot, const uint8_t* sh, int version,
                                          int sh_values, float fixed, float* __restrict__ out, size_t st) {
  if (version == 1) {
#pragma unroll
    for (int k = 0; k < 3; k++) out[k * st] = half_to_float(cen[2 * k] | (cen[2 * k + 1] << 8));
  } else {
    const uint8_t* p = cen;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int32_t v = (int32_t)(((uint32_t)p[3 * k] | ((uint32_t)p[3 * k + 1] << 8) | ((uint32_t)p[3 * k + 2] << 16)) << 8) >> 8;
      out[k * st] = (float)v * fixed;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) out[(3 + k) * st] = (float)sc[k] / 16.0f - 10.0f;
  // The rotation and colour maps in double, rounded once: w = sqrt(1 - |xyz|^2)
  // is ill-conditioned near w = 0, where a float evaluation strays by many ulps
  // from readSpz's double one. One splat's few FP64 operations are free next to
  // its bytes.
  double q[4];
  if (version >= 3) {
    const uint8_t* p = rot;
    uint32_t rest = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
    const uint32_t largest = rest >> 30;
    double sum = 0.0;
#pragma unroll
    for (int k = 3; k >= 0; k--) {
      if ((uint32_t)k != largest) {
        const double v = 0.70710678118654752 * ((double)(rest & 511u) / 511.0) * ((rest >> 9) & 1u ? -1.0 : 1.0);
        rest >>= 10;
        q[k] = v;
        sum += v * v;
      }
    }
    q[largest] = sqrt(fmax(1.0 - sum, 0.0));
  } else {
    const uint8_t* p = rot;
    q[0] = (double)p[0] / 127.5 - 1.0;
    q[1] = (double)p[1] / 127.5 - 1.0;
    q[2] = (double)p[2] / 127.5 - 1.0;
    q[3] = sqrt(fmax(0.0, 1.0 - q[0] * q[0] - q[1] * q[1] - q[2] * q[2]));
  }
#pragma unroll
  for (int k = 0; k < 4; k++) out[(6 + k) * st] = (float)q[k];
  out[10 * st] = (float)alpha[0] / 255.0f;
#pragma unroll
  for (int k = 0; k < 3; k++) out[(11 + k) * st] = (float)(((double)col[k] / 255.0 - 0.5) * SPZ_COLOR_SCALE + 0.5);
  for (int v = 0; v < sh_values; v++) out[(14 + v) * st] = ((float)sh[v] - 128.0f) / 128.0f;
}

// The byte path, for a body at any alignment: each thread reads its own
// bytes from global memory.
static __global__ __launch_bounds__(256) void k_spz_decode(const uint8_t* __restrict__ body, SpzCols c, uint32_t count,
                                                           int version, int sh_values, float fixed,
                                                           float* __restrict__ planes, size_t st) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= count) return;
  const size_t s = i;
  spz_splat(body + c.centers + s * (version == 1 ? 6 : 9), body + c.alphas + s, body + c.colors + s * 3,
            body + c.scales + s * 3, body + c.rotations + s * (version >= 3 ? 4 : 3), body + c.sh + s * sh_values,
            version, sh_values, fixed, planes + i, st);
}

// The staged path, for a 4-byte-aligned body. A wave's byte loads in the byte
// path span up to 45 cache lines each (45 SH bytes per splat). Here block b
// first copies its 256 splats' span of every column into LDS, one dword per
// lane over consecutive addresses, and each thread then reads its bytes
// there.
constexpr uint32_t SPZ_STAGE_WORDS = 64u * 65u + 6u * 2u;  // 256 splats x 65 bytes (9+1+3+3+4+45) + a head and tail word per column

// Copies body[s, s + len) into the stage from word `w` on (the aligned words
// that cover it; only bytes below body_bytes are read) and advances `w`;
// returns the stage byte offset of body[s].
__device__ __forceinline__ uint32_t spz_stage(const uint8_t* __restrict__ body, size_t body_bytes, size_t s,
                                              uint32_t len, uint32_t* stage, uint32_t& w) {
  const size_t a = s & ~(size_t)3;
  const uint32_t head = (uint32_t)(s - a);
  const uint32_t off = 4u * w + head;
  if (len == 0) return off;
  const uint32_t words = (head + len + 3u) / 4u;
  for (uint32_t q = threadIdx.x; q < words; q += 256u) {
    const size_t at = a + 4u * (size_t)q;
    uint32_t v;
    if (at + 4u <= body_bytes) {
      v = *reinterpret_cast<const uint32_t*>(body + at);
    } else {  // the body's last, partial word
      v = 0u;
      for (uint32_t k = 0; k < 4u; k++)
        if (at + k < body_bytes) v |= (uint32_t)body[at + k] << (8u * k);
    }
    stage[w + q] = v;
  }
  w += words;
  return off;
}

static __global__ __launch_bounds__(256) void k_spz_decode_staged(const uint8_t* __restrict__ body, size_t body_bytes,
                                                                  SpzCols c, uint32_t count, int version, int sh_values,
                                                                  float fixed, float* __restrict__ planes, size_t st) {
  __shared__ uint32_t stage[SPZ_STAGE_WORDS];
  const uint32_t b0 = blockIdx.x * 256u;
  const uint32_t nb = min(256u, count - b0);
  const uint32_t bc = version == 1 ? 6u : 9u, br = version >= 3 ? 4u : 3u, bs = (uint32_t)sh_values;
  const size_t s0 = b0;
  uint32_t w = 0;
  const uint32_t o_cen = spz_stage(body, body_bytes, c.centers + s0 * bc, nb * bc, stage, w);
  const uint32_t o_alpha = spz_stage(body, body_bytes, c.alphas + s0, nb, stage, w);
  const uint32_t o_col = spz_stage(body, body_bytes, c.colors + s0 * 3u, nb * 3u, stage, w);
  const uint32_t o_sc = spz_stage(body, body_bytes, c.scales + s0 * 3u, nb * 3u, stage, w);
  const uint32_t o_rot = spz_stage(body, body_bytes, c.rotations + s0 * br, nb * br, stage, w);
  const uint32_t o_sh = spz_stage(body, body_bytes, c.sh + s0 * bs, nb * bs, stage, w);
  __syncthreads();
  const uint32_t t = threadIdx.x;
  if (t >= nb) return;
  const uint8_t* sb = reinterpret_cast<const uint8_t*>(stage);
  spz_splat(sb + o_cen + t * bc, sb + o_alpha + t, sb + o_col + 3u * t, sb + o_sc + 3u * t, sb + o_rot + t * br,
            sb + o_sh + t * bs, version, sh_values, fixed, planes + b0 + t, st);
}

}  // namespace rdx

extern "C" {

int rdx_spz_decode(void* stream, const uint8_t* body_dev, size_t body_bytes, uint32_t count, int version, int sh_degree,
                   int fractional_bits, float* planes_dev, size_t plane_stride) {
  if (version < 1 || version > 4 || sh_degree < 0 || sh_degree > 3)
    return fail(version >= 1 && version <= 4 && sh_degree == 4 ? RDX_E_UNSUPPORTED : RDX_E_INVALID,
                "Unsupported SPZ version %d or SH degree %d", version, sh_degree);
  if (fractional_bits < 0 || fractional_bits > 30) return fail(RDX_E_INVALID, "SPZ fractional bits %d", fractional_bits);
  if (count == 0) return RDX_OK;
  if (!body_dev || !planes_dev) return fail(RDX_E_INVALID, "rdx_spz_decode: NULL body or planes");
  if (plane_stride < count) return fail(RDX_E_INVALID, "rdx_spz_decode: plane stride %zu below %u", plane_stride, count);
  static const int SH_VECTORS[4] = {0, 3, 8, 15};
  const int sh_values = SH_VECTORS[sh_degree] * 3;
  SpzCols c;
  c.centers = 0;
  const size_t n = count;
  c.alphas = c.centers + n * (version == 1 ? 6 : 9);
  c.colors = c.alphas + n;
  c.scales = c.colors + n * 3;
  c.rotations = c.scales + n * 3;
  c.sh = c.rotations + n * (version >= 3 ? 4 : 3);
  const size_t need = c.sh + n * sh_values;
  if (body_bytes < need) return fail(RDX_E_INVALID, "SPZ body holds %zu bytes, %zu needed for %u splats", body_bytes, need, count);
  const float fixed = 1.0f / (float)(1u << fractional_bits);
  const dim3 grid((uint32_t)(((uint64_t)count + 255) / 256));
  if (((uintptr_t)body_dev & 3u) == 0)
    hipLaunchKernelGGL(k_spz_decode_staged, grid, dim3(256), 0, (hipStream_t)stream, body_dev, body_bytes, c, count,
                       version, sh_values, fixed, planes_dev, plane_stride);
  else
    hipLaunchKernelGGL(k_spz_decode, grid, dim3(256), 0, (hipStream_t)stream, body_dev, c, count, version, sh_values,
                       fixed, planes_dev, plane_stride);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_spz_decode");
}

}  // extern "C"

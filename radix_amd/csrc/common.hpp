// Shared device helpers for the gfx950 splat pipeline: wave64 ballots and
// lane masks, DPP prefix sums, agent-scope atomics for cross-XCD look-back,
// and the float/half bit helpers the codecs use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RDX_WAVE 64

namespace rdx {

// ---- Lane masks and ballots (wave64) ----------------------------------------

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Set bits of `m` in lanes below this lane.
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// ---- DPP inclusive prefix sum over a wave64 ------------------------------------
// Rows of 16: row_shr 1, 2, 4, 8 with zeros shifted in (Hillis-Steele), then
// row_bcast:15 carries each even row's last lane into the next row and
// row_bcast:31 carries lane 31 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); // row_bcast:31
  return x;
}

// Value of lane `src` (uniform) broadcast to the wave.
__device__ __forceinline__ uint32_t wave_read_lane(uint32_t x, int src) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, src);
}

// Exclusive prefix sum over a block of BLOCK threads of one value each;
// `scratch` holds BLOCK/64 words of LDS. Returns the exclusive sum and the
// block total in *total. Contains two barriers.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t x, uint32_t* scratch, uint32_t* total,
                                                        int tid = (int)threadIdx.x) {
  constexpr int WAVES = BLOCK / RDX_WAVE;
  const int w = tid / RDX_WAVE;
  uint32_t inc = wave_inclusive_sum(x);
  if ((tid & 63) == 63) scratch[w] = inc;
  __syncthreads();
  uint32_t before = 0, sum = 0;
#pragma unroll
  for (int i = 0; i < WAVES; i++) {
    uint32_t v = scratch[i];
    before += i < w ? v : 0u;
    sum += v;
  }
  __syncthreads();
  *total = sum;
  return before + inc - x;
}

// ---- Agent-scope atomics (coherent across the XCDs' L2 caches) ---------------

__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_agent(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- Float bits ---------------------------------------------------------------

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

// Half bits (low 16 of h) to float: exact.
__device__ __forceinline__ float half_to_float(uint32_t h) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)(h & 0xffffu));
}

// Sort keys (src/core/sort.ts:46-58): the float32 bits f of a distance; a
// pattern not below +infinity's (NaN, infinity, negative) gives `mask`.
__device__ __forceinline__ uint32_t sort_key(uint32_t f, int bits) {
  const uint32_t mask = bits == 32 ? 0xffffffffu : ((1u << bits) - 1u);
  if (!(f < 0x7f800000u)) return mask;
  uint32_t q;
  if (bits == 32) {
    q = f;
  } else if (bits == 24) {
    q = f >> 7;
  } else {
    const uint32_t lo = 95u << 23, hi = (159u << 23) - 1u;
    q = (min(max(f, lo), hi) - lo) >> 13;
  }
  const uint32_t vmax = (uint32_t)(0xffull << (bits - 8)) - 1u;
  return min(mask ^ q, vmax);
}

__host__ __device__ __forceinline__ uint32_t key_invalid(int bits) {
  return bits == 32 ? 0xffffffffu : ((1u << bits) - 1u);
}

}  // namespace rdx

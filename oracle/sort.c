/*
 * Oracle: sort keys and the stable CPU radix sort, restated from
 * src/core/sort.ts. Test infrastructure only (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* src/core/sort.ts:29-31: the float bit range 16-bit keys keep */
#define KEY16_LO (95u << 23)
#define KEY16_HI ((159u << 23) - 1u)
#define FLOAT_BITS_INFINITY 0x7f800000u

/* src/core/sort.ts:37-39 */
uint32_t orc_key_invalid(int bits) { return bits == 32 ? 0xffffffffu : ((1u << bits) - 1u); }

/* src/core/sort.ts:41-43: 0xff * 2^(bits - 8) - 1 */
uint32_t orc_key_valid_max(int bits) { return (uint32_t)(0xffull << (bits - 8)) - 1u; }

/* src/core/sort.ts:46-58. The distance goes through a Float32Array slot
 * (round to nearest even, as `(float)` does); a bit pattern not below
 * +infinity's (NaN, infinity, any negative value) gives the invalid key. */
uint32_t orc_key_compute(double distance, int bits) {
  const uint32_t mask = orc_key_invalid(bits);
  float d32 = (float)distance;
  uint32_t f;
  memcpy(&f, &d32, 4);
  if (!(f < FLOAT_BITS_INFINITY)) return mask;
  uint32_t q;
  if (bits == 32) {
    q = f;
  } else if (bits == 24) {
    q = f >> 7;
  } else {
    uint32_t c = f < KEY16_LO ? KEY16_LO : (f > KEY16_HI ? KEY16_HI : f);
    q = (c - KEY16_LO) >> 13;
  }
  uint32_t key = mask ^ q;
  uint32_t vmax = orc_key_valid_max(bits);
  return key < vmax ? key : vmax;
}

void orc_keys_compute(const double* distance, size_t n, int bits, uint32_t* keys) {
  for (size_t i = 0; i < n; i++) keys[i] = orc_key_compute(distance[i], bits);
}

/* src/core/sort.ts:84-91: exclusive prefix sum of histogram[0..last] */
static void exclusive_sum(uint32_t* h, uint32_t last) {
  uint32_t sum = 0;
  for (uint32_t d = 0; d <= last; d++) {
    uint32_t c = h[d];
    h[d] = sum;
    sum += c;
  }
}

/* src/core/sort.ts:97-110: one counting pass over the low 16 bits */
static int64_t sort_16(const uint32_t* keys, uint32_t count, uint32_t* ordering) {
  uint32_t* h = (uint32_t*)calloc(65536, 4);
  for (uint32_t i = 0; i < count; i++) h[keys[i] & 0xffffu]++;
  int64_t visible = (int64_t)count - h[0xffff];
  exclusive_sum(h, 0xffff);
  for (uint32_t i = 0; i < count; i++) ordering[h[keys[i] & 0xffffu]++] = i;
  free(h);
  return visible;
}

/* Two LSD passes: `lo_bits` then the rest of the key (sort.ts:113-137 with
 * 11 + 13, sort.ts:139-163 with 16 + 16). The first pass records each
 * position's index and high digit, the second scatters by the high digit. */
static int64_t sort_two_pass(const uint32_t* keys, uint32_t count, unsigned lo_bits, unsigned hi_bins,
                             uint32_t* ordering) {
  const uint32_t lo_mask = (1u << lo_bits) - 1u;
  uint32_t* lo = (uint32_t*)calloc((size_t)1 << lo_bits, 4);
  uint32_t* hi = (uint32_t*)calloc(hi_bins, 4);
  uint32_t* index = (uint32_t*)malloc((size_t)(count ? count : 1) * 4);
  uint32_t* digit = (uint32_t*)malloc((size_t)(count ? count : 1) * 4);
  int64_t visible = -1;
  for (uint32_t i = 0; i < count; i++) {
    uint32_t k = keys[i];
    if ((k >> lo_bits) >= hi_bins) goto out; /* outside the reference's contract (keys < 2^bits) */
    lo[k & lo_mask]++;
    hi[k >> lo_bits]++;
  }
  visible = (int64_t)count - hi[hi_bins - 1];
  exclusive_sum(lo, lo_mask);
  exclusive_sum(hi, hi_bins - 1);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t k = keys[i];
    uint32_t pos = lo[k & lo_mask]++;
    index[pos] = i;
    digit[pos] = k >> lo_bits;
  }
  for (uint32_t j = 0; j < count; j++) ordering[hi[digit[j]]++] = index[j];
out:
  free(lo);
  free(hi);
  free(index);
  free(digit);
  return visible;
}

/* src/core/sort.ts:64-82: the visible count, or -1 for keys above 2^24 at
 * 24 bits (which the reference's typed arrays drop silently). */
int64_t orc_sort(const uint32_t* keys, uint32_t count, int bits, uint32_t* ordering) {
  if (bits == 16) return sort_16(keys, count, ordering);
  if (bits == 24) return sort_two_pass(keys, count, 11, 1u << 13, ordering);
  if (bits == 32) return sort_two_pass(keys, count, 16, 1u << 16, ordering);
  return -1;
}

// Stable LSD radix sort of (uint32 key, uint32 value) for gfx950.
//
// A sort of P digit passes (P = ceil(bits / RB), RB = 4, 6 or 8 bits) is:
//   clear      zero the digit histograms, tile tickets and the first status rows
//   histogram  one read of the keys: per-digit counts of every pass at once
//              (LDS-privatised, flushed with agent-scope atomics); the fused
//              projection kernel produces the same histograms instead
//   P passes   one kernel each ("onesweep"): a tile of BLOCK*KPT keys is
//              ranked in registers with wave64 ballots (match-any over the
//              digit's bits, mbcnt for the rank), the per-digit tile counts
//              are published and a single-pass decoupled look-back over the
//              previous tiles gives each digit's prefix; keys are staged in
//              LDS in digit order and written out in runs.
// Tiles are taken in order from an atomic ticket, so every tile a tile waits
// on is already running: forward progress holds for any grid size.
//
// Status words: bits 30-31 flag (1 aggregate, 2 inclusive prefix), bits
// 0-29 the count. Two status regions alternate between passes; each pass
// zeroes the other region's rows for the next pass.
#pragma once

#include "common.hpp"

namespace rdx {

constexpr uint32_t ST_AGG = 1u << 30;
constexpr uint32_t ST_INC = 2u << 30;
constexpr uint32_t ST_MASK = (1u << 30) - 1u;

// Digit histograms of a sort: hist[pass * 256 + digit] (256 bins per pass
// whatever the digit width), then counters.
constexpr int HIST_BINS = 256;
constexpr int MAX_PASSES = 8;
// hist words: MAX_PASSES * 256 bins, then [tickets: MAX_PASSES], [visible_rule_count], [n_input]
constexpr int HIST_TICKETS = MAX_PASSES * HIST_BINS;
constexpr int HIST_RULE = HIST_TICKETS + MAX_PASSES;
constexpr int HIST_ERROR = HIST_RULE + 1;  // set when a look-back wait ran out
constexpr int HIST_WORDS = HIST_RULE + 8;
// Look-back polls before giving up (each poll sleeps ~64 clocks): far above
// any real wait, a bound that keeps a broken run from hanging the device.
constexpr uint32_t SPIN_LIMIT = 1u << 22;
constexpr int LOOKBACK = 8;  // predecessors read together in the look-back

struct PassArgs {
  const uint32_t* keys_in;
  const uint32_t* vals_in;   // NULL: the value is the key's input index
  uint32_t* keys_out;        // NULL: keys are not written (last pass of an index sort)
  uint32_t* vals_out;
  const uint32_t* n_dev;     // input count on the device (NULL: n below)
  uint32_t n;                // input count (or the cap with n_dev)
  const uint32_t* hist;      // this pass's digit counts (256 words)
  uint32_t* ticket;          // this pass's tile ticket
  uint32_t* status;          // this pass's status rows [tile * BINS + digit]
  uint32_t* status_next;     // the other region: rows zeroed here for the next pass
  uint32_t invalid;          // with drop: keys equal to this are dropped
  int drop;                  // first pass of a compacting sort
  int use_hist_total;        // n = sum of this pass's histogram (passes after a compaction)
  int shift;
  uint32_t key_mask;         // the sorted bits (keys above 2^bits keep their high bits, unsorted)
  uint32_t* count_out;       // tile 0 writes the sorted count here (may be NULL)
  uint32_t* draw_args;       // tile 0 writes [6 * 1024, ceil(count / 1024), 0, 0, 0] (may be NULL)
  uint32_t* error;           // set to 1 if a look-back wait ran out
  // Segments (independent sorts in one launch; NULL: one segment): tile t
  // belongs to segment seg_of_tile[t], whose keys start at seg[4 s + 1] in
  // the input and output (both: the same offset), first tile seg[4 s],
  // count seg[4 s + 2]; its histograms at hist + s * hist_stride. count_out
  // and draw_args are per segment (1 and 5 words apart).
  const uint32_t* seg_of_tile;
  const uint32_t* seg;
  uint32_t hist_stride;
  uint32_t dst_cap;          // output capacity (a guard on every write)
};

// Zero the histograms, tickets and counters, and the first `rows` status
// rows (rows * bins words) of region 0.
static __global__ void k_clear(uint32_t* hist, uint32_t* status, uint32_t words) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < HIST_WORDS; i += stride) hist[i] = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) status[i] = 0u;
}

// Add a key's digits of every pass (digit width rb) to the LDS histograms
// (one ds_add per pass).
__device__ __forceinline__ void hist_add_key(uint32_t* s_hist, uint32_t key, int passes, int rb) {
  const uint32_t mask = (1u << rb) - 1u;
  for (int p = 0; p < passes; p++) atomicAdd(&s_hist[p * HIST_BINS + ((key >> (p * rb)) & mask)], 1u);
}

// Flush LDS histograms to the global ones (nonzero bins only).
template <int BLOCK>
__device__ __forceinline__ void hist_flush(const uint32_t* s_hist, uint32_t* hist, int passes) {
  for (int i = threadIdx.x; i < passes * HIST_BINS; i += BLOCK) {
    uint32_t c = s_hist[i];
    if (c) add_agent(&hist[i], c);
  }
}

// The standalone histogram pass over device keys: counts per digit of every
// pass for the kept keys, and the reference CPU sort's hidden count (keys
// whose top part is all ones: src/core/sort.ts:104, 128, 154).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_histogram(const uint32_t* __restrict__ keys, const uint32_t* n_dev,
                                                     uint32_t n_cap, uint32_t* hist, int passes, int rb, int drop,
                                                     uint32_t invalid, int bits) {
  __shared__ uint32_t s_hist[MAX_PASSES * HIST_BINS];
  __shared__ uint32_t s_rule;
  for (int i = threadIdx.x; i < passes * HIST_BINS; i += BLOCK) s_hist[i] = 0u;
  if (threadIdx.x == 0) s_rule = 0u;
  __syncthreads();
  const uint32_t n = n_dev ? min(*n_dev, n_cap) : n_cap;
  const int rule_shift = bits == 16 ? 0 : (bits == 24 ? 11 : 16);
  const uint32_t rule_all = bits == 16 ? 0xffffu : (bits == 24 ? 0x1fffu : 0xffffu);
  const uint32_t key_mask = key_invalid(bits);
  uint32_t rule = 0;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
    uint32_t k = keys[i];
    uint32_t top = bits == 16 ? (k & 0xffffu) : (k >> rule_shift);
    rule += top == rule_all ? 1u : 0u;
    if (drop && k == invalid) continue;
    hist_add_key(s_hist, k & key_mask, passes, rb);
  }
  if (rule) atomicAdd(&s_rule, rule);
  __syncthreads();
  hist_flush<BLOCK>(s_hist, hist, passes);
  if (threadIdx.x == 0 && s_rule) add_agent(&hist[HIST_RULE], s_rule);
}

// One digit pass. Tile t (from the ticket) covers input keys
// [t * TILE, (t + 1) * TILE): wave w holds [w * 64 * KPT, ...), as KPT
// rounds of 64 consecutive keys, so (round, lane) order is input order.
template <int RB, int BLOCK, int KPT>
__global__ __launch_bounds__(BLOCK) void k_onesweep(PassArgs a) {
  constexpr int BINS = 1 << RB;
  constexpr uint32_t DMASK = BINS - 1;
  constexpr int WAVES = BLOCK / RDX_WAVE;
  constexpr int WTILE = RDX_WAVE * KPT;
  constexpr int TILE = BLOCK * KPT;
  static_assert(BLOCK % RDX_WAVE == 0, "block of whole waves");

  __shared__ uint32_t s_wave[WAVES][BINS];  // per-wave digit counts, then exclusive offsets
  __shared__ uint32_t s_digit_base[BINS];   // global start of each digit (scan of the histogram)
  __shared__ uint32_t s_local[BINS];        // start of each digit within the tile
  __shared__ uint32_t s_out[BINS];          // global start - local start, for the write-out
  __shared__ uint32_t s_keys[TILE];
  __shared__ uint32_t s_vals[TILE];
  __shared__ uint32_t s_scan[WAVES];
  __shared__ uint32_t s_misc[4];

  const int tid = threadIdx.x;
  const int w = tid / RDX_WAVE;
  const int lane = tid & (RDX_WAVE - 1);

  if (tid == 0) {
    const uint32_t t = add_agent(a.ticket, 1u);
    s_misc[0] = t;
    s_misc[1] = a.seg_of_tile ? a.seg_of_tile[t] : 0u;
  }
  for (int d = tid; d < WAVES * BINS; d += BLOCK) (&s_wave[0][0])[d] = 0u;
  __syncthreads();
  const uint32_t tile = s_misc[0];
  const uint32_t sg = s_misc[1];
  const uint32_t tile0 = a.seg ? a.seg[4 * sg] : 0u;
  const uint32_t off = a.seg ? a.seg[4 * sg + 1] : 0u;
  const uint32_t* hist = a.hist + (size_t)sg * a.hist_stride;

  // Digit starts of the segment: exclusive scan of this pass's histogram.
  uint32_t total_valid = 0;
  for (int base = 0; base < BINS; base += BLOCK) {
    const int d = base + tid;
    uint32_t c = d < BINS ? hist[d] : 0u;
    uint32_t tot;
    uint32_t ex = block_exclusive_sum<BLOCK>(c, s_scan, &tot);
    if (d < BINS) s_digit_base[d] = off + total_valid + ex;
    total_valid += tot;
  }
  __syncthreads();

  // Zero this tile's row of the other status region for the next pass.
  if (a.status_next)
    for (int d = tid; d < BINS; d += BLOCK) a.status_next[(size_t)tile * BINS + d] = 0u;

  const uint32_t seg_n = a.seg ? a.seg[4 * sg + 2] : a.n;
  uint32_t n = a.n_dev ? min(*a.n_dev, seg_n) : seg_n;
  if (a.use_hist_total) n = min(total_valid, seg_n);
  const uint32_t sorted_count = a.drop || a.use_hist_total ? total_valid : n;
  const uint32_t local_tile = tile - tile0;
  if (local_tile == 0 && tid == 0) {
    if (a.count_out) a.count_out[sg] = sorted_count;
    if (a.draw_args) {
      uint32_t* args = a.draw_args + 5 * sg;
      args[0] = 6u * 1024u;
      args[1] = (sorted_count + 1023u) / 1024u;
      args[2] = 0u;
      args[3] = 0u;
      args[4] = 0u;
    }
  }
  const uint32_t tile_base = local_tile * (uint32_t)TILE;
  if (tile_base >= n) return;
  const uint32_t* keys_in = a.keys_in + off;
  const uint32_t* vals_in = a.vals_in ? a.vals_in + off : nullptr;

  // Load and rank.
  uint32_t keys[KPT], vals[KPT], rank[KPT];
  const uint32_t wbase = tile_base + (uint32_t)(w * WTILE) + (uint32_t)lane;
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t i = wbase + (uint32_t)(j * RDX_WAVE);
    const bool in = i < n;
    keys[j] = in ? keys_in[i] : 0u;
    vals[j] = in ? (vals_in ? vals_in[i] : i) : 0u;
    rank[j] = in && !(a.drop && keys[j] == a.invalid) ? 0u : 0xffffffffu;  // 0xffffffff: not kept
  }
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const bool keep = rank[j] == 0u;
    const uint32_t d = ((keys[j] & a.key_mask) >> a.shift) & DMASK;
    uint64_t m = ballot(keep);
#pragma unroll
    for (int b = 0; b < RB; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = ballot(bit);
      m &= bit ? bb : ~bb;
    }
    if (keep) {
      const uint32_t below = count_below(m);
      const uint32_t cnt = (uint32_t)__builtin_popcountll(m);
      const uint32_t old = s_wave[w][d];
      if (below == cnt - 1u) s_wave[w][d] = old + cnt;
      rank[j] = old + below;
    }
  }
  __syncthreads();

  // Per digit: exclusive offsets across the waves, the tile's count.
  uint32_t my_count = 0;
  for (int d = tid; d < BINS; d += BLOCK) {
    uint32_t run = 0;
#pragma unroll
    for (int v = 0; v < WAVES; v++) {
      const uint32_t c = s_wave[v][d];
      s_wave[v][d] = run;
      run += c;
    }
    if (d == tid) my_count = run;
    s_local[d] = run;  // the count for now
  }
  __syncthreads();

  // Publish the tile's counts, then look back for each digit's prefix.
  for (int d = tid; d < BINS; d += BLOCK) {
    const uint32_t cnt = s_local[d];
    uint32_t* st = a.status + (size_t)tile * BINS + d;
    uint32_t prefix = 0;
    if (local_tile == 0) {
      store_agent(st, ST_INC | cnt);
    } else {
      store_agent(st, ST_AGG | cnt);
      // Look back LOOKBACK predecessors at a time (their loads in flight
      // together): every tile is resident at once, so the walk to the
      // nearest inclusive prefix is the latency chain of the pass.
      int k = (int)tile - 1;
      uint32_t spins = 0;
      bool done = false;
      while (!done && k >= (int)tile0) {
        uint32_t v[LOOKBACK];
#pragma unroll
        for (int j = 0; j < LOOKBACK; j++)
          v[j] = k - j >= (int)tile0 ? load_agent(a.status + (size_t)(k - j) * BINS + d) : ST_INC;
        int used = 0;
#pragma unroll
        for (int j = 0; j < LOOKBACK; j++) {
          if (done || used < j) continue;  // stop at the first not-yet-published status
          const uint32_t f = v[j] & ~ST_MASK;
          if (f == 0u) continue;
          prefix += v[j] & ST_MASK;
          used = j + 1;
          if (f == ST_INC) done = true;
        }
        k -= used;
        if (!done && used == 0) {
          // Bounded wait: a predecessor that never publishes (a broken
          // invariant, not a normal run) flags an error instead of hanging.
          if (++spins > SPIN_LIMIT) {
            store_agent(a.error, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      store_agent(st, ST_INC | (prefix + cnt));
    }
    s_out[d] = s_digit_base[d] + prefix;  // the digit's global start for this tile
  }
  (void)my_count;

  // Local starts of the digits within the tile (exclusive scan of the counts).
  uint32_t tile_kept = 0;
  for (int base = 0; base < BINS; base += BLOCK) {
    const int d = base + tid;
    uint32_t c = d < BINS ? s_local[d] : 0u;
    uint32_t tot;
    uint32_t ex = block_exclusive_sum<BLOCK>(c, s_scan, &tot);
    if (d < BINS) {
      s_local[d] = tile_kept + ex;
      s_out[d] -= tile_kept + ex;
    }
    tile_kept += tot;
  }
  __syncthreads();

  // Stage in digit order.
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    if (rank[j] != 0xffffffffu) {
      const uint32_t d = ((keys[j] & a.key_mask) >> a.shift) & DMASK;
      const uint32_t p = s_local[d] + s_wave[w][d] + rank[j];
      s_keys[p] = keys[j];
      s_vals[p] = vals[j];
    }
  }
  __syncthreads();

  // Write out: runs of one digit go to consecutive addresses.
  for (uint32_t i = tid; i < tile_kept; i += BLOCK) {
    const uint32_t k = s_keys[i];
    const uint32_t dst = s_out[((k & a.key_mask) >> a.shift) & DMASK] + i;
    if (dst < a.dst_cap) {  // always true when the histograms are consistent
      if (a.keys_out) a.keys_out[dst] = k;
      a.vals_out[dst] = s_vals[i];
    }
  }
}

// Keys of device distances (GpuSortKeys.compute with the CPU's rule for
// non-finite and negative patterns).
static __global__ void k_compute_keys(const float* __restrict__ d, uint32_t n, int bits, uint32_t* __restrict__ keys) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    keys[i] = sort_key(fbits(d[i]), bits);
}

// Counts and draw arguments for an empty sort.
static __global__ void k_empty_outputs(uint32_t* count_out, uint32_t* draw_args) {
  if (threadIdx.x == 0) {
    if (count_out) *count_out = 0u;
    if (draw_args) {
      draw_args[0] = 6u * 1024u;
      draw_args[1] = 0u;
      draw_args[2] = 0u;
      draw_args[3] = 0u;
      draw_args[4] = 0u;
    }
  }
}

}  // namespace rdx

// Stable LSD radix sort of (uint32 key, uint32 value) for gfx950.
//
// A sort of P digit passes (P = ceil(bits / RB), RB = 4, 6 or 8 bits) runs,
// per pass, three kernels over tiles of TILE = BLOCK * KPT keys:
//   count    per tile, the number of keys of each digit -> cnt[tile][digit]
//   scan     per (segment, group of 16 digits), an exclusive scan of each
//            digit's column over the segment's tiles, in place, and the
//            digit's total -> tot
//   scatter  per tile: rank every key within the tile (wave64 ballots for a
//            round's most present digits, LDS lane masks for the rest, mbcnt
//            for the rank), stage the tile in LDS in digit order and write it
//            out in runs starting at base(digit) + cnt[tile][digit]
// No kernel waits on another block: a pass costs one extra read of the keys
// (4 B per key) instead of a cross-tile look-back. Every count word is
// written before it is read, so nothing needs clearing between sorts.
//
// Segments (independent sorts in one launch, e.g. one per view): tile t
// belongs to segment seg_of_tile[t], whose keys are [off, off + n) in both
// the input and the output; its tiles are [tile0, tile0 + ntiles).
#pragma once

#include "common.hpp"

namespace rdx {

constexpr int MAX_BINS = 512;
constexpr int TOT_STRIDE = MAX_BINS + 1;  // tot words per segment: digits, then the hidden-rule count
constexpr int SCAN_BLOCK = 1024;
// Tile counts are stored tile-major, cnt[tile][cstride] (a count block
// writes its row contiguously, a scatter block reads it so): the scan walks
// them in column groups of CGROUP digits, one 64-byte line per tile.
constexpr int CGROUP = 16;
constexpr int SCAN_CHUNKS = SCAN_BLOCK / CGROUP;  // tile ranges per scan block and round
__host__ __device__ constexpr uint32_t count_stride(int rows) { return (uint32_t)((rows + CGROUP - 1) / CGROUP * CGROUP); }
// Digits a wave ranks in a round with one ballot each before the LDS lane
// masks take the rest: a round whose keys share a few digits (the top
// digit of clustered distances) would otherwise serialise its LDS atomics
// on one address. Peeling stops at the first digit held by fewer than
// PEEL_MIN lanes (a round of spread digits then costs one ballot). Only the
// last digit pass peels (PassArgs::peel): the lower digits of distance keys
// are spread, and there the ballots cost more than they save.
constexpr int PEEL = 3;
constexpr uint32_t PEEL_MIN = 8;

// Ranged sorts: a compacting 32-bit sort sorts key - base, base = the
// segment's smallest kept key rounded down to a multiple of 2^RANGED_RB, by
// RANGED_RB-bit digits. The first pass's digit is then the key's own low
// digit (it needs no base; its count kernel finds each tile's smallest and
// largest kept key and its scan reduces them per segment); the later passes
// read the base. Three passes cover RANGED_BITS bits above the base, i.e.
// kept distances within a factor 2^16 of each other; a fourth pass runs only
// for segments whose keys reach further (the third pass then writes to the
// scratch instead of the output).
constexpr int RANGED_RB = 9;
constexpr int RANGED_BITS = 3 * RANGED_RB;
constexpr uint32_t RANGED_LOW = (1u << RANGED_RB) - 1u;
// The fourth (gated) pass's digits: bits RANGED_BITS..31 of key - base.
constexpr int GATED_BINS = 1 << (32 - RANGED_BITS);

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs
// on XCD b % 8), each with its own L2. The scatter maps them so that each
// XCD takes one contiguous range of tiles: the digit runs of neighbouring
// tiles meet inside shared cache lines, which then fill in one L2 instead of
// being written back in pieces from two. A bijection of [0, tiles).
constexpr uint32_t XCDS = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t tiles) {
  const uint32_t x = b % XCDS, q = tiles / XCDS, r = tiles % XCDS;
  return x * q + (x < r ? x : r) + b / XCDS;
}

// One pass of a sort: what the count, scan and scatter kernels share.
struct PassArgs {
  const uint32_t* keys_in;
  const uint32_t* vals_in;   // NULL: the value is the key's input index
  uint32_t* keys_out;        // NULL: keys are not written (last pass of an index sort)
  uint32_t* vals_out;
  const uint32_t* n_dev;     // input count on the device (NULL: n below); one segment only
  uint32_t n;                // input count (or the cap with n_dev)
  uint32_t* cnt;             // [tiles][cstride] tile counts, scanned in place (per column, over tiles)
  uint32_t cstride;          // count_stride(rows)
  uint32_t* tot;             // [segments][TOT_STRIDE] digit totals
  uint32_t* nvalid;          // [segments] kept keys (written by a dropping pass)
  uint32_t tiles;            // tiles of all segments (the row length of cnt)
  int rows;                  // count rows: BINS, or BINS + 1 with the hidden-rule row
  int shift;
  uint32_t key_mask;         // the sorted bits (keys above 2^bits keep their high bits, unsorted)
  uint32_t invalid;          // with drop: keys equal to this are dropped
  int drop;                  // this pass drops invalid keys (first pass of a compacting sort)
  int compacted;             // an earlier pass dropped: a segment's count is nvalid[s]
  int bits;                  // key width (the hidden rule of the host sort)
  uint32_t* count_out;       // last pass: the sorted count per segment (may be NULL)
  uint32_t* draw_args;       // last pass: [6 * 1024, ceil(count / 1024), 0, 0, 0] per segment (may be NULL)
  const uint32_t* seg_of_tile;  // NULL: one segment of all tiles
  const uint32_t* seg;          // [segments][4]: tile0, off, n, ntiles
  uint32_t dst_cap;          // output capacity (a guard on every write)
  uint32_t tile_keys;        // keys per tile (the scan covers the tiles that hold keys)
  // ranged sorts
  int ranged;                // 0: digits of the key; 1: first ranged pass; 2: later ranged pass (digits of key - base)
  uint32_t* tile_mm;         // [tiles][2] smallest and largest kept key (first ranged pass)
  uint32_t* seg_mm;          // [segments][2] the same per segment (the first pass's scan)
  int gate;                  // run only for segments whose keys reach beyond RANGED_BITS above the base
  int peel;                  // rank a round's most present digits by ballots first (the top digit pass)
  int alt;                   // segments that reach beyond write to keys_alt / vals_alt instead, and
                             // count their gated digits per output tile into cnt_g
  uint32_t* keys_alt;
  uint32_t* vals_alt;
  // The gated (fourth) pass has no count or scan launch of its own. For the
  // segments that need it, the second ranged pass's count blocks clear their
  // tile's row of cnt_g (zero_g), the third pass's scatter counts each key's
  // gated digit into the row of the tile it lands in, and the segment's last
  // scatter block to finish (its ticket in done, cleared by the first pass's
  // scan) scans the rows in place and writes the digit totals to tot_g.
  int zero_g;
  uint32_t* cnt_g;           // [tiles][GATED_BINS]
  uint32_t* tot_g;           // [segments][GATED_BINS]
  uint32_t* done;            // [segments]
};

struct SegInfo {
  uint32_t s, tile0, off, cap, ntiles;  // cap: the segment's input size before compaction
};

__device__ __forceinline__ SegInfo seg_info(const PassArgs& a, uint32_t s) {
  SegInfo g;
  g.s = s;
  if (a.seg) {
    g.tile0 = a.seg[4 * s];
    g.off = a.seg[4 * s + 1];
    g.cap = a.seg[4 * s + 2];
    g.ntiles = a.seg[4 * s + 3];
  } else {
    g.tile0 = 0;
    g.off = 0;
    g.cap = a.n;
    g.ntiles = a.tiles;
  }
  return g;
}

// The keys a segment holds in this pass (the kept count after a dropping
// pass). Loads of keys below the capacity may be issued before it is known.
__device__ __forceinline__ uint32_t seg_count(const PassArgs& a, const SegInfo& g) {
  uint32_t n = g.cap;
  if (a.n_dev) n = min(*a.n_dev, n);
  if (a.compacted) n = min(a.nvalid[g.s], n);
  return n;
}

// The base of a segment's digits (0 but in a later ranged pass) and whether
// its kept keys reach beyond RANGED_BITS above the base.
__device__ __forceinline__ uint32_t seg_base(const PassArgs& a, uint32_t s) {
  return a.ranged == 2 ? (a.seg_mm[2 * s] & ~RANGED_LOW) : 0u;
}
__device__ __forceinline__ bool seg_wide(const PassArgs& a, uint32_t s) {
  const uint32_t mn = a.seg_mm[2 * s], mx = a.seg_mm[2 * s + 1];
  return mx >= mn && mx - (mn & ~RANGED_LOW) >= (1u << RANGED_BITS);
}
__device__ __forceinline__ uint32_t digit_of(const PassArgs& a, uint32_t k, uint32_t base, uint32_t dmask) {
  return (((k & a.key_mask) - base) >> a.shift) & dmask;
}

// The hidden rule of the reference's CPU sort: keys whose top part is all
// ones sort last and are not counted as visible (src/core/sort.ts:104, 128,
// 154).
__device__ __forceinline__ bool hidden_rule(uint32_t k, int bits) {
  if (bits == 16) return (k & 0xffffu) == 0xffffu;
  if (bits == 24) return (k >> 11) == 0x1fffu;
  return (k >> 16) == 0xffffu;
}

// ---- count ---------------------------------------------------------------------

// Counts that the blocks of one kernel add up and one block of the same
// kernel then reads (the gated pass's counts below) are device-scope atomic
// adds, which are performed at the device's coherence point, not in an
// XCD's L2. Each block waits for its adds to complete before it takes a
// ticket; the block holding the last ticket reads the counts with
// agent-scope atomic loads, which miss the L2. An agent-scope fence in
// their place would write back the whole L2 of the writer's XCD in every
// block.
__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void wait_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// The gated pass's scan of one segment, by the segment's last third-pass
// scatter block: thread (c, d) sums column d of cnt_g over the c-th run of
// the segment's tiles; the runs' sums are scanned per column; each thread
// writes its run's exclusive prefixes in place, and the column totals go to
// tot_g. The rows were counted with device-scope atomics, so they are read
// at the coherence point.
template <int BLOCK>
__device__ __forceinline__ void gated_scan(const PassArgs& a, const SegInfo& g, uint32_t used, uint32_t* lds) {
  constexpr int CH = BLOCK / GATED_BINS;
  static_assert(CH * GATED_BINS == BLOCK && CH >= 1, "whole runs of columns");
  const int tid = threadIdx.x;
  const int d = tid % GATED_BINS, c = tid / GATED_BINS;
  const uint32_t per = (used + CH - 1) / CH;
  const uint32_t t0 = min(used, (uint32_t)c * per), t1 = min(used, t0 + per);
  uint32_t* col = a.cnt_g + (size_t)g.tile0 * GATED_BINS + d;
  constexpr uint32_t CHUNK = 16;  // loads in flight per thread
  uint32_t sum = 0;
  for (uint32_t t = t0; t < t1; t += CHUNK) {
    uint32_t v[CHUNK];
#pragma unroll
    for (uint32_t k = 0; k < CHUNK; k++) v[k] = t + k < t1 ? ld_agent(col + (size_t)(t + k) * GATED_BINS) : 0u;
#pragma unroll
    for (uint32_t k = 0; k < CHUNK; k++) sum += v[k];
  }
  lds[c * GATED_BINS + d] = sum;
  __syncthreads();
  if (c == 0) {
    uint32_t run = 0;
    for (int k = 0; k < CH; k++) {
      const uint32_t x = lds[k * GATED_BINS + d];
      lds[k * GATED_BINS + d] = run;
      run += x;
    }
    a.tot_g[(size_t)g.s * GATED_BINS + d] = run;
  }
  __syncthreads();
  uint32_t run = lds[c * GATED_BINS + d];
  for (uint32_t t = t0; t < t1; t += CHUNK) {
    uint32_t v[CHUNK];
#pragma unroll
    for (uint32_t k = 0; k < CHUNK; k++) v[k] = t + k < t1 ? ld_agent(col + (size_t)(t + k) * GATED_BINS) : 0u;
#pragma unroll
    for (uint32_t k = 0; k < CHUNK; k++) {
      if (t + k < t1) col[(size_t)(t + k) * GATED_BINS] = run;
      run += v[k];
    }
  }
}

// Tile counts of one pass's digit. Wave w holds keys [w * 64 * KPT, ...) of
// the tile and counts them with LDS atomics into SUBS private rows per wave
// (by lane), so a round whose keys share a digit (the top digit of
// clustered distances) contends SUBS ways less on one address; the rows
// (padded by one word: distinct banks) are summed at the end. A 1024-thread
// block of 512 digits keeps one row per wave: zeroing and summing 32 rows
// cost more than the contention they save (plain_4m step 267.6 -> 265.7 us
// in interleaved runs).
template <int RB, int BLOCK>
struct CountLds {
  static constexpr int BINS = 1 << RB;
  static constexpr int WAVES = BLOCK / RDX_WAVE;
  static constexpr int SUBS = BINS >= 512 ? (BLOCK >= 1024 ? 1 : 2) : 4;
  static constexpr int ROW = BINS + 1;
  uint32_t rows[WAVES * SUBS * ROW];
  uint32_t rule[WAVES];
  uint32_t mm[2];
};

// One tile's counts by a block of BLOCK threads (the count kernel's block,
// or a persistent block taking tiles in turn; L is free on entry and may be
// reused when it returns).
template <int RB, int BLOCK, int KPT>
__device__ __forceinline__ void count_tile(const PassArgs& a, uint32_t tile, CountLds<RB, BLOCK>& L) {
  using CL = CountLds<RB, BLOCK>;
  constexpr int BINS = CL::BINS;
  constexpr uint32_t DMASK = BINS - 1;
  constexpr int WAVES = CL::WAVES;
  constexpr int WTILE = RDX_WAVE * KPT;
  constexpr int TILE = BLOCK * KPT;
  constexpr int SUBS = CL::SUBS;
  constexpr int ROW = CL::ROW;
  uint32_t* s_rows = L.rows;
  uint32_t* s_rule = L.rule;
  uint32_t* s_mm = L.mm;

  const int tid = threadIdx.x;
  const int w = tid / RDX_WAVE;
  const int lane = tid & (RDX_WAVE - 1);
  for (int d = tid; d < WAVES * SUBS * ROW; d += BLOCK) s_rows[d] = 0u;
  if (tid < WAVES) s_rule[tid] = 0u;
  if (tid == 0) {
    s_mm[0] = 0xffffffffu;
    s_mm[1] = 0u;
  }
  const SegInfo g = seg_info(a, a.seg_of_tile ? a.seg_of_tile[tile] : 0u);
  // clear this tile's row of the gated pass's counts (its segment needs that pass)
  if (a.zero_g && tid < GATED_BINS && seg_wide(a, g.s)) a.cnt_g[(size_t)tile * GATED_BINS + tid] = 0u;
  const uint32_t kbase = seg_base(a, g.s);
  const uint32_t base = (tile - g.tile0) * (uint32_t)TILE;
  const uint32_t* keys = a.keys_in + g.off;
  const bool rule = a.rows > BINS;

  // a tile past the segment's keys writes nothing: the scan stops before it
  // (a sort sized by a capacity, its count on the device, launches them all,
  // and those past the count load nothing)
  const uint32_t n = seg_count(a, g);
  if (base >= n && base > 0) return;  // uniform per block; a segment's first tile always writes
  const uint32_t wbase = base + (uint32_t)(w * WTILE) + (uint32_t)lane;
  uint32_t k[KPT];
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t i = wbase + (uint32_t)(j * RDX_WAVE);
    k[j] = i < n ? keys[i] : 0u;
  }
  __syncthreads();  // rows zeroed
  uint32_t* row = s_rows + (w * SUBS + (lane & (SUBS - 1))) * ROW;
  uint32_t hidden = 0, kmin = 0xffffffffu, kmax = 0u;
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t i = wbase + (uint32_t)(j * RDX_WAVE);
    const bool in = i < n;
    const bool kept = in && !(a.drop && k[j] == a.invalid);
    if (rule && in && hidden_rule(k[j], a.bits)) hidden++;
    if (kept) {
      atomicAdd(&row[digit_of(a, k[j], kbase, DMASK)], 1u);
      kmin = min(kmin, k[j]);
      kmax = max(kmax, k[j]);
    }
  }
  if (rule && hidden) atomicAdd(&s_rule[w], hidden);
  if (a.ranged == 1) {
    for (int off = 32; off >= 1; off >>= 1) {
      kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, off));
      kmax = max(kmax, (uint32_t)__shfl_xor((int)kmax, off));
    }
    if (lane == 0) {
      atomicMin(&s_mm[0], kmin);
      atomicMax(&s_mm[1], kmax);
    }
  }
  __syncthreads();
  if (a.ranged == 1 && tid == 0) {
    a.tile_mm[2 * (size_t)tile] = s_mm[0];
    a.tile_mm[2 * (size_t)tile + 1] = s_mm[1];
  }
  for (int d = tid; d < BINS; d += BLOCK) {
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < WAVES * SUBS; r++) c += s_rows[r * ROW + d];
    a.cnt[(size_t)tile * a.cstride + d] = c;
  }
  if (rule && tid == 0) {
    uint32_t c = 0;
    for (int v = 0; v < WAVES; v++) c += s_rule[v];
    a.cnt[(size_t)tile * a.cstride + BINS] = c;
  }
}

template <int RB, int BLOCK, int KPT>
__global__ __launch_bounds__(BLOCK) void k_count(PassArgs a) {
  __shared__ CountLds<RB, BLOCK> L;
  count_tile<RB, BLOCK, KPT>(a, blockIdx.x, L);
}

// ---- scan ----------------------------------------------------------------------

// Block (s, g): for the CGROUP digit columns [g * CGROUP, ...) of segment
// s, an exclusive scan of each column over the segment's tiles, in place,
// and the column's total to tot[s][r] (r = bins: the hidden-rule row, to
// tot[s][MAX_BINS]). A round covers SCAN_CHUNKS x ITEMS tiles: thread
// (c, d) loads column d of the c-th run of ITEMS tiles (the 16 threads
// of a run read one 64-byte line per tile); wave d scans column d's 64 run
// sums (one per lane) through LDS; the thread writes its run's prefixes.
// A scan block of NT threads takes NT / 64 columns: a wave per column, a
// lane per run of tiles (the scan kernel: 1024 threads, CGROUP columns).
template <int NT>
struct ScanLds {
  static constexpr int CG = NT / RDX_WAVE;
  uint32_t sum[SCAN_CHUNKS][CG];
  uint32_t tot[CG];
};

// Scan items per segment: the column groups, and the range reduction in a
// ranged sort's first pass.
template <int NT>
__host__ __device__ __forceinline__ uint32_t scan_items_per_seg(int rows, int ranged) {
  return (uint32_t)(rows + ScanLds<NT>::CG - 1) / ScanLds<NT>::CG + (ranged == 1 ? 1u : 0u);
}

template <int ITEMS, int NT>
__device__ __forceinline__ void scan_item(const PassArgs& a, int bins, uint32_t item, ScanLds<NT>& L) {
  constexpr int CG = ScanLds<NT>::CG;
  static_assert(NT / CG == SCAN_CHUNKS && NT / RDX_WAVE == CG, "a wave per column, a lane per run");
  auto& s_sum = L.sum;
  auto& s_tot = L.tot;
  const uint32_t groups = (uint32_t)(a.rows + CG - 1) / CG;
  const uint32_t per_seg = scan_items_per_seg<NT>(a.rows, a.ranged);
  const uint32_t s = item / per_seg;
  const int g = (int)(item % per_seg);
  const int d = threadIdx.x % CG;
  const int c = threadIdx.x / CG;
  const int r = g * CG + d;
  uint32_t tile0 = 0, ntiles = a.tiles;
  if (a.seg) {
    tile0 = a.seg[4 * s];
    ntiles = a.seg[4 * s + 3];
  }
  {
    // the tiles that hold keys (at least the first): the count kernel wrote only theirs
    const uint32_t nk = seg_count(a, seg_info(a, s));
    ntiles = min(ntiles, max(1u, (nk + a.tile_keys - 1) / a.tile_keys));
  }
  if (g == (int)groups) {
    // the segment's smallest and largest kept key
    uint32_t mn = 0xffffffffu, mx = 0u;
    for (uint32_t t = threadIdx.x; t < ntiles; t += NT) {
      mn = min(mn, a.tile_mm[2 * (size_t)(tile0 + t)]);
      mx = max(mx, a.tile_mm[2 * (size_t)(tile0 + t) + 1]);
    }
    for (int off = 32; off >= 1; off >>= 1) {
      mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
      mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    }
    if (threadIdx.x % RDX_WAVE == 0) {
      s_sum[threadIdx.x / RDX_WAVE][0] = mn;
      s_sum[threadIdx.x / RDX_WAVE][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int v = 1; v < NT / RDX_WAVE; v++) {
        mn = min(mn, s_sum[v][0]);
        mx = max(mx, s_sum[v][1]);
      }
      a.seg_mm[2 * s] = mn;
      a.seg_mm[2 * s + 1] = mx;
      if (a.done) a.done[s] = 0u;  // the gated pass's count blocks count from zero
    }
    return;
  }
  uint32_t* col = a.cnt + (size_t)tile0 * a.cstride + r;
  const bool live = r < a.rows;
  uint32_t carry = 0;
  for (uint32_t r0 = 0; r0 < ntiles; r0 += SCAN_CHUNKS * ITEMS) {
    const uint32_t t0 = r0 + (uint32_t)c * ITEMS;
    uint32_t v[ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      v[k] = live && t0 + k < ntiles ? col[(size_t)(t0 + k) * a.cstride] : 0u;
      sum += v[k];
    }
    __syncthreads();  // the previous round's s_sum reads are done
    s_sum[c][d] = sum;
    __syncthreads();
    {
      // wave w scans column w: lane l holds run l's sum
      const int w = threadIdx.x / RDX_WAVE, l = threadIdx.x % RDX_WAVE;
      const uint32_t x = s_sum[l][w];
      const uint32_t inc = wave_inclusive_sum(x);
      __syncthreads();
      s_sum[l][w] = inc - x;
      if (l == RDX_WAVE - 1) s_tot[w] = inc;
    }
    __syncthreads();
    uint32_t run = carry + s_sum[c][d];
    const uint32_t total = s_tot[d];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      if (live && t0 + k < ntiles) col[(size_t)(t0 + k) * a.cstride] = run;
      run += v[k];
    }
    carry += total;
  }
  if (live && c == 0) a.tot[(size_t)s * TOT_STRIDE + (r < bins ? r : MAX_BINS)] = carry;
}

template <int ITEMS>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan(PassArgs a, int bins) {
  __shared__ ScanLds<SCAN_BLOCK> L;
  scan_item<ITEMS, SCAN_BLOCK>(a, bins, blockIdx.x, L);
}

// ---- scatter --------------------------------------------------------------------

// One digit pass's write-out. Tile t covers input keys [t * TILE, (t + 1) *
// TILE) of its segment: wave w holds [w * 64 * KPT, ...), as KPT rounds of 64
// consecutive keys, so (round, lane) order is input order.
// VALS: the pass reads values (a.vals_in); otherwise a key's value is its
// input index. A compile-time choice: a branch around the value loads would
// hide the count of loads in flight from the ranking's waits.
template <int RB, int BLOCK, int KPT>
struct ScatterLds {
  static constexpr int BINS = 1 << RB;
  static constexpr int WAVES = BLOCK / RDX_WAVE;
  static constexpr int TILE = BLOCK * KPT;
  uint16_t wave[WAVES][BINS];   // per-wave digit counts, then exclusive offsets (< TILE)
  uint32_t local[BINS];         // start of each digit within the tile
  uint32_t out[BINS];           // global start - local start, for the write-out
  uint64_t stage[TILE];         // ranking: lane masks; then the staged (key, value) pairs
  uint32_t scan[WAVES];
  uint32_t last;
};

// One tile's scatter by a block of BLOCK threads; `bid` is the block's
// place in the pass (the kernel's blockIdx.x, or a persistent block's turn;
// workgroup b runs on XCD b % 8 either way) and tid the thread's index in
// the block. L is free on entry and may be reused when it returns.
template <int RB, int BLOCK, int KPT, bool VALS>
__device__ __forceinline__ void scatter_tile(const PassArgs& a, uint32_t bid, ScatterLds<RB, BLOCK, KPT>& L, int tid) {
  constexpr int BINS = 1 << RB;
  constexpr uint32_t DMASK = BINS - 1;
  constexpr int WAVES = BLOCK / RDX_WAVE;
  constexpr int WTILE = RDX_WAVE * KPT;
  constexpr int TILE = BLOCK * KPT;
  static_assert(BLOCK % RDX_WAVE == 0, "block of whole waves");
  static_assert(BINS <= BLOCK, "a thread per digit");

  static_assert(WAVES * BINS * 8 <= TILE * 8, "lane masks fit in the staging area");
  static_assert(TILE < 65536, "16-bit wave counts");
  auto& s_wave = L.wave;
  uint32_t* s_local = L.local;
  uint32_t* s_out = L.out;
  uint64_t* s_stage = L.stage;
  uint32_t* s_scan = L.scan;
  uint64_t (*s_match)[BINS] = reinterpret_cast<uint64_t (*)[BINS]>(s_stage);  // per-wave lane masks of a round

  const int w = tid / RDX_WAVE;
  const int lane = tid & (RDX_WAVE - 1);
  // One segment: the tiles that hold keys are mapped over all XCDs. Its count
  // can be far below the capacity the grid covers (a count on the device,
  // the kept keys after a dropping pass), and a map over the whole grid would
  // give all of them to the first XCDs. The tiles past them keep their own
  // index and return below. A bijection of [0, tiles) either way.
  // The gated pass: most sorts are done after three passes, and their blocks
  // return before anything else (uniform per block).
  if (a.gate && !a.seg_of_tile && !seg_wide(a, 0u)) return;
  uint32_t tile;
  if (!a.seg && !a.seg_of_tile) {
    const uint32_t n0 = seg_count(a, seg_info(a, 0u));
    const uint32_t used = max(1u, (uint32_t)(((uint64_t)n0 + TILE - 1) / TILE));
    tile = bid < used ? xcd_tile(bid, used) : bid;
  } else {
    tile = xcd_tile(bid, a.tiles);
  }
  const SegInfo g = seg_info(a, a.seg_of_tile ? a.seg_of_tile[tile] : 0u);
  if (a.gate && !seg_wide(a, g.s)) return;  // (segments of a batch, each on its own)
  for (int d = tid; d < WAVES * BINS; d += BLOCK) {
    (&s_wave[0][0])[d] = 0u;
    (&s_match[0][0])[d] = 0ull;
  }
  const uint32_t kbase = seg_base(a, g.s);
  uint32_t* keys_out = a.keys_out;
  uint32_t* vals_out = a.vals_out;
  const bool to_gated = a.alt && seg_wide(a, g.s);  // this segment goes on to the gated pass
  if (to_gated) {
    keys_out = a.keys_alt;
    vals_out = a.vals_alt;
  }
  const uint32_t local_tile = tile - g.tile0;
  const uint32_t tile_base = local_tile * (uint32_t)TILE;
  // A tile past the segment's count (a sort sized by a capacity, its count
  // on the device) has nothing to write; a segment's first tile still writes
  // its count and draw arguments below.
  const uint32_t n = seg_count(a, g);
  if (tile_base >= n && local_tile != 0) return;  // uniform per block

  // Every load of the tile at once: first the digit totals and the tile's
  // scanned counts, then the keys, then the values (below the count).
  // Vector loads complete in issue order, so the digit-start scan waits only
  // for the first two and round j of the ranking only for keys[j]: the
  // ranking runs while the rest of the tile arrives.
  // (the gated pass: its digits' totals and scanned rows from the third pass)
  const uint32_t* tot = a.gate ? a.tot_g + (size_t)g.s * GATED_BINS : a.tot + (size_t)g.s * TOT_STRIDE;
  const uint32_t* pre = a.gate ? a.cnt_g + (size_t)tile * GATED_BINS : a.cnt + (size_t)tile * a.cstride;
  const int rows = a.gate ? GATED_BINS : BINS;
  const uint32_t c_tot = tid < rows ? tot[tid] : 0u;
  const uint32_t c_pre = tid < rows ? pre[tid] : 0u;
  uint32_t keys[KPT], vals[KPT], rank[KPT];
  const uint32_t* keys_in = a.keys_in + g.off;
  const uint32_t* vals_in = VALS ? a.vals_in + g.off : nullptr;
  const uint32_t wbase = tile_base + (uint32_t)(w * WTILE) + (uint32_t)lane;
  // The loads take no branch, or the compiler could not count the loads in
  // flight and would wait for all of them before the first use: a lane past
  // the segment's count reads its last key again (an empty segment: a word
  // of the totals) and drops it. The pass's last reads of its input:
  // nontemporal.
  const uint32_t last = n ? n - 1u : 0u;
  const uint32_t* ksrc = n ? keys_in : tot;
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t i = wbase + (uint32_t)(j * RDX_WAVE);
    const uint32_t k = __builtin_nontemporal_load(ksrc + min(i, last));
    keys[j] = i < n ? k : 0u;
  }
  if constexpr (VALS) {
    const uint32_t* vsrc = n ? vals_in : tot;
    // (a lane past the count is never kept: its value is not read)
#pragma unroll
    for (int j = 0; j < KPT; j++) {
      const uint32_t i = wbase + (uint32_t)(j * RDX_WAVE);
      vals[j] = __builtin_nontemporal_load(vsrc + min(i, last));
    }
  } else {
#pragma unroll
    for (int j = 0; j < KPT; j++) vals[j] = wbase + (uint32_t)(j * RDX_WAVE);
  }

  // Digit starts of the segment: exclusive scan of the digit totals; with
  // the tile's prefix from the scanned counts, each digit's global start.
  uint32_t kept;
  const uint32_t ex = block_exclusive_sum<BLOCK>(c_tot, s_scan, &kept, tid);
  if (tid < BINS) s_out[tid] = g.off + ex + c_pre;
  const uint32_t sorted_count = a.drop || a.compacted ? kept : n;
  if (local_tile == 0 && tid == 0) {
    if (a.drop && a.nvalid) a.nvalid[g.s] = kept;
    if (a.count_out) a.count_out[g.s] = sorted_count;
    if (a.draw_args) {
      uint32_t* args = a.draw_args + 5 * g.s;
      args[0] = 6u * 1024u;
      args[1] = (sorted_count + 1023u) / 1024u;
      args[2] = 0u;
      args[3] = 0u;
      args[4] = 0u;
    }
  }
  if (tile_base >= n) return;  // uniform per block

  // Rank.
  uint32_t dig[KPT];
  __syncthreads();  // s_wave, s_match zeroed
  // Rank each round of 64 keys in input order. Up to PEEL digits of the
  // round (those of its first remaining lanes) are ranked by one ballot each.
  // The other lanes OR their bit into their digit's LDS mask, read the mask
  // back (the lanes of the round with that digit) and rank themselves by the
  // set lanes below them; the lowest adds the round's count to the wave's
  // digit count and clears the mask. A wave's LDS operations complete in
  // order, so the read sees every lane's OR.
  const uint64_t lane_bit = 1ull << lane;
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    const uint32_t i = wbase + (uint32_t)(j * RDX_WAVE);
    const bool keep = i < n && !(a.drop && keys[j] == a.invalid);
    rank[j] = keep ? 0u : 0xffffffffu;  // 0xffffffff: not kept
    const uint32_t d = digit_of(a, keys[j], kbase, DMASK);
    dig[j] = d;
    bool done = !keep;
    uint64_t rem = a.peel ? ballot(keep) : 0ull;
#pragma unroll
    for (int q = 0; q < PEEL; q++) {
      if (rem == 0ull) break;
      const int first = __builtin_ctzll(rem);
      const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)d, first);
      const uint64_t m = ballot(!done && d == d0);
      const uint32_t old0 = s_wave[w][d0];
      if (!done && d == d0) {
        rank[j] = old0 + count_below(m);
        done = true;
      }
      const uint32_t c = (uint32_t)__builtin_popcountll(m);
      if (lane == first) s_wave[w][d0] = old0 + c;
      rem &= ~m;
      if (c < PEEL_MIN) break;
    }
    if (!done) {
      __hip_atomic_fetch_or(&s_match[w][d], lane_bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const uint64_t m = __hip_atomic_load(&s_match[w][d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const uint32_t below = count_below(m);
      const uint32_t old = s_wave[w][d];
      if (below == 0u) {
        s_wave[w][d] = old + (uint32_t)__builtin_popcountll(m);
        __hip_atomic_store(&s_match[w][d], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      rank[j] = old + below;
    }
  }
  __syncthreads();

  // Per digit: exclusive offsets across the waves and the tile's count.
  for (int d = tid; d < BINS; d += BLOCK) {
    uint32_t run = 0;
#pragma unroll
    for (int v = 0; v < WAVES; v++) {
      const uint32_t c = s_wave[v][d];
      s_wave[v][d] = run;
      run += c;
    }
    s_local[d] = run;  // the count for now
  }
  __syncthreads();

  // Local starts of the digits within the tile (exclusive scan of the counts).
  uint32_t tile_kept = 0;
  for (int b0 = 0; b0 < BINS; b0 += BLOCK) {
    const int d = b0 + tid;
    const uint32_t c = d < BINS ? s_local[d] : 0u;
    uint32_t sum;
    const uint32_t ex = block_exclusive_sum<BLOCK>(c, s_scan, &sum, tid);
    if (d < BINS) {
      s_local[d] = tile_kept + ex;
      s_out[d] -= tile_kept + ex;
    }
    tile_kept += sum;
  }
  __syncthreads();

  // Stage in digit order.
#pragma unroll
  for (int j = 0; j < KPT; j++) {
    if (rank[j] != 0xffffffffu) {
      const uint32_t d = dig[j];
      const uint32_t p = s_local[d] + s_wave[w][d] + rank[j];
      s_stage[p] = (uint64_t)keys[j] | ((uint64_t)vals[j] << 32);  // one 8-byte LDS write per key
    }
  }
  __syncthreads();

  // Write out: runs of one digit go to consecutive addresses. Unrolled over
  // the KPT rounds so the LDS reads of several rounds are in flight at once.
  uint32_t ok[KPT], ov[KPT], od[KPT];
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    const uint32_t i = (uint32_t)(r * BLOCK + tid);
    const uint64_t kv = i < tile_kept ? s_stage[i] : 0ull;
    ok[r] = (uint32_t)kv;
    ov[r] = (uint32_t)(kv >> 32);
  }
#pragma unroll
  for (int r = 0; r < KPT; r++) od[r] = s_out[digit_of(a, ok[r], kbase, DMASK)] + (uint32_t)(r * BLOCK + tid);
#pragma unroll
  for (int r = 0; r < KPT; r++) {
    const uint32_t i = (uint32_t)(r * BLOCK + tid);
    if (i < tile_kept && od[r] < a.dst_cap) {  // always below when the counts are consistent
      if (keys_out) keys_out[od[r]] = ok[r];
      vals_out[od[r]] = ov[r];
    }
  }
  if (to_gated) {
    // Count each key's gated digit into the row of the tile it lands in; the
    // segment's last block to finish scans the rows for the gated pass.
#pragma unroll
    for (int r = 0; r < KPT; r++) {
      const uint32_t i = (uint32_t)(r * BLOCK + tid);
      if (i < tile_kept && od[r] < a.dst_cap) {
        const uint32_t t = g.tile0 + (od[r] - g.off) / (uint32_t)TILE;
        const uint32_t d = ((ok[r] & a.key_mask) - kbase) >> RANGED_BITS;
        atomicAdd(&a.cnt_g[(size_t)t * GATED_BINS + d], 1u);
      }
    }
    const uint32_t used = (uint32_t)(((uint64_t)n + TILE - 1) / TILE);  // the blocks that reach here
    wait_stores();
    __syncthreads();  // (also: every read of s_stage is done)
    if (tid == 0) L.last = atomicAdd(&a.done[g.s], 1u) == used - 1u ? 1u : 0u;
    __syncthreads();
    if (L.last) gated_scan<BLOCK>(a, g, used, reinterpret_cast<uint32_t*>(s_stage));
  }
}

template <int RB, int BLOCK, int KPT, bool VALS>
__global__ __launch_bounds__(BLOCK) void k_scatter(PassArgs a) {
  __shared__ ScatterLds<RB, BLOCK, KPT> L;
  scatter_tile<RB, BLOCK, KPT, VALS>(a, blockIdx.x, L, threadIdx.x);
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

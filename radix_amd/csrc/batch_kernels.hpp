// Batched projection: many views over many tables in one launch. A block
// takes 256 splats of one table and projects each of them for every view of
// that table in its group (up to BATCH_MAX_VIEWS), so a table is read once
// for all the views that draw it. Each view is a segment of the batch's keys.
#pragma once

#include "project_kernels.hpp"

namespace rdx {

constexpr int BATCH_MAX_VIEWS = 8;  // views per group

struct BatchView {
  EyeU eye;          // object (its table) -> eye
  double h[3];       // the eye in object space: the SH origin
  uint32_t seg_off;  // the view's first key in the batch
  uint32_t pad;
  uint4* words;      // its ewa words (NULL: not written)
};

struct BatchGroup {
  uint32_t table;       // table index
  uint32_t first_view;  // into the view list
  uint32_t num_views;   // <= BATCH_MAX_VIEWS
  uint32_t block0;      // first block of the group in its launch
  uint32_t chunk0;      // first chunk of the group in the persistent launch (PERSIST_BLOCK splats a chunk)
};

// Uniforms shared by every view of a batch (one rdx_params).
struct BatchU {
  int lod_inflate;
  float ndc_z_min, clip_xy;
  float max_std_dev, min_alpha, blur, pre_blur, min_pr, max_pr, min_contrib;
  float focal_distance, aperture_tan;
  int sort_bits, cull_keys;
};

struct BatchDesc {
  const BatchView* views;
  const BatchGroup* groups;
  const TableP* tables;
  const int* table_sh_degree;
  const float* table_enc;  // 4 per table
  int num_groups;
  uint32_t num_segments;   // the batch's views (the persistent form's scans)
};

// Splat i of group G's table, projected for each of the group's views: its
// ewa words and one key per view (view v's at its segment offset + i).
// The views' colours come first, each SH-evaluated from its own origin and
// kept encoded in LDS (s_rgb[v * BLOCK + thread]), so the SH coefficients
// are dead before the views' projections: the thread then holds no more
// than the plain projection does (the colour of a view that culls the
// splat is computed and not used). i >= the table's count: nothing.
template <int LAYOUT, int SHC, int DEG, int BLOCK>
__device__ __forceinline__ void project_batch_splat(const BatchDesc& D, const BatchU& u, uint32_t* keys,
                                                    const BatchGroup& G, uint32_t i, uint32_t* s_rgb) {
  const TableP& t = D.tables[G.table];
  if (i >= t.count) return;
  const int deg = DEG >= 0 ? DEG : D.table_sh_degree[G.table];
  const uint32_t nv = G.num_views;
  const float* enc = D.table_enc + 4 * G.table;
  const uint32_t invalid = key_invalid(u.sort_bits);

  F32Rsrc rs;
  if constexpr (LAYOUT == LAYOUT_F32) rs = f32_rsrc<SHC>(t, deg);

  Splat sp;
  float op, std_dev;
  float3 s;
  bool invisible;
  {
    float3 coef[15];
    float band[3];
    load_splat<LAYOUT, SHC, DEG>(t, rs, enc, i, deg, sp, coef, band);
    prepare<LAYOUT == LAYOUT_COV>(sp, u, op, s, std_dev, invisible);
    if (!invisible) {
      for (uint32_t v = 0; v < nv; v++) {
        const float3 rgb = sh_colour(sp, D.views[G.first_view + v].h, deg, coef, SHC == SHC_PACKED ? band : nullptr);
        s_rgb[v * BLOCK + threadIdx.x] = encode_ext_rgb(rgb.x, rgb.y, rgb.z);
      }
    }
  }
  float R[9];
  if constexpr (LAYOUT != LAYOUT_COV) {
    if (!invisible) quat_matrix(sp.q, R);
  }
  for (uint32_t v = 0; v < nv; v++) {
    const BatchView& V = D.views[G.first_view + v];
    uint4 words = make_uint4(0u, 0u, 0u, 0u);
    bool visible = false;
    EyeCentre C = eye_centre(V.eye, sp.c, u.ndc_z_min, u.clip_xy);
    C.pass = C.pass && !invisible;
    if (C.pass) {
      const uint32_t rgb_word = s_rgb[v * BLOCK + threadIdx.x];
      EwaResult r;
      if constexpr (LAYOUT == LAYOUT_COV) {
        r = ewa_eye_cov(V.eye, C, true, sp.cv, op, std_dev, rgb_word, u);
      } else {
        float Q[9];
        eye_axes(V.eye, R, Q);
        r = ewa_eye(V.eye, C, true, R, Q, s, op, std_dev, rgb_word, u);
      }
      if (r.ok) {
        words = r.words;
        visible = true;
      }
    }
    if (V.words) st_stream(V.words + i, words);
    // Mono views: the head is the eye, the key distance the eye's.
    const bool keyed = u.cull_keys ? visible : !invisible;
    const uint32_t key = keyed ? sort_key(fbits((float)C.dist), u.sort_bits) : invalid;
    keys[V.seg_off + i] = key;
  }
}

template <int LAYOUT, int SHC, int DEG>
__global__ __launch_bounds__(256) void k_project_batch(BatchDesc D, BatchU u, uint32_t* keys) {
  constexpr int BLOCK = 256;
  __shared__ uint32_t s_group;
  __shared__ uint32_t s_rgb[BATCH_MAX_VIEWS * BLOCK];

  // The block's group: the last group whose first block is <= blockIdx.x.
  if (threadIdx.x == 0) {
    uint32_t g = 0;
    for (int k = 1; k < D.num_groups; k++)
      if (D.groups[k].block0 <= blockIdx.x) g = (uint32_t)k;
    s_group = g;
  }
  __syncthreads();
  // uniform across the block: the group's and its views' fields load through
  // the scalar cache into SGPRs instead of once per lane
  const BatchGroup G = D.groups[__builtin_amdgcn_readfirstlane(s_group)];
  project_batch_splat<LAYOUT, SHC, DEG, BLOCK>(D, u, keys, G, (blockIdx.x - G.block0) * BLOCK + threadIdx.x, s_rgb);
}

// ---- the persistent form: projection and segmented sort in one launch ----------
//
// A grid of as many 512-thread blocks as are resident at once (a multiple
// of the 8 XCDs) runs the batch's phases in order, each block taking work
// items in turn (item b, b + grid, ...): the projection's chunks of
// PERSIST_BLOCK splats, then for each digit pass of the ranged sort (9-bit
// digits, 16 keys per thread: 8,192-key tiles) the tiles' counts, the
// segments' column scans and the tiles' scatters, with a grid-wide barrier
// between phases. A barrier releases the block's writes at agent scope
// (they reach the coherence point past the XCD's L2), counts the block in,
// waits for the grid's count, and acquires. Every wait is bounded by the
// wall clock: past PERSIST_WAIT_TICKS (about 1.3 s of the 100 MHz
// s_memrealtime) the block sets the error word and goes on, so a grid that
// is not all resident finishes with wrong output and the error set instead
// of hanging. The last block to finish resets the barrier words for the
// next launch (a graph replay).

constexpr int PERSIST_BLOCK = 512;
constexpr int PERSIST_KPT = 12;
constexpr uint64_t PERSIST_WAIT_TICKS = 1ull << 27;

constexpr int PERSIST_PHASES = 11;  // the projection, 3 x (count, scan, scatter), the gated scatter

struct PersistArgs {
  PassArgs pass[4];     // the ranged sort's passes (the fourth gated)
  uint32_t items;       // projection chunks of the kind's groups
  uint32_t blocks;      // the grid
  uint32_t* bar;        // [2] barrier arrivals and finished blocks, zero between launches
  uint32_t* err;        // set by a block whose wait ran out
  // Profiling (NULL: off): per block, the s_memrealtime of its start and of
  // its arrival at and departure from each barrier; [PERSIST_PHASES + 1][2][blocks]
  // (phase k starts at departure k and ends at arrival k + 1).
  uint64_t* stamps;
};

template <bool STAMPS>
__device__ __forceinline__ void persist_stamp(const PersistArgs& P, int k, int side) {
  if (STAMPS && threadIdx.x == 0)
    P.stamps[((size_t)k * 2 + side) * P.blocks + blockIdx.x] = __builtin_amdgcn_s_memrealtime();
}

// threadIdx.x through an opaque move, taken once per work item: the values
// a tile's scatter derives from the thread index (its keys' and outputs'
// offsets) are then computed per item instead of being hoisted out of the
// item loop and held in registers across it (165 VGPRs against 110).
__device__ __forceinline__ int item_tid() {
  int t;
  asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((int)threadIdx.x));
  return t;
}

template <bool STAMPS>
__device__ __forceinline__ void grid_sync(const PersistArgs& P, uint32_t& epoch) {
  __syncthreads();
  epoch++;
  persist_stamp<STAMPS>(P, (int)epoch, 0);
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(P.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t target = epoch * P.blocks;
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    while (__hip_atomic_load(P.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(P.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
      if (__builtin_amdgcn_s_memrealtime() - t0 > PERSIST_WAIT_TICKS) {
        __hip_atomic_store(P.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  persist_stamp<STAMPS>(P, (int)epoch, 1);
  __syncthreads();
}

// STAMPS: the profiling instantiation (P.stamps set), so that the stamps'
// registers never touch the production kernel's allocation.
template <int LAYOUT, int SHC, int DEG, bool STAMPS>
__global__ __launch_bounds__(PERSIST_BLOCK) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_batch_persistent(BatchDesc D, BatchU u, uint32_t* keys, PersistArgs P) {
  constexpr int RB = RANGED_RB, BLOCK = PERSIST_BLOCK, KPT = PERSIST_KPT;
  __shared__ union {
    CountLds<RB, BLOCK> count;
    ScanLds<BLOCK> scan;
    ScatterLds<RB, BLOCK, KPT> scatter;
    uint32_t rgb[BATCH_MAX_VIEWS * BLOCK];  // the projection's view colours
  } L;
  const uint32_t G = P.blocks;
  uint32_t epoch = 0;
  persist_stamp<STAMPS>(P, 0, 1);

  // projection: chunk c of the kind belongs to the last group whose first chunk is <= c
  for (uint32_t c = blockIdx.x; c < P.items; c += G) {
    uint32_t g = 0;
    for (int k = 1; k < D.num_groups; k++)
      if (D.groups[k].chunk0 <= c) g = (uint32_t)k;
    const BatchGroup Gr = D.groups[__builtin_amdgcn_readfirstlane(g)];
    project_batch_splat<LAYOUT, SHC, DEG, BLOCK>(D, u, keys, Gr, (c - Gr.chunk0) * BLOCK + threadIdx.x, L.rgb);
  }
  grid_sync<STAMPS>(P, epoch);

  const int bins = 1 << RB;
#pragma unroll 1
  for (int pass = 0; pass < 3; pass++) {
    const PassArgs& a = P.pass[pass];
    for (uint32_t t = blockIdx.x; t < a.tiles; t += G) {
      count_tile<RB, BLOCK, KPT>(a, t, L.count);
      __syncthreads();
    }
    grid_sync<STAMPS>(P, epoch);
    const uint32_t items = D.num_segments * scan_items_per_seg<BLOCK>(a.rows, a.ranged);
    for (uint32_t it = blockIdx.x; it < items; it += G) {
      scan_item<8, BLOCK>(a, bins, it, L.scan);
      __syncthreads();
    }
    grid_sync<STAMPS>(P, epoch);
    for (uint32_t t = blockIdx.x; t < a.tiles; t += G) {
      if (pass == 0)
        scatter_tile<RB, BLOCK, KPT, false>(a, t, L.scatter, item_tid());
      else
        scatter_tile<RB, BLOCK, KPT, true>(a, t, L.scatter, item_tid());
      __syncthreads();
    }
    grid_sync<STAMPS>(P, epoch);
  }
  // the gated fourth pass (its blocks return at once for segments within three passes)
  for (uint32_t t = blockIdx.x; t < P.pass[3].tiles; t += G) {
    scatter_tile<RB, BLOCK, KPT, true>(P.pass[3], t, L.scatter, item_tid());
    __syncthreads();
  }
  __syncthreads();
  persist_stamp<STAMPS>(P, PERSIST_PHASES, 0);
  // the last block out resets the barrier words
  if (threadIdx.x == 0) {
    const uint32_t out = __hip_atomic_fetch_add(P.bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (out == G - 1u) {
      __hip_atomic_store(P.bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(P.bar + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

}  // namespace rdx

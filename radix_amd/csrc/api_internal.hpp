// Host-side internals shared by the C ABI's translation units: device
// buffers, the context's scratch, sort plans and the launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/radix_hip.h"
#include "sort_kernels.hpp"

namespace rdx {

struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need_bytes);  // grows (contents not kept)
  void release();
};

struct SortPlan {
  int rb, block, kpt, tile, passes, bits;
  bool ranged = false;  // digits of key - base (compacting 32-bit sorts; the last pass gated)
};

struct StageRecord {
  std::string name;
  hipEvent_t start, end;
};

// Words of a ranged sort's own scratch: [tiles][2] then [segments][2]
// smallest and largest kept key, [segments] gated-pass tickets,
// [tiles][GATED_BINS] gated-pass tile rows and [segments][GATED_BINS]
// gated-pass digit totals.
constexpr size_t ranged_words(size_t tiles, size_t segments) {
  return 2 * tiles + 3 * (segments ? segments : 1) + (size_t)GATED_BINS * (tiles + (segments ? segments : 1));
}

// Device scratch of sorts: the passes' alternating (key, value) buffers
// and the tile counts, digit totals and kept counts.
struct SortScratch {
  Buffer keys[2], vals[2];
  Buffer cnt;     // [tiles][count_stride(MAX_BINS + 1)]
  Buffer tot;     // [segments][TOT_STRIDE]
  Buffer nvalid;  // [segments]
  Buffer mm;      // ranged sorts: ranged_words(tiles, segments)
  Buffer seg;     // segment table [segments][4] and tile map [tiles] (small segmented sorts)
  int reserve(uint32_t n, uint32_t segments, uint32_t tiles);
  void release() {
    for (int i = 0; i < 2; i++) {
      keys[i].release();
      vals[i].release();
    }
    cnt.release();
    tot.release();
    nvalid.release();
    mm.release();
    seg.release();
  }
};

struct Context {
  rdx_sort_tuning tuning;
  bool profiling = false;
  std::vector<StageRecord> records;  // recorded, not yet read
  std::vector<hipEvent_t> event_pool;
  SortScratch sort;
  Buffer io_keys, io_vals;  // host-array sorts: device copies
  Buffer proj;              // projection scratch (keys of the fused path)
  // The stream of host-array sorts (rdx_sort_keys), made on first use: the
  // sorts of contexts on different threads overlap on the device.
  hipStream_t host_stream = nullptr;
  void release() {
    if (host_stream) (void)hipStreamDestroy(host_stream);
    host_stream = nullptr;
    sort.release();
    io_keys.release();
    io_vals.release();
    proj.release();
    for (auto& r : records) {
      (void)hipEventDestroy(r.start);
      (void)hipEventDestroy(r.end);
    }
    records.clear();
    for (auto e : event_pool) (void)hipEventDestroy(e);
    event_pool.clear();
  }
  hipEvent_t take_event();
};

// Segments of a sort on the device (NULL tables: one segment of all tiles).
struct SegLayout {
  const uint32_t* seg_of_tile = nullptr;
  const uint32_t* seg = nullptr;  // [segments][4]: tile0, off, n, ntiles
  uint32_t segments = 1;
  uint32_t tiles = 1;
};

// Brackets the kernels launched in its scope with events when profiling.
struct Stage {
  Context* ctx;
  hipStream_t stream;
  StageRecord rec;
  bool on;
  Stage(Context* c, hipStream_t s, const char* name);
  ~Stage();
};

// The tuning cache: a sort tuning and its time per (gfx arch, form, key bits,
// ceil(log2) of the keys, ceil(log2) of the segments), kept for the process
// and, when a cache file is set (rdx_tuning_cache_file or the environment's
// RDX_TUNING_CACHE), read from and appended to that file.
struct TuneKey {
  std::string arch, form;
  int bits;
  int n_bucket, seg_bucket;
};
TuneKey tune_key(const char* form, uint64_t n, uint32_t segments, int bits);
bool tuning_cache_get(const TuneKey& k, rdx_sort_tuning* t, double* ms);
void tuning_cache_put(const TuneKey& k, const rdx_sort_tuning& t, double ms);

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int check_bits(int bits, const char* who);
SortPlan make_plan(const rdx_sort_tuning& t, int bits);
// The plan of a compacting sort (invalid keys dropped): ranged (9-bit
// digits of key - base, three passes plus one gated) for 32-bit keys under
// an 8-bit, 512-thread tuning; make_plan otherwise.
SortPlan compact_plan(const rdx_sort_tuning& t, int bits);
bool tuning_supported(const rdx_sort_tuning& t);
// Tiles of a segment of n keys (at least one, which writes its counts).
uint32_t tiles_of(const SortPlan& p, uint32_t n);
// A segment table of up to 8 segments written on the device by a kernel
// (safe under stream capture) into the scratch's seg buffer.
constexpr int SMALL_SEGMENTS = 8;
int small_segments(SortScratch& sc, hipStream_t s, const SortPlan& p, const uint32_t* off, const uint32_t* n, int count,
                   SegLayout* out);
// The arguments of digit pass `pass` of the sort run_sort runs: kin / vin
// are the pass's input (the sort's input for pass 0, the previous pass's
// keys_out / vals_out after it).
PassArgs pass_args(SortScratch& sc, const SortPlan& p, const SegLayout& L, int pass, const uint32_t* kin,
                   const uint32_t* vin, uint32_t* final_keys, uint32_t* final_vals, uint32_t n_cap, const uint32_t* n_dev,
                   bool drop, uint32_t invalid, bool rule, uint32_t* count_out, uint32_t* draw_args);
// The P digit passes: pass 0 reads (keys_in, vals_in or the index), the last
// writes (final_keys, final_vals); between them the scratch pairs alternate.
// rule: count the host sort's hidden keys into tot[s][MAX_BINS].
int run_sort(Context* ctx, SortScratch& sc, hipStream_t s, const SortPlan& p, const SegLayout& L,
             const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* final_keys, uint32_t* final_vals, uint32_t n_cap,
             const uint32_t* n_dev, bool drop, uint32_t invalid, bool rule, uint32_t* count_out,
             uint32_t* draw_args);

}  // namespace rdx

// Host-side internals shared by the C ABI's translation units: device
// buffers, the context's scratch, sort plans and the launch helpers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/radix_hip.h"

namespace rdx {

struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need_bytes);  // grows (contents not kept)
  void release();
};

struct SortPlan {
  int rb, block, kpt, tile, passes, bits;
};

struct StageRecord {
  std::string name;
  hipEvent_t start, end;
};

struct Context {
  rdx_sort_tuning tuning;
  bool profiling = false;
  std::vector<StageRecord> records;  // recorded, not yet read
  std::vector<hipEvent_t> event_pool;
  Buffer keys[2], vals[2];  // the passes' alternating scratch pairs
  Buffer hist;              // digit histograms, tickets, counters
  Buffer status;            // two look-back status regions
  size_t status_rows = 0;   // rows per region
  Buffer io_keys, io_vals;  // host-array sorts: device copies
  Buffer proj;              // projection scratch (keys of the fused path)
  int reserve(const SortPlan& p, uint32_t n, int segments);
  void release() {
    for (int i = 0; i < 2; i++) {
      keys[i].release();
      vals[i].release();
    }
    hist.release();
    status.release();
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

// Brackets the kernels launched in its scope with events when profiling.
struct Stage {
  Context* ctx;
  hipStream_t stream;
  StageRecord rec;
  bool on;
  Stage(Context* c, hipStream_t s, const char* name);
  ~Stage();
};

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
int check_bits(int bits, const char* who);
SortPlan make_plan(const rdx_sort_tuning& t, int bits);
bool tuning_supported(const rdx_sort_tuning& t);
uint32_t pass_grid(const SortPlan& p, uint32_t n_cap);
void launch_clear(Context* ctx, hipStream_t s, uint32_t* hist, uint32_t rows);
void launch_histogram(Context* ctx, hipStream_t s, const SortPlan& p, const uint32_t* keys, const uint32_t* n_dev, uint32_t n_cap,
                      uint32_t* hist, int drop, uint32_t invalid);
int run_passes(Context* ctx, hipStream_t s, const SortPlan& p, const uint32_t* keys_in, const uint32_t* vals_in,
               uint32_t* final_keys, uint32_t* final_vals, uint32_t n_cap, const uint32_t* n_dev, bool drop,
               uint32_t invalid, uint32_t* hist, uint32_t* count_out, uint32_t* draw_args);

}  // namespace rdx

// The C ABI of include/radix_hip.h: contexts, the sorts, the projection and
// the fused project + sort, each a short stream-ordered sequence of the
// kernels in sort_kernels.hpp and project_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "sort_kernels.hpp"

using namespace rdx;

namespace {
thread_local std::string g_error;
}

namespace rdx {

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(RDX_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---- device buffers --------------------------------------------------------

int Buffer::ensure(size_t need_bytes) {
  if (need_bytes <= bytes) return RDX_OK;
  if (ptr) {
    hipError_t e = hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    if (e != hipSuccess) return hip_fail(e, "hipFree");
  }
  size_t grow = std::max(need_bytes, (size_t)4096);
  hipError_t e = hipMalloc(&ptr, grow);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    return fail(RDX_E_NOMEM, "device allocation of %zu bytes failed: %s", grow, hipGetErrorString(e));
  }
  bytes = grow;
  return RDX_OK;
}

void Buffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

// ---- stage timing ---------------------------------------------------------------

hipEvent_t Context::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

Stage::Stage(Context* c, hipStream_t s, const char* name) : ctx(c), stream(s), on(c && c->profiling) {
  if (!on) return;
  rec.name = name;
  rec.start = ctx->take_event();
  rec.end = ctx->take_event();
  if (!rec.start || !rec.end) {
    on = false;
    return;
  }
  (void)hipEventRecord(rec.start, stream);
}

Stage::~Stage() {
  if (!on) return;
  (void)hipEventRecord(rec.end, stream);
  ctx->records.push_back(rec);
}

// ---- sort plans ---------------------------------------------------------------

SortPlan make_plan(const rdx_sort_tuning& t, int bits) {
  SortPlan p;
  p.rb = t.digit_bits;
  p.block = t.block_threads;
  p.kpt = t.keys_per_thread;
  p.tile = p.block * p.kpt;
  p.passes = (bits + p.rb - 1) / p.rb;
  p.bits = bits;
  return p;
}

SortPlan compact_plan(const rdx_sort_tuning& t, int bits) {
  SortPlan p = make_plan(t, bits);
  if (bits == 32 && t.digit_bits == 8 && t.block_threads == 512) {
    p.rb = RANGED_RB;
    p.passes = 4;  // the fourth gated on the range
    p.ranged = true;
  }
  return p;
}

bool tuning_supported(const rdx_sort_tuning& t) {
  const bool rb = t.digit_bits == 4 || t.digit_bits == 6 || t.digit_bits == 8;
  const bool shape = (t.block_threads == 256 && (t.keys_per_thread == 8 || t.keys_per_thread == 12 ||
                                                 t.keys_per_thread == 16 || t.keys_per_thread == 20)) ||
                     (t.block_threads == 512 && (t.keys_per_thread == 8 || t.keys_per_thread == 12 ||
                                                 t.keys_per_thread == 16));
  return rb && shape;
}

// ---- launches of the pass kernels ---------------------------------------------------

enum class PassKernel { count, scatter };

// The count of a 512 x 16 tile runs as 1024 threads x 8 keys (the same
// 8,192-key tile; twice the waves in flight per block: measured 0.3 % off a
// plain_4m step, in interleaved runs on one box).
template <int RB, int BLOCK, int KPT>
static void launch_kpass(PassKernel k, uint32_t grid, hipStream_t s, const PassArgs& a) {
  if (k == PassKernel::count) {
    if constexpr (BLOCK == 512 && KPT == 16)
      hipLaunchKernelGGL((k_count<RB, 1024, 8>), dim3(grid), dim3(1024), 0, s, a);
    else
      hipLaunchKernelGGL((k_count<RB, BLOCK, KPT>), dim3(grid), dim3(BLOCK), 0, s, a);
  } else if (a.vals_in)
    hipLaunchKernelGGL((k_scatter<RB, BLOCK, KPT, true>), dim3(grid), dim3(BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL((k_scatter<RB, BLOCK, KPT, false>), dim3(grid), dim3(BLOCK), 0, s, a);
}

// 256-thread blocks for digits of up to 8 bits (a thread per digit), 512 for all.
template <int RB>
static bool launch_rb(const SortPlan& p, PassKernel k, uint32_t grid, hipStream_t s, const PassArgs& a) {
  if constexpr (RB <= 8) {
    if (p.block == 256) {
      switch (p.kpt) {
        case 8: launch_kpass<RB, 256, 8>(k, grid, s, a); return true;
        case 12: launch_kpass<RB, 256, 12>(k, grid, s, a); return true;
        case 16: launch_kpass<RB, 256, 16>(k, grid, s, a); return true;
        case 20: launch_kpass<RB, 256, 20>(k, grid, s, a); return true;
      }
      return false;
    }
  }
  if (p.block == 512) {
    switch (p.kpt) {
      case 8: launch_kpass<RB, 512, 8>(k, grid, s, a); return true;
      case 12: launch_kpass<RB, 512, 12>(k, grid, s, a); return true;
      case 16: launch_kpass<RB, 512, 16>(k, grid, s, a); return true;
    }
  }
  return false;
}

static bool launch_pass_kernel(const SortPlan& p, PassKernel k, uint32_t grid, hipStream_t s, const PassArgs& a) {
  switch (p.rb) {
    case 4: return launch_rb<4>(p, k, grid, s, a);
    case 6: return launch_rb<6>(p, k, grid, s, a);
    case 8: return launch_rb<8>(p, k, grid, s, a);
    case RANGED_RB: return launch_rb<RANGED_RB>(p, k, grid, s, a);
  }
  return false;
}

uint32_t tiles_of(const SortPlan& p, uint32_t n) {
  return std::max<uint32_t>(1u, (uint32_t)(((uint64_t)n + p.tile - 1) / p.tile));
}

int SortScratch::reserve(uint32_t n, uint32_t segments, uint32_t tiles) {
  const size_t words = std::max<size_t>(n, 1);
  int rc;
  for (int i = 0; i < 2; i++) {
    if ((rc = keys[i].ensure(words * 4))) return rc;
    if ((rc = vals[i].ensure(words * 4))) return rc;
  }
  if ((rc = cnt.ensure((size_t)count_stride(MAX_BINS + 1) * tiles * 4))) return rc;
  if ((rc = tot.ensure((size_t)TOT_STRIDE * std::max<uint32_t>(segments, 1) * 4))) return rc;
  if ((rc = nvalid.ensure((size_t)std::max<uint32_t>(segments, 1) * 4))) return rc;
  if ((rc = mm.ensure(ranged_words(tiles, segments) * 4))) return rc;
  return RDX_OK;
}

struct SmallSegs {
  uint32_t count;
  uint32_t tile0[SMALL_SEGMENTS], off[SMALL_SEGMENTS], n[SMALL_SEGMENTS], ntiles[SMALL_SEGMENTS];
};

static __global__ void k_small_segments(SmallSegs q, uint32_t* seg, uint32_t* seg_of_tile, uint32_t tiles) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < q.count) {
    seg[4 * i] = q.tile0[i];
    seg[4 * i + 1] = q.off[i];
    seg[4 * i + 2] = q.n[i];
    seg[4 * i + 3] = q.ntiles[i];
  }
  for (uint32_t t = i; t < tiles; t += gridDim.x * blockDim.x) {
    uint32_t s = 0;
    for (uint32_t k = 1; k < q.count; k++)
      if (q.tile0[k] <= t) s = k;
    seg_of_tile[t] = s;
  }
}

int small_segments(SortScratch& sc, hipStream_t s, const SortPlan& p, const uint32_t* off, const uint32_t* n, int count,
                   SegLayout* out) {
  if (count < 1 || count > SMALL_SEGMENTS) return fail(RDX_E_INVALID, "small_segments: %d segments", count);
  SmallSegs q;
  memset(&q, 0, sizeof q);
  q.count = (uint32_t)count;
  uint32_t tiles = 0;
  for (int k = 0; k < count; k++) {
    q.tile0[k] = tiles;
    q.off[k] = off[k];
    q.n[k] = n[k];
    q.ntiles[k] = tiles_of(p, n[k]);
    tiles += q.ntiles[k];
  }
  if (int rc = sc.seg.ensure((size_t)(4 * SMALL_SEGMENTS + tiles) * 4)) return rc;
  uint32_t* seg = (uint32_t*)sc.seg.ptr;
  const uint32_t grid = std::min<uint32_t>((tiles + 255) / 256, 1024u);
  hipLaunchKernelGGL(k_small_segments, dim3(grid), dim3(256), 0, s, q, seg, seg + 4 * SMALL_SEGMENTS, tiles);
  out->seg = seg;
  out->seg_of_tile = seg + 4 * SMALL_SEGMENTS;
  out->segments = (uint32_t)count;
  out->tiles = tiles;
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "segment table");
}

PassArgs pass_args(SortScratch& sc, const SortPlan& p, const SegLayout& L, int pass, const uint32_t* kin,
                   const uint32_t* vin, uint32_t* final_keys, uint32_t* final_vals, uint32_t n_cap, const uint32_t* n_dev,
                   bool drop, uint32_t invalid, bool rule, uint32_t* count_out, uint32_t* draw_args) {
  const int bins = 1 << p.rb;
  // the ranged sort's words (ranged_words): tile and segment ranges, the
  // gated pass's done tickets, tile rows and digit totals
  const size_t nseg = std::max<uint32_t>(L.segments, 1);
  uint32_t* tile_mm = (uint32_t*)sc.mm.ptr;
  uint32_t* seg_mm = tile_mm + 2 * (size_t)L.tiles;
  uint32_t* done = seg_mm + 2 * nseg;
  uint32_t* cnt_g = done + nseg;
  uint32_t* tot_g = cnt_g + (size_t)GATED_BINS * L.tiles;
  const bool last = pass == p.passes - 1;
  PassArgs a;
  memset(&a, 0, sizeof a);
  a.keys_in = kin;
  a.vals_in = vin;
  a.keys_out = last ? final_keys : (uint32_t*)sc.keys[pass & 1].ptr;
  a.vals_out = last ? final_vals : (uint32_t*)sc.vals[pass & 1].ptr;
  a.n_dev = n_dev;
  a.n = n_cap;
  a.cnt = (uint32_t*)sc.cnt.ptr;
  a.tot = (uint32_t*)sc.tot.ptr;
  a.nvalid = (uint32_t*)sc.nvalid.ptr;
  a.tiles = L.tiles;
  a.rows = bins + (rule && pass == 0 ? 1 : 0);
  a.cstride = count_stride(a.rows);
  a.shift = pass * p.rb;
  a.key_mask = key_invalid(p.bits);
  a.invalid = invalid;
  a.drop = pass == 0 && drop;
  a.compacted = pass > 0 && drop;
  a.bits = p.bits;
  a.count_out = last ? count_out : nullptr;
  a.draw_args = last ? draw_args : nullptr;
  a.seg_of_tile = L.seg_of_tile;
  a.seg = L.seg;
  a.dst_cap = n_cap;
  a.tile_keys = (uint32_t)p.tile;
  // the top digit (skewed for distance keys) peels its common digits by ballots
  a.peel = pass == p.passes - 1 || (p.ranged && pass == 2);
  if (p.ranged) {
    // pass 0 -> scratch 0, 1 -> scratch 1, 2 -> the output (scratch 0 for
    // segments whose keys reach beyond three passes), 3 (those segments
    // only): scratch 0 -> the output. The kept count is final after pass 0.
    a.ranged = pass == 0 ? 1 : 2;
    a.tile_mm = tile_mm;
    a.seg_mm = seg_mm;
    a.gate = pass == 3;
    a.alt = pass == 2;
    a.zero_g = pass == 1;
    a.cnt_g = cnt_g;
    a.tot_g = tot_g;
    a.done = done;
    if (pass < 2) {
      a.keys_out = (uint32_t*)sc.keys[pass].ptr;
      a.vals_out = (uint32_t*)sc.vals[pass].ptr;
    } else {
      a.keys_out = final_keys;
      a.vals_out = final_vals;
    }
    if (pass == 2) {
      a.keys_alt = (uint32_t*)sc.keys[0].ptr;
      a.vals_alt = (uint32_t*)sc.vals[0].ptr;
    }
    if (pass == 3) {
      a.keys_in = (const uint32_t*)sc.keys[0].ptr;
      a.vals_in = (const uint32_t*)sc.vals[0].ptr;
    }
    a.count_out = pass == 0 ? count_out : nullptr;
    a.draw_args = pass == 0 ? draw_args : nullptr;
  }
  return a;
}

int run_sort(Context* ctx, SortScratch& sc, hipStream_t s, const SortPlan& p, const SegLayout& L,
             const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* final_keys, uint32_t* final_vals, uint32_t n_cap,
             const uint32_t* n_dev, bool drop, uint32_t invalid, bool rule, uint32_t* count_out,
             uint32_t* draw_args) {
  const int bins = 1 << p.rb;
  const uint32_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  for (int pass = 0; pass < p.passes; pass++) {
    const PassArgs a = pass_args(sc, p, L, pass, kin, vin, final_keys, final_vals, n_cap, n_dev, drop, invalid, rule,
                                 count_out, draw_args);
    char name[32];
    // the gated pass is a scatter alone: the third pass counted and scanned its digits
    if (!a.gate) {
      snprintf(name, sizeof name, "sort_count%d", pass);
      Stage st(ctx, s, name);
      if (!launch_pass_kernel(p, PassKernel::count, L.tiles, s, a)) return fail(RDX_E_INVALID, "unsupported sort tuning");
    }
    if (!a.gate) {
      snprintf(name, sizeof name, "sort_scan%d", pass);
      Stage st(ctx, s, name);
      // column groups per segment, and the range reduction in a ranged sort's first pass
      const uint32_t per_seg = (uint32_t)(a.rows + CGROUP - 1) / CGROUP + (a.ranged == 1 ? 1u : 0u);
      // 32 tiles per thread and round from 2048 tiles (one round for a
      // segment of up to 16M keys at 8192-key tiles), else 8
      if (L.tiles >= 2048)
        hipLaunchKernelGGL(k_scan<32>, dim3(L.segments * per_seg), dim3(SCAN_BLOCK), 0, s, a, bins);
      else
        hipLaunchKernelGGL(k_scan<8>, dim3(L.segments * per_seg), dim3(SCAN_BLOCK), 0, s, a, bins);
    }
    {
      snprintf(name, sizeof name, "sort_scatter%d", pass);
      Stage st(ctx, s, name);
      if (!launch_pass_kernel(p, PassKernel::scatter, L.tiles, s, a)) return fail(RDX_E_INVALID, "unsupported sort tuning");
    }
    kin = a.keys_out;
    vin = a.vals_out;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "radix sort launch");
  return RDX_OK;
}

int check_bits(int bits, const char* who) {
  if (bits != 16 && bits != 24 && bits != 32) return fail(RDX_E_INVALID, "%s: bits must be 16, 24 or 32 (got %d)", who, bits);
  return RDX_OK;
}

}  // namespace rdx

// ============================================================================
// extern "C"
// ============================================================================

extern "C" {

const char* rdx_last_error(void) { return g_error.c_str(); }

const char* rdx_version(void) { return "radix_hip 0.1 gfx950"; }

uint32_t rdx_key_invalid(int bits) { return key_invalid(bits); }

uint32_t rdx_key_valid_max(int bits) { return (uint32_t)(0xffull << (bits - 8)) - 1u; }

uint32_t rdx_key_compute(float distance, int bits) {
  uint32_t f;
  memcpy(&f, &distance, 4);
  const uint32_t mask = key_invalid(bits);
  if (!(f < 0x7f800000u)) return mask;
  uint32_t q;
  if (bits == 32) {
    q = f;
  } else if (bits == 24) {
    q = f >> 7;
  } else {
    const uint32_t lo = 95u << 23, hi = (159u << 23) - 1u;
    q = (std::min(std::max(f, lo), hi) - lo) >> 13;
  }
  return std::min(mask ^ q, rdx_key_valid_max(bits));
}

int rdx_compute_keys(void* stream, const float* d, uint32_t n, int bits, uint32_t* keys) {
  if (int rc = check_bits(bits, "SortKeys.compute")) return rc;
  if (n == 0) return RDX_OK;
  const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)n + 255) / 256, 4096u);
  hipLaunchKernelGGL(k_compute_keys, dim3(grid), dim3(256), 0, (hipStream_t)stream, d, n, bits, keys);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_compute_keys");
}

int rdx_context_create(rdx_context** out) {
  if (!out) return fail(RDX_E_INVALID, "rdx_context_create: out is NULL");
  Context* c = new Context();
  c->tuning.digit_bits = 8;
  c->tuning.keys_per_thread = 16;
  c->tuning.block_threads = 512;
  *out = reinterpret_cast<rdx_context*>(c);
  return RDX_OK;
}

int rdx_context_destroy(rdx_context* ctx) {
  if (!ctx) return RDX_OK;
  Context* c = reinterpret_cast<Context*>(ctx);
  c->release();
  delete c;
  return RDX_OK;
}

int rdx_context_reserve(rdx_context* ctx, uint32_t max_keys) {
  if (!ctx) return fail(RDX_E_INVALID, "rdx_context_reserve: no context");
  Context* c = reinterpret_cast<Context*>(ctx);
  return c->sort.reserve(max_keys, 1, tiles_of(make_plan(c->tuning, 32), max_keys));
}

int rdx_context_set_tuning(rdx_context* ctx, const rdx_sort_tuning* t) {
  if (!ctx || !t) return fail(RDX_E_INVALID, "rdx_context_set_tuning: NULL argument");
  Context* c = reinterpret_cast<Context*>(ctx);
  rdx_sort_tuning v = *t;
  if (!v.digit_bits) v.digit_bits = 8;
  if (!v.keys_per_thread) v.keys_per_thread = 16;
  if (!v.block_threads) v.block_threads = 512;
  if (!tuning_supported(v))
    return fail(RDX_E_INVALID, "unsupported sort tuning: %d-bit digits, %d keys per thread, %d threads", v.digit_bits,
                v.keys_per_thread, v.block_threads);
  c->tuning = v;
  return RDX_OK;
}

int rdx_context_get_tuning(const rdx_context* ctx, rdx_sort_tuning* out) {
  if (!ctx || !out) return fail(RDX_E_INVALID, "rdx_context_get_tuning: NULL argument");
  *out = reinterpret_cast<const Context*>(ctx)->tuning;
  return RDX_OK;
}

static __global__ void k_fill_u32(uint32_t* dst, uint32_t value, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = value;
}

int rdx_device_alloc(size_t bytes, void** out) {
  if (!out) return fail(RDX_E_INVALID, "rdx_device_alloc: out is NULL");
  *out = nullptr;
  hipError_t e = hipMalloc(out, std::max<size_t>(bytes, 4));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(RDX_E_NOMEM, "device allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  return RDX_OK;
}

int rdx_device_free(void* ptr) {
  if (!ptr) return RDX_OK;
  hipError_t e = hipFree(ptr);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_device_free");
}

int rdx_device_upload(void* stream, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_device_upload");
}

int rdx_device_download(void* stream, void* dst, const void* src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_device_download");
}

int rdx_device_fill_u32(void* stream, uint32_t* dst, uint32_t value, size_t count) {
  if (count == 0) return RDX_OK;
  const uint32_t grid = (uint32_t)std::min<size_t>((count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid), dim3(256), 0, (hipStream_t)stream, dst, value, count);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_device_fill_u32");
}

int rdx_stream_create(void** out) {
  if (!out) return fail(RDX_E_INVALID, "rdx_stream_create: out is NULL");
  hipStream_t s = nullptr;
  hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "rdx_stream_create");
  *out = s;
  return RDX_OK;
}

int rdx_stream_destroy(void* stream) {
  if (!stream) return RDX_OK;
  hipError_t e = hipStreamDestroy((hipStream_t)stream);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_stream_destroy");
}

int rdx_stream_synchronize(void* stream) {
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_stream_synchronize");
}

int rdx_context_profile(rdx_context* ctx, int enable) {
  if (!ctx) return fail(RDX_E_INVALID, "rdx_context_profile: no context");
  reinterpret_cast<Context*>(ctx)->profiling = enable != 0;
  return RDX_OK;
}

int rdx_context_stage_times(rdx_context* ctx, char* names, double* ms, int* launches, int max_stages) {
  if (!ctx) return fail(RDX_E_INVALID, "rdx_context_stage_times: no context");
  Context* c = reinterpret_cast<Context*>(ctx);
  std::vector<std::string> order;
  std::vector<double> total;
  std::vector<int> count;
  for (auto& r : c->records) {
    hipError_t e = hipEventSynchronize(r.end);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, r.start, r.end);
    if (e != hipSuccess) return hip_fail(e, "rdx_context_stage_times");
    size_t k = 0;
    while (k < order.size() && order[k] != r.name) k++;
    if (k == order.size()) {
      order.push_back(r.name);
      total.push_back(0.0);
      count.push_back(0);
    }
    total[k] += t;
    count[k] += 1;
    c->event_pool.push_back(r.start);
    c->event_pool.push_back(r.end);
  }
  c->records.clear();
  const int n = std::min<int>((int)order.size(), max_stages);
  for (int k = 0; k < n; k++) {
    if (names) {
      memset(names + 32 * k, 0, 32);
      strncpy(names + 32 * k, order[k].c_str(), 31);
    }
    if (ms) ms[k] = total[k];
    if (launches) launches[k] = count[k];
  }
  return n;
}

int rdx_sort_keys(rdx_context* ctx, const uint32_t* keys, uint32_t count, int bits, uint32_t* ordering,
                  size_t ordering_len, uint32_t* visible) {
  if (!ctx) return fail(RDX_E_INVALID, "SortKeys.sort: no context");
  if (ordering_len < count)
    return fail(RDX_E_CAPACITY, "SortKeys.sort: ordering holds %zu of %u indices", ordering_len, count);
  if (int rc = check_bits(bits, "SortKeys.sort")) return rc;
  Context* c = reinterpret_cast<Context*>(ctx);
  if (count == 0) {
    if (visible) *visible = 0;
    return RDX_OK;
  }
  const SortPlan p = make_plan(c->tuning, bits);
  SegLayout L;
  L.tiles = tiles_of(p, count);
  int rc;
  if ((rc = c->sort.reserve(count, 1, L.tiles))) return rc;
  if ((rc = c->io_keys.ensure((size_t)count * 4))) return rc;
  if ((rc = c->io_vals.ensure((size_t)count * 4))) return rc;
  if (!c->host_stream) {
    hipError_t se = hipStreamCreateWithFlags(&c->host_stream, hipStreamNonBlocking);
    if (se != hipSuccess) {
      c->host_stream = nullptr;
      return hip_fail(se, "SortKeys.sort: stream");
    }
  }
  hipStream_t s = c->host_stream;
  uint32_t* dkeys = (uint32_t*)c->io_keys.ptr;
  uint32_t* dord = (uint32_t*)c->io_vals.ptr;
  hipError_t e = hipMemcpyAsync(dkeys, keys, (size_t)count * 4, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hip_fail(e, "SortKeys.sort upload");
  if ((rc = run_sort(c, c->sort, s, p, L, dkeys, nullptr, nullptr, dord, count, nullptr, false, 0u, true, nullptr,
                     nullptr)))
    return rc;
  uint32_t hidden = 0;
  e = hipMemcpyAsync(ordering, dord, (size_t)count * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&hidden, (uint32_t*)c->sort.tot.ptr + MAX_BINS, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "SortKeys.sort download");
  if (visible) *visible = count - hidden;
  return RDX_OK;
}

static int sort_pairs_impl(rdx_context* ctx, void* stream, uint32_t* keys, uint32_t* vals, const uint32_t* count_dev,
                           uint32_t n, int bits) {
  if (!ctx) return fail(RDX_E_INVALID, "WgpuRadixSort: no context");
  if (int rc = check_bits(bits, "WgpuRadixSort")) return rc;
  Context* c = reinterpret_cast<Context*>(ctx);
  if (n == 0) return RDX_OK;
  const SortPlan p = make_plan(c->tuning, bits);
  SegLayout L;
  L.tiles = tiles_of(p, n);
  if (int rc = c->sort.reserve(n, 1, L.tiles)) return rc;
  return run_sort(c, c->sort, (hipStream_t)stream, p, L, keys, vals, keys, vals, n, count_dev, false, 0u, false, nullptr,
                  nullptr);
}

int rdx_sort_pairs(rdx_context* ctx, void* stream, uint32_t* keys, uint32_t* vals, uint32_t count, int bits) {
  return sort_pairs_impl(ctx, stream, keys, vals, nullptr, count, bits);
}

int rdx_sort_pairs_indirect(rdx_context* ctx, void* stream, uint32_t* keys, uint32_t* vals, const uint32_t* count_dev,
                            uint32_t capacity, int bits) {
  if (!count_dev) return fail(RDX_E_INVALID, "WgpuRadixSort.sortIndirect: no count buffer");
  return sort_pairs_impl(ctx, stream, keys, vals, count_dev, capacity, bits);
}

int rdx_sort_splats(rdx_context* ctx, void* stream, const uint32_t* keys, uint32_t n, int bits, uint32_t* out_indices,
                    uint32_t* out_count, uint32_t* out_args) {
  if (!ctx) return fail(RDX_E_INVALID, "WgpuSplatSorter: no context");
  if (int rc = check_bits(bits, "WgpuRadixSort")) return rc;
  Context* c = reinterpret_cast<Context*>(ctx);
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    hipLaunchKernelGGL(k_empty_outputs, dim3(1), dim3(64), 0, s, out_count, out_args);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RDX_OK : hip_fail(e, "WgpuSplatSorter");
  }
  const SortPlan p = compact_plan(c->tuning, bits);
  SegLayout L;
  L.tiles = tiles_of(p, n);
  if (int rc = c->sort.reserve(n, 1, L.tiles)) return rc;
  return run_sort(c, c->sort, s, p, L, keys, nullptr, nullptr, out_indices, n, nullptr, true, key_invalid(bits), false,
                  out_count, out_args);
}

// ---- the splat sorter without allocation ----------------------------------------

// Scratch of rdx_sort_splats_static, one copy per GPU for the life of the
// process: the passes' two (key, value) buffer pairs of up to
// RDX_STATIC_SORT_MAX_KEYS words, the tile counts, digit totals, kept count
// and the ranged sort's tile and segment ranges and done count, each region
// at a 256-byte boundary. Host code never names `scratch` as an address: its
// device address comes from hipGetSymbolAddress, per device.
namespace {
constexpr uint32_t kStaticTile = 512 * 16;  // the default tuning's tile
constexpr uint32_t kStaticTiles = (RDX_STATIC_SORT_MAX_KEYS + kStaticTile - 1) / kStaticTile;
constexpr size_t round64(size_t w) { return (w + 63) / 64 * 64; }
constexpr size_t kStaticPair = round64(RDX_STATIC_SORT_MAX_KEYS);
constexpr size_t kStaticCnt = round64((size_t)count_stride(MAX_BINS + 1) * kStaticTiles);
constexpr size_t kStaticTot = round64(TOT_STRIDE);
constexpr size_t kStaticNvalid = 64;
constexpr size_t kStaticMm = round64(ranged_words(kStaticTiles, 1));
constexpr size_t kStaticWords = 4 * kStaticPair + kStaticCnt + kStaticTot + kStaticNvalid + kStaticMm;
std::mutex g_static_mu;
constexpr int kMaxDevices = 64;
void* g_static_ptr[kMaxDevices];
}  // namespace

static __device__ __align__(256) uint32_t scratch[kStaticWords];

// The current device's copy of `scratch` (resolved on its first use there).
static int static_scratch(uint32_t** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "rdx_sort_splats_static: hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(RDX_E_UNSUPPORTED, "rdx_sort_splats_static: device %d", dev);
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(RDX_E_HIP, "rdx_sort_splats_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
    }
    g_static_ptr[dev] = p;
  }
  *out = (uint32_t*)g_static_ptr[dev];
  return RDX_OK;
}

static void view(Buffer& b, uint32_t*& at, size_t words) {
  b.ptr = at;
  b.bytes = words * 4;
  at += words;
}

int rdx_sort_splats_static(void* stream, const uint32_t* keys, uint32_t n, int bits, uint32_t* out_indices,
                           uint32_t* out_count, uint32_t* out_args) {
  if (int rc = check_bits(bits, "WgpuRadixSort")) return rc;
  if (n > RDX_STATIC_SORT_MAX_KEYS)
    return fail(RDX_E_CAPACITY, "rdx_sort_splats_static: %u keys exceed its static scratch of %u keys", n,
                (unsigned)RDX_STATIC_SORT_MAX_KEYS);
  if (n && (!keys || !out_indices)) return fail(RDX_E_INVALID, "rdx_sort_splats_static: NULL keys or indices");
  hipStream_t s = (hipStream_t)stream;
  if (n == 0) {
    hipLaunchKernelGGL(k_empty_outputs, dim3(1), dim3(64), 0, s, out_count, out_args);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_sort_splats_static");
  }
  uint32_t* base = nullptr;
  if (int rc = static_scratch(&base)) return rc;
  const rdx_sort_tuning t = {8, 16, 512};
  const SortPlan p = compact_plan(t, bits);
  SegLayout L;
  L.tiles = tiles_of(p, n);
  // views into the static array (never released: SortScratch has no destructor)
  SortScratch sc;
  uint32_t* at = base;
  for (int i = 0; i < 2; i++) {
    view(sc.keys[i], at, kStaticPair);
    view(sc.vals[i], at, kStaticPair);
  }
  view(sc.cnt, at, kStaticCnt);
  view(sc.tot, at, kStaticTot);
  view(sc.nvalid, at, kStaticNvalid);
  view(sc.mm, at, kStaticMm);
  return run_sort(nullptr, sc, s, p, L, keys, nullptr, nullptr, out_indices, n, nullptr, true, key_invalid(bits), false,
                  out_count, out_args);
}

}  // extern "C"

// ---- the tuning cache --------------------------------------------------------------

namespace rdx {
namespace {
struct TuneEntry {
  TuneKey key;
  rdx_sort_tuning tuning;
  double ms;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_tune;
std::string g_tune_file;
bool g_tune_loaded = false;

bool same_key(const TuneKey& a, const TuneKey& b) {
  return a.arch == b.arch && a.form == b.form && a.bits == b.bits && a.n_bucket == b.n_bucket &&
         a.seg_bucket == b.seg_bucket;
}

int ceil_log2(uint64_t x) {
  int b = 0;
  while (b < 63 && (1ull << b) < x) b++;
  return b;
}

// The file's lines: arch form bits n_bucket seg_bucket digit_bits keys_per_thread block_threads ms
void load_file_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  if (g_tune_file.empty()) {
    const char* env = getenv("RDX_TUNING_CACHE");
    if (env && *env) g_tune_file = env;
  }
  if (g_tune_file.empty()) return;
  FILE* f = fopen(g_tune_file.c_str(), "r");
  if (!f) return;
  char arch[128], form[32];
  TuneEntry e;
  while (fscanf(f, "%127s %31s %d %d %d %d %d %d %lf", arch, form, &e.key.bits, &e.key.n_bucket, &e.key.seg_bucket,
                &e.tuning.digit_bits, &e.tuning.keys_per_thread, &e.tuning.block_threads, &e.ms) == 9) {
    e.key.arch = arch;
    e.key.form = form;
    if (!tuning_supported(e.tuning)) continue;
    bool found = false;
    for (auto& x : g_tune)
      if (same_key(x.key, e.key)) x = e, found = true;
    if (!found) g_tune.push_back(e);
  }
  fclose(f);
}
}  // namespace

TuneKey tune_key(const char* form, uint64_t n, uint32_t segments, int bits) {
  TuneKey k;
  k.form = form;
  k.bits = bits;
  k.n_bucket = ceil_log2(std::max<uint64_t>(n, 1));
  k.seg_bucket = ceil_log2(std::max<uint32_t>(segments, 1));
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    std::string a = prop.gcnArchName;
    k.arch = a.substr(0, a.find(':'));  // gfx950, without the feature suffixes
  } else {
    (void)hipGetLastError();
    k.arch = "unknown";
  }
  return k;
}

bool tuning_cache_get(const TuneKey& k, rdx_sort_tuning* t, double* ms) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  load_file_locked();
  for (const auto& e : g_tune)
    if (same_key(e.key, k)) {
      if (t) *t = e.tuning;
      if (ms) *ms = e.ms;
      return true;
    }
  return false;
}

void tuning_cache_put(const TuneKey& k, const rdx_sort_tuning& t, double ms) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  load_file_locked();
  bool found = false;
  for (auto& e : g_tune)
    if (same_key(e.key, k)) e.tuning = t, e.ms = ms, found = true;
  if (!found) g_tune.push_back(TuneEntry{k, t, ms});
  if (g_tune_file.empty()) return;
  FILE* f = fopen(g_tune_file.c_str(), "a");
  if (!f) return;
  fprintf(f, "%s %s %d %d %d %d %d %d %.6f\n", k.arch.c_str(), k.form.c_str(), k.bits, k.n_bucket, k.seg_bucket,
          t.digit_bits, t.keys_per_thread, t.block_threads, ms);
  fclose(f);
}

}  // namespace rdx

extern "C" {

int rdx_tuning_cache_file(const char* path) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune_file = path ? path : "";
  g_tune_loaded = false;
  load_file_locked();
  return RDX_OK;
}

int rdx_tuning_cache_lookup(const char* form, uint64_t n, uint32_t segments, int bits, rdx_sort_tuning* out,
                            double* ms) {
  if (!form || (strcmp(form, "sort") && strcmp(form, "batch")))
    return fail(RDX_E_INVALID, "rdx_tuning_cache_lookup: form \"sort\" or \"batch\"");
  return tuning_cache_get(tune_key(form, n, segments, bits), out, ms) ? 1 : 0;
}

static const rdx_sort_tuning k_candidates[] = {
    {8, 16, 256}, {8, 8, 256}, {8, 12, 256}, {8, 20, 256}, {8, 8, 512}, {8, 12, 512}, {8, 16, 512},
    {6, 16, 256}, {6, 8, 256}, {6, 12, 256}, {6, 20, 256}, {6, 8, 512}, {6, 12, 512}, {6, 16, 512},
    {4, 16, 256}, {4, 8, 256}, {4, 12, 256}, {4, 20, 256}, {4, 8, 512}, {4, 12, 512}, {4, 16, 512},
};

int rdx_sort_tuning_candidate(int index, rdx_sort_tuning* out) {
  const int n = (int)(sizeof k_candidates / sizeof k_candidates[0]);
  if (index < 0 || index >= n || !out) return fail(RDX_E_INVALID, "rdx_sort_tuning_candidate: %d of %d", index, n);
  *out = k_candidates[index];
  return RDX_OK;
}

static __global__ void k_random_keys(uint32_t* keys, uint32_t n, uint32_t mask, uint32_t seed) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    // splitmix32-style mixing of (seed, index)
    uint32_t x = i * 0x9e3779b9u + seed;
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    keys[i] = x & mask;
  }
}

int rdx_context_autotune(rdx_context* ctx, void* stream, const uint32_t* keys_dev, uint32_t n, int bits, int reps,
                         rdx_sort_tuning* out, double* ms_out) {
  if (!ctx) return fail(RDX_E_INVALID, "rdx_context_autotune: no context");
  if (int rc = check_bits(bits, "WgpuRadixSort")) return rc;
  if (n == 0) return fail(RDX_E_INVALID, "rdx_context_autotune: no keys");
  Context* c = reinterpret_cast<Context*>(ctx);
  hipStream_t s = (hipStream_t)stream;
  reps = std::max(1, std::min(reps, 100));
  Buffer keys, out_idx, counts;
  int rc;
  if ((rc = out_idx.ensure((size_t)n * 4)) || (rc = counts.ensure(64))) {
    keys.release();
    out_idx.release();
    counts.release();
    return rc;
  }
  if (!keys_dev) {
    if ((rc = keys.ensure((size_t)n * 4))) {
      out_idx.release();
      counts.release();
      return rc;
    }
    // below the invalid key: every key is sorted
    hipLaunchKernelGGL(k_random_keys, dim3((uint32_t)std::min<uint64_t>(((uint64_t)n + 255) / 256, 4096u)), dim3(256), 0, s,
                       (uint32_t*)keys.ptr, n, key_invalid(bits) >> 1, 0x2545f491u);
    keys_dev = (const uint32_t*)keys.ptr;
  }
  const rdx_sort_tuning saved = c->tuning;
  const bool prof = c->profiling;
  c->profiling = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double best = 1e30;
  rdx_sort_tuning best_t = saved;
  const int ncand = (int)(sizeof k_candidates / sizeof k_candidates[0]);
  for (int k = 0; k < ncand && e == hipSuccess && rc == RDX_OK; k++) {
    c->tuning = k_candidates[k];
    rc = rdx_sort_splats(ctx, stream, keys_dev, n, bits, (uint32_t*)out_idx.ptr, (uint32_t*)counts.ptr, nullptr);
    if (rc) break;
    e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == RDX_OK && e == hipSuccess; r++)
      rc = rdx_sort_splats(ctx, stream, keys_dev, n, bits, (uint32_t*)out_idx.ptr, (uint32_t*)counts.ptr, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && rc == RDX_OK && ms / reps < best) {
      best = ms / reps;
      best_t = k_candidates[k];
    }
  }
  c->profiling = prof;
  c->tuning = rc == RDX_OK && e == hipSuccess ? best_t : saved;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  keys.release();
  out_idx.release();
  counts.release();
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "rdx_context_autotune");
  tuning_cache_put(tune_key("sort", n, 1, bits), best_t, best);
  if (out) *out = best_t;
  if (ms_out) *ms_out = best;
  return RDX_OK;
}

}  // extern "C"

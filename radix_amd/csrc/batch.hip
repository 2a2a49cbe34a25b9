// The batched form: every (view, table) pair of a batch is a segment of one
// keys array. One projection launch per table kind (layout, SH codec and
// degree) covers all its tables and views; the radix passes sort all
// segments at once. The sequence can be captured into a HIP graph and
// replayed; nothing in it clears memory with a memset, and every buffer it
// touches belongs to the batch, so a replay never sees a reallocation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/radix_hip.h"
#include "api_internal.hpp"
#include "batch_kernels.hpp"
#include "project_host.hpp"

namespace rdx {

struct BatchKind {
  int layout, shc, deg;  // deg: the f32 tables' degree, -1 for word tables
  std::vector<BatchGroup> groups;
  uint32_t blocks = 0;   // 256-splat blocks of the projection launch
  uint32_t chunks = 0;   // PERSIST_BLOCK-splat chunks of the persistent launch
  Buffer d_groups;
};

struct Batch {
  Context* ctx = nullptr;
  rdx_params params;
  SortPlan plan;
  std::vector<rdx_splat_table> tables;
  std::vector<int> table_deg;
  std::vector<rdx_batch_view> views;
  std::vector<uint64_t> seg_off;
  std::vector<uint32_t> tile0;
  uint32_t total = 0, tiles = 0;
  BatchU u;
  std::vector<BatchView> host_views;  // in group order
  std::vector<int> slot_of_view;      // view -> index in host_views
  std::vector<BatchKind> kinds;
  Buffer d_views, d_tables, d_deg, d_enc, d_seg, keys_own;
  uint32_t* keys = nullptr;
  SortScratch scratch;
  SegLayout layout;
  // graph
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::tuple<uint32_t*, uint32_t*, uint32_t*> graph_outputs{nullptr, nullptr, nullptr};
  bool graph_persistent = false;
  // the persistent form (one kind, the ranged sort plan): its grid and its
  // barrier words [arrivals, finished blocks, error]
  bool persist_ok = false;
  uint32_t persist_grid = 0;
  Buffer d_sync;
  Buffer d_stamps;  // the persistent run's phase stamps while the context profiles
  bool stamped = false;
  hipEvent_t last_run = nullptr;
  bool ran = false;

  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    if (last_run) (void)hipEventDestroy(last_run);
    exec = nullptr;
    graph = nullptr;
    capture_stream = nullptr;
    last_run = nullptr;
    for (auto& k : kinds) k.d_groups.release();
    d_views.release();
    d_tables.release();
    d_deg.release();
    d_enc.release();
    d_seg.release();
    d_sync.release();
    d_stamps.release();
    keys_own.release();
    scratch.release();
  }
};

static int upload(Buffer& b, const void* src, size_t bytes) {
  if (int rc = b.ensure(std::max<size_t>(bytes, 4))) return rc;
  hipError_t e = hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "batch upload");
}

// View uniforms: the table seen by the view's (mono) camera.
static int view_uniforms(Batch* b, int v) {
  const rdx_batch_view& V = b->views[v];
  ProjU pu;
  if (int rc = make_uniforms(&b->tables[V.table], &V.camera, &b->params, &pu)) return rc;
  BatchView& bv = b->host_views[b->slot_of_view[v]];
  bv.eye = pu.eye[0];
  for (int k = 0; k < 3; k++) bv.h[k] = pu.h[k];
  bv.seg_off = (uint32_t)b->seg_off[v];
  bv.pad = 0;
  bv.words = (uint4*)V.words_dev;
  return RDX_OK;
}

template <int LAYOUT, int SHC, int DEG>
static void launch_kind(hipStream_t s, uint32_t blocks, const BatchDesc& d, const BatchU& u, uint32_t* keys) {
  hipLaunchKernelGGL((k_project_batch<LAYOUT, SHC, DEG>), dim3(blocks), dim3(256), 0, s, d, u, keys);
}

static bool launch_batch_kind(hipStream_t s, const BatchKind& k, const BatchDesc& d, const BatchU& u, uint32_t* keys) {
  if (k.layout == LAYOUT_F32 && k.shc == SHC_F32) {
    switch (k.deg) {
      case 0: launch_kind<LAYOUT_F32, SHC_F32, 0>(s, k.blocks, d, u, keys); return true;
      case 1: launch_kind<LAYOUT_F32, SHC_F32, 1>(s, k.blocks, d, u, keys); return true;
      case 2: launch_kind<LAYOUT_F32, SHC_F32, 2>(s, k.blocks, d, u, keys); return true;
      case 3: launch_kind<LAYOUT_F32, SHC_F32, 3>(s, k.blocks, d, u, keys); return true;
    }
    return false;
  }
  if (k.layout == LAYOUT_F32 && k.shc == SHC_F32_WIDE) {
    if (k.deg == 2) launch_kind<LAYOUT_F32, SHC_F32_WIDE, 2>(s, k.blocks, d, u, keys);
    else launch_kind<LAYOUT_F32, SHC_F32_WIDE, 3>(s, k.blocks, d, u, keys);
    return true;
  }
  if (k.layout == LAYOUT_EXT) {
    if (k.shc == SHC_PACKED) launch_kind<LAYOUT_EXT, SHC_PACKED, -1>(s, k.blocks, d, u, keys);
    else launch_kind<LAYOUT_EXT, SHC_EXT, -1>(s, k.blocks, d, u, keys);
    return true;
  }
  if (k.layout == LAYOUT_COV) {
    if (k.shc == SHC_PACKED) launch_kind<LAYOUT_COV, SHC_PACKED, -1>(s, k.blocks, d, u, keys);
    else launch_kind<LAYOUT_COV, SHC_EXT, -1>(s, k.blocks, d, u, keys);
    return true;
  }
  if (k.shc == SHC_PACKED) launch_kind<LAYOUT_PACKED, SHC_PACKED, -1>(s, k.blocks, d, u, keys);
  else launch_kind<LAYOUT_PACKED, SHC_EXT, -1>(s, k.blocks, d, u, keys);
  return true;
}

template <int LAYOUT, int SHC, int DEG>
static const void* persist_fn(bool stamps) {
  return stamps ? reinterpret_cast<const void*>(&k_batch_persistent<LAYOUT, SHC, DEG, true>)
                : reinterpret_cast<const void*>(&k_batch_persistent<LAYOUT, SHC, DEG, false>);
}

// The persistent kernel of a kind.
static const void* persistent_kernel(const BatchKind& k, bool stamps = false) {
  if (k.layout == LAYOUT_F32 && k.shc == SHC_F32) {
    switch (k.deg) {
      case 0: return persist_fn<LAYOUT_F32, SHC_F32, 0>(stamps);
      case 1: return persist_fn<LAYOUT_F32, SHC_F32, 1>(stamps);
      case 2: return persist_fn<LAYOUT_F32, SHC_F32, 2>(stamps);
      case 3: return persist_fn<LAYOUT_F32, SHC_F32, 3>(stamps);
    }
    return nullptr;
  }
  if (k.layout == LAYOUT_F32 && k.shc == SHC_F32_WIDE)
    return k.deg == 2 ? persist_fn<LAYOUT_F32, SHC_F32_WIDE, 2>(stamps) : persist_fn<LAYOUT_F32, SHC_F32_WIDE, 3>(stamps);
  if (k.layout == LAYOUT_EXT)
    return k.shc == SHC_PACKED ? persist_fn<LAYOUT_EXT, SHC_PACKED, -1>(stamps) : persist_fn<LAYOUT_EXT, SHC_EXT, -1>(stamps);
  if (k.layout == LAYOUT_COV)
    return k.shc == SHC_PACKED ? persist_fn<LAYOUT_COV, SHC_PACKED, -1>(stamps) : persist_fn<LAYOUT_COV, SHC_EXT, -1>(stamps);
  return k.shc == SHC_PACKED ? persist_fn<LAYOUT_PACKED, SHC_PACKED, -1>(stamps)
                             : persist_fn<LAYOUT_PACKED, SHC_EXT, -1>(stamps);
}

static BatchDesc batch_desc(const Batch* b, const BatchKind& k) {
  BatchDesc d;
  d.views = (const BatchView*)b->d_views.ptr;
  d.groups = (const BatchGroup*)k.d_groups.ptr;
  d.tables = (const TableP*)b->d_tables.ptr;
  d.table_sh_degree = (const int*)b->d_deg.ptr;
  d.table_enc = (const float*)b->d_enc.ptr;
  d.num_groups = (int)k.groups.size();
  d.num_segments = (uint32_t)b->views.size();
  return d;
}

// The whole batch in one launch of the persistent kernel.
static int run_batch_persistent(Batch* b, hipStream_t s, uint32_t* out_indices, uint32_t* out_counts,
                                uint32_t* out_args) {
  const BatchKind& k = b->kinds[0];
  PersistArgs P;
  memset(&P, 0, sizeof P);
  const uint32_t* kin = b->keys;
  const uint32_t* vin = nullptr;
  for (int pass = 0; pass < 4; pass++) {
    P.pass[pass] = pass_args(b->scratch, b->plan, b->layout, pass, kin, vin, nullptr, out_indices, b->total, nullptr,
                             true, key_invalid(b->params.sort_bits), false, out_counts, out_args);
    kin = P.pass[pass].keys_out;
    vin = P.pass[pass].vals_out;
  }
  P.items = k.chunks;
  P.blocks = b->persist_grid;
  P.bar = (uint32_t*)b->d_sync.ptr;
  P.err = P.bar + 2;
  P.stamps = nullptr;
  b->stamped = false;
  if (b->ctx->profiling) {
    const size_t words = (size_t)(PERSIST_PHASES + 1) * 2 * b->persist_grid;
    if (int rc = b->d_stamps.ensure(words * sizeof(uint64_t))) return rc;
    P.stamps = (uint64_t*)b->d_stamps.ptr;
    b->stamped = true;
  }
  BatchDesc d = batch_desc(b, k);
  BatchU u = b->u;
  uint32_t* keys = b->keys;
  void* args[] = {&d, &u, &keys, &P};
  Stage st(b->ctx, s, "batch_persistent");
  hipError_t e = hipLaunchKernel(persistent_kernel(k, P.stamps != nullptr), dim3(b->persist_grid), dim3(PERSIST_BLOCK),
                                 args, 0, s);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "batch persistent launch");
}

// The whole batch on one stream: projections, then the segmented sort.
static int run_batch(Batch* b, hipStream_t s, uint32_t* out_indices, uint32_t* out_counts, uint32_t* out_args,
                     bool persistent) {
  if (persistent) return run_batch_persistent(b, s, out_indices, out_counts, out_args);
  {
    Stage st(b->ctx, s, "batch_project");
    for (auto& k : b->kinds) {
      const BatchDesc d = batch_desc(b, k);
      if (!launch_batch_kind(s, k, d, b->u, b->keys))
        return fail(RDX_E_INVALID, "batch: no kernel for layout %d", k.layout);
    }
  }
  return run_sort(b->ctx, b->scratch, s, b->plan, b->layout, b->keys, nullptr, nullptr, out_indices, b->total, nullptr,
                  true, key_invalid(b->params.sort_bits), false, out_counts, out_args);
}

}  // namespace rdx

using namespace rdx;

extern "C" {

int rdx_batch_create(rdx_context* ctx, const rdx_splat_table* tables, int num_tables, const rdx_batch_view* views,
                     int num_views, const rdx_params* params, uint32_t* keys_dev, rdx_batch** out) {
  if (!out) return fail(RDX_E_INVALID, "rdx_batch_create: out is NULL");
  *out = nullptr;
  if (!ctx || !tables || !views || !params) return fail(RDX_E_INVALID, "rdx_batch_create: NULL argument");
  if (num_tables < 1 || num_views < 1 || num_views > RDX_BATCH_MAX_VIEWS)
    return fail(RDX_E_INVALID, "rdx_batch_create: %d tables, %d views (1..%d views)", num_tables, num_views,
                RDX_BATCH_MAX_VIEWS);
  if (int rc = check_bits(params->sort_bits, "WgpuRadixSort")) return rc;
  Context* c = reinterpret_cast<Context*>(ctx);
  Batch* b = new Batch();
  b->ctx = c;
  b->params = *params;
  b->plan = compact_plan(c->tuning, params->sort_bits);
  b->tables.assign(tables, tables + num_tables);
  b->views.assign(views, views + num_views);
  auto bail = [&](int rc) {
    b->release();
    delete b;
    return rc;
  };
  for (int t = 0; t < num_tables; t++)
    if (int rc = check_table(&tables[t])) return bail(rc);
  // Segments: view order.
  uint64_t off = 0;
  uint32_t tiles = 0;
  for (int v = 0; v < num_views; v++) {
    const rdx_batch_view& V = views[v];
    if (V.table >= (uint32_t)num_tables) return bail(fail(RDX_E_INVALID, "batch view %d: table %u of %d", v, V.table, num_tables));
    if (V.camera.eyes != 1) return bail(fail(RDX_E_INVALID, "batch view %d: one eye per view (got %d)", v, V.camera.eyes));
    b->seg_off.push_back(off);
    b->tile0.push_back(tiles);
    off += tables[V.table].count;
    tiles += tiles_of(b->plan, tables[V.table].count);
  }
  if (off >= 0x80000000ull) return bail(fail(RDX_E_CAPACITY, "batch: %llu splat-views (limit 2^31)", (unsigned long long)off));
  b->total = (uint32_t)off;
  b->tiles = tiles;
  // Tables: kernel pointers, degrees, encodings.
  std::vector<TableP> tp;
  std::vector<float> enc;
  for (int t = 0; t < num_tables; t++) {
    tp.push_back(make_table(&tables[t]));
    ProjU pu;
    rdx_camera probe;
    memset(&probe, 0, sizeof probe);
    probe.eyes = 1;
    for (int k = 0; k < 16; k += 5) probe.head[k] = 1.0, probe.eye[0].view[k] = 1.0, probe.eye[0].projection[k] = 1.0;
    probe.size[0] = probe.size[1] = 1.0;
    if (int rc = make_uniforms(&tables[t], &probe, params, &pu)) return bail(rc);
    b->table_deg.push_back(pu.sh_degree);
    for (int k = 0; k < 4; k++) enc.push_back(pu.enc[k]);
  }
  // Groups: per kind, per table, up to BATCH_MAX_VIEWS views each.
  std::map<std::tuple<int, int, int>, std::vector<int>> by_kind;  // kind -> views
  for (int v = 0; v < num_views; v++) {
    const rdx_splat_table& t = tables[views[v].table];
    const int layout = t.layout == RDX_LAYOUT_F32_SOA ? LAYOUT_F32
                       : t.layout == RDX_LAYOUT_EXT   ? LAYOUT_EXT
                       : t.layout == RDX_LAYOUT_COV   ? LAYOUT_COV
                                                      : LAYOUT_PACKED;
    const int deg = layout == LAYOUT_F32 ? b->table_deg[views[v].table] : -1;
    const uint64_t stride = t.plane_stride ? t.plane_stride : t.count;
    const int shc = layout == LAYOUT_F32 ? (f32_sh_wide(stride, deg) ? SHC_F32_WIDE : SHC_F32)
                                         : (t.sh_codec == RDX_SH_PACKED ? SHC_PACKED : SHC_EXT);
    by_kind[std::make_tuple(layout, shc, deg)].push_back(v);
  }
  b->slot_of_view.assign(num_views, -1);
  for (auto& kv : by_kind) {
    BatchKind k;
    std::tie(k.layout, k.shc, k.deg) = kv.first;
    std::vector<int> vs = kv.second;
    std::stable_sort(vs.begin(), vs.end(), [&](int x, int y) { return views[x].table < views[y].table; });
    size_t i = 0;
    while (i < vs.size()) {
      const uint32_t t = views[vs[i]].table;
      BatchGroup g;
      g.table = t;
      g.first_view = (uint32_t)b->host_views.size();
      g.num_views = 0;
      g.block0 = k.blocks;
      g.chunk0 = k.chunks;
      while (i < vs.size() && views[vs[i]].table == t && g.num_views < (uint32_t)BATCH_MAX_VIEWS) {
        b->slot_of_view[vs[i]] = (int)b->host_views.size();
        b->host_views.push_back(BatchView());
        g.num_views++;
        i++;
      }
      k.blocks += std::max<uint32_t>(1u, (tables[t].count + 255u) / 256u);
      k.chunks += std::max<uint32_t>(1u, (tables[t].count + PERSIST_BLOCK - 1u) / PERSIST_BLOCK);
      k.groups.push_back(g);
    }
    b->kinds.push_back(std::move(k));
  }
  for (int v = 0; v < num_views; v++)
    if (int rc = view_uniforms(b, v)) return bail(rc);
  // Shared uniforms (the first table's; everything but the eye is per params).
  {
    ProjU pu;
    if (int rc = make_uniforms(&tables[views[0].table], &views[0].camera, params, &pu)) return bail(rc);
    BatchU& u = b->u;
    u.lod_inflate = pu.lod_inflate;
    u.ndc_z_min = pu.ndc_z_min;
    u.clip_xy = pu.clip_xy;
    u.max_std_dev = pu.max_std_dev;
    u.min_alpha = pu.min_alpha;
    u.blur = pu.blur;
    u.pre_blur = pu.pre_blur;
    u.min_pr = pu.min_pr;
    u.max_pr = pu.max_pr;
    u.min_contrib = pu.min_contrib;
    u.focal_distance = pu.focal_distance;
    u.aperture_tan = pu.aperture_tan;
    u.sort_bits = pu.sort_bits;
    u.cull_keys = pu.cull_keys;
  }
  for (int v = 1; v < num_views; v++)
    if (views[v].camera.ndc_z_min != views[0].camera.ndc_z_min)
      return bail(fail(RDX_E_INVALID, "batch: views differ in clip convention (ndc_z_min)"));
  // Device copies: views, groups, tables, segment table.
  int rc;
  if ((rc = upload(b->d_views, b->host_views.data(), b->host_views.size() * sizeof(BatchView)))) return bail(rc);
  for (auto& k : b->kinds)
    if ((rc = upload(k.d_groups, k.groups.data(), k.groups.size() * sizeof(BatchGroup)))) return bail(rc);
  if ((rc = upload(b->d_tables, tp.data(), tp.size() * sizeof(TableP)))) return bail(rc);
  if ((rc = upload(b->d_deg, b->table_deg.data(), b->table_deg.size() * sizeof(int)))) return bail(rc);
  if ((rc = upload(b->d_enc, enc.data(), enc.size() * sizeof(float)))) return bail(rc);
  std::vector<uint32_t> seg(4 * (size_t)num_views + tiles);
  for (int v = 0; v < num_views; v++) {
    const uint32_t n = tables[views[v].table].count;
    seg[4 * v] = b->tile0[v];
    seg[4 * v + 1] = (uint32_t)b->seg_off[v];
    seg[4 * v + 2] = n;
    seg[4 * v + 3] = tiles_of(b->plan, n);
    for (uint32_t t = 0; t < seg[4 * v + 3]; t++) seg[4 * (size_t)num_views + b->tile0[v] + t] = (uint32_t)v;
  }
  if ((rc = upload(b->d_seg, seg.data(), seg.size() * 4))) return bail(rc);
  b->layout.seg = (const uint32_t*)b->d_seg.ptr;
  b->layout.seg_of_tile = (const uint32_t*)b->d_seg.ptr + 4 * num_views;
  b->layout.segments = (uint32_t)num_views;
  b->layout.tiles = tiles;
  if ((rc = b->scratch.reserve(b->total, (uint32_t)num_views, tiles))) return bail(rc);
  if (keys_dev) {
    b->keys = keys_dev;
  } else {
    if ((rc = b->keys_own.ensure((size_t)std::max<uint32_t>(b->total, 1) * 4))) return bail(rc);
    b->keys = (uint32_t*)b->keys_own.ptr;
  }
  hipError_t e = hipEventCreateWithFlags(&b->last_run, hipEventDisableTiming);
  if (e != hipSuccess) return bail(hip_fail(e, "batch event"));
  // The persistent form: one kind of table and the ranged sort plan; a grid
  // of the blocks resident at once, a multiple of the 8 XCDs.
  if (b->kinds.size() == 1 && b->plan.ranged && b->plan.block == PERSIST_BLOCK && b->plan.kpt == PERSIST_KPT &&
      persistent_kernel(b->kinds[0])) {
    int dev = 0, cus = 0, per_cu = 0;
    e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess)
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_kernel(b->kinds[0]), PERSIST_BLOCK, 0);
    if (e != hipSuccess) return bail(hip_fail(e, "batch persistent occupancy"));
    const uint32_t grid = (uint32_t)(cus * per_cu) / XCDS * XCDS;
    if (grid >= XCDS) {
      if ((rc = b->d_sync.ensure(4 * sizeof(uint32_t)))) return bail(rc);
      e = hipMemset(b->d_sync.ptr, 0, 4 * sizeof(uint32_t));
      if (e != hipSuccess) return bail(hip_fail(e, "batch persistent barrier"));
      b->persist_grid = grid;
      b->persist_ok = true;
    }
  }
  *out = reinterpret_cast<rdx_batch*>(b);
  return RDX_OK;
}

int rdx_batch_destroy(rdx_batch* batch) {
  if (!batch) return RDX_OK;
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (b->ran) (void)hipEventSynchronize(b->last_run);
  b->release();
  delete b;
  return RDX_OK;
}

int rdx_batch_set_cameras(rdx_batch* batch, const rdx_camera* cameras, int num_views) {
  if (!batch || !cameras) return fail(RDX_E_INVALID, "rdx_batch_set_cameras: NULL argument");
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (num_views != (int)b->views.size())
    return fail(RDX_E_INVALID, "rdx_batch_set_cameras: %d cameras for %zu views", num_views, b->views.size());
  for (int v = 0; v < num_views; v++) {
    if (cameras[v].eyes != 1) return fail(RDX_E_INVALID, "batch view %d: one eye per view (got %d)", v, cameras[v].eyes);
    if (cameras[v].ndc_z_min != b->views[0].camera.ndc_z_min)
      return fail(RDX_E_INVALID, "batch: views differ in clip convention (ndc_z_min)");
  }
  for (int v = 0; v < num_views; v++) b->views[v].camera = cameras[v];
  for (int v = 0; v < num_views; v++)
    if (int rc = view_uniforms(b, v)) return rc;
  if (b->ran) {
    hipError_t e = hipEventSynchronize(b->last_run);
    if (e != hipSuccess) return hip_fail(e, "rdx_batch_set_cameras");
  }
  hipError_t e = hipMemcpy(b->d_views.ptr, b->host_views.data(), b->host_views.size() * sizeof(BatchView),
                           hipMemcpyHostToDevice);
  return e == hipSuccess ? RDX_OK : hip_fail(e, "rdx_batch_set_cameras");
}

uint64_t rdx_batch_total(const rdx_batch* batch) {
  return batch ? reinterpret_cast<const Batch*>(batch)->total : 0;
}

uint64_t rdx_batch_segment_offset(const rdx_batch* batch, int view) {
  if (!batch) return 0;
  const Batch* b = reinterpret_cast<const Batch*>(batch);
  if (view < 0 || view >= (int)b->seg_off.size()) return b->total;
  return b->seg_off[view];
}

const uint32_t* rdx_batch_keys(const rdx_batch* batch) {
  return batch ? reinterpret_cast<const Batch*>(batch)->keys : nullptr;
}

int rdx_batch_run(rdx_batch* batch, void* stream, uint32_t* out_indices, uint32_t* out_counts, uint32_t* out_args,
                  int flags) {
  if (!batch || !out_indices) return fail(RDX_E_INVALID, "rdx_batch_run: NULL argument");
  Batch* b = reinterpret_cast<Batch*>(batch);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  const bool persistent = (flags & RDX_BATCH_PERSISTENT) != 0;
  if (persistent && !b->persist_ok)
    return fail(RDX_E_UNSUPPORTED,
                "rdx_batch_run: the persistent form needs tables of one kind (layout, SH codec and degree) and the 8-bit, "
                "12-key, 512-thread tuning at 32-bit keys");
  if (flags & RDX_BATCH_GRAPH) {
    const auto outs = std::make_tuple(out_indices, out_counts, out_args);
    if (!b->exec || b->graph_outputs != outs || b->graph_persistent != persistent) {
      if (b->exec) (void)hipGraphExecDestroy(b->exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->exec = nullptr;
      b->graph = nullptr;
      if (!b->capture_stream) {
        e = hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking);
        if (e != hipSuccess) return hip_fail(e, "batch capture stream");
      }
      // Stage events stay out of the graph.
      const bool prof = b->ctx->profiling;
      b->ctx->profiling = false;
      e = hipStreamBeginCapture(b->capture_stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) {
        b->ctx->profiling = prof;
        return hip_fail(e, "batch capture");
      }
      const int rc = run_batch(b, b->capture_stream, out_indices, out_counts, out_args, persistent);
      e = hipStreamEndCapture(b->capture_stream, &b->graph);
      b->ctx->profiling = prof;
      if (rc) return rc;
      if (e != hipSuccess) return hip_fail(e, "batch capture");
      e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
      if (e != hipSuccess) return hip_fail(e, "batch graph instantiate");
      b->graph_outputs = outs;
      b->graph_persistent = persistent;
    }
    Stage st(b->ctx, s, "batch_graph");
    e = hipGraphLaunch(b->exec, s);
    if (e != hipSuccess) return hip_fail(e, "batch graph launch");
  } else {
    if (int rc = run_batch(b, s, out_indices, out_counts, out_args, persistent)) return rc;
  }
  e = hipEventRecord(b->last_run, s);
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_run");
  b->ran = true;
  return RDX_OK;
}

int rdx_batch_persistent_grid(const rdx_batch* batch) {
  if (!batch) return 0;
  const Batch* b = reinterpret_cast<const Batch*>(batch);
  return b->persist_ok ? (int)b->persist_grid : 0;
}

int rdx_batch_persistent_phases(rdx_batch* batch, double* phase_ms, double* release_ms, int max_phases) {
  if (!batch || !phase_ms || !release_ms) return fail(RDX_E_INVALID, "rdx_batch_persistent_phases: NULL argument");
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b->stamped) return 0;
  hipError_t e = hipEventSynchronize(b->last_run);
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_persistent_phases");
  const uint32_t G = b->persist_grid;
  std::vector<uint64_t> st((size_t)(PERSIST_PHASES + 1) * 2 * G);
  e = hipMemcpy(st.data(), b->d_stamps.ptr, st.size() * sizeof(uint64_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_persistent_phases");
  auto at = [&](int k, int side, uint32_t blk) { return st[((size_t)k * 2 + side) * G + blk]; };
  const int n = std::min(max_phases, PERSIST_PHASES);
  // s_memrealtime counts at 100 MHz
  for (int k = 0; k < n; k++) {
    uint64_t first_dep = ~0ull, last_arr = 0, last_dep_next = 0;
    for (uint32_t blk = 0; blk < G; blk++) {
      first_dep = std::min(first_dep, at(k, 1, blk));
      last_arr = std::max(last_arr, at(k + 1, 0, blk));
      if (k + 1 < PERSIST_PHASES) last_dep_next = std::max(last_dep_next, at(k + 1, 1, blk));
    }
    phase_ms[k] = (double)(last_arr - first_dep) * 1e-5;
    release_ms[k] = k + 1 < PERSIST_PHASES ? (double)(last_dep_next - last_arr) * 1e-5 : 0.0;
  }
  return n;
}

int rdx_batch_status(rdx_batch* batch) {
  if (!batch) return fail(RDX_E_INVALID, "rdx_batch_status: NULL batch");
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (b->ran) {
    hipError_t e = hipEventSynchronize(b->last_run);
    if (e != hipSuccess) return hip_fail(e, "rdx_batch_status");
  }
  if (!b->persist_ok) return RDX_OK;
  uint32_t err = 0;
  hipError_t e = hipMemcpy(&err, (uint32_t*)b->d_sync.ptr + 2, sizeof err, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_status");
  if (!err) return RDX_OK;
  const uint32_t zero = 0;
  e = hipMemcpy((uint32_t*)b->d_sync.ptr + 2, &zero, sizeof zero, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_status");
  return fail(RDX_E_HIP, "batch: a persistent run's grid barrier waited past its limit (blocks not all resident); "
                         "its output is not valid");
}

int rdx_batch_choose_form(rdx_batch* batch, void* stream, int reps, int* persistent_out, double* ms_graph,
                          double* ms_persistent) {
  if (!batch) return fail(RDX_E_INVALID, "rdx_batch_choose_form: NULL batch");
  Batch* b = reinterpret_cast<Batch*>(batch);
  reps = std::max(1, std::min(reps, 100));
  hipStream_t s = (hipStream_t)stream;
  Buffer idx, counts, args;
  int rc = RDX_OK;
  if ((rc = idx.ensure((size_t)std::max<uint32_t>(b->total, 1) * 4)) || (rc = counts.ensure(b->views.size() * 4)) ||
      (rc = args.ensure(b->views.size() * 20))) {
    idx.release();
    counts.release();
    args.release();
    return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  double ms[2] = {1e30, 1e30};
  for (int form = 0; form < (b->persist_ok ? 2 : 1) && rc == RDX_OK && e == hipSuccess; form++) {
    const int flags = RDX_BATCH_GRAPH | (form ? RDX_BATCH_PERSISTENT : 0);
    uint32_t* ip = (uint32_t*)idx.ptr;
    uint32_t* cp = (uint32_t*)counts.ptr;
    uint32_t* ap = (uint32_t*)args.ptr;
    rc = rdx_batch_run(batch, stream, ip, cp, ap, flags);  // capture, and one replay to warm
    if (rc == RDX_OK) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == RDX_OK && e == hipSuccess; r++) rc = rdx_batch_run(batch, stream, ip, cp, ap, flags);
    if (rc == RDX_OK && e == hipSuccess) e = hipEventRecord(e1, s);
    if (rc == RDX_OK && e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (rc == RDX_OK && e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (rc == RDX_OK && e == hipSuccess) ms[form] = t / reps;
    if (rc == RDX_OK && e == hipSuccess && form == 1) rc = rdx_batch_status(batch);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  idx.release();
  counts.release();
  args.release();
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_choose_form");
  if (persistent_out) *persistent_out = ms[1] < ms[0] ? 1 : 0;
  if (ms_graph) *ms_graph = ms[0];
  if (ms_persistent) *ms_persistent = b->persist_ok ? ms[1] : -1.0;
  return RDX_OK;
}

// Candidates in the order of rdx_sort_tuning_candidate; the batch sequence
// (projection + segmented sort) timed per candidate as graph replays.
int rdx_batch_autotune(rdx_context* ctx, void* stream, const rdx_splat_table* tables, int num_tables,
                       const rdx_batch_view* views, int num_views, const rdx_params* params, int reps, int flags,
                       rdx_sort_tuning* out, double* ms_out, int* cached_out) {
  if (!ctx || !tables || !views || !params) return fail(RDX_E_INVALID, "rdx_batch_autotune: NULL argument");
  if (int rc = check_bits(params->sort_bits, "WgpuRadixSort")) return rc;
  if (num_views < 1 || num_views > RDX_BATCH_MAX_VIEWS || num_tables < 1)
    return fail(RDX_E_INVALID, "rdx_batch_autotune: %d tables, %d views", num_tables, num_views);
  Context* c = reinterpret_cast<Context*>(ctx);
  uint64_t total = 0;
  for (int v = 0; v < num_views; v++) {
    if (views[v].table >= (uint32_t)num_tables)
      return fail(RDX_E_INVALID, "batch view %d: table %u of %d", v, views[v].table, num_tables);
    total += tables[views[v].table].count;
  }
  const TuneKey key = tune_key("batch", total, (uint32_t)num_views, params->sort_bits);
  if (cached_out) *cached_out = 0;
  rdx_sort_tuning best_t = c->tuning;
  double best = 1e30;
  if ((flags & RDX_TUNE_CACHED) && tuning_cache_get(key, &best_t, &best)) {
    c->tuning = best_t;
    if (out) *out = best_t;
    if (ms_out) *ms_out = best;
    if (cached_out) *cached_out = 1;
    return RDX_OK;
  }
  reps = std::max(1, std::min(reps, 100));
  hipStream_t s = (hipStream_t)stream;
  const rdx_sort_tuning saved = c->tuning;
  const bool prof = c->profiling;
  c->profiling = false;
  Buffer idx, counts, args;
  int rc = RDX_OK;
  if ((rc = idx.ensure((size_t)std::max<uint64_t>(total, 1) * 4)) || (rc = counts.ensure((size_t)num_views * 4)) ||
      (rc = args.ensure((size_t)num_views * 20))) {
    idx.release();
    counts.release();
    args.release();
    c->profiling = prof;
    return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  rdx_sort_tuning cand;
  for (int k = 0; rc == RDX_OK && e == hipSuccess && rdx_sort_tuning_candidate(k, &cand) == RDX_OK; k++) {
    c->tuning = cand;
    rdx_batch* b = nullptr;
    if ((rc = rdx_batch_create(ctx, tables, num_tables, views, num_views, params, nullptr, &b))) break;
    uint32_t* ip = (uint32_t*)idx.ptr;
    uint32_t* cp = (uint32_t*)counts.ptr;
    uint32_t* ap = (uint32_t*)args.ptr;
    rc = rdx_batch_run(b, stream, ip, cp, ap, RDX_BATCH_GRAPH);  // capture, and one replay to warm
    if (rc == RDX_OK) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == RDX_OK && e == hipSuccess; r++) rc = rdx_batch_run(b, stream, ip, cp, ap, RDX_BATCH_GRAPH);
    if (rc == RDX_OK && e == hipSuccess) e = hipEventRecord(e1, s);
    if (rc == RDX_OK && e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (rc == RDX_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (rc == RDX_OK && e == hipSuccess && ms / reps < best) {
      best = ms / reps;
      best_t = cand;
    }
    const int drc = rdx_batch_destroy(b);
    if (rc == RDX_OK) rc = drc;
  }
  c->profiling = prof;
  c->tuning = rc == RDX_OK && e == hipSuccess ? best_t : saved;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  idx.release();
  counts.release();
  args.release();
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "rdx_batch_autotune");
  tuning_cache_put(key, best_t, best);
  if (out) *out = best_t;
  if (ms_out) *ms_out = best;
  return RDX_OK;
}

}  // extern "C"

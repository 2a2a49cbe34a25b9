// N-API binding of libradix_hip.so for Node: the thin layer TypeScript hosts
// call (N-API version 6, no node-addon-api). Each function converts JS
// values to the C ABI's plain pointers and sizes, checks every device
// buffer against the bytes the call will touch (a short buffer throws
// before anything runs: the device never sees an access out of bounds), and
// maps a failure to a JS Error carrying the library's message (the
// reference's own text where it has one).
//
// Handles are Externals with a tag: device memory (rdx_device_alloc), a
// stream with its own context (rdx_stream_create; calls given one are
// ordered on it and use its scratch, so calls on different streams are in
// flight together), and the compositor, draw and batch objects. Every GPU
// call takes an optional stream as its last argument; without one it runs
// on the null stream with the calling thread's context.
#define NAPI_VERSION 6
#include <node_api.h>

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/radix_hip.h"

namespace {

enum Tag : uint32_t { T_MEM = 0x4d454d31, T_STREAM = 0x5354524d, T_RASTER = 0x52415354, T_DRAW = 0x44524157,
                      T_BATCH = 0x42415443 };

struct DeviceMem {
  uint32_t tag = T_MEM;
  void* ptr;
  size_t bytes;
};

struct GpuStream {
  uint32_t tag = T_STREAM;
  void* stream = nullptr;
  rdx_context* ctx = nullptr;
};

struct RasterH {
  uint32_t tag = T_RASTER;
  rdx_raster* h = nullptr;
  uint32_t n = 0, width = 0, height = 0;  // the last forward's
  bool forward = false;
};

struct DrawH {
  uint32_t tag = T_DRAW;
  rdx_draw* h = nullptr;
};

struct BatchH {
  uint32_t tag = T_BATCH;
  rdx_batch* h = nullptr;
  uint64_t total = 0;
  int views = 0;
};

thread_local rdx_context* t_ctx = nullptr;

rdx_context* thread_context() {
  if (!t_ctx) rdx_context_create(&t_ctx);
  return t_ctx;
}

napi_value throw_error(napi_env env, const char* msg) {
  napi_throw_error(env, nullptr, msg);
  return nullptr;
}

bool check(napi_env env, int rc) {
  if (rc == RDX_OK) return true;
  napi_throw_error(env, nullptr, rdx_last_error());
  return false;
}

#define ARGS(n)                   \
  size_t argc = n;                \
  napi_value argv[n];             \
  napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr)

double num(napi_env env, napi_value v, double fallback = 0.0) {
  double d = fallback;
  napi_valuetype t;
  if (v && napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_double(env, v, &d);
  return d;
}

napi_value make_num(napi_env env, double d) {
  napi_value r;
  napi_create_double(env, d, &r);
  return r;
}

napi_value make_u32(napi_env env, uint32_t u) {
  napi_value r;
  napi_create_uint32(env, u, &r);
  return r;
}

bool absent(napi_env env, napi_value v) {
  napi_valuetype t;
  return !v || napi_typeof(env, v, &t) != napi_ok || t == napi_undefined || t == napi_null;
}

template <typename T>
T* external(napi_env env, napi_value v, uint32_t tag) {
  napi_valuetype t;
  if (!v || napi_typeof(env, v, &t) != napi_ok || t != napi_external) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  if (!p || *(const uint32_t*)p != tag) return nullptr;
  return (T*)p;
}

DeviceMem* device(napi_env env, napi_value v) { return external<DeviceMem>(env, v, T_MEM); }

// A Uint32Array's data and element count
bool u32_array(napi_env env, napi_value v, uint32_t** data, size_t* length) {
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return false;
  napi_typedarray_type type;
  size_t len;
  void* ptr;
  napi_value buf;
  size_t off;
  napi_get_typedarray_info(env, v, &type, &len, &ptr, &buf, &off);
  if (type != napi_uint32_array) return false;
  *data = (uint32_t*)ptr;
  *length = len;
  return true;
}

// Any typed array's bytes
bool bytes_of(napi_env env, napi_value v, void** data, size_t* bytes) {
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_ta) return false;
  napi_typedarray_type type;
  size_t len;
  void* ptr;
  napi_value buf;
  size_t off;
  napi_get_typedarray_info(env, v, &type, &len, &ptr, &buf, &off);
  size_t elem = 1;
  switch (type) {
    case napi_int16_array:
    case napi_uint16_array: elem = 2; break;
    case napi_int32_array:
    case napi_uint32_array:
    case napi_float32_array: elem = 4; break;
    case napi_float64_array: elem = 8; break;
    default: elem = 1;
  }
  *data = ptr;
  *bytes = len * elem;
  return true;
}

// One N-API call: its name for messages, its context and stream (the
// optional stream argument's, else the thread's and the null stream), and
// the checks of its device buffers. After the first failed check `ok` is
// false and the call returns with the pending JS Error.
struct Call {
  napi_env env;
  const char* who;
  bool ok = true;
  rdx_context* ctx = nullptr;
  void* stream = nullptr;

  Call(napi_env e, const char* name, napi_value stream_arg) : env(e), who(name) {
    if (!absent(env, stream_arg)) {
      GpuStream* s = external<GpuStream>(env, stream_arg, T_STREAM);
      if (!s || !s->stream) {
        fail("%s: the last argument is not a stream", who);
        return;
      }
      ctx = s->ctx;
      stream = s->stream;
    } else {
      ctx = thread_context();
    }
  }

  bool fail(const char* fmt, ...) {
    if (ok) {
      char buf[512];
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(buf, sizeof buf, fmt, ap);
      va_end(ap);
      napi_throw_error(env, nullptr, buf);
      ok = false;
    }
    return false;
  }

  // A device buffer argument the call reads or writes `bytes` of.
  // Optional ones may be undefined or null (NULL to the library).
  void* buf(napi_value v, const char* what, uint64_t bytes, bool optional = false) {
    if (!ok) return nullptr;
    if (absent(env, v)) {
      if (!optional) fail("%s: %s is required", who, what);
      return nullptr;
    }
    DeviceMem* m = device(env, v);
    if (!m || !m->ptr) {
      fail("%s: %s is not a device buffer", who, what);
      return nullptr;
    }
    if (bytes > m->bytes) {
      fail("%s: %s holds %zu bytes, %llu needed (the range is out of bounds)", who, what, m->bytes,
           (unsigned long long)bytes);
      return nullptr;
    }
    return m->ptr;
  }

  bool result(int rc) {
    if (!ok) return false;
    if (rc == RDX_OK) return true;
    napi_throw_error(env, nullptr, rdx_last_error());
    ok = false;
    return false;
  }
};

napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v = nullptr;
  bool has = false;
  if (obj && napi_has_named_property(env, obj, name, &has) == napi_ok && has) napi_get_named_property(env, obj, name, &v);
  return v;
}

double prop_num(napi_env env, napi_value obj, const char* name, double fallback) {
  napi_value v = prop(env, obj, name);
  return v ? num(env, v, fallback) : fallback;
}

bool prop_bool(napi_env env, napi_value obj, const char* name, bool fallback) {
  napi_value v = prop(env, obj, name);
  if (!v) return fallback;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b;
  }
  if (t == napi_number) return num(env, v) != 0.0;
  return fallback;
}

napi_value element(napi_env env, napi_value arr, uint32_t i) {
  bool has = false;
  napi_value e = nullptr;
  if (arr && napi_has_element(env, arr, i, &has) == napi_ok && has && napi_get_element(env, arr, i, &e) == napi_ok)
    return e;
  return nullptr;
}

uint32_t length_of(napi_env env, napi_value arr) {
  uint32_t n = 0;
  if (arr) napi_get_array_length(env, arr, &n);
  return n;
}

// 16 numbers (a Matrix4's elements, column-major) from an array or typed array
bool matrix16(napi_env env, napi_value v, double* out) {
  if (!v) return false;
  bool is_array = false;
  napi_is_array(env, v, &is_array);
  bool is_ta = false;
  napi_is_typedarray(env, v, &is_ta);
  if (!is_array && !is_ta) return false;
  for (uint32_t i = 0; i < 16; i++) {
    napi_value e;
    if (napi_get_element(env, v, i, &e) != napi_ok) return false;
    out[i] = num(env, e);
  }
  return true;
}

void identity16(double* m) {
  memset(m, 0, 16 * sizeof(double));
  m[0] = m[5] = m[10] = m[15] = 1.0;
}

napi_value make_external(napi_env env, void* data, napi_finalize fin) {
  napi_value ext;
  napi_create_external(env, data, fin, nullptr, &ext);
  return ext;
}

// ---- keys ----------------------------------------------------------------

napi_value KeyInvalid(napi_env env, napi_callback_info info) {
  ARGS(1);
  return make_u32(env, rdx_key_invalid((int)num(env, argv[0], 24)));
}

napi_value KeyValidMax(napi_env env, napi_callback_info info) {
  ARGS(1);
  return make_u32(env, rdx_key_valid_max((int)num(env, argv[0], 24)));
}

napi_value KeyCompute(napi_env env, napi_callback_info info) {
  ARGS(2);
  return make_u32(env, rdx_key_compute((float)num(env, argv[0]), (int)num(env, argv[1], 24)));
}

// sortKeys(keys, count, bits, ordering) -> visible (synchronous)
napi_value SortKeys(napi_env env, napi_callback_info info) {
  ARGS(4);
  uint32_t *keys, *ordering;
  size_t nk, no;
  if (!u32_array(env, argv[0], &keys, &nk)) return throw_error(env, "SortKeys.sort: keys must be a Uint32Array");
  if (!u32_array(env, argv[3], &ordering, &no)) return throw_error(env, "SortKeys.sort: ordering must be a Uint32Array");
  const double count = num(env, argv[1], (double)nk);
  if (!(count >= 0) || count > (double)nk) return throw_error(env, "SortKeys.sort: count above the keys given");
  uint32_t visible = 0;
  if (!check(env, rdx_sort_keys(thread_context(), keys, (uint32_t)count, (int)num(env, argv[2], 24), ordering, no,
                                &visible)))
    return nullptr;
  return make_u32(env, visible);
}

// sortKeysAsync(keys, count, bits, ordering) -> Promise<visible>, on the libuv pool
struct SortWork {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref keys_ref, ordering_ref;
  uint32_t* keys;
  uint32_t* ordering;
  size_t ordering_len;
  uint32_t count;
  int bits;
  uint32_t visible;
  int rc;
  std::string error;
};

void sort_execute(napi_env, void* data) {
  SortWork* w = (SortWork*)data;
  // the pool thread's own context: its own scratch and stream
  w->rc = rdx_sort_keys(thread_context(), w->keys, w->count, w->bits, w->ordering, w->ordering_len, &w->visible);
  if (w->rc != RDX_OK) w->error = rdx_last_error();
}

void sort_complete(napi_env env, napi_status, void* data) {
  SortWork* w = (SortWork*)data;
  if (w->rc == RDX_OK) {
    napi_resolve_deferred(env, w->deferred, make_u32(env, w->visible));
  } else {
    napi_value msg, err;
    napi_create_string_utf8(env, w->error.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_reference(env, w->keys_ref);
  napi_delete_reference(env, w->ordering_ref);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value SortKeysAsync(napi_env env, napi_callback_info info) {
  ARGS(4);
  uint32_t *keys, *ordering;
  size_t nk, no;
  if (!u32_array(env, argv[0], &keys, &nk) || !u32_array(env, argv[3], &ordering, &no))
    return throw_error(env, "sortSplatKeys: keys and ordering must be Uint32Arrays");
  const double count = num(env, argv[1], (double)nk);
  if (!(count >= 0) || count > (double)nk) return throw_error(env, "SortKeys.sort: count above the keys given");
  SortWork* w = new SortWork();
  w->keys = keys;
  w->ordering = ordering;
  w->ordering_len = no;
  w->count = (uint32_t)count;
  w->bits = (int)num(env, argv[2], 24);
  napi_create_reference(env, argv[0], 1, &w->keys_ref);
  napi_create_reference(env, argv[3], 1, &w->ordering_ref);
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_string_utf8(env, "radix.sortKeys", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, sort_execute, sort_complete, w, &w->work);
  napi_queue_async_work(env, w->work);
  return promise;
}

// ---- streams ---------------------------------------------------------------

void stream_finalize(napi_env, void* data, void*) {
  GpuStream* s = (GpuStream*)data;
  if (s->ctx) rdx_context_destroy(s->ctx);
  if (s->stream) rdx_stream_destroy(s->stream);
  delete s;
}

napi_value StreamCreate(napi_env env, napi_callback_info) {
  GpuStream* s = new GpuStream();
  if (!check(env, rdx_stream_create(&s->stream)) || !check(env, rdx_context_create(&s->ctx))) {
    stream_finalize(env, s, nullptr);
    return nullptr;
  }
  return make_external(env, s, stream_finalize);
}

napi_value StreamSynchronize(napi_env env, napi_callback_info info) {
  ARGS(1);
  Call c(env, "GpuStream.synchronize", argv[0]);
  if (c.ok) c.result(rdx_stream_synchronize(c.stream));
  return nullptr;
}

// streamSynchronizeAsync(stream) -> Promise resolved when the stream's work is done (waited on the pool)
struct SyncWork {
  napi_async_work work;
  napi_deferred deferred;
  napi_ref stream_ref;
  void* stream;
  int rc;
  std::string error;
};

void sync_execute(napi_env, void* data) {
  SyncWork* w = (SyncWork*)data;
  w->rc = rdx_stream_synchronize(w->stream);
  if (w->rc != RDX_OK) w->error = rdx_last_error();
}

void sync_complete(napi_env env, napi_status, void* data) {
  SyncWork* w = (SyncWork*)data;
  if (w->rc == RDX_OK) {
    napi_value undef;
    napi_get_undefined(env, &undef);
    napi_resolve_deferred(env, w->deferred, undef);
  } else {
    napi_value msg, err;
    napi_create_string_utf8(env, w->error.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &err);
    napi_reject_deferred(env, w->deferred, err);
  }
  napi_delete_reference(env, w->stream_ref);
  napi_delete_async_work(env, w->work);
  delete w;
}

napi_value StreamSynchronizeAsync(napi_env env, napi_callback_info info) {
  ARGS(1);
  GpuStream* s = external<GpuStream>(env, argv[0], T_STREAM);
  if (!s) return throw_error(env, "GpuStream.done: not a stream");
  SyncWork* w = new SyncWork();
  w->stream = s->stream;
  napi_create_reference(env, argv[0], 1, &w->stream_ref);
  napi_value promise, name;
  napi_create_promise(env, &w->deferred, &promise);
  napi_create_string_utf8(env, "radix.streamSync", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, sync_execute, sync_complete, w, &w->work);
  napi_queue_async_work(env, w->work);
  return promise;
}

// ---- device memory ---------------------------------------------------------

void device_finalize(napi_env, void* data, void*) {
  DeviceMem* m = (DeviceMem*)data;
  if (m->ptr) rdx_device_free(m->ptr);
  delete m;
}

napi_value DeviceAlloc(napi_env env, napi_callback_info info) {
  ARGS(1);
  const double bytes = num(env, argv[0], -1);
  if (!(bytes >= 0)) return throw_error(env, "deviceAlloc: bytes must be >= 0");
  DeviceMem* m = new DeviceMem{T_MEM, nullptr, (size_t)bytes};
  if (!check(env, rdx_device_alloc(m->bytes, &m->ptr))) {
    delete m;
    return nullptr;
  }
  return make_external(env, m, device_finalize);
}

napi_value DeviceFree(napi_env env, napi_callback_info info) {
  ARGS(1);
  DeviceMem* m = device(env, argv[0]);
  if (m && m->ptr) {
    rdx_device_free(m->ptr);
    m->ptr = nullptr;
    m->bytes = 0;
  }
  return nullptr;
}

// deviceUpload(mem, typedArray, byteOffset, stream?)
napi_value DeviceUpload(napi_env env, napi_callback_info info) {
  ARGS(4);
  Call c(env, "deviceUpload", argv[3]);
  void* src;
  size_t bytes;
  if (!bytes_of(env, argv[1], &src, &bytes)) return throw_error(env, "deviceUpload: (buffer, typedArray, byteOffset)");
  const double off = num(env, argv[2], 0);
  if (!(off >= 0)) return throw_error(env, "deviceUpload: byteOffset must be >= 0");
  char* dst = (char*)c.buf(argv[0], "buffer", (uint64_t)off + bytes);
  if (c.ok) c.result(rdx_device_upload(c.stream, dst + (size_t)off, src, bytes));
  return nullptr;
}

// deviceDownload(mem, byteOffset, typedArray, stream?)
napi_value DeviceDownload(napi_env env, napi_callback_info info) {
  ARGS(4);
  Call c(env, "deviceDownload", argv[3]);
  void* dst;
  size_t bytes;
  if (!bytes_of(env, argv[2], &dst, &bytes)) return throw_error(env, "deviceDownload: (buffer, byteOffset, typedArray)");
  const double off = num(env, argv[1], 0);
  if (!(off >= 0)) return throw_error(env, "deviceDownload: byteOffset must be >= 0");
  const char* src = (const char*)c.buf(argv[0], "buffer", (uint64_t)off + bytes);
  if (c.ok) c.result(rdx_device_download(c.stream, dst, src + (size_t)off, bytes));
  return nullptr;
}

// deviceFillU32(mem, value, count, stream?)
napi_value DeviceFillU32(napi_env env, napi_callback_info info) {
  ARGS(4);
  Call c(env, "deviceFillU32", argv[3]);
  const double count = num(env, argv[2], -1);
  if (!(count >= 0)) return throw_error(env, "deviceFillU32: count must be >= 0");
  uint32_t* dst = (uint32_t*)c.buf(argv[0], "buffer", 4 * (uint64_t)count);
  if (c.ok) c.result(rdx_device_fill_u32(c.stream, dst, (uint32_t)num(env, argv[1]), (size_t)count));
  return nullptr;
}

// ---- GPU sorts ---------------------------------------------------------------

uint32_t count_arg(Call& c, napi_value v, const char* what) {
  const double d = num(c.env, v, -1);
  if (!(d >= 0) || d > 4294967295.0) c.fail("%s: %s must be a count (0 .. 2^32 - 1)", c.who, what);
  return c.ok ? (uint32_t)d : 0u;
}

// sortPairs(keys, values, count, bits, stream?)
napi_value SortPairs(napi_env env, napi_callback_info info) {
  ARGS(5);
  Call c(env, "WgpuRadixSort.sort", argv[4]);
  const uint32_t n = count_arg(c, argv[2], "count");
  uint32_t* k = (uint32_t*)c.buf(argv[0], "keys", 4ull * n);
  uint32_t* v = (uint32_t*)c.buf(argv[1], "values", 4ull * n);
  if (c.ok) c.result(rdx_sort_pairs(c.ctx, c.stream, k, v, n, (int)num(env, argv[3], 32)));
  return nullptr;
}

// sortPairsIndirect(keys, values, countBuffer, capacity, bits, stream?)
napi_value SortPairsIndirect(napi_env env, napi_callback_info info) {
  ARGS(6);
  Call c(env, "WgpuRadixSort.sortIndirect", argv[5]);
  const uint32_t cap = count_arg(c, argv[3], "capacity");
  uint32_t* k = (uint32_t*)c.buf(argv[0], "keys", 4ull * cap);
  uint32_t* v = (uint32_t*)c.buf(argv[1], "values", 4ull * cap);
  const uint32_t* cnt = (const uint32_t*)c.buf(argv[2], "countBuffer", 4);
  if (c.ok) c.result(rdx_sort_pairs_indirect(c.ctx, c.stream, k, v, cnt, cap, (int)num(env, argv[4], 32)));
  return nullptr;
}

// sortSplats(keys, n, bits, indices, count, drawArgs, stream?)
napi_value SortSplats(napi_env env, napi_callback_info info) {
  ARGS(7);
  Call c(env, "WgpuSplatSorter.sort", argv[6]);
  const uint32_t n = count_arg(c, argv[1], "count");
  const uint32_t* k = (const uint32_t*)c.buf(argv[0], "keys", 4ull * n);
  uint32_t* idx = (uint32_t*)c.buf(argv[3], "values", 4ull * n);
  uint32_t* cnt = (uint32_t*)c.buf(argv[4], "count", 4, true);
  uint32_t* args = (uint32_t*)c.buf(argv[5], "indirect", 20, true);
  if (c.ok) c.result(rdx_sort_splats(c.ctx, c.stream, k, n, (int)num(env, argv[2], 24), idx, cnt, args));
  return nullptr;
}

// ---- projection ----------------------------------------------------------------

// A table descriptor, every buffer checked for the table's count.
void read_table(Call& c, napi_value t, rdx_splat_table* out) {
  napi_env env = c.env;
  memset(out, 0, sizeof *out);
  identity16(out->object_to_world);
  if (absent(env, t)) {
    c.fail("%s: no splat table", c.who);
    return;
  }
  out->layout = (int)prop_num(env, t, "layout", 0);
  const double count = prop_num(env, t, "count", -1);
  if (!(count >= 0) || count > 4294967295.0) {
    c.fail("%s: table.count must be a count", c.who);
    return;
  }
  out->count = (uint32_t)count;
  out->plane_stride = (size_t)prop_num(env, t, "planeStride", 0);
  out->sh_codec = (int)prop_num(env, t, "shCodec", 0);
  out->sh_degree = (int)prop_num(env, t, "shDegree", 0);
  const uint64_t n = out->count;
  const int deg = out->sh_codec == RDX_SH_NONE ? 0 : out->sh_degree;
  const int K = deg >= 3 ? 15 : (deg == 2 ? 8 : (deg == 1 ? 3 : 0));
  if (out->layout == RDX_LAYOUT_F32_SOA) {
    // one device buffer of f32 planes: centre 3, log-scale 3, quat 4, opacity, rgb 3, SH 3K
    const uint64_t st = (out->plane_stride ? out->plane_stride : n) * 4;
    const int planes = 14 + (out->sh_codec == RDX_SH_F32 ? 3 * K : 0);
    char* base = (char*)c.buf(prop(env, t, "planes"), "table.planes", planes * st, n == 0);
    if (base) {
      out->center = (const float*)base;
      out->log_scale = (const float*)(base + 3 * st);
      out->quat = (const float*)(base + 6 * st);
      out->opacity = (const float*)(base + 10 * st);
      out->rgb = (const float*)(base + 11 * st);
      out->sh = (const float*)(base + 14 * st);
    }
  } else if (out->layout == RDX_LAYOUT_EXT) {
    out->ext_a = (const uint32_t*)c.buf(prop(env, t, "extA"), "table.extA", 16 * n, n == 0);
    out->ext_b = (const uint32_t*)c.buf(prop(env, t, "extB"), "table.extB", 16 * n, n == 0);
  } else if (out->layout == RDX_LAYOUT_PACKED) {
    out->packed = (const uint32_t*)c.buf(prop(env, t, "packed"), "table.packed", 16 * n, n == 0);
  } else if (out->layout == RDX_LAYOUT_COV) {
    out->ext_a = (const uint32_t*)c.buf(prop(env, t, "covA"), "table.covA", 16 * n, n == 0);
    out->ext_b = (const uint32_t*)c.buf(prop(env, t, "covB"), "table.covB", 16 * n, n == 0);
  }
  out->packed_encoding[0] = (float)prop_num(env, t, "rgbMin", 0.0);
  out->packed_encoding[1] = (float)prop_num(env, t, "rgbMax", 1.0);
  out->packed_encoding[2] = (float)prop_num(env, t, "lnScaleMin", -12.0);
  out->packed_encoding[3] = (float)prop_num(env, t, "lnScaleMax", 9.0);
  if (out->layout != RDX_LAYOUT_F32_SOA && deg > 0) {
    const int texels = out->sh_codec == RDX_SH_EXT ? (deg == 1 ? 1 : deg == 2 ? 2 : 4) : deg;
    napi_value ch = prop(env, t, "shChannels");
    for (int k = 0; k < texels && k < 4; k++) {
      char name[32];
      snprintf(name, sizeof name, "table.shChannels[%d]", k);
      out->sh_channels[k] = (const uint32_t*)c.buf(element(env, ch, (uint32_t)k), name, 16 * n, n == 0);
    }
  }
  if (!matrix16(env, prop(env, t, "objectToWorld"), out->object_to_world)) identity16(out->object_to_world);
}

bool read_camera(Call& c, napi_value cam, rdx_camera* out) {
  napi_env env = c.env;
  memset(out, 0, sizeof *out);
  if (!c.ok) return false;
  napi_value eyes = prop(env, cam, "eyes");
  const uint32_t ne = length_of(env, eyes);
  if (ne < 1 || ne > 2) return c.fail("SplatCamera: 1 or 2 eyes");
  out->eyes = (int)ne;
  for (uint32_t e = 0; e < ne; e++) {
    napi_value eye = element(env, eyes, e);
    if (!matrix16(env, prop(env, eye, "view"), out->eye[e].view) ||
        !matrix16(env, prop(env, eye, "projection"), out->eye[e].projection))
      return c.fail("camera eyes need view and projection (16 elements each)");
    out->eye[e].ortho = prop_bool(env, eye, "ortho", false) ? 1 : 0;
  }
  if (!matrix16(env, prop(env, cam, "head"), out->head)) memcpy(out->head, out->eye[0].view, sizeof out->head);
  napi_value size = prop(env, cam, "size");
  out->size[0] = num(env, element(env, size, 0), 1);
  out->size[1] = num(env, element(env, size, 1), 1);
  out->ndc_z_min = prop_num(env, cam, "ndcZMin", 0.0);
  return true;
}

void read_params(napi_env env, napi_value p, rdx_params* out) {
  rdx_params_default(out);
  if (absent(env, p)) return;
  out->max_std_dev = prop_num(env, p, "maxStdDev", out->max_std_dev);
  out->min_alpha = prop_num(env, p, "minAlpha", out->min_alpha);
  out->falloff = prop_num(env, p, "falloff", out->falloff);
  out->lod_inflate = prop_bool(env, p, "lodInflate", out->lod_inflate != 0) ? 1 : 0;
  out->blur_amount = prop_num(env, p, "blurAmount", out->blur_amount);
  out->pre_blur_amount = prop_num(env, p, "preBlurAmount", out->pre_blur_amount);
  out->clip_xy = prop_num(env, p, "clipXY", out->clip_xy);
  out->min_pixel_radius = prop_num(env, p, "minPixelRadius", out->min_pixel_radius);
  out->max_pixel_radius = prop_num(env, p, "maxPixelRadius", out->max_pixel_radius);
  out->min_contribution = prop_num(env, p, "minContribution", out->min_contribution);
  out->focal_distance = prop_num(env, p, "focalDistance", out->focal_distance);
  out->aperture_tan = prop_num(env, p, "apertureTan", out->aperture_tan);
  out->sort_guard = prop_num(env, p, "sortGuard", out->sort_guard);
  out->max_sh_degree = (int)prop_num(env, p, "maxShDegree", out->max_sh_degree);
  out->sort_bits = (int)prop_num(env, p, "sortBits", out->sort_bits);
  out->cull_keys = prop_bool(env, p, "cullKeys", out->cull_keys != 0) ? 1 : 0;
  out->key_per_eye = prop_bool(env, p, "keyPerEye", out->key_per_eye != 0) ? 1 : 0;
}

// projectEwa(table, camera, params, words[], keys, at, stream?)
napi_value ProjectEwa(napi_env env, napi_callback_info info) {
  ARGS(7);
  Call c(env, "SplatProjector.project", argv[6]);
  rdx_splat_table t;
  rdx_camera cam;
  rdx_params p;
  read_table(c, argv[0], &t);
  read_camera(c, argv[1], &cam);
  read_params(env, argv[2], &p);
  const uint32_t at = count_arg(c, argv[5], "at");
  const uint64_t end = (uint64_t)at + t.count;
  uint32_t* words[2] = {nullptr, nullptr};
  for (int e = 0; c.ok && e < cam.eyes; e++)
    words[e] = (uint32_t*)c.buf(element(env, argv[3], (uint32_t)e), e ? "words[1]" : "words[0]", 16 * end, true);
  // keys of eye 1 (per-eye keys, stereo) at keys + count: key_stride = count
  const bool two = p.key_per_eye && cam.eyes == 2;
  uint32_t* keys = (uint32_t*)c.buf(argv[4], "keys", 4 * (end + (two ? t.count : 0)), true);
  if (c.ok) c.result(rdx_project_ewa(c.ctx, c.stream, &t, &cam, &p, at, words, keys, t.count, nullptr));
  return nullptr;
}

// projectSort(table, camera, params, words[], keys, indices, count, drawArgs, stream?)
napi_value ProjectSort(napi_env env, napi_callback_info info) {
  ARGS(9);
  Call c(env, "SplatProjector.projectSort", argv[8]);
  rdx_splat_table t;
  rdx_camera cam;
  rdx_params p;
  read_table(c, argv[0], &t);
  read_camera(c, argv[1], &cam);
  read_params(env, argv[2], &p);
  const uint64_t segs = p.key_per_eye && cam.eyes == 2 ? 2 : 1, n = t.count;
  uint32_t* words[2] = {nullptr, nullptr};
  for (int e = 0; c.ok && e < cam.eyes; e++)
    words[e] = (uint32_t*)c.buf(element(env, argv[3], (uint32_t)e), e ? "words[1]" : "words[0]", 16 * n, true);
  uint32_t* keys = (uint32_t*)c.buf(argv[4], "keys", 4 * n * segs, true);
  uint32_t* idx = (uint32_t*)c.buf(argv[5], "indices", 4 * n * segs);
  uint32_t* cnt = (uint32_t*)c.buf(argv[6], "count", 4 * segs, true);
  uint32_t* args = (uint32_t*)c.buf(argv[7], "indirect", 20 * segs, true);
  if (c.ok) c.result(rdx_project_sort(c.ctx, c.stream, &t, &cam, &p, words, keys, idx, cnt, args, nullptr));
  return nullptr;
}

// projectKeys(table, camera, params, keys, at, stream?): the keys-only pass
napi_value ProjectKeys(napi_env env, napi_callback_info info) {
  ARGS(6);
  Call c(env, "SplatProjector.projectKeys", argv[5]);
  rdx_splat_table t;
  rdx_camera cam;
  rdx_params p;
  read_table(c, argv[0], &t);
  read_camera(c, argv[1], &cam);
  read_params(env, argv[2], &p);
  const uint32_t at = count_arg(c, argv[4], "at");
  uint32_t* keys = (uint32_t*)c.buf(argv[3], "keys", 4 * ((uint64_t)at + t.count));
  if (c.ok) c.result(rdx_project_keys(c.ctx, c.stream, &t, &cam, &p, at, keys));
  return nullptr;
}

// projectRay(table, camera, params, method, channels[], keys, at, stream?): 3dgut (1), htgs (2), raygs (3)
napi_value ProjectRay(napi_env env, napi_callback_info info) {
  ARGS(8);
  Call c(env, "SplatProjector.projectRay", argv[7]);
  rdx_splat_table t;
  rdx_camera cam;
  rdx_params p;
  read_table(c, argv[0], &t);
  read_camera(c, argv[1], &cam);
  read_params(env, argv[2], &p);
  const uint32_t at = count_arg(c, argv[6], "at");
  const uint64_t end = (uint64_t)at + t.count;
  // channels w0, w1 (uint32x4) and with two eyes w2 (uint32x2)
  uint32_t* channels[3] = {nullptr, nullptr, nullptr};
  const char* names[3] = {"channels[0]", "channels[1]", "channels[2]"};
  for (int k = 0; c.ok && k < (cam.eyes == 2 ? 3 : 2); k++)
    channels[k] = (uint32_t*)c.buf(element(env, argv[4], (uint32_t)k), names[k], (k == 2 ? 8 : 16) * end, true);
  uint32_t* keys = (uint32_t*)c.buf(argv[5], "keys", 4 * end, true);
  if (c.ok) c.result(rdx_project_ray(c.ctx, c.stream, &t, &cam, &p, (int)num(env, argv[3], 0), at, channels, keys));
  return nullptr;
}

// untransposeBytes(planes, count, elementBytes, out, outOffsetBytes, stream?): an RDX payload's planes back to elements
napi_value UntransposeBytes(napi_env env, napi_callback_info info) {
  ARGS(6);
  Call c(env, "untransposeBytes", argv[5]);
  const uint32_t n = count_arg(c, argv[1], "count");
  const uint32_t eb = count_arg(c, argv[2], "elementBytes");
  const double off = num(env, argv[4], 0);
  if (c.ok && !(off >= 0)) c.fail("untransposeBytes: outOffset must be >= 0");
  const uint8_t* planes = (const uint8_t*)c.buf(argv[0], "planes", (uint64_t)n * eb);
  char* out = (char*)c.buf(argv[3], "out", (uint64_t)off + (uint64_t)n * eb);
  if (c.ok) c.result(rdx_untranspose_bytes(c.stream, planes, n, eb, out + (size_t)off));
  return nullptr;
}

// spzDecode(body, bodyBytes, count, version, shDegree, fractionalBits, planes, planeStride, stream?):
// an SPZ body (the gunzipped file after its 16-byte header) to f32 planes
napi_value SpzDecode(napi_env env, napi_callback_info info) {
  ARGS(9);
  Call c(env, "readSpz", argv[8]);
  const double body_bytes = num(env, argv[1], -1);
  if (!(body_bytes >= 0)) c.fail("readSpz: bodyBytes must be >= 0");
  const uint32_t n = count_arg(c, argv[2], "count");
  const int deg = (int)num(env, argv[4], 0);
  const uint64_t stride = (uint64_t)num(env, argv[7], 0);
  if (c.ok && stride < n) c.fail("readSpz: planeStride %llu below the count %u", (unsigned long long)stride, n);
  const int K = deg >= 3 ? 15 : (deg == 2 ? 8 : (deg == 1 ? 3 : 0));
  const uint8_t* body = (const uint8_t*)c.buf(argv[0], "body", (uint64_t)body_bytes);
  float* planes = (float*)c.buf(argv[6], "planes", 4 * (uint64_t)(14 + 3 * K) * stride);
  if (c.ok)
    c.result(rdx_spz_decode(c.stream, body, (size_t)body_bytes, n, (int)num(env, argv[3], 0), deg,
                            (int)num(env, argv[5], 0), planes, (size_t)stride));
  return nullptr;
}

// ---- the differentiable compositor -------------------------------------------

void raster_finalize(napi_env, void* data, void*) {
  RasterH* r = (RasterH*)data;
  if (r->h) rdx_raster_destroy(r->h);
  delete r;
}

// rasterCreate({shDegree, tStop, maxEntriesPerSplat}, stream?)
napi_value RasterCreate(napi_env env, napi_callback_info info) {
  ARGS(2);
  Call c(env, "Rasterizer", argv[1]);
  if (!c.ok) return nullptr;
  rdx_raster_settings s;
  s.sh_degree = (int)prop_num(env, argv[0], "shDegree", 3);
  s.t_stop = (float)prop_num(env, argv[0], "tStop", 1e-4);
  s.max_entries_per_splat = (uint32_t)prop_num(env, argv[0], "maxEntriesPerSplat", 0);
  RasterH* r = new RasterH();
  if (!c.ok || !c.result(rdx_raster_create(c.ctx, &s, &r->h))) {
    delete r;
    return nullptr;
  }
  return make_external(env, r, raster_finalize);
}

// rasterForward(raster, params, n, camera, renderParams, image, alpha, stream?): params 59 planes of n floats
napi_value RasterForward(napi_env env, napi_callback_info info) {
  ARGS(8);
  Call c(env, "Rasterizer.forward", argv[7]);
  RasterH* r = external<RasterH>(env, argv[0], T_RASTER);
  if (!r) return throw_error(env, "Rasterizer.forward: not a rasterizer");
  const uint32_t n = count_arg(c, argv[2], "count");
  rdx_camera cam;
  rdx_params p;
  read_camera(c, argv[3], &cam);
  read_params(env, argv[4], &p);
  const uint64_t px = (uint64_t)(cam.size[0] > 0 ? cam.size[0] : 0) * (uint64_t)(cam.size[1] > 0 ? cam.size[1] : 0);
  const float* params = (const float*)c.buf(argv[1], "params", 59ull * 4 * n, n == 0);
  float* image = (float*)c.buf(argv[5], "image", 12 * px);
  float* alpha = (float*)c.buf(argv[6], "alpha", 4 * px, true);
  r->forward = false;
  if (!c.ok || !c.result(rdx_raster_forward(r->h, c.stream, params, n, &cam, &p, image, alpha))) return nullptr;
  r->n = n;
  r->width = (uint32_t)cam.size[0];
  r->height = (uint32_t)cam.size[1];
  r->forward = true;
  return nullptr;
}

// rasterBackward(raster, gradImage, gradParams, stream?): the adjoint of the last forward
napi_value RasterBackward(napi_env env, napi_callback_info info) {
  ARGS(4);
  Call c(env, "Rasterizer.backward", argv[3]);
  RasterH* r = external<RasterH>(env, argv[0], T_RASTER);
  if (!r) return throw_error(env, "Rasterizer.backward: not a rasterizer");
  if (!r->forward) return throw_error(env, "rdx_raster_backward: no forward to differentiate");
  const float* g = (const float*)c.buf(argv[1], "gradImage", 12ull * r->width * r->height);
  float* out = (float*)c.buf(argv[2], "gradParams", 59ull * 4 * std::max<uint32_t>(r->n, 1));
  if (c.ok) c.result(rdx_raster_backward(r->h, c.stream, g, out));
  return nullptr;
}

// rasterStats(raster, stream?) -> {visible, entries, capacity}
napi_value RasterStats(napi_env env, napi_callback_info info) {
  ARGS(2);
  Call c(env, "Rasterizer.stats", argv[1]);
  RasterH* r = external<RasterH>(env, argv[0], T_RASTER);
  if (!r) return throw_error(env, "Rasterizer.stats: not a rasterizer");
  uint32_t v = 0, e = 0, cap = 0;
  int ov = 0;
  if (!c.ok || !c.result(rdx_raster_stats(r->h, c.stream, &v, &e, &cap, &ov))) return nullptr;
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "visible", make_u32(env, v));
  napi_set_named_property(env, o, "entries", make_u32(env, e));
  napi_set_named_property(env, o, "capacity", make_u32(env, cap));
  return o;
}

napi_value RasterDestroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  RasterH* r = external<RasterH>(env, argv[0], T_RASTER);
  if (r && r->h) {
    rdx_raster_destroy(r->h);
    r->h = nullptr;
    r->forward = false;
  }
  return nullptr;
}

// ---- the draw -----------------------------------------------------------------

void draw_finalize(napi_env, void* data, void*) {
  DrawH* d = (DrawH*)data;
  if (d->h) rdx_draw_destroy(d->h);
  delete d;
}

// drawCreate({tStop, linearOutput, maxEntriesPerSplat}, stream?)
napi_value DrawCreate(napi_env env, napi_callback_info info) {
  ARGS(2);
  Call c(env, "SplatDrawer", argv[1]);
  if (!c.ok) return nullptr;
  rdx_draw_settings s;
  s.t_stop = (float)prop_num(env, argv[0], "tStop", 0.0);
  s.linear_output = prop_bool(env, argv[0], "linearOutput", false) ? 1 : 0;
  s.max_entries_per_splat = (uint32_t)prop_num(env, argv[0], "maxEntriesPerSplat", 0);
  DrawH* d = new DrawH();
  if (!c.ok || !c.result(rdx_draw_create(c.ctx, &s, &d->h))) {
    delete d;
    return nullptr;
  }
  return make_external(env, d, draw_finalize);
}

// drawSplats(draw, words, n, order, count, capacity, width, height, params, rgba, stream?)
napi_value DrawSplats(napi_env env, napi_callback_info info) {
  ARGS(11);
  Call c(env, "SplatDrawer.draw", argv[10]);
  DrawH* d = external<DrawH>(env, argv[0], T_DRAW);
  if (!d || !d->h) return throw_error(env, "SplatDrawer.draw: not a drawer");
  const uint32_t n = count_arg(c, argv[2], "n");
  const uint32_t cap = count_arg(c, argv[5], "capacity");
  const uint32_t w = count_arg(c, argv[6], "width"), h = count_arg(c, argv[7], "height");
  rdx_params p;
  read_params(env, argv[8], &p);
  const uint32_t* words = (const uint32_t*)c.buf(argv[1], "words", 16ull * n, n == 0);
  const uint32_t* order = (const uint32_t*)c.buf(argv[3], "order", 4ull * cap, cap == 0);
  const uint32_t* cnt = (const uint32_t*)c.buf(argv[4], "count", 4);
  float* rgba = (float*)c.buf(argv[9], "rgba", 16ull * w * h);
  if (c.ok) c.result(rdx_draw_splats(d->h, c.stream, words, n, order, cnt, cap, w, h, &p, rgba));
  return nullptr;
}

// drawRay(draw, method, channels[], n, order, count, capacity, camera, eye, params, rgba, stream?): one eye of
// the 3dgut (1), htgs (2) or raygs (3) words at the camera's render size
napi_value DrawRay(napi_env env, napi_callback_info info) {
  ARGS(12);
  Call c(env, "SplatDrawer.drawRay", argv[11]);
  DrawH* d = external<DrawH>(env, argv[0], T_DRAW);
  if (!d || !d->h) return throw_error(env, "SplatDrawer.drawRay: not a drawer");
  rdx_camera cam;
  rdx_params p;
  read_camera(c, argv[7], &cam);
  read_params(env, argv[9], &p);
  const uint32_t n = count_arg(c, argv[3], "n");
  const uint32_t cap = count_arg(c, argv[6], "capacity");
  const double w = cam.size[0], h = cam.size[1];
  if (c.ok && !(w >= 1 && h >= 1 && w <= 16384 && h <= 16384))
    c.fail("SplatDrawer.drawRay: render size %gx%g (whole pixels, 1..16384)", w, h);
  uint32_t* channels[3] = {nullptr, nullptr, nullptr};
  const char* names[3] = {"channels[0]", "channels[1]", "channels[2]"};
  for (int k = 0; c.ok && k < (cam.eyes == 2 ? 3 : 2); k++)
    channels[k] = (uint32_t*)c.buf(element(env, argv[2], (uint32_t)k), names[k], (k == 2 ? 8ull : 16ull) * n, n == 0);
  const uint32_t* order = (const uint32_t*)c.buf(argv[4], "order", 4ull * cap, cap == 0);
  const uint32_t* cnt = (const uint32_t*)c.buf(argv[5], "count", 4);
  float* rgba = c.ok ? (float*)c.buf(argv[10], "rgba", 16ull * (uint64_t)w * (uint64_t)h) : nullptr;
  if (c.ok)
    c.result(rdx_draw_ray(d->h, c.stream, (int)num(env, argv[1], 0), channels, n, order, cnt, cap, &cam,
                          (int)num(env, argv[8], 0), &p, rgba));
  return nullptr;
}

// drawStats(draw, stream?) -> {drawn, entries, capacity}
napi_value DrawStats(napi_env env, napi_callback_info info) {
  ARGS(2);
  Call c(env, "SplatDrawer.stats", argv[1]);
  DrawH* d = external<DrawH>(env, argv[0], T_DRAW);
  if (!d || !d->h) return throw_error(env, "SplatDrawer.stats: not a drawer");
  uint32_t dr = 0, e = 0, cap = 0;
  int ov = 0;
  if (!c.ok || !c.result(rdx_draw_stats(d->h, c.stream, &dr, &e, &cap, &ov))) return nullptr;
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "drawn", make_u32(env, dr));
  napi_set_named_property(env, o, "entries", make_u32(env, e));
  napi_set_named_property(env, o, "capacity", make_u32(env, cap));
  return o;
}

napi_value DrawDestroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  DrawH* d = external<DrawH>(env, argv[0], T_DRAW);
  if (d && d->h) {
    rdx_draw_destroy(d->h);
    d->h = nullptr;
  }
  return nullptr;
}

// ---- batched views -----------------------------------------------------------------

void batch_finalize(napi_env, void* data, void*) {
  BatchH* b = (BatchH*)data;
  if (b->h) rdx_batch_destroy(b->h);
  delete b;
}

// batchCreate(tables[], views[{table, camera, words?}], params, keys?, stream?, tuning?)
// tuning: [digitBits, keysPerThread, blockThreads] for this batch's sort (the
// persistent form needs [8, 12, 512]); the context's own tuning is restored.
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  ARGS(6);
  Call c(env, "SplatBatch", argv[4]);
  const uint32_t nt = length_of(env, argv[0]), nv = length_of(env, argv[1]);
  if (c.ok && (nt < 1 || nv < 1 || nv > RDX_BATCH_MAX_VIEWS))
    c.fail("SplatBatch: 1 or more tables and 1 to %d views", RDX_BATCH_MAX_VIEWS);
  std::vector<rdx_splat_table> tables(nt);
  for (uint32_t k = 0; c.ok && k < nt; k++) read_table(c, element(env, argv[0], k), &tables[k]);
  std::vector<rdx_batch_view> views(nv);
  uint64_t total = 0;
  for (uint32_t v = 0; c.ok && v < nv; v++) {
    napi_value view = element(env, argv[1], v);
    memset(&views[v], 0, sizeof views[v]);
    const double t = prop_num(env, view, "table", -1);
    if (!(t >= 0) || t >= nt) {
      c.fail("SplatBatch: view %u names table %g of %u", v, t, nt);
      break;
    }
    views[v].table = (uint32_t)t;
    if (!read_camera(c, prop(env, view, "camera"), &views[v].camera)) break;
    const uint64_t n = tables[(uint32_t)t].count;
    views[v].words_dev = (uint32_t*)c.buf(prop(env, view, "words"), "view.words", 16 * n, true);
    total += n;
  }
  rdx_params p;
  read_params(env, argv[2], &p);
  uint32_t* keys = (uint32_t*)c.buf(argv[3], "keys", 4 * total, true);
  if (!c.ok) return nullptr;
  rdx_sort_tuning saved, want;
  const bool tuned = !absent(env, argv[5]);
  if (tuned) {
    if (length_of(env, argv[5]) != 3) return throw_error(env, "SplatBatch: tuning is [digitBits, keysPerThread, blockThreads]");
    want.digit_bits = (int)num(env, element(env, argv[5], 0), 0.0);
    want.keys_per_thread = (int)num(env, element(env, argv[5], 1), 0.0);
    want.block_threads = (int)num(env, element(env, argv[5], 2), 0.0);
    if (!c.result(rdx_context_get_tuning(c.ctx, &saved)) || !c.result(rdx_context_set_tuning(c.ctx, &want)))
      return nullptr;
  }
  BatchH* b = new BatchH();
  const bool made = c.result(rdx_batch_create(c.ctx, tables.data(), (int)nt, views.data(), (int)nv, &p, keys, &b->h));
  if (tuned) (void)rdx_context_set_tuning(c.ctx, &saved);
  if (!made) {
    delete b;
    return nullptr;
  }
  b->total = rdx_batch_total(b->h);
  b->views = (int)nv;
  napi_value ext = make_external(env, b, batch_finalize);
  return ext;
}

// batchInfo(batch) -> {total, offsets[]}
napi_value BatchInfo(napi_env env, napi_callback_info info) {
  ARGS(1);
  BatchH* b = external<BatchH>(env, argv[0], T_BATCH);
  if (!b || !b->h) return throw_error(env, "SplatBatch: not a batch");
  napi_value o, offs;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "total", make_num(env, (double)b->total));
  napi_create_array_with_length(env, (size_t)b->views, &offs);
  for (int v = 0; v < b->views; v++)
    napi_set_element(env, offs, (uint32_t)v, make_num(env, (double)rdx_batch_segment_offset(b->h, v)));
  napi_set_named_property(env, o, "offsets", offs);
  return o;
}

// batchSetCameras(batch, cameras[])
napi_value BatchSetCameras(napi_env env, napi_callback_info info) {
  ARGS(2);
  Call c(env, "SplatBatch.setCameras", nullptr);
  BatchH* b = external<BatchH>(env, argv[0], T_BATCH);
  if (!b || !b->h) return throw_error(env, "SplatBatch.setCameras: not a batch");
  const uint32_t n = length_of(env, argv[1]);
  if ((int)n != b->views) return throw_error(env, "SplatBatch.setCameras: one camera per view");
  std::vector<rdx_camera> cams(n);
  for (uint32_t v = 0; c.ok && v < n; v++) read_camera(c, element(env, argv[1], v), &cams[v]);
  if (c.ok) c.result(rdx_batch_set_cameras(b->h, cams.data(), (int)n));
  return nullptr;
}

// batchRun(batch, indices, counts, drawArgs, graph, persistent, stream?)
napi_value BatchRun(napi_env env, napi_callback_info info) {
  ARGS(7);
  Call c(env, "SplatBatch.run", argv[6]);
  BatchH* b = external<BatchH>(env, argv[0], T_BATCH);
  if (!b || !b->h) return throw_error(env, "SplatBatch.run: not a batch");
  uint32_t* idx = (uint32_t*)c.buf(argv[1], "indices", 4 * std::max<uint64_t>(b->total, 1));
  uint32_t* cnt = (uint32_t*)c.buf(argv[2], "counts", 4ull * b->views, true);
  uint32_t* args = (uint32_t*)c.buf(argv[3], "drawArgs", 20ull * b->views, true);
  bool graph = false;
  napi_valuetype t;
  if (!absent(env, argv[4]) && napi_typeof(env, argv[4], &t) == napi_ok && t == napi_boolean)
    napi_get_value_bool(env, argv[4], &graph);
  bool persistent = false;
  if (!absent(env, argv[5]) && napi_typeof(env, argv[5], &t) == napi_ok && t == napi_boolean)
    napi_get_value_bool(env, argv[5], &persistent);
  if (c.ok)
    c.result(rdx_batch_run(b->h, c.stream, idx, cnt, args,
                           (graph ? RDX_BATCH_GRAPH : 0) | (persistent ? RDX_BATCH_PERSISTENT : 0)));
  return nullptr;
}

// batchStatus(batch): waits for the last run; throws if a persistent run's barrier wait ran out
napi_value BatchStatus(napi_env env, napi_callback_info info) {
  ARGS(1);
  Call c(env, "SplatBatch.status", nullptr);
  BatchH* b = external<BatchH>(env, argv[0], T_BATCH);
  if (!b || !b->h) return throw_error(env, "SplatBatch.status: not a batch");
  if (c.ok) c.result(rdx_batch_status(b->h));
  return nullptr;
}

napi_value BatchDestroy(napi_env env, napi_callback_info info) {
  ARGS(1);
  BatchH* b = external<BatchH>(env, argv[0], T_BATCH);
  if (b && b->h) {
    rdx_batch_destroy(b->h);
    b->h = nullptr;
  }
  return nullptr;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value s;
  napi_create_string_utf8(env, rdx_version(), NAPI_AUTO_LENGTH, &s);
  return s;
}

napi_value Init(napi_env env, napi_value exports) {
  const struct {
    const char* name;
    napi_callback fn;
  } fns[] = {
      {"version", Version},
      {"keyInvalid", KeyInvalid},
      {"keyValidMax", KeyValidMax},
      {"keyCompute", KeyCompute},
      {"sortKeys", SortKeys},
      {"sortKeysAsync", SortKeysAsync},
      {"streamCreate", StreamCreate},
      {"streamSynchronize", StreamSynchronize},
      {"streamSynchronizeAsync", StreamSynchronizeAsync},
      {"deviceAlloc", DeviceAlloc},
      {"deviceFree", DeviceFree},
      {"deviceUpload", DeviceUpload},
      {"deviceDownload", DeviceDownload},
      {"deviceFillU32", DeviceFillU32},
      {"sortPairs", SortPairs},
      {"sortPairsIndirect", SortPairsIndirect},
      {"sortSplats", SortSplats},
      {"projectEwa", ProjectEwa},
      {"projectSort", ProjectSort},
      {"projectKeys", ProjectKeys},
      {"projectRay", ProjectRay},
      {"untransposeBytes", UntransposeBytes},
      {"spzDecode", SpzDecode},
      {"rasterCreate", RasterCreate},
      {"rasterForward", RasterForward},
      {"rasterBackward", RasterBackward},
      {"rasterStats", RasterStats},
      {"rasterDestroy", RasterDestroy},
      {"drawCreate", DrawCreate},
      {"drawSplats", DrawSplats},
      {"drawRay", DrawRay},
      {"drawStats", DrawStats},
      {"drawDestroy", DrawDestroy},
      {"batchCreate", BatchCreate},
      {"batchInfo", BatchInfo},
      {"batchSetCameras", BatchSetCameras},
      {"batchRun", BatchRun},
      {"batchStatus", BatchStatus},
      {"batchDestroy", BatchDestroy},
  };
  for (const auto& f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)

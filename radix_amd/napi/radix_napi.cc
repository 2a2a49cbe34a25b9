// N-API binding of libradix_hip.so for Node: the thin layer TypeScript hosts
// call (N-API version 6, no node-addon-api). Each function converts JS
// values to the C ABI's plain pointers and sizes and maps a failure to a JS
// Error carrying the library's message (the reference's own text where it
// has one). Device memory is an External around a rdx_device_alloc pointer.
#define NAPI_VERSION 6
#include <node_api.h>

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/radix_hip.h"

namespace {

struct DeviceMem {
  void* ptr;
  size_t bytes;
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
  if (napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_double(env, v, &d);
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

DeviceMem* device(napi_env env, napi_value v) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t != napi_external) return nullptr;
  void* p = nullptr;
  napi_get_value_external(env, v, &p);
  return (DeviceMem*)p;
}

void* device_ptr(napi_env env, napi_value v) {
  napi_valuetype t;
  if (napi_typeof(env, v, &t) != napi_ok || t == napi_undefined || t == napi_null) return nullptr;
  DeviceMem* m = device(env, v);
  return m ? m->ptr : nullptr;
}

napi_value prop(napi_env env, napi_value obj, const char* name) {
  napi_value v = nullptr;
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) == napi_ok && has) napi_get_named_property(env, obj, name, &v);
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
  if (count < 0 || count > (double)nk) return throw_error(env, "SortKeys.sort: count above the keys given");
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
  SortWork* w = new SortWork();
  size_t nk;
  if (!u32_array(env, argv[0], &w->keys, &nk) || !u32_array(env, argv[3], &w->ordering, &w->ordering_len)) {
    delete w;
    return throw_error(env, "sortSplatKeys: keys and ordering must be Uint32Arrays");
  }
  w->count = (uint32_t)num(env, argv[1], (double)nk);
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

// ---- device memory ---------------------------------------------------------

void device_finalize(napi_env, void* data, void*) {
  DeviceMem* m = (DeviceMem*)data;
  if (m->ptr) rdx_device_free(m->ptr);
  delete m;
}

napi_value DeviceAlloc(napi_env env, napi_callback_info info) {
  ARGS(1);
  const double bytes = num(env, argv[0]);
  if (!(bytes >= 0)) return throw_error(env, "deviceAlloc: bytes must be >= 0");
  DeviceMem* m = new DeviceMem{nullptr, (size_t)bytes};
  if (!check(env, rdx_device_alloc(m->bytes, &m->ptr))) {
    delete m;
    return nullptr;
  }
  napi_value ext;
  napi_create_external(env, m, device_finalize, nullptr, &ext);
  return ext;
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

// deviceUpload(mem, typedArray, byteOffset)
napi_value DeviceUpload(napi_env env, napi_callback_info info) {
  ARGS(3);
  DeviceMem* m = device(env, argv[0]);
  void* src;
  size_t bytes;
  if (!m || !bytes_of(env, argv[1], &src, &bytes)) return throw_error(env, "deviceUpload: (buffer, typedArray, byteOffset)");
  const size_t off = (size_t)num(env, argv[2], 0);
  if (off + bytes > m->bytes) return throw_error(env, "deviceUpload: past the end of the device buffer");
  if (!check(env, rdx_device_upload(nullptr, (char*)m->ptr + off, src, bytes))) return nullptr;
  return nullptr;
}

// deviceDownload(mem, byteOffset, typedArray)
napi_value DeviceDownload(napi_env env, napi_callback_info info) {
  ARGS(3);
  DeviceMem* m = device(env, argv[0]);
  void* dst;
  size_t bytes;
  if (!m || !bytes_of(env, argv[2], &dst, &bytes)) return throw_error(env, "deviceDownload: (buffer, byteOffset, typedArray)");
  const size_t off = (size_t)num(env, argv[1], 0);
  if (off + bytes > m->bytes) return throw_error(env, "deviceDownload: past the end of the device buffer");
  if (!check(env, rdx_device_download(nullptr, dst, (const char*)m->ptr + off, bytes))) return nullptr;
  return nullptr;
}

// deviceFillU32(mem, value, count)
napi_value DeviceFillU32(napi_env env, napi_callback_info info) {
  ARGS(3);
  DeviceMem* m = device(env, argv[0]);
  if (!m) return throw_error(env, "deviceFillU32: not a device buffer");
  const size_t count = (size_t)num(env, argv[2]);
  if (count * 4 > m->bytes) return throw_error(env, "deviceFillU32: past the end of the device buffer");
  check(env, rdx_device_fill_u32(nullptr, (uint32_t*)m->ptr, (uint32_t)num(env, argv[1]), count));
  return nullptr;
}

// ---- GPU sorts ---------------------------------------------------------------

// sortPairs(keys, values, count, bits)
napi_value SortPairs(napi_env env, napi_callback_info info) {
  ARGS(4);
  check(env, rdx_sort_pairs(thread_context(), nullptr, (uint32_t*)device_ptr(env, argv[0]),
                            (uint32_t*)device_ptr(env, argv[1]), (uint32_t)num(env, argv[2]), (int)num(env, argv[3], 32)));
  return nullptr;
}

// sortPairsIndirect(keys, values, countBuffer, capacity, bits)
napi_value SortPairsIndirect(napi_env env, napi_callback_info info) {
  ARGS(5);
  check(env, rdx_sort_pairs_indirect(thread_context(), nullptr, (uint32_t*)device_ptr(env, argv[0]),
                                     (uint32_t*)device_ptr(env, argv[1]), (const uint32_t*)device_ptr(env, argv[2]),
                                     (uint32_t)num(env, argv[3]), (int)num(env, argv[4], 32)));
  return nullptr;
}

// sortSplats(keys, n, bits, indices, count, drawArgs)
napi_value SortSplats(napi_env env, napi_callback_info info) {
  ARGS(6);
  check(env, rdx_sort_splats(thread_context(), nullptr, (const uint32_t*)device_ptr(env, argv[0]),
                             (uint32_t)num(env, argv[1]), (int)num(env, argv[2], 24), (uint32_t*)device_ptr(env, argv[3]),
                             (uint32_t*)device_ptr(env, argv[4]), (uint32_t*)device_ptr(env, argv[5])));
  return nullptr;
}

// ---- projection ----------------------------------------------------------------

bool read_table(napi_env env, napi_value t, rdx_splat_table* out) {
  memset(out, 0, sizeof *out);
  out->layout = (int)prop_num(env, t, "layout", 0);
  out->count = (uint32_t)prop_num(env, t, "count", 0);
  out->plane_stride = (size_t)prop_num(env, t, "planeStride", 0);
  napi_value planes = prop(env, t, "planes");
  if (planes) {
    // one device buffer of f32 planes: centre 3, log-scale 3, quat 4, opacity, rgb 3, SH
    const size_t st = (out->plane_stride ? out->plane_stride : out->count) * 4;
    char* base = (char*)device_ptr(env, planes);
    out->center = (const float*)base;
    out->log_scale = (const float*)(base + 3 * st);
    out->quat = (const float*)(base + 6 * st);
    out->opacity = (const float*)(base + 10 * st);
    out->rgb = (const float*)(base + 11 * st);
    out->sh = (const float*)(base + 14 * st);
  }
  napi_value v;
  if ((v = prop(env, t, "extA"))) out->ext_a = (const uint32_t*)device_ptr(env, v);
  if ((v = prop(env, t, "extB"))) out->ext_b = (const uint32_t*)device_ptr(env, v);
  if ((v = prop(env, t, "packed"))) out->packed = (const uint32_t*)device_ptr(env, v);
  out->packed_encoding[0] = (float)prop_num(env, t, "rgbMin", 0.0);
  out->packed_encoding[1] = (float)prop_num(env, t, "rgbMax", 1.0);
  out->packed_encoding[2] = (float)prop_num(env, t, "lnScaleMin", -12.0);
  out->packed_encoding[3] = (float)prop_num(env, t, "lnScaleMax", 9.0);
  out->sh_codec = (int)prop_num(env, t, "shCodec", 0);
  out->sh_degree = (int)prop_num(env, t, "shDegree", 0);
  if ((v = prop(env, t, "shChannels"))) {
    for (uint32_t k = 0; k < 4; k++) {
      napi_value e;
      bool has = false;
      napi_has_element(env, v, k, &has);
      if (has && napi_get_element(env, v, k, &e) == napi_ok) out->sh_channels[k] = (const uint32_t*)device_ptr(env, e);
    }
  }
  if (!matrix16(env, prop(env, t, "objectToWorld"), out->object_to_world)) identity16(out->object_to_world);
  return true;
}

bool read_camera(napi_env env, napi_value c, rdx_camera* out) {
  memset(out, 0, sizeof *out);
  napi_value eyes = prop(env, c, "eyes");
  uint32_t ne = 0;
  if (eyes) napi_get_array_length(env, eyes, &ne);
  if (ne < 1 || ne > 2) {
    napi_throw_error(env, nullptr, "SplatCamera: 1 or 2 eyes");
    return false;
  }
  out->eyes = (int)ne;
  for (uint32_t e = 0; e < ne; e++) {
    napi_value eye;
    napi_get_element(env, eyes, e, &eye);
    if (!matrix16(env, prop(env, eye, "view"), out->eye[e].view) ||
        !matrix16(env, prop(env, eye, "projection"), out->eye[e].projection)) {
      napi_throw_error(env, nullptr, "camera eyes need view and projection (16 elements each)");
      return false;
    }
    out->eye[e].ortho = prop_bool(env, eye, "ortho", false) ? 1 : 0;
  }
  if (!matrix16(env, prop(env, c, "head"), out->head)) memcpy(out->head, out->eye[0].view, sizeof out->head);
  napi_value size = prop(env, c, "size");
  if (size) {
    napi_value w, h;
    napi_get_element(env, size, 0, &w);
    napi_get_element(env, size, 1, &h);
    out->size[0] = num(env, w, 1);
    out->size[1] = num(env, h, 1);
  }
  out->ndc_z_min = prop_num(env, c, "ndcZMin", 0.0);
  return true;
}

void read_params(napi_env env, napi_value p, rdx_params* out) {
  rdx_params_default(out);
  if (!p) return;
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

// projectEwa(table, camera, params, words[], keys, at)
napi_value ProjectEwa(napi_env env, napi_callback_info info) {
  ARGS(6);
  rdx_splat_table t;
  rdx_camera c;
  rdx_params p;
  read_table(env, argv[0], &t);
  if (!read_camera(env, argv[1], &c)) return nullptr;
  read_params(env, argv[2], &p);
  uint32_t* words[2] = {nullptr, nullptr};
  for (uint32_t e = 0; e < 2; e++) {
    napi_value w;
    bool has = false;
    napi_has_element(env, argv[3], e, &has);
    if (has && napi_get_element(env, argv[3], e, &w) == napi_ok) words[e] = (uint32_t*)device_ptr(env, w);
  }
  check(env, rdx_project_ewa(thread_context(), nullptr, &t, &c, &p, (uint32_t)num(env, argv[5], 0), words,
                             (uint32_t*)device_ptr(env, argv[4]), t.count, nullptr));
  return nullptr;
}

// projectSort(table, camera, params, words[], keys, indices, count, drawArgs)
napi_value ProjectSort(napi_env env, napi_callback_info info) {
  ARGS(8);
  rdx_splat_table t;
  rdx_camera c;
  rdx_params p;
  read_table(env, argv[0], &t);
  if (!read_camera(env, argv[1], &c)) return nullptr;
  read_params(env, argv[2], &p);
  uint32_t* words[2] = {nullptr, nullptr};
  for (uint32_t e = 0; e < 2; e++) {
    napi_value w;
    bool has = false;
    napi_has_element(env, argv[3], e, &has);
    if (has && napi_get_element(env, argv[3], e, &w) == napi_ok) words[e] = (uint32_t*)device_ptr(env, w);
  }
  check(env, rdx_project_sort(thread_context(), nullptr, &t, &c, &p, words, (uint32_t*)device_ptr(env, argv[4]),
                              (uint32_t*)device_ptr(env, argv[5]), (uint32_t*)device_ptr(env, argv[6]),
                              (uint32_t*)device_ptr(env, argv[7]), nullptr));
  return nullptr;
}

// projectKeys(table, camera, params, keys, at): the keys-only pass
napi_value ProjectKeys(napi_env env, napi_callback_info info) {
  ARGS(5);
  rdx_splat_table t;
  rdx_camera c;
  rdx_params p;
  read_table(env, argv[0], &t);
  if (!read_camera(env, argv[1], &c)) return nullptr;
  read_params(env, argv[2], &p);
  check(env, rdx_project_keys(thread_context(), nullptr, &t, &c, &p, (uint32_t)num(env, argv[4], 0),
                              (uint32_t*)device_ptr(env, argv[3])));
  return nullptr;
}

// projectRay(table, camera, params, method, channels[], keys, at): 3dgut (1), htgs (2), raygs (3)
napi_value ProjectRay(napi_env env, napi_callback_info info) {
  ARGS(7);
  rdx_splat_table t;
  rdx_camera c;
  rdx_params p;
  read_table(env, argv[0], &t);
  if (!read_camera(env, argv[1], &c)) return nullptr;
  read_params(env, argv[2], &p);
  uint32_t* channels[3] = {nullptr, nullptr, nullptr};
  for (uint32_t k = 0; k < 3; k++) {
    napi_value w;
    bool has = false;
    napi_has_element(env, argv[4], k, &has);
    if (has && napi_get_element(env, argv[4], k, &w) == napi_ok) channels[k] = (uint32_t*)device_ptr(env, w);
  }
  check(env, rdx_project_ray(thread_context(), nullptr, &t, &c, &p, (int)num(env, argv[3], 0), (uint32_t)num(env, argv[6], 0),
                             channels, (uint32_t*)device_ptr(env, argv[5])));
  return nullptr;
}

// untransposeBytes(planes, count, elementBytes, out, outOffsetBytes): an RDX payload's planes back to elements
napi_value UntransposeBytes(napi_env env, napi_callback_info info) {
  ARGS(5);
  char* out = (char*)device_ptr(env, argv[3]);
  check(env, rdx_untranspose_bytes(nullptr, (const uint8_t*)device_ptr(env, argv[0]), (uint32_t)num(env, argv[1], 0),
                                   (uint32_t)num(env, argv[2], 0), out ? out + (size_t)num(env, argv[4], 0) : nullptr));
  return nullptr;
}

// spzDecode(body, bodyBytes, count, version, shDegree, fractionalBits, planes, planeStride):
// an SPZ body (the gunzipped file after its 16-byte header) to f32 planes
napi_value SpzDecode(napi_env env, napi_callback_info info) {
  ARGS(8);
  check(env, rdx_spz_decode(nullptr, (const uint8_t*)device_ptr(env, argv[0]), (size_t)num(env, argv[1], 0),
                            (uint32_t)num(env, argv[2], 0), (int)num(env, argv[3], 0), (int)num(env, argv[4], 0),
                            (int)num(env, argv[5], 0), (float*)device_ptr(env, argv[6]), (size_t)num(env, argv[7], 0)));
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

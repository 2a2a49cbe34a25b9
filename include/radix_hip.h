/*
 * radix_hip.h — C ABI of the MI355X (gfx950) splat pipeline: sort keys,
 * stable LSD radix sort of splat indices, fused EWA projection with SH
 * colour, the batched (many views, many collections) form and the
 * differentiable compositor with its adjoint.
 *
 * Plain pointers and sizes only. Device pointers are HIP device memory;
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).
 * Every entry point returns 0 or a negative RDX_E* code; the message the
 * reference would throw (same text) is then in rdx_last_error(), per thread.
 *
 * Matrices are 16 doubles in column-major order, the layout of three.js
 * Matrix4.elements, so a JavaScript binding passes `matrix.elements` as is.
 *
 * Citations name files of the reference (sparkjsdev/radix) whose interface
 * an entry point replaces.
 */
#ifndef RADIX_HIP_H
#define RADIX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Error codes -------------------------------------------------------- */
#define RDX_OK 0
#define RDX_E_INVALID (-1)   /* bad argument (bits, counts, layout)        */
#define RDX_E_CAPACITY (-2)  /* an output or the context is too small      */
#define RDX_E_HIP (-3)       /* a HIP runtime call failed                  */
#define RDX_E_NOMEM (-4)     /* device allocation failed                   */
#define RDX_E_UNSUPPORTED (-5)

/* The message of the calling thread's last failed call ("" if none). */
const char* rdx_last_error(void);
/* Library version string and the gfx target it was built for. */
const char* rdx_version(void);

/* ---- Sort keys (src/core/sort.ts:37-58, src/gpu/sortKey.tsl.ts:17-26) ---- */

/* SortKeys.invalid(bits): all ones (src/core/sort.ts:37-39). */
uint32_t rdx_key_invalid(int bits);
/* SortKeys.validMax(bits) (src/core/sort.ts:41-43). */
uint32_t rdx_key_valid_max(int bits);
/* SortKeys.compute(distance, bits) on the host, for one float32 distance
 * (src/core/sort.ts:46-58). */
uint32_t rdx_key_compute(float distance, int bits);
/* SortKeys.compute over `n` device float32 distances into device keys
 * (GpuSortKeys.compute, src/gpu/sortKey.tsl.ts:17-26, with the CPU rule
 * that non-finite and negative bit patterns give the invalid key). */
int rdx_compute_keys(void* stream, const float* distances_dev, uint32_t n, int bits, uint32_t* keys_dev);

/* ---- Context: device scratch, reused across calls on one stream ---------- */

/* A context owns the device scratch of the sorts and projections run
 * through it (no module-level scratch: src/core/sort.ts:22-25 is re-entrant
 * only per thread; a context is re-entrant per stream). Calls on one
 * context must be ordered on one stream; use one context per stream. */
typedef struct rdx_context rdx_context;
int rdx_context_create(rdx_context** out);
int rdx_context_destroy(rdx_context* ctx);
/* Grow scratch now for sorts of up to `max_keys` keys (optional; calls grow it). */
int rdx_context_reserve(rdx_context* ctx, uint32_t max_keys);

/* Radix sort tuning: digit width 4, 6 or 8 bits, keys per thread and
 * workgroup size (threads). Zeros pick the built-in default (8, 16, 512). */
typedef struct rdx_sort_tuning {
  int digit_bits;
  int keys_per_thread;
  int block_threads;
} rdx_sort_tuning;
int rdx_context_set_tuning(rdx_context* ctx, const rdx_sort_tuning* tuning);
int rdx_context_get_tuning(const rdx_context* ctx, rdx_sort_tuning* out);
/* Autotuner (no reference counterpart: the reference fixes 8-bit digits,
 * WgpuRadixSort.ts:13-15, and a keysPerThread per tier): times
 * rdx_sort_splats of `n` keys under every supported tuning (the first
 * `count` of them, max 64, in the order of rdx_sort_tuning_candidate) and
 * keeps the fastest on the context. keys_dev: the keys to sort (NULL:
 * uniform random keys of `bits` bits made on the device). Writes the chosen
 * tuning and its time per sort in milliseconds, and stores them in the
 * tuning cache (form "sort", one segment). Synchronous. */
int rdx_sort_tuning_candidate(int index, rdx_sort_tuning* out);  /* RDX_E_INVALID past the last */
int rdx_context_autotune(rdx_context* ctx, void* stream, const uint32_t* keys_dev, uint32_t n, int bits, int reps,
                         rdx_sort_tuning* out, double* ms_out);
/* The tuning cache: the autotuners' results per (gfx arch, form "sort" or
 * "batch", key bits, ceil(log2) of the keys, ceil(log2) of the segments),
 * kept for the process and, with a cache file (this call, or the
 * environment's RDX_TUNING_CACHE), read from it and appended to it, one
 * line per result: arch form bits n_bucket segment_bucket digit_bits
 * keys_per_thread block_threads ms. NULL or "" keeps the cache in memory. */
int rdx_tuning_cache_file(const char* path);
/* 1 and the cached tuning (and its time) when the cache holds one, 0 when not. */
int rdx_tuning_cache_lookup(const char* form, uint64_t n, uint32_t segments, int bits, rdx_sort_tuning* out, double* ms);

/* Stage timing: with profiling on, each kernel a call on this context
 * launches is bracketed by HIP events on that call's stream. No reference
 * counterpart: it replaces GpuProfiler's timestamp phases
 * (src/gpu/GpuProfiler.ts:28-144) for measurement. */
int rdx_context_profile(rdx_context* ctx, int enable);
/* Totals per stage name since the last read (waits for the recorded
 * events): up to max_stages entries of names (32 chars each, NUL-padded),
 * milliseconds and launch counts; returns the number of stages, or < 0. */
int rdx_context_stage_times(rdx_context* ctx, char* names, double* ms, int* launches, int max_stages);

/* ---- Device memory for bindings (Node N-API) without a GPU runtime -------
 * The storage a WgpuBuffer gives the reference (src/wgpu/WgpuBuffer.ts:15-201:
 * allocation, upload, read back); plain HIP device memory here. */
int rdx_device_alloc(size_t bytes, void** out);
int rdx_device_free(void* ptr);
/* Copies wait for completion (the host buffer may be reused on return). */
int rdx_device_upload(void* stream, void* dst_dev, const void* src_host, size_t bytes);
int rdx_device_download(void* stream, void* dst_host, const void* src_dev, size_t bytes);
int rdx_device_fill_u32(void* stream, uint32_t* dst_dev, uint32_t value, size_t count);
int rdx_stream_synchronize(void* stream);
/* A HIP stream of its own for a binding's in-flight call (default priority,
 * not synchronizing with the null stream): the queue a WebGPU device gives
 * each submission (src/wgpu/WgpuCompute.ts:64-73), one per concurrent call. */
int rdx_stream_create(void** out);
int rdx_stream_destroy(void* stream);

/* ---- The CPU-sort API, backed by the GPU --------------------------------- */

/* SortKeys.sort({keys, count, bits, ordering}) (src/core/sort.ts:64-82) on
 * HOST arrays: stable ascending order of the low `bits` bits of `keys`
 * (keys must be below 2^bits at 24 bits, as the reference requires) into
 * `ordering[0..count)`; *visible = count minus the keys whose top part is
 * all ones, exactly as the reference counts it (16 bits: key & 0xffff ==
 * 0xffff; 24: key >> 11 == 0x1fff; 32: key >> 16 == 0xffff).
 * ordering_len < count fails with the reference's message
 * "SortKeys.sort: ordering holds <len> of <count> indices" (sort.ts:70-72).
 * Synchronous: returns when `ordering` holds the result; the copies and
 * kernels run on a stream the context owns, so sorts through contexts on
 * different threads overlap on the device. Also the body of
 * sortSplatKeys (src/worker/splatSort.ts:17-30) and the worker's sortKeys
 * handler (src/worker/coreWorker.ts:63-65). */
int rdx_sort_keys(rdx_context* ctx, const uint32_t* keys, uint32_t count, int bits,
                  uint32_t* ordering, size_t ordering_len, uint32_t* visible);

/* ---- The GPU-sort API (device buffers, stream-ordered) ------------------- */

/* WgpuRadixSort.sort(count, {bits}) (src/wgpu/WgpuRadixSort.ts:118-128):
 * sorts the first `count` device keys and values in place, ascending and
 * stable, on the low `bits` bits (16, 24 or 32; others fail with
 * "WgpuRadixSort: bits must be 16, 24 or 32 (got <bits>)", WgpuRadixSort.ts:150-152). */
int rdx_sort_pairs(rdx_context* ctx, void* stream, uint32_t* keys_dev, uint32_t* values_dev,
                   uint32_t count, int bits);

/* WgpuRadixSort.sortIndirect(countBuffer, {bits}) (WgpuRadixSort.ts:132-141):
 * as rdx_sort_pairs on the first min(*count_dev, capacity) pairs, the count
 * read on the device (nothing read back). */
int rdx_sort_pairs_indirect(rdx_context* ctx, void* stream, uint32_t* keys_dev, uint32_t* values_dev,
                            const uint32_t* count_dev, uint32_t capacity, int bits);

/* WgpuSplatSorter.sort(projected, count, bits) (src/wgpu/WgpuSplatSorter.ts:64-71):
 * drops keys equal to SortKeys.invalid(bits) (WgpuCompact, keeping input
 * order), sorts the rest back to front, and writes
 *   out_indices_dev[0..visible)  the splat indices in draw order,
 *   *out_count_dev               visible (WgpuSplatSorter.count),
 *   out_draw_args_dev[0..5)      [6 * 1024, ceil(visible / 1024), 0, 0, 0]
 * (WgpuSplatSorter.ts:48-58). Any of count/args may be NULL. With 32-bit
 * keys and an 8-bit, 512-thread tuning (the default) the kept keys are
 * sorted by 9-bit digits above their smallest value: three passes when they
 * span less than 2^27 (distances within a factor 2^16), four otherwise; the
 * order is the same stable order either way. */
int rdx_sort_splats(rdx_context* ctx, void* stream, const uint32_t* keys_dev, uint32_t n, int bits,
                    uint32_t* out_indices_dev, uint32_t* out_count_dev, uint32_t* out_draw_args_dev);

/* The splat sorter without allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It does
 * what rdx_sort_splats does under the default tuning, with its scratch in a
 * static device array of about 17 MB per GPU, held until the process ends,
 * instead of a context's buffers: at most RDX_STATIC_SORT_MAX_KEYS keys
 * (more fail with RDX_E_CAPACITY). Nothing is allocated or freed. Calls on
 * one GPU share its array: calls that overlap on one GPU (two streams, two
 * threads) race and must be ordered by the caller. */
#define RDX_STATIC_SORT_MAX_KEYS 1048576u
int rdx_sort_splats_static(void* stream, const uint32_t* keys_dev, uint32_t n, int bits, uint32_t* out_indices_dev,
                           uint32_t* out_count_dev, uint32_t* out_draw_args_dev);

/* ---- Projection (ewa) ---------------------------------------------------- */

#define RDX_LAYOUT_F32_SOA 0  /* float planes: center, log-scale, quat, opacity, rgb, SH */
#define RDX_LAYOUT_EXT 1      /* extSplats: extA, extB uint32x4 (src/core/splats.ts:10-38) */
#define RDX_LAYOUT_PACKED 2   /* packedSplats: packed uint32x4 (src/core/splats.ts:271-309) */
#define RDX_LAYOUT_COV 3      /* covSplats: covA, covB uint32x4 in ext_a, ext_b (src/core/splats.ts:25-27,
                                 219-266); the ewa method reads the covariance as stored */

#define RDX_SH_NONE 0
#define RDX_SH_F32 1     /* float planes 3k+c (k = coefficient past DC, c = channel) */
#define RDX_SH_EXT 2     /* extSh channels A.. (src/core/sh.ts:114-135)          */
#define RDX_SH_PACKED 3  /* packedSh channels A.. (src/core/sh.ts:137-174)       */

/* One splat table (a collection object's data) in device memory. */
typedef struct rdx_splat_table {
  int layout;            /* RDX_LAYOUT_* */
  uint32_t count;        /* splats */
  /* F32_SOA: planes of `plane_stride` floats each */
  const float* center;     /* 3 planes x, y, z */
  const float* log_scale;  /* 3 planes, natural log of the scales */
  const float* quat;       /* 4 planes x, y, z, w (any length; normalised here) */
  const float* opacity;    /* 1 plane, 0..1 (LoD values above 1 allowed) */
  const float* rgb;        /* 3 planes: the base colour (SH DC) */
  size_t plane_stride;     /* elements per plane (>= count); 0 = count; at most 71,582,788
                              (planes are read in groups of at most 15 within 4 GiB) */
  /* EXT: two uint32x4 arrays; PACKED: one, with its encoding */
  const uint32_t* ext_a;
  const uint32_t* ext_b;
  const uint32_t* packed;
  float packed_encoding[4];  /* rgbMin, rgbMax, lnScaleMin, lnScaleMax (splats.ts:90-97) */
  /* SH */
  int sh_codec;              /* RDX_SH_* */
  int sh_degree;             /* 0..3 */
  const float* sh;           /* F32: 3*K planes of plane_stride floats */
  const uint32_t* sh_channels[4]; /* EXT / PACKED: uint32x4 arrays, one per texel */
  /* Object to world (matrixWorld), column-major */
  double object_to_world[16];
} rdx_splat_table;

/* One eye: world to eye (camera.matrixWorldInverse) and its projection. */
typedef struct rdx_eye {
  double view[16];
  double projection[16];
  int ortho;
} rdx_eye;

/* SplatCamera.setViews({head, eyes, size, ndcZMin}) (src/render/splatCamera.ts:119-138). */
typedef struct rdx_camera {
  double head[16];   /* world to head */
  int eyes;          /* 1 or 2 */
  rdx_eye eye[2];
  double size[2];    /* render size, pixels (width, height) */
  double ndc_z_min;  /* 0: WebGPU clip convention, -1: WebGL */
} rdx_camera;

/* SplatViewParams + SplatObjectParams (src/render/splatParams.ts:8-87) and
 * the key uniforms (SplatProjection.setKeys, SplatProjection.ts:24-26, 157-160). */
typedef struct rdx_params {
  double max_std_dev;      /* sqrt(8) */
  double min_alpha;        /* 0.5 / 255 */
  double falloff;          /* 1 */
  int lod_inflate;         /* 1 */
  double blur_amount;      /* 0.3 */
  double pre_blur_amount;  /* 0 */
  double clip_xy;          /* 1.4 */
  double min_pixel_radius; /* 0 */
  double max_pixel_radius; /* 512 */
  double min_contribution; /* 0 */
  double focal_distance;   /* 0 */
  double aperture_tan;     /* 0 */
  double sort_guard;       /* 0 */
  int max_sh_degree;       /* 4 */
  int sort_bits;           /* 16, 24 or 32 */
  int cull_keys;           /* 1: keys of splats culled in every eye are invalid */
  int key_per_eye;         /* 0: one key from the head (reference); 1: per eye */
} rdx_params;

/* The reference defaults (splatParams.ts:68-72), sort_bits 24, cull_keys 1. */
void rdx_params_default(rdx_params* out);

/* Float side output of the projection for checks (NULL members skipped):
 * planes of `stride` floats; per-eye planes at [eye * planes * stride]. */
typedef struct rdx_float_out {
  size_t stride;
  float* conic;     /* per eye 3 planes: inverse 2D covariance (A, B, C), pixels^-2 */
  float* cov2d;     /* per eye 3 planes: blurred 2D covariance (a, b, d), pixels^2 */
  float* opacity;   /* per eye 1 plane: blur-compensated opacity */
  float* ndc;       /* per eye 2 planes */
  float* view_z;    /* per eye 1 plane: eye-space z (negative ahead) */
  float* bound;     /* per eye 3 planes: scale1, scale2 (pixels), angle */
  float* rgb;       /* 3 planes: base colour plus SH, clamped at 0 */
  float* distance;  /* 1 plane: the key distance (per eye with key_per_eye) */
} rdx_float_out;

/* SplatProjection.write for the ewa method (SplatProjection.ts:141-153,
 * projections.tsl.ts:95-148), run as one kernel over the table:
 *   words_dev[e]  eye e's 4 words per splat at [4 * (at + i) ...] (the
 *                 projected bundle's channel w<e>, SplatProjection.ts:48-54)
 *   keys_dev      the sort key per splat at [at + i] (sortKeys.key, little-endian)
 *                 (with key_per_eye, eye e's keys at keys_dev + e * key_stride)
 *   side          optional float outputs (NULL: none). */
int rdx_project_ewa(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                    const rdx_params* params, uint32_t at, uint32_t* const* words_dev, uint32_t* keys_dev,
                    size_t key_stride, const rdx_float_out* side);

/* The keys-only pass for CPU sorts (SplatProjection.writeKeys,
 * src/render/SplatProjection.ts:165-172, run by projectCollection's "keys"
 * pass, SplatProjector.ts:184-189, for SplatView.sort, SplatView.ts:137-185):
 * keys[at + i] = compute(|head-space centre|, sort_bits), or invalid for
 * splats cullInvisible drops or the sortGuard excludes (params->sort_guard;
 * cull_keys does not apply). Reads the centre, scales, rotation and opacity
 * only (no SH, no EWA); keys are the ones rdx_project_ewa writes with
 * cull_keys 0. The host then reads them back for rdx_sort_keys. */
int rdx_project_keys(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                     const rdx_params* params, uint32_t at, uint32_t* keys_dev);

/* The ray-evaluated projection methods' words (SplatProjection.write for
 * "3dgut", "htgs" and "raygs", src/render/SplatProjection.ts:141-151 with
 * projections.tsl.ts:170-429). 3dgut and htgs: 6 head words, then 2 per eye,
 *   head [centre direction (octahedral 16 + 16), centre distance (float
 *         bits), rgb, rotation QuatOct32, half2 ln(scales.xy),
 *         half ln(scales.z) | opacity8 << 16 | direction's low 4 + 4 bits << 24]
 *   eye  [bound ndc 2x16, angle10 | log2 scale (11 bits, outward) << 10 | << 21]
 * raygs: 4 head words, then 3 per eye,
 *   head [centre direction, centre distance, rgb, opacity8 | low 4 + 4 bits << 24]
 *   eye  [quad frame QuatOct32, c^2 (float bits), half2 ln(axis lengths)]
 * in the projected bundle's channels (SplatProjection.ts:46-53):
 * channels_dev[0], [1] uint32x4 and, with two eyes, channels_dev[2]
 * uint32x2, each at element at + i. A culled eye's words are 0; a splat
 * culled in every eye keeps its head words (zero when invisible). keys_dev
 * (NULL: none) gets the head-distance key at [at + i] (cullKeys: visible in
 * an eye; else !invisible and the sortGuard). The object transform must be
 * a rotation times a uniform scale (the reference's fast path,
 * splatTransform.tsl.ts:142-152; others fail with RDX_E_UNSUPPORTED).
 * method: RDX_METHOD_3DGUT, RDX_METHOD_HTGS or RDX_METHOD_RAYGS. */
#define RDX_METHOD_EWA 0
#define RDX_METHOD_3DGUT 1
#define RDX_METHOD_HTGS 2
#define RDX_METHOD_RAYGS 3
int rdx_project_ray(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                    const rdx_params* params, int method, uint32_t at, uint32_t* const* channels_dev,
                    uint32_t* keys_dev);

/* Project and sort in one stream-ordered sequence: the projection kernel
 * writes words and keys; the first radix pass drops invalid keys. Output
 * as rdx_sort_splats for the keys it made (keys_dev NULL: context scratch). The headline path
 * (SplatRenderer.update's project + sort, SplatRenderer.ts:412-456).
 * With params->key_per_eye on a stereo camera each eye sorts its own keys
 * (two segments of n = table->count): eye e's keys at keys_dev + e n, its
 * indices at out_indices_dev + e n, its count at out_count_dev[e] and its
 * draw arguments at out_draw_args_dev + 5 e; the buffers hold 2 n, 2 n, 2
 * and 10 words. */
int rdx_project_sort(rdx_context* ctx, void* stream, const rdx_splat_table* table, const rdx_camera* camera,
                     const rdx_params* params, uint32_t* const* words_dev, uint32_t* keys_dev,
                     uint32_t* out_indices_dev, uint32_t* out_count_dev, uint32_t* out_draw_args_dev,
                     const rdx_float_out* side);

/* ---- File decode onto the device ------------------------------------------
 * RDX chunks store each channel's elements byte-transposed, then gzipped
 * (src/core/rdx.ts:8-17, Compression.compressTransposed, compress.ts:192-195).
 * After the host gunzips a payload, this undoes the transpose on the device
 * (the second step of Compression.decompressTransposed, compress.ts:197-204,
 * transposeBytes 6-19): planes_dev holds element_bytes planes of `count`
 * bytes (plane k = byte k of every element); out_dev gets `count` elements
 * of element_bytes bytes (a channel buffer at the chunk's start element).
 * Stream-ordered; element_bytes 1 .. 64. */
int rdx_untranspose_bytes(void* stream, const uint8_t* planes_dev, uint32_t count, uint32_t element_bytes, void* out_dev);

/* SPZ (Niantic's format in its gzip container, src/core/spz.ts:140-175,
 * 260-349) decoded on the device into f32 planes: body_dev is the unzipped
 * file after its 16-byte header (columns of centres, alphas, colours,
 * scales, rotations, SH); planes_dev gets 14 + 3K planes of plane_stride
 * floats (centre 3, log-scale 3, quaternion 4, opacity, rgb 3, SH 3K:
 * rdx_splat_table's f32 layout) with the values readSpz computes: centres
 * int24 x 2^-fractional_bits (version 1: half floats), opacity a / 255,
 * colour (b / 255 - 0.5) SH_C0 / 0.15 + 0.5, log-scale b / 16 - 10,
 * rotations smallest-three (versions 3 and 4) or xyz with w >= 0 (1, 2),
 * SH (b - 128) / 128 coefficient-major RGB. Versions 1-4 and SH degrees 0-3
 * ("Unsupported SPZ version v or SH degree d" otherwise; degree 4 is
 * RDX_E_UNSUPPORTED); no coordinate-system conversion. */
int rdx_spz_decode(void* stream, const uint8_t* body_dev, size_t body_bytes, uint32_t count, int version, int sh_degree,
                   int fractional_bits, float* planes_dev, size_t plane_stride);

/* ---- Batched views: many (view, collection) segments in one sequence -----
 * No reference counterpart: the reference re-projects and re-sorts per
 * render and per object (SplatProjector.ts:127-137); here every view of a
 * batch is a segment of one keys array, each table is read once per group
 * of up to 8 views drawing it, one projection launch covers every table of
 * a kind and each radix pass sorts every segment at once. Segments follow
 * the offsets of SplatCollection (SplatCollection.ts:54-65): view v's keys
 * and indices are [off_v, off_v + n_v) with n_v its table's count. */
typedef struct rdx_batch rdx_batch;
#define RDX_BATCH_MAX_VIEWS 4096
typedef struct rdx_batch_view {
  uint32_t* words_dev;  /* the view's ewa words, 4 per splat of its table (NULL: not written) */
  uint32_t table;       /* index into the batch's tables */
  uint32_t reserved;
  rdx_camera camera;    /* one eye */
} rdx_batch_view;
/* A batch over `tables` and `views` with one rdx_params (sortGuard is not
 * applied: keys follow cullKeys) and the context's sort tuning at creation.
 * keys_dev: rdx_batch_total() words for the keys (NULL: the batch owns
 * them). The tables' device memory must outlive the batch. */
int rdx_batch_create(rdx_context* ctx, const rdx_splat_table* tables, int num_tables, const rdx_batch_view* views,
                     int num_views, const rdx_params* params, uint32_t* keys_dev, rdx_batch** out);
int rdx_batch_destroy(rdx_batch* batch);
/* New cameras for every view (one each, in view order); waits for the
 * batch's last run to finish before the device copy changes. */
int rdx_batch_set_cameras(rdx_batch* batch, const rdx_camera* cameras, int num_views);
/* Sum of the views' splat counts: the length of the keys and indices. */
uint64_t rdx_batch_total(const rdx_batch* batch);
/* View v's first key and index (its segment offset). */
uint64_t rdx_batch_segment_offset(const rdx_batch* batch, int view);
/* The device keys of the last run (view v's at rdx_batch_segment_offset(v)). */
const uint32_t* rdx_batch_keys(const rdx_batch* batch);
/* Project and sort every view: view v's visible indices back to front at
 * out_indices_dev + off_v, its count at out_counts_dev[v] and its draw
 * arguments at out_draw_args_dev + 5 v (either may be NULL). flags
 * RDX_BATCH_GRAPH: capture the sequence once into a HIP graph (per output
 * set and form) and replay it. RDX_BATCH_PERSISTENT: the whole batch in one
 * launch of a persistent kernel (the projection and every radix pass of
 * every segment, with grid-wide barriers between the phases), available
 * when the batch's tables are of one kind (layout, SH codec and degree)
 * and it was created under the 8-bit, 12-key, 512-thread tuning with
 * 32-bit keys (RDX_E_UNSUPPORTED otherwise); the same output as the launch
 * sequence.
 * Runs of one batch must not overlap on the device. */
#define RDX_BATCH_GRAPH 1
#define RDX_BATCH_PERSISTENT 2
int rdx_batch_run(rdx_batch* batch, void* stream, uint32_t* out_indices_dev, uint32_t* out_counts_dev,
                  uint32_t* out_draw_args_dev, int flags);
/* The persistent form's grid (blocks resident at once, a multiple of 8),
 * or 0 when the batch has no persistent form. */
int rdx_batch_persistent_grid(const rdx_batch* batch);
/* Phase times of the batch's last persistent run made while its context
 * profiled (rdx_context_profile), from per-block s_memrealtime stamps:
 * phase_ms[k] from the first block starting phase k to the last finishing
 * it, release_ms[k] from then until the last block left the barrier after
 * it. Phases: the projection, count / scan / scatter of passes 0-2, the
 * gated scatter (11). Returns the count written (0: no such run). */
int rdx_batch_persistent_phases(rdx_batch* batch, double* phase_ms, double* release_ms, int max_phases);
/* Waits for the batch's last run; RDX_E_HIP (and clears the flag) if a
 * persistent run's barrier wait ran past its limit of about 1.3 s (a grid
 * whose blocks were not all resident at once: that run's output is not
 * valid). */
int rdx_batch_status(rdx_batch* batch);
/* Times `reps` graph replays of the batch in each form it has (the launch
 * sequence, and the persistent launch when available) on `stream` and
 * reports the faster in *persistent_out (1: the persistent launch) with
 * both times (ms_persistent -1 without that form). Synchronous. */
int rdx_batch_choose_form(rdx_batch* batch, void* stream, int reps, int* persistent_out, double* ms_graph,
                          double* ms_persistent);
/* The autotuner of a batch (no reference counterpart: the reference fixes
 * 8-bit digits and its keys per thread, WgpuRadixSort.ts:13-15, 80-95):
 * for every supported tuning (rdx_sort_tuning_candidate) it creates the
 * batch these arguments describe on the context, times `reps` graph replays
 * of its projection and segmented sort on `stream`, and keeps the fastest
 * tuning on the context, so a batch created next uses it. The result goes
 * into the tuning cache under (gfx arch, "batch", key bits, ceil(log2) of
 * the total keys, ceil(log2) of the views); with RDX_TUNE_CACHED a cached
 * tuning is taken without timing (*cached_out = 1). Synchronous. */
#define RDX_TUNE_CACHED 1
int rdx_batch_autotune(rdx_context* ctx, void* stream, const rdx_splat_table* tables, int num_tables,
                       const rdx_batch_view* views, int num_views, const rdx_params* params, int reps, int flags,
                       rdx_sort_tuning* out, double* ms_out, int* cached_out);

/* ---- Differentiable compositor ------------------------------------------------
 * No reference counterpart: the reference draws the sorted splats with
 * rasterised quads and has no backward (splatDraw.ts). This composites the
 * ewa splats front to back over 16 x 16 pixel tiles with the draw's fragment
 * (alpha = opacity exp(-z^2 / 2), discarded beyond the cutoff and below
 * minAlpha: projection.tsl.ts:444-455, splatDraw.ts:136-174) and
 * premultiplied "over" (splatDraw.ts:62-78), stopping before the splat that
 * would take transmittance below t_stop, and differentiates it.
 * Parameters: 59 planes of n floats (params_dev[k * n + i]): centre 3,
 * log-scale 3, quaternion 4 (x, y, z, w), opacity LOGIT 1, base colour 3,
 * SH bands 1-3 45 (coefficient-major). Images: 3 planes of width x height
 * floats, row 0 at the bottom (pixel centres (x + 0.5, y + 0.5), y up);
 * alpha = 1 - final transmittance. */
typedef struct rdx_raster rdx_raster;
typedef struct rdx_raster_settings {
  int sh_degree;                  /* SH bands used, 0..3 */
  float t_stop;                   /* transmittance stop (0: 1e-4) */
  uint32_t max_entries_per_splat; /* limit on (splat, tile) entries per splat (0: none); the entry
                                     buffers are sized to each forward's count */
} rdx_raster_settings;
int rdx_raster_create(rdx_context* ctx, const rdx_raster_settings* settings, rdx_raster** out);
int rdx_raster_destroy(rdx_raster* raster);
/* The forward on a mono camera (eye 0; renderSize = camera->size) with the
 * params' maxStdDev, minAlpha, blurAmount and clipXY. Keeps what the
 * backward needs; params_dev must stay unchanged until then. Synchronizes
 * the stream once, after binning, to read the frame's (splat, tile) entry
 * count and size the entry buffers to it; a count above the settings' limit
 * fails with RDX_E_CAPACITY (nothing is drawn). */
int rdx_raster_forward(rdx_raster* raster, void* stream, const float* params_dev, uint32_t n, const rdx_camera* camera,
                       const rdx_params* params, float* image_dev, float* alpha_dev);
/* The adjoint of the last forward: dL/dimage (3 planes) -> dL/dparams (59
 * planes of n, overwritten). Transmittance is recomputed back to front over
 * each pixel's contributors, by division, from the value the forward kept
 * at the end of each 256-entry chunk of its tile's list (one per chunk, not
 * per contributor); a forward that failed leaves nothing to differentiate
 * (RDX_E_INVALID). */
int rdx_raster_backward(rdx_raster* raster, void* stream, const float* grad_image_dev, float* grad_params_dev);
/* The last forward's drawn splats, (splat, tile) entries and entry capacity;
 * overflowed is 0 (a count past the limit fails the forward). Synchronizes. */
int rdx_raster_stats(rdx_raster* raster, void* stream, uint32_t* visible, uint32_t* entries, uint32_t* capacity,
                     int* overflowed);

/* ---- The draw (production form) -------------------------------------------- */

/* The splat draw of the ewa method (src/render/splatDraw.ts:98-181 with
 * projections.tsl.ts:150-167 and fragmentAlpha, projection.tsl.ts:444-455):
 * for each splat of the sorted order, back to front, the fragments of its
 * quad, from the eye's ewa words (ndc centre, angle, log scales, extRgb,
 * opacity byte; zero byte = culled): z^2 = |uv|^2, discarded beyond the
 * cutoff and below minAlpha, alpha = fragmentAlpha(z^2, opacity, falloff)
 * (opacities above 1 are LoD stretches), blended premultiplied "over"
 * (src = color * alpha, dst = src + (1 - alpha) dst). Composited per 16x16
 * tile front to back (equal to the blend up to rounding). No counterpart
 * for the rasteriser itself: it replaces the reference's drawIndexed of
 * quads (WebGPU) for native hosts. */
typedef struct rdx_draw rdx_draw;
typedef struct rdx_draw_settings {
  float t_stop;                   /* 0: every fragment (the reference's blend); t > 0: stop a pixel at transmittance t */
  int linear_output;              /* 1: colours sRGB-decoded first (an output storing linear colour, splatDraw.ts:176) */
  uint32_t max_entries_per_splat; /* limit on (splat, tile) entries per draw position (0: none); the
                                     entry buffers are sized to each draw's count */
} rdx_draw_settings;
int rdx_draw_create(rdx_context* ctx, const rdx_draw_settings* settings, rdx_draw** out);
int rdx_draw_destroy(rdx_draw* draw);
/* words_dev: the eye's ewa words of n splats (uint32x4 each, as
 * rdx_project_ewa writes them); order_dev: `capacity` splat indices, back
 * to front (rdx_sort_splats' indices), of which *count_dev (device, clamped
 * to capacity) are drawn; indices >= n are skipped (splatDraw.ts:121-122).
 * params: max_std_dev, min_alpha, falloff. rgba_dev: 4 planes of width x
 * height floats, premultiplied r, g, b and alpha over a transparent target,
 * row 0 at the top (framebuffer rows, pixel centres at + 0.5). Synchronizes
 * the stream once, after binning, to size the entry buffers to the frame's
 * (splat, tile) entries; a count above the settings' limit fails with
 * RDX_E_CAPACITY (nothing is drawn). */
int rdx_draw_splats(rdx_draw* draw, void* stream, const uint32_t* words_dev, uint32_t n, const uint32_t* order_dev,
                    const uint32_t* count_dev, uint32_t capacity, uint32_t width, uint32_t height,
                    const rdx_params* params, float* rgba_dev);
/* The draw of the ray-evaluated methods' words, one eye (splatDraw.ts:98-181
 * with the methods' vertex and density stages: 3dgut and htgs,
 * projections.tsl.ts:310-330 and rayVertex / rayDensity,
 * projection.tsl.ts:412-442; raygs, projections.tsl.ts:432-466): the words
 * rdx_project_ray wrote (channels w0, w1 and, with two eyes, w2; n splats)
 * with the camera they were projected for, `eye` of its eyes, whose
 * SplatCamera fromHead (view * head^-1), projection and render size place
 * them. 3dgut and htgs: the quad is the eye's screen bound (its corner
 * shrunk to cut / bound for 3dgut when the bound lies on screen), each pixel
 * evaluating the Gaussian along its ray in the splat's whitened frame,
 * z^2 = |localCross|^2 / |localDir|^2; raygs: the facing quad of the eye's
 * frame, z^2 = |z|^2 c^2 / (c^2 + |z|^2) at the ray's hit. Cutoff, alpha,
 * blending, order, count and rgba as rdx_draw_splats. method:
 * RDX_METHOD_3DGUT, RDX_METHOD_HTGS or RDX_METHOD_RAYGS. */
int rdx_draw_ray(rdx_draw* draw, void* stream, int method, uint32_t* const* channels_dev, uint32_t n,
                 const uint32_t* order_dev, const uint32_t* count_dev, uint32_t capacity, const rdx_camera* camera,
                 int eye, const rdx_params* params, float* rgba_dev);
/* The last draw's drawn positions, (splat, tile) entries and entry
 * capacity; overflowed is 0 (a count past the limit fails the draw).
 * Synchronizes. */
int rdx_draw_stats(rdx_draw* draw, void* stream, uint32_t* drawn, uint32_t* entries, uint32_t* capacity,
                   int* overflowed);

#ifdef __cplusplus
}
#endif
#endif /* RADIX_HIP_H */

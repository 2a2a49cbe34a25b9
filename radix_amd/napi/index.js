'use strict';
// The Node host side of the splat pipeline: the reference's interfaces
// (src/core/sort.ts SortKeys, src/worker/splatSort.ts sortSplatKeys,
// src/wgpu/WgpuRadixSort.ts, src/wgpu/WgpuSplatSorter.ts and the
// projection writer of src/render/SplatProjection.ts) over the N-API
// binding of libradix_hip.so, plus the grown forms with no reference
// counterpart: the differentiable compositor and its adjoint (Rasterizer),
// the draw for native hosts (SplatDrawer) and many views in one sequence
// (SplatBatch). GPU calls take an optional { stream } (a GpuStream): calls on
// different streams are in flight together, each with its own scratch; a
// stream's done() resolves when its work has finished. Plain ES2018 so it
// runs on Node 12 as well as under TypeScript (types in index.d.ts).
const path = require('path');

const native = require(path.join(__dirname, 'radix_napi.node'));

const ROW_ELEMENTS = 1024 * 4;

// TextureLayout.alignCapacity (src/core/layout.ts:70-72)
function alignCapacity(count) {
  return Math.max(1, Math.ceil(count / ROW_ELEMENTS)) * ROW_ELEMENTS;
}

// src/core/sort.ts: keys of distances and their stable sort (on the GPU)
class SortKeys {
  static invalid(bits) {
    return native.keyInvalid(bits);
  }

  static validMax(bits) {
    return native.keyValidMax(bits);
  }

  static compute(distance, bits) {
    return native.keyCompute(distance, bits);
  }

  // {keys, count, bits, ordering} -> {ordering, visible}; throws the
  // reference's Error messages (sort.ts:70-72)
  static sort({ keys, count, bits, ordering }) {
    const n = count === undefined ? keys.length : count;
    const out = ordering || new Uint32Array(alignCapacity(n));
    const visible = native.sortKeys(keys, n, bits, out);
    return { ordering: out, visible };
  }
}
SortKeys.KEY16_LO = 95 << 23;
SortKeys.KEY16_HI = (159 << 23) - 1;

// sortSplatKeys (src/worker/splatSort.ts:17-30): off the main thread (the
// libuv pool runs the call), the arrays handed back for reuse
function sortSplatKeys(keys, { bits, ordering } = {}) {
  const count = keys.length;
  const reuse = ordering && ordering.length >= alignCapacity(count) ? ordering : new Uint32Array(alignCapacity(count));
  return native.sortKeysAsync(keys, count, bits, reuse).then((visible) => ({ ordering: reuse, visible, keys }));
}

// A HIP stream with its own scratch: the calls given it run in its order,
// overlapping with calls on other streams (one per in-flight call)
class GpuStream {
  constructor() {
    this.handle = native.streamCreate();
  }

  // Waits (blocking this thread) for the stream's work
  synchronize() {
    native.streamSynchronize(this.handle);
  }

  // A Promise resolved when the stream's work has finished (waited on the libuv pool)
  done() {
    return native.streamSynchronizeAsync(this.handle);
  }
}

const streamOf = (options) => (options && options.stream ? options.stream.handle : undefined);
const handleOf = (b) => (b && b.handle !== undefined ? b.handle : b === undefined ? null : b);
// The ray-evaluated methods' codes (RDX_METHOD_3DGUT, _HTGS, _RAYGS)
const rayMethod = (method, who = 'SplatProjection') => {
  const code = { '3dgut': 1, htgs: 2, raygs: 3 }[method];
  if (code === undefined) throw new Error(`${who}: unknown method ${method}`);
  return code;
};

// Device memory (the storage of a WgpuBuffer, src/wgpu/WgpuBuffer.ts)
class DeviceBuffer {
  constructor(bytes) {
    this.bytes = bytes;
    this.handle = native.deviceAlloc(bytes);
  }

  static fromArray(array, options) {
    const b = new DeviceBuffer(array.byteLength);
    b.upload(array, 0, options);
    return b;
  }

  // Waits for the copy (and, with a stream, for that stream's earlier work)
  upload(array, byteOffset = 0, options) {
    native.deviceUpload(this.handle, array, byteOffset, streamOf(options));
  }

  // The first `count` uint32 from element `start`, after the stream's earlier work
  read(start = 0, count = this.bytes / 4 - start, Type = Uint32Array, options) {
    const out = new Type(count);
    native.deviceDownload(this.handle, start * out.BYTES_PER_ELEMENT, out, streamOf(options));
    return out;
  }

  fill(value, count = this.bytes / 4, options) {
    native.deviceFillU32(this.handle, value, count, streamOf(options));
  }

  dispose() {
    native.deviceFree(this.handle);
  }
}

// WgpuRadixSort: stable LSD sort of device keys and values, in place
class RadixSort {
  constructor({ keys, values }) {
    this.keys = keys;
    this.values = values;
  }

  get capacity() {
    return Math.floor(Math.min(this.keys.bytes, this.values.bytes) / 4);
  }

  sort(count, { bits = 32, stream } = {}) {
    if (count <= 0) return;
    if (count > this.capacity) {
      throw new Error(`WgpuRadixSort: ${count} keys past the buffers' capacity of ${this.capacity}`);
    }
    native.sortPairs(this.keys.handle, this.values.handle, count, bits, streamOf({ stream }));
  }

  sortIndirect(countBuffer, { bits = 32, stream } = {}) {
    native.sortPairsIndirect(this.keys.handle, this.values.handle, countBuffer.handle, this.capacity, bits,
      streamOf({ stream }));
  }

  dispose() {}
}

// WgpuSplatSorter: drop invalid keys, order the rest back to front, write
// the drawIndexedIndirect arguments
class SplatSorter {
  constructor(capacity = 1) {
    this.values = new DeviceBuffer(Math.max(capacity, 1) * 4);
    this.count = new DeviceBuffer(4);
    this.indirect = new DeviceBuffer(20);
  }

  sort(keys, count, bits, options) {
    if (count * 4 > keys.bytes) {
      throw new Error(`WgpuSplatSorter: ${count} keys past the buffer's ${Math.floor(keys.bytes / 4)}`);
    }
    if (this.values.bytes < count * 4) {
      this.values.dispose();
      this.values = new DeviceBuffer(count * 4);
    }
    native.sortSplats(keys.handle, count, bits, this.values.handle, this.count.handle, this.indirect.handle,
      streamOf(options));
  }

  readCount(options) {
    return Promise.resolve(this.count.read(0, 1, Uint32Array, options)[0]);
  }

  dispose() {
    [this.values, this.count, this.indirect].forEach((b) => b.dispose());
  }
}
SplatSorter.QUADS_PER_INSTANCE = 1024;

const LAYOUTS = { f32: 0, extSplats: 1, packedSplats: 2, covSplats: 3 };
const SH_CODECS = { none: 0, f32: 1, ext: 2, packed: 3 };

function tableDesc(table) {
  const d = Object.assign({}, table);
  d.layout = typeof table.layout === 'string' ? LAYOUTS[table.layout] : table.layout || 0;
  d.shCodec = typeof table.shCodec === 'string' ? SH_CODECS[table.shCodec] : table.shCodec || 0;
  ['planes', 'extA', 'extB', 'covA', 'covB', 'packed'].forEach((k) => {
    if (d[k]) d[k] = handleOf(d[k]);
  });
  if (d.shChannels) d.shChannels = d.shChannels.map(handleOf);
  return d;
}

// The ewa projection writer (SplatProjection.write) and the per-frame
// project + GPU sort (SplatRenderer.update): camera as SplatCamera.setViews
// takes it ({head, eyes: [{view, projection, ortho}], size, ndcZMin},
// matrices as Matrix4.elements), params as SplatParams' fields
class SplatProjector {
  project(table, camera, params, { words, keys, at = 0, stream }) {
    native.projectEwa(tableDesc(table), camera, params || {}, words.map(handleOf), handleOf(keys), at,
      streamOf({ stream }));
  }

  projectSort(table, camera, params, { words, keys, indices, count, indirect, stream }) {
    native.projectSort(tableDesc(table), camera, params || {}, (words || []).map(handleOf), handleOf(keys),
      handleOf(indices), handleOf(count), handleOf(indirect), streamOf({ stream }));
  }

  // The keys-only pass (SplatProjection.writeKeys, src/render/SplatProjection.ts:165-172)
  projectKeys(table, camera, params, { keys, at = 0, stream }) {
    native.projectKeys(tableDesc(table), camera, params || {}, handleOf(keys), at, streamOf({ stream }));
  }

  // The words of "3dgut", "htgs" or "raygs" into the projected bundle's
  // channels w0, w1 (and w2 with two eyes) (SplatProjection.ts:46-53, 141-151)
  projectRay(table, camera, params, method, { channels, keys, at = 0, stream }) {
    native.projectRay(tableDesc(table), camera, params || {}, rayMethod(method), channels.map(handleOf), handleOf(keys), at,
      streamOf({ stream }));
  }
}

// The differentiable compositor (no reference counterpart): the ewa splats
// of 59 parameter planes (centre 3, log-scale 3, quaternion 4, opacity
// logit, base colour 3, SH 45) composited front to back over 16x16 tiles
// with the draw's fragment alpha and premultiplied "over"
// (src/render/splatDraw.ts:62-174, projection.tsl.ts:444-455), and its
// hand-written adjoint. Images are 3 planes of width x height floats, row 0
// at the bottom. One forward's state at a time: backward() differentiates
// the last forward (forward() returns its generation; a backward given an
// older one throws).
class Rasterizer {
  constructor({ shDegree = 3, tStop = 1e-4, maxEntriesPerSplat = 0, stream } = {}) {
    this.handle = native.rasterCreate({ shDegree, tStop, maxEntriesPerSplat }, streamOf({ stream }));
    this.generation = 0;
  }

  forward(params, count, camera, renderParams, { image, alpha, stream } = {}) {
    native.rasterForward(this.handle, handleOf(params), count, camera, renderParams || {}, handleOf(image),
      handleOf(alpha), streamOf({ stream }));
    this.generation += 1;
    return this.generation;
  }

  backward(gradImage, gradParams, { generation, stream } = {}) {
    if (generation !== undefined && generation !== this.generation) {
      throw new Error(`Rasterizer: the backward of forward ${generation} was asked after forward ${this.generation} replaced its state`);
    }
    native.rasterBackward(this.handle, handleOf(gradImage), handleOf(gradParams), streamOf({ stream }));
  }

  stats(options) {
    return native.rasterStats(this.handle, streamOf(options));
  }

  dispose() {
    native.rasterDestroy(this.handle);
  }
}

// The splat draw for native hosts (splatDraw.ts:98-181 without a
// rasteriser): one eye's ewa words in a back-to-front order composited into
// premultiplied RGBA (4 planes of width x height floats, row 0 at the top)
class SplatDrawer {
  constructor({ tStop = 0, linearOutput = false, maxEntriesPerSplat = 0, stream } = {}) {
    this.handle = native.drawCreate({ tStop, linearOutput, maxEntriesPerSplat }, streamOf({ stream }));
  }

  draw(words, n, order, count, capacity, width, height, params, rgba, options) {
    native.drawSplats(this.handle, handleOf(words), n, handleOf(order), handleOf(count), capacity, width, height,
      params || {}, handleOf(rgba), streamOf(options));
  }

  // One eye of the "3dgut", "htgs" or "raygs" words projectRay wrote
  // (channels w0, w1 and, with two eyes, w2) with the camera they were
  // projected for, at its render size
  drawRay(method, channels, n, order, count, capacity, camera, params, rgba, { eye = 0, stream } = {}) {
    native.drawRay(this.handle, rayMethod(method, 'SplatDrawer.drawRay'), channels.map(handleOf), n, handleOf(order), handleOf(count),
      capacity, camera, eye, params || {}, handleOf(rgba), streamOf({ stream }));
  }

  stats(options) {
    return native.drawStats(this.handle, streamOf(options));
  }

  dispose() {
    native.drawDestroy(this.handle);
  }
}

// Many views of many collections projected and segment-sorted in one
// sequence (no reference counterpart: the reference projects and sorts per
// object and render, SplatProjector.ts:127-137). views: [{table, camera,
// words?}] (table: an index into tables; one eye per camera). View v's
// keys and indices are at offsets[v] (SplatCollection's offsets,
// SplatCollection.ts:54-65).
class SplatBatch {
  // tuning: [digitBits, keysPerThread, blockThreads] for this batch's sort
  // (the persistent form needs [8, 12, 512]); default: the context's.
  constructor(tables, views, params, { keys, stream, tuning } = {}) {
    this.tables = tables;
    this.views = views;
    this.handle = native.batchCreate(tables.map(tableDesc), views.map((v) => ({
      table: v.table, camera: v.camera, words: handleOf(v.words),
    })), params || {}, handleOf(keys), streamOf({ stream }), tuning);
    const info = native.batchInfo(this.handle);
    this.total = info.total;
    this.offsets = info.offsets;
  }

  setCameras(cameras) {
    native.batchSetCameras(this.handle, cameras);
  }

  // indices: total uint32; counts: a uint32 per view; drawArgs: 5 per view.
  // graph: capture the sequence into a HIP graph once and replay it.
  // persistent: the whole batch as one launch of the persistent kernel
  // (tables of one kind, created with tuning [8, 12, 512], 32-bit keys).
  run({ indices, counts, drawArgs, graph = true, persistent = false, stream }) {
    native.batchRun(this.handle, handleOf(indices), handleOf(counts), handleOf(drawArgs), graph, persistent,
      streamOf({ stream }));
  }

  // Waits for the last run; throws if a persistent run's barrier wait ran out
  // (that run's output is then not valid).
  status() {
    native.batchStatus(this.handle);
  }

  dispose() {
    native.batchDestroy(this.handle);
  }
}

// An RDX payload, gunzipped into `planes` (a DeviceBuffer), back to `count`
// elements of `elementBytes` in `out` from byte `outOffset` (the second step
// of Compression.decompressTransposed, src/core/compress.ts:197-204)
function untransposeBytes(planes, count, elementBytes, out, outOffset = 0, options) {
  native.untransposeBytes(handleOf(planes), count, elementBytes, handleOf(out), outOffset, streamOf(options));
}

const SPZ_MAGIC = 0x5053474e;
const SH_VECTORS = [0, 3, 8, 15, 24];

// An SPZ file (Niantic's gzip container) decoded onto the device as an f32
// table for SplatProjector (SpzFile.read, src/core/spz.ts:140-170, 260-349):
// Node's zlib gunzips, the header is read here, one upload puts the columns
// in device memory and the decode kernel writes the planes (centre 3,
// log-scale 3, quaternion 4, opacity, rgb 3, SH 3K; stride `count`).
// Returns { table, meta }; table.planes is the caller's to dispose.
function readSpz(bytes) {
  const data = Buffer.from(bytes.buffer, bytes.byteOffset, bytes.byteLength);
  if (data.length >= 4 && data.readUInt32LE(0) === SPZ_MAGIC) {
    throw new Error("SPZ 4's NGSP (zstd) container is not supported here");
  }
  const raw = require('zlib').gunzipSync(data);
  if (raw.length < 16 || raw.readUInt32LE(0) !== SPZ_MAGIC) throw new Error('Invalid SPZ file');
  const version = raw.readUInt32LE(4);
  const count = raw.readUInt32LE(8);
  const [shDegree, fractionalBits, flags] = [raw[12], raw[13], raw[14]];
  if (version < 1 || version > 4 || !(shDegree in SH_VECTORS) || shDegree > 3) {
    throw new Error(`Unsupported SPZ version ${version} or SH degree ${shDegree}`);
  }
  const body = raw.subarray(16);
  const n = Math.max(count, 1);
  const planes = new DeviceBuffer(4 * n * (14 + 3 * SH_VECTORS[shDegree]));
  const staged = DeviceBuffer.fromArray(body.length ? body : new Uint8Array(1));
  try {
    native.spzDecode(staged.handle, body.length, count, version, shDegree, fractionalBits, planes.handle, n);
  } catch (e) {
    planes.dispose();
    throw e;
  } finally {
    staged.dispose();
  }
  const table = { layout: 'f32', count, planes, planeStride: n, shCodec: shDegree > 0 ? 'f32' : 'none', shDegree };
  return { table, meta: { format: 'spz', version, count, shDegree, fractionalBits, antialiased: (flags & 1) !== 0 } };
}

module.exports = {
  untransposeBytes, readSpz,
  alignCapacity, SortKeys, sortSplatKeys, GpuStream, DeviceBuffer, RadixSort, SplatSorter, SplatProjector,
  Rasterizer, SplatDrawer, SplatBatch,
  version: native.version,
};

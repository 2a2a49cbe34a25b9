'use strict';
// The Node host side of the splat pipeline: the reference's interfaces
// (src/core/sort.ts SortKeys, src/worker/splatSort.ts sortSplatKeys,
// src/wgpu/WgpuRadixSort.ts, src/wgpu/WgpuSplatSorter.ts and the
// projection writer of src/render/SplatProjection.ts) over the N-API
// binding of libradix_hip.so. Plain ES2018 so it runs on Node 12 as well
// as under TypeScript (types in index.d.ts).
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

// Device memory (the storage of a WgpuBuffer, src/wgpu/WgpuBuffer.ts)
class DeviceBuffer {
  constructor(bytes) {
    this.bytes = bytes;
    this.handle = native.deviceAlloc(bytes);
  }

  static fromArray(array) {
    const b = new DeviceBuffer(array.byteLength);
    b.upload(array);
    return b;
  }

  upload(array, byteOffset = 0) {
    native.deviceUpload(this.handle, array, byteOffset);
  }

  // The first `count` uint32 from element `start`
  read(start = 0, count = this.bytes / 4 - start, Type = Uint32Array) {
    const out = new Type(count);
    native.deviceDownload(this.handle, start * out.BYTES_PER_ELEMENT, out);
    return out;
  }

  fill(value, count = this.bytes / 4) {
    native.deviceFillU32(this.handle, value, count);
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
    return Math.min(this.keys.bytes, this.values.bytes) / 4;
  }

  sort(count, { bits = 32 } = {}) {
    if (count <= 0) return;
    native.sortPairs(this.keys.handle, this.values.handle, count, bits);
  }

  sortIndirect(countBuffer, { bits = 32 } = {}) {
    native.sortPairsIndirect(this.keys.handle, this.values.handle, countBuffer.handle, this.capacity, bits);
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

  sort(keys, count, bits) {
    if (this.values.bytes < count * 4) {
      this.values.dispose();
      this.values = new DeviceBuffer(count * 4);
    }
    native.sortSplats(keys.handle, count, bits, this.values.handle, this.count.handle, this.indirect.handle);
  }

  readCount() {
    return Promise.resolve(this.count.read(0, 1)[0]);
  }

  dispose() {
    [this.values, this.count, this.indirect].forEach((b) => b.dispose());
  }
}
SplatSorter.QUADS_PER_INSTANCE = 1024;

const LAYOUTS = { f32: 0, extSplats: 1, packedSplats: 2 };
const SH_CODECS = { none: 0, f32: 1, ext: 2, packed: 3 };

function tableDesc(table) {
  const d = Object.assign({}, table);
  d.layout = typeof table.layout === 'string' ? LAYOUTS[table.layout] : table.layout || 0;
  d.shCodec = typeof table.shCodec === 'string' ? SH_CODECS[table.shCodec] : table.shCodec || 0;
  const unwrap = (b) => (b && b.handle !== undefined ? b.handle : b);
  ['planes', 'extA', 'extB', 'packed'].forEach((k) => {
    if (d[k]) d[k] = unwrap(d[k]);
  });
  if (d.shChannels) d.shChannels = d.shChannels.map(unwrap);
  return d;
}

// The ewa projection writer (SplatProjection.write) and the per-frame
// project + GPU sort (SplatRenderer.update): camera as SplatCamera.setViews
// takes it ({head, eyes: [{view, projection, ortho}], size, ndcZMin},
// matrices as Matrix4.elements), params as SplatParams' fields
class SplatProjector {
  project(table, camera, params, { words, keys, at = 0 }) {
    native.projectEwa(tableDesc(table), camera, params || {}, words.map((w) => w.handle), keys ? keys.handle : null, at);
  }

  projectSort(table, camera, params, { words, keys, indices, count, indirect }) {
    native.projectSort(tableDesc(table), camera, params || {}, (words || []).map((w) => w.handle),
      keys ? keys.handle : null, indices.handle, count ? count.handle : null, indirect ? indirect.handle : null);
  }

  // The keys-only pass (SplatProjection.writeKeys, src/render/SplatProjection.ts:165-172)
  projectKeys(table, camera, params, { keys, at = 0 }) {
    native.projectKeys(tableDesc(table), camera, params || {}, keys.handle, at);
  }

  // The words of "3dgut", "htgs" or "raygs" into the projected bundle's
  // channels w0, w1 (and w2 with two eyes) (SplatProjection.ts:46-53, 141-151)
  projectRay(table, camera, params, method, { channels, keys, at = 0 }) {
    const code = { '3dgut': 1, htgs: 2, raygs: 3 }[method];
    if (code === undefined) throw new Error(`SplatProjection: unknown method ${method}`);
    native.projectRay(tableDesc(table), camera, params || {}, code, channels.map((c) => c.handle), keys ? keys.handle : null, at);
  }
}

// An RDX payload, gunzipped into `planes` (a DeviceBuffer), back to `count`
// elements of `elementBytes` in `out` from byte `outOffset` (the second step
// of Compression.decompressTransposed, src/core/compress.ts:197-204)
function untransposeBytes(planes, count, elementBytes, out, outOffset = 0) {
  native.untransposeBytes(planes.handle, count, elementBytes, out.handle, outOffset);
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
  alignCapacity, SortKeys, sortSplatKeys, DeviceBuffer, RadixSort, SplatSorter, SplatProjector,
  version: native.version,
};

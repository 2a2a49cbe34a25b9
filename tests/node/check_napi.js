'use strict';
// Checks of the Node binding, written like the reference's own checks
// (test/gpu/renderTests.ts, scripts/check-core.ts): each check returns
// undefined or a failure message. `node check_napi.js cpu` runs the checks
// that make no GPU call; `gpu` adds the GPU sorts and projection.
const path = require('path');
const radix = require(path.join(__dirname, '..', '..', 'radix_amd', 'napi', 'index.js'));
const { SortKeys, sortSplatKeys, DeviceBuffer, RadixSort, SplatSorter, SplatProjector, alignCapacity, readSpz, GpuStream,
  Rasterizer, SplatDrawer, SplatBatch, untransposeBytes } = radix;
const native = require(path.join(__dirname, '..', '..', 'radix_amd', 'napi', 'radix_napi.node'));
const fs = require('fs');
const zlib = require('zlib');

const failures = [];
async function check(name, fn) {
  let detail;
  try {
    detail = await fn();
  } catch (error) {
    detail = String(error && error.message ? error.message : error);
  }
  console.log(`${detail === undefined ? 'ok  ' : 'FAIL'}  ${name}${detail ? `  (${detail})` : ''}`);
  if (detail !== undefined) failures.push(name);
}

// mulberry32 (test/gpu/renderUtils.ts:23-30)
function rng(seed) {
  return () => {
    seed = (seed + 0x6d2b79f5) | 0;
    let t = Math.imul(seed ^ (seed >>> 15), 1 | seed);
    t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t;
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

// The stable order of keys (the low `bits` bits; 16 bits ignore the rest) by
// Array.prototype.sort, stable since V8 7.0
function stableOrder(keys, count, bits) {
  const mask = bits === 32 ? 0xffffffff : 2 ** bits - 1;
  const idx = Array.from({ length: count }, (_, i) => i);
  idx.sort((a, b) => ((keys[a] & mask) >>> 0) - ((keys[b] & mask) >>> 0));
  return idx;
}

function firstMismatch(actual, expected, n) {
  for (let i = 0; i < n; i += 1) {
    if (actual[i] !== expected[i]) return `index ${i} is ${actual[i]}, expected ${expected[i]}`;
  }
  return undefined;
}

async function cpuChecks() {
  await check('sort keys: known answer and invalid keys', () => {
    if (SortKeys.compute(1, 32) !== 0xc07fffff) return `key of 1: ${SortKeys.compute(1, 32)}`;
    for (const bits of [16, 24, 32]) {
      if ([Infinity, NaN, -Infinity].some((d) => SortKeys.compute(d, bits) !== SortKeys.invalid(bits))) return `${bits} bits`;
    }
    return SortKeys.validMax(32) === 0xfeffffff && alignCapacity(4097) === 8192 ? undefined : 'validMax / alignCapacity';
  });
  await check('SortKeys.sort: the reference errors', () => {
    try {
      SortKeys.sort({ keys: new Uint32Array(10), count: 10, bits: 24, ordering: new Uint32Array(3) });
      return 'no error';
    } catch (e) {
      return e.message === 'SortKeys.sort: ordering holds 3 of 10 indices' ? undefined : e.message;
    }
  });
}

// The error a call throws, or a failure message when it does not throw that one
function throwsWith(fn, text) {
  try {
    fn();
  } catch (e) {
    return e.message.includes(text) ? undefined : `threw "${e.message}", expected "${text}"`;
  }
  return `no error (expected "${text}")`;
}

// Argument checks that throw before any GPU call
async function cpuBoundsChecks() {
  await check('bounds: counts past the arrays and buffers throw before any device work', () => {
    return throwsWith(() => native.sortKeysAsync(new Uint32Array(4), 10, 24, new Uint32Array(8)), 'count above the keys given')
      || throwsWith(() => new RadixSort({ keys: { bytes: 16, handle: null }, values: { bytes: 64, handle: null } }).sort(10),
        "past the buffers' capacity of 4")
      || throwsWith(() => SplatSorter.prototype.sort.call({ values: { bytes: 0 } }, { bytes: 8 }, 10, 32),
        "past the buffer's 2")
      || throwsWith(() => native.sortPairs({}, {}, 10, 32), 'WgpuRadixSort.sort: keys is not a device buffer')
      || throwsWith(() => native.sortPairs(null, null, 0, 32, 5), 'the last argument is not a stream')
      || throwsWith(() => native.sortSplats(undefined, 10, 32), 'WgpuSplatSorter.sort: keys is required')
      || throwsWith(() => native.untransposeBytes({}, 4, 16, {}, 0), 'untransposeBytes: planes is not a device buffer')
      || throwsWith(() => native.projectSort({ layout: 0, count: 10 }, { eyes: [] }, {}, [], null, null),
        'table.planes is required')
      || throwsWith(() => native.rasterForward({}, null, 1), 'Rasterizer.forward: not a rasterizer');
  });
}

// Bounds checks against real device buffers: each call throws, the device untouched
async function gpuBoundsChecks() {
  await check('bounds: short device buffers throw (sorts, projection, loaders, compositor, draws)', () => {
    const small = new DeviceBuffer(40);
    const big = new DeviceBuffer(4096);
    const planes = new DeviceBuffer(14 * 4 * 8);
    const eye = { view: [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1], projection: [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, -1, -1, 0, 0, -0.1, 0] };
    const camera = { eyes: [eye], size: [16, 16], ndcZMin: 0 };
    const rast = new Rasterizer();
    const drawer = new SplatDrawer();
    const out = throwsWith(() => native.sortPairs(small.handle, big.handle, 100, 32), 'keys holds 40 bytes, 400 needed')
      || throwsWith(() => native.sortSplats(big.handle, 100, 32, small.handle, null, null), 'values holds 40 bytes, 400 needed')
      || throwsWith(() => native.projectSort({ layout: 0, count: 16, planes: planes.handle }, camera, {}, [], null, big.handle),
        'table.planes holds 448 bytes, 896 needed')
      || throwsWith(() => untransposeBytes(big, 100, 16, big, 3000), 'out holds 4096 bytes, 4600 needed')
      || throwsWith(() => native.spzDecode(big.handle, 100, 10, 3, 0, 12, small.handle, 10), 'planes holds 40 bytes, 560 needed')
      || throwsWith(() => rast.forward(small, 100, camera, {}, { image: big }), 'params holds 40 bytes, 23600 needed')
      || throwsWith(() => drawer.draw(big, 10, big, big, 10, 16, 16, {}, small), 'rgba holds 40 bytes, 4096 needed')
      || throwsWith(() => drawer.drawRay('htgs', [big, small], 10, big, big, 10, camera, {}, big), 'channels[1] holds 40 bytes, 160 needed')
      || throwsWith(() => drawer.drawRay('htgs', [big, big], 10, big, big, 10, camera, {}, small), 'rgba holds 40 bytes, 4096 needed');
    [small, big, planes].forEach((b) => b.dispose());
    rast.dispose();
    drawer.dispose();
    return out;
  });
}

async function gpuChecks() {
  await check('SortKeys.sort equals the stable order: 16/24/32 bits, counts 0, 1, block edges, large', () => {
    const random = rng(7);
    for (const bits of [16, 24, 32]) {
      for (const count of [0, 1, 4095, 4096, 4097, 70000]) {
        const keys = new Uint32Array(Math.max(count, 1)).map(() => Math.floor(random() * 2 ** bits) >>> 0);
        const { ordering } = SortKeys.sort({ keys, count, bits });
        const m = firstMismatch(ordering, stableOrder(keys, count, bits), count);
        if (m) return `${bits} bits, ${count} keys: ${m}`;
      }
    }
    return undefined;
  });

  await check('RadixSort (WgpuRadixSort) equals the stable order', () => {
    const random = rng(8);
    for (const bits of [16, 24, 32]) {
      for (const count of [1, 4097, 70000]) {
        const keys = new Uint32Array(count).map(() => Math.floor(random() * 2 ** bits) >>> 0);
        const kb = DeviceBuffer.fromArray(keys);
        const vb = DeviceBuffer.fromArray(new Uint32Array(count).map((_, i) => i));
        new RadixSort({ keys: kb, values: vb }).sort(count, { bits });
        const m = firstMismatch(vb.read(0, count), stableOrder(keys, count, bits), count);
        kb.dispose();
        vb.dispose();
        if (m) return `${bits} bits, ${count}: ${m}`;
      }
    }
    return undefined;
  });

  await check('SplatSorter (WgpuSplatSorter): invalid keys dropped, draw arguments', async () => {
    const random = rng(11);
    for (const bits of [16, 24, 32]) {
      const count = 20000;
      const keys = new Uint32Array(count).map(() => (random() < 0.3 ? SortKeys.invalid(bits) : Math.floor(random() * 0xff * 2 ** (bits - 8))) >>> 0);
      const kb = DeviceBuffer.fromArray(keys);
      const sorter = new SplatSorter(count);
      sorter.sort(kb, count, bits);
      const { ordering, visible } = SortKeys.sort({ keys, count, bits });
      const got = await sorter.readCount();
      const args = Array.from(sorter.indirect.read(0, 5));
      const expected = [6 * 1024, Math.ceil(visible / 1024), 0, 0, 0];
      const m = firstMismatch(sorter.values.read(0, visible), ordering, visible);
      kb.dispose();
      sorter.dispose();
      if (got !== visible || args.join() !== expected.join()) return `${bits}: count ${got} vs ${visible}, args ${args}`;
      if (m) return `${bits}: ${m}`;
    }
    return undefined;
  });

  await check('sortSplatKeys: the pool and this thread agree', async () => {
    const random = rng(9);
    const keys = new Uint32Array(50000).map(() => (random() < 0.1 ? 0xffffff : Math.floor(random() * 0xfeffff)));
    const here = SortKeys.sort({ keys: keys.slice(), count: keys.length, bits: 24 });
    const there = await sortSplatKeys(keys.slice(), { bits: 24 });
    if (here.visible !== there.visible) return `visible ${here.visible} vs ${there.visible}`;
    return firstMismatch(there.ordering, here.ordering, keys.length) || (there.keys.length === keys.length ? undefined : 'keys not returned');
  });

  await check('SplatProjector.projectSort: the order of SortKeys.sort on its keys', () => {
    // 4096 splats in front of a camera at the origin (renderUtils.ts:49-62 ranges)
    const random = rng(3);
    const n = 4096;
    const planes = new Float32Array(14 * n);
    for (let i = 0; i < n; i += 1) {
      planes[i] = (random() - 0.5) * 4;
      planes[n + i] = (random() - 0.5) * 3;
      planes[2 * n + i] = -2 - random() * 6;
      for (let k = 3; k < 6; k += 1) planes[k * n + i] = Math.log(0.01 + random() * 0.1);
      planes[9 * n + i] = 1; // quaternion (0, 0, 0, 1)
      planes[10 * n + i] = 0.2 + random() * 0.8;
      for (let k = 11; k < 14; k += 1) planes[k * n + i] = random();
    }
    const f = 1 / Math.tan(Math.PI / 6);
    const near = 0.1;
    const far = 100;
    const aspect = 2;
    // three.js makePerspective, WebGPU clip convention, column-major
    const projection = [f / aspect, 0, 0, 0, 0, f, 0, 0, 0, 0, -far / (far - near), -1, 0, 0, (-far * near) / (far - near), 0];
    const view = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1];
    const camera = { head: view, eyes: [{ view, projection }], size: [128, 64], ndcZMin: 0 };
    const table = { layout: 'f32', count: n, planes: DeviceBuffer.fromArray(planes) };
    const words = [new DeviceBuffer(16 * n)];
    const keys = new DeviceBuffer(4 * n);
    const indices = new DeviceBuffer(4 * n);
    const count = new DeviceBuffer(4);
    const indirect = new DeviceBuffer(20);
    new SplatProjector().projectSort(table, camera, { sortBits: 32 }, { words, keys, indices, count, indirect });
    const k = keys.read(0, n);
    const { ordering, visible } = SortKeys.sort({ keys: k, count: n, bits: 32 });
    const got = count.read(0, 1)[0];
    if (got !== visible || visible < n / 2) return `count ${got}, visible ${visible}`;
    return firstMismatch(indices.read(0, visible), ordering, visible);
  });
}

async function gpuChecksKeysAndRays() {
  await check('SplatProjector.projectKeys and projectRay: the keys of projectEwa with cullKeys off', () => {
    const random = rng(5);
    const n = 4096;
    const planes = new Float32Array(14 * n);
    for (let i = 0; i < n; i += 1) {
      planes[i] = (random() - 0.5) * 4;
      planes[n + i] = (random() - 0.5) * 3;
      planes[2 * n + i] = -2 - random() * 6;
      for (let k = 3; k < 6; k += 1) planes[k * n + i] = Math.log(0.01 + random() * 0.1);
      planes[9 * n + i] = 1;
      planes[10 * n + i] = i % 7 === 0 ? 0.0005 : 0.2 + random() * 0.8; // some below minAlpha
      for (let k = 11; k < 14; k += 1) planes[k * n + i] = random();
    }
    const f = 1 / Math.tan(Math.PI / 6);
    const near = 0.1;
    const far = 100;
    const projection = [f / 2, 0, 0, 0, 0, f, 0, 0, 0, 0, -far / (far - near), -1, 0, 0, (-far * near) / (far - near), 0];
    const view = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1];
    const camera = { head: view, eyes: [{ view, projection }], size: [128, 64], ndcZMin: 0 };
    const table = { layout: 'f32', count: n, planes: DeviceBuffer.fromArray(planes) };
    const params = { sortBits: 32, cullKeys: false };
    const projector = new SplatProjector();
    const ewaKeys = new DeviceBuffer(4 * n);
    projector.project(table, camera, params, { words: [new DeviceBuffer(16 * n)], keys: ewaKeys });
    const keys = new DeviceBuffer(4 * n);
    projector.projectKeys(table, camera, params, { keys });
    const rayKeys = new DeviceBuffer(4 * n);
    const channels = [new DeviceBuffer(16 * n), new DeviceBuffer(16 * n)];
    projector.projectRay(table, camera, params, '3dgut', { channels, keys: rayKeys });
    const a = ewaKeys.read(0, n);
    const b = keys.read(0, n);
    const c = rayKeys.read(0, n);
    const invalid = a.filter((k) => k === 0xffffffff).length;
    if (invalid < n / 8) return `only ${invalid} invalid keys`;
    const w = channels[0].read(0, 4 * n);
    if (!w.some((x) => x !== 0)) return 'no 3dgut words';
    return firstMismatch(b, a, n) || firstMismatch(c, a, n);
  });
}

// readSpz's per-splat values in doubles (src/core/spz.ts:297-347, 571-586),
// plane by plane, over an unzipped body
function spzPlanes(body, n, version, shDegree, fractionalBits) {
  const K = [0, 3, 8, 15][shDegree];
  const out = Array.from({ length: 14 + 3 * K }, () => new Float64Array(n));
  const alphas = n * (version === 1 ? 6 : 9);
  const colors = alphas + n;
  const scales = colors + 3 * n;
  const rotations = scales + 3 * n;
  const sh = rotations + n * (version >= 3 ? 4 : 3);
  const dv = new DataView(body.buffer, body.byteOffset, body.byteLength);
  for (let i = 0; i < n; i += 1) {
    for (let k = 0; k < 3; k += 1) {
      const at = i * 9 + 3 * k;
      out[k][i] = (((body[at] | (body[at + 1] << 8) | (body[at + 2] << 16)) << 8) >> 8) / 2 ** fractionalBits;
      out[3 + k][i] = body[scales + 3 * i + k] / 16 - 10;
      out[11 + k][i] = (body[colors + 3 * i + k] / 255 - 0.5) * (0.28209479177387814 / 0.15) + 0.5;
    }
    const q = [0, 0, 0, 0];
    if (version >= 3) {
      let rest = dv.getUint32(rotations + 4 * i, true);
      const largest = rest >>> 30;
      let sum = 0;
      for (let k = 3; k >= 0; k -= 1) {
        if (k !== largest) {
          q[k] = Math.SQRT1_2 * ((rest & 511) / 511) * ((rest >>> 9) & 1 ? -1 : 1);
          rest >>>= 10;
          sum += q[k] * q[k];
        }
      }
      q[largest] = Math.sqrt(Math.max(1 - sum, 0));
    } else {
      for (let k = 0; k < 3; k += 1) q[k] = body[rotations + 3 * i + k] / 127.5 - 1;
      q[3] = Math.sqrt(Math.max(0, 1 - q[0] * q[0] - q[1] * q[1] - q[2] * q[2]));
    }
    for (let k = 0; k < 4; k += 1) out[6 + k][i] = q[k];
    out[10][i] = body[alphas + i] / 255;
    for (let v = 0; v < 3 * K; v += 1) out[14 + v][i] = (body[sh + 3 * K * i + v] - 128) / 128;
  }
  return out;
}

async function gpuChecksSpz() {
  await check('readSpz: planes equal readSpz\'s values (versions 2, 3; SH 0, 1, 3), its errors', () => {
    const random = rng(13);
    for (const [version, shDegree, n] of [[3, 1, 1000], [2, 3, 257], [3, 0, 1]]) {
      const K = [0, 3, 8, 15][shDegree];
      const bodyBytes = n * (9 + 1 + 3 + 3 + (version >= 3 ? 4 : 3) + 3 * K);
      const raw = new Uint8Array(16 + bodyBytes);
      const dv = new DataView(raw.buffer);
      dv.setUint32(0, 0x5053474e, true);
      dv.setUint32(4, version, true);
      dv.setUint32(8, n, true);
      raw[12] = shDegree;
      raw[13] = 12;
      raw[14] = 1;
      for (let b = 16; b < raw.length; b += 1) raw[b] = Math.floor(random() * 256);
      const { table, meta } = readSpz(zlib.gzipSync(raw));
      const expected = spzPlanes(raw.subarray(16), n, version, shDegree, 12);
      const got = table.planes.read(0, n * expected.length, Float32Array);
      table.planes.dispose();
      if (meta.count !== n || meta.version !== version || !meta.antialiased) return `meta ${JSON.stringify(meta)}`;
      for (let p = 0; p < expected.length; p += 1) {
        for (let i = 0; i < n; i += 1) {
          const e = Math.fround(expected[p][i]);
          const g = got[p * n + i];
          // one float ulp: the kernel rounds its double maps once
          if (Math.abs(g - e) > Math.abs(e) * 2 ** -23 + 1e-45) return `v${version} plane ${p} splat ${i}: ${g} vs ${e}`;
        }
      }
    }
    const errors = [
      [Buffer.from('NGSP0000000000000000'), 'NGSP'],
      [zlib.gzipSync(Buffer.alloc(32)), 'Invalid SPZ file'],
    ];
    for (const [bytes, text] of errors) {
      try {
        readSpz(bytes);
        return `no error for ${text}`;
      } catch (e) {
        if (!e.message.includes(text)) return e.message;
      }
    }
    return undefined;
  });
}

// A random f32 table in front of a camera at the origin (renderUtils.ts:49-62 ranges)
function randomPlanes(seed, n) {
  const random = rng(seed);
  const planes = new Float32Array(14 * n);
  for (let i = 0; i < n; i += 1) {
    planes[i] = (random() - 0.5) * 4;
    planes[n + i] = (random() - 0.5) * 3;
    planes[2 * n + i] = -2 - random() * 6;
    for (let k = 3; k < 6; k += 1) planes[k * n + i] = Math.log(0.01 + random() * 0.1);
    planes[9 * n + i] = 1;
    planes[10 * n + i] = 0.2 + random() * 0.8;
    for (let k = 11; k < 14; k += 1) planes[k * n + i] = random();
  }
  return planes;
}

function cameraAt(x, width = 128, height = 64) {
  const f = 1 / Math.tan(Math.PI / 6);
  const near = 0.1;
  const far = 100;
  const projection = [f / (width / height), 0, 0, 0, 0, f, 0, 0, 0, 0, -far / (far - near), -1, 0, 0, (-far * near) / (far - near), 0];
  const view = [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, -x, 0, 0, 1];
  return { head: view, eyes: [{ view, projection }], size: [width, height], ndcZMin: 0 };
}

async function gpuChecksStreams() {
  await check('GpuStream: two projectSort calls in flight on their own streams, each the order of SortKeys.sort', async () => {
    const runs = [[31, 6000, 0.0], [32, 9000, 0.4]].map(([seed, n, x]) => {
      const stream = new GpuStream();
      const table = { layout: 'f32', count: n, planes: DeviceBuffer.fromArray(randomPlanes(seed, n), { stream }) };
      const out = { keys: new DeviceBuffer(4 * n), indices: new DeviceBuffer(4 * n), count: new DeviceBuffer(4), stream };
      new SplatProjector().projectSort(table, cameraAt(x), { sortBits: 32 }, out);
      return { n, stream, out, table };
    });
    await Promise.all(runs.map((r) => r.stream.done()));
    for (const { n, stream, out, table } of runs) {
      const keys = out.keys.read(0, n, Uint32Array, { stream });
      const { ordering, visible } = SortKeys.sort({ keys, count: n, bits: 32 });
      const got = out.count.read(0, 1, Uint32Array, { stream })[0];
      if (got !== visible || visible < n / 2) return `count ${got}, visible ${visible}`;
      const m = firstMismatch(out.indices.read(0, visible, Uint32Array, { stream }), ordering, visible);
      [out.keys, out.indices, out.count, table.planes].forEach((b) => b.dispose());
      if (m) return m;
    }
    return undefined;
  });
  await check('sortSplatKeys: three promises in flight, each the order of SortKeys.sort', async () => {
    const random = rng(33);
    const sets = [20000, 50000, 3000].map((n) => new Uint32Array(n).map(() => (random() < 0.1 ? 0xffffff : Math.floor(random() * 0xfeffff))));
    const results = await Promise.all(sets.map((keys) => sortSplatKeys(keys.slice(), { bits: 24 })));
    for (let k = 0; k < sets.length; k += 1) {
      const here = SortKeys.sort({ keys: sets[k], count: sets[k].length, bits: 24 });
      if (results[k].visible !== here.visible) return `set ${k}: visible ${results[k].visible} vs ${here.visible}`;
      const m = firstMismatch(results[k].ordering, here.ordering, sets[k].length);
      if (m) return `set ${k}: ${m}`;
    }
    return undefined;
  });
}

async function gpuChecksBatch() {
  await check('SplatBatch: each view segment is the order of SortKeys.sort of its keys (graph and plain runs)', () => {
    const sizes = [3000, 5000];
    const tables = sizes.map((n, k) => ({ layout: 'f32', count: n, planes: DeviceBuffer.fromArray(randomPlanes(40 + k, n)) }));
    const views = [[0, 0.0], [1, 0.3], [0, -0.5]].map(([table, x]) => ({ table, camera: cameraAt(x) }));
    const total = views.reduce((s, v) => s + sizes[v.table], 0);
    const keys = new DeviceBuffer(4 * total);
    const batch = new SplatBatch(tables, views, { sortBits: 32 }, { keys });
    if (batch.total !== total) return `total ${batch.total} vs ${total}`;
    const indices = new DeviceBuffer(4 * total);
    const counts = new DeviceBuffer(4 * views.length);
    const drawArgs = new DeviceBuffer(20 * views.length);
    for (const graph of [true, false, true]) {
      batch.run({ indices, counts, drawArgs, graph });
      if (graph) batch.status();
      const k = keys.read(0, total);
      const idx = indices.read(0, total);
      const c = counts.read(0, views.length);
      const args = drawArgs.read(0, 5 * views.length);
      for (let v = 0; v < views.length; v += 1) {
        const off = batch.offsets[v];
        const n = sizes[views[v].table];
        const { ordering, visible } = SortKeys.sort({ keys: k.subarray(off, off + n), count: n, bits: 32 });
        if (c[v] !== visible || visible < n / 2) return `view ${v}: count ${c[v]}, visible ${visible}`;
        if (args[5 * v + 1] !== Math.ceil(visible / 1024)) return `view ${v}: draw arguments ${args.subarray(5 * v, 5 * v + 5)}`;
        const m = firstMismatch(idx.subarray(off, off + visible), ordering, visible);
        if (m) return `view ${v} (graph ${graph}): ${m}`;
      }
    }
    batch.dispose();
    [keys, indices, counts, drawArgs].concat(tables.map((t) => t.planes)).forEach((b) => b.dispose());
    return undefined;
  });

  await check('SplatBatch: the persistent launch equals the launch sequence (tuning [8, 12, 512], graph and eager)', () => {
    const sizes = [7000, 4000];
    const tables = sizes.map((n, k) => ({ layout: 'f32', count: n, planes: DeviceBuffer.fromArray(randomPlanes(50 + k, n)) }));
    const views = [[0, 0.1], [1, -0.2], [1, 0.4], [0, -0.3]].map(([table, x]) => ({ table, camera: cameraAt(x) }));
    const total = views.reduce((s, v) => s + sizes[v.table], 0);
    const keys = new DeviceBuffer(4 * total);
    const batch = new SplatBatch(tables, views, { sortBits: 32 }, { keys, tuning: [8, 12, 512] });
    const indices = new DeviceBuffer(4 * total);
    const counts = new DeviceBuffer(4 * views.length);
    const drawArgs = new DeviceBuffer(20 * views.length);
    batch.run({ indices, counts, drawArgs, graph: false });
    const refIdx = indices.read(0, total);
    const refCounts = counts.read(0, views.length);
    const refArgs = drawArgs.read(0, 5 * views.length);
    for (const graph of [false, true, true]) {
      indices.fill(0);
      counts.fill(0);
      batch.run({ indices, counts, drawArgs, graph, persistent: true });
      batch.status();
      const idx = indices.read(0, total);
      const c = counts.read(0, views.length);
      const args = drawArgs.read(0, 5 * views.length);
      for (let v = 0; v < views.length; v += 1) {
        if (c[v] !== refCounts[v]) return `view ${v} (graph ${graph}): count ${c[v]} vs ${refCounts[v]}`;
        const m = firstMismatch(idx.subarray(batch.offsets[v], batch.offsets[v] + c[v]),
          refIdx.subarray(batch.offsets[v], batch.offsets[v] + c[v]), c[v]);
        if (m) return `view ${v} (graph ${graph}): ${m}`;
        for (let k = 0; k < 5; k += 1) if (args[5 * v + k] !== refArgs[5 * v + k]) return `view ${v}: draw arguments`;
      }
    }
    const wrong = new SplatBatch(tables, views, { sortBits: 32 }, { keys, tuning: [8, 16, 512] });
    let threw = false;
    try {
      wrong.run({ indices, counts, drawArgs, persistent: true });
    } catch (e) {
      threw = /persistent form/.test(String(e.message));
    }
    wrong.dispose();
    if (!threw) return 'the persistent form ran under a tuning without it';
    batch.dispose();
    [keys, indices, counts, drawArgs].concat(tables.map((t) => t.planes)).forEach((b) => b.dispose());
    return undefined;
  });
}

function readF32(file) {
  const b = fs.readFileSync(file);
  return new Float32Array(new Uint8Array(b).buffer);
}

function readU32(file) {
  const b = fs.readFileSync(file);
  return new Uint32Array(new Uint8Array(b).buffer);
}

// The compositor's forward and adjoint, and the draw, on fixtures written by
// tests/test_napi.py (which compares what this writes with the FP64 oracle)
async function gpuFixture(dir) {
  await check('Rasterizer: forward and adjoint of the L2 loss on the fixture scene (on a stream)', () => {
    const spec = JSON.parse(fs.readFileSync(path.join(dir, 'scene.json'), 'utf8'));
    const theta = readF32(path.join(dir, 'theta.f32'));
    const target = readF32(path.join(dir, 'target.f32'));
    const { n, camera } = spec;
    const [width, height] = camera.size;
    const px = width * height;
    const stream = new GpuStream();
    const params = DeviceBuffer.fromArray(theta, { stream });
    const image = new DeviceBuffer(12 * px);
    const rast = new Rasterizer({ shDegree: 3, stream });
    const generation = rast.forward(params, n, camera, {}, { image, stream });
    const img = image.read(0, 3 * px, Float32Array, { stream });
    // L = mean((image - target)^2): dL/dimage = 2 (image - target) / (3 width height)
    const g = new Float32Array(3 * px);
    for (let i = 0; i < g.length; i += 1) g[i] = (2 * (img[i] - target[i])) / g.length;
    const gradImage = DeviceBuffer.fromArray(g, { stream });
    const gradParams = new DeviceBuffer(59 * 4 * n);
    rast.backward(gradImage, gradParams, { generation, stream });
    fs.writeFileSync(path.join(dir, 'image.f32'), Buffer.from(img.buffer));
    fs.writeFileSync(path.join(dir, 'grad.f32'), Buffer.from(gradParams.read(0, 59 * n, Float32Array, { stream }).buffer));
    const stats = rast.stats({ stream });
    if (!(stats.entries > 0 && stats.capacity >= stats.entries)) return `stats ${JSON.stringify(stats)}`;
    rast.forward(params, n, camera, {}, { image, stream });
    const stale = throwsWith(() => rast.backward(gradImage, gradParams, { generation, stream }), 'replaced its state');
    rast.dispose();
    [params, image, gradImage, gradParams].forEach((b) => b.dispose());
    return stale;
  });
  await check('SplatDrawer: the fixture frame drawn', () => {
    const spec = JSON.parse(fs.readFileSync(path.join(dir, 'draw.json'), 'utf8'));
    const words = DeviceBuffer.fromArray(readU32(path.join(dir, 'words.u32')));
    const order = DeviceBuffer.fromArray(readU32(path.join(dir, 'order.u32')));
    const count = DeviceBuffer.fromArray(new Uint32Array([spec.count]));
    const rgba = new DeviceBuffer(16 * spec.width * spec.height);
    const drawer = new SplatDrawer();
    drawer.draw(words, spec.n, order, count, spec.capacity, spec.width, spec.height, {}, rgba);
    fs.writeFileSync(path.join(dir, 'rgba.f32'), Buffer.from(rgba.read(0, 4 * spec.width * spec.height, Float32Array).buffer));
    const stats = drawer.stats();
    drawer.dispose();
    [words, order, count, rgba].forEach((b) => b.dispose());
    return stats.drawn === spec.count ? undefined : `drawn ${stats.drawn} of ${spec.count}`;
  });
  await check('SplatDrawer.drawRay: the fixture raygs frame, second eye of a stereo camera', () => {
    const spec = JSON.parse(fs.readFileSync(path.join(dir, 'ray.json'), 'utf8'));
    const channels = [0, 1, 2].map((k) => DeviceBuffer.fromArray(readU32(path.join(dir, `ray_w${k}.u32`))));
    const order = DeviceBuffer.fromArray(readU32(path.join(dir, 'ray_order.u32')));
    const count = DeviceBuffer.fromArray(new Uint32Array([spec.count]));
    const [width, height] = spec.camera.size;
    const rgba = new DeviceBuffer(16 * width * height);
    const drawer = new SplatDrawer();
    drawer.drawRay(spec.method, channels, spec.n, order, count, spec.count, spec.camera, {}, rgba, { eye: spec.eye });
    fs.writeFileSync(path.join(dir, 'ray_rgba.f32'), Buffer.from(rgba.read(0, 4 * width * height, Float32Array).buffer));
    const stats = drawer.stats();
    const unknown = throwsWith(() => drawer.drawRay('ewa', channels, spec.n, order, count, spec.count, spec.camera, {},
      rgba), 'unknown method');
    drawer.dispose();
    [...channels, order, count, rgba].forEach((b) => b.dispose());
    if (unknown) return unknown;
    return stats.drawn === spec.count ? undefined : `drawn ${stats.drawn} of ${spec.count}`;
  });
  await check('SplatProjector: projectSort of the fixture covSplats table (packedSh bundles)', () => {
    const spec = JSON.parse(fs.readFileSync(path.join(dir, 'cov.json'), 'utf8'));
    const n = spec.n;
    const covA = DeviceBuffer.fromArray(readU32(path.join(dir, 'covA.u32')));
    const covB = DeviceBuffer.fromArray(readU32(path.join(dir, 'covB.u32')));
    const shChannels = [0, 1, 2].map((t) => DeviceBuffer.fromArray(readU32(path.join(dir, `sh${t}.u32`))));
    const table = { layout: 'covSplats', count: n, covA, covB, shCodec: 'packed', shDegree: 3, shChannels };
    const words = [new DeviceBuffer(16 * n)];
    const keys = new DeviceBuffer(4 * n);
    const indices = new DeviceBuffer(4 * n);
    const count = new DeviceBuffer(4);
    const indirect = new DeviceBuffer(20);
    new SplatProjector().projectSort(table, spec.camera, { sortBits: 32 }, { words, keys, indices, count, indirect });
    const visible = count.read(0, 1, Uint32Array)[0];
    fs.writeFileSync(path.join(dir, 'cov_keys.u32'), Buffer.from(keys.read(0, n, Uint32Array).buffer));
    fs.writeFileSync(path.join(dir, 'cov_indices.u32'), Buffer.from(indices.read(0, visible, Uint32Array).buffer));
    fs.writeFileSync(path.join(dir, 'cov_words.u32'), Buffer.from(words[0].read(0, 4 * n, Uint32Array).buffer));
    [covA, covB, ...shChannels, ...words, keys, indices, count, indirect].forEach((b) => b.dispose());
    return visible > 0 ? undefined : 'nothing visible';
  });
}

(async () => {
  await cpuChecks();
  await cpuBoundsChecks();
  if (process.argv[2] === 'gpu') {
    await gpuChecks();
    await gpuChecksKeysAndRays();
    await gpuChecksSpz();
    await gpuBoundsChecks();
    await gpuChecksStreams();
    await gpuChecksBatch();
  } else if (process.argv[2] === 'fixture') {
    await gpuFixture(process.argv[3]);
  }
  console.log(failures.length ? `${failures.length} failed` : 'all passed');
  process.exit(failures.length ? 1 : 0);
})();

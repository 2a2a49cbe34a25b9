'use strict';
// Checks of the Node binding, written like the reference's own checks
// (test/gpu/renderTests.ts, scripts/check-core.ts): each check returns
// undefined or a failure message. `node check_napi.js cpu` runs the checks
// that make no GPU call; `gpu` adds the GPU sorts and projection.
const path = require('path');
const radix = require(path.join(__dirname, '..', '..', 'radix_amd', 'napi', 'index.js'));
const { SortKeys, sortSplatKeys, DeviceBuffer, RadixSort, SplatSorter, SplatProjector, alignCapacity, readSpz } = radix;
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

(async () => {
  await cpuChecks();
  if (process.argv[2] === 'gpu') {
    await gpuChecks();
    await gpuChecksKeysAndRays();
    await gpuChecksSpz();
  }
  console.log(failures.length ? `${failures.length} failed` : 'all passed');
  process.exit(failures.length ? 1 : 0);
})();

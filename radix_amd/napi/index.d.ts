// Types of the Node binding (index.js): the reference's sort and projection
// interfaces over libradix_hip.so, the compositor and its adjoint, the draw
// and the batched views.
export type SortKeyBits = 16 | 24 | 32;

/** A HIP stream with its own scratch: calls given it are ordered on it and overlap with other streams. */
export class GpuStream {
  constructor();
  synchronize(): void;
  done(): Promise<void>;
}

export interface OnStream {
  stream?: GpuStream;
}

export function alignCapacity(count: number): number;

export class SortKeys {
  static readonly KEY16_LO: number;
  static readonly KEY16_HI: number;
  static invalid(bits: SortKeyBits): number;
  static validMax(bits: SortKeyBits): number;
  static compute(distance: number, bits: SortKeyBits): number;
  static sort<O extends Uint32Array = Uint32Array>(args: { keys: Uint32Array; count?: number; bits: SortKeyBits; ordering?: O }): {
    ordering: O;
    visible: number;
  };
}

export interface SortedSplats {
  ordering: Uint32Array;
  visible: number;
  keys: Uint32Array;
}

export function sortSplatKeys(keys: Uint32Array, options: { bits: SortKeyBits; ordering?: Uint32Array }): Promise<SortedSplats>;

export class DeviceBuffer {
  readonly bytes: number;
  constructor(bytes: number);
  static fromArray(array: ArrayBufferView, options?: OnStream): DeviceBuffer;
  upload(array: ArrayBufferView, byteOffset?: number, options?: OnStream): void;
  read<T extends Uint32Array | Float32Array = Uint32Array>(start?: number, count?: number,
    Type?: { new (n: number): T }, options?: OnStream): T;
  fill(value: number, count?: number, options?: OnStream): void;
  dispose(): void;
}

export class RadixSort {
  readonly keys: DeviceBuffer;
  readonly values: DeviceBuffer;
  constructor(args: { keys: DeviceBuffer; values: DeviceBuffer });
  readonly capacity: number;
  sort(count: number, options?: { bits?: SortKeyBits } & OnStream): void;
  sortIndirect(countBuffer: DeviceBuffer, options?: { bits?: SortKeyBits } & OnStream): void;
  dispose(): void;
}

export class SplatSorter {
  static readonly QUADS_PER_INSTANCE: number;
  readonly values: DeviceBuffer;
  readonly count: DeviceBuffer;
  readonly indirect: DeviceBuffer;
  constructor(capacity?: number);
  sort(keys: DeviceBuffer, count: number, bits: SortKeyBits, options?: OnStream): void;
  readCount(options?: OnStream): Promise<number>;
  dispose(): void;
}

export interface SplatTable {
  layout: "f32" | "extSplats" | "packedSplats" | "covSplats";
  count: number;
  planeStride?: number;
  planes?: DeviceBuffer;
  extA?: DeviceBuffer;
  extB?: DeviceBuffer;
  covA?: DeviceBuffer;
  covB?: DeviceBuffer;
  packed?: DeviceBuffer;
  rgbMin?: number;
  rgbMax?: number;
  lnScaleMin?: number;
  lnScaleMax?: number;
  shCodec?: "none" | "f32" | "ext" | "packed";
  shDegree?: number;
  shChannels?: DeviceBuffer[];
  objectToWorld?: ArrayLike<number>;
}

export interface SplatCameraViews {
  head?: ArrayLike<number>;
  eyes: { view: ArrayLike<number>; projection: ArrayLike<number>; ortho?: boolean }[];
  size: [number, number];
  ndcZMin: number;
}

export interface SplatProjectionParams {
  maxStdDev?: number;
  minAlpha?: number;
  falloff?: number;
  lodInflate?: boolean;
  blurAmount?: number;
  preBlurAmount?: number;
  clipXY?: number;
  minPixelRadius?: number;
  maxPixelRadius?: number;
  minContribution?: number;
  focalDistance?: number;
  apertureTan?: number;
  sortGuard?: number;
  maxShDegree?: number;
  sortBits?: SortKeyBits;
  cullKeys?: boolean;
  keyPerEye?: boolean;
}

export class SplatProjector {
  project(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams,
    out: { words: DeviceBuffer[]; keys?: DeviceBuffer; at?: number } & OnStream): void;
  projectSort(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams,
    out: { words?: DeviceBuffer[]; keys?: DeviceBuffer; indices: DeviceBuffer; count?: DeviceBuffer; indirect?: DeviceBuffer } & OnStream): void;
  projectKeys(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams,
    out: { keys: DeviceBuffer; at?: number } & OnStream): void;
  projectRay(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams, method: '3dgut' | 'htgs' | 'raygs',
    out: { channels: DeviceBuffer[]; keys?: DeviceBuffer; at?: number } & OnStream): void;
}

/** The differentiable compositor: 59 parameter planes of `count` floats in, 3 image planes out; the adjoint back. */
export class Rasterizer {
  readonly generation: number;
  constructor(settings?: { shDegree?: number; tStop?: number; maxEntriesPerSplat?: number } & OnStream);
  /** Returns the forward's generation (backward of an older one throws). */
  forward(params: DeviceBuffer, count: number, camera: SplatCameraViews, renderParams: SplatProjectionParams,
    out: { image: DeviceBuffer; alpha?: DeviceBuffer } & OnStream): number;
  backward(gradImage: DeviceBuffer, gradParams: DeviceBuffer, options?: { generation?: number } & OnStream): void;
  stats(options?: OnStream): { visible: number; entries: number; capacity: number };
  dispose(): void;
}

/** The splat draw for native hosts: ewa words in a back-to-front order into premultiplied RGBA planes. */
export class SplatDrawer {
  constructor(settings?: { tStop?: number; linearOutput?: boolean; maxEntriesPerSplat?: number } & OnStream);
  draw(words: DeviceBuffer, n: number, order: DeviceBuffer, count: DeviceBuffer, capacity: number, width: number,
    height: number, params: SplatProjectionParams, rgba: DeviceBuffer, options?: OnStream): void;
  drawRay(method: '3dgut' | 'htgs' | 'raygs', channels: DeviceBuffer[], n: number, order: DeviceBuffer,
    count: DeviceBuffer, capacity: number, camera: SplatCameraViews, params: SplatProjectionParams, rgba: DeviceBuffer,
    options?: { eye?: number } & OnStream): void;
  stats(options?: OnStream): { drawn: number; entries: number; capacity: number };
  dispose(): void;
}

/** Many (view, collection) segments projected and sorted in one sequence. */
export class SplatBatch {
  readonly total: number;
  readonly offsets: number[];
  constructor(tables: SplatTable[], views: { table: number; camera: SplatCameraViews; words?: DeviceBuffer }[],
    params: SplatProjectionParams, options?: { keys?: DeviceBuffer; tuning?: [number, number, number] } & OnStream);
  setCameras(cameras: SplatCameraViews[]): void;
  run(out: { indices: DeviceBuffer; counts?: DeviceBuffer; drawArgs?: DeviceBuffer; graph?: boolean;
    persistent?: boolean } & OnStream): void;
  status(): void;
  dispose(): void;
}

export function untransposeBytes(planes: DeviceBuffer, count: number, elementBytes: number, out: DeviceBuffer,
  outOffset?: number, options?: OnStream): void;

export interface SpzMeta {
  format: 'spz'; version: number; count: number; shDegree: number; fractionalBits: number; antialiased: boolean;
}
export function readSpz(bytes: Uint8Array): {
  table: { layout: 'f32'; count: number; planes: DeviceBuffer; planeStride: number; shCodec: 'f32' | 'none'; shDegree: number };
  meta: SpzMeta;
};

export function version(): string;

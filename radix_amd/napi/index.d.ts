// Types of the Node binding (index.js): the reference's sort and projection
// interfaces over libradix_hip.so.
export type SortKeyBits = 16 | 24 | 32;

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
  static fromArray(array: ArrayBufferView): DeviceBuffer;
  upload(array: ArrayBufferView, byteOffset?: number): void;
  read(start?: number, count?: number): Uint32Array;
  fill(value: number, count?: number): void;
  dispose(): void;
}

export class RadixSort {
  readonly keys: DeviceBuffer;
  readonly values: DeviceBuffer;
  constructor(args: { keys: DeviceBuffer; values: DeviceBuffer });
  readonly capacity: number;
  sort(count: number, options?: { bits?: SortKeyBits }): void;
  sortIndirect(countBuffer: DeviceBuffer, options?: { bits?: SortKeyBits }): void;
  dispose(): void;
}

export class SplatSorter {
  static readonly QUADS_PER_INSTANCE: number;
  readonly values: DeviceBuffer;
  readonly count: DeviceBuffer;
  readonly indirect: DeviceBuffer;
  constructor(capacity?: number);
  sort(keys: DeviceBuffer, count: number, bits: SortKeyBits): void;
  readCount(): Promise<number>;
  dispose(): void;
}

export interface SplatTable {
  layout: "f32" | "extSplats" | "packedSplats";
  count: number;
  planeStride?: number;
  planes?: DeviceBuffer;
  extA?: DeviceBuffer;
  extB?: DeviceBuffer;
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
    out: { words: DeviceBuffer[]; keys?: DeviceBuffer; at?: number }): void;
  projectSort(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams,
    out: { words?: DeviceBuffer[]; keys?: DeviceBuffer; indices: DeviceBuffer; count?: DeviceBuffer; indirect?: DeviceBuffer }): void;
  projectKeys(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams,
    out: { keys: DeviceBuffer; at?: number }): void;
  projectRay(table: SplatTable, camera: SplatCameraViews, params: SplatProjectionParams, method: '3dgut' | 'htgs' | 'raygs',
    out: { channels: DeviceBuffer[]; keys?: DeviceBuffer; at?: number }): void;
}

export function untransposeBytes(planes: DeviceBuffer, count: number, elementBytes: number, out: DeviceBuffer,
  outOffset?: number): void;

export interface SpzMeta {
  format: 'spz'; version: number; count: number; shDegree: number; fractionalBits: number; antialiased: boolean;
}
export function readSpz(bytes: Uint8Array): {
  table: { layout: 'f32'; count: number; planes: DeviceBuffer; planeStride: number; shCodec: 'f32' | 'none'; shDegree: number };
  meta: SpzMeta;
};

export function version(): string;

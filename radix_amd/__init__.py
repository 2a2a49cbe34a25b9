"""radix_amd: the per-frame splat pipeline of sparkjsdev/radix on MI355X (gfx950).

EWA projection with SH colour, back-to-front sort keys and a stable LSD
radix sort of splat indices, as hand-written HIP kernels behind the C ABI of
include/radix_hip.h (libradix_hip.so), with host mirrors of the reference's
interfaces: SortKeys, RadixSort (WgpuRadixSort), SplatSorter
(WgpuSplatSorter), sort_splat_keys (sortSplatKeys) and Projector
(SplatProjection.write + the GPU sort of SplatRenderer.update), and Batch,
many views over many tables projected and sorted as segments of one
sequence.
"""
from .sort import Context, RadixSort, SortKeys, SplatSorter, align_capacity, default_context, sort_splat_keys
from .projection import DeviceTable, Params, Projector
from .batch import Batch

__all__ = [
    "Context", "RadixSort", "SortKeys", "SplatSorter", "align_capacity", "default_context", "sort_splat_keys",
    "DeviceTable", "Params", "Projector", "Batch",
]

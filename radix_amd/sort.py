"""Sort keys and the GPU radix sort, mirroring the reference's sort API.

SortKeys      src/core/sort.ts (compute, invalid, validMax, sort)
RadixSort     src/wgpu/WgpuRadixSort.ts (sort, sortIndirect), device tensors
SplatSorter   src/wgpu/WgpuSplatSorter.ts (compaction + sort + draw arguments)
sort_splat_keys  src/worker/splatSort.ts (sortSplatKeys)

Every sort runs the HIP kernels of libradix_hip.so; host arrays are copied
to the device and back. Device arrays are torch tensors of dtype int32
holding the uint32 bits.
"""
from __future__ import annotations

import ctypes as C
import threading
from typing import Optional, Tuple

import numpy as np

from . import _native as N

ROW_ELEMENTS = 1024 * 4  # TextureLayout.ROW_ELEMENTS (src/core/layout.ts:28)


def align_capacity(count: int) -> int:
    """TextureLayout.alignCapacity (src/core/layout.ts:70-72)."""
    return max(1, -(-count // ROW_ELEMENTS)) * ROW_ELEMENTS


def on_stream(fn):
    """Runs a method with a `stream` argument under that stream, when one is
    given: its allocations, fills and launches are then ordered on the
    caller's stream rather than split between it and torch's current one."""
    import functools
    import inspect

    # the position of `stream` among the parameters, found once (binding the
    # signature on every call cost about 10 us of host time per launch)
    names = list(inspect.signature(fn).parameters)
    at = names.index("stream")

    @functools.wraps(fn)
    def run(*args, **kwargs):
        stream = kwargs.get("stream", args[at] if len(args) > at else None)
        if stream is None:
            return fn(*args, **kwargs)
        import torch

        with torch.cuda.stream(stream):
            return fn(*args, **kwargs)

    return run


class Context:
    """A native context: the device scratch of sorts on one stream."""

    def __init__(self, tuning: Optional[Tuple[int, int, int]] = None):
        lib = N.lib()
        h = C.c_void_p()
        N.check(lib.rdx_context_create(C.byref(h)))
        self._h = h
        self._lib = lib
        if tuning is not None:
            self.set_tuning(*tuning)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def set_tuning(self, digit_bits: int = 8, keys_per_thread: int = 16, block_threads: int = 512) -> None:
        t = N.SortTuning(digit_bits, keys_per_thread, block_threads)
        N.check(self._lib.rdx_context_set_tuning(self._h, C.byref(t)))

    def tuning(self) -> Tuple[int, int, int]:
        t = N.SortTuning()
        N.check(self._lib.rdx_context_get_tuning(self._h, C.byref(t)))
        return (t.digit_bits, t.keys_per_thread, t.block_threads)

    @on_stream
    def autotune(self, count: int, bits: int = 32, keys=None, reps: int = 5, stream=None,
                 cache: Optional[str] = None) -> Tuple[Tuple[int, int, int], float]:
        """Time every supported tuning on `count` keys (`keys`: a device
        tensor, or None for uniform random keys) and keep the fastest.
        Returns ((digit_bits, keys_per_thread, block_threads), ms per sort).
        cache: a JSON file of earlier choices per (device, bits, count
        rounded up to a power of two), read before and written after."""
        import json
        import os

        import torch

        slot = f"{torch.cuda.get_device_name()}|{bits}|{1 << max(0, (int(count) - 1).bit_length())}"
        if cache and keys is None and os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
            if slot in known:
                t, ms = known[slot]
                self.set_tuning(*t)
                return tuple(t), float(ms)
        out = N.SortTuning()
        ms = C.c_double()
        N.check(self._lib.rdx_context_autotune(self._h, _stream(stream), _ptr(keys), count, bits, reps,
                                               C.byref(out), C.byref(ms)))
        chosen = (out.digit_bits, out.keys_per_thread, out.block_threads)
        if cache and keys is None:
            known = {}
            if os.path.exists(cache):
                with open(cache) as f:
                    known = json.load(f)
            known[slot] = [list(chosen), ms.value]
            with open(cache, "w") as f:
                json.dump(known, f, indent=1, sort_keys=True)
        return chosen, ms.value

    @staticmethod
    def candidates():
        """The tunings the library supports, in autotuner order."""
        lib = N.lib()
        out, i = [], 0
        t = N.SortTuning()
        while lib.rdx_sort_tuning_candidate(i, C.byref(t)) == N.RDX_OK:
            out.append((t.digit_bits, t.keys_per_thread, t.block_threads))
            i += 1
        return out

    def reserve(self, max_keys: int) -> None:
        N.check(self._lib.rdx_context_reserve(self._h, max_keys))

    def close(self) -> None:
        if self._h:
            self._lib.rdx_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_local = threading.local()


def default_context() -> Context:
    """One context per thread (the reference's sort scratch is per thread)."""
    ctx = getattr(_local, "ctx", None)
    if ctx is None:
        ctx = Context()
        _local.ctx = ctx
    return ctx


def _stream(stream=None):
    """The HIP stream to launch on: `stream`, else torch's current stream
    (read as its raw handle where torch exposes it, without building a
    Stream object)."""
    if stream is not None:
        return C.c_void_p(stream.cuda_stream)
    import torch

    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return C.c_void_p(raw(torch._C._cuda_getDevice()))
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p()


class SortKeys:
    """src/core/sort.ts: back-to-front keys of distances and their stable sort."""

    KEY16_LO = 95 << 23
    KEY16_HI = (159 << 23) - 1

    @staticmethod
    def invalid(bits: int) -> int:
        return int(N.lib().rdx_key_invalid(bits))

    @staticmethod
    def valid_max(bits: int) -> int:
        return int(N.lib().rdx_key_valid_max(bits))

    validMax = valid_max

    @staticmethod
    def compute(distance: float, bits: int) -> int:
        """The key of one distance (rounded to float32 first, as a Float32Array slot does)."""
        return int(N.lib().rdx_key_compute(C.c_float(distance), bits))

    @staticmethod
    @on_stream
    def compute_device(distances, bits: int, out=None, stream=None):
        """Keys of a float32 device tensor of distances (GpuSortKeys.compute)."""
        import torch

        if distances.dtype != torch.float32 or not distances.is_cuda:
            raise TypeError("distances must be a float32 device tensor")
        d = distances.contiguous()
        out = out if out is not None else torch.empty(d.numel(), dtype=torch.int32, device=d.device)
        N.check(N.lib().rdx_compute_keys(_stream(stream), _ptr(d), d.numel(), bits, _ptr(out)))
        return out

    @staticmethod
    def sort(keys: np.ndarray, count: Optional[int] = None, bits: int = 24, ordering: Optional[np.ndarray] = None,
             context: Optional[Context] = None) -> Tuple[np.ndarray, int]:
        """SortKeys.sort({keys, count, bits, ordering}) -> (ordering, visible).

        Stable ascending order of `count` host keys on the GPU; the default
        `ordering` is a whole capacity (alignCapacity(count)) of uint32."""
        keys = np.ascontiguousarray(keys, dtype=np.uint32)
        if count is None:
            count = keys.size
        if count > keys.size:
            raise ValueError(f"count {count} above the {keys.size} keys given")
        if ordering is None:
            ordering = np.zeros(align_capacity(count), dtype=np.uint32)
        if ordering.dtype != np.uint32 or not ordering.flags.c_contiguous:
            raise TypeError("ordering must be a contiguous uint32 array")
        ctx = context or default_context()
        visible = C.c_uint32(0)
        N.check(N.lib().rdx_sort_keys(ctx.handle, keys.ctypes.data_as(C.c_void_p), count, bits,
                                      ordering.ctypes.data_as(C.c_void_p), ordering.size, C.byref(visible)))
        return ordering, int(visible.value)


def sort_splat_keys(keys: np.ndarray, bits: int, ordering: Optional[np.ndarray] = None):
    """sortSplatKeys (src/worker/splatSort.ts:17-30): {ordering, visible, keys}.

    `ordering` is reused when it holds a whole capacity for the keys."""
    count = keys.size
    reuse = ordering if ordering is not None and ordering.size >= align_capacity(count) else None
    ordering, visible = SortKeys.sort(keys, count=count, bits=bits, ordering=reuse)
    return {"ordering": ordering, "visible": visible, "keys": keys}


class RadixSort:
    """WgpuRadixSort: stable LSD sort of device keys and values, in place."""

    def __init__(self, keys, values, context: Optional[Context] = None, keys_per_thread: Optional[int] = None,
                 digit_bits: Optional[int] = None, block_threads: Optional[int] = None):
        """keys_per_thread (WgpuRadixSort's keysPerThread), digit_bits and
        block_threads override the context's tuning where given."""
        import torch

        for name, t in (("keys", keys), ("values", values)):
            if t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous int32 device tensor (uint32 bits)")
        self.keys = keys
        self.values = values
        self.context = context or Context()
        if keys_per_thread is not None or digit_bits is not None or block_threads is not None:
            rb, kpt, bt = self.context.tuning()
            self.context.set_tuning(digit_bits or rb, keys_per_thread or kpt, block_threads or bt)

    @property
    def capacity(self) -> int:
        return min(self.keys.numel(), self.values.numel())

    @on_stream
    def sort(self, count: int, bits: int = 32, stream=None) -> None:
        if count <= 0:
            return
        if count > self.capacity:
            raise ValueError(f"WgpuRadixSort: {count} keys above the capacity {self.capacity}")
        N.check(N.lib().rdx_sort_pairs(self.context.handle, _stream(stream), _ptr(self.keys), _ptr(self.values),
                                       count, bits))

    @on_stream
    def sort_indirect(self, count_buffer, bits: int = 32, stream=None) -> None:
        N.check(N.lib().rdx_sort_pairs_indirect(self.context.handle, _stream(stream), _ptr(self.keys),
                                                _ptr(self.values), _ptr(count_buffer), self.capacity, bits))

    sortIndirect = sort_indirect


class SplatSorter:
    """WgpuSplatSorter: drop invalid keys, sort back to front, draw arguments."""

    QUADS_PER_INSTANCE = 1024

    def __init__(self, capacity: int = 0, context: Optional[Context] = None, device=None):
        import torch

        self.device = device or torch.device("cuda")
        self.context = context or Context()
        self.values = torch.empty(max(capacity, 1), dtype=torch.int32, device=self.device)
        self.count = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.indirect = torch.zeros(5, dtype=torch.int32, device=self.device)

    @on_stream
    def sort(self, keys, count: int, bits: int, stream=None) -> None:
        import torch

        if self.values.numel() < max(count, 1):
            self.values = torch.empty(count, dtype=torch.int32, device=self.device)
        N.check(N.lib().rdx_sort_splats(self.context.handle, _stream(stream), _ptr(keys), count, bits,
                                        _ptr(self.values), _ptr(self.count), _ptr(self.indirect)))

    def read_count(self) -> int:
        return int(self.count.cpu().item())

    readCount = read_count

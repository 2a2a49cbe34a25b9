"""RDX files decoded onto the device (src/core/rdx.ts).

An RDX file is a header container ("RDXH": the count, chunk size, bundles
and chunk sizes as JSON) followed by chunk containers ("RDXC": {start,
count} and one payload per serialized channel, in header order). Each
payload is the channel's elements of the chunk, byte-transposed and gzipped
(rdx.ts:8-17, 40-118). Here the host reads the containers and gunzips the
payloads (zlib, several threads), and the device undoes the transpose
straight into the channel's buffer (rdx_untranspose_bytes), so a file
becomes a splat table in HBM without a CPU pass over its elements.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import json
import os
import re
import struct
import zlib
from concurrent.futures import ThreadPoolExecutor
from typing import Dict, Optional, Tuple

from . import _native as N
from .sort import _stream

RDX_MAGIC = 0x48584452  # "RDXH"
CHUNK_MAGIC = 0x43584452  # "RDXC"
RDX_VERSION = 1

_SCALAR_BYTES = {"uint8": 1, "sint8": 1, "unorm8": 1, "snorm8": 1, "uint16": 2, "sint16": 2, "float16": 2,
                 "uint32": 4, "sint32": 4, "float32": 4}


def format_bytes(fmt: str) -> int:
    """Bytes per element of an array format ("uint32x4" -> 16; formats.ts:43-92)."""
    m = re.fullmatch(r"([a-z]+[0-9]+)(?:x([1-4]))?", fmt)
    if not m or m.group(1) not in _SCALAR_BYTES:
        raise ValueError(f"Unknown array format {fmt}")
    return _SCALAR_BYTES[m.group(1)] * int(m.group(2) or 1)


def _align16(n: int) -> int:
    return (n + 15) // 16 * 16


def decode_container(buf, magic: int):
    """(meta, payload memoryviews, byte length) of one container (rdx.ts:62-81)."""
    mv = memoryview(buf)
    if len(mv) < 16 or struct.unpack_from("<I", mv, 0)[0] != magic:
        raise ValueError("Not an RDX piece of the expected type")
    meta_len, count = struct.unpack_from("<II", mv, 4)
    meta = json.loads(bytes(mv[16:16 + meta_len]).decode())
    sizes_at = 16 + _align16(meta_len)
    offset = sizes_at + _align16(count * 4)
    payloads = []
    for i in range(count):
        size = struct.unpack_from("<I", mv, sizes_at + 4 * i)[0]
        payloads.append(mv[offset:offset + size])
        offset += _align16(size)
    return meta, payloads, offset


def decode_header(buf) -> Tuple[dict, int]:
    """The header at the start of a whole file (rdx.ts:156-164)."""
    meta, _, length = decode_container(buf, RDX_MAGIC)
    if meta.get("version") != RDX_VERSION:
        raise ValueError(f"RDX version {meta.get('version')} is not {RDX_VERSION}")
    return meta, length


class RdxDeviceData:
    """An ArrayData whose channels live in device memory: bundles[name] =
    {"meta": ..., "channels": {label: (format, uint8 tensor of count x bytes)}}."""

    def __init__(self, header: dict, device):
        import torch

        self.header = header
        self.count = int(header["count"])
        self.meta = header.get("meta", {})
        self.bundles: Dict[str, dict] = {}
        for name, b in header["bundles"].items():
            chans = {}
            for label, fmt in b["def"]["channels"].items():
                eb = format_bytes(fmt)
                chans[label] = (fmt, torch.zeros(max(self.count, 1) * eb, dtype=torch.uint8, device=device))
            self.bundles[name] = {"meta": b.get("meta", {}), "channels": chans}

    def words(self, bundle: str, label: str):
        """A uint32xN channel as an (n, N) int32 tensor view."""
        fmt, t = self.bundles[bundle]["channels"][label]
        return t.view(-1).view(dtype=__import__("torch").int32).view(max(self.count, 1), format_bytes(fmt) // 4)[: self.count]


def decode_chunk(header: dict, buf, into: RdxDeviceData, stream=None, pool: Optional[ThreadPoolExecutor] = None):
    """Writes a chunk's elements [start, start + count) into `into` (rdx.ts:176-199)."""
    import torch

    meta, payloads, _ = decode_container(buf, CHUNK_MAGIC)
    start, count = int(meta["start"]), int(meta["count"])
    if start + count > into.count:
        raise ValueError(f"Chunk at {start} holds {count} elements past the header's {into.count}")
    order = [(name, label, fmt) for name, b in header["bundles"].items() for label, fmt in b["def"]["channels"].items()]
    if len(payloads) != len(order):
        raise ValueError(f"Chunk at {start}: {len(payloads)} payloads for {len(order)} channels")
    unzip = lambda p: zlib.decompress(bytes(p), 31)  # gzip
    planes = list(pool.map(unzip, payloads)) if pool else [unzip(p) for p in payloads]
    dev = next(iter(next(iter(into.bundles.values()))["channels"].values()))[1].device
    staged = []  # the uploaded planes, alive until the transposes have run
    for (name, label, fmt), raw in zip(order, planes):
        eb = format_bytes(fmt)
        if len(raw) != count * eb:
            raise ValueError(f"Chunk at {start}: {name}.{label} holds {len(raw)} bytes, expected {count * eb}")
        if count == 0:
            continue
        # the upload and the transpose on one stream (the caller's, if given)
        with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            src = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev, non_blocking=False)
            staged.append(src)
            _, out = into.bundles[name]["channels"][label]
            N.check(N.lib().rdx_untranspose_bytes(_stream(stream), C.c_void_p(src.data_ptr()), count, eb,
                                                  C.c_void_p(out.data_ptr() + start * eb)))
    if staged:
        torch.cuda.synchronize(dev)


def read_rdx(source, device="cuda", stream=None, threads: int = 8) -> RdxDeviceData:
    """A whole RDX file (bytes or a path) decoded onto the device (RdxFile.decode, rdx.ts:201-213)."""
    import torch

    if isinstance(source, (str, os.PathLike)):
        with open(source, "rb") as f:
            source = f.read()
    header, offset = decode_header(source)
    data = RdxDeviceData(header, device)
    mv = memoryview(source)
    with ThreadPoolExecutor(max_workers=threads) as pool:
        for size in header["chunks"]:
            decode_chunk(header, mv[offset:offset + size], data, stream, pool)
            offset += size
    torch.cuda.synchronize()
    return data


def splat_table(data: RdxDeviceData, object_to_world=None):
    """The splat table of an RDX file's extSplats or packedSplats bundle and
    its extSh / packedSh bundle (splats.ts:69-97, 335-423; sh.ts:225-240),
    over the decoded device channels (no copy)."""
    from .projection import DeviceTable

    b = data.bundles
    sh_name = "extSh" if "extSh" in b else ("packedSh" if "packedSh" in b else None)
    tensors = {}
    codec, degree = N.SH_NONE, 0
    if sh_name:
        degree = int(b[sh_name]["meta"].get("degree", 0))
        labels = sorted(b[sh_name]["channels"])  # <codec>ShA, <codec>ShB, ...
        if degree > 0:
            tensors["sh"] = [data.words(sh_name, label) for label in labels]
            codec = N.SH_EXT if sh_name == "extSh" else N.SH_PACKED
    if "extSplats" in b:
        tensors = {"ext_a": data.words("extSplats", "extA"), "ext_b": data.words("extSplats", "extB"), **tensors}
        tensors["rdx"] = data  # keeps the decoded channels alive
        return DeviceTable(N.LAYOUT_EXT, data.count, tensors, sh_codec=codec, sh_degree=degree,
                           object_to_world=object_to_world)
    if "packedSplats" in b:
        enc = {"rgbMin": 0.0, "rgbMax": 1.0, "lnScaleMin": -12.0, "lnScaleMax": 9.0}
        enc.update(b["packedSplats"]["meta"].get("encoding", {}))
        tensors = {"packed": data.words("packedSplats", "packed"), **tensors}
        tensors["rdx"] = data
        return DeviceTable(N.LAYOUT_PACKED, data.count, tensors, sh_codec=codec, sh_degree=degree,
                           packed_encoding=(enc["rgbMin"], enc["rgbMax"], enc["lnScaleMin"], enc["lnScaleMax"]),
                           object_to_world=object_to_world)
    raise ValueError("No extSplats or packedSplats bundle")

"""SPZ files decoded onto the device (Niantic's format in its gzip container,
src/core/spz.ts): the host unzips the file and reads its 16-byte header
(magic "NGSP", version, count, SH degree, fractional bits, flags), one upload
puts the columns in HBM, and rdx_spz_decode writes the f32 splat table."""
from __future__ import annotations

import contextlib
import ctypes as C
import gzip
import os
import struct

from . import _native as N
from .sort import _stream

SPZ_MAGIC = 0x5053474E  # "NGSP"
SH_VECTORS = {0: 0, 1: 3, 2: 8, 3: 15, 4: 24}


def unpack(data: bytes):
    """(header dict, body bytes) of a gzip-container SPZ (spz.ts:140-160)."""
    if data[:4] == b"NGSP":
        raise ValueError("SPZ 4's NGSP (zstd) container is not supported here")
    raw = gzip.decompress(data)
    if len(raw) < 16 or struct.unpack_from("<I", raw, 0)[0] != SPZ_MAGIC:
        raise ValueError("Invalid SPZ file")
    version, count = struct.unpack_from("<II", raw, 4)
    sh_degree, fractional_bits, flags = raw[12], raw[13], raw[14]
    if version < 1 or version > 4 or sh_degree not in SH_VECTORS:
        raise ValueError(f"Unsupported SPZ version {version} or SH degree {sh_degree}")
    return {"version": version, "count": count, "sh_degree": sh_degree, "fractional_bits": fractional_bits,
            "flags": flags, "antialiased": bool(flags & 1)}, raw[16:]


def read_spz(source, device="cuda", stream=None, object_to_world=None):
    """A DeviceTable (f32 planes) of an SPZ file's splats (readSpz, spz.ts:260-349)."""
    import torch

    from .projection import DeviceTable

    if isinstance(source, (str, os.PathLike)):
        with open(source, "rb") as f:
            source = f.read()
    head, body = unpack(source)
    n, deg = head["count"], head["sh_degree"]
    if deg > 3:
        raise ValueError(f"Unsupported SPZ version {head['version']} or SH degree {deg}")
    # the upload and the decode on one stream (the caller's, if given)
    with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
        # (an empty file still gets a one-byte buffer: frombuffer refuses none)
        dev_body = torch.frombuffer(bytearray(body) or bytearray(1), dtype=torch.uint8).to(device)
        planes = torch.zeros((14 + 3 * SH_VECTORS[deg], max(n, 1)), dtype=torch.float32, device=device)
        N.check(N.lib().rdx_spz_decode(_stream(stream), C.c_void_p(dev_body.data_ptr()), len(body), n, head["version"],
                                       deg, head["fractional_bits"], C.c_void_p(planes.data_ptr()), planes.shape[1]))
    torch.cuda.synchronize()
    t = DeviceTable.from_f32_planes(planes, n, deg, device=device, object_to_world=object_to_world)
    t.meta = {"format": "spz", **head}
    return t

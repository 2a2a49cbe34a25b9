"""An RDX writer for the tests' fixtures, restating the container of
src/core/rdx.ts:40-118: each piece is u32 magic, u32 metaLength, u32
payloadCount, u32 0, JSON meta, u32 payloadSize[payloadCount], payloads,
each section 16-byte aligned; a chunk's payload per channel is its elements
byte-transposed (compress.ts:6-19) and gzipped."""
import gzip
import json
import struct

import numpy as np

RDX_MAGIC = 0x48584452
CHUNK_MAGIC = 0x43584452


def align16(n):
    return (n + 15) // 16 * 16


def container(magic, meta, payloads):
    mb = json.dumps(meta).encode()
    sizes_at = 16 + align16(len(mb))
    payloads_at = sizes_at + align16(len(payloads) * 4)
    total = payloads_at + sum(align16(len(p)) for p in payloads)
    out = bytearray(total)
    struct.pack_into("<III", out, 0, magic, len(mb), len(payloads))
    out[16:16 + len(mb)] = mb
    off = payloads_at
    for i, p in enumerate(payloads):
        struct.pack_into("<I", out, sizes_at + 4 * i, len(p))
        out[off:off + len(p)] = p
        off += align16(len(p))
    return bytes(out)


def transpose_bytes(data: bytes, stride: int) -> bytes:
    """transposeBytes (compress.ts:6-19), byte by byte as the reference loops."""
    assert len(data) % stride == 0
    out = bytearray(len(data))
    o = 0
    for offset in range(stride):
        for i in range(offset, len(data), stride):
            out[o] = data[i]
            o += 1
    return bytes(out)


def encode_rdx(bundles, count, chunk_size=65536, meta=None, version=1):
    """bundles: {name: {"channels": {label: fmt}, "meta": {...}, "data": {label: ndarray (count, ...)}}}."""
    from radix_amd.rdx import format_bytes

    assert chunk_size % 4096 == 0
    chunks = []
    for start in range(0, count, chunk_size):
        n = min(chunk_size, count - start)
        payloads = []
        for b in bundles.values():
            for label, fmt in b["channels"].items():
                eb = format_bytes(fmt)
                raw = np.ascontiguousarray(b["data"][label]).view(np.uint8).reshape(-1)[start * eb:(start + n) * eb]
                # numpy's transpose of the (n, eb) byte matrix: the same bytes as transposeBytes
                payloads.append(gzip.compress(np.ascontiguousarray(raw.reshape(n, eb).T).tobytes(), compresslevel=1))
        chunks.append(container(CHUNK_MAGIC, {"start": start, "count": n}, payloads))
    header = {"version": version, "count": count, "chunkSize": chunk_size, "meta": meta or {},
              "bundles": {name: {"def": {"channels": b["channels"], "serialize": True}, "meta": b.get("meta", {})}
                          for name, b in bundles.items()},
              "chunks": [len(c) for c in chunks]}
    return container(RDX_MAGIC, header, []) + b"".join(chunks)

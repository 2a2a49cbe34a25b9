"""An SPZ writer (gzip container, versions 1-3) for the tests' fixtures,
restating the reference's writer (src/core/spz.ts:380-480; version 1's
half-float centres, which it does not write, as readSpz reads them, 302-305): a 16-byte header
(magic "NGSP", version, count, SH degree, fractional bits, flags), then the
columns; rotations by encodeSmallestThree (549-569) in versions 3+."""
import gzip
import math
import struct

import numpy as np

from oracle.spz import SH_C0, SH_VECTORS


def encode_smallest_three(q):
    n = math.sqrt(sum(x * x for x in q))
    q = [x / n for x in q] if n > 0 else [0.0, 0.0, 0.0, 1.0]
    largest = 0  # the first index of the largest magnitude (a strict > scan)
    for i in range(1, 4):
        if abs(q[i]) > abs(q[largest]):
            largest = i
    negate = 1 if q[largest] < 0 else 0
    packed = largest
    for i in range(4):
        if i != largest:
            sign = (1 if q[i] < 0 else 0) ^ negate
            mag = max(0, min(511, int(math.floor(511 * abs(q[i]) / math.sqrt(0.5) + 0.5))))
            packed = (packed << 10) | (sign << 9) | mag
    return packed & 0xFFFFFFFF


def byte(x):
    return int(max(0, min(255, math.floor(x + 0.5))))


def encode_spz(centers, log_scales, quats, opacity, rgb, sh=None, sh_degree=0, version=3, fractional_bits=12, flags=1):
    n = len(opacity)
    k = SH_VECTORS[sh_degree]
    out = bytearray(struct.pack("<IIIBBBB", 0x5053474E, version, n, sh_degree, fractional_bits, flags, 0))
    if version == 1:
        out += np.asarray(centers, dtype=np.float16).tobytes()
    else:
        fixed = 1 << fractional_bits
        for c in centers:
            for x in c:
                v = max(-0x7FFFFF, min(0x7FFFFF, int(math.floor(x * fixed + 0.5)))) & 0xFFFFFF
                out += bytes([v & 255, (v >> 8) & 255, (v >> 16) & 255])
    out += bytes(byte(o * 255) for o in opacity)
    out += bytes(byte(((c - 0.5) / (SH_C0 / 0.15) + 0.5) * 255) for row in rgb for c in row)
    out += bytes(byte((s + 10) * 16) for row in log_scales for s in row)
    for q in quats:
        if version >= 3:
            out += struct.pack("<I", encode_smallest_three(q))
        else:
            nq = math.sqrt(sum(x * x for x in q))
            q = [x / nq for x in q]
            if q[3] < 0:
                q = [-x for x in q]
            out += bytes(byte((x + 1) * 127.5) for x in q[:3])
    if k:
        out += bytes(byte(v * 128 + 128) for row in sh for v in row[: 3 * k])
    return gzip.compress(bytes(out), compresslevel=1)

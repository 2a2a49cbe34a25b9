"""The oracle's SPZ decode: readSpz's per-splat values (src/core/spz.ts:297-347)
and decodeSmallestThree (571-586), restated in numpy over the unzipped body
after the 16-byte header. TEST INFRASTRUCTURE ONLY: the tests' checker of
rdx_spz_decode; the product never imports it.

Pinned by tests/test_spz.py: int24 sign extension and half-float centres,
the smallest-three codec's round trip through a restatement of the
reference's encoder (spz.ts:549-569) within its 9-bit step, the colour and
scale maps' closed forms at their end points, and SH bytes' exact values."""
import numpy as np

SH_C0 = 0.28209479177387814  # ply.ts:15
SH_VECTORS = {0: 0, 1: 3, 2: 8, 3: 15, 4: 24}


def columns(count, version, sh_degree):
    """Byte offsets of the columns (spz.ts:152-160)."""
    off = {"centers": 0}
    off["alphas"] = off["centers"] + count * (6 if version == 1 else 9)
    off["colors"] = off["alphas"] + count
    off["scales"] = off["colors"] + count * 3
    off["rotations"] = off["scales"] + count * 3
    off["sh"] = off["rotations"] + count * (4 if version >= 3 else 3)
    off["end"] = off["sh"] + count * SH_VECTORS[sh_degree] * 3
    return off


def decode_smallest_three(packed: int):
    """decodeSmallestThree (spz.ts:571-586), in double."""
    largest = packed >> 30
    q = [0.0, 0.0, 0.0, 0.0]
    rest = packed
    s = 0.0
    for i in range(3, -1, -1):
        if i != largest:
            v = np.sqrt(0.5) * ((rest & 511) / 511) * (-1.0 if (rest >> 9) & 1 else 1.0)
            rest >>= 10
            q[i] = v
            s += v * v
    q[largest] = np.sqrt(max(1.0 - s, 0.0))
    return q


def decode(body: bytes, count: int, version: int, sh_degree: int, fractional_bits: int) -> np.ndarray:
    """The f32 planes (14 + 3K, count): centre, log-scale, quaternion, opacity, rgb, SH."""
    b = np.frombuffer(body, dtype=np.uint8)
    c = columns(count, version, sh_degree)
    k = SH_VECTORS[sh_degree]
    out = np.zeros((14 + 3 * k, count), dtype=np.float64)
    if version == 1:
        h = b[c["centers"]:c["alphas"]].view("<u2").reshape(count, 3)
        out[0:3] = h.view(np.float16).astype(np.float64).T
    else:
        t = b[c["centers"]:c["alphas"]].reshape(count, 3, 3).astype(np.int64)
        v = t[:, :, 0] | (t[:, :, 1] << 8) | (t[:, :, 2] << 16)
        v = np.where(v >= 1 << 23, v - (1 << 24), v)
        out[0:3] = (v * (1.0 / (1 << fractional_bits))).T
    out[3:6] = (b[c["scales"]:c["rotations"]].reshape(count, 3) / 16.0 - 10.0).T
    if version >= 3:
        r = b[c["rotations"]:c["sh"]].reshape(count, 4).astype(np.uint64)
        packed = r[:, 0] | (r[:, 1] << 8) | (r[:, 2] << 16) | (r[:, 3] << 24)
        out[6:10] = np.array([decode_smallest_three(int(p)) for p in packed]).T
    else:
        r = b[c["rotations"]:c["sh"]].reshape(count, 3) / 127.5 - 1.0
        out[6:9] = r.T
        out[9] = np.sqrt(np.maximum(0.0, 1.0 - (r * r).sum(axis=1)))
    out[10] = b[c["alphas"]:c["colors"]] / 255.0
    out[11:14] = ((b[c["colors"]:c["scales"]].reshape(count, 3) / 255.0 - 0.5) * (SH_C0 / 0.15) + 0.5).T
    if k:
        out[14:] = ((b[c["sh"]:c["end"]].reshape(count, 3 * k).astype(np.float64) - 128.0) / 128.0).T
    return out

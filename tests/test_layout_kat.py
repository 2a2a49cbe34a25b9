"""Known-answer words of the splat and SH layouts (no GPU).

Every word below is derived by hand from the bit layouts the reference
documents, field by field, each field holding a distinct value so that a
swapped or shifted field fails:
- packedSplats, src/core/splats.ts:29-34 and the codec at :271-307;
- extSplats, src/core/splats.ts:12-21 (LoD opacity byte: encoding.ts:209-215;
  extRgb: encoding.ts:219-231; QuatOct32: encoding.ts:85-98 with 10 + 10
  axis and 12 angle bits);
- extSh and packedSh, src/core/sh.ts:7-25 (bands in order, coefficient-major
  RGB, one extRgb word per coefficient; per band a scale byte then sint8
  codes, bytes little-endian into words).
The oracle's encoders and decoders (oracle/codec.c) and the synthetic
table packers the GPU tests use (radix_amd/csrc/synth.cpp) must reproduce
them. The ext codec's round-trip budgets follow scripts/check-core.ts:341-400,
the cov codec's (src/core/splats.ts:219-266) scripts/check-core.ts:402-427.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from radix_amd import workloads as W
from tests.helpers import Rng

L = O.lib()


def d(v):
    return np.ascontiguousarray(v, dtype=np.float64)


# The splat of every known answer: centre (1, -2, -4); rotation pi / 3 about
# the axis (0.6, 0, 0.8), i.e. q = (0.3, 0, 0.4, cos(pi / 6)).
CENTER = [1.0, -2.0, -4.0]
QUAT = [0.3, 0.0, 0.4, math.cos(math.pi / 6)]
# Octahedral axis of (0.3, 0, 0.4): |.|_1 = 0.7, so (px, py) = (3/7, 0) and
# code = round((p / 2 + 1/2) max): x = 182.14 -> 182 at 8 bits, 730.71 -> 731
# at 10; y = 127.5 -> 128, 511.5 -> 512. Angle pi / 3 of pi: 85 of 255, 1365 of 4095.

# packedSplats, encoding rgb [0, 1], ln scale [-12, 9] (the default, splats.ts:97),
# rgb (0, 1, 0.2), opacity 0.6, scales e^-12, e^9, e^-1.5:
#   x = r 0 | g 255 << 8 | b 51 << 16 | opacity 153 << 24
#   y = half(1) 0x3c00 | half(-2) 0xc000 << 16
#   z = half(-4) 0xc400 | axis (182 | 128 << 8) << 16
#   w = scale codes round((ln s + 12) / (21 / 254)) + 1 = 1, 255, 128 | angle 85 << 24
PACKED_ENC = [0.0, 1.0, -12.0, 9.0]
PACKED_RGB, PACKED_OPACITY, PACKED_LN = [0.0, 1.0, 0.2], 0.6, [-12.0, 9.0, -1.5]
PACKED_WORDS = [0x9933FF00, 0xC0003C00, 0x80B6C400, 0x5580FF01]

# extSplats, rgb (0.25, -1, 0.5), LoD opacity 1.5, ln scales (-4, -2.5, 0.75):
#   A = float bits of the centre 0x3f800000, 0xc0000000, 0xc0800000 | normal none 0 | round(1.5 * 128) = 192 << 24
#   B.x = extRgb: max 1, base ceil(log2 1) + 15 = 15, magnitudes 0.25 * 255 = 63.75 -> 64, 255, 127.5 -> 128,
#         signs g (2): 64 | 255 << 8 | 128 << 16 | (15 << 3 | 2) << 24
#   B.y = half(-4) 0xc400 | half(-2.5) 0xc100 << 16;  B.z = half(0.75) 0x3a00 (no roughness, metallic)
#   B.w = QuatOct32 731 | 512 << 10 | 1365 << 20
EXT_RGB, EXT_OPACITY, EXT_LN = [0.25, -1.0, 0.5], 1.5, [-4.0, -2.5, 0.75]
EXT_A = [0x3F800000, 0xC0000000, 0xC0800000, 0xC0000000]
EXT_B = [0x7A80FF40, 0xC100C400, 0x00003A00, 0x555802DB]

# extSh, degree 3: coefficient k = ((k + 1) / 16, -(k + 1) / 32, 0), word k. Three by hand:
#   k = 0: max 1/16, base -4 + 15 = 11, scale 255 * 2^4: 255, 127.5 -> 128, 0, signs 2:
#          255 | 128 << 8 | (11 << 3 | 2) << 24
#   k = 7: max 1/2, base 14: 255, 128, 0 | (14 << 3 | 2) << 24
#   k = 14: max 15/16, base 15: 239.06 -> 239, 119.53 -> 120 | (15 << 3 | 2) << 24
EXT_SH_KNOWN = {0: 0x5A0080FF, 7: 0x720080FF, 14: 0x7A0078EF}

# packedSh, degree 3 (48 bytes, 12 words, 3 texels):
#   band 1 (values 0..8): max 1 -> scale code ceil(log2(1) 16) + 192 = 192; codes round(v * 127)
#   band 2 (values 9..23, v = 0.5 (j - 7) / 7): max 0.5 -> code 176; codes round((j - 7) / 7 * 127)
#   band 3 (values 24..44, v = 2 (j - 10) / 10): max 2 -> code 208; codes round((j - 10) / 10 * 127)
# Math.round takes halves up: 63.5 -> 64, -63.5 -> -63.
SH_BAND1 = [1.0, -1.0, 0.5, 0.25, 0.0, -0.5, 0.75, -0.25, 0.125]
SH_BAND2 = [0.5 * (j - 7) / 7 for j in range(15)]
SH_BAND3 = [2.0 * (j - 10) / 10 for j in range(21)]
PACKED_SH_BYTES = (
    [192, 127, -127, 64, 32, 0, -63, 95, -32, 16]
    + [176, -127, -109, -91, -73, -54, -36, -18, 0, 18, 36, 54, 73, 91, 109, 127]
    + [208, -127, -114, -102, -89, -76, -63, -51, -38, -25, -13, 0, 13, 25, 38, 51, 64, 76, 89, 102, 114, 127])


def packed_sh_words():
    b = np.array(PACKED_SH_BYTES, dtype=np.int64) & 0xFF
    assert b.size == 48
    return (b.reshape(12, 4) << np.array([0, 8, 16, 24])).sum(1).astype(np.uint32)


def sh_coeffs():
    return d(SH_BAND1 + SH_BAND2 + SH_BAND3)


def ext_sh_coeffs():
    return d([v for k in range(15) for v in ((k + 1) / 16, -(k + 1) / 32, 0.0)])


def test_packed_words_known_answer():
    w = np.zeros(4, dtype=np.uint32)
    L.orc_packed_encode(d(CENTER).ctypes.data_as(C.POINTER(C.c_double)), O.dptr(d(np.exp(PACKED_LN))), O.dptr(d(QUAT)),
                        O.dptr(d(PACKED_RGB)), PACKED_OPACITY, O.dptr(d(PACKED_ENC)), O.uptr(w))
    assert [hex(x) for x in w] == [hex(x) for x in PACKED_WORDS]
    # decoded field by field (splats.ts:288-306)
    c, s, q, rgb = (d(np.zeros(k)) for k in (3, 3, 4, 3))
    op = C.c_double()
    L.orc_packed_decode(O.uptr(np.array(PACKED_WORDS, dtype=np.uint32)), O.dptr(d(PACKED_ENC)), O.dptr(c), O.dptr(s),
                        O.dptr(q), O.dptr(rgb), C.byref(op))
    assert list(c) == CENTER
    assert list(rgb) == [0.0, 1.0, 51 / 255] and op.value == 153 / 255
    step = 21 / 254
    assert np.allclose(np.log(s), [-12.0, -12.0 + 254 * step, -12.0 + 127 * step], rtol=0, atol=1e-12)
    # axis (182, 128) of 255: x = 182/255*2-1, y = 128/255*2-1, z = 1-|x|-|y|; angle 85/255 pi
    ax = np.array([182 / 255 * 2 - 1, 128 / 255 * 2 - 1, 0.0])
    ax[2] = 1 - abs(ax[0]) - abs(ax[1])
    ax /= np.linalg.norm(ax)
    half = 85 / 255 * math.pi / 2
    assert np.allclose(q, list(ax * math.sin(half)) + [math.cos(half)], atol=1e-12)


def test_packed_words_from_the_synthetic_packer():
    planes = np.zeros((14, 1), dtype=np.float32)
    planes[0:3, 0] = CENTER
    planes[3:6, 0] = PACKED_LN
    planes[6:10, 0] = QUAT
    planes[10, 0] = PACKED_OPACITY
    planes[11:14, 0] = PACKED_RGB
    words, _ = W.pack(planes, 0, encoding=tuple(PACKED_ENC))
    assert [hex(x) for x in words[0]] == [hex(x) for x in PACKED_WORDS]


def test_ext_words_known_answer():
    a = np.zeros(4, dtype=np.uint32)
    b = np.zeros(4, dtype=np.uint32)
    L.orc_ext_encode(O.dptr(d(CENTER)), O.dptr(d(np.exp(EXT_LN))), O.dptr(d(QUAT)), O.dptr(d(EXT_RGB)), EXT_OPACITY,
                     O.uptr(a), O.uptr(b))
    assert [hex(x) for x in a] == [hex(x) for x in EXT_A]
    assert [hex(x) for x in b] == [hex(x) for x in EXT_B]
    c, s, q, rgb = (d(np.zeros(k)) for k in (3, 3, 4, 3))
    op = C.c_double()
    L.orc_ext_decode(O.uptr(np.array(EXT_A, dtype=np.uint32)), O.uptr(np.array(EXT_B, dtype=np.uint32)), O.dptr(c),
                     O.dptr(s), O.dptr(q), O.dptr(rgb), C.byref(op))
    assert list(c) == CENTER and op.value == 1.5
    assert list(rgb) == [64 / 255, -1.0, 128 / 255]
    assert list(np.log(s)) == pytest.approx(EXT_LN, abs=1e-15)
    ax = np.array([731 / 1023 * 2 - 1, 512 / 1023 * 2 - 1, 0.0])
    ax[2] = 1 - abs(ax[0]) - abs(ax[1])
    ax /= np.linalg.norm(ax)
    half = 1365 / 4095 * math.pi / 2
    assert np.allclose(q, list(ax * math.sin(half)) + [math.cos(half)], atol=1e-12)


def test_ext_sh_words_known_answer():
    words = np.zeros(16, dtype=np.uint32)
    L.orc_sh_encode(2, O.dptr(ext_sh_coeffs()), 3, O.uptr(words))
    for k, w in EXT_SH_KNOWN.items():
        assert hex(words[k]) == hex(w), k
    # word k is coefficient k's extRgb: every k, against the extRgb codec pinned by its own known answers
    for k in range(15):
        assert words[k] == L.orc_encode_ext_rgb((k + 1) / 16, -(k + 1) / 32, 0.0)
    back = d(np.zeros(45))
    L.orc_sh_decode(2, O.uptr(words), 3, O.dptr(back))
    assert list(back[0:3]) == [255 / 255 / 16, -128 / 255 / 16, 0.0]


def test_packed_sh_words_known_answer():
    exp = packed_sh_words()
    words = np.zeros(16, dtype=np.uint32)
    L.orc_sh_encode(3, O.dptr(sh_coeffs()), 3, O.uptr(words))
    assert [hex(x) for x in words[:12]] == [hex(x) for x in exp]
    back = d(np.zeros(45))
    L.orc_sh_decode(3, O.uptr(np.concatenate([exp, np.zeros(4, dtype=np.uint32)])), 3, O.dptr(back))
    codes = [b for i, b in enumerate(PACKED_SH_BYTES) if i not in (0, 10, 26)]
    scales = [1.0] * 9 + [0.5] * 15 + [2.0] * 21
    assert list(back) == [q * (sc / 127) for q, sc in zip(codes, scales)]  # code * (scale / 127), sh.ts:165
    # the synthetic packer's texels (channel t holds words 4t .. 4t + 3)
    planes = np.zeros((14 + 45, 1), dtype=np.float32)
    planes[14:, 0] = sh_coeffs()
    _, chans = W.pack(planes, 3)
    assert [hex(x) for x in np.concatenate([c[0] for c in chans])] == [hex(x) for x in exp]


def lnhalf_bound(v):
    # check-core.ts:86
    return max(abs(math.log(v)) * 2 ** -11, 2 ** -25) + 1e-12


def quat_angle(a, b):
    return 2 * math.acos(min(1.0, abs(float(np.dot(a / np.linalg.norm(a), b / np.linalg.norm(b))))))


def test_ext_codec_budgets():
    # check-core.ts:341-376 (no normal, roughness or metallic here: the projection reads none)
    r = Rng(77)
    worst = 0.0
    c, s, q, rgb = (d(np.zeros(k)) for k in (3, 3, 4, 3))
    op = C.c_double()
    a = np.zeros(4, dtype=np.uint32)
    b = np.zeros(4, dtype=np.uint32)
    for _ in range(2000):
        u = r.random(14)
        center = d((u[0:3] - 0.5) * 200)
        scales = d(np.exp(u[3:6] * 9 - 6))
        quat = d(u[6:10] * 2 - 1)
        colour = d(u[10:13] * 4 - 1)
        opacity = float(u[13] * 255 / 128)
        L.orc_ext_encode(O.dptr(center), O.dptr(scales), O.dptr(quat), O.dptr(colour), opacity, O.uptr(a), O.uptr(b))
        L.orc_ext_decode(O.uptr(a), O.uptr(b), O.dptr(c), O.dptr(s), O.dptr(q), O.dptr(rgb), C.byref(op))
        assert np.array_equal(c, center.astype(np.float32).astype(np.float64))
        assert abs(op.value - opacity) <= 1 / 256 + 1e-12
        for k in range(3):
            assert abs(math.log(s[k]) - math.log(scales[k])) <= lnhalf_bound(scales[k])
        assert np.max(np.abs(rgb - colour)) <= np.max(np.abs(colour)) / 255 + 1e-9
        worst = max(worst, quat_angle(q, quat))
    assert math.degrees(worst) <= 0.45


def test_ext_small_turns_keep_the_angle_precision():
    # check-core.ts:378-391: turns of up to 2 degrees within 0.03 degrees
    r = Rng(78)
    a = np.zeros(4, dtype=np.uint32)
    b = np.zeros(4, dtype=np.uint32)
    c, s, q, rgb = (d(np.zeros(k)) for k in (3, 3, 4, 3))
    op = C.c_double()
    worst = 0.0
    for _ in range(1000):
        u = r.random(4)
        axis = u[:3] * 2 - 1
        axis /= np.linalg.norm(axis)
        half = math.radians(u[3] * 2) / 2
        quat = d(list(axis * math.sin(half)) + [math.cos(half)])
        L.orc_ext_encode(O.dptr(d([0, 0, 0])), O.dptr(d([1, 1, 1])), O.dptr(quat), O.dptr(d([0, 0, 0])), 1.0, O.uptr(a),
                         O.uptr(b))
        L.orc_ext_decode(O.uptr(a), O.uptr(b), O.dptr(c), O.dptr(s), O.dptr(q), O.dptr(rgb), C.byref(op))
        worst = max(worst, quat_angle(q, quat))
    assert math.degrees(worst) <= 0.03


def test_ext_identity_and_zero_quaternions():
    # check-core.ts:393-399
    a = np.zeros(4, dtype=np.uint32)
    b = np.zeros(4, dtype=np.uint32)
    for quat in ([0, 0, 0, 1], [0, 0, 0, 0]):
        L.orc_ext_encode(O.dptr(d([0, 0, 0])), O.dptr(d([1, 1, 1])), O.dptr(d(quat)), O.dptr(d([0, 0, 0])), 1.0,
                         O.uptr(a), O.uptr(b))
        q = d(np.zeros(4))
        L.orc_decode_quat_oct(int(b[3]), 10, 12, O.dptr(q))
        assert list(q) == [0, 0, 0, 1]


# covSplats (src/core/splats.ts:25-27 and the codec at :219-266), A and B.x as
# ext; rgb (0.25, -1, 0.5) as the ext answer, LoD opacity 1.25; covariance
# diagonal e^-4, e^-2.5, e^0.75 and correlations c_xy = -e^-0.5, c_xz = e^-2,
# c_yz = 0, i.e. xy = -e^-0.5 sqrt(xx yy) = -e^-3.75, xz = e^-2 sqrt(xx zz) =
# e^-3.625, yz = 0. A correlation c is stored as sign(c) ln|c|:
#   A = the centre's float bits | round(1.25 * 128) = 160 << 24 (normal none)
#   B.x = 0x7A80FF40 (the ext answer's extRgb)
#   B.y = half(-4) 0xc400 | half(-2.5) 0xc100 << 16
#   B.z = half(0.75) 0x3a00 | half(-1 * -0.5 = 0.5) 0x3800 << 16
#   B.w = half(+1 * -2 = -2) 0xc000 | half(0) 0 << 16
COV_OPACITY = 1.25
COV6 = [math.exp(-4), math.exp(-2.5), math.exp(0.75), -math.exp(-3.75), math.exp(-3.625), 0.0]
COV_A = [0x3F800000, 0xC0000000, 0xC0800000, 0xA0000000]
COV_B = [0x7A80FF40, 0xC100C400, 0x38003A00, 0x0000C000]


def test_cov_words_known_answer():
    a = np.zeros(4, dtype=np.uint32)
    b = np.zeros(4, dtype=np.uint32)
    L.orc_cov_encode(O.dptr(d(CENTER)), O.dptr(d(COV6)), O.dptr(d(EXT_RGB)), COV_OPACITY, O.uptr(a), O.uptr(b))
    assert [hex(x) for x in a] == [hex(x) for x in COV_A]
    assert [hex(x) for x in b] == [hex(x) for x in COV_B]
    c, cov, rgb = (d(np.zeros(k)) for k in (3, 6, 3))
    op = C.c_double()
    L.orc_cov_decode(O.uptr(np.array(COV_A, dtype=np.uint32)), O.uptr(np.array(COV_B, dtype=np.uint32)), O.dptr(c),
                     O.dptr(cov), O.dptr(rgb), C.byref(op))
    assert list(c) == CENTER and op.value == 1.25
    assert list(rgb) == [64 / 255, -1.0, 128 / 255]
    # every field decoded from its own bits: ln of the diagonal, c = -sign(e) e^-|e| times sqrt(a b)
    assert list(np.log(cov[:3])) == pytest.approx([-4.0, -2.5, 0.75], abs=1e-15)
    assert cov[3] == pytest.approx(-math.exp(-0.5) * math.sqrt(cov[0] * cov[1]), rel=1e-15)
    assert cov[4] == pytest.approx(math.exp(-2.0) * math.sqrt(cov[0] * cov[2]), rel=1e-15)
    assert cov[5] == 0.0


def test_cov_correlation_edges():
    # splats.ts:221-232: |c| = 1 keeps -1e-7 (a half subnormal, 2 x 2^-24);
    # the clamp to [-1, 1] comes first, so an infinite ratio is a full
    # correlation and only a NaN (0 / 0) stores 0, as c = 0 does; |c| below
    # e^-100 keeps -100
    enc, dec = L.orc_cov_encode_correlation, L.orc_cov_decode_correlation
    assert enc(1.0, 1.0, 1.0) == -1e-7 and enc(-2.0, 1.0, 4.0) == 1e-7
    assert L.orc_encode_half(enc(1.0, 1.0, 1.0)) == 0x8002
    assert dec(L.orc_decode_half(0x8002)) == pytest.approx(1.0 - 2.0 * 2 ** -24, rel=1e-15)
    assert enc(0.0, 1.0, 1.0) == 0.0 and enc(0.0, 0.0, 0.0) == 0.0
    assert enc(1.0, 0.0, 0.0) == -1e-7 and enc(-1.0, 0.0, 1.0) == 1e-7
    assert enc(1e-50, 1.0, 1.0) == -100.0 and enc(-1e-50, 1.0, 1.0) == 100.0
    assert L.orc_encode_half(-100.0) == 0xD640
    assert dec(-100.0) == pytest.approx(math.exp(-100), rel=1e-15) and dec(100.0) == -dec(-100.0)
    assert dec(0.0) == 0.0


def test_cov_codec_budgets():
    # check-core.ts:402-427: a random splat's covariance (covFromScaleQuat) through the words
    r = Rng(79)
    a = np.zeros(4, dtype=np.uint32)
    b = np.zeros(4, dtype=np.uint32)
    c, cov, rgb = (d(np.zeros(k)) for k in (3, 6, 3))
    op = C.c_double()
    ref = d(np.zeros(6))
    for _ in range(2000):
        u = r.random(14)
        center = d((u[0:3] - 0.5) * 200)
        scales = d(np.exp(u[3:6] * 9 - 6))
        quat = d(u[6:10] * 2 - 1)
        quat /= np.linalg.norm(quat)
        colour = d(u[10:13] * 4 - 1)
        opacity = float(u[13] * 255 / 128)
        L.orc_cov_from_scale_quat(O.dptr(scales), O.dptr(quat), O.dptr(ref))
        L.orc_cov_encode(O.dptr(center), O.dptr(ref), O.dptr(colour), opacity, O.uptr(a), O.uptr(b))
        L.orc_cov_decode(O.uptr(a), O.uptr(b), O.dptr(c), O.dptr(cov), O.dptr(rgb), C.byref(op))
        assert np.array_equal(c, center.astype(np.float32).astype(np.float64))
        assert abs(op.value - opacity) <= 1 / 256 + 1e-12
        assert np.max(np.abs(rgb - colour)) <= np.max(np.abs(colour)) / 255 + 1e-9
        for k in range(3):
            assert abs(math.log(cov[k]) - math.log(ref[k])) <= lnhalf_bound(ref[k])
        for k, i, j in ((3, 0, 1), (4, 0, 2), (5, 1, 2)):
            ab = math.sqrt(ref[i] * ref[j])
            corr = abs(ref[k]) / ab
            bound = ab * (2e-4 + corr * (lnhalf_bound(ref[i]) + lnhalf_bound(ref[j])) / 2)
            assert abs(cov[k] - ref[k]) <= bound, (k, ref[k], cov[k])

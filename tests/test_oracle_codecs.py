"""The oracle's codecs, SH and covariance, pinned (no GPU).

Pins: numpy's float16 (IEEE round to nearest even) for halves, the
reference's bit budgets (scripts/check-core.ts:330-470, 1265-1300), known
answers (check-core.ts:1666-1678), scipy's associated Legendre functions for
the SH basis (the textbook real SH with the Condon-Shortley phase, as
check-core.ts:1680-1722 restates them), and a rotation matrix built from the
quaternion by scipy for the covariance.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation
from scipy.special import lpmv

from oracle import oracle as O
from tests.helpers import Rng

L = O.lib()


def d3(v):
    return np.ascontiguousarray(v, dtype=np.float64)


def test_half_matches_ieee_round_to_nearest_even():
    f = np.arange(0, 0xFFFFFFFF, 99991, dtype=np.uint64).astype(np.uint32).view(np.float32)
    f = f[np.isfinite(f)]
    with np.errstate(over="ignore"):
        exp = f.astype(np.float16).view(np.uint16)
    got = np.array([L.orc_encode_half(float(x)) for x in f], dtype=np.uint32)
    assert np.array_equal(got, exp.astype(np.uint32))
    h = np.arange(0, 0x10000, dtype=np.uint32)
    h = h[(h & 0x7C00) != 0x7C00]
    dec = np.array([L.orc_decode_half(int(x)) for x in h])
    assert np.array_equal(dec, h.astype(np.uint16).view(np.float16).astype(np.float64))


def test_ext_rgb_known_answers_and_budget():
    out = d3(np.zeros(3))
    # 1.0: ceil(log2 1) = 0, base 15, magnitude 255 exactly
    e = L.orc_encode_ext_rgb(1.0, 0.5, -0.25)
    assert e >> 27 == 15 and (e & 0xFF) == 255 and ((e >> 24) & 7) == 4
    L.orc_decode_ext_rgb(e, O.dptr(out))
    assert abs(out[0] - 1.0) < 1e-15 and abs(out[1] - 128 / 255) < 1e-15 and abs(out[2] + 64 / 255) < 1e-15
    r = Rng(3)
    v = (r.random(3000) - 0.5) * 8.0
    for i in range(0, v.size, 3):
        rgb = v[i:i + 3]
        L.orc_decode_ext_rgb(L.orc_encode_ext_rgb(*rgb), O.dptr(out))
        # check-core.ts:362-365: within maxAbs / 255
        assert np.max(np.abs(out - rgb)) <= np.max(np.abs(rgb)) / 255 + 1e-9


def test_oct_codecs_zero_and_minus_z():
    # check-core.ts:330-339
    out = d3(np.ones(3))
    L.orc_decode_oct(L.orc_encode_oct(O.dptr(d3([0, 0, 0])), 12), 12, O.dptr(out))
    assert np.all(out == 0)
    code = L.orc_encode_oct(O.dptr(d3([0, 0, -1])), 12)
    L.orc_decode_oct(code, 12, O.dptr(out))
    assert code != 0 and abs(out[2] + 1) < 1e-9


def _quat_angle(a, b):
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    return 2 * math.degrees(math.acos(min(1.0, abs(float(np.dot(a, b))))))


@pytest.mark.parametrize("axis_bits,angle_bits,limit", [(10, 12, 0.45), (8, 8, 2.0)])
def test_quat_oct_budgets(axis_bits, angle_bits, limit):
    # check-core.ts:393-395 (QuatOct32 within 0.45 degrees) and 467-469 (QuatOct24 within 2)
    r = Rng(17)
    g = r.random(4 * 2000).reshape(-1, 4) - 0.5
    out = d3(np.zeros(4))
    worst = 0.0
    for q in g:
        q = d3(q)
        L.orc_decode_quat_oct(L.orc_encode_quat_oct(O.dptr(q), axis_bits, angle_bits), axis_bits, angle_bits,
                              O.dptr(out))
        worst = max(worst, _quat_angle(out, q))
    assert worst <= limit
    # identity and zero quaternions decode to identity (check-core.ts:414-420)
    for q in ([0, 0, 0, 1], [0, 0, 0, 0]):
        L.orc_decode_quat_oct(L.orc_encode_quat_oct(O.dptr(d3(q)), axis_bits, angle_bits), axis_bits, angle_bits,
                              O.dptr(out))
        assert list(out) == [0, 0, 0, 1]


def test_packed_codec_within_ranges():
    # check-core.ts:442-470, with the quantized stereo workload's encoding [-10, 2] for log-scales
    enc = d3([0.0, 1.0, -10.0, 2.0])
    r = Rng(23)
    w = np.zeros(4, dtype=np.uint32)
    c, s, q, rgb, op = (d3(np.zeros(k)) for k in (3, 3, 4, 3, 1))
    opv = d3([0.0])
    for i in range(1000):
        u = r.random(14)
        center = d3((u[0:3] - 0.5) * 20)
        scales = d3(np.exp(-10 + u[3:6] * 12))
        quat = d3(u[6:10] - 0.5)
        color = d3(u[10:13])
        opacity = float(u[13])
        L.orc_packed_encode(O.dptr(center), O.dptr(scales), O.dptr(quat), O.dptr(color), opacity, O.dptr(enc),
                            O.uptr(w))
        L.orc_packed_decode(O.uptr(w), O.dptr(enc), O.dptr(c), O.dptr(s), O.dptr(q), O.dptr(rgb), O.dptr(opv))
        assert np.max(np.abs(rgb - color)) <= 1 / 510 + 1e-9
        assert abs(opv[0] - opacity) <= 1 / 510 + 1e-9
        assert np.all(np.abs(c - center) <= np.abs(center) * 2 ** -11 + 2 ** -25)
        assert np.max(np.abs(np.log(s) - np.log(scales))) <= 12 / 254 / 2 + 1e-9
        assert _quat_angle(q, quat) <= 2.0


def test_sh_band1_axes_known_answer():
    # check-core.ts:1666-1678
    c = d3(np.zeros(72))
    c[0], c[4], c[8] = 1, 2, 3
    out = d3(np.zeros(3))
    k = 0.4886025
    L.orc_sh_evaluate(O.dptr(c), O.dptr(d3([0, 1, 0])), 1, O.dptr(out))
    assert abs(out[0] + k) < 1e-9
    L.orc_sh_evaluate(O.dptr(c), O.dptr(d3([0, 0, 1])), 1, O.dptr(out))
    assert abs(out[1] - 2 * k) < 1e-9
    L.orc_sh_evaluate(O.dptr(c), O.dptr(d3([1, 0, 0])), 1, O.dptr(out))
    assert abs(out[2] + 3 * k) < 1e-9


def real_sh(l, m, d):
    """Real SH from scipy's associated Legendre P_l^|m| (Condon-Shortley phase)."""
    am = abs(m)
    theta_cos = d[2]
    phi = math.atan2(d[1], d[0])
    k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    p = lpmv(am, l, theta_cos)
    if m == 0:
        return k * p
    if m > 0:
        return math.sqrt(2) * k * math.cos(m * phi) * p
    return math.sqrt(2) * k * math.sin(am * phi) * p


def test_sh_basis_is_real_sh_bands_1_to_4():
    r = Rng(31)
    basis = d3(np.zeros(24))
    worst = 0.0
    for _ in range(200):
        v = r.random(3) * 2 - 1
        d = d3(v / np.linalg.norm(v))
        L.orc_sh_basis(O.dptr(d), O.dptr(basis))
        for l in range(1, 5):
            for m in range(-l, l + 1):
                k = l * l - 1 + m + l
                worst = max(worst, abs(basis[k] - real_sh(l, m, d)))
    assert worst < 2e-6


def test_sh_evaluate_caps_degree():
    c = d3(np.ones(72))
    out = d3(np.zeros(3))
    d = d3([0.3, 0.4, math.sqrt(1 - 0.25)])
    basis = d3(np.zeros(24))
    L.orc_sh_basis(O.dptr(d), O.dptr(basis))
    for deg, n in ((0, 0), (1, 3), (2, 8), (3, 15)):
        L.orc_sh_evaluate(O.dptr(c), O.dptr(d), deg, O.dptr(out))
        assert abs(out[1] - basis[:n].sum()) < 1e-12


@pytest.mark.parametrize("codec,budget", [(2, 0.2), (3, 0.4)])
def test_sh_codec_rms_budgets(codec, budget):
    # check-core.ts:1265-1300: RMS over directions in 8-bit steps
    r = Rng(41)
    for degree in (1, 2, 3):
        n = 3 * [0, 3, 8, 15][degree]
        words = np.zeros(16, dtype=np.uint32)
        back = d3(np.zeros(45))
        total = 0.0
        count = 1500
        for i in range(count):
            u = r.random(2 * n)
            vals = (u[:n] - 0.5) * (10 if i % 211 == 0 else 0.4 * u[n:2 * n])
            vals = d3(np.concatenate([vals, np.zeros(45 - n)]))
            L.orc_sh_encode(codec, O.dptr(vals), degree, O.uptr(words))
            L.orc_sh_decode(codec, O.uptr(words), degree, O.dptr(back))
            total += float(np.sum((back[:n] - vals[:n]) ** 2))
        rms = math.sqrt(total / count / 3 / (4 * math.pi)) * 255
        assert rms <= budget, (codec, degree, rms)


def test_sh_packed_scale_code_covers_max():
    for m in (1e-6, 0.01, 0.3, 1.0, 7.5):
        code = L.orc_sh_scale_encode(m)
        assert L.orc_sh_scale_decode(code) >= m
        assert code == 1 or L.orc_sh_scale_decode(code - 1) < m
    assert L.orc_sh_scale_encode(0.0) == 0


def test_cov_from_scale_quat_is_r_s2_rt():
    r = Rng(5)
    cov = d3(np.zeros(6))
    for _ in range(200):
        u = r.random(7)
        s = d3(np.exp(u[:3] * 4 - 3))
        q = d3(u[3:] - 0.5)
        L.orc_cov_from_scale_quat(O.dptr(s), O.dptr(q), O.dptr(cov))
        R = Rotation.from_quat(q / np.linalg.norm(q)).as_matrix()  # scipy: (x, y, z, w)
        S = R @ np.diag(s ** 2) @ R.T
        exp = [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]]
        assert np.allclose(cov, exp, rtol=1e-12, atol=1e-15 * np.max(s) ** 2)


def test_sh_apply_table_is_sh_evaluate_per_splat():
    # the CPU baseline's table-wide Sh.evaluate: base colour plus evaluate() along normalize(centre - origin)
    r = Rng(61)
    n = 50
    centre = (r.random(3 * n).reshape(3, n) - 0.5).astype(np.float32) * 4
    sh = (r.random(45 * n).reshape(45, n) - 0.5).astype(np.float32)
    rgb = r.random(3 * n).reshape(3, n).astype(np.float32)
    origin = d3([0.1, -0.2, 0.3])
    got = np.ascontiguousarray(rgb.copy())
    L.orc_sh_apply_table(n, n, np.ascontiguousarray(centre).ctypes.data, np.ascontiguousarray(sh).ctypes.data, 3,
                         O.dptr(origin), got.ctypes.data)
    out = d3(np.zeros(3))
    for i in range(n):
        v = centre[:, i].astype(np.float64) - origin
        L.orc_sh_evaluate(O.dptr(d3(sh[:, i])), O.dptr(d3(v / np.linalg.norm(v))), 3, O.dptr(out))
        assert np.allclose(got[:, i], (rgb[:, i] + out).astype(np.float32), rtol=0, atol=1e-6)

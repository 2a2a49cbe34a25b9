"""The draw oracle (oracle/draw.c, the ewa splat draw of
src/render/splatDraw.ts:98-181) pinned by what fixes it independently: a
circular and a rotated elliptical 2D Gaussian in pixels (Sigma = R diag(s^2)
R^T with s the bound's pixel scale over maxStdDev), the premultiplied "over"
operator for two splats, the cutoff and minAlpha discards, flat discs at
falloff 0, LoD opacities (alpha 1 at the centre, the disc out to the bound),
the sRGB curve's known values, and the skipped entries (culled words,
indices past n, counts past the order)."""
import math

import numpy as np
import pytest

from oracle import oracle as O

MSD = math.sqrt(8.0)
MIN_ALPHA = 0.5 / 255.0


def log_scale_code(s):
    """The 12-bit log2 pixel-scale code nearest to s, and the scale it decodes to."""
    code = int(math.floor((math.log2(s) + 4.0) / 16.0 * 4094.0 + 0.5)) + 1
    return code, 2.0 ** ((code - 1) / 4094.0 * 16.0 - 4.0)


CENTRE = {}


def word(ndc_x, ndc_y, angle_code, sx, sy, byte, rgb=(0.5, 0.25, 1.0)):
    """The four ewa words of one eye (projection.tsl.ts:277-311), from an ndc
    centre (to the nearest fixed16 code), an angle code (angle = code pi /
    4096), pixel scales and the opacity byte. Returns (words, decoded sx,
    decoded sy); CENTRE["ndc"] holds the decoded centre."""
    fx = int(round((ndc_x + 2.0) / 4.0 * 65535.0))
    fy = int(round((ndc_y + 2.0) / 4.0 * 65535.0))
    CENTRE["ndc"] = (fx / 65535.0 * 4.0 - 2.0, fy / 65535.0 * 4.0 - 2.0)
    cx, dsx = log_scale_code(sx)
    cy, dsy = log_scale_code(sy)
    w = np.array([fx | (fy << 16), (1000 & 0xFFFFF) | (angle_code << 20), O.lib().orc_encode_ext_rgb(*rgb),
                  byte | (cx << 8) | (cy << 20)], dtype=np.uint32)
    return w, dsx, dsy


def pixel_ndc(x, y, W, H):
    # framebuffer pixel (x, y), row 0 at the top
    return (x + 0.5) * 2.0 / W - 1.0, 1.0 - (y + 0.5) * 2.0 / H


def centred_word(W, H, px, py, **kw):
    # a centre at (the fixed16 code nearest) pixel (px, py)'s centre
    nx, ny = pixel_ndc(px, py, W, H)
    return word(nx, ny, **kw)


def offset(x, y, W, H):
    """Pixel (x, y)'s centre minus the decoded splat centre, in pixels, y up."""
    nx, ny = pixel_ndc(x, y, W, H)
    cx, cy = CENTRE["ndc"]
    return (nx - cx) * 0.5 * W, (ny - cy) * 0.5 * H


def ext_rgb(rgb):
    out = (C := __import__("ctypes")).c_double * 3
    buf = out()
    O.lib().orc_decode_ext_rgb(O.lib().orc_encode_ext_rgb(*rgb), buf)
    return np.array(buf[:])


def test_circular_gaussian_in_pixels():
    W = H = 64
    w, sx, _ = centred_word(W, H, 20, 30, angle_code=0, sx=9.0, sy=9.0, byte=96)
    img = O.draw(w[None], [0], 1, W, H)
    sigma = sx / MSD  # the bound's pixel radius is maxStdDev standard deviations
    op = 96 / 128.0
    rgb = ext_rgb((0.5, 0.25, 1.0))
    cut = min(MSD, math.sqrt(2.0 * math.log(op / MIN_ALPHA)))
    for (x, y) in [(20, 30), (23, 30), (20, 26), (25, 33), (30, 30)]:
        dx, dy = offset(x, y, W, H)
        r2 = dx * dx + dy * dy
        a = op * math.exp(-r2 / (2.0 * sigma * sigma))
        inside = r2 / (sigma * sigma) <= cut * cut and a >= MIN_ALPHA
        exp_a = a if inside else 0.0
        assert img[3, y, x] == pytest.approx(exp_a, abs=1e-12), (x, y)
        assert np.allclose(img[:3, y, x], rgb * exp_a, atol=1e-12)
    assert img[3, 30, 20] == pytest.approx(op, rel=1e-6)


def test_rotated_elliptical_gaussian():
    W, H = 96, 64
    code = 700  # angle = 700 pi / 4096
    w, sx, sy = centred_word(W, H, 40, 20, angle_code=code, sx=14.0, sy=5.0, byte=128)
    img = O.draw(w[None], [0], 1, W, H, falloff=1.0)
    th = code * math.pi / 4096.0
    R = np.array([[math.cos(th), math.sin(th)], [math.sin(th), -math.cos(th)]]).T  # columns: axis1, axis2 (y up)
    Sig = R @ np.diag([(sx / MSD) ** 2, (sy / MSD) ** 2]) @ R.T
    Si = np.linalg.inv(Sig)
    cut = min(MSD, math.sqrt(2.0 * math.log(1.0 / MIN_ALPHA)))
    checked = 0
    for y in range(0, H):
        for x in range(0, W):
            d = np.array(offset(x, y, W, H), dtype=np.float64)  # y up
            m2 = d @ Si @ d
            a = math.exp(-0.5 * m2)
            exp_a = a if (m2 <= cut * cut and a >= MIN_ALPHA) else 0.0
            if abs(m2 - cut * cut) < 1e-6:
                continue
            assert img[3, y, x] == pytest.approx(exp_a, abs=1e-9), (x, y)
            checked += exp_a > 0
    assert checked > 100


def test_two_splats_blend_over_in_draw_order():
    W = H = 32
    w1, sx, _ = centred_word(W, H, 10, 10, angle_code=0, sx=40.0, sy=40.0, byte=64, rgb=(1.0, 0.0, 0.0))
    w2, _, _ = centred_word(W, H, 10, 10, angle_code=0, sx=40.0, sy=40.0, byte=100, rgb=(0.0, 0.0, 1.0))
    words = np.stack([w1, w2])
    img = O.draw(words, [0, 1], 2, W, H)  # splat 0 first (far), splat 1 over it
    dx, dy = offset(10, 10, W, H)
    g = math.exp(-(dx * dx + dy * dy) / (2.0 * (sx / MSD) ** 2))
    a1, a2 = 64 / 128.0 * g, 100 / 128.0 * g
    c1, c2 = ext_rgb((1.0, 0.0, 0.0)), ext_rgb((0.0, 0.0, 1.0))
    assert np.allclose(img[:3, 10, 10], c2 * a2 + (1.0 - a2) * c1 * a1, atol=1e-12)
    assert img[3, 10, 10] == pytest.approx(a2 + (1.0 - a2) * a1, abs=1e-12)
    rev = O.draw(words, [1, 0], 2, W, H)
    assert np.allclose(rev[:3, 10, 10], c1 * a1 + (1.0 - a1) * c2 * a2, atol=1e-12)


def test_flat_disc_at_falloff_zero():
    W = H = 48
    w, sx, _ = centred_word(W, H, 24, 24, angle_code=0, sx=10.0, sy=10.0, byte=64)
    img = O.draw(w[None], [0], 1, W, H, falloff=0.0)
    # falloff < 1: the cutoff stays maxStdDev, the disc reaches the bound's pixel radius
    for y in range(H):
        for x in range(W):
            r = math.hypot(*offset(x, y, W, H))
            if abs(r - sx) < 1e-6:
                continue
            assert img[3, y, x] == pytest.approx(0.5 if r < sx else 0.0, abs=1e-12)


def test_lod_opacity_above_one():
    W = H = 48
    byte = 200  # 1.5625: an LoD stretch, alpha 4 x 1.5625 - 3 = 3.25
    w, sx, _ = centred_word(W, H, 24, 24, angle_code=0, sx=8.0, sy=8.0, byte=byte)
    img = O.draw(w[None], [0], 1, W, H)
    assert img[3, 24, 24] == pytest.approx(1.0, abs=1e-9)  # 1 - (1 - g)^a, g within 1e-7 of 1
    # bound = cut: fragments out to the bound's pixel radius, none beyond
    assert img[3, 24, 24 + 7] > 0.0
    assert img[3, 24, 24 + 9] == 0.0
    flat = O.draw(w[None], [0], 1, W, H, falloff=0.0)
    assert flat[3, 24, 24 + 7] == pytest.approx(1.0, abs=1e-12)


def test_cutoff_and_min_alpha_discards():
    W = H = 64
    op_byte = 2  # 1/64: the cutoff sqrt(2 ln(op / minAlpha)) is well inside maxStdDev
    w, sx, _ = centred_word(W, H, 32, 32, angle_code=0, sx=30.0, sy=30.0, byte=op_byte)
    img = O.draw(w[None], [0], 1, W, H)
    op = op_byte / 128.0
    cut = math.sqrt(2.0 * math.log(op / MIN_ALPHA))
    sigma = sx / MSD
    drawn = img[3] > 0
    ys, xs = np.nonzero(drawn)
    r = np.array([math.hypot(*offset(x, y, W, H)) for x, y in zip(xs, ys)])
    assert r.max() <= cut * sigma + 1e-9
    assert img[3][drawn].min() >= MIN_ALPHA


def test_srgb_decode_for_linear_outputs():
    W = H = 16
    w, _, _ = centred_word(W, H, 8, 8, angle_code=0, sx=20.0, sy=20.0, byte=128, rgb=(0.5, 0.04, 1.0))
    lin = O.draw(w[None], [0], 1, W, H, linear=True)
    c = ext_rgb((0.5, 0.04, 1.0))
    srgb = lambda v: v / 12.92 if v <= 0.04045 else ((v + 0.055) / 1.055) ** 2.4
    assert lin[0, 8, 8] == pytest.approx(srgb(c[0]), rel=1e-6)
    assert lin[1, 8, 8] == pytest.approx(srgb(c[1]), rel=1e-6)
    assert srgb(0.5) == pytest.approx(0.2140411, rel=1e-6)


def test_skipped_entries():
    W = H = 16
    w, _, _ = centred_word(W, H, 8, 8, angle_code=0, sx=20.0, sy=20.0, byte=128)
    culled = w.copy()
    culled[3] &= ~np.uint32(0xFF)
    words = np.stack([w, culled])
    assert O.draw(words, [1], 1, W, H).sum() == 0.0          # a culled eye
    assert O.draw(words, [5, 9], 2, W, H).sum() == 0.0       # indices past n
    assert O.draw(words, [0], 0, W, H).sum() == 0.0          # nothing counted
    assert np.array_equal(O.draw(words, [0], 7, W, H), O.draw(words, [0], 1, W, H))  # count past the order

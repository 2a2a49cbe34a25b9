"""The oracle's draw of the ray-evaluated methods (orc_draw_ray, oracle/draw.c),
pinned (no GPU).

One splat is projected by the pinned ray-word oracle (orc_project_ray) and
drawn alone over a transparent target, so a pixel's alpha is op exp(-z^2 / 2)
(falloff 1, opacity op <= 1) and z^2 = -2 ln(alpha / op) is read back. The
pins are what the methods' densities mean, computed here from geometry:
- 3dgut and htgs evaluate the Gaussian at its peak along the pixel's ray, so
  z^2 is the ray's smallest squared Mahalanobis distance to the centre. For a
  sphere of radius s at distance L seen at angle theta from its centre that
  is (L sin theta / s)^2; for a rotated, off-axis ellipsoid the minimum is
  found here by a bracketed search along the ray (not the closed form the
  oracle uses); under an orthographic camera the ray is parallel to -z and
  the distance is the in-plane offset.
- raygs evaluates 1 / (1 / c^2 + 1 / |z|^2) at the ray's hit on the quad
  facing the eye; for a sphere (lengths s, c^2 = L^2 / s^2) that is again
  (L sin theta / s)^2, up to the words' half-float lengths.
- htgs's quad is the exact outline of the std-dev ellipsoid, so every pixel
  whose density is within the cutoff is drawn; pixels beyond the cutoff are
  not.
- Two splats blend back to front.
The decoded centre, scales and opacity come from the pinned word codecs
(orc_decode_center, orc_decode_half)."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from radix_amd import camera as cam

W, H = 160, 120
FOV = 50.0
CUT = math.sqrt(8.0)


def params(**kw):
    base = dict(blur_amount=0.0, pre_blur_amount=0.0, sort_bits=32, cull_keys=True, max_pixel_radius=1e6)
    base.update(kw)
    return O.params(**base)


def one(method, rig, center, scales, quat, op=0.8, prm=None):
    view = O.view_from_rig(rig)
    o = O.project_ray(method, center, scales, quat, [0.6, 0.3, 0.2], op, view, prm or params())
    assert o.eye_visible[0]
    words = np.array(list(o.words), dtype=np.uint32)[None, :]
    img = O.draw_ray(method, words, np.array([0], dtype=np.uint32), 1, rig, falloff=1.0)
    return o, img


def decoded(o, method):
    w = list(o.words)
    c = np.zeros(3)
    last = w[3] if method == "raygs" else w[5]
    O.lib().orc_decode_center(w[0], w[1], last, O.dptr(c))
    byte = (w[3] & 0xFF) if method == "raygs" else ((w[5] >> 16) & 0xFF)
    op = byte / 128.0
    if method == "raygs":
        return c, None, op
    s = np.array([math.exp(O.lib().orc_decode_half(w[4] & 0xFFFF)), math.exp(O.lib().orc_decode_half(w[4] >> 16)),
                  math.exp(O.lib().orc_decode_half(w[5] & 0xFFFF))])
    return c, s, op


def pixel_rays(rig):
    """Each pixel centre's ray: (origin, direction) in the eye, perspective."""
    P = np.asarray(rig.eyes[0].projection)
    xs = (np.arange(W) + 0.5) * 2.0 / W - 1.0
    ys = 1.0 - (np.arange(H) + 0.5) * 2.0 / H
    nx, ny = np.meshgrid(xs, ys)
    # at view depth -1: x = (nx w - P02 z - P03) / P00 with w = P32 z + P33
    z = -1.0
    w = P[3, 2] * z + P[3, 3]
    dx = (nx * w - P[0, 2] * z - P[0, 3]) / P[0, 0]
    dy = (ny * w - P[1, 2] * z - P[1, 3]) / P[1, 1]
    return np.stack([dx, dy, np.full_like(dx, z)], axis=-1)


def read_z2(img, op):
    a = img[3]
    with np.errstate(divide="ignore"):
        z2 = -2.0 * np.log(a / op)
    return np.where(a > 0, z2, np.inf)


@pytest.mark.parametrize("method", ["3dgut", "htgs", "raygs"])
def test_sphere_density_is_the_closest_approach(method):
    rig = cam.mono(np.eye(4), FOV, W, H, near=0.05, far=100)
    center, s = [0.35, -0.2, -4.0], 0.35
    q = Rotation.from_euler("xyz", [0.4, -0.9, 0.2]).as_quat()
    o, img = one(method, rig, center, [s] * 3, q)
    c, sc, op = decoded(o, method)
    if sc is not None:
        s_dec = sc[0]
        assert np.allclose(sc, s_dec)
    else:
        s_dec = s
    d = pixel_rays(rig)
    dn = d / np.linalg.norm(d, axis=-1, keepdims=True)
    L = np.linalg.norm(c)
    cos_t = dn @ (c / L)
    z2_true = (L * L * (1.0 - cos_t ** 2)) / (s_dec * s_dec)
    z2 = read_z2(img, op)
    drawn = np.isfinite(z2)
    print(method, "drawn pixels", int(drawn.sum()), flush=True)
    assert drawn.sum() > 200
    rel = 1e-9 if method != "raygs" else 3e-3  # raygs: half-float lengths and a float c^2 in its words
    err = np.abs(z2[drawn] - z2_true[drawn]) / np.maximum(z2_true[drawn], 1e-3)
    assert err.max() < rel, err.max()


def closest_mahalanobis(o, d, c, R, s):
    """min over t of |S^-1 R^T (o + t d - c)|^2 by a bracketed golden search."""
    A = (R / s).T  # rows: the whitened axes
    f = lambda t: float(np.sum((A @ (o + t * d - c)) ** 2))
    lo, hi = 0.0, 40.0
    g = (math.sqrt(5.0) - 1.0) / 2.0
    a, b = hi - g * (hi - lo), lo + g * (hi - lo)
    fa, fb = f(a), f(b)
    for _ in range(200):
        if fa < fb:
            hi, b, fb = b, a, fa
            a = hi - g * (hi - lo)
            fa = f(a)
        else:
            lo, a, fa = a, b, fb
            b = lo + g * (hi - lo)
            fb = f(b)
    return min(fa, fb)


@pytest.mark.parametrize("method", ["3dgut", "htgs"])
def test_ellipsoid_density_is_the_ray_minimum(method):
    rig = cam.mono(np.eye(4), FOV, W, H, near=0.05, far=100)
    center = [-0.3, 0.25, -3.5]
    scales = [0.45, 0.12, 0.25]
    q = Rotation.from_euler("zyx", [0.8, -0.5, 1.3]).as_quat()
    o, img = one(method, rig, center, scales, q)
    c, s, op = decoded(o, method)
    hq = np.zeros(4)
    O.lib().orc_decode_quat_oct(o.words[3], 10, 12, O.dptr(hq))
    R = Rotation.from_quat(hq).as_matrix()
    d = pixel_rays(rig)
    z2 = read_z2(img, op)
    ys, xs = np.nonzero(np.isfinite(z2))
    assert ys.size > 100
    pick = np.linspace(0, ys.size - 1, 60).astype(int)
    for k in pick:
        y, x = ys[k], xs[k]
        ref = closest_mahalanobis(np.zeros(3), d[y, x], c, R, s)
        assert abs(z2[y, x] - ref) < 1e-7 * max(ref, 1.0), (method, y, x, z2[y, x], ref)


@pytest.mark.parametrize("method", ["3dgut", "htgs", "raygs"])
def test_orthographic_sphere(method):
    rig = cam.mono_ortho(np.eye(4), 1.5, W, H)
    center, s = [0.3, -0.15, -5.0], 0.3
    o, img = one(method, rig, center, [s] * 3, [0, 0, 0, 1])
    c, sc, op = decoded(o, method)
    s_dec = sc[0] if sc is not None else s
    P = np.asarray(rig.eyes[0].projection)
    xs = ((np.arange(W) + 0.5) * 2.0 / W - 1.0 - P[0, 3]) / P[0, 0]
    ys = (1.0 - (np.arange(H) + 0.5) * 2.0 / H - P[1, 3]) / P[1, 1]
    gx, gy = np.meshgrid(xs, ys)
    z2_true = ((gx - c[0]) ** 2 + (gy - c[1]) ** 2) / (s_dec * s_dec)
    z2 = read_z2(img, op)
    drawn = np.isfinite(z2)
    assert drawn.sum() > 200
    rel = 1e-9 if method != "raygs" else 3e-3
    err = np.abs(z2[drawn] - z2_true[drawn]) / np.maximum(z2_true[drawn], 1e-3)
    assert err.max() < rel, err.max()


def test_htgs_quad_covers_the_cutoff():
    # htgs's bound is the exact outline of the maxStdDev ellipsoid: every
    # pixel within the cutoff is drawn, none beyond it
    rig = cam.mono(np.eye(4), FOV, W, H, near=0.05, far=100)
    center = [0.2, 0.1, -3.0]
    q = Rotation.from_euler("xyz", [1.0, 0.3, -0.6]).as_quat()
    o, img = one("htgs", rig, center, [0.3, 0.15, 0.2], q, op=0.9)
    c, s, op = decoded(o, "htgs")
    hq = np.zeros(4)
    O.lib().orc_decode_quat_oct(o.words[3], 10, 12, O.dptr(hq))
    R = Rotation.from_quat(hq).as_matrix()
    cut = min(CUT, math.sqrt(2.0 * math.log(op / (0.5 / 255.0))))
    d = pixel_rays(rig).reshape(-1, 3)
    drawn = img[3] > 0
    # every pixel's ray minimum, by the bracketed search run on all rays at once
    A = (R / s).T
    f = lambda t: np.sum(((t[:, None] * d - c) @ A.T) ** 2, axis=1)
    lo, hi = np.zeros(len(d)), np.full(len(d), 40.0)
    g = (math.sqrt(5.0) - 1.0) / 2.0
    a, b = hi - g * (hi - lo), lo + g * (hi - lo)
    fa, fb = f(a), f(b)
    for _ in range(200):
        left = fa < fb
        hi = np.where(left, b, hi)
        lo = np.where(left, lo, a)
        na = np.where(left, hi - g * (hi - lo), b)
        nb = np.where(left, a, lo + g * (hi - lo))
        a, b = na, nb
        fa, fb = f(a), f(b)
    inside = (np.minimum(fa, fb) <= cut * cut * (1 - 1e-6)).reshape(H, W)
    assert inside.sum() > 100
    assert np.all(drawn[inside]), int((inside & ~drawn).sum())
    assert not np.any(drawn & ~inside & (read_z2(img, op) > cut * cut))


def test_two_splats_blend_back_to_front():
    rig = cam.mono(np.eye(4), FOV, W, H, near=0.05, far=100)
    view = O.view_from_rig(rig)
    far = O.project_ray("htgs", [0.0, 0.0, -5.0], [0.5] * 3, [0, 0, 0, 1], [0.0, 0.0, 1.0], 0.9, view, params())
    near = O.project_ray("htgs", [0.0, 0.0, -3.0], [0.2] * 3, [0, 0, 0, 1], [1.0, 0.0, 0.0], 0.8, view, params())
    words = np.array([list(far.words), list(near.words)], dtype=np.uint32)
    img = O.draw_ray("htgs", words, np.array([0, 1], dtype=np.uint32), 2, rig)
    y, x = H // 2, W // 2
    a_far = O.draw_ray("htgs", words, np.array([0], dtype=np.uint32), 1, rig)[:, y, x]
    a_near = O.draw_ray("htgs", words[1:], np.array([0], dtype=np.uint32), 1, rig)[:, y, x]
    # "over": near on top of far
    expect = a_near + (1.0 - a_near[3]) * a_far
    assert np.allclose(img[:, y, x], expect, rtol=1e-12, atol=1e-15)
    assert img[0, y, x] > img[2, y, x]

"""The oracle's 3dgut, htgs and raygs projection words (oracle/ray.c), pinned (no GPU).

Pins: closed forms of the unscented transform for an axis-aligned ellipsoid
seen head-on (sigma points c +- sqrt(3) s_i e_i project to +-sqrt(3) s_i f /
z, so the pixel covariance is diag(s_x^2, s_y^2) f^2 / z^2, the EWA result);
the exact outline of a sphere (a cone of half angle asin(r / z): radius f r /
sqrt(z^2 - r^2)); for any rotated, off-axis ellipsoid, tangency of the htgs
outline (every boundary pixel's ray touches the std-dev ellipsoid: min over
the ray of (x - c)^T Sigma^-1 (x - c) equals stdDev^2); rotation invariance
of a sphere; the MIP inflation's closed form for a sphere (scales sqrt(s^2 +
blur fp), opacity times s^2 / (s^2 + blur fp), fp the pixel footprint); the
centre and rotation codecs' round trips; the screen bound when the camera is
inside the ellipsoid; and the culls. raygs's quad frame by what defines
it: its two axes are orthonormal in the ellipsoid's metric (Sigma^-1) and
conjugate to the centre (axis^T Sigma^-1 c = 0: the plane through the centre
facing the eye), orthogonal to each other (the principal axes of the section),
the second the longer; c^2 the centre's squared Mahalanobis distance; for a
sphere both lengths are its scale; the frame rotation's columns are the
normalised axes and their cross product; the camera inside the support and a
quad off screen are culled.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from radix_amd import camera as cam

W, H = 640, 480
FOV = 50.0


def rig_and_view():
    rig = cam.mono(np.eye(4), FOV, W, H, near=0.05, far=100)
    return rig, O.view_from_rig(rig)


def focal(rig):
    P = np.asarray(rig.eyes[0].projection).reshape(4, 4).T  # column-major elements
    return W * 0.5 * P[0, 0], H * 0.5 * P[1, 1], P


def params(**kw):
    base = dict(blur_amount=0.0, pre_blur_amount=0.0, sort_bits=32, cull_keys=True, max_pixel_radius=1e6)
    base.update(kw)
    return O.params(**base)


def decode_bound(w0, w1):
    fx = lambda c: c / 65535.0 * 4.0 - 2.0
    ls = lambda c: 2.0 ** ((c - 1) / 2046.0 * 16.0 - 4.0)
    return (fx(w0 & 0xFFFF), fx(w0 >> 16), (w1 & 1023) / 1024.0 * math.pi, ls((w1 >> 10) & 0x7FF), ls(w1 >> 21))


@pytest.mark.parametrize("sx,sy", [(0.3, 0.3), (0.4, 0.15), (0.1, 0.25)])
def test_3dgut_axis_aligned_head_on(sx, sy):
    rig, view = rig_and_view()
    fx, fy, _ = focal(rig)
    z = 6.0
    o = O.project_ray("3dgut", [0, 0, -z], [sx, sy, 0.2], [0, 0, 0, 1], [0.5, 0.5, 0.5], 0.9, view, params())
    std = math.sqrt(8.0)
    nx, ny, angle, s1, s2 = o.bound[0]
    assert o.visible and o.eye_visible[0]
    assert abs(nx) < 1e-12 and abs(ny) < 1e-12
    big, small = (sx * fx, sy * fy) if sx * fx >= sy * fy else (sy * fy, sx * fx)
    assert s1 == pytest.approx(std * big / z, rel=1e-12)
    assert s2 == pytest.approx(std * small / z, rel=1e-12)
    if sx != sy:
        assert angle == pytest.approx(0.0 if sx * fx > sy * fy else math.pi / 2, abs=1e-12)
    # the words: outward-rounded scales cover the bound, within one code
    d = decode_bound(o.words[6], o.words[7])
    step = 2.0 ** (16.0 / 2046.0)
    assert s1 <= d[3] <= s1 * step * 1.01 and s2 <= d[4] <= s2 * step * 1.01


def test_htgs_sphere_outline_closed_form():
    rig, view = rig_and_view()
    fx, fy, _ = focal(rig)
    assert fx == pytest.approx(fy, rel=1e-12)
    z, s = 5.0, 0.5
    std = math.sqrt(8.0)
    o = O.project_ray("htgs", [0, 0, -z], [s, s, s], [0, 0, 0, 1], [0.5, 0.5, 0.5], 0.9, view, params())
    r = std * s
    expected = fx * r / math.sqrt(z * z - r * r)
    nx, ny, angle, s1, s2 = o.bound[0]
    assert s1 == pytest.approx(expected, rel=1e-10) and s2 == pytest.approx(expected, rel=1e-10)
    assert abs(nx) < 1e-12 and abs(ny) < 1e-12


def tangency(center, scales, quat, std, bound, P):
    """Largest |m - std^2| over the htgs outline's boundary pixels, m the
    smallest (x - c)^T Sigma^-1 (x - c) along each pixel's ray."""
    R = Rotation.from_quat(quat).as_matrix()
    A = R @ np.diag(1.0 / np.asarray(scales) ** 2) @ R.T
    c = np.asarray(center, dtype=np.float64)
    nx, ny, angle, s1, s2 = bound
    worst = 0.0
    for t in np.linspace(0, 2 * math.pi, 64, endpoint=False):
        u = np.array([s1 * math.cos(t), s2 * math.sin(t)])
        rot = np.array([[math.cos(angle), -math.sin(angle)], [math.sin(angle), math.cos(angle)]])
        px = rot @ u + np.array([nx * W / 2, ny * H / 2])
        ndc = px / np.array([W / 2, H / 2])
        d = np.array([ndc[0] / P[0, 0], ndc[1] / P[1, 1], -1.0])
        m = c @ A @ c - (d @ A @ c) ** 2 / (d @ A @ d)
        worst = max(worst, abs(m - std * std) / (std * std))
    return worst


@pytest.mark.parametrize("seed", range(6))
def test_htgs_outline_is_tangent(seed):
    rig, view = rig_and_view()
    _, _, P = focal(rig)
    rng = np.random.default_rng(seed)
    center = [rng.uniform(-1.5, 1.5), rng.uniform(-1.0, 1.0), -rng.uniform(4.0, 9.0)]
    scales = rng.uniform(0.05, 0.4, 3)
    quat = Rotation.random(random_state=seed).as_quat()
    o = O.project_ray("htgs", center, scales, quat, [0.5, 0.5, 0.5], 0.9, view, params())
    assert o.eye_visible[0]
    assert tangency(center, scales, quat, math.sqrt(8.0), o.bound[0], P) < 1e-9


@pytest.mark.parametrize("method,rel", [("htgs", 1e-9), ("3dgut", 5e-3)])
def test_sphere_rotation_invariant(method, rel):
    # the outline of a sphere does not depend on its rotation; the unscented
    # transform's sigma points do, by a second-order term in s / z
    _, view = rig_and_view()
    a = O.project_ray(method, [0.7, -0.3, -5], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    q = Rotation.from_euler("xyz", [0.3, 1.1, -0.7]).as_quat()
    b = O.project_ray(method, [0.7, -0.3, -5], [0.2] * 3, q, [0.5] * 3, 0.9, view, params())
    for k in (0, 1, 3, 4):
        assert a.bound[0][k] == pytest.approx(b.bound[0][k], rel=rel, abs=1e-12)


def test_mip_inflation_sphere_closed_form():
    rig, view = rig_and_view()
    _, _, P = focal(rig)
    z, s, blur, op = 7.0, 0.01, 0.3, 0.8
    o = O.project_ray("3dgut", [0, 0, -z], [s, s, s], [0, 0, 0, 1], [0.5] * 3, op, view, params(blur_amount=blur))
    fp = (2.0 / (H * P[1, 1])) ** 2 * z * z
    assert o.head_scales[0] == pytest.approx(math.sqrt(s * s + blur * fp), rel=1e-12)
    assert o.head_opacity == pytest.approx(op * s * s / (s * s + blur * fp), rel=1e-12)
    # the stored opacity byte and the half-float log scales of the head words
    assert (o.words[5] >> 16) & 0xFF == max(1, int(math.floor(o.head_opacity * 128 + 0.5)))
    assert O.lib().orc_decode_half(o.words[4] & 0xFFFF) == pytest.approx(math.log(o.head_scales[0]), abs=2e-3)


def test_centre_codec_round_trip():
    rng = np.random.default_rng(3)
    import ctypes as C
    for _ in range(200):
        c = rng.normal(size=3) * rng.uniform(0.1, 50.0)
        c = np.ascontiguousarray(c)
        d, dist, low = C.c_uint32(), C.c_uint32(), C.c_uint32()
        O.lib().orc_encode_center(O.dptr(c), C.byref(d), C.byref(dist), C.byref(low))
        assert low.value & 0x00FFFFFF == 0
        out = np.zeros(3)
        O.lib().orc_decode_center(d.value, dist.value, low.value, O.dptr(out))
        assert np.linalg.norm(out) == pytest.approx(np.linalg.norm(c), rel=1e-7)
        cosang = out @ c / (np.linalg.norm(out) * np.linalg.norm(c))
        assert math.acos(min(1.0, cosang)) < 4e-6  # 20-bit octahedral codes
    # +z: the octahedron's centre code
    c = np.array([0.0, 0.0, 3.0])
    d, dist, low = C.c_uint32(), C.c_uint32(), C.c_uint32()
    O.lib().orc_encode_center(O.dptr(c), C.byref(d), C.byref(dist), C.byref(low))
    q = 524288  # floor(0.5 * 0xfffff + 0.5)
    assert d.value == (q >> 4) | ((q >> 4) << 16) and low.value == ((q & 15) << 24) | ((q & 15) << 28)


def test_quat_oct32_matches_the_cpu_codec_round_trip():
    errs = []
    for k in range(300):
        q = Rotation.random(random_state=k).as_quat()
        q = np.ascontiguousarray(q)
        code = O.lib().orc_encode_quat_oct32_gpu(O.dptr(q))
        # the shader's encoder equals the CPU codec (encoding.ts:85-98), pinned on its own
        assert code == O.lib().orc_encode_quat_oct(O.dptr(q), 10, 12)
        out = np.zeros(4)
        O.lib().orc_decode_quat_oct(code, 10, 12, O.dptr(out))
        # the rotation angle between them: 10-bit octahedral axes (cells of
        # about 2e-3 rad), a 12-bit angle (7.7e-4 rad)
        dot = abs(out @ q) / np.linalg.norm(out)
        errs.append(2 * math.acos(min(1.0, dot)))
    assert max(errs) < 8e-3 and np.mean(errs) < 3e-3
    assert O.lib().orc_encode_quat_oct32_gpu(O.dptr(np.zeros(4))) == 0


def test_camera_inside_gets_the_screen_bound():
    rig, view = rig_and_view()
    # the camera inside the std-dev ellipsoid: htgs has no outline
    o = O.project_ray("htgs", [0, 0, -1.0], [0.5, 0.5, 0.5], [0, 0, 0, 1], [0.5] * 3, 0.9, view,
                      params(max_pixel_radius=512))
    nx, ny, angle, s1, s2 = o.bound[0]
    half = math.hypot(W, H) * 0.5
    # min(maxPixelRadius, |size| / 2) = 400, then clipped to the 640 x 480 screen
    assert min(512.0, half) == 400.0
    assert (nx, ny, angle, s1, s2) == (0.0, 0.0, 0.0, W / 2, H / 2)


def test_culls():
    _, view = rig_and_view()
    # behind the camera: the head words stay, the eye's are zero, not visible
    o = O.project_ray("3dgut", [0, 0, 3.0], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    assert not o.visible and o.words[0] != 0 and o.words[6] == 0 and o.words[7] == 0
    assert o.key == O.key_invalid(32)  # cullKeys: not visible
    o2 = O.project_ray("3dgut", [0, 0, 3.0], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params(cull_keys=False))
    assert o2.key == O.key_compute(3.0, 32)
    # below minAlpha: nothing at all
    o3 = O.project_ray("htgs", [0, 0, -3.0], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.001, view, params())
    assert o3.invisible and not any(o3.words)
    # a visible splat's key is the head distance's
    o4 = O.project_ray("htgs", [0, 0, -3.0], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    assert o4.visible and o4.key == O.key_compute(3.0, 32)


def sigma_inv(scales, quat):
    R = Rotation.from_quat(quat).as_matrix()
    return R @ np.diag(1.0 / np.asarray(scales, dtype=np.float64) ** 2) @ R.T


@pytest.mark.parametrize("seed", range(8))
def test_raygs_frame_invariants(seed):
    _, view = rig_and_view()
    rng = np.random.default_rng(100 + seed)
    center = np.array([rng.uniform(-1.5, 1.5), rng.uniform(-1.0, 1.0), -rng.uniform(3.0, 9.0)])
    scales = rng.uniform(0.05, 0.4, 3)
    quat = Rotation.random(random_state=100 + seed).as_quat()
    o = O.project_ray("raygs", center, scales, quat, [0.5] * 3, 0.9, view, params())
    assert o.eye_visible[0] and o.visible
    A = sigma_inv(scales, quat)
    a0, a1 = np.array(o.frame_axes[0][0][:]), np.array(o.frame_axes[0][1][:])
    G = np.array([[a0 @ A @ a0, a0 @ A @ a1], [a1 @ A @ a0, a1 @ A @ a1]])
    assert np.allclose(G, np.eye(2), atol=1e-10)
    assert abs(a0 @ A @ center) < 1e-9 * np.linalg.norm(A @ center) and abs(a1 @ A @ center) < 1e-9 * np.linalg.norm(A @ center)
    assert abs(a0 @ a1) < 1e-10 * np.linalg.norm(a0) * np.linalg.norm(a1)
    lx, ly = o.frame_lengths[0][:]
    assert lx == pytest.approx(np.linalg.norm(a0), rel=1e-12) and ly >= lx
    assert o.frame_c2[0] == pytest.approx(center @ A @ center, rel=1e-12)
    # the words: c^2 as float bits, the lengths as half-float logs, the frame rotation
    w = list(o.words)
    assert np.float32(np.array([w[5]], dtype=np.uint32).view(np.float32)[0]) == np.float32(center @ A @ center)
    assert O.lib().orc_decode_half(w[6] & 0xFFFF) == pytest.approx(math.log(lx), abs=2e-3)
    assert O.lib().orc_decode_half(w[6] >> 16) == pytest.approx(math.log(ly), abs=2e-3)
    q = np.zeros(4)
    O.lib().orc_decode_quat_oct(w[4], 10, 12, O.dptr(q))
    M = Rotation.from_quat(q).as_matrix()
    n0, n1 = a0 / lx, a1 / ly
    for col, ref in ((0, n0), (1, n1), (2, np.cross(n0, n1))):
        assert np.linalg.norm(M[:, col] - ref) < 1e-2  # the codec's 10-bit axis, 12-bit angle


def test_raygs_sphere():
    _, view = rig_and_view()
    c, s = np.array([0.4, -0.2, -5.0]), 0.3
    o = O.project_ray("raygs", c, [s, s, s], [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    assert o.frame_lengths[0][0] == pytest.approx(s, rel=1e-12) and o.frame_lengths[0][1] == pytest.approx(s, rel=1e-12)
    assert o.frame_c2[0] == pytest.approx(c @ c / (s * s), rel=1e-12)
    # the quad faces the eye: its normal along the centre direction
    a0, a1 = np.array(o.frame_axes[0][0][:]), np.array(o.frame_axes[0][1][:])
    nrm = np.cross(a0, a1)
    assert abs(abs(nrm @ c) / (np.linalg.norm(nrm) * np.linalg.norm(c)) - 1.0) < 1e-12


def test_raygs_culls():
    _, view = rig_and_view()
    # the camera inside the support (c^2 <= stdDev^2): no eye words, head words kept
    o = O.project_ray("raygs", [0, 0, -1.0], [0.5] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    assert not o.visible and o.words[0] != 0 and o.words[4] == 0 and o.words[5] == 0
    # far off to the side: every corner beyond the right edge
    o = O.project_ray("raygs", [40.0, 0, -5.0], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    assert not o.visible and o.words[5] == 0
    # head words: opacity byte and the centre's low bits share word 3
    o = O.project_ray("raygs", [0.3, 0.2, -4.0], [0.2] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, view, params())
    assert o.visible and (o.words[3] & 0xFF) == max(1, int(math.floor(o.head_opacity * 128 + 0.5)))
    assert o.words[3] & 0x00FFFF00 == 0


def from_cov(c6):
    s, q = np.zeros(3), np.zeros(4)
    O.lib().orc_from_cov(O.dptr(np.ascontiguousarray(c6, dtype=np.float64)), O.dptr(s), O.dptr(q))
    return s, q


def test_from_cov_single_rotation_known_answer():
    # symEigens (math.tsl.ts:46-113) on xx = yy = 2, xy = 1, zz = 1/2: one
    # rotation in the (0, 1) plane, phi = atan2(2, 0) / 2 = pi / 4, leaves
    # diag(1, 3, 1/2) with rows (1, -1, 0) / sqrt2, (1, 1, 0) / sqrt2, (0, 0, 1);
    # sorted descending: values 3, 1, 1/2, v0 = (1, 1, 0) / sqrt2, v1 = (1, -1, 0) / sqrt2,
    # and v0 x v1 = (0, 0, -1) flips v2 to (0, 0, -1). basisToQuat (z.z <= 0,
    # y.y - x.x <= 0): four = 2 + 2a (a = 1 / sqrt2), q = (1 + a, a, 0, 0) / |.|:
    # the half turn about (cos 22.5, sin 22.5, 0), which takes x to v0 and z to -z.
    s, q = from_cov([2.0, 2.0, 0.5, 1.0, 0.0, 0.0])
    a = math.sqrt(0.5)
    assert s == pytest.approx([math.sqrt(3.0), 1.0, math.sqrt(0.5)], rel=1e-15)
    assert q == pytest.approx(np.array([1 + a, a, 0.0, 0.0]) / math.sqrt(2 + math.sqrt(2)), abs=1e-15)
    R = Rotation.from_quat(q).as_matrix()
    assert R[:, 0] == pytest.approx([a, a, 0], abs=1e-15) and R[:, 2] == pytest.approx([0, 0, -1], abs=1e-15)


def test_from_cov_round_trip():
    # check-core.ts:429-435's round trip covFromScaleQuat(fromCov(c)), to the
    # GPU solver's stopping rule: it drops off-diagonal entries within 1e-6
    # of the trace (math.tsl.ts:49; the CPU twin of check-core stops at 1e-9)
    rng = np.random.default_rng(5)
    cov, back = np.zeros(6), np.zeros(6)
    for _ in range(2000):
        scales = np.exp(rng.random(3) * 9 - 6)
        quat = rng.random(4) * 2 - 1
        quat /= np.linalg.norm(quat)
        O.lib().orc_cov_from_scale_quat(O.dptr(scales), O.dptr(quat), O.dptr(cov))
        s, q = from_cov(cov)
        assert list(s) == sorted(s, reverse=True) and q[3] >= 0
        O.lib().orc_cov_from_scale_quat(O.dptr(np.ascontiguousarray(s)), O.dptr(q), O.dptr(back))
        assert np.max(np.abs(back - cov)) <= 1e-6 * np.sum(cov[:3])
        # the scales are the square roots of the eigenvalues
        S = np.array([[cov[0], cov[3], cov[4]], [cov[3], cov[1], cov[5]], [cov[4], cov[5], cov[2]]])
        ev = np.sort(np.linalg.eigvalsh(S))[::-1]
        assert s ** 2 == pytest.approx(ev, abs=1e-6 * np.sum(cov[:3]))


@pytest.mark.parametrize("method", ["3dgut", "htgs", "raygs"])
def test_non_uniform_transform_through_the_covariance(method):
    # transformSplat's general case (splatTransform.tsl.ts:154-165): a sphere of
    # radius s under diag(2, 1, 1/2) has the head covariance diag(4, 1, 1/4) s^2,
    # whose fromCov needs no rotation: scales (2s, s, s/2), identity. The words
    # equal those of that ellipsoid under the pure translation.
    s, t = 0.05, [0.4, -0.2, -6.0]
    rig = cam.mono(np.eye(4), FOV, W, H, near=0.05, far=100)
    M = cam.translation(t)
    M[:3, :3] = np.diag([2.0, 1.0, 0.5])
    a = O.project_ray(method, [0, 0, 0], [s] * 3, [0, 0, 0, 1], [0.5] * 3, 0.9, O.view_from_rig(rig, M), params())
    b = O.project_ray(method, [0, 0, 0], [2 * s, s, s / 2], [0, 0, 0, 1], [0.5] * 3, 0.9,
                      O.view_from_rig(rig, cam.translation(t)), params())
    assert a.visible and list(a.words) == list(b.words)
    assert list(a.head_scales) == pytest.approx(list(b.head_scales), rel=1e-15)

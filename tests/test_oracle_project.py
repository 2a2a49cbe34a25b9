"""The oracle's ewa projection, pinned (no GPU).

Pins: an independent numpy EWA (J W Sigma W^T J^T, eigenvalues by
numpy.linalg.eigvalsh) for the bound axes, as renderTests.ts:242-277 checks
the GPU against a CPU EWA; the culling known answers of renderTests.ts:302-321
and the sortGuard ones of renderTests.ts:341-367; keys from
SortKeys.compute of the true distance (renderTests.ts:323-339); the word
codecs' closed forms and round trips; and stereo with the head at an eye
against mono from that eye (renderTests.ts:221-240).
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O
from radix_amd import camera as cam
from tests.helpers import Rng

W, H = 128, 64


def planes_of(centers, scales, quats=None, opacity=None, rgb=None):
    n = len(centers)
    p = np.zeros((14, n), dtype=np.float32)
    p[0:3] = np.asarray(centers, dtype=np.float64).T
    p[3:6] = np.log(np.broadcast_to(np.asarray(scales, dtype=np.float64).reshape(n, -1), (n, 3))).T
    q = np.tile([0, 0, 0, 1.0], (n, 1)) if quats is None else np.asarray(quats)
    p[6:10] = q.T
    p[10] = 1.0 if opacity is None else np.asarray(opacity)
    p[11:14] = 1.0 if rgb is None else np.asarray(rgb).T
    return p


def test_culling_known_answers():
    # renderTests.ts:302-321: ahead; behind; far off screen; below minAlpha
    rig = cam.mono(np.eye(4), 60, W, H, near=0.1, far=100)
    p = planes_of([[0, 0, -3], [0, 0, 3], [40, 0, -3], [0, 0, -3]], [0.1] * 4, opacity=[1, 1, 1, 0.001])
    for cull_keys, expected in ((True, [False, True, True, True]), (False, [False, False, False, True])):
        out = O.project_f32(p, 0, O.view_from_rig(rig), O.params(sort_bits=32, cull_keys=cull_keys))
        assert list(out["visible"][0] == 0) == [False, True, True, True]
        assert list(out["keys"] == O.key_invalid(32)) == expected


def test_sort_guard_known_answers():
    # renderTests.ts:341-367: ahead, just past the right edge, far to the side, behind
    rig = cam.mono(np.eye(4), 60, W, H, near=0.1, far=100)
    p = planes_of([[0, 0, -4], [5.5, 0, -4], [20, 0, -4], [0, 0, 4]], [0.05] * 4)
    view = O.view_from_rig(rig)
    for guard, expected in ((0.0, [True] * 4), (0.5, [True, True, False, False])):
        out = O.project_f32(p, 0, view, O.params(sort_bits=24, cull_keys=False, sort_guard=guard))
        assert list(out["keys"] != O.key_invalid(24)) == expected, guard


def random_scene(n, seed):
    """Random splats in front of a camera at the origin (renderUtils.ts:49-62 ranges)."""
    r = Rng(seed)
    u = r.random(n * 14).reshape(n, 14)
    centers = np.stack([(u[:, 0] - 0.5) * 4, (u[:, 1] - 0.5) * 3, -2 - u[:, 2] * 6], 1)
    scales = np.stack([0.01 + u[:, 3] * 0.1, 0.01 + u[:, 4] * 0.1, 0.01 + u[:, 5] * 0.05], 1)
    axis = u[:, 6:9] - 0.5
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = u[:, 9] * 2 * math.pi
    quats = np.concatenate([axis * np.sin(ang / 2)[:, None], np.cos(ang / 2)[:, None]], 1)
    return centers, scales, quats, 0.2 + u[:, 13] * 0.8


def test_ewa_axes_match_numpy_ewa():
    # blur 0 (renderTests.ts:248-274): bound axes = stdDev sqrt(eigenvalues of J W S W^T J^T)
    centers, scales, quats, op = random_scene(300, 3)
    world = cam.look_at_world([0.3, -0.2, 0.5], [0, 0, -5])
    rig = cam.mono(world, 60, W, H, near=0.1, far=100)
    p = planes_of(centers, scales, quats, op)
    prm = O.params(sort_bits=32, blur_amount=0.0)
    v = O.view_from_rig(rig)
    checked = 0
    for i in range(len(centers)):
        s = O.OrcSplat()
        s.center[:] = list(map(float, p[0:3, i]))
        cov = np.zeros(6)
        L = O.lib()
        sc = np.exp(p[3:6, i].astype(np.float64))
        L.orc_cov_from_scale_quat(O.dptr(np.ascontiguousarray(sc)), O.dptr(np.ascontiguousarray(p[6:10, i].astype(np.float64))), O.dptr(cov))
        s.cov[:] = list(cov)
        s.rgb[:] = [1, 1, 1]
        s.opacity = float(p[10, i])
        out = O.OrcOut()
        L.orc_project_ewa(s, v, prm, out)
        e = out.eye[0]
        if not e.visible:
            continue
        # independent numpy EWA
        view = np.linalg.inv(world)
        c = view[:3, :3] @ p[0:3, i].astype(np.float64) + view[:3, 3]
        R = Rotation.from_quat(p[6:10, i].astype(np.float64)).as_matrix()
        M = view[:3, :3] @ R @ np.diag(sc)
        S = M @ M.T
        P = rig.eyes[0].projection
        fx, fy = W / 2 * P[0, 0], H / 2 * P[1, 1]
        J = np.array([[fx / c[2], 0, -fx * c[0] / c[2] ** 2], [0, fy / c[2], -fy * c[1] / c[2] ** 2]])
        lam = np.linalg.eigvalsh(J @ S @ J.T)
        major, minor = math.sqrt(8 * lam[1]), math.sqrt(8 * max(lam[0], 0))
        assert abs(e.scale1 / major - 1) < 1e-9 and abs(e.scale2 / minor - 1) < 1e-7, i
        # the stored depth is view z, the ndc the projected centre
        assert abs(e.view_z - c[2]) < 1e-12
        clip = P @ np.append(c, 1)
        assert abs(e.ndc[0] - clip[0] / clip[3]) < 1e-12
        checked += 1
    assert checked > 200


def test_opacity_compensation_and_conic():
    centers, scales, quats, op = random_scene(100, 8)
    rig = cam.mono(np.eye(4), 60, W, H, near=0.1, far=100)
    p = planes_of(centers, scales, quats, op)
    out0 = O.project_f32(p, 0, O.view_from_rig(rig), O.params(sort_bits=32, blur_amount=0.0))
    out1 = O.project_f32(p, 0, O.view_from_rig(rig), O.params(sort_bits=32, blur_amount=0.3))
    for i in range(100):
        if not (out0["visible"][0, i] and out1["visible"][0, i]):
            continue
        A, B, Cc = out0["conic"][0, :, i]
        det0 = 1.0 / (A * Cc - B * B)  # det of the unblurred cov2D
        a0, d0 = Cc * det0, A * det0
        b0 = -B * det0
        det1 = (a0 + 0.3) * (d0 + 0.3) - b0 * b0
        o = float(p[10, i])  # the stored float32 opacity
        assert abs(out1["opacity"][0, i] - o * math.sqrt(det0 / det1)) < 1e-12 * o + 1e-15
        A1, B1, C1 = out1["conic"][0, :, i]
        assert np.allclose([A1, B1, C1], [(d0 + 0.3) / det1, -b0 / det1, (a0 + 0.3) / det1], rtol=1e-9)


def test_keys_are_sortkeys_of_true_distance():
    centers, scales, quats, op = random_scene(500, 13)
    eye = np.array([0.03, 0.0, 0.0])
    rig = cam.mono(cam.translation(eye), 60, W, H, near=0.1, far=100)
    p = planes_of(centers, scales, quats, op)
    for bits in (16, 24, 32):
        out = O.project_f32(p, 0, O.view_from_rig(rig), O.params(sort_bits=bits, cull_keys=False))
        exp = [O.key_compute(float(np.linalg.norm(p[0:3, i].astype(np.float64) - eye)), bits) for i in range(500)]
        assert list(out["keys"]) == exp


def test_word_codecs_closed_forms():
    rig = cam.mono(np.eye(4), 60, W, H, near=0.1, far=100)
    p = planes_of([[0, 0, -1], [0, 0, -2.5]], [0.01, 0.02], opacity=[1.0, 0.5], rgb=[[1, 0.5, 0.25], [0, 0, 0]])
    out = O.project_f32(p, 0, O.view_from_rig(rig), O.params(sort_bits=32, blur_amount=0.0))
    w = out["words"][0]
    # on-axis: ndc 0 in both 16-bit halves
    assert w[0, 0] == 0x8000 | (0x8000 << 16)
    # depth20: float bits of -z above 2^-14, rounded to 20 bits: z = -1 -> 14 << 15
    assert w[0, 1] & 0xFFFFF == 14 << 15
    z = -2.5
    code = w[1, 1] & 0xFFFFF
    back = np.array([(code << 8) + (113 << 23)], dtype=np.uint32).view(np.float32)[0]
    assert abs(back - 2.5) <= 2.5 * 2 ** -15
    # rgb (1, 0.5, 0.25) in extRgb
    assert w[0, 2] == O.lib().orc_encode_ext_rgb(1.0, 0.5, 0.25)
    # opacity byte: round(opacity * 128), >= 1
    assert w[0, 3] & 0xFF == 128 and w[1, 3] & 0xFF == 64
    # log2 scale codes: decode(code) within half a step of the scale
    for i in range(2):
        s1 = (w[i, 3] >> 8) & 0xFFF
        dec = 2 ** ((s1 - 1) / 4094 * 16 - 4)
        exact = math.sqrt(8) * (0.01 if i == 0 else 0.02) * (W / 2 * rig.eyes[0].projection[0, 0]) / abs([-1, -2.5][i])
        assert abs(math.log2(dec) - math.log2(exact)) <= 16 / 4094 / 2 + 1e-9


def test_stereo_eye_equals_mono_from_that_eye():
    centers, scales, quats, op = random_scene(400, 21)
    p = planes_of(centers, scales, quats, op)
    world = cam.look_at_world([0, 0.1, 0.2], [0, 0, -4])
    st = cam.stereo(world, 0.064, 60, W, H, near=0.1, far=100)
    prm = O.params(sort_bits=32)
    s = O.project_f32(p, 0, O.view_from_rig(st), prm)
    for v in range(2):
        eye_world = np.linalg.inv(st.eyes[v].view)
        m = O.project_f32(p, 0, O.view_from_rig(cam.mono(eye_world, 60, W, H, near=0.1, far=100)), prm)
        diff = np.abs(s["words"][v].astype(np.int64) - m["words"][0].astype(np.int64))
        assert diff.max() <= 32  # renderTests.ts:229-236 allows 32 per word
        assert np.array_equal(s["visible"][v], m["visible"][0])

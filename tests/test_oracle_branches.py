"""The oracle's LoD, depth-of-field and orthographic branches, pinned by
closed forms (no GPU).

An isotropic splat of scale sigma straight ahead at depth D has, under a
pinhole of focal length f pixels, the 2D covariance A = (f sigma / D)^2 on
both axes and no cross term, so each branch reduces to scalar formulas
derived from the reference's text:
- LoD (projection.tsl.ts:42-64, storedOpacity :72): an opacity o > 1 is
  stretched to o' = min(4 o - 3, 5); with lodInflate the scales grow by
  exp((o'^2 - 1) / e)^(1/3) and the opacity becomes 1, without it the
  opacity stays o' and the bound's std devs become maxStdDev + 0.7 (o' - 1);
  the opacity byte stores (o + 3) / 4 above 1.
- Depth of field (projection.tsl.ts:117-125): the blur becomes the circle
  of confusion's square, coc = |D - F| apertureTan f / D (f without / D for
  an orthographic eye), clamped to [blurAmount, maxPixelRadius^2].
- Orthographic J (projections.tsl.ts:111): J = diag(fx, fy), so the 2D
  covariance is fx^2 Sxx, fy^2 Syy, fx fy Sxy whatever the splat's depth
  and position, and the ndc is P (x, y) with w = 1.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from radix_amd import camera as cam

W, H = 256, 128
SQRT8 = math.sqrt(8.0)


def one(center, sigma, quat, opacity, rig, **prm):
    """orc_project_ewa of one splat; the covariance is R diag(sigma^2) R^T in numpy."""
    x, y, z, w = np.asarray(quat, dtype=np.float64) / np.linalg.norm(quat)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    S = R @ np.diag(np.asarray(sigma, dtype=np.float64) ** 2) @ R.T
    s = O.OrcSplat()
    s.center[:] = [float(v) for v in center]
    s.cov[:] = [S[0, 0], S[1, 1], S[2, 2], S[0, 1], S[0, 2], S[1, 2]]
    s.rgb[:] = [0.5, 0.5, 0.5]
    s.opacity = float(opacity)
    out = O.OrcOut()
    O.lib().orc_project_ewa(s, O.view_from_rig(rig), O.params(sort_bits=32, **prm), out)
    return out


def focal(rig):
    return H / 2 * rig.eyes[0].projection[1, 1]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("o_in", [1.5, 2.0, 3.0])
@pytest.mark.parametrize("inflate_on", [True, False])
def test_lod_stretch_and_inflate(o_in, inflate_on):
    rig = cam.mono(np.eye(4), 60, W, H)
    f, D, sigma = focal(rig), 4.0, 0.01
    o2 = min(4 * o_in - 3, 5.0)
    if inflate_on:
        inflate, op_in, std = math.exp((o2 * o2 - 1) / math.e) ** (1 / 3), 1.0, SQRT8
    else:
        inflate, op_in, std = 1.0, o2, SQRT8 + 0.7 * (o2 - 1)
    A = (f * sigma * inflate / D) ** 2
    op = op_in * A / (A + 0.3)
    e = one([0, 0, -D], [sigma] * 3, [0, 0, 0, 1], o_in, rig, lod_inflate=inflate_on).eye[0]
    assert e.visible
    assert rel(e.conic[0], 1 / (A + 0.3)) < 1e-12 and rel(e.conic[2], 1 / (A + 0.3)) < 1e-12
    assert abs(e.conic[1]) < 1e-15 / (A + 0.3)
    assert rel(e.opacity, op) < 1e-12
    assert rel(e.scale1, std * math.sqrt(A + 0.3)) < 1e-12 and rel(e.scale2, std * math.sqrt(A + 0.3)) < 1e-12
    stored = (op + 3) / 4 if op > 1 else op
    assert e.words[3] & 0xFF == max(1, math.floor(min(stored * 128, 255) + 0.5))


def test_lod_opacity_at_one_and_below_is_untouched():
    rig = cam.mono(np.eye(4), 60, W, H)
    f, D, sigma = focal(rig), 4.0, 0.01
    A = (f * sigma / D) ** 2
    for o_in in (1.0, 0.4):
        e = one([0, 0, -D], [sigma] * 3, [0, 0, 0, 1], o_in, rig).eye[0]
        assert rel(e.opacity, o_in * A / (A + 0.3)) < 1e-12
        assert rel(e.scale1, SQRT8 * math.sqrt(A + 0.3)) < 1e-12


@pytest.mark.parametrize("depth,focus,aperture,sigma", [(4.0, 4.0, 0.02, 0.02), (4.0, 2.0, 0.02, 0.02),
                                                         (9.0, 2.0, 0.01, 0.02), (3.0, 40.0, 2.0, 1.0)])
def test_depth_of_field_blur(depth, focus, aperture, sigma):
    # the last case's circle of confusion passes maxPixelRadius: blur 512^2
    rig = cam.mono(np.eye(4), 60, W, H)
    f = focal(rig)
    coc = abs(depth - focus) * aperture * f / depth
    blur = min(max(coc * coc, 0.3), 512.0 ** 2)
    A = (f * sigma / depth) ** 2
    e = one([0, 0, -depth], [sigma] * 3, [0, 0, 0, 1], 1.0, rig, focal_distance=focus, aperture_tan=aperture).eye[0]
    assert e.visible
    assert rel(e.conic[0], 1 / (A + blur)) < 1e-12
    assert rel(e.opacity, A / (A + blur)) < 1e-12
    assert rel(e.scale1, min(512.0, SQRT8 * math.sqrt(A + blur))) < 1e-12


def test_depth_of_field_needs_both_params():
    rig = cam.mono(np.eye(4), 60, W, H)
    f, D, sigma = focal(rig), 4.0, 0.02
    A = (f * sigma / D) ** 2
    for fd, ap in ((2.0, 0.0), (0.0, 0.02)):
        e = one([0, 0, -D], [sigma] * 3, [0, 0, 0, 1], 1.0, rig, focal_distance=fd, aperture_tan=ap).eye[0]
        assert rel(e.conic[0], 1 / (A + 0.3)) < 1e-12


def ortho_rig():
    """An orthographic eye whose pixels are not square (fx != fy): x in [-3, 3], y in [-1, 1]."""
    rig = cam.mono_ortho(np.eye(4), 1.0, W, H, near=0.1, far=100)
    rig.eyes[0].projection = cam.orthographic(-3.0, 3.0, 1.0, -1.0, 0.1, 100.0)
    return rig


@pytest.mark.parametrize("theta", [0.0, 0.4, 1.1])
def test_orthographic_j_is_depth_free(theta):
    rig = ortho_rig()
    P = rig.eyes[0].projection
    fx, fy = W / 2 * P[0, 0], H / 2 * P[1, 1]
    sx, sy, sz = 0.05, 0.02, 0.03
    c, s = math.cos(theta), math.sin(theta)
    Sxx, Syy, Sxy = sx * sx * c * c + sy * sy * s * s, sx * sx * s * s + sy * sy * c * c, (sx * sx - sy * sy) * c * s
    a, d, b = fx * fx * Sxx + 0.3, fy * fy * Syy + 0.3, fx * fy * Sxy
    det = a * d - b * b
    q = [0, 0, math.sin(theta / 2), math.cos(theta / 2)]
    for center in ([0, 0, -1], [0.7, -0.4, -3], [-1.5, 0.6, -50]):
        out = one(center, [sx, sy, sz], q, 0.8, rig)
        e = out.eye[0]
        assert e.visible, center
        assert rel(e.a, a) < 1e-12 and rel(e.d, d) < 1e-12
        assert abs(e.b - b) < 1e-12 * max(a, d)
        assert rel(e.conic[0], d / det) < 1e-10 and rel(e.conic[2], a / det) < 1e-10
        assert rel(e.opacity, 0.8 * math.sqrt((a - 0.3) * (d - 0.3) - b * b) / math.sqrt(det)) < 1e-10
        assert abs(e.ndc[0] - center[0] * P[0, 0]) < 1e-15 and abs(e.ndc[1] - center[1] * P[1, 1]) < 1e-15
        assert e.view_z == center[2]


def test_orthographic_depth_of_field_has_no_perspective_divide():
    rig = ortho_rig()
    P = rig.eyes[0].projection
    fx, fy = W / 2 * P[0, 0], H / 2 * P[1, 1]
    sigma = 0.02
    for depth in (2.0, 8.0):
        # the circle of confusion's pixels are vertical ones (renderSize.y / 2 P11)
        coc = abs(depth - 5.0) * 0.05 * fy
        blur = min(max(coc * coc, 0.3), 512.0 ** 2)
        e = one([0, 0, -depth], [sigma] * 3, [0, 0, 0, 1], 1.0, rig, focal_distance=5.0, aperture_tan=0.05).eye[0]
        assert rel(e.conic[0], 1 / ((fx * sigma) ** 2 + blur)) < 1e-12, depth
        assert rel(e.conic[2], 1 / ((fy * sigma) ** 2 + blur)) < 1e-12, depth

"""The oracle's keys-only pass (SplatProjection.writeKeys), pinned (no GPU).

The keys program keeps a stored splat in the splat form
(SplatProjector.ts:152-194 with no wanted form), so inSortGuard's extent is
its largest head-space scale (projection.tsl.ts:145): the scales times the
uniform scale of a rotation-times-scale transform, or the square root of the
largest eigenvalue of M Sigma M^T otherwise (transformSplat,
splatTransform.tsl.ts:154-165; fromCov, splats.tsl.ts:189-198). The full
ewa projection's key uses the covariance form, sqrt of the largest diagonal
entry, which is never larger: an elongated splat turned 45 degrees near the
guard's edge is keyed by the keys pass and not by the full projection.
Pins: numpy's eigvalsh for the extent, the guard's edge placed between the
two extents, the reference's isotropic sortGuard known answers
(renderTests.ts:341-367), and agreement with the full projection wherever
the two extents are equal.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from radix_amd import camera as cam
from tests.test_oracle_project import planes_of, random_scene

W, H, GUARD = 128, 64, 0.5
SQ8 = math.sqrt(8.0)


def rig():
    return cam.mono(np.eye(4), 60, W, H, near=0.1, far=100)


def edge_x(radius):
    """Head-space x (at depth 4, y 0) of the guard's edge for a splat of `radius`:
    |P00 x / w| - P00 r / w = guard + 1 with w = 4."""
    P00 = rig().eyes[0].projection[0, 0]
    return 4 * (GUARD + 1) / P00 + radius


def keys(planes, obj=None, guard=GUARD, bits=32):
    view = O.view_from_rig(rig(), obj)
    prm = O.params(sort_bits=bits, cull_keys=False, sort_guard=guard)
    return O.project_keys_f32(planes, view, prm)["keys"], O.project_f32(planes, 0, view, prm)["keys"]


def turned(n):
    s = math.sin(math.pi / 8)
    return [[0, 0, s, math.cos(math.pi / 8)]] * n  # 45 degrees about z


def test_reference_sort_guard_cases_match_the_full_projection():
    # renderTests.ts:341-367 (isotropic: both extents agree)
    p = planes_of([[0, 0, -4], [5.5, 0, -4], [20, 0, -4], [0, 0, 4]], [0.05] * 4)
    for guard, expected in ((0.0, [True] * 4), (0.5, [True, True, False, False])):
        k, full = keys(p, guard=guard, bits=24)
        assert list(k != 0xFFFFFF) == expected
        assert np.array_equal(k, full)


def test_elongated_turned_splat_uses_its_largest_scale():
    # scales (1, 0.01, 0.01) turned 45 degrees about z: splat extent 1, covariance extent sqrt(0.5 + 0.5e-4)
    e_cov = math.sqrt(0.5 + 0.5e-4)
    xs = [edge_x(SQ8 * e_cov) - 0.3, 0.5 * (edge_x(SQ8 * e_cov) + edge_x(SQ8)), edge_x(SQ8) + 0.3]
    p = planes_of([[x, 0, -4] for x in xs], [[1.0, 0.01, 0.01]] * 3, turned(3))
    k, full = keys(p)
    inv = O.key_invalid(32)
    assert list(k != inv) == [True, True, False]
    assert list(full != inv) == [True, False, False]
    assert k[0] == O.key_compute(math.hypot(xs[0], 4.0), 32)


def test_uniform_transform_scales_the_extent():
    # object = rotation about y times a uniform scale 2: extent 2 x max scale
    c, s = math.cos(0.3), math.sin(0.3)
    obj = np.array([[2 * c, 0, 2 * s, 0], [0, 2, 0, 0], [-2 * s, 0, 2 * c, 0], [0, 0, 0, 1.0]])
    r = 2 * 0.3 * SQ8
    head_x = [edge_x(r) - 0.1, edge_x(r) + 0.1]
    # object-space centres of head (x, 0, -4)
    inv = np.linalg.inv(obj)
    centres = [list((inv @ np.array([x, 0, -4, 1]))[:3]) for x in head_x]
    p = planes_of(centres, [[0.3, 0.2, 0.1]] * 2)
    k, _ = keys(p, obj)
    assert list(k != O.key_invalid(32)) == [True, False]


def test_non_uniform_transform_takes_the_largest_eigenvalue():
    obj = np.diag([1.5, 1.0, 1.0, 1.0])
    sc = np.array([1.0, 0.01, 0.01])
    q = turned(1)[0]
    from scipy.spatial.transform import Rotation
    R = Rotation.from_quat(q).as_matrix()
    M = obj[:3, :3] @ R @ np.diag(sc)
    lam = np.linalg.eigvalsh(M @ M.T)[-1]
    e_splat = math.sqrt(lam)
    e_cov = math.sqrt(max(np.diag(M @ M.T)))
    assert e_splat > e_cov + 0.1
    head_x = [edge_x(SQ8 * e_cov) - 0.05, 0.5 * (edge_x(SQ8 * e_cov) + edge_x(SQ8 * e_splat)), edge_x(SQ8 * e_splat) + 0.05]
    p = planes_of([[x / 1.5, 0, -4] for x in head_x], [sc] * 3, [q] * 3)
    k, full = keys(p, obj)
    inv = O.key_invalid(32)
    assert list(k != inv) == [True, True, False]
    assert list(full != inv) == [True, False, False]


def test_isotropic_scenes_agree_with_the_full_projection():
    centers, scales, quats, op = random_scene(2000, 31)
    iso = np.repeat(scales[:, :1], 3, axis=1)
    p = planes_of(centers * [3, 3, 1], iso, quats, op)
    for guard in (0.0, 0.2, 0.8):
        k, full = keys(p, guard=guard)
        assert np.array_equal(k, full), guard
        assert 0 < np.count_nonzero(k == O.key_invalid(32)) < 2000 or guard == 0.0

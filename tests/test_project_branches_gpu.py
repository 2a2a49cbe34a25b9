"""GPU parity of the ewa path's LoD, depth-of-field and orthographic
branches against the oracle (whose branches are pinned by closed forms in
tests/test_oracle_branches.py), to the parity bar of projection_check."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector, camera as cam, workloads as W
from tests.helpers import Rng
from tests.projection_check import compare, to_host
from tests.test_oracle_project import planes_of, random_scene

pytestmark = pytest.mark.gpu


def run(table, rig, params):
    out = Projector().project(table, rig, params, side=True)
    torch.cuda.synchronize()
    return to_host(out, table.count, len(rig.eyes))


def scene(n, seed, lod=False, sh=0):
    centers, scales, quats, op = random_scene(n, seed)
    if lod:
        # LoD opacities up to 255 / 128 (extSplats' byte), a third of them above 1
        op = 0.2 + Rng(seed + 1).random(n) * (255 / 128 - 0.2)
    p = planes_of(centers, scales, quats, op)
    if sh:
        K = [0, 3, 8, 15][sh]
        coef = (Rng(seed + 2).random(3 * K * n).reshape(3 * K, n) - 0.5) * 0.2
        p = np.concatenate([p, coef.astype(np.float32)], 0)
    return p


@pytest.mark.parametrize("lod_inflate", [True, False])
def test_lod_opacity_parity(cuda, lod_inflate):
    n = 20000
    p = scene(n, 4, lod=True, sh=3)
    rig = cam.mono(cam.look_at_world([0.2, 0.1, 0.4], [0, 0, -5]), 60, 640, 360)
    print(f"LoD opacities, lodInflate {lod_inflate}", flush=True)
    g = run(DeviceTable.from_f32_planes(p, n, 3), rig, Params(sort_bits=32, lod_inflate=lod_inflate))
    ref = O.project_f32(p, 3, O.view_from_rig(rig), O.params(sort_bits=32, lod_inflate=lod_inflate))
    rep = compare(g, ref, 1)
    above = np.count_nonzero((p[10] > 1) & ref["visible"][0].astype(bool))
    assert rep["eye0_compared"] > n // 2 and above > n // 5, (rep, above)


@pytest.mark.parametrize("ortho", [False, True])
def test_depth_of_field_parity(cuda, ortho):
    n = 20000
    p = scene(n, 7, sh=1)
    if ortho:
        rig = cam.mono_ortho(np.eye(4), 2.0, 640, 360, near=0.1, far=100)
    else:
        rig = cam.mono(np.eye(4), 60, 640, 360)
    prm = dict(sort_bits=32, focal_distance=5.0, aperture_tan=0.01)
    print(f"depth of field, ortho {ortho}", flush=True)
    g = run(DeviceTable.from_f32_planes(p, n, 1), rig, Params(**prm))
    ref = O.project_f32(p, 1, O.view_from_rig(rig), O.params(**prm))
    rep = compare(g, ref, 1)
    assert rep["eye0_compared"] > n // 2, rep


@pytest.mark.parametrize("square", [True, False])
def test_orthographic_parity(cuda, square):
    n = 20000
    p = scene(n, 11, lod=True, sh=2)
    rig = cam.mono_ortho(cam.look_at_world([0.3, -0.2, 0.5], [0, 0, -5]), 2.0, 640, 360, near=0.1, far=100)
    if not square:
        rig.eyes[0].projection = cam.orthographic(-3.0, 3.0, 1.5, -1.5, 0.1, 100.0)
    print(f"orthographic, square pixels {square}", flush=True)
    for bits in (32, 24):
        g = run(DeviceTable.from_f32_planes(p, n, 2), rig, Params(sort_bits=bits))
        ref = O.project_f32(p, 2, O.view_from_rig(rig), O.params(sort_bits=bits))
        rep = compare(g, ref, 1)
        assert rep["eye0_compared"] > n // 2, rep


def test_orthographic_stereo_and_sort(cuda):
    n = 30000
    p = scene(n, 13)
    rig = cam.stereo(cam.look_at_world([0, 0, 0.5], [0, 0, -5]), 0.064, 60, 512, 544)
    for e in rig.eyes:
        e.projection = cam.orthographic(-2.0, 2.0, 2.0, -2.0, 0.1, 100.0)
        e.ortho = True
    prm = Params(sort_bits=32, key_per_eye=True)
    table = DeviceTable.from_f32_planes(p, n, 0)
    g = run(table, rig, prm)
    ref = O.project_f32(p, 0, O.view_from_rig(rig), O.params(sort_bits=32, key_per_eye=True))
    g["keys"] = g["keys"].reshape(2, -1)
    compare(g, ref, 2, per_eye=True)

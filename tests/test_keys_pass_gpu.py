"""The keys-only pass (rdx_project_keys: SplatProjection.writeKeys,
src/render/SplatProjection.ts:165-172, the "keys" pass of
SplatProjector.ts:184-189) on the GPU: its keys equal the keys the full
projection writes with cullKeys off (the same head-space distance, cull and
sortGuard rule) bit for bit, for f32, ext and packed tables, mono and stereo
heads, with and without the guard and an object transform, and equal the FP64
oracle's away from a cull threshold; the reference's sortGuard case
(renderTests.ts:341-367) and the write offset."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector, camera as cam, workloads as W
from tests.projection_check import TOL
from tests.test_oracle_project import planes_of
from tests.test_project_gpu import ext_table

pytestmark = pytest.mark.gpu


def keys_both(table, rig, prm):
    """(keys pass, full projection's keys) as uint32 arrays."""
    p = Projector()
    k = p.project_keys(table, rig, prm)
    full = p.project(table, rig, prm)
    torch.cuda.synchronize()
    return k.cpu().numpy().view(np.uint32), full["keys"].cpu().numpy().view(np.uint32)


def against_oracle(keys, ref):
    bad = np.nonzero(keys != ref["keys"])[0]
    assert np.all(ref["margin"][bad] < TOL), f"{bad.size} keys differ away from a cull threshold"
    return bad.size


@pytest.mark.parametrize("bits,guard", [(32, 0.0), (24, 0.0), (16, 0.6)])
def test_f32_equals_projection_and_oracle(cuda, bits, guard):
    spec = W.plain_small_spec(seed=4)
    planes = W.generate(spec)
    # a 20 degree view: the guard excludes the splats well outside it
    rig = W.plain_small_camera() if guard == 0 else cam.mono(np.eye(4), 20, 320, 240, near=0.1, far=100)
    prm = Params(sort_bits=bits, cull_keys=False, sort_guard=guard)
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    k, full = keys_both(table, rig, prm)
    assert np.array_equal(k, full)
    ref = O.project_f32(planes, 0, O.view_from_rig(rig), O.params(sort_bits=bits, cull_keys=False, sort_guard=guard))
    excused = against_oracle(k, ref)
    invalid = np.count_nonzero(k == O.lib().orc_key_invalid(bits))
    print(f"bits {bits} guard {guard}: {invalid} invalid, {excused} excused", flush=True)
    assert invalid < spec.count and (invalid > 0) == (guard > 0)


def test_f32_sh3_object_transform_4m(cuda):
    spec = W.plain_4m_spec(seed=21)
    planes = W.generate(spec)
    obj = cam.translation([0.5, -1.0, 2.0])
    obj[:3, :3] = 1.3 * np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    rig = W.plain_4m_camera()
    table = DeviceTable.from_f32_planes(planes, spec.count, 3, object_to_world=obj)
    print("plain_4m keys pass, object transform", flush=True)
    for guard in (0.0, 0.8):
        k, full = keys_both(table, rig, Params(sort_bits=32, cull_keys=False, sort_guard=guard))
        assert np.array_equal(k, full)


@pytest.mark.parametrize("sh_codec", [2, 3])
def test_ext_equals_projection(cuda, sh_codec):
    n = 8000
    a, b, chans = ext_table(n, 14, sh_codec)
    rig = W.orbit_camera(6, 7.0, 1024, 768)
    table = DeviceTable.from_ext(a, b, n, chans, sh_codec=sh_codec, sh_degree=3)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=0.7)
    k, full = keys_both(table, rig, prm)
    assert np.array_equal(k, full)
    ref = O.project_ext(a, b, chans, sh_codec, 3, O.view_from_rig(rig),
                        O.params(sort_bits=32, cull_keys=False, sort_guard=0.7))
    against_oracle(k, ref)


@pytest.mark.parametrize("guard", [0.0, 0.5])
def test_packed_stereo_head_equals_projection(cuda, guard):
    spec = W.stereo_16m_spec(count=65536, seed=3)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = (0.0, 1.0, -10.0, 2.0)
    rig = W.stereo_camera(width=512, height=544)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=guard)
    table = DeviceTable.from_packed(words, spec.count, enc, chans, 3)
    k, full = keys_both(table, rig, prm)
    assert np.array_equal(k, full)
    ref = O.project_packed(words, enc, chans, 3, O.view_from_rig(rig),
                           O.params(sort_bits=32, cull_keys=False, sort_guard=guard))
    against_oracle(k, ref)


def test_sort_guard_known_answers(cuda):
    # renderTests.ts:341-367
    rig = cam.mono(np.eye(4), 60, 128, 64, near=0.1, far=100)
    p = planes_of([[0, 0, -4], [5.5, 0, -4], [20, 0, -4], [0, 0, 4]], [0.05] * 4)
    table = DeviceTable.from_f32_planes(p, 4, 0)
    for guard, expected in ((0.0, [True] * 4), (0.5, [True, True, False, False])):
        k = Projector().project_keys(table, rig, Params(sort_bits=24, sort_guard=guard))
        assert list(k.cpu().numpy().view(np.uint32) != 0xFFFFFF) == expected


def test_cull_invisible_and_offset(cuda):
    # the near-transparent splat is culled (renderTests.ts:302-321), the others keep their keys
    rig = cam.mono(np.eye(4), 60, 128, 64, near=0.1, far=100)
    p = planes_of([[0, 0, -3], [0, 0, 3], [40, 0, -3], [0, 0, -3]], [0.1] * 4, opacity=[1, 1, 1, 0.001])
    table = DeviceTable.from_f32_planes(p, 4, 0)
    keys = torch.full((10,), 7, dtype=torch.int32, device="cuda")
    Projector().project_keys(table, rig, Params(sort_bits=32), keys=keys, at=5)
    h = keys.cpu().numpy().view(np.uint32)
    assert list(h[:5]) == [7] * 5 and h[9] == 7  # only [at, at + n) written
    assert list(h[5:9] != 0xFFFFFFFF) == [True, True, True, False]
    d = np.array([3.0, 3.0, np.hypot(40.0, 3.0)], dtype=np.float32)
    assert list(h[5:8]) == [O.lib().orc_key_compute(float(x), 32) for x in d]


def test_empty_table_and_bad_bits(cuda):
    rig = cam.mono(np.eye(4), 60, 128, 64, near=0.1, far=100)
    p = planes_of([[0, 0, -3]], [0.1])
    table = DeviceTable.from_f32_planes(p, 1, 0)
    with pytest.raises(Exception, match="bits"):
        Projector().project_keys(table, rig, Params(sort_bits=12))

"""The keys-only pass (rdx_project_keys: SplatProjection.writeKeys,
src/render/SplatProjection.ts:165-172, the "keys" pass of
SplatProjector.ts:184-189) on the GPU, against the oracle's keys pass
(orc_project_keys_table_f32, whose splat-form sortGuard extent is pinned in
tests/test_oracle_keys_pass.py) away from a cull threshold, for f32, ext and
packed tables (the word tables through their decoded splats), mono and
stereo heads, uniform and non-uniform object transforms; without the guard
its keys equal the keys the full projection writes with cullKeys off bit for
bit (the same head-space distance and cull rule); the reference's sortGuard
case (renderTests.ts:341-367) and the write offset."""
import ctypes as C

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


def keys_oracle(planes, rig, bits, guard, obj=None):
    return O.project_keys_f32(planes, O.view_from_rig(rig, obj),
                              O.params(sort_bits=bits, cull_keys=False, sort_guard=guard))


def decoded_planes(layout, words, enc=None):
    """f32 planes of a word table's splats as the oracle decodes them (log-scales, opacity, colour)."""
    L = O.lib()
    n = words[0].shape[0]
    p = np.zeros((14, n), dtype=np.float32)
    c, s, q, rgb = (np.zeros(k) for k in (3, 3, 4, 3))
    op = C.c_double()
    e = np.asarray(enc if enc is not None else [0, 1, 0, 1], dtype=np.float64)
    for i in range(n):
        if layout == "ext":
            L.orc_ext_decode(O.uptr(np.ascontiguousarray(words[0][i])), O.uptr(np.ascontiguousarray(words[1][i])),
                             O.dptr(c), O.dptr(s), O.dptr(q), O.dptr(rgb), C.byref(op))
        else:
            L.orc_packed_decode(O.uptr(np.ascontiguousarray(words[0][i])), O.dptr(e), O.dptr(c), O.dptr(s), O.dptr(q),
                                O.dptr(rgb), C.byref(op))
        p[0:3, i] = c
        with np.errstate(divide="ignore"):
            p[3:6, i] = np.log(s)
        p[6:10, i] = q
        p[10, i] = op.value
        p[11:14, i] = rgb
    return p


@pytest.mark.parametrize("bits,guard", [(32, 0.0), (24, 0.0), (16, 0.6), (32, 0.3)])
def test_f32_keys_pass_against_oracle(cuda, bits, guard):
    spec = W.plain_small_spec(seed=4)
    planes = W.generate(spec)
    # a 20 degree view: the guard excludes the splats well outside it
    rig = W.plain_small_camera() if guard == 0 else cam.mono(np.eye(4), 20, 320, 240, near=0.1, far=100)
    prm = Params(sort_bits=bits, cull_keys=False, sort_guard=guard)
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    k, full = keys_both(table, rig, prm)
    if guard == 0:
        assert np.array_equal(k, full)
    excused = against_oracle(k, keys_oracle(planes, rig, bits, guard))
    invalid = np.count_nonzero(k == O.lib().orc_key_invalid(bits))
    print(f"bits {bits} guard {guard}: {invalid} invalid, {excused} excused", flush=True)
    assert invalid < spec.count and (invalid > 0) == (guard > 0)


def test_elongated_splats_near_the_guard(cuda):
    # scales (1, 0.01, 0.01) turned 45 degrees about z along the guard's edge: the
    # keys pass keys those within their largest scale's reach, the full projection
    # only those within the covariance diagonal's (tests/test_oracle_keys_pass.py)
    n = 4096
    u = np.random.default_rng(5).random((n, 3))
    rig = cam.mono(np.eye(4), 60, 128, 64, near=0.1, far=100)
    s = np.sin(np.pi / 8)
    planes = planes_of(np.stack([6.5 + u[:, 0] * 5, (u[:, 1] - 0.5) * 2, -4 - u[:, 2]], 1), [[1.0, 0.01, 0.01]] * n,
                       [[0, 0, s, np.cos(np.pi / 8)]] * n)
    table = DeviceTable.from_f32_planes(planes, n, 0)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=0.5)
    k, full = keys_both(table, rig, prm)
    against_oracle(k, keys_oracle(planes, rig, 32, 0.5))
    inv = O.key_invalid(32)
    only_keys = np.count_nonzero((k != inv) & (full == inv))
    assert only_keys > n // 20 and np.count_nonzero((full != inv) & (k == inv)) == 0, only_keys


@pytest.mark.parametrize("uniform", [True, False])
def test_f32_sh3_object_transform_4m(cuda, uniform):
    spec = W.plain_4m_spec(seed=21)
    planes = W.generate(spec)
    obj = cam.translation([0.5, -1.0, 2.0])
    rot = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    obj[:3, :3] = 1.3 * rot if uniform else rot @ np.diag([1.4, 0.9, 1.1])
    rig = W.plain_4m_camera()
    table = DeviceTable.from_f32_planes(planes, spec.count, 3, object_to_world=obj)
    print(f"plain_4m keys pass, object transform (uniform {uniform})", flush=True)
    for guard in (0.0, 0.8):
        k, full = keys_both(table, rig, Params(sort_bits=32, cull_keys=False, sort_guard=guard))
        if guard == 0:
            assert np.array_equal(k, full)
        excused = against_oracle(k, keys_oracle(planes[:14], rig, 32, guard, obj))
        print(f"guard {guard}: {excused} excused", flush=True)


@pytest.mark.parametrize("sh_codec", [2, 3])
def test_ext_keys_pass(cuda, sh_codec):
    n = 8000
    a, b, chans = ext_table(n, 14, sh_codec)
    rig = W.orbit_camera(6, 7.0, 1024, 768)
    table = DeviceTable.from_ext(a, b, n, chans, sh_codec=sh_codec, sh_degree=3)
    for guard in (0.0, 0.7):
        k, full = keys_both(table, rig, Params(sort_bits=32, cull_keys=False, sort_guard=guard))
        if guard == 0:
            assert np.array_equal(k, full)
        against_oracle(k, keys_oracle(decoded_planes("ext", (a, b)), rig, 32, guard))


def cov_decoded_planes(a, b):
    """f32 planes of a covSplats table's splats in the splat form: the
    decoded covariance (orc_cov_decode) through fromCov (splats.tsl.ts:189-198)
    as an eigendecomposition in double: scales sqrt(eigenvalues), rotation the
    right-handed eigenbasis."""
    from scipy.spatial.transform import Rotation

    L = O.lib()
    n = a.shape[0]
    p = np.zeros((14, n), dtype=np.float32)
    c, cov, rgb = (np.zeros(k) for k in (3, 6, 3))
    op = C.c_double()
    for i in range(n):
        L.orc_cov_decode(O.uptr(np.ascontiguousarray(a[i])), O.uptr(np.ascontiguousarray(b[i])), O.dptr(c), O.dptr(cov),
                         O.dptr(rgb), C.byref(op))
        S = np.array([[cov[0], cov[3], cov[4]], [cov[3], cov[1], cov[5]], [cov[4], cov[5], cov[2]]])
        w, V = np.linalg.eigh(S)
        if np.linalg.det(V) < 0:
            V[:, 0] = -V[:, 0]
        p[0:3, i] = c
        with np.errstate(divide="ignore"):
            p[3:6, i] = 0.5 * np.log(np.maximum(w, 0.0))
        p[6:10, i] = Rotation.from_matrix(V).as_quat()
        p[10, i] = op.value
        p[11:14, i] = rgb
    return p


def test_cov_keys_pass(cuda):
    # a covSplats source in the keys pass (no wanted form) is read in the splat
    # form (SplatProjector.ts:152-167): its extent is the largest scale of
    # fromCov, sqrt of the head-space covariance's largest eigenvalue
    from tests.test_project_gpu import cov_table

    n = 8000
    a, b, chans = cov_table(n, 16, 2, lod=True)
    rig = W.orbit_camera(8, 7.0, 1024, 768)
    table = DeviceTable.from_cov(a, b, n, chans, sh_codec=2, sh_degree=3)
    planes = cov_decoded_planes(a, b)
    for guard in (0.0, 0.7):
        k, full = keys_both(table, rig, Params(sort_bits=32, cull_keys=False, sort_guard=guard))
        if guard == 0:
            assert np.array_equal(k, full)
        excused = against_oracle(k, keys_oracle(planes, rig, 32, guard))
        print(f"guard {guard}: {excused} excused", flush=True)


@pytest.mark.parametrize("guard", [0.0, 0.5])
def test_packed_stereo_head_keys_pass(cuda, guard):
    spec = W.stereo_16m_spec(count=65536, seed=3)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = (0.0, 1.0, -10.0, 2.0)
    rig = W.stereo_camera(width=512, height=544)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=guard)
    table = DeviceTable.from_packed(words, spec.count, enc, chans, 3)
    k, full = keys_both(table, rig, prm)
    if guard == 0:
        assert np.array_equal(k, full)
    against_oracle(k, keys_oracle(decoded_planes("packed", (words,), enc), rig, 32, guard))


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

"""GPU projection parity against the oracle (FP64 host restatement):
words, keys, float outputs and the fused project + sort order, on the
workloads and on the reference's own test cases (renderTests.ts)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector, camera as cam, workloads as W
from tests.projection_check import _fields, compare, to_host
from tests.test_oracle_project import planes_of, random_scene

pytestmark = pytest.mark.gpu


def run(table, rig, params, side=True):
    out = Projector().project(table, rig, params, side=side)
    torch.cuda.synchronize()
    return to_host(out, table.count, len(rig.eyes))


def check_order(table, rig, params, keys_expected=None):
    """project_sort's indices equal SortKeys.sort of the keys it wrote."""
    out = Projector().project_sort(table, rig, params)
    torch.cuda.synchronize()
    n = table.count
    keys = out["keys"].cpu().numpy().view(np.uint32)[:n]
    count = int(out["count"].cpu().item())
    order = out["indices"].cpu().numpy().view(np.uint32)[:count]
    ordering, visible = O.sort(keys, params.sort_bits)
    assert count == visible
    assert np.array_equal(order, ordering[:visible])
    args = out["args"].cpu().numpy()
    assert list(args) == [6144, -(-visible // 1024), 0, 0, 0]
    if keys_expected is not None:
        assert np.array_equal(keys, keys_expected)
    return keys, visible


def test_plain_small_parity(cuda):
    spec = W.plain_small_spec()
    planes = W.generate(spec)
    rig = W.plain_small_camera()
    prm = Params(sort_bits=32)
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    print("plain_small 65,536 SH0 32-bit", flush=True)
    g = run(table, rig, prm)
    ref = O.project_f32(planes, 0, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare(g, ref, 1, exact_rgb=True)
    assert rep["eye0_compared"] > 30000
    check_order(table, rig, prm, keys_expected=g["keys"][: spec.count])


@pytest.mark.parametrize("bits", [16, 24])
def test_plain_small_key_widths(cuda, bits):
    spec = W.plain_small_spec(seed=9)
    planes = W.generate(spec)
    rig = W.plain_small_camera()
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    g = run(table, rig, Params(sort_bits=bits))
    ref = O.project_f32(planes, 0, O.view_from_rig(rig), O.params(sort_bits=bits))
    compare(g, ref, 1)
    check_order(table, rig, Params(sort_bits=bits))


def test_plain_4m_sh3_parity(cuda):
    spec = W.plain_4m_spec()
    planes = W.generate(spec)
    rig = W.plain_4m_camera()
    prm = Params(sort_bits=32)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3)
    print("plain_4m 4,194,304 SH3 32-bit", flush=True)
    g = run(table, rig, prm)
    ref = O.project_f32(planes, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare(g, ref, 1, rgb_sample=50000)
    print(rep, flush=True)
    assert rep["eye0_compared"] > 1_000_000
    check_order(table, rig, prm, keys_expected=g["keys"][: spec.count])


def test_packed_stereo_per_eye_parity(cuda):
    spec = W.stereo_16m_spec(count=65536)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = (0.0, 1.0, -10.0, 2.0)
    rig = W.stereo_camera(width=512, height=544)
    prm = Params(sort_bits=32, key_per_eye=True)
    table = DeviceTable.from_packed(words, spec.count, enc, chans, 3)
    g = run(table, rig, prm)
    ref = O.project_packed(words, enc, chans, 3, O.view_from_rig(rig),
                           O.params(sort_bits=32, key_per_eye=True))
    g["keys"] = g["keys"].reshape(2, -1)
    compare(g, ref, 2, per_eye=True)


def test_packed_mono_shared_key_parity(cuda):
    spec = W.stereo_16m_spec(seed=8, count=50000)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = (0.0, 1.0, -10.0, 2.0)
    rig = W.orbit_camera(77, 6.0, 800, 600)
    table = DeviceTable.from_packed(words, spec.count, enc, chans, 3)
    g = run(table, rig, Params(sort_bits=24))
    ref = O.project_packed(words, enc, chans, 3, O.view_from_rig(rig), O.params(sort_bits=24))
    compare(g, ref, 1)
    check_order(table, rig, Params(sort_bits=24))


def ext_table(n, seed, sh_codec):
    """ext + extSh (codec 2) or ext + packedSh (codec 3) words built by the oracle's encoders."""
    L = O.lib()
    spec = W.plain_4m_spec(seed=seed, count=n)
    p = W.generate(spec).astype(np.float64)
    a = np.zeros((n, 4), dtype=np.uint32)
    b = np.zeros((n, 4), dtype=np.uint32)
    texels = 4 if sh_codec == 2 else 3
    chans = [np.zeros((n, 4), dtype=np.uint32) for _ in range(texels)]
    words = np.zeros(16, dtype=np.uint32)
    for i in range(n):
        c = np.ascontiguousarray(p[0:3, i])
        s = np.ascontiguousarray(np.exp(p[3:6, i]))
        q = np.ascontiguousarray(p[6:10, i])
        rgb = np.ascontiguousarray(p[11:14, i])
        L.orc_ext_encode(O.dptr(c), O.dptr(s), O.dptr(q), O.dptr(rgb), float(p[10, i]), O.uptr(a[i]), O.uptr(b[i]))
        words[:] = 0
        L.orc_sh_encode(sh_codec, O.dptr(np.ascontiguousarray(p[14:59, i])), 3, O.uptr(words))
        for t in range(texels):
            chans[t][i] = words[4 * t:4 * t + 4]
    return a, b, chans


@pytest.mark.parametrize("sh_codec", [2, 3])
def test_ext_layout_parity(cuda, sh_codec):
    n = 20000
    a, b, chans = ext_table(n, 12, sh_codec)
    rig = W.orbit_camera(5, 7.0, 1024, 768)
    table = DeviceTable.from_ext(a, b, n, chans, sh_codec=sh_codec, sh_degree=3)
    g = run(table, rig, Params(sort_bits=32))
    ref = O.project_ext(a, b, chans, sh_codec, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    compare(g, ref, 1)


def cov_table(n, seed, sh_codec, lod=False):
    """covSplats words (splats.ts:219-266) of a synthetic cloud's covariances
    (covFromScaleQuat) with extSh (codec 2) or packedSh (codec 3) bundles,
    built by the oracle's encoders (pinned by the known answers of
    tests/test_layout_kat.py); `lod`: a fifth of the splats get LoD
    opacities in (1, 255/128]."""
    L = O.lib()
    spec = W.plain_4m_spec(seed=seed, count=n)
    p = W.generate(spec).astype(np.float64)
    if lod:
        p[10, ::5] = 1.0 + (255 / 128 - 1.0) * np.linspace(0.01, 1.0, p[10, ::5].size)
    a = np.zeros((n, 4), dtype=np.uint32)
    b = np.zeros((n, 4), dtype=np.uint32)
    texels = 4 if sh_codec == 2 else 3
    chans = [np.zeros((n, 4), dtype=np.uint32) for _ in range(texels)]
    words = np.zeros(16, dtype=np.uint32)
    cov = np.zeros(6)
    for i in range(n):
        s = np.ascontiguousarray(np.exp(p[3:6, i]))
        q = np.ascontiguousarray(p[6:10, i])
        L.orc_cov_from_scale_quat(O.dptr(s), O.dptr(q), O.dptr(cov))
        L.orc_cov_encode(O.dptr(np.ascontiguousarray(p[0:3, i])), O.dptr(cov), O.dptr(np.ascontiguousarray(p[11:14, i])),
                         float(p[10, i]), O.uptr(a[i]), O.uptr(b[i]))
        words[:] = 0
        L.orc_sh_encode(sh_codec, O.dptr(np.ascontiguousarray(p[14:59, i])), 3, O.uptr(words))
        for t in range(texels):
            chans[t][i] = words[4 * t:4 * t + 4]
    return a, b, chans


@pytest.mark.parametrize("sh_codec", [2, 3])
def test_cov_layout_parity(cuda, sh_codec):
    # the ewa method reads a covSplats splat's covariance as stored (GpuCovSplatReader.loadCovSplat,
    # splatAccess.ts:107-109; decodeCov, splats.tsl.ts:105-121), LoD opacities included
    n = 20000
    a, b, chans = cov_table(n, 14, sh_codec, lod=True)
    rig = W.orbit_camera(9, 7.0, 1024, 768)
    table = DeviceTable.from_cov(a, b, n, chans, sh_codec=sh_codec, sh_degree=3)
    g = run(table, rig, Params(sort_bits=32))
    ref = O.project_cov(a, b, chans, sh_codec, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare(g, ref, 1)
    assert rep["eye0_compared"] > n // 4
    check_order(table, rig, Params(sort_bits=32), keys_expected=g["keys"][:n])


def test_cov_layout_transform_stereo_guard(cuda):
    # a non-uniform object transform, two eyes, and the CPU sort's keys
    # (cullKeys off, sortGuard on: the cov form's extent, projection.tsl.ts:144-155)
    n = 12000
    a, b, chans = cov_table(n, 15, 2)
    obj = cam.translation([0.3, 0.2, -0.5])
    obj[:3, :3] = np.array([[1.3, 0.2, 0.0], [0.0, 0.8, 0.1], [0.1, 0.0, 1.1]])
    rig = cam.stereo(cam.orbit_world(cam.random_direction(23), 7.0), 0.064, 70.0, 900, 700)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=0.1)
    table = DeviceTable.from_cov(a, b, n, chans, sh_codec=2, sh_degree=3, object_to_world=obj)
    g = run(table, rig, prm)
    ref = O.project_cov(a, b, chans, 2, 3, O.view_from_rig(rig, object_to_world=obj), O.params(sort_bits=32, cull_keys=False,
                                                                                           sort_guard=0.1))
    compare(g, ref, 2)


def test_object_transform_parity(cuda):
    spec = W.plain_4m_spec(seed=31, count=40000)
    planes = W.generate(spec)
    obj = cam.translation([0.5, -1.0, 2.0])
    obj[:3, :3] = 1.7 * np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    rig = W.orbit_camera(13, 10.0, 1280, 720)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3, object_to_world=obj)
    g = run(table, rig, Params(sort_bits=32))
    ref = O.project_f32(planes, 3, O.view_from_rig(rig, object_to_world=obj), O.params(sort_bits=32))
    compare(g, ref, 1)


def test_culling_known_answers_gpu(cuda):
    # renderTests.ts:302-321
    rig = cam.mono(np.eye(4), 60, 128, 64, near=0.1, far=100)
    p = planes_of([[0, 0, -3], [0, 0, 3], [40, 0, -3], [0, 0, -3]], [0.1] * 4, opacity=[1, 1, 1, 0.001])
    table = DeviceTable.from_f32_planes(p, 4, 0)
    for cull_keys, expected in ((True, [False, True, True, True]), (False, [False, False, False, True])):
        g = run(table, rig, Params(sort_bits=32, cull_keys=cull_keys), side=False)
        assert list((g["words"][0][:, 3] & 0xFF) == 0) == [False, True, True, True]
        assert list(g["keys"][:4] == 0xFFFFFFFF) == expected


def test_sort_guard_known_answers_gpu(cuda):
    # renderTests.ts:341-367
    rig = cam.mono(np.eye(4), 60, 128, 64, near=0.1, far=100)
    p = planes_of([[0, 0, -4], [5.5, 0, -4], [20, 0, -4], [0, 0, 4]], [0.05] * 4)
    table = DeviceTable.from_f32_planes(p, 4, 0)
    for guard, expected in ((0.0, [True] * 4), (0.5, [True, True, False, False])):
        g = run(table, rig, Params(sort_bits=24, cull_keys=False, sort_guard=guard), side=False)
        assert list(g["keys"][:4] != 0xFFFFFF) == expected


def test_two_objects_at_offsets_equal_one(cuda):
    # renderTests.ts:202-219: objects of one collection at offsets project as one
    spec = W.plain_4m_spec(seed=17, count=30000)
    planes = W.generate(spec)
    rig = W.orbit_camera(3, 8.0, 800, 600)
    whole = run(DeviceTable.from_f32_planes(planes, spec.count, 3), rig, Params(sort_bits=32), side=False)
    import ctypes as C

    from radix_amd import _native as N
    from radix_amd.sort import _stream

    h = spec.count // 2
    words = torch.zeros((spec.count, 4), dtype=torch.int32, device="cuda")
    keys = torch.zeros(spec.count, dtype=torch.int32, device="cuda")
    proj = Projector()
    for lo, hi in ((0, h), (h, spec.count)):
        t = DeviceTable.from_f32_planes(np.ascontiguousarray(planes[:, lo:hi]), hi - lo, 3)
        wp = (C.c_void_p * 2)(words.data_ptr() + 0, None)
        N.check(N.lib().rdx_project_ewa(proj.context.handle, _stream(), C.byref(t.desc()), C.byref(rig.desc()),
                                        C.byref(Params(sort_bits=32).desc()), lo, wp, C.c_void_p(keys.data_ptr()),
                                        0, None))
    torch.cuda.synchronize()
    assert np.array_equal(words.cpu().numpy().view(np.uint32), whole["words"][0])
    assert np.array_equal(keys.cpu().numpy().view(np.uint32), whole["keys"][: spec.count])


def test_stereo_eye_equals_mono_from_eye_gpu(cuda):
    # renderTests.ts:221-240
    centers, scales, quats, op = random_scene(3000, 3)
    p = planes_of(centers, scales, quats, op)
    table = DeviceTable.from_f32_planes(p, 3000, 0)
    st = cam.stereo(cam.look_at_world([0, 0.1, 0.2], [0, 0, -4]), 0.06, 60, 128, 64, near=0.1, far=100)
    s = run(table, st, Params(sort_bits=32), side=False)
    for v in range(2):
        m = run(table, cam.mono(np.linalg.inv(st.eyes[v].view), 60, 128, 64, near=0.1, far=100),
                Params(sort_bits=32), side=False)
        # The two routes to the eye transform differ in the last bits, so a
        # quantized field may land one code apart (angle: modulo its wrap).
        fs, fm = _fields(s["words"][v]), _fields(m["words"][0])
        for name in fs:
            d = np.abs(fs[name] - fm[name])
            if name == "angle":
                d = np.minimum(d, 4096 - d)
            assert d.max() <= 1, name
        assert np.array_equal(s["words"][v][:, 2], m["words"][0][:, 2])


def test_project_sort_deterministic(cuda):
    spec = W.plain_4m_spec(seed=23, count=500000)
    planes = W.generate(spec)
    rig = W.plain_4m_camera(seed=23)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3)
    outs = []
    for _ in range(3):
        o = Projector().project_sort(table, rig, Params(sort_bits=32))
        torch.cuda.synchronize()
        outs.append((o["indices"].cpu().numpy().copy(), int(o["count"].cpu().item())))
    assert all(np.array_equal(outs[0][0], o[0]) and outs[0][1] == o[1] for o in outs)


def test_stereo_per_eye_segmented_sort(cuda):
    # Per-eye keys sort as two segments in one sequence; each eye's order is
    # SortKeys.sort of that eye's keys (the mono-from-eye oracle per segment).
    spec = W.stereo_16m_spec(seed=4, count=70000)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = (0.0, 1.0, -10.0, 2.0)
    rig = W.stereo_camera(width=640, height=680)
    prm = Params(sort_bits=32, key_per_eye=True)
    table = DeviceTable.from_packed(words, spec.count, enc, chans, 3)
    n = spec.count
    out = Projector().project_sort(table, rig, prm)
    torch.cuda.synchronize()
    keys = out["keys"].cpu().numpy().view(np.uint32).reshape(2, n)
    ref = O.project_packed(words, enc, chans, 3, O.view_from_rig(rig), O.params(sort_bits=32, key_per_eye=True))
    counts = out["count"].cpu().numpy()
    args = out["args"].cpu().numpy().reshape(2, 5)
    idx = out["indices"].cpu().numpy().view(np.uint32).reshape(2, n)
    g = run(table, rig, prm, side=False)
    assert np.array_equal(g["keys"].reshape(2, -1)[:, :n], keys)
    for e in range(2):
        ordering, visible = O.sort(keys[e], 32)
        assert counts[e] == visible
        assert np.array_equal(idx[e, :visible], ordering[:visible])
        assert list(args[e]) == [6144, -(-visible // 1024), 0, 0, 0]
        # keys within the parity bar of the oracle's per-eye keys
        near = ref["margin"] < 1e-5
        bad = np.nonzero(keys[e] != ref["keys"][e])[0]
        assert np.all(near[bad])


def test_project_sort_on_a_caller_stream(cuda):
    # Outputs allocated, filled and written on the caller's stream (default
    # priority) equal those of torch's current stream.
    spec = W.plain_4m_spec(seed=29, count=300000)
    table = DeviceTable.from_f32_planes(W.generate(spec), spec.count, 3)
    rig = W.plain_4m_camera(seed=29)
    a = Projector().project_sort(table, rig, Params(sort_bits=32))
    side = torch.cuda.Stream()
    b = Projector().project_sort(table, rig, Params(sort_bits=32), stream=side)
    side.synchronize()
    torch.cuda.synchronize()
    n = int(a["count"].item())
    assert n == int(b["count"].item()) and n > spec.count // 2
    assert torch.equal(a["indices"][:n], b["indices"][:n]) and torch.equal(a["keys"], b["keys"])
    assert torch.equal(a["words"][0], b["words"][0])


def test_stereo_16m_full_size_order(cuda):
    # The quantized stereo workload at its full size (16,777,216 packedSplats
    # + packedSh SH3, two 2064x2208 eyes, per-eye keys): each eye's indices,
    # count and draw arguments are the oracle's SortKeys.sort of the keys the
    # GPU wrote. Size-independent: the keys' own parity is held at 70,000
    # splats above.
    spec = W.stereo_16m_spec()
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    del planes
    table = DeviceTable.from_packed(words, spec.count, (0.0, 1.0, -10.0, 2.0), chans, 3)
    del words, chans
    n = spec.count
    out = Projector().project_sort(table, W.stereo_camera(), Params(sort_bits=32, key_per_eye=True))
    torch.cuda.synchronize()
    keys = out["keys"].cpu().numpy().view(np.uint32).reshape(2, n)
    idx = out["indices"].cpu().numpy().view(np.uint32).reshape(2, n)
    counts = out["count"].cpu().numpy()
    args = out["args"].cpu().numpy().reshape(2, 5)
    for e in range(2):
        ordering, visible = O.sort(keys[e], 32)
        assert visible > n // 2
        assert counts[e] == visible
        assert np.array_equal(idx[e, :visible], ordering[:visible])
        assert list(args[e]) == [6144, -(-visible // 1024), 0, 0, 0]


def test_plain_small_against_golden_fixture(cuda):
    # The committed vectors (tests/golden/plain_small.npz): keys bit-exact
    # away from cull thresholds, and then the order and visible count exactly.
    import os

    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plain_small.npz"))
    spec = W.plain_small_spec()
    planes = W.generate(spec)
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    out = Projector().project_sort(table, W.plain_small_camera(), Params(sort_bits=32))
    torch.cuda.synchronize()
    keys = out["keys"].cpu().numpy().view(np.uint32)[:spec.count]
    bad = np.nonzero(keys != g["keys"])[0]
    assert np.all(g["near"][bad]), f"{np.count_nonzero(~g['near'][bad])} keys differ away from cull thresholds"
    c = int(out["count"].cpu().item())
    if bad.size == 0:
        assert c == int(g["visible"][0])
        assert np.array_equal(out["indices"].cpu().numpy().view(np.uint32)[:c], g["ordering"][:c])
    else:  # a flipped cull moves the order: check it against the oracle sort of these keys
        ordering, visible = O.sort(keys, 32)
        assert c == visible and np.array_equal(out["indices"].cpu().numpy().view(np.uint32)[:c], ordering[:c])
    words = out["words"][0].cpu().numpy().view(np.uint32)[:4096]
    assert np.array_equal((words[:, 3] & 0xFF) != 0, (g["words"][:, 3] & 0xFF) != 0)

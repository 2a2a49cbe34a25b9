"""GPU 3dgut, htgs and raygs projection words (rdx_project_ray) against the FP64
oracle (oracle/ray.c, the reference's projections.tsl.ts:170-429 restated)
on the same tables: f32 planes with SH, two eyes, a uniformly scaled and
rotated object, a non-uniform object transform (through fromCov), ext and
covSplats words, and the keys with cullKeys and the sortGuard.

Tolerance: the kernel computes in f32 as the reference's shader does, the
oracle in f64. Codes may differ by one step (the 16-bit ndc, the 20-bit
centre direction, the opacity byte, the 11-bit log2 scales, the 10-bit angle
of elongated bounds, the half-float log scales), and a few splats may flip a
decision near a threshold (the oracle's margin says which); htgs's outline
comes from differences of nearly equal products, so a small fraction of its
scales may differ by more than a step."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector, camera as cam, workloads as W
from tests.test_project_gpu import ext_table

pytestmark = pytest.mark.gpu
TOL = 1e-4


def half(h):
    return np.array([O.lib().orc_decode_half(int(x)) for x in h])


def quat_angle(a, b):
    out = []
    for x, y in zip(a, b):
        qa, qb = np.zeros(4), np.zeros(4)
        O.lib().orc_decode_quat_oct(int(x), 10, 12, O.dptr(qa))
        O.lib().orc_decode_quat_oct(int(y), 10, 12, O.dptr(qb))
        out.append(2 * math.acos(min(1.0, abs(qa @ qb))))
    return np.array(out)


def head_cov(quat_word, log_words):
    """R diag(S^2) R^T of a head rotation word (QuatOct32) and its half-float log scales."""
    S = np.exp(np.array([O.lib().orc_decode_half(int(h)) for h in log_words]))
    R = quat_matrix(quat_word)
    return R @ np.diag(S * S) @ R.T


def compare(g, ref, eyes, scale_frac=0.0005, scale_codes=1, basis_flips=0.0):
    """Assert word parity; returns counts. basis_flips: the fraction of the
    sampled head rotations that may differ when the head basis comes from
    fromCov (an eigenbasis is unique up to the signs of two axes, which f32
    and f64 Jacobi sweeps may pick differently near a tie); those must still
    describe the same ellipsoid."""
    o, margin, em = ref["words"], ref["margin"], ref["eye_margin"]
    n = o.shape[0]
    rep = {}
    # head words: written when the splat is kept (not invisible, opacity >= minAlpha after MIP)
    gh, oh = g[:, 0] != 0, o[:, 0] != 0
    flips = np.nonzero(gh != oh)[0]
    assert np.all(margin[flips] < TOL), f"{np.count_nonzero(margin[flips] >= TOL)} head flips away from thresholds"
    both = np.nonzero(gh & oh)[0]
    rep["heads"] = both.size
    qx = lambda w: ((w[:, 0] & 0xFFFF) << 4) | ((w[:, 5] >> 24) & 15)
    qy = lambda w: ((w[:, 0] >> 16) << 4) | (w[:, 5] >> 28)
    gb, ob = g[both], o[both]
    for f in (qx, qy):
        d = np.abs(f(gb).astype(np.int64) - f(ob).astype(np.int64))
        assert np.mean(d <= 1) > 0.999 and d.max() <= 4, f"centre direction codes: max {d.max()}"
    gd, od = gb[:, 1].view(np.float32), ob[:, 1].view(np.float32)
    assert np.allclose(gd, od, rtol=2e-6), "centre distance"
    # rgb: one step of the coarser code
    buf = (C.c_double * 3)()
    sel = np.arange(both.size)[:: max(1, both.size // 3000)]
    for i in sel:
        O.lib().orc_decode_ext_rgb(int(gb[i, 2]), buf)
        a = np.array(buf[:])
        O.lib().orc_decode_ext_rgb(int(ob[i, 2]), buf)
        b = np.array(buf[:])
        step = 2.0 ** (max(gb[i, 2] >> 27, ob[i, 2] >> 27) - 15.0) / 255.0
        assert np.all(np.abs(a - b) <= step * 1.0001 + 1e-12), "rgb word"
    ang = quat_angle(gb[sel, 3], ob[sel, 3])
    off = np.nonzero(ang >= 1.2e-2)[0]
    rep["basis_flips"] = int(off.size)
    assert off.size <= basis_flips * sel.size, f"rotation codes differ by {ang.max()} rad"
    for k in off:
        i = sel[k]
        cg = head_cov(gb[i, 3], [gb[i, 4] & 0xFFFF, gb[i, 4] >> 16, gb[i, 5] & 0xFFFF])
        co = head_cov(ob[i, 3], [ob[i, 4] & 0xFFFF, ob[i, 4] >> 16, ob[i, 5] & 0xFFFF])
        assert np.max(np.abs(cg - co)) <= 2e-2 * np.max(np.abs(co)), "a flipped head basis is another ellipsoid"
    for lo, hi, word in ((0, 16, 4), (16, 32, 4), (0, 16, 5)):
        a = half((gb[sel, word] >> lo) & 0xFFFF)
        b = half((ob[sel, word] >> lo) & 0xFFFF)
        assert np.all(np.abs(a - b) <= 2e-3 * np.maximum(1.0, np.abs(b))), "half-float log scales"
    dop = np.abs(((gb[:, 5] >> 16) & 0xFF).astype(np.int64) - ((ob[:, 5] >> 16) & 0xFF))
    assert dop.max() <= 1, "opacity byte"
    # eye words
    for e in range(eyes):
        w0, w1 = 6 + 2 * e, 7 + 2 * e
        gv, ov = g[:, w1] != 0, o[:, w1] != 0
        fl = np.nonzero(gv != ov)[0]
        assert np.all(em[e][fl] < TOL), f"eye {e}: {np.count_nonzero(em[e][fl] >= TOL)} visibility flips away from thresholds"
        bv = np.nonzero(gv & ov & (em[e] >= TOL))[0]
        rep[f"eye{e}_visible"] = bv.size
        rep[f"eye{e}_flips"] = fl.size
        ga, oa = g[bv], o[bv]
        for sh in (0, 16):
            d = np.abs(((ga[:, w0] >> sh) & 0xFFFF).astype(np.int64) - ((oa[:, w0] >> sh) & 0xFFFF))
            assert np.mean(d <= 1) >= 1 - scale_frac and d.max() <= 64, f"eye {e} ndc codes: max {d.max()}"
        for sh, mask in ((10, 0x7FF), (21, 0x7FF)):
            d = np.abs(((ga[:, w1] >> sh) & mask).astype(np.int64) - ((oa[:, w1] >> sh) & mask))
            frac = np.mean(d > scale_codes)
            rep[f"eye{e}_scale{sh}_off"] = float(frac)
            assert frac <= scale_frac, f"eye {e} log-scale codes: {frac:.5f} off by more than {scale_codes}"
        # the angle, where the bound is clearly elongated
        s1 = (oa[:, w1] >> 10) & 0x7FF
        s2 = oa[:, w1] >> 21
        elong = s1 > s2 + 4
        da = np.abs((ga[:, w1] & 1023).astype(np.int64) - (oa[:, w1] & 1023))
        da = np.minimum(da, 1024 - da)[elong]
        assert da.size == 0 or np.mean(da <= 1) >= 1 - scale_frac, f"eye {e} angle codes: max {da.max()}"
    return rep


def quat_matrix(code):
    from scipy.spatial.transform import Rotation

    q = np.zeros(4)
    O.lib().orc_decode_quat_oct(int(code), 10, 12, O.dptr(q))
    return Rotation.from_quat(q).as_matrix()


def compare_raygs(g, ref, eyes, frac=0.0005):
    """raygs: head [dir, distance, rgb, opacity8 | low], eye [frame quat, c^2, half2 ln lengths]."""
    o, margin, em = ref["words"], ref["margin"], ref["eye_margin"]
    rep = {}
    gh, oh = g[:, 0] != 0, o[:, 0] != 0
    flips = np.nonzero(gh != oh)[0]
    assert np.all(margin[flips] < TOL)
    both = np.nonzero(gh & oh)[0]
    gb, ob = g[both], o[both]
    for f in (lambda w: ((w[:, 0] & 0xFFFF) << 4) | ((w[:, 3] >> 24) & 15), lambda w: ((w[:, 0] >> 16) << 4) | (w[:, 3] >> 28)):
        d = np.abs(f(gb).astype(np.int64) - f(ob).astype(np.int64))
        assert np.mean(d <= 1) > 0.999 and d.max() <= 4
    assert np.allclose(gb[:, 1].view(np.float32), ob[:, 1].view(np.float32), rtol=2e-6)
    assert np.abs((gb[:, 3] & 0xFF).astype(np.int64) - (ob[:, 3] & 0xFF)).max() <= 1
    rep["heads"] = both.size
    for e in range(eyes):
        wq, wc, wl = 4 + 3 * e, 5 + 3 * e, 6 + 3 * e
        gv, ov = g[:, wc] != 0, o[:, wc] != 0
        fl = np.nonzero(gv != ov)[0]
        assert np.all(em[e][fl] < TOL), f"eye {e}: {np.count_nonzero(em[e][fl] >= TOL)} flips away from thresholds"
        bv = np.nonzero(gv & ov & (em[e] >= TOL))[0]
        rep[f"eye{e}_visible"] = bv.size
        ga, oa = g[bv], o[bv]
        assert np.allclose(ga[:, wc].view(np.float32), oa[:, wc].view(np.float32), rtol=2e-5), "c^2"
        lg = np.stack([half(ga[::50, wl] & 0xFFFF), half(ga[::50, wl] >> 16)])
        lo = np.stack([half(oa[::50, wl] & 0xFFFF), half(oa[::50, wl] >> 16)])
        assert np.all(np.abs(lg - lo) <= 2e-3 * np.maximum(1.0, np.abs(lo))), "half-float log lengths"
        # the quad's normal always; its in-plane rotation where the section is elongated
        elong = (lo[1] - lo[0]) > 2e-2
        bad = 0
        for k, i in enumerate(range(0, bv.size, 50)):
            Mg, Mo = quat_matrix(ga[i, wq]), quat_matrix(oa[i, wq])
            if abs(abs(Mg[:, 2] @ Mo[:, 2]) - 1.0) > 1e-4:
                bad += 1
            elif elong[k] and np.linalg.norm(Mg - Mo) > 3e-2:
                bad += 1
        rep[f"eye{e}_frame_off"] = bad
        assert bad <= max(1, int(frac * bv.size / 50)), f"eye {e}: {bad} quad frames differ"
    return rep


def run(table, rig, prm, method):
    out = Projector().project_ray(table, rig, prm, method)
    torch.cuda.synchronize()
    return out["words"].cpu().numpy().view(np.uint32), out["keys"].cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("method", ["3dgut", "htgs"])
def test_plain_small_mono(cuda, method):
    spec = W.plain_small_spec(seed=2)
    planes = W.generate(spec)
    rig = W.plain_small_camera()
    prm = Params(sort_bits=32)
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    g, keys = run(table, rig, prm, method)
    ref = O.project_ray_f32(method, planes, 0, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare(g, ref, 1)
    print(method, rep, flush=True)
    assert rep["eye0_visible"] > 30000
    bad = np.nonzero(keys != ref["keys"])[0]
    assert np.all(ref["margin"][bad] < TOL), f"{bad.size} keys differ"


@pytest.mark.parametrize("method", ["3dgut", "htgs"])
def test_sh3_orbit_and_object_transform(cuda, method):
    spec = W.plain_4m_spec(seed=17, count=40000)
    planes = W.generate(spec)
    obj = cam.translation([0.3, -0.8, 1.5])
    rot = np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    obj[:3, :3] = 1.6 * rot
    rig = W.orbit_camera(9, 9.0, 1280, 720)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=0.3)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3, object_to_world=obj)
    g, keys = run(table, rig, prm, method)
    ref = O.project_ray_f32(method, planes, 3, O.view_from_rig(rig, object_to_world=obj),
                            O.params(sort_bits=32, cull_keys=False, sort_guard=0.3))
    rep = compare(g, ref, 1)
    print(method, rep, flush=True)
    assert rep["eye0_visible"] > 10000
    bad = np.nonzero(keys != ref["keys"])[0]
    assert np.all(ref["margin"][bad] < TOL), f"{bad.size} keys differ"


@pytest.mark.parametrize("method", ["3dgut", "htgs"])
def test_stereo_two_eyes(cuda, method):
    spec = W.stereo_16m_spec(count=30000, seed=4)
    planes = W.generate(spec)
    rig = W.stereo_camera(width=512, height=544)
    prm = Params(sort_bits=32)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3)
    g, _ = run(table, rig, prm, method)
    assert g.shape[1] == 10
    ref = O.project_ray_f32(method, planes, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare(g, ref, 2)
    print(method, rep, flush=True)
    assert rep["eye0_visible"] > 5000 and rep["eye1_visible"] > 5000


def test_ext_words_decode_like_planes(cuda):
    # ext words: the kernel decodes them; the oracle takes the decoded planes
    n = 6000
    a, b, chans = ext_table(n, 21, 2)
    planes = np.zeros((59, n), dtype=np.float32)
    c, s, q, rgb, op = (np.zeros(3), np.zeros(3), np.zeros(4), np.zeros(3), C.c_double())
    for i in range(n):
        ai = np.ascontiguousarray(a[i])
        bi = np.ascontiguousarray(b[i])
        O.lib().orc_ext_decode(O.uptr(ai), O.uptr(bi), O.dptr(c), O.dptr(s), O.dptr(q), O.dptr(rgb), C.byref(op))
        planes[0:3, i], planes[3:6, i], planes[6:10, i], planes[10, i], planes[11:14, i] = c, np.log(s), q, op.value, rgb
    rig = W.orbit_camera(5, 7.0, 1024, 768)
    table = DeviceTable.from_ext(a, b, n, [], sh_codec=0, sh_degree=0)
    for method in ("3dgut", "htgs"):
        g, _ = run(table, rig, Params(sort_bits=32), method)
        ref = O.project_ray_f32(method, planes, 0, O.view_from_rig(rig), O.params(sort_bits=32))
        rep = compare(g, ref, 1)
        print(method, rep, flush=True)


# A transform that is not a rotation times a uniform scale (shear and
# anisotropy): transformSplat's general case, fromCov of H Σ H^T
# (splatTransform.tsl.ts:154-165; fromCov splats.tsl.ts:189-198).
def non_uniform_obj():
    obj = cam.translation([0.3, 0.2, -0.5])
    obj[:3, :3] = np.array([[1.3, 0.2, 0.0], [0.0, 0.8, 0.1], [0.1, 0.0, 1.1]])
    return obj


@pytest.mark.parametrize("method", ["3dgut", "htgs"])
def test_non_uniform_object_transform(cuda, method):
    spec = W.plain_4m_spec(seed=19, count=30000)
    planes = W.generate(spec)
    obj = non_uniform_obj()
    rig = W.orbit_camera(11, 8.0, 1024, 768)
    prm = Params(sort_bits=32, cull_keys=False, sort_guard=0.3)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3, object_to_world=obj)
    g, keys = run(table, rig, prm, method)
    ref = O.project_ray_f32(method, planes, 3, O.view_from_rig(rig, object_to_world=obj),
                            O.params(sort_bits=32, cull_keys=False, sort_guard=0.3))
    rep = compare(g, ref, 1, basis_flips=0.002)
    print(method, rep, flush=True)
    assert rep["eye0_visible"] > 5000
    bad = np.nonzero(keys != ref["keys"])[0]
    assert np.all(ref["margin"][bad] < TOL), f"{bad.size} keys differ"


def cov_ray_planes(a, b, chans, sh_codec):
    """f32 planes of a covSplats table in the splat form: the decoded
    covariance through the oracle's fromCov (pinned in test_oracle_ray),
    the SH bundles decoded to coefficients."""
    L = O.lib()
    n = a.shape[0]
    planes = np.zeros((59, n), dtype=np.float32)
    c, cov, rgb, op = np.zeros(3), np.zeros(6), np.zeros(3), C.c_double()
    sc, q = np.zeros(3), np.zeros(4)
    words = np.zeros(16, dtype=np.uint32)
    coeffs = np.zeros(45)
    for i in range(n):
        L.orc_cov_decode(O.uptr(np.ascontiguousarray(a[i])), O.uptr(np.ascontiguousarray(b[i])), O.dptr(c), O.dptr(cov),
                         O.dptr(rgb), C.byref(op))
        L.orc_from_cov(O.dptr(cov), O.dptr(sc), O.dptr(q))
        with np.errstate(divide="ignore"):
            planes[3:6, i] = np.log(sc)
        planes[0:3, i], planes[6:10, i], planes[10, i], planes[11:14, i] = c, q, op.value, rgb
        for t, ch in enumerate(chans):
            words[4 * t:4 * t + 4] = ch[i]
        L.orc_sh_decode(sh_codec, O.uptr(words), 3, O.dptr(coeffs))
        planes[14:59, i] = coeffs
    return planes


@pytest.mark.parametrize("method", ["3dgut", "htgs", "raygs"])
def test_cov_table(cuda, method):
    # a covSplats table in the splat form (GpuCovSplatReader.loadSplat = fromCov,
    # splatAccess.ts:103-105), under no transform and under a non-uniform one
    from tests.test_project_gpu import cov_table

    n = 12000
    a, b, chans = cov_table(n, 23, 2, lod=True)
    planes = cov_ray_planes(a, b, chans, 2)
    rig = W.orbit_camera(12, 7.0, 1024, 768)
    for obj in (None, non_uniform_obj()):
        table = DeviceTable.from_cov(a, b, n, chans, sh_codec=2, sh_degree=3, object_to_world=obj)
        g, keys = run(table, rig, Params(sort_bits=32), method)
        ref = O.project_ray_f32(method, planes, 3, O.view_from_rig(rig, object_to_world=obj), O.params(sort_bits=32))
        rep = compare_raygs(g, ref, 1, frac=0.002) if method == "raygs" else compare(g, ref, 1, basis_flips=0.002)
        print(method, "non-uniform" if obj is not None else "plain", rep, flush=True)
        bad = np.nonzero(keys != ref["keys"])[0]
        assert np.all(ref["margin"][bad] < TOL), f"{bad.size} keys differ"


def test_raygs_mono_and_object_transform(cuda):
    spec = W.plain_small_spec(seed=3)
    planes = W.generate(spec)
    rig = W.plain_small_camera()
    table = DeviceTable.from_f32_planes(planes, spec.count, 0)
    g, keys = run(table, rig, Params(sort_bits=32), "raygs")
    assert g.shape[1] == 7
    ref = O.project_ray_f32("raygs", planes, 0, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare_raygs(g, ref, 1)
    print("raygs", rep, flush=True)
    assert rep["eye0_visible"] > 30000
    bad = np.nonzero(keys != ref["keys"])[0]
    assert np.all(ref["margin"][bad] < TOL)
    spec = W.plain_4m_spec(seed=19, count=30000)
    planes = W.generate(spec)
    obj = cam.translation([0.3, -0.8, 1.5])
    obj[:3, :3] = 0.7 * np.array([[0.36, 0.48, -0.8], [-0.8, 0.6, 0.0], [0.48, 0.64, 0.6]])
    rig = W.orbit_camera(12, 8.0, 1280, 720)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3, object_to_world=obj)
    g, _ = run(table, rig, Params(sort_bits=32), "raygs")
    ref = O.project_ray_f32("raygs", planes, 3, O.view_from_rig(rig, object_to_world=obj), O.params(sort_bits=32))
    rep = compare_raygs(g, ref, 1)
    print("raygs sh3 transform", rep, flush=True)
    assert rep["eye0_visible"] > 8000


def test_raygs_stereo(cuda):
    spec = W.stereo_16m_spec(count=20000, seed=6)
    planes = W.generate(spec)
    rig = W.stereo_camera(width=512, height=544)
    table = DeviceTable.from_f32_planes(planes, spec.count, 3)
    g, _ = run(table, rig, Params(sort_bits=32), "raygs")
    assert g.shape[1] == 10
    ref = O.project_ray_f32("raygs", planes, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare_raygs(g, ref, 2)
    print("raygs stereo", rep, flush=True)
    assert rep["eye0_visible"] > 5000 and rep["eye1_visible"] > 5000

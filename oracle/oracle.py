"""ctypes binding of the oracle (oracle/liboracle.so) for the tests.

TEST INFRASTRUCTURE ONLY: the checker of the HIP path. Only tests/, the
smoke check and the bench's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_d = C.c_double
_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)


class OrcParams(C.Structure):
    _fields_ = [
        ("max_std_dev", _d), ("min_alpha", _d), ("falloff", _d), ("lod_inflate", C.c_int),
        ("blur_amount", _d), ("pre_blur_amount", _d), ("clip_xy", _d), ("min_pixel_radius", _d),
        ("max_pixel_radius", _d), ("min_contribution", _d), ("focal_distance", _d), ("aperture_tan", _d),
        ("sort_guard", _d), ("max_sh_degree", C.c_int), ("sort_bits", C.c_int), ("cull_keys", C.c_int),
        ("key_per_eye", C.c_int),
    ]


class OrcEye(C.Structure):
    _fields_ = [("from_head", _d * 16), ("projection", _d * 16), ("ortho", C.c_int)]


class OrcView(C.Structure):
    _fields_ = [("object_to_head", _d * 16), ("view_origin", _d * 3), ("eyes", C.c_int), ("eye", OrcEye * 2),
                ("size", _d * 2), ("ndc_z_min", _d)]


class OrcSplat(C.Structure):
    _fields_ = [("center", _d * 3), ("cov", _d * 6), ("rgb", _d * 3), ("opacity", _d)]


class OrcEyeOut(C.Structure):
    _fields_ = [("visible", C.c_int), ("words", C.c_uint32 * 4), ("ndc", _d * 2), ("view_z", _d), ("a", _d),
                ("b", _d), ("d", _d), ("conic", _d * 3), ("opacity", _d), ("scale1", _d), ("scale2", _d),
                ("angle", _d), ("lambda1", _d), ("lambda2", _d), ("margin", _d), ("distance", _d),
                ("key", C.c_uint32)]


class OrcOut(C.Structure):
    _fields_ = [("eye", OrcEyeOut * 2), ("rgb", _d * 3), ("distance", _d), ("invisible", C.c_int),
                ("guarded", C.c_int), ("visible", C.c_int), ("key", C.c_uint32), ("margin", _d)]


class OrcTableF32(C.Structure):
    _fields_ = [("count", C.c_uint32), ("stride", C.c_size_t), ("center", C.c_void_p), ("log_scale", C.c_void_p),
                ("quat", C.c_void_p), ("opacity", C.c_void_p), ("rgb", C.c_void_p), ("sh", C.c_void_p),
                ("sh_degree", C.c_int)]


class OrcBatchOut(C.Structure):
    _fields_ = [("words", C.c_void_p), ("keys", C.c_void_p), ("conic", C.c_void_p), ("opacity", C.c_void_p),
                ("ndc", C.c_void_p), ("view_z", C.c_void_p), ("rgb", C.c_void_p), ("distance", C.c_void_p),
                ("margin", C.c_void_p), ("visible", C.c_void_p)]


class OrcRaster(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("view", _d * 16), ("proj", _d * 16),
                ("ndc_z_min", _d), ("max_std_dev", _d), ("min_alpha", _d), ("blur", _d), ("clip_xy", _d),
                ("t_stop", _d), ("sh_degree", C.c_int)]


DIFF_PARAMS = 59


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        h = C.CDLL(LIB)
        sig = {
            "orc_key_invalid": (C.c_uint32, [C.c_int]),
            "orc_key_valid_max": (C.c_uint32, [C.c_int]),
            "orc_key_compute": (C.c_uint32, [_d, C.c_int]),
            "orc_keys_compute": (None, [C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]),
            "orc_sort": (C.c_int64, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
            "orc_encode_half": (C.c_uint32, [_d]),
            "orc_decode_half": (_d, [C.c_uint32]),
            "orc_encode_ext_rgb": (C.c_uint32, [_d, _d, _d]),
            "orc_decode_ext_rgb": (None, [C.c_uint32, _dp]),
            "orc_encode_oct": (C.c_uint32, [_dp, C.c_int]),
            "orc_decode_oct": (None, [C.c_uint32, C.c_int, _dp]),
            "orc_encode_quat_oct": (C.c_uint32, [_dp, C.c_int, C.c_int]),
            "orc_decode_quat_oct": (None, [C.c_uint32, C.c_int, C.c_int, _dp]),
            "orc_packed_encode": (None, [_dp, _dp, _dp, _dp, _d, _dp, _u32p]),
            "orc_packed_decode": (None, [_u32p, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_ext_encode": (None, [_dp, _dp, _dp, _dp, _d, _u32p, _u32p]),
            "orc_ext_decode": (None, [_u32p, _u32p, _dp, _dp, _dp, _dp, _dp]),
            "orc_cov_encode": (None, [_dp, _dp, _dp, _d, _u32p, _u32p]),
            "orc_cov_decode": (None, [_u32p, _u32p, _dp, _dp, _dp, _dp]),
            "orc_cov_encode_correlation": (_d, [_d, _d, _d]),
            "orc_cov_decode_correlation": (_d, [_d]),
            "orc_from_cov": (None, [_dp, _dp, _dp]),
            "orc_sh_basis": (None, [_dp, _dp]),
            "orc_sh_evaluate": (None, [_dp, _dp, C.c_int, _dp]),
            "orc_raster_binning": (None, [C.c_int]),
            "orc_sh_apply_table": (None, [C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_void_p]),
            "orc_sh_scale_encode": (C.c_uint32, [_d]),
            "orc_sh_scale_decode": (_d, [C.c_uint32]),
            "orc_sh_words": (C.c_int, [C.c_int, C.c_int]),
            "orc_sh_encode": (None, [C.c_int, _dp, C.c_int, _u32p]),
            "orc_sh_decode": (None, [C.c_int, _u32p, C.c_int, _dp]),
            "orc_cov_from_scale_quat": (None, [_dp, _dp, _dp]),
            "orc_params_default": (None, [C.POINTER(OrcParams)]),
            "orc_project_ewa": (None, [C.POINTER(OrcSplat), C.POINTER(OrcView), C.POINTER(OrcParams),
                                       C.POINTER(OrcOut)]),
            "orc_project_keys_table_f32": (None, [C.POINTER(OrcTableF32), C.POINTER(OrcView), C.POINTER(OrcParams),
                                                   C.c_void_p, C.c_void_p]),
            "orc_project_table_f32": (None, [C.POINTER(OrcTableF32), C.POINTER(OrcView), C.POINTER(OrcParams),
                                             C.POINTER(OrcBatchOut)]),
            "orc_project_table_packed": (None, [C.c_uint32, C.c_void_p, _dp, C.c_int, C.c_int, C.c_void_p,
                                                C.POINTER(OrcView), C.POINTER(OrcParams), C.POINTER(OrcBatchOut)]),
            "orc_project_table_ext": (None, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                             C.POINTER(OrcView), C.POINTER(OrcParams), C.POINTER(OrcBatchOut)]),
            "orc_project_table_cov": (None, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                             C.POINTER(OrcView), C.POINTER(OrcParams), C.POINTER(OrcBatchOut)]),
            "orc_view_from_camera": (None, [_dp, _dp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, _d, _d, _d,
                                            C.POINTER(OrcView)]),
            "orc_diff_render": (C.c_int, [_dp, C.c_int, C.POINTER(OrcRaster), _dp, _dp]),
            "orc_diff_loss": (_d, [_dp, C.c_int, C.POINTER(OrcRaster), _dp]),
            "orc_diff_grad": (_d, [_dp, C.c_int, C.POINTER(OrcRaster), _dp, _dp]),
            "orc_diff_freeze": (C.c_void_p, [_dp, C.c_int, C.POINTER(OrcRaster), _dp]),
            "orc_diff_loss_frozen": (_d, [C.c_void_p, _dp, C.c_int, C.POINTER(OrcRaster), _dp]),
            "orc_diff_free": (None, [C.c_void_p]),
            "orc_sh_basis_grad": (None, [_dp, _dp]),
            "orc_draw": (None, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_int, C.c_int, _d, _d, _d, C.c_int,
                                _dp]),
            "orc_draw_ray": (None, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                    C.c_uint32, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _d, _d, _d, C.c_int, _dp]),
            "orc_project_ray": (None, [C.c_int, C.POINTER(OrcRaySplat), C.POINTER(OrcView), C.POINTER(OrcParams),
                                       C.POINTER(OrcRayOut)]),
            "orc_project_ray_table_f32": (None, [C.c_int, C.POINTER(OrcTableF32), C.POINTER(OrcView),
                                                 C.POINTER(OrcParams), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
            "orc_encode_center": (None, [_dp, _u32p, _u32p, _u32p]),
            "orc_decode_center": (None, [C.c_uint32, C.c_uint32, C.c_uint32, _dp]),
            "orc_encode_quat_oct32_gpu": (C.c_uint32, [_dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _arr(x, dtype):
    a = np.ascontiguousarray(x, dtype=dtype)
    return a


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def uptr(a: np.ndarray):
    return a.ctypes.data_as(_u32p)


# ---- sort keys and the CPU sort ----------------------------------------------

def key_invalid(bits: int) -> int:
    return int(lib().orc_key_invalid(bits))


def key_valid_max(bits: int) -> int:
    return int(lib().orc_key_valid_max(bits))


def key_compute(distance: float, bits: int) -> int:
    return int(lib().orc_key_compute(float(distance), bits))


def keys_compute(distances, bits: int) -> np.ndarray:
    d = _arr(distances, np.float64)
    out = np.empty(d.size, dtype=np.uint32)
    lib().orc_keys_compute(d.ctypes.data, d.size, bits, out.ctypes.data)
    return out


def sort(keys, bits: int, count=None):
    """SortKeys.sort: (ordering[:count], visible)."""
    k = _arr(keys, np.uint32)
    n = k.size if count is None else count
    ordering = np.zeros(max(n, 1), dtype=np.uint32)
    visible = lib().orc_sort(k.ctypes.data, n, bits, ordering.ctypes.data)
    if visible < 0:
        raise ValueError("keys outside the reference's contract (above 2^24 at 24 bits)")
    return ordering[:n], int(visible)


# ---- cameras and projection ----------------------------------------------------

def params(**kw) -> OrcParams:
    p = OrcParams()
    lib().orc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, int(v) if isinstance(v, bool) else v)
    return p


def view_from_rig(rig, object_to_world=None) -> OrcView:
    """An OrcView from a radix_amd.camera.CameraRig (column-major inputs)."""
    from radix_amd.camera import cm

    obj = cm(np.eye(4) if object_to_world is None else object_to_world)
    head = cm(rig.head)
    eyes = len(rig.eyes)
    views = [cm(e.view) for e in rig.eyes]
    projs = [cm(e.projection) for e in rig.eyes]
    vp = (C.c_void_p * 2)(*[v.ctypes.data for v in views] + [None] * (2 - eyes))
    pp = (C.c_void_p * 2)(*[p.ctypes.data for p in projs] + [None] * (2 - eyes))
    orth = (C.c_int * 2)(*[1 if e.ortho else 0 for e in rig.eyes] + [0] * (2 - eyes))
    out = OrcView()
    lib().orc_view_from_camera(dptr(obj), dptr(head), eyes, vp, pp, orth, float(rig.width), float(rig.height),
                               float(rig.ndc_z_min), C.byref(out))
    out._keep = (obj, head, views, projs)
    return out


def project_f32(planes: np.ndarray, sh_degree: int, view: OrcView, prm: OrcParams):
    """The oracle over an f32-planes table: dict of per-splat arrays."""
    n = planes.shape[1]
    eyes = view.eyes
    per_eye = bool(prm.key_per_eye) and eyes == 2
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    base = planes.ctypes.data
    st = n * 4
    t = OrcTableF32(n, n, base, base + 3 * st, base + 6 * st, base + 10 * st, base + 11 * st,
                    base + 14 * st if sh_degree > 0 else None, sh_degree)
    return _run_batch(lambda o: lib().orc_project_table_f32(C.byref(t), C.byref(view), C.byref(prm), C.byref(o)),
                      n, eyes, per_eye)


def project_keys_f32(planes: np.ndarray, view: OrcView, prm: OrcParams):
    """The keys-only pass over an f32-planes table: keys (n,) and margins (n,)."""
    n = planes.shape[1]
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    base = planes.ctypes.data
    st = n * 4
    t = OrcTableF32(n, n, base, base + 3 * st, base + 6 * st, base + 10 * st, base + 11 * st, None, 0)
    keys = np.zeros(n, dtype=np.uint32)
    margin = np.zeros(n, dtype=np.float64)
    lib().orc_project_keys_table_f32(C.byref(t), C.byref(view), C.byref(prm), keys.ctypes.data, margin.ctypes.data)
    return {"keys": keys, "margin": margin}


class OrcRaySplat(C.Structure):
    _fields_ = [("center", _d * 3), ("scales", _d * 3), ("quat", _d * 4), ("rgb", _d * 3), ("opacity", _d)]


class OrcRayOut(C.Structure):
    _fields_ = [("words", C.c_uint32 * 10), ("visible", C.c_int), ("invisible", C.c_int), ("guarded", C.c_int),
                ("eye_visible", C.c_int * 2), ("key", C.c_uint32), ("distance", _d), ("head_scales", _d * 3),
                ("head_opacity", _d), ("bound", (_d * 5) * 2), ("frame_axes", ((_d * 3) * 2) * 2), ("frame_c2", _d * 2),
                ("frame_lengths", (_d * 2) * 2), ("eye_margin", _d * 2), ("margin", _d)]


METHODS = {"3dgut": 1, "htgs": 2, "raygs": 3}


def project_ray(method: str, center, scales, quat, rgb, opacity, view: OrcView, prm: OrcParams) -> OrcRayOut:
    """One object-space splat (scales, not log-scales) through orc_project_ray."""
    s = OrcRaySplat((_d * 3)(*center), (_d * 3)(*scales), (_d * 4)(*quat), (_d * 3)(*rgb), float(opacity))
    o = OrcRayOut()
    lib().orc_project_ray(METHODS[method], C.byref(s), C.byref(view), C.byref(prm), C.byref(o))
    return o


def project_ray_f32(method: str, planes: np.ndarray, sh_degree: int, view: OrcView, prm: OrcParams):
    """The ray oracle over an f32-planes table: words (n, 6 + 2 eyes; raygs 4 + 3 eyes), keys, margins (n,) and (eyes, n)."""
    n = planes.shape[1]
    eyes = view.eyes
    planes = np.ascontiguousarray(planes, dtype=np.float32)
    base = planes.ctypes.data
    st = n * 4
    t = OrcTableF32(n, n, base, base + 3 * st, base + 6 * st, base + 10 * st, base + 11 * st,
                    base + 14 * st if sh_degree > 0 else None, sh_degree)
    words = np.zeros((n, (4 + 3 * eyes) if method == "raygs" else (6 + 2 * eyes)), dtype=np.uint32)
    keys = np.zeros(n, dtype=np.uint32)
    margin = np.zeros(n, dtype=np.float64)
    eye_margin = np.zeros((eyes, n), dtype=np.float64)
    lib().orc_project_ray_table_f32(METHODS[method], C.byref(t), C.byref(view), C.byref(prm), words.ctypes.data,
                                    keys.ctypes.data, margin.ctypes.data, eye_margin.ctypes.data)
    return {"words": words, "keys": keys, "margin": margin, "eye_margin": eye_margin}


def project_packed(words: np.ndarray, enc, sh_channels, sh_degree: int, view: OrcView, prm: OrcParams):
    n = words.shape[0]
    eyes = view.eyes
    per_eye = bool(prm.key_per_eye) and eyes == 2
    w = np.ascontiguousarray(words, dtype=np.uint32)
    e = np.asarray(enc, dtype=np.float64)
    chans = [np.ascontiguousarray(c, dtype=np.uint32) for c in (sh_channels or [])]
    cp = (C.c_void_p * 4)(*[c.ctypes.data for c in chans] + [None] * (4 - len(chans)))
    codec = 3 if chans and sh_degree > 0 else 0
    return _run_batch(lambda o: lib().orc_project_table_packed(n, w.ctypes.data, dptr(e), codec, sh_degree, cp,
                                                               C.byref(view), C.byref(prm), C.byref(o)),
                      n, eyes, per_eye)


def project_ext(ext_a, ext_b, sh_channels, sh_codec: int, sh_degree: int, view: OrcView, prm: OrcParams):
    n = ext_a.shape[0]
    eyes = view.eyes
    per_eye = bool(prm.key_per_eye) and eyes == 2
    a = np.ascontiguousarray(ext_a, dtype=np.uint32)
    b = np.ascontiguousarray(ext_b, dtype=np.uint32)
    chans = [np.ascontiguousarray(c, dtype=np.uint32) for c in (sh_channels or [])]
    cp = (C.c_void_p * 4)(*[c.ctypes.data for c in chans] + [None] * (4 - len(chans)))
    codec = sh_codec if chans and sh_degree > 0 else 0
    return _run_batch(lambda o: lib().orc_project_table_ext(n, a.ctypes.data, b.ctypes.data, codec, sh_degree, cp,
                                                            C.byref(view), C.byref(prm), C.byref(o)),
                      n, eyes, per_eye)


def project_cov(cov_a, cov_b, sh_channels, sh_codec: int, sh_degree: int, view: OrcView, prm: OrcParams):
    """A covSplats table (covA, covB words) with ext (2) or packed (3) SH bundles."""
    n = cov_a.shape[0]
    eyes = view.eyes
    per_eye = bool(prm.key_per_eye) and eyes == 2
    a = np.ascontiguousarray(cov_a, dtype=np.uint32)
    b = np.ascontiguousarray(cov_b, dtype=np.uint32)
    chans = [np.ascontiguousarray(c, dtype=np.uint32) for c in (sh_channels or [])]
    cp = (C.c_void_p * 4)(*[c.ctypes.data for c in chans] + [None] * (4 - len(chans)))
    codec = sh_codec if chans and sh_degree > 0 else 0
    return _run_batch(lambda o: lib().orc_project_table_cov(n, a.ctypes.data, b.ctypes.data, codec, sh_degree, cp,
                                                            C.byref(view), C.byref(prm), C.byref(o)),
                      n, eyes, per_eye)


def _run_batch(call, n, eyes, per_eye):
    out = {
        "words": np.zeros((eyes, n, 4), dtype=np.uint32),
        "keys": np.zeros(((2 if per_eye else 1), n), dtype=np.uint32),
        "conic": np.zeros((eyes, 3, n)),
        "opacity": np.zeros((eyes, n)),
        "ndc": np.zeros((eyes, 2, n)),
        "view_z": np.zeros((eyes, n)),
        "rgb": np.zeros((3, n)),
        "distance": np.zeros(((2 if per_eye else 1), n)),
        "margin": np.zeros(n),
        "visible": np.zeros((eyes, n), dtype=np.uint8),
    }
    o = OrcBatchOut(*[out[k].ctypes.data for k in
                      ("words", "keys", "conic", "opacity", "ndc", "view_z", "rgb", "distance", "margin", "visible")])
    call(o)
    if not per_eye:
        out["keys"] = out["keys"][0]
        out["distance"] = out["distance"][0]
    return out


# ---- the differentiable compositor (raster.c) ---------------------------------


def raster(width: int, height: int, view, proj, sh_degree: int = 3, ndc_z_min: float = 0.0,
           max_std_dev: float = float(np.sqrt(8.0)), min_alpha: float = 0.5 / 255.0, blur: float = 0.3,
           clip_xy: float = 1.4, t_stop: float = 1e-4) -> OrcRaster:
    """view: world -> view 4x4 (row-major numpy); proj: 4x4 (row-major)."""
    r = OrcRaster()
    r.width, r.height = width, height
    r.view[:] = list(np.asarray(view, dtype=np.float64).T.reshape(16))
    r.proj[:] = list(np.asarray(proj, dtype=np.float64).T.reshape(16))
    r.ndc_z_min, r.max_std_dev, r.min_alpha, r.blur, r.clip_xy, r.t_stop = ndc_z_min, max_std_dev, min_alpha, blur, clip_xy, t_stop
    r.sh_degree = sh_degree
    return r


def diff_render(theta: np.ndarray, r: OrcRaster):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    n = theta.shape[0]
    img = np.zeros((3, r.height, r.width))
    alpha = np.zeros((r.height, r.width))
    drawn = lib().orc_diff_render(dptr(theta), n, C.byref(r), dptr(img), dptr(alpha))
    return img, alpha, drawn


def diff_loss(theta: np.ndarray, r: OrcRaster, target: np.ndarray) -> float:
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    return lib().orc_diff_loss(dptr(theta), theta.shape[0], C.byref(r), dptr(np.ascontiguousarray(target)))


def diff_grad(theta: np.ndarray, r: OrcRaster, target: np.ndarray):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    loss = lib().orc_diff_grad(dptr(theta), theta.shape[0], C.byref(r), dptr(np.ascontiguousarray(target)), dptr(g))
    return loss, g


class Frozen:
    """The loss with every pixel's contributor list fixed at theta0."""

    def __init__(self, theta0: np.ndarray, r: OrcRaster, target: np.ndarray):
        self.r = r
        self.target = np.ascontiguousarray(target, dtype=np.float64)
        t0 = np.ascontiguousarray(theta0, dtype=np.float64)
        self.h = lib().orc_diff_freeze(dptr(t0), t0.shape[0], C.byref(r), dptr(self.target))

    def loss(self, theta: np.ndarray) -> float:
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        return lib().orc_diff_loss_frozen(self.h, dptr(theta), theta.shape[0], C.byref(self.r), dptr(self.target))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_diff_free(self.h)
            self.h = None


def draw(words, order, count: int, width: int, height: int, max_std_dev: float = float(np.sqrt(8.0)),
         min_alpha: float = 0.5 / 255.0, falloff: float = 1.0, linear: bool = False) -> np.ndarray:
    """The ewa splat draw (src/render/splatDraw.ts:98-181): words (n, 4)
    uint32 of one eye, order (back to front) of which `count` are drawn.
    Returns (4, height, width) float64: premultiplied rgb and alpha, row 0
    at the top."""
    w = np.ascontiguousarray(words, dtype=np.uint32).reshape(-1, 4)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    count = int(min(count, o.size))
    out = np.zeros((4, height, width), dtype=np.float64)
    lib().orc_draw(w.ctypes.data, w.shape[0], o.ctypes.data if o.size else None, count, width, height,
                   max_std_dev, min_alpha, falloff, 1 if linear else 0, dptr(out))
    return out


def eye_from_head(rig, eye: int):
    """The eye's head -> eye matrix (row-major 4x4) and its rotation as a
    quaternion (x, y, z, w): SplatCamera.setViews' fromHead and the
    quaternion of its decomposition (src/render/splatCamera.ts:119-138)."""
    m = np.asarray(rig.eyes[eye].view, dtype=np.float64) @ np.linalg.inv(np.asarray(rig.head, dtype=np.float64))
    r = m[:3, :3].copy()
    sx = np.linalg.norm(r[:, 0]) * (-1.0 if np.linalg.det(r) < 0 else 1.0)
    r[:, 0] /= sx
    r[:, 1] /= np.linalg.norm(r[:, 1])
    r[:, 2] /= np.linalg.norm(r[:, 2])
    tr = r[0, 0] + r[1, 1] + r[2, 2]
    if tr > 0:
        k = 0.5 / np.sqrt(tr + 1.0)
        q = [(r[2, 1] - r[1, 2]) * k, (r[0, 2] - r[2, 0]) * k, (r[1, 0] - r[0, 1]) * k, 0.25 / k]
    elif r[0, 0] > r[1, 1] and r[0, 0] > r[2, 2]:
        k = 2.0 * np.sqrt(1.0 + r[0, 0] - r[1, 1] - r[2, 2])
        q = [0.25 * k, (r[0, 1] + r[1, 0]) / k, (r[0, 2] + r[2, 0]) / k, (r[2, 1] - r[1, 2]) / k]
    elif r[1, 1] > r[2, 2]:
        k = 2.0 * np.sqrt(1.0 + r[1, 1] - r[0, 0] - r[2, 2])
        q = [(r[0, 1] + r[1, 0]) / k, 0.25 * k, (r[1, 2] + r[2, 1]) / k, (r[0, 2] - r[2, 0]) / k]
    else:
        k = 2.0 * np.sqrt(1.0 + r[2, 2] - r[0, 0] - r[1, 1])
        q = [(r[0, 2] + r[2, 0]) / k, (r[1, 2] + r[2, 1]) / k, 0.25 * k, (r[1, 0] - r[0, 1]) / k]
    return m, np.asarray(q, dtype=np.float64)


def ray_channels(words: np.ndarray):
    """A ray method's words (n, head + eyes x eye words) as the projected
    bundle's channels w0, w1 (n, 4) and w2 (n, 2) (SplatProjection.ts:46-53)."""
    w = np.zeros((words.shape[0], 10), dtype=np.uint32)
    w[:, :words.shape[1]] = words
    return np.ascontiguousarray(w[:, :4]), np.ascontiguousarray(w[:, 4:8]), np.ascontiguousarray(w[:, 8:10])


def draw_ray(method: str, words, order, count: int, rig, eye: int = 0, max_std_dev: float = float(np.sqrt(8.0)),
             min_alpha: float = 0.5 / 255.0, falloff: float = 1.0, linear: bool = False) -> np.ndarray:
    """The draw of a ray method's words (n, head + eyes x eye words) for one
    eye of `rig`, splats order[:count] back to front. Returns (4, height,
    width) float64: premultiplied rgb and alpha, row 0 at the top."""
    c0, c1, c2 = ray_channels(np.asarray(words, dtype=np.uint32))
    o = np.ascontiguousarray(order, dtype=np.uint32)
    count = int(min(count, o.size))
    m, q = eye_from_head(rig, eye)
    fh = np.ascontiguousarray(m.T.reshape(-1))
    pr = np.ascontiguousarray(np.asarray(rig.eyes[eye].projection, dtype=np.float64).T.reshape(-1))
    out = np.zeros((4, rig.height, rig.width), dtype=np.float64)
    lib().orc_draw_ray(METHODS[method], c0.ctypes.data, c1.ctypes.data, c2.ctypes.data if len(rig.eyes) == 2 else None,
                       c0.shape[0], eye, o.ctypes.data if o.size else None, count, dptr(fh), dptr(q), dptr(pr),
                       1 if rig.eyes[eye].ortho else 0, rig.width, rig.height, max_std_dev, min_alpha, falloff,
                       1 if linear else 0, dptr(out))
    return out

"""ctypes binding of libradix_hip.so (include/radix_hip.h) and libradix_synth.so.

The product path runs only through these libraries: if libradix_hip.so is
missing or fails to load, every entry point raises; nothing falls back to a
CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libradix_hip.so")
SYNTH_PATH = os.path.join(HERE, "libradix_synth.so")

RDX_OK = 0
RDX_E_INVALID = -1
RDX_E_CAPACITY = -2
RDX_E_HIP = -3
RDX_E_NOMEM = -4
RDX_E_UNSUPPORTED = -5

LAYOUT_F32_SOA = 0
LAYOUT_EXT = 1
LAYOUT_PACKED = 2
LAYOUT_COV = 3
SH_NONE = 0
SH_F32 = 1
SH_EXT = 2
SH_PACKED = 3

_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)


class SortTuning(C.Structure):
    _fields_ = [("digit_bits", C.c_int), ("keys_per_thread", C.c_int), ("block_threads", C.c_int)]


class SplatTable(C.Structure):
    _fields_ = [
        ("layout", C.c_int),
        ("count", C.c_uint32),
        ("center", C.c_void_p),
        ("log_scale", C.c_void_p),
        ("quat", C.c_void_p),
        ("opacity", C.c_void_p),
        ("rgb", C.c_void_p),
        ("plane_stride", C.c_size_t),
        ("ext_a", C.c_void_p),
        ("ext_b", C.c_void_p),
        ("packed", C.c_void_p),
        ("packed_encoding", C.c_float * 4),
        ("sh_codec", C.c_int),
        ("sh_degree", C.c_int),
        ("sh", C.c_void_p),
        ("sh_channels", C.c_void_p * 4),
        ("object_to_world", C.c_double * 16),
    ]


class Eye(C.Structure):
    _fields_ = [("view", C.c_double * 16), ("projection", C.c_double * 16), ("ortho", C.c_int)]


class Camera(C.Structure):
    _fields_ = [
        ("head", C.c_double * 16),
        ("eyes", C.c_int),
        ("eye", Eye * 2),
        ("size", C.c_double * 2),
        ("ndc_z_min", C.c_double),
    ]


class Params(C.Structure):
    _fields_ = [
        ("max_std_dev", C.c_double),
        ("min_alpha", C.c_double),
        ("falloff", C.c_double),
        ("lod_inflate", C.c_int),
        ("blur_amount", C.c_double),
        ("pre_blur_amount", C.c_double),
        ("clip_xy", C.c_double),
        ("min_pixel_radius", C.c_double),
        ("max_pixel_radius", C.c_double),
        ("min_contribution", C.c_double),
        ("focal_distance", C.c_double),
        ("aperture_tan", C.c_double),
        ("sort_guard", C.c_double),
        ("max_sh_degree", C.c_int),
        ("sort_bits", C.c_int),
        ("cull_keys", C.c_int),
        ("key_per_eye", C.c_int),
    ]


class FloatOut(C.Structure):
    _fields_ = [
        ("stride", C.c_size_t),
        ("conic", C.c_void_p),
        ("cov2d", C.c_void_p),
        ("opacity", C.c_void_p),
        ("ndc", C.c_void_p),
        ("view_z", C.c_void_p),
        ("bound", C.c_void_p),
        ("rgb", C.c_void_p),
        ("distance", C.c_void_p),
    ]


class BatchView(C.Structure):
    _fields_ = [
        ("words_dev", C.c_void_p),
        ("table", C.c_uint32),
        ("reserved", C.c_uint32),
        ("camera", Camera),
    ]


class RasterSettings(C.Structure):
    _fields_ = [("sh_degree", C.c_int), ("t_stop", C.c_float), ("max_entries_per_splat", C.c_uint32)]


class DrawSettings(C.Structure):
    _fields_ = [("t_stop", C.c_float), ("linear_output", C.c_int), ("max_entries_per_splat", C.c_uint32)]


BATCH_GRAPH = 1
BATCH_PERSISTENT = 2
TUNE_CACHED = 1
STATIC_SORT_MAX_KEYS = 1048576
BATCH_MAX_VIEWS = 4096


class RadixError(RuntimeError):
    """An error of the native library, with the reference's message text."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


# (name, restype, argtypes)
_SIGNATURES = [
    ("rdx_last_error", C.c_char_p, []),
    ("rdx_version", C.c_char_p, []),
    ("rdx_key_invalid", C.c_uint32, [C.c_int]),
    ("rdx_key_valid_max", C.c_uint32, [C.c_int]),
    ("rdx_key_compute", C.c_uint32, [C.c_float, C.c_int]),
    ("rdx_compute_keys", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p]),
    ("rdx_context_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("rdx_context_destroy", C.c_int, [C.c_void_p]),
    ("rdx_context_reserve", C.c_int, [C.c_void_p, C.c_uint32]),
    ("rdx_context_set_tuning", C.c_int, [C.c_void_p, C.POINTER(SortTuning)]),
    ("rdx_context_get_tuning", C.c_int, [C.c_void_p, C.POINTER(SortTuning)]),
    ("rdx_sort_tuning_candidate", C.c_int, [C.c_int, C.POINTER(SortTuning)]),
    ("rdx_context_autotune", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.POINTER(SortTuning),
      C.POINTER(C.c_double)]),
    ("rdx_device_alloc", C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    ("rdx_device_free", C.c_int, [C.c_void_p]),
    ("rdx_device_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("rdx_device_download", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("rdx_device_fill_u32", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_size_t]),
    ("rdx_stream_synchronize", C.c_int, [C.c_void_p]),
    ("rdx_stream_create", C.c_int, [C.POINTER(C.c_void_p)]),
    ("rdx_stream_destroy", C.c_int, [C.c_void_p]),
    ("rdx_context_profile", C.c_int, [C.c_void_p, C.c_int]),
    ("rdx_context_stage_times", C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int]),
    ("rdx_sort_keys", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_size_t, _u32p]),
    ("rdx_sort_pairs", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]),
    ("rdx_sort_pairs_indirect", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int]),
    ("rdx_sort_splats", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rdx_sort_splats_static", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rdx_params_default", None, [C.POINTER(Params)]),
    ("rdx_project_ewa", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(SplatTable), C.POINTER(Camera), C.POINTER(Params), C.c_uint32,
      C.POINTER(C.c_void_p), C.c_void_p, C.c_size_t, C.POINTER(FloatOut)]),
    ("rdx_project_sort", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(SplatTable), C.POINTER(Camera), C.POINTER(Params),
      C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(FloatOut)]),
    ("rdx_batch_create", C.c_int,
     [C.c_void_p, C.POINTER(SplatTable), C.c_int, C.POINTER(BatchView), C.c_int, C.POINTER(Params), C.c_void_p,
      C.POINTER(C.c_void_p)]),
    ("rdx_batch_destroy", C.c_int, [C.c_void_p]),
    ("rdx_batch_autotune", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(SplatTable), C.c_int, C.POINTER(BatchView), C.c_int, C.POINTER(Params), C.c_int,
      C.c_int, C.POINTER(SortTuning), C.POINTER(C.c_double), C.POINTER(C.c_int)]),
    ("rdx_tuning_cache_file", C.c_int, [C.c_char_p]),
    ("rdx_tuning_cache_lookup", C.c_int,
     [C.c_char_p, C.c_uint64, C.c_uint32, C.c_int, C.POINTER(SortTuning), C.POINTER(C.c_double)]),
    ("rdx_batch_set_cameras", C.c_int, [C.c_void_p, C.POINTER(Camera), C.c_int]),
    ("rdx_batch_total", C.c_uint64, [C.c_void_p]),
    ("rdx_batch_segment_offset", C.c_uint64, [C.c_void_p, C.c_int]),
    ("rdx_batch_keys", C.c_void_p, [C.c_void_p]),
    ("rdx_batch_run", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    ("rdx_batch_persistent_grid", C.c_int, [C.c_void_p]),
    ("rdx_draw_ray", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_void_p), C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
      C.POINTER(Camera), C.c_int, C.POINTER(Params), C.c_void_p]),
    ("rdx_batch_status", C.c_int, [C.c_void_p]),
    ("rdx_batch_persistent_phases", C.c_int,
     [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]),
    ("rdx_batch_choose_form", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("rdx_raster_create", C.c_int, [C.c_void_p, C.POINTER(RasterSettings), C.POINTER(C.c_void_p)]),
    ("rdx_raster_destroy", C.c_int, [C.c_void_p]),
    ("rdx_raster_forward", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(Camera), C.POINTER(Params), C.c_void_p, C.c_void_p]),
    ("rdx_raster_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rdx_raster_stats", C.c_int,
     [C.c_void_p, C.c_void_p, _u32p, _u32p, _u32p, C.POINTER(C.c_int)]),
    ("rdx_project_keys", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(SplatTable), C.POINTER(Camera), C.POINTER(Params), C.c_uint32, C.c_void_p]),
    ("rdx_project_ray", C.c_int,
     [C.c_void_p, C.c_void_p, C.POINTER(SplatTable), C.POINTER(Camera), C.POINTER(Params), C.c_int, C.c_uint32,
      C.c_void_p, C.c_void_p]),
    ("rdx_untranspose_bytes", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    ("rdx_spz_decode", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t]),
    ("rdx_draw_create", C.c_int, [C.c_void_p, C.POINTER(DrawSettings), C.POINTER(C.c_void_p)]),
    ("rdx_draw_destroy", C.c_int, [C.c_void_p]),
    ("rdx_draw_splats", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
      C.POINTER(Params), C.c_void_p]),
    ("rdx_draw_stats", C.c_int,
     [C.c_void_p, C.c_void_p, _u32p, _u32p, _u32p, C.POINTER(C.c_int)]),
]

_lib = None


def lib() -> C.CDLL:
    """The loaded libradix_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -m radix_amd.build` (the HIP path has no fallback)")
        handle = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, res, args in _SIGNATURES:
            f = getattr(handle, name)
            f.restype = res
            f.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != RDX_OK:
        msg = lib().rdx_last_error().decode()
        raise RadixError(rc, msg)


_synth = None


class CloudDesc(C.Structure):
    _fields_ = [
        ("seed", C.c_uint32),
        ("count", C.c_uint32),
        ("sh_degree", C.c_int),
        ("centers", C.c_int),
        ("offset", C.c_double * 3),
        ("log_scale_mean", C.c_double),
        ("log_scale_std", C.c_double),
        ("sh_std", C.c_double),
        ("clusters", C.c_int),
        ("extent", C.c_double),
        ("std_lo", C.c_double),
        ("std_hi", C.c_double),
    ]


def synth() -> C.CDLL:
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} is not built: run `python -m radix_amd.build`")
        h = C.CDLL(SYNTH_PATH)
        h.rdxs_draws_per_splat.restype = C.c_int
        h.rdxs_draws_per_splat.argtypes = [C.POINTER(CloudDesc)]
        h.rdxs_generate.restype = C.c_int
        h.rdxs_generate.argtypes = [C.POINTER(CloudDesc)] + [C.c_void_p] * 6 + [C.c_size_t]
        h.rdxs_pack_splats.restype = C.c_int
        h.rdxs_pack_splats.argtypes = [C.c_uint32] + [C.c_void_p] * 5 + [C.c_size_t, C.c_void_p, C.c_void_p]
        h.rdxs_pack_sh.restype = C.c_int
        h.rdxs_pack_sh.argtypes = [C.c_uint32, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p]
        h.rdxs_uniforms.restype = C.c_int
        h.rdxs_uniforms.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
        h.rdxs_normals.restype = C.c_int
        h.rdxs_normals.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p]
        _synth = h
    return _synth

"""The five synthetic workloads (seeded), built on libradix_synth.so.

plain_small   65,536 splats, centres uniform in [-1,1]^3 + (0,0,-3), log-scales
              N(-4, 0.5), SH0; camera at the origin looking down -z, 60 deg,
              1280x720, near 0.01, far 1000; 32-bit keys
plain_4m      4,194,304 splats from 64 anisotropic clusters, log-scales
              N(-5, 0.7), SH3 (bands N(0, 0.05)); 1920x1080 camera on a
              random orbit of radius 8, 60 deg
batched       8 collections of log-uniform counts in [1e4, 6e6], 32 views at
              1280x720 on orbits of radius 4-12
diff          262,144 splats (SH3, plain_4m generators, centres in
              [-1,1]^3 + (0,0,-3)) at 512x512
stereo_16m    16,777,216 splats quantized to packedSplats + packedSh SH3
              (lnScale range [-10, 2]); two 2064x2208 eyes 64 mm apart
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _native as N
from . import camera as cam

SH_COEFFS = [0, 3, 8, 15]


def planes_for(sh_degree: int) -> int:
    return 14 + 3 * SH_COEFFS[sh_degree]


@dataclass
class CloudSpec:
    seed: int
    count: int
    sh_degree: int
    centers: int  # 0 cube, 1 clusters
    offset: tuple = (0.0, 0.0, 0.0)
    log_scale_mean: float = -5.0
    log_scale_std: float = 0.7
    sh_std: float = 0.05
    clusters: int = 64
    extent: float = 5.0
    std_lo: float = 0.05
    std_hi: float = 1.0

    def desc(self) -> N.CloudDesc:
        d = N.CloudDesc()
        d.seed = self.seed & 0xFFFFFFFF
        d.count = self.count
        d.sh_degree = self.sh_degree
        d.centers = self.centers
        d.offset[:] = list(self.offset)
        d.log_scale_mean = self.log_scale_mean
        d.log_scale_std = self.log_scale_std
        d.sh_std = self.sh_std
        d.clusters = self.clusters
        d.extent = self.extent
        d.std_lo = self.std_lo
        d.std_hi = self.std_hi
        return d


def generate(spec: CloudSpec, out: Optional[np.ndarray] = None) -> np.ndarray:
    """float32 planes (14 + 3K, n): center 3, log-scale 3, quat 4, opacity, rgb 3, SH."""
    n = spec.count
    P = planes_for(spec.sh_degree)
    if out is None:
        out = np.empty((P, n), dtype=np.float32)
    base = out.ctypes.data
    st = n * 4
    lib = N.synth()
    d = spec.desc()
    rc = lib.rdxs_generate(C.byref(d), base, base + 3 * st, base + 6 * st, base + 10 * st, base + 11 * st,
                           base + 14 * st if spec.sh_degree > 0 else None, n)
    if rc != 0:
        raise RuntimeError("rdxs_generate failed")
    return out


def pack(planes: np.ndarray, sh_degree: int, encoding=(0.0, 1.0, -10.0, 2.0)):
    """packedSplats words (n, 4) uint32 and packedSh channels (sh_degree of (n, 4))."""
    n = planes.shape[1]
    base = planes.ctypes.data
    st = n * 4
    words = np.empty((n, 4), dtype=np.uint32)
    enc = (C.c_double * 4)(*encoding)
    lib = N.synth()
    lib.rdxs_pack_splats(n, base, base + 3 * st, base + 6 * st, base + 10 * st, base + 11 * st, n, enc,
                         words.ctypes.data)
    channels = []
    if sh_degree > 0:
        channels = [np.empty((n, 4), dtype=np.uint32) for _ in range(sh_degree)]
        ptrs = (C.c_void_p * 4)(*[c.ctypes.data for c in channels] + [None] * (4 - sh_degree))
        lib.rdxs_pack_sh(n, sh_degree, base + 14 * st, n, ptrs)
    return words, channels


def plain_small_spec(seed: int = 1, count: int = 65536) -> CloudSpec:
    return CloudSpec(seed=seed, count=count, sh_degree=0, centers=0, offset=(0.0, 0.0, -3.0), log_scale_mean=-4.0,
                     log_scale_std=0.5)


def plain_small_camera() -> cam.CameraRig:
    return cam.mono(np.eye(4), 60.0, 1280, 720, near=0.01, far=1000.0)


def plain_4m_spec(seed: int = 2, count: int = 4194304) -> CloudSpec:
    return CloudSpec(seed=seed, count=count, sh_degree=3, centers=1)


def orbit_camera(seed: int, radius: float, width: int, height: int, fov: float = 60.0) -> cam.CameraRig:
    d = cam.random_direction(seed)
    return cam.mono(cam.orbit_world(d, radius), fov, width, height, near=0.01, far=1000.0)


def plain_4m_camera(seed: int = 2) -> cam.CameraRig:
    return orbit_camera(seed * 7919 + 11, 8.0, 1920, 1080)


def diff_spec(seed: int = 4, count: int = 262144) -> CloudSpec:
    return CloudSpec(seed=seed, count=count, sh_degree=3, centers=0, offset=(0.0, 0.0, -3.0))


def diff_camera(size: int = 512) -> cam.CameraRig:
    return cam.mono(np.eye(4), 60.0, size, size, near=0.01, far=1000.0)


@dataclass
class BatchedWorkload:
    specs: List[CloudSpec]
    cameras: List[cam.CameraRig]
    view_table: List[int]


def batched(seed: int = 3, collections: int = 8, views: int = 32, lo: float = 1e4, hi: float = 6e6,
            width: int = 1280, height: int = 720, scale: float = 1.0) -> BatchedWorkload:
    """Counts log-uniform in [lo, hi] (times `scale`, for smaller test runs)."""
    u = np.zeros(collections + 3 * views, dtype=np.float64)
    N.synth().rdxs_uniforms(seed, u.size, u.ctypes.data_as(C.c_void_p))
    counts = [max(1, int(round(math.exp(math.log(lo) + u[c] * (math.log(hi) - math.log(lo))) * scale)))
              for c in range(collections)]
    specs = [CloudSpec(seed=seed * 1000 + c + 1, count=counts[c], sh_degree=3, centers=1) for c in range(collections)]
    cameras, table = [], []
    for v in range(views):
        a, b, c = u[collections + 3 * v: collections + 3 * v + 3]
        table.append(min(int(a * collections), collections - 1))
        radius = 4.0 + 8.0 * b
        cameras.append(orbit_camera(seed * 100000 + v + 1, radius, width, height))
    return BatchedWorkload(specs, cameras, table)


def stereo_16m_spec(seed: int = 5, count: int = 16777216) -> CloudSpec:
    return CloudSpec(seed=seed, count=count, sh_degree=3, centers=1)


def stereo_camera(seed: int = 5, width: int = 2064, height: int = 2208, ipd: float = 0.064,
                  fov: float = 90.0) -> cam.CameraRig:
    d = cam.random_direction(seed * 7919 + 11)
    return cam.stereo(cam.orbit_world(d, 8.0), ipd, fov, width, height)

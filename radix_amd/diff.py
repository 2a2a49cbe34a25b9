"""Host side of the differentiable compositor: parameter tensors in the
f32 table's plane order (centre 3, log-scale 3, quaternion 4, opacity
LOGIT 1, base colour 3, SH bands 1-3 45 = 59 per splat) and the training
workload (262,144 splats at 512x512: the same splats as target with means
+ N(0, 0.01) and base colours + N(0, 0.05)). No reference counterpart: the
reference draws sorted splats with rasterised quads and has no backward
(src/render/splatDraw.ts); the compositor follows its fragment alpha and
premultiplied "over" blending.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from . import workloads as W

PARAMS = 59


def theta_from_planes(planes: np.ndarray) -> np.ndarray:
    """(n, 59) float64 parameters from the generator's float32 planes (the
    opacity plane holds sigmoid(N(0, 1)); its logit is the parameter)."""
    p = np.asarray(planes, dtype=np.float64)
    theta = np.zeros((p.shape[1], PARAMS))
    rows = min(p.shape[0], PARAMS)
    theta[:, :rows] = p[:rows].T
    op = np.clip(p[10], 1e-7, 1.0 - 1e-7)
    theta[:, 10] = np.log(op / (1.0 - op))
    return theta


def perturbed(theta: np.ndarray, seed: int, mean_std: float = 0.01, dc_std: float = 0.05) -> np.ndarray:
    """The target scene: means + N(0, mean_std), base colours + N(0, dc_std)
    (seeded Box-Muller normals of the generator)."""
    n = theta.shape[0]
    z = np.zeros(6 * n)
    N.synth().rdxs_normals(seed, z.size, z.ctypes.data_as(C.c_void_p))
    out = theta.copy()
    out[:, 0:3] += mean_std * z[:3 * n].reshape(n, 3)
    out[:, 11:14] += dc_std * z[3 * n:].reshape(n, 3)
    return out


def workload(count: int = 262144, seed: int = 4):
    """theta (n, 59), target theta and the camera of the training workload."""
    spec = W.diff_spec(seed=seed, count=count)
    theta = theta_from_planes(W.generate(spec))
    return theta, perturbed(theta, seed * 31 + 7), W.diff_camera()

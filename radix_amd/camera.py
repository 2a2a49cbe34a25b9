"""Cameras as the reference builds them with three.js: column-major 4x4
matrices (Matrix4.elements), perspective projections in the WebGPU or WebGL
clip convention, look-at poses, orbits and stereo rigs, and the rdx_camera
descriptor (SplatCamera.setViews, src/render/splatCamera.ts:119-138).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _native as N


def cm(m: np.ndarray) -> np.ndarray:
    """Row-major 4x4 to column-major elements."""
    return np.ascontiguousarray(np.asarray(m, dtype=np.float64).T.reshape(16))


def rm(e: np.ndarray) -> np.ndarray:
    """Column-major elements to a row-major 4x4."""
    return np.asarray(e, dtype=np.float64).reshape(4, 4).T.copy()


def perspective(fov_deg: float, aspect: float, near: float, far: float, webgpu: bool = True) -> np.ndarray:
    """PerspectiveCamera.updateProjectionMatrix + Matrix4.makePerspective (row-major)."""
    top = near * math.tan(math.radians(0.5 * fov_deg))
    height = 2 * top
    width = aspect * height
    left = -0.5 * width
    right = left + width
    bottom = top - height
    x = 2 * near / (right - left)
    y = 2 * near / (top - bottom)
    a = (right + left) / (right - left)
    b = (top + bottom) / (top - bottom)
    if webgpu:
        c = -far / (far - near)
        d = -far * near / (far - near)
    else:
        c = -(far + near) / (far - near)
        d = -2 * far * near / (far - near)
    return np.array([[x, 0, a, 0], [0, y, b, 0], [0, 0, c, d], [0, 0, -1, 0]], dtype=np.float64)


def orthographic(left: float, right: float, top: float, bottom: float, near: float, far: float,
                 webgpu: bool = True) -> np.ndarray:
    """OrthographicCamera.updateProjectionMatrix + Matrix4.makeOrthographic (row-major)."""
    w, h, p = 1.0 / (right - left), 1.0 / (top - bottom), 1.0 / (far - near)
    x, y = (right + left) * w, (top + bottom) * h
    if webgpu:
        z, z_inv = near * p, -p
    else:
        z, z_inv = (far + near) * p, -2.0 * p
    return np.array([[2 * w, 0, 0, -x], [0, 2 * h, 0, -y], [0, 0, z_inv, -z], [0, 0, 0, 1]], dtype=np.float64)


def look_at_world(eye, target, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """A camera's world matrix looking from `eye` at `target` (Matrix4.lookAt)."""
    eye = np.asarray(eye, dtype=np.float64)
    z = eye - np.asarray(target, dtype=np.float64)
    if np.dot(z, z) == 0:
        z = np.array([0.0, 0.0, 1.0])
    z = z / np.linalg.norm(z)
    up = np.asarray(up, dtype=np.float64)
    x = np.cross(up, z)
    if np.dot(x, x) == 0:
        z = z + np.array([1e-4, 0.0, 0.0]) if abs(up[2]) == 1 else z + np.array([0.0, 0.0, 1e-4])
        z = z / np.linalg.norm(z)
        x = np.cross(up, z)
    x = x / np.linalg.norm(x)
    y = np.cross(z, x)
    w = np.eye(4)
    w[:3, 0], w[:3, 1], w[:3, 2], w[:3, 3] = x, y, z, eye
    return w


def translation(t) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = t
    return m


@dataclass
class View:
    """One eye: world-to-eye view and projection (row-major)."""

    view: np.ndarray
    projection: np.ndarray
    ortho: bool = False


@dataclass
class CameraRig:
    """A render's camera: the head (world to head) and 1 or 2 eyes."""

    head: np.ndarray
    eyes: List[View]
    width: int
    height: int
    ndc_z_min: float = 0.0
    _desc: Optional[N.Camera] = field(default=None, repr=False)

    def desc(self) -> N.Camera:
        c = N.Camera()
        c.head[:] = list(cm(self.head))
        c.eyes = len(self.eyes)
        for e, v in enumerate(self.eyes):
            c.eye[e].view[:] = list(cm(v.view))
            c.eye[e].projection[:] = list(cm(v.projection))
            c.eye[e].ortho = 1 if v.ortho else 0
        c.size[0] = float(self.width)
        c.size[1] = float(self.height)
        c.ndc_z_min = float(self.ndc_z_min)
        return c


def mono(world: np.ndarray, fov_deg: float, width: int, height: int, near: float = 0.01, far: float = 1000.0,
         webgpu: bool = True) -> CameraRig:
    """A plain camera: one eye that is also the head (exactly)."""
    view = np.linalg.inv(world)
    proj = perspective(fov_deg, width / height, near, far, webgpu)
    return CameraRig(head=view, eyes=[View(view.copy(), proj)], width=width, height=height,
                     ndc_z_min=0.0 if webgpu else -1.0)


def mono_ortho(world: np.ndarray, half_height: float, width: int, height: int, near: float = 0.01,
               far: float = 1000.0) -> CameraRig:
    """A plain orthographic camera (frustum [-h a, h a] x [-h, h], a = width / height)."""
    view = np.linalg.inv(world)
    a = width / height
    proj = orthographic(-half_height * a, half_height * a, half_height, -half_height, near, far, True)
    return CameraRig(head=view, eyes=[View(view.copy(), proj, ortho=True)], width=width, height=height, ndc_z_min=0.0)


def stereo(head_world: np.ndarray, ipd: float, fov_deg: float, width: int, height: int, near: float = 0.01,
           far: float = 1000.0) -> CameraRig:
    """Two eyes at -ipd/2 and +ipd/2 along the head's x axis."""
    proj = perspective(fov_deg, width / height, near, far, True)
    eyes = [View(np.linalg.inv(head_world @ translation((s * ipd / 2, 0.0, 0.0))), proj) for s in (-1.0, 1.0)]
    return CameraRig(head=np.linalg.inv(head_world), eyes=eyes, width=width, height=height, ndc_z_min=0.0)


def orbit_world(direction, radius: float, target=(0.0, 0.0, 0.0)) -> np.ndarray:
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    return look_at_world(np.asarray(target) + radius * d, target)


def random_direction(seed: int) -> np.ndarray:
    """A uniform direction from a seeded stream (normalised 3D normal)."""
    import ctypes as C

    out = np.zeros(3)
    N.synth().rdxs_normals(seed, 3, out.ctypes.data_as(C.c_void_p))
    return out / np.linalg.norm(out)

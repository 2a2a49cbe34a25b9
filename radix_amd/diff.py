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
from .sort import on_stream

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


class Rasterizer:
    """rdx_raster_*: the tile compositor and its adjoint on the device."""

    def __init__(self, sh_degree: int = 3, t_stop: float = 1e-4, entries_per_splat: int = 0, context=None):
        """entries_per_splat: a limit on (splat, tile) entries per splat (0:
        none; the entry buffers follow each forward's count); a forward past
        it raises RadixError instead of drawing."""
        from .sort import default_context

        self.context = context or default_context()
        # One forward state lives on the device: each forward bumps the
        # generation, and a backward of an older one is refused.
        self.generation = 0
        s = N.RasterSettings(sh_degree, t_stop, entries_per_splat)
        h = C.c_void_p()
        N.check(N.lib().rdx_raster_create(self.context.handle, C.byref(s), C.byref(h)))
        self._h = h

    @on_stream
    def forward(self, params, rig, prm, stream=None):
        """params: (59, n) float32 device tensor. Returns (image (3, H, W), alpha (H, W))."""
        import torch

        from .sort import _stream

        if params.dtype != torch.float32 or not params.is_cuda or params.dim() != 2 or params.shape[0] != PARAMS:
            raise TypeError("params must be a (59, n) float32 device tensor")
        params = params.contiguous()
        n = params.shape[1]
        W, H = rig.width, rig.height
        image = torch.empty((3, H, W), dtype=torch.float32, device=params.device)
        alpha = torch.empty((H, W), dtype=torch.float32, device=params.device)
        self._params = params  # the backward reads the forward's parameters
        self.generation += 1
        N.check(N.lib().rdx_raster_forward(self._h, _stream(stream), C.c_void_p(params.data_ptr()), n,
                                           C.byref(rig.desc()), C.byref(prm.desc()), C.c_void_p(image.data_ptr()),
                                           C.c_void_p(alpha.data_ptr())))
        return image, alpha

    @on_stream
    def backward(self, grad_image, generation=None, stream=None):
        """dL/dparams (59, n) of the last forward (of `generation`, if given:
        a later forward on this Rasterizer raises instead of differentiating
        the wrong image)."""
        import torch

        from .sort import _stream

        check_generation(self, generation)
        g = grad_image.contiguous().float()
        out = torch.empty_like(self._params)
        N.check(N.lib().rdx_raster_backward(self._h, _stream(stream), C.c_void_p(g.data_ptr()),
                                            C.c_void_p(out.data_ptr())))
        return out

    def stats(self, stream=None):
        from .sort import _stream

        v, e, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        o = C.c_int()
        N.check(N.lib().rdx_raster_stats(self._h, _stream(stream), C.byref(v), C.byref(e), C.byref(c), C.byref(o)))
        return {"visible": v.value, "entries": e.value, "capacity": c.value, "overflowed": bool(o.value)}

    def close(self):
        if getattr(self, "_h", None):
            N.lib().rdx_raster_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_generation(rasterizer, generation) -> None:
    if generation is not None and generation != rasterizer.generation:
        raise RuntimeError(
            f"Rasterizer: the backward of forward {generation} was asked after forward {rasterizer.generation} "
            "replaced its state; composite each image that is differentiated together with its own Rasterizer")


def composite(params, rig, prm, rasterizer: "Rasterizer"):
    """The image of `params` (59, n) as a differentiable torch function: a
    torch.autograd.Function whose backward is the hand-written adjoint.

    The rasterizer keeps one forward's state: composite two images with one
    Rasterizer before backward and the first backward raises (its forward's
    generation is no longer current). The parameters are saved for backward,
    so an in-place change to them before backward raises too (autograd's
    version counter)."""
    import torch

    class _Composite(torch.autograd.Function):
        @staticmethod
        def forward(ctx, p):
            image, _ = rasterizer.forward(p, rig, prm)
            ctx.generation = rasterizer.generation
            ctx.save_for_backward(p)
            return image

        @staticmethod
        def backward(ctx, grad_image):
            ctx.saved_tensors  # raises if the parameters were changed in place
            return rasterizer.backward(grad_image, generation=ctx.generation)

    return _Composite.apply(params)


def params_tensor(theta: np.ndarray, device="cuda"):
    """(n, 59) float64 -> (59, n) float32 device planes."""
    import torch

    return torch.from_numpy(np.ascontiguousarray(theta.T, dtype=np.float32)).to(device)

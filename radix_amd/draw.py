"""Host side of the draw in production form: the sorted splats' ewa words
composited into a premultiplied RGBA image, as the reference's splat
material draws them (src/render/splatDraw.ts:98-181): for each splat of the
order, back to front, the fragments of its quad (z^2 = |uv|^2, discarded
beyond the cutoff and below minAlpha, alpha = fragmentAlpha) blended
premultiplied "over". It consumes what Projector.project_sort returns.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

from . import _native as N
from .sort import on_stream


# RDX_METHOD_3DGUT, RDX_METHOD_HTGS, RDX_METHOD_RAYGS
RAY_METHODS = {"3dgut": 1, "htgs": 2, "raygs": 3}


class Drawer:
    """rdx_draw_*: composite sorted splats into an RGBA image on the device."""

    def __init__(self, t_stop: float = 0.0, linear_output: bool = False, entries_per_splat: int = 0, context=None):
        """t_stop 0 composites every fragment (the reference's blend); t_stop
        > 0 stops a pixel once its transmittance would fall below it.
        linear_output sRGB-decodes the colours first (an output that stores
        linear values, splatDraw.ts:176). entries_per_splat: a limit on
        (splat, tile) entries per draw position (0: none; the entry buffers
        follow each draw's count); a draw past it raises RadixError."""
        from .sort import default_context

        self.context = context or default_context()
        s = N.DrawSettings(float(t_stop), 1 if linear_output else 0, int(entries_per_splat))
        h = C.c_void_p()
        N.check(N.lib().rdx_draw_create(self.context.handle, C.byref(s), C.byref(h)))
        self._h = h

    @on_stream
    def draw(self, words, order, count, width: int, height: int, params, out=None, stream=None):
        """words: (n, 4) int32 device tensor of one eye's ewa words; order:
        int32 device tensor of splat indices, back to front; count: int32
        device tensor whose first element is the number drawn (clamped to
        order's length). Returns (4, height, width) float32: premultiplied
        r, g, b and alpha, row 0 at the top."""
        import torch

        from .sort import _stream

        for name, t in (("words", words), ("order", order), ("count", count)):
            if t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous int32 device tensor")
        n = words.numel() // 4
        if out is None:
            out = torch.empty((4, height, width), dtype=torch.float32, device=words.device)
        elif out.dtype != torch.float32 or out.numel() < 4 * width * height or not out.is_contiguous():
            raise TypeError("out must be a contiguous float32 tensor of 4 x height x width")
        N.check(N.lib().rdx_draw_splats(self._h, _stream(stream), C.c_void_p(words.data_ptr()), n,
                                        C.c_void_p(order.data_ptr()), C.c_void_p(count.data_ptr()), order.numel(),
                                        width, height, C.byref(params.desc()), C.c_void_p(out.data_ptr())))
        return out

    @on_stream
    def draw_ray(self, method: str, channels, order, count, rig, params, eye: int = 0, out=None, stream=None):
        """The draw of a ray method's words ("3dgut", "htgs" or "raygs"):
        channels as Projector.project_ray returns them (w0, w1 of (n, 4) and
        with two eyes w2 of (n, 2) int32 device tensors), projected for
        `rig`, whose eye `eye` is drawn at rig's render size; order and count
        as in draw(). Returns (4, height, width) float32, premultiplied, row
        0 at the top."""
        import torch

        from .sort import _stream

        if method not in RAY_METHODS:
            raise ValueError(f"draw_ray: method {method!r} is not one of {sorted(RAY_METHODS)}")
        m = RAY_METHODS[method]
        eyes = len(rig.eyes)
        need = 3 if eyes == 2 else 2
        if len(channels) < need:
            raise ValueError(f"draw_ray: {need} word channels for {eyes} eye(s)")
        for name, t in [("order", order), ("count", count)] + [(f"w{k}", c) for k, c in enumerate(channels[:need])]:
            if t.dtype != torch.int32 or not t.is_cuda or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous int32 device tensor")
        n = channels[0].numel() // 4
        if channels[1].numel() < 4 * n or (eyes == 2 and channels[2].numel() < 2 * n):
            raise ValueError("draw_ray: word channels shorter than w0")
        width, height = int(rig.width), int(rig.height)
        if out is None:
            out = torch.empty((4, height, width), dtype=torch.float32, device=channels[0].device)
        elif out.dtype != torch.float32 or out.numel() < 4 * width * height or not out.is_contiguous():
            raise TypeError("out must be a contiguous float32 tensor of 4 x height x width")
        cp = (C.c_void_p * 3)(*[c.data_ptr() for c in channels[:need]] + [None] * (3 - need))
        N.check(N.lib().rdx_draw_ray(self._h, _stream(stream), m, cp, n, C.c_void_p(order.data_ptr()),
                                     C.c_void_p(count.data_ptr()), order.numel(), C.byref(rig.desc()), eye,
                                     C.byref(params.desc()), C.c_void_p(out.data_ptr())))
        return out

    def stats(self, stream=None) -> dict:
        from .sort import _stream

        drawn, entries, cap = C.c_uint32(), C.c_uint32(), C.c_uint32()
        ov = C.c_int()
        N.check(N.lib().rdx_draw_stats(self._h, _stream(stream), C.byref(drawn), C.byref(entries), C.byref(cap),
                                       C.byref(ov)))
        return {"drawn": drawn.value, "entries": entries.value, "capacity": cap.value, "overflowed": bool(ov.value)}

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().rdx_draw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def draw_projected(out: dict, width: int, height: int, params, eye: int = 0, drawer: Optional[Drawer] = None):
    """Draw what Projector.project_sort returned (its words of `eye`, its
    indices and count): the reference's per-frame project, sort and draw."""
    drawer = drawer or Drawer()
    words, indices, count = out["words"][eye], out["indices"], out["count"]
    if count.numel() > 1:  # per-eye sorts: eye e's order at [e n, (e + 1) n)
        n = words.shape[0]
        indices, count = indices[eye * n:(eye + 1) * n], count[eye:eye + 1]
    return drawer.draw(words, indices, count, width, height, params)

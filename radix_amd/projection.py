"""The ewa projection and the fused project + sort, mirroring the reference's
SplatProjection.write (src/render/SplatProjection.ts:141-153) and the
per-frame GPU path of SplatRenderer.update (src/render/SplatRenderer.ts:412-456).

Tables live on the device as torch tensors: f32 planes (a structure-of-arrays
layout), or the reference's extSplats / packedSplats words with their SH
bundles. Outputs keep the reference layouts: per eye a uint32x4 word array
(channel w<e> of the projected bundle), one uint32 key per splat, and the
ordered uint32 indices of the visible splats.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields
from typing import Dict, List, Optional

import numpy as np

from . import _native as N
from .camera import CameraRig
from .sort import Context, _ptr, _stream, default_context, on_stream

SH_COEFFS = [0, 3, 8, 15]


@dataclass
class Params:
    """SplatViewParams + SplatObjectParams (src/render/splatParams.ts:68-72) and the key uniforms."""

    max_std_dev: float = float(np.sqrt(8.0))
    min_alpha: float = 0.5 / 255.0
    falloff: float = 1.0
    lod_inflate: bool = True
    blur_amount: float = 0.3
    pre_blur_amount: float = 0.0
    clip_xy: float = 1.4
    min_pixel_radius: float = 0.0
    max_pixel_radius: float = 512.0
    min_contribution: float = 0.0
    focal_distance: float = 0.0
    aperture_tan: float = 0.0
    sort_guard: float = 0.0
    max_sh_degree: int = 4
    sort_bits: int = 24
    cull_keys: bool = True
    key_per_eye: bool = False

    def desc(self) -> N.Params:
        p = N.Params()
        for f in fields(self):
            v = getattr(self, f.name)
            setattr(p, f.name, int(v) if isinstance(v, bool) else v)
        return p


class DeviceTable:
    """A splat table on the device and its rdx_splat_table descriptor."""

    def __init__(self, layout: int, count: int, tensors: Dict[str, object], sh_codec: int = N.SH_NONE,
                 sh_degree: int = 0, packed_encoding=(0.0, 1.0, -12.0, 9.0), plane_stride: int = 0,
                 object_to_world: Optional[np.ndarray] = None):
        self.layout = layout
        self.count = count
        self.tensors = tensors  # keeps the device memory alive
        self.sh_codec = sh_codec
        self.sh_degree = sh_degree
        self.packed_encoding = tuple(packed_encoding)
        self.plane_stride = plane_stride or count
        self.object_to_world = np.eye(4) if object_to_world is None else np.asarray(object_to_world, dtype=np.float64)

    @classmethod
    def from_f32_planes(cls, planes, count: int, sh_degree: int, device="cuda", object_to_world=None):
        """`planes`: float32 (14 + 3K, n): center 3, log-scale 3, quat 4,
        opacity 1, rgb 3, SH 3K (plane 3k + c)."""
        import torch

        t = planes if hasattr(planes, "is_cuda") else torch.from_numpy(np.ascontiguousarray(planes))
        t = t.to(device=device, dtype=torch.float32).contiguous()
        return cls(N.LAYOUT_F32_SOA, count, {"planes": t}, sh_codec=N.SH_F32 if sh_degree > 0 else N.SH_NONE,
                   sh_degree=sh_degree, plane_stride=t.shape[1], object_to_world=object_to_world)

    @classmethod
    def from_packed(cls, words, count: int, encoding, sh_channels=None, sh_degree: int = 0, device="cuda",
                    object_to_world=None):
        import torch

        w = torch.from_numpy(np.ascontiguousarray(words).view(np.int32)).to(device)
        tensors = {"packed": w}
        codec = N.SH_NONE
        if sh_channels is not None and sh_degree > 0:
            tensors["sh"] = [torch.from_numpy(np.ascontiguousarray(c).view(np.int32)).to(device) for c in sh_channels]
            codec = N.SH_PACKED
        return cls(N.LAYOUT_PACKED, count, tensors, sh_codec=codec, sh_degree=sh_degree, packed_encoding=encoding,
                   object_to_world=object_to_world)

    @classmethod
    def from_ext(cls, ext_a, ext_b, count: int, sh_channels=None, sh_codec: int = N.SH_EXT, sh_degree: int = 0,
                 device="cuda", object_to_world=None):
        import torch

        tensors = {"ext_a": torch.from_numpy(np.ascontiguousarray(ext_a).view(np.int32)).to(device),
                   "ext_b": torch.from_numpy(np.ascontiguousarray(ext_b).view(np.int32)).to(device)}
        codec = N.SH_NONE
        if sh_channels is not None and sh_degree > 0:
            tensors["sh"] = [torch.from_numpy(np.ascontiguousarray(c).view(np.int32)).to(device) for c in sh_channels]
            codec = sh_codec
        return cls(N.LAYOUT_EXT, count, tensors, sh_codec=codec, sh_degree=sh_degree, object_to_world=object_to_world)

    @classmethod
    def from_cov(cls, cov_a, cov_b, count: int, sh_channels=None, sh_codec: int = N.SH_EXT, sh_degree: int = 0,
                 device="cuda", object_to_world=None):
        """covSplats words (covA, covB uint32x4 per splat, src/core/splats.ts:25-27) with ext or packed SH."""
        t = cls.from_ext(cov_a, cov_b, count, sh_channels, sh_codec=sh_codec, sh_degree=sh_degree, device=device,
                         object_to_world=object_to_world)
        t.layout = N.LAYOUT_COV
        return t

    def desc(self) -> N.SplatTable:
        d = N.SplatTable()
        d.layout = self.layout
        d.count = self.count
        d.plane_stride = self.plane_stride
        t = self.tensors
        if self.layout == N.LAYOUT_F32_SOA:
            base = t["planes"].data_ptr()
            st = self.plane_stride * 4
            d.center = base
            d.log_scale = base + 3 * st
            d.quat = base + 6 * st
            d.opacity = base + 10 * st
            d.rgb = base + 11 * st
            d.sh = base + 14 * st if self.sh_degree > 0 else None
        elif self.layout in (N.LAYOUT_EXT, N.LAYOUT_COV):  # covA, covB travel as ext_a, ext_b
            d.ext_a = t["ext_a"].data_ptr()
            d.ext_b = t["ext_b"].data_ptr()
        else:
            d.packed = t["packed"].data_ptr()
        d.packed_encoding[:] = list(self.packed_encoding)
        d.sh_codec = self.sh_codec
        d.sh_degree = self.sh_degree
        if "sh" in t:
            for k, c in enumerate(t["sh"]):
                d.sh_channels[k] = c.data_ptr()
        d.object_to_world[:] = list(np.asarray(self.object_to_world, dtype=np.float64).T.reshape(16))
        return d


SIDE_PLANES = {"conic": 3, "cov2d": 3, "opacity": 1, "ndc": 2, "view_z": 1, "bound": 3}


class Projector:
    """Runs rdx_project_ewa / rdx_project_sort for tables and cameras."""

    def __init__(self, context: Optional[Context] = None):
        self.context = context or default_context()

    def _side(self, n: int, eyes: int, key_per_eye: bool, device):
        import torch

        out = {k: torch.empty(eyes * v * max(n, 1), dtype=torch.float32, device=device) for k, v in SIDE_PLANES.items()}
        out["rgb"] = torch.empty(3 * max(n, 1), dtype=torch.float32, device=device)
        out["distance"] = torch.empty((eyes if key_per_eye else 1) * max(n, 1), dtype=torch.float32, device=device)
        f = N.FloatOut()
        f.stride = n
        for k, v in out.items():
            setattr(f, k, v.data_ptr())
        return out, f

    @on_stream
    def project(self, table: DeviceTable, rig: CameraRig, params: Params, side: bool = False, stream=None):
        """Words (per eye, (n, 4) int32), keys ((n,) or (eyes, n) int32) and optional float outputs."""
        import torch

        dev = table.tensors[next(iter(table.tensors))]
        dev = dev[0].device if isinstance(dev, list) else dev.device
        n = table.count
        eyes = len(rig.eyes)
        words = [torch.zeros((max(n, 1), 4), dtype=torch.int32, device=dev) for _ in range(eyes)]
        per_eye = params.key_per_eye and eyes == 2
        keys = torch.zeros((2 if per_eye else 1) * max(n, 1), dtype=torch.int32, device=dev)
        wp = (C.c_void_p * 2)(*[w.data_ptr() for w in words] + [None] * (2 - eyes))
        out = {"words": [w[:n] for w in words]}
        fo = None
        if side:
            sd, fo = self._side(n, eyes, per_eye, dev)
            out["side"] = sd
        N.check(N.lib().rdx_project_ewa(self.context.handle, _stream(stream), C.byref(table.desc()),
                                        C.byref(rig.desc()), C.byref(params.desc()), 0, wp, _ptr(keys),
                                        max(n, 1) if per_eye else 0, C.byref(fo) if fo is not None else None))
        out["keys"] = keys.view(2, -1)[:, :n] if per_eye else keys[:n]
        return out

    @on_stream
    def project_keys(self, table: DeviceTable, rig: CameraRig, params: Params, keys=None, at: int = 0, stream=None):
        """The keys-only pass (SplatProjection.writeKeys, src/render/SplatProjection.ts:165-172): the
        guarded head-distance key of each splat into keys[at:at + n] ((n,) int32 when keys is None)."""
        import torch

        n = table.count
        if keys is None:
            keys = torch.zeros(max(at + n, 1), dtype=torch.int32, device="cuda")
        if keys.numel() < at + n:
            raise ValueError(f"project_keys: keys hold {keys.numel()} of {at + n}")
        N.check(N.lib().rdx_project_keys(self.context.handle, _stream(stream), C.byref(table.desc()),
                                         C.byref(rig.desc()), C.byref(params.desc()), at, _ptr(keys)))
        return keys[at:at + n]

    METHODS = {"3dgut": 1, "htgs": 2, "raygs": 3}

    @on_stream
    def project_ray(self, table: DeviceTable, rig: CameraRig, params: Params, method: str, keys: bool = True,
                    stream=None):
        """The 3dgut, htgs or raygs words (SplatProjection.write, src/render/SplatProjection.ts:141-151):
        {"channels": [w0 (n, 4), w1 (n, 4)(, w2 (n, 2))] int32, "words": (n, 6 + 2 eyes; raygs 4 + 3 eyes), "keys"}."""
        import torch

        if method not in self.METHODS:
            raise ValueError(f"project_ray: method {method!r} is not one of {sorted(self.METHODS)}")
        n = table.count
        eyes = len(rig.eyes)
        dev = torch.device("cuda")
        m = max(n, 1)
        chans = [torch.zeros((m, 4), dtype=torch.int32, device=dev), torch.zeros((m, 4), dtype=torch.int32, device=dev)]
        if eyes == 2:
            chans.append(torch.zeros((m, 2), dtype=torch.int32, device=dev))
        kd = torch.zeros(m, dtype=torch.int32, device=dev) if keys else None
        cp = (C.c_void_p * 3)(*[c.data_ptr() for c in chans] + [None] * (3 - len(chans)))
        N.check(N.lib().rdx_project_ray(self.context.handle, _stream(stream), C.byref(table.desc()), C.byref(rig.desc()),
                                        C.byref(params.desc()), self.METHODS[method], 0, cp, _ptr(kd) if keys else None))
        words = torch.cat(chans, dim=1)[:n, : (4 + 3 * eyes) if method == "raygs" else (6 + 2 * eyes)]
        return {"channels": [c[:n] for c in chans], "words": words, "keys": kd[:n] if keys else None}

    @on_stream
    def project_sort(self, table: DeviceTable, rig: CameraRig, params: Params, words=None, keys=None, indices=None,
                     count=None, args=None, side: bool = False, stream=None):
        """The fused path: words, keys, the visible indices back to front, count and draw arguments.

        With per-eye keys (stereo, params.key_per_eye) each eye is its own
        sort: eye e's keys and indices at [e n, (e + 1) n), its count at
        count[e] and its draw arguments at args[5 e:5 e + 5]."""
        import torch

        n = table.count
        eyes = len(rig.eyes)
        segs = 2 if (eyes == 2 and params.key_per_eye) else 1
        dev = torch.device("cuda")
        if words is None:
            words = [torch.zeros((max(n, 1), 4), dtype=torch.int32, device=dev) for _ in range(eyes)]
        if keys is None:
            keys = torch.zeros(max(n * segs, 1), dtype=torch.int32, device=dev)
        if indices is None:
            indices = torch.zeros(max(n * segs, 1), dtype=torch.int32, device=dev)
        if count is None:
            count = torch.zeros(segs, dtype=torch.int32, device=dev)
        if args is None:
            args = torch.zeros(5 * segs, dtype=torch.int32, device=dev)
        if keys.numel() < n * segs or indices.numel() < n * segs or count.numel() < segs or args.numel() < 5 * segs:
            raise ValueError(f"project_sort: outputs too small for {segs} segment(s) of {n} splats")
        wp = (C.c_void_p * 2)(*[w.data_ptr() for w in words] + [None] * (2 - eyes))
        fo = None
        out = {}
        if side:
            sd, fo = self._side(n, eyes, segs == 2, dev)
            out["side"] = sd
        N.check(N.lib().rdx_project_sort(self.context.handle, _stream(stream), C.byref(table.desc()),
                                         C.byref(rig.desc()), C.byref(params.desc()), wp, _ptr(keys), _ptr(indices),
                                         _ptr(count), _ptr(args), C.byref(fo) if fo is not None else None))
        out.update(words=words, keys=keys, indices=indices, count=count, args=args)
        return out

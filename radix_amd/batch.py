"""Batched views: many (view, table) segments projected and sorted in one
sequence (rdx_batch_*). No reference counterpart: the reference re-projects
and re-sorts per render and per object (src/render/SplatProjector.ts:127-137);
a segment is what one SplatProjector.project + WgpuSplatSorter.sort of that
view would produce, at the offsets SplatCollection gives its objects
(src/render/SplatCollection.ts:54-65).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

from . import _native as N
from .camera import CameraRig
from .projection import DeviceTable, Params
from .sort import Context, _stream, default_context


class Batch:
    """views: (table index, mono CameraRig) pairs. Outputs per view v:
    words[v] (its table's count x 4 int32, if words=True), and after run()
    indices[off_v : off_v + counts[v]] (back to front), counts[v] and
    args[5 v : 5 v + 5]; keys[off_v : off_v + n_v] hold its sort keys."""

    def __init__(self, tables: Sequence[DeviceTable], views: Sequence[Tuple[int, CameraRig]], params: Params,
                 words: bool = True, context: Optional[Context] = None, device="cuda", tuning=None,
                 tune_reps: int = 5, stream=None):
        """tuning: None (the context's), a (digit bits, keys per thread,
        threads) tuple, or "auto": rdx_batch_autotune picks it by timing this
        batch under every supported tuning (a cached result for the same
        arch, key bits and sizes is taken without timing). The choice is in
        self.tuning, its time per run in self.tuning_ms and "autotuned",
        "cache", "given" or "context" in self.tuning_source."""
        import torch

        if tuning is not None and tuning != "auto":
            context = Context(tuple(tuning))
        self.context = context or (Context() if tuning == "auto" else default_context())
        self.tables = list(tables)
        self.views = [(int(t), rig) for t, rig in views]
        self.params = params
        dev = torch.device(device)
        self.counts_of_view = [self.tables[t].count for t, _ in self.views]
        self.total = sum(self.counts_of_view)
        self.keys = torch.zeros(max(self.total, 1), dtype=torch.int32, device=dev)
        self.words = ([torch.zeros((max(n, 1), 4), dtype=torch.int32, device=dev) for n in self.counts_of_view]
                      if words else None)
        self._tdesc = (N.SplatTable * len(self.tables))(*[t.desc() for t in self.tables])
        bv = (N.BatchView * len(self.views))()
        for v, (t, rig) in enumerate(self.views):
            if len(rig.eyes) != 1:
                raise ValueError(f"batch view {v}: one eye per view")
            bv[v].words_dev = self.words[v].data_ptr() if words else None
            bv[v].table = t
            bv[v].camera = rig.desc()
        self._views_desc = bv
        self.tuning_ms = None
        self.tuning_source = "given" if tuning not in (None, "auto") else "context"
        if tuning == "auto":
            t = N.SortTuning()
            ms = C.c_double()
            cached = C.c_int()
            N.check(N.lib().rdx_batch_autotune(self.context.handle, _stream(stream), self._tdesc, len(self.tables), bv,
                                               len(self.views), C.byref(params.desc()), tune_reps, N.TUNE_CACHED,
                                               C.byref(t), C.byref(ms), C.byref(cached)))
            self.tuning_ms = ms.value
            self.tuning_source = "cache" if cached.value else "autotuned"
        self.tuning = self.context.tuning()
        h = C.c_void_p()
        N.check(N.lib().rdx_batch_create(self.context.handle, self._tdesc, len(self.tables), bv, len(self.views),
                                          C.byref(params.desc()), C.c_void_p(self.keys.data_ptr()), C.byref(h)))
        self._h = h
        lib = N.lib()
        assert lib.rdx_batch_total(h) == self.total
        self.offsets = [int(lib.rdx_batch_segment_offset(h, v)) for v in range(len(self.views))]
        self.indices = torch.zeros(max(self.total, 1), dtype=torch.int32, device=dev)
        self.counts = torch.zeros(len(self.views), dtype=torch.int32, device=dev)
        self.args = torch.zeros(5 * len(self.views), dtype=torch.int32, device=dev)
        self.persistent = False

    def segment(self, v: int) -> Tuple[int, int]:
        """(offset, splat count) of view v's segment."""
        return self.offsets[v], self.counts_of_view[v]

    def set_cameras(self, rigs: Sequence[CameraRig]) -> None:
        cams = (N.Camera * len(rigs))(*[r.desc() for r in rigs])
        N.check(N.lib().rdx_batch_set_cameras(self._h, cams, len(rigs)))
        self.views = [(t, r) for (t, _), r in zip(self.views, rigs)]

    def run(self, graph: bool = True, stream=None, persistent: Optional[bool] = None) -> None:
        """persistent: True runs the whole batch as one launch of the
        persistent kernel (tables of one kind, tuning (8, 12, 512), 32-bit keys);
        None takes self.persistent (False unless choose_form() picked it)."""
        if persistent is None:
            persistent = self.persistent
        flags = (N.BATCH_GRAPH if graph else 0) | (N.BATCH_PERSISTENT if persistent else 0)
        N.check(N.lib().rdx_batch_run(self._h, _stream(stream), C.c_void_p(self.indices.data_ptr()),
                                      C.c_void_p(self.counts.data_ptr()), C.c_void_p(self.args.data_ptr()), flags))

    @property
    def persistent_grid(self) -> int:
        """Blocks of the persistent launch (0: this batch has no persistent form)."""
        return int(N.lib().rdx_batch_persistent_grid(self._h))

    PHASES = ["project"] + [f"{k}{p}" for p in range(3) for k in ("count", "scan", "scatter")] + ["scatter3"]

    def persistent_phases(self) -> dict:
        """{phase: (ms, barrier release ms)} of the last persistent run made
        while the batch's context profiled ({} if none)."""
        ph = (C.c_double * 16)()
        rel = (C.c_double * 16)()
        n = N.lib().rdx_batch_persistent_phases(self._h, ph, rel, 16)
        if n < 0:
            N.check(n)
        return {self.PHASES[k]: (ph[k], rel[k]) for k in range(n)}

    def status(self) -> None:
        """Waits for the last run; raises if a persistent run's barrier wait
        ran out (its output is then not valid)."""
        N.check(N.lib().rdx_batch_status(self._h))

    def choose_form(self, reps: int = 5, stream=None) -> Tuple[bool, float, Optional[float]]:
        """Times graph replays of the launch sequence and of the persistent
        launch (when this batch has one) and keeps the faster in
        self.persistent: (persistent, ms of the sequence, ms persistent or None)."""
        p = C.c_int()
        mg = C.c_double()
        mp = C.c_double()
        N.check(N.lib().rdx_batch_choose_form(self._h, _stream(stream), reps, C.byref(p), C.byref(mg), C.byref(mp)))
        self.persistent = bool(p.value)
        return self.persistent, mg.value, (mp.value if mp.value >= 0 else None)

    def close(self) -> None:
        if getattr(self, "_h", None):
            N.lib().rdx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Benchmark: project + SH + sort keys + stable radix sort of 4,194,304 splats
(SH degree 3, 1920x1080, 32-bit keys) on one MI355X, as Gsplats/s and as a
fraction of HBM peak. Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

A step is one rdx_project_sort over the table already resident in HBM:
projection kernel (words, keys), then the radix
passes (compaction in the first), writing the visible indices back to front.
"""
from __future__ import annotations

import argparse
import contextlib
import ctypes as C
import gc
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
B_IN_SH3 = 236                 # f32 planes: centre 12, log-scale 12, quat 16, opacity 4, colour + SH 192
B_OUT = 20                     # ewa words 16 + key 4 (one eye)
B_SORT = 60                    # 4 passes of 32-bit keys at 8-bit digits
METRIC = "Gsplats/s project+SH+sort @4M SH3, % HBM peak; batched vs plain; bwd Gsplats/s"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def stage_times(ctx) -> dict:
    from radix_amd import _native as N

    names = C.create_string_buffer(32 * 32)
    ms = (C.c_double * 32)()
    cnt = (C.c_int * 32)()
    k = N.lib().rdx_context_stage_times(ctx.handle, names, ms, cnt, 32)
    if k < 0:
        N.check(k)
    out = {}
    for i in range(k):
        name = names.raw[32 * i: 32 * i + 32].split(b"\0", 1)[0].decode()
        out[name] = {"ms_total": ms[i], "launches": cnt[i], "ms_avg": ms[i] / max(cnt[i], 1)}
    return out


def host_cpu() -> dict:
    """The host the CPU baselines ran on: its model, logical CPUs and those this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"model": model, "cpu_count": os.cpu_count(), "usable_cpus": usable}


def distances_of(keys):
    """Distances whose 32-bit keys are `keys`: the float bits of ~key, infinity for the invalid key."""
    import numpy as np

    k = np.ascontiguousarray(keys, dtype=np.uint32)
    dist = (~k).view(np.float32).astype(np.float64)
    dist[k == 0xFFFFFFFF] = np.inf
    return dist


def cpu_baseline(keys_u32, n: int, planes, rig, reps: int = 15, warm: int = 3) -> dict:
    """The oracle's restatement of src/core's SortKeys.compute + SortKeys.sort
    (32 bits) on the host, one thread (a sort is single-threaded in the
    reference), median of `reps` after `warm` warm-ups
    (scripts/bench-core.ts:261-311); and the "+SH" figure: the same plus
    Sh.evaluate of every splat at SH3 (src/core/sh.ts:362-373) from the
    head, timed over the same 4M table."""
    import numpy as np

    from oracle import oracle as O

    lib = O.lib()
    keys = np.ascontiguousarray(keys_u32[:n], dtype=np.uint32)
    dist = distances_of(keys)
    out_keys = np.empty(n, dtype=np.uint32)
    ordering = np.empty(n, dtype=np.uint32)
    times = []
    for r in range(warm + reps):
        t0 = time.perf_counter()
        lib.orc_keys_compute(dist.ctypes.data, n, 32, out_keys.ctypes.data)
        lib.orc_sort(out_keys.ctypes.data, n, 32, ordering.ctypes.data)
        dt = time.perf_counter() - t0
        if r >= warm:
            times.append(dt)
    med = statistics.median(times)
    # Sh.evaluate per splat: centres and the 45 SH planes of the table, colours accumulated in a copy
    p = np.ascontiguousarray(planes[:, :n], dtype=np.float32)
    rgb = np.ascontiguousarray(p[11:14])
    origin = np.linalg.inv(rig.head)[:3, 3].astype(np.float64)
    sh_times = []
    for r in range(4):
        t0 = time.perf_counter()
        lib.orc_sh_apply_table(n, n, p[0:3].ctypes.data, p[14:59].ctypes.data, 3, O.dptr(np.ascontiguousarray(origin)),
                               rgb.ctypes.data)
        sh_times.append(time.perf_counter() - t0)
    sh = statistics.median(sh_times[1:])
    return {"value": n / med / 1e9, "unit": "Gsplats/s", "cores": 1, "kind": "port",
            "host": host_cpu(),
            "sample": f"oracle/sort.c restating src/core SortKeys.compute + SortKeys.sort (32-bit, 2x16-bit LSD) "
                      f"over the {n:,} keys of one step, median of {reps} after {warm} warm-ups, 1 thread; "
                      f"median {med * 1e3:.1f} ms",
            "plus_sh": {"value": n / (med + sh) / 1e9, "unit": "Gsplats/s", "cores": 1, "kind": "port",
                        "sample": f"the same plus oracle/codec.c's Sh.evaluate (sh.ts:362-373) of all {n:,} splats "
                                  f"at SH3 from the head, median {sh * 1e3:.1f} ms of 3 after a warm-up"}}


def cpu_baseline_views(keys_u32, segments, reps: int = 3) -> dict:
    """The reference's CPU sort path for many views: SortKeys.compute + sort
    of each view's keys on a pool of min(32, usable CPUs) threads, one view
    per task (SplatView.sort on the worker pool, SplatRenderer.ts:106-116;
    the oracle's C restatement releases the GIL under ctypes)."""
    import concurrent.futures as cf

    import numpy as np

    from oracle import oracle as O

    lib = O.lib()
    host = host_cpu()
    threads = max(1, min(32, host["usable_cpus"] or 1))
    work = []
    for off, n in segments:
        k = np.ascontiguousarray(keys_u32[off:off + n], dtype=np.uint32)
        work.append((distances_of(k), np.empty(n, dtype=np.uint32), np.empty(n, dtype=np.uint32)))

    def one(w):
        dist, out_keys, ordering = w
        n = dist.size
        lib.orc_keys_compute(dist.ctypes.data, n, 32, out_keys.ctypes.data)
        lib.orc_sort(out_keys.ctypes.data, n, 32, ordering.ctypes.data)

    total = sum(n for _, n in segments)
    times = []
    with cf.ThreadPoolExecutor(threads) as pool:
        for r in range(reps + 1):
            t0 = time.perf_counter()
            list(pool.map(one, work))
            if r:
                times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": total / med / 1e9, "unit": "Gsplat-views/s", "cores": threads, "kind": "port", "host": host,
            "sample": f"oracle/sort.c's SortKeys.compute + sort of each of the {len(segments)} views' keys "
                      f"({total:,} in all), one view per task on {threads} threads, median {med * 1e3:.1f} ms of {reps}"}


def oracle_projection_rate(planes, rig, count: int) -> dict:
    """The FP64 oracle's projection + SH + key on a sample (labelled oracle: src/core has no CPU projection)."""
    import numpy as np

    from oracle import oracle as O

    sample = np.ascontiguousarray(planes[:, :count])
    t0 = time.perf_counter()
    O.project_f32(sample, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    dt = time.perf_counter() - t0
    return {"value": count / dt / 1e9, "unit": "Gsplats/s", "cores": 1, "kind": "oracle",
            "sample": f"{count:,} splats of the same table, FP64 projection + SH3 + key, 1 thread"}


@contextlib.contextmanager
def no_gc():
    """Python's cyclic garbage collector held off inside a timed loop (a full
    collection, tens of ms with torch loaded, would otherwise land between
    two launches at a point fixed by the process's allocation count and
    idle the GPU); collected before and after."""
    gc.collect()
    gc.disable()
    try:
        yield
    finally:
        gc.enable()


def time_events(fn, steps: int, warmup: int = 2) -> float:
    """ms per call of fn() on the current stream, HIP events around `steps` calls."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with no_gc():
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def bench_batched(steps: int, with_cpu: bool = True) -> dict:
    """Batched views (8 collections, 32 views at 1280x720): one graph replay of
    the batch against 32 plain rdx_project_sort calls, both as splat-views/s
    and as a fraction of HBM peak over sum_c 236 n_c + sum_v 80 n_v bytes."""
    import torch

    from radix_amd import Batch, DeviceTable, Params, Projector, workloads as W
    from radix_amd.sort import Context

    wl = W.batched()
    tables = [DeviceTable.from_f32_planes(W.generate(sp), sp.count, 3) for sp in wl.specs]
    views = list(zip(wl.view_table, wl.cameras))
    prm = Params(sort_bits=32)
    ctx = Context()
    # the batch's sort tuning from the autotuner (every supported tuning timed on this batch)
    b = Batch(tables, views, prm, tuning="auto", tune_reps=3)
    tuning = {"digit_bits": b.tuning[0], "keys_per_thread": b.tuning[1], "block_threads": b.tuning[2],
              "source": b.tuning_source, "tuned_ms": round(b.tuning_ms, 4)}
    splat_views = b.total
    bytes_ = 236 * sum(sp.count for sp in wl.specs) + 80 * splat_views
    ms_b = time_events(lambda: b.run(graph=True), steps)
    # the persistent form: the whole batch in one launch (its own 12-key tiles), replayed from a graph
    bp = Batch(tables, views, prm, tuning=(8, 12, 512))
    ms_pers = time_events(lambda: bp.run(graph=True, persistent=True), steps)
    bp.status()
    ms_pers_seq = time_events(lambda: bp.run(graph=True, persistent=False), steps)
    pers_grid = bp.persistent_grid
    bp.close()
    del bp
    proj = Projector(ctx)
    outs = []
    for t, rig in views:
        n = tables[t].count
        outs.append(dict(words=[torch.empty((n, 4), dtype=torch.int32, device="cuda")],
                         keys=torch.empty(n, dtype=torch.int32, device="cuda"),
                         indices=torch.empty(n, dtype=torch.int32, device="cuda"),
                         count=torch.zeros(1, dtype=torch.int32, device="cuda"),
                         args=torch.zeros(5, dtype=torch.int32, device="cuda")))

    def plain_all():
        for (t, rig), o in zip(views, outs):
            proj.project_sort(tables[t], rig, prm, **o)

    ms_p = time_events(plain_all, steps)
    plain_bytes = sum((236 + 80) * tables[t].count for t, _ in views)
    cpu = None
    if with_cpu:
        torch.cuda.synchronize()
        keys_h = b.keys.cpu().numpy().view("uint32")
        cpu = cpu_baseline_views(keys_h, [b.segment(v) for v in range(len(views))])
    b.close()
    return {"workload": "8 collections (log-uniform 1e4..6e6 splats, SH3 f32), 32 views 1280x720 on radius 4-12 orbits",
            "collections": [sp.count for sp in wl.specs], "splat_views": splat_views,
            "batched": {"ms": round(ms_b, 4), "Gsplat_views_per_s": round(splat_views / ms_b / 1e6, 3),
                        "hbm_frac": round(bytes_ / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "bytes": bytes_, "tuning": tuning,
                        "form": "one HIP graph: 1 projection launch + the segmented sort's passes"},
            "persistent": {"ms": round(ms_pers, 4), "Gsplat_views_per_s": round(splat_views / ms_pers / 1e6, 3),
                           "hbm_frac": round(bytes_ / (ms_pers * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "grid": pers_grid,
                           "tuning": [8, 12, 512], "same_tuning_sequence_ms": round(ms_pers_seq, 4),
                           "form": "one HIP graph of one launch: projection and every radix pass with grid barriers"},
            "best_ms": round(min(ms_b, ms_pers), 4),
            "plain": {"ms": round(ms_p, 4), "Gsplat_views_per_s": round(splat_views / ms_p / 1e6, 3),
                      "hbm_frac_of_batched_bytes": round(bytes_ / (ms_p * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "bytes_read_and_written": plain_bytes, "form": "32 rdx_project_sort calls"},
            "speedup": round(ms_p / min(ms_b, ms_pers), 3), "cpu_baseline": cpu}


def bench_stereo(steps: int, count: int = 16777216) -> dict:
    """Quantized stereo: packedSplats + packedSh SH3, two 2064x2208 eyes,
    per-eye keys sorted as two segments; 64 + 2 x (20 + 60) B per splat."""
    from radix_amd import DeviceTable, Params, Projector, workloads as W

    spec = W.stereo_16m_spec(count=count)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    del planes
    table = DeviceTable.from_packed(words, spec.count, (0.0, 1.0, -10.0, 2.0), chans, 3)
    del words, chans
    rig = W.stereo_camera()
    prm = Params(sort_bits=32, key_per_eye=True)
    proj = Projector()
    out = proj.project_sort(table, rig, prm)
    ms = time_events(lambda: proj.project_sort(table, rig, prm, words=out["words"], keys=out["keys"],
                                               indices=out["indices"], count=out["count"], args=out["args"]), steps)
    n = spec.count
    bytes_ = (64 + 2 * (20 + 60)) * n
    from radix_amd import _native as N

    N.lib().rdx_context_profile(proj.context.handle, 1)
    stage_times(proj.context)
    for _ in range(3):
        proj.project_sort(table, rig, prm, words=out["words"], keys=out["keys"], indices=out["indices"],
                          count=out["count"], args=out["args"])
    stages = stage_times(proj.context)
    N.lib().rdx_context_profile(proj.context.handle, 0)
    return {"workload": "16,777,216 packedSplats + packedSh SH3, stereo 2064x2208 (64 mm), per-eye segmented sort",
            "stages_ms": {k: round(v["ms_avg"], 4) for k, v in stages.items()},
            "ms": round(ms, 4), "Gsplats_per_s": round(n / ms / 1e6, 3),
            "hbm_frac": round(bytes_ / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "bytes_per_splat": 224,
            "visible_per_eye": [int(v) for v in out["count"].cpu().tolist()]}


def bench_draw(steps: int, frame, words, idx, cnt, width: int, height: int, params, ctx) -> dict:
    """The production draw of the plain_4m frame: its sorted, visible splats'
    ewa words composited into a 1920x1080 premultiplied RGBA image (every
    fragment, as the reference's blend). Also a whole frame: project, sort
    and draw."""
    from radix_amd import _native as N
    from radix_amd.draw import Drawer

    drawer = Drawer(context=ctx)
    image = drawer.draw(words, idx, cnt, width, height, params)
    ms = time_events(lambda: drawer.draw(words, idx, cnt, width, height, params, out=image), steps)

    def whole():
        frame()
        drawer.draw(words, idx, cnt, width, height, params, out=image)

    ms_frame = time_events(whole, steps)
    st = drawer.stats()
    N.lib().rdx_context_profile(ctx.handle, 1)
    stage_times(ctx)
    for _ in range(3):
        drawer.draw(words, idx, cnt, width, height, params, out=image)
    stages = stage_times(ctx)
    N.lib().rdx_context_profile(ctx.handle, 0)
    drawer.close()
    return {"workload": f"the plain_4m frame's {st['drawn']:,} sorted splats drawn at {width}x{height} "
                        "(ewa fragments, premultiplied over, no early stop)",
            "ms": round(ms, 4), "Gsplats_per_s": round(st["drawn"] / ms / 1e6, 3), "entries": st["entries"],
            "frame_ms": round(ms_frame, 4), "frame_fps": round(1e3 / ms_frame, 1),
            "stages_ms": {(k if k.startswith("draw") else "entries_" + k): round(v["ms_avg"], 4)
                          for k, v in stages.items()}}


def bench_keys(steps: int, proj, table, rig) -> dict:
    """The keys-only pass (the CPU sort's keys) over the plain_4m table:
    centre, log-scales, quaternion and opacity in (44 B), one key out (4 B)."""
    import torch

    from radix_amd import Params
    from radix_amd import _native as N

    prm = Params(sort_bits=32)
    n = table.count
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    proj.project_keys(table, rig, prm, keys=keys)
    ms = time_events(lambda: proj.project_keys(table, rig, prm, keys=keys), steps)
    N.lib().rdx_context_profile(proj.context.handle, 1)
    stage_times(proj.context)
    for _ in range(steps):
        proj.project_keys(table, rig, prm, keys=keys)
    k_ms = stage_times(proj.context).get("project_keys", {}).get("ms_avg", float("nan"))
    N.lib().rdx_context_profile(proj.context.handle, 0)
    bytes_ = 48 * n
    return {"workload": "keys-only pass over the plain_4m table (4,194,304 f32 splats), 32-bit keys",
            "ms": round(ms, 4), "kernel_ms": round(k_ms, 4), "Gsplats_per_s": round(n / ms / 1e6, 3),
            "bytes_per_splat": 48, "hbm_frac": round(bytes_ / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def bench_loaders(steps: int, n: int = 4194304) -> dict:
    """The loaders' device kernels on synthetic bytes (any bytes are valid
    input): the RDX untranspose of an extSplats chunk (two 16-byte channels:
    32 B in, 32 B out per splat) and the SPZ decode of a version-3 SH3 body
    (65 B in, the 59 f32 planes out: 236 B), HIP events on the launch stream."""
    import ctypes as C

    import torch

    from radix_amd import _native as N
    from radix_amd.sort import _stream

    lib = N.lib()
    g = torch.Generator(device="cuda").manual_seed(5)
    planes = [torch.randint(0, 256, (16 * n,), dtype=torch.uint8, device="cuda", generator=g) for _ in range(2)]
    outs = [torch.empty(16 * n, dtype=torch.uint8, device="cuda") for _ in range(2)]

    def untranspose():
        for p, o in zip(planes, outs):
            N.check(lib.rdx_untranspose_bytes(_stream(), C.c_void_p(p.data_ptr()), n, 16, C.c_void_p(o.data_ptr())))

    rdx_ms = time_events(untranspose, steps)
    del planes, outs
    body = torch.randint(0, 256, (65 * n,), dtype=torch.uint8, device="cuda", generator=g)
    out = torch.empty((59, n), dtype=torch.float32, device="cuda")

    def spz():
        N.check(lib.rdx_spz_decode(_stream(), C.c_void_p(body.data_ptr()), body.numel(), n, 3, 3, 12,
                                   C.c_void_p(out.data_ptr()), n))

    spz_ms = time_events(spz, steps)
    frac = lambda b, ms: round(b * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return {"workload": "4,194,304 splats of synthetic bytes: RDX untranspose of an extSplats chunk (extA, extB), "
                        "SPZ v3 SH3 body decode to f32 planes",
            "rdx_untranspose": {"ms": round(rdx_ms, 4), "bytes_per_splat": 64, "hbm_frac": frac(64, rdx_ms)},
            "spz_decode": {"ms": round(spz_ms, 4), "bytes_per_splat": 301, "hbm_frac": frac(301, spz_ms)}}


def bench_ray(steps: int, proj, table, rig) -> dict:
    """The ray methods' projection words over the plain_4m table: the f32
    SH3 splat in (236 B), a key and 8 words out (36 B; raygs writes 7 of them
    and a zero). Then each method's frame sorted (rdx_sort_splats) and drawn
    (rdx_draw_ray) at 1920x1080."""
    import torch

    from radix_amd import Params
    from radix_amd import _native as N
    from radix_amd.draw import Drawer
    from radix_amd.sort import SplatSorter

    import ctypes as C

    prm = Params(sort_bits=32)
    sorter = SplatSorter(table.count, context=proj.context)
    drawer = Drawer(context=proj.context)
    image = torch.empty((4, int(rig.height), int(rig.width)), dtype=torch.float32, device="cuda")
    n = table.count
    chans = [torch.empty((n, 4), dtype=torch.int32, device="cuda") for _ in range(2)]
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    cp = (C.c_void_p * 3)(chans[0].data_ptr(), chans[1].data_ptr(), None)
    out = {}
    for name, m in (("3dgut", 1), ("htgs", 2), ("raygs", 3)):
        def call():
            N.check(N.lib().rdx_project_ray(proj.context.handle, None, C.byref(table.desc()), C.byref(rig.desc()),
                                            C.byref(prm.desc()), m, 0, cp, C.c_void_p(keys.data_ptr())))
        call()
        ms = time_events(call, steps)
        N.lib().rdx_context_profile(proj.context.handle, 1)
        stage_times(proj.context)
        for _ in range(steps):
            call()
        k_ms = stage_times(proj.context).get(f"project_{name}", {}).get("ms_avg", float("nan"))
        N.lib().rdx_context_profile(proj.context.handle, 0)
        out[name] = {"ms": round(ms, 4), "kernel_ms": round(k_ms, 4), "Gsplats_per_s": round(n / ms / 1e6, 3),
                     "hbm_frac": round(272 * n / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        sorter.sort(keys, n, 32)
        order = sorter.values
        draw = lambda: drawer.draw_ray(name, chans, order, sorter.count, rig, prm, out=image)  # noqa: E731
        draw()
        draw_ms = time_events(draw, steps)
        st = drawer.stats()
        out[name].update({"draw_ms": round(draw_ms, 4), "drawn": st["drawn"], "draw_entries": st["entries"]})
    drawer.close()
    out["workload"] = ("3dgut, htgs and raygs words of the plain_4m table (4,194,304 f32 SH3 splats, 1920x1080), "
                       "32-bit keys; draw_ms: the sorted frame drawn (every fragment, premultiplied over)")
    out["bytes_per_splat"] = 272
    return out


def bench_backward(steps: int) -> dict:
    """Differentiable compositor: 262,144 splats, 512x512, 16x16 tiles, L2
    loss against the perturbed scene; bwd Gsplats/s = n / backward time."""
    import torch

    from radix_amd import Params
    from radix_amd import diff as D

    theta, target_theta, rig = D.workload()
    rast = D.Rasterizer(sh_degree=3)
    target, _ = rast.forward(D.params_tensor(target_theta), rig, Params())
    target = target.clone()
    p = D.params_tensor(theta)
    prm = Params()
    img, _ = rast.forward(p, rig, prm)
    gi = 2.0 * (img - target) / img.numel()
    ms_f = time_events(lambda: rast.forward(p, rig, prm), steps)
    rast.forward(p, rig, prm)
    ms_b = time_events(lambda: rast.backward(gi), steps)
    st = rast.stats()
    from radix_amd import _native as N

    N.lib().rdx_context_profile(rast.context.handle, 1)
    stage_times(rast.context)
    for _ in range(3):
        rast.forward(p, rig, prm)
        rast.backward(gi)
    stages = stage_times(rast.context)
    N.lib().rdx_context_profile(rast.context.handle, 0)
    n = p.shape[1]
    px = rig.width * rig.height
    bytes_ = (236 + 236 + 40) * n + 20 * px
    return {"workload": "262,144 splats (SH3), 512x512, 16x16 tiles, T < 1e-4 stop, L2 loss; step = rdx_raster_backward",
            "bwd_ms": round(ms_b, 4), "bwd_Gsplats_per_s": round(n / ms_b / 1e6, 3),
            "fwd_ms": round(ms_f, 4), "fwd_Gsplats_per_s": round(n / ms_f / 1e6, 3),
            "entries": st["entries"], "hbm_frac_lower_bound": round(bytes_ / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "stages_ms": {k: round(v["ms_avg"], 4) for k, v in stages.items()}}


def measured_traffic(n: int):
    """HBM bytes per k_project launch at this size, from the committed
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, each scaled by the
    same counter over a copy of known size), or None if not measured."""
    path = os.path.join(ROOT, "profiles", "plain_4m_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if t.get("splats") != n:
        return None
    return t.get("k_project_bytes")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--count", type=int, default=4194304)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning", default=None,
                    help="the headline step's sort tuning as digit_bits,keys_per_thread,block_threads (default: the library's)")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and stereo workloads")
    ap.add_argument("--only", choices=["batched", "stereo", "differentiable"], default=None,
                    help="run one extra workload alone and print its object (for profiling its kernels)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of 512 MiB five times (the reference kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU design is described, not built)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if args.copy_calibration:
        src = torch.ones(512 << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "copies": 5}))
        return 0

    from radix_amd import DeviceTable, Params, Projector, workloads as W
    from radix_amd.sort import Context

    if args.only:
        steps = max(3, args.steps // 5)
        if args.only == "batched":
            res = bench_batched(steps, with_cpu=False)
        elif args.only == "stereo":
            res = bench_stereo(steps)
        else:
            res = bench_backward(steps)
        print(json.dumps({args.only: res}), flush=True)
        return 0

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log(f"generating plain_4m ({args.count:,} splats, SH3)")
    spec = W.plain_4m_spec(count=args.count)
    planes = W.generate(spec)
    rig = W.plain_4m_camera()
    table = DeviceTable.from_f32_planes(planes, spec.count, 3)
    params = Params(sort_bits=32)
    ctx = Context(tuple(int(x) for x in args.tuning.split(",")) if args.tuning else None)
    proj = Projector(ctx)
    n = spec.count
    words = [torch.empty((n, 4), dtype=torch.int32, device=dev)]
    keys = torch.empty(n, dtype=torch.int32, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    args5 = torch.zeros(5, dtype=torch.int32, device=dev)

    def step():
        proj.project_sort(table, rig, params, words=words, keys=keys, indices=idx, count=cnt, args=args5)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    with no_gc():
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            step()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    visible = int(cnt.cpu().item())

    # Per-kernel times with HIP events on the launch stream (second timed region)
    N_lib = __import__("radix_amd._native", fromlist=["lib"]).lib()
    N_lib.rdx_context_profile(ctx.handle, 1)
    stage_times(ctx)
    with no_gc():
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
    stages = stage_times(ctx)
    N_lib.rdx_context_profile(ctx.handle, 0)
    proj_ms = stages.get("project", {}).get("ms_avg", float("nan"))
    kernel_bytes = (B_IN_SH3 + B_OUT) * n
    achieved = kernel_bytes / (proj_ms * 1e-3) / 1e9
    pipeline_bytes = (B_IN_SH3 + B_OUT + B_SORT) * n
    value = n / (ms * 1e-3) / 1e9
    log(f"{ms:.3f} ms/step, {value:.2f} Gsplats/s, project {proj_ms:.3f} ms ({achieved:.0f} GB/s), visible {visible:,}")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gsplats/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 64-cluster cloud, random orbit camera)",
        "config": {"workload": "plain_4m: 4,194,304 splats, SH3 f32 planes (236 B/splat), 1920x1080 camera on a "
                               "radius-8 orbit, 60 deg, 32-bit keys, cullKeys on; step = rdx_project_sort",
                   "splats": n, "visible": visible, "sort_bits": 32, "sort_tuning": list(ctx.tuning())},
        "hbm_frac_pipeline": round(pipeline_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "pipeline_bytes_per_splat": B_IN_SH3 + B_OUT + B_SORT,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(n),
                     "kernel": "k_project (projection + SH + key)",
                     "bytes_per_splat": B_IN_SH3 + B_OUT},
        "stages_ms": {k: round(v["ms_avg"], 5) for k, v in stages.items()},
        "wall_s_timed": round(wall, 4),
    }
    if not args.no_extra:
        log("draw of the plain_4m frame")
        out["draw_4m"] = bench_draw(max(3, args.steps // 5), step, words[0], idx, cnt, int(rig.width), int(rig.height),
                                    params, ctx)
        log("ray projection methods")
        out["ray_4m"] = bench_ray(max(3, args.steps // 5), proj, table, rig)
        log("keys-only pass")
        out["keys_pass_4m"] = bench_keys(max(3, args.steps // 5), proj, table, rig)
        torch.cuda.empty_cache()
        log("loaders")
        out["loaders_4m"] = bench_loaders(max(3, args.steps // 5))
        torch.cuda.empty_cache()
        log("batched workload")
        out["batched_vs_plain"] = bench_batched(max(3, args.steps // 5), with_cpu=not args.no_cpu)
        torch.cuda.empty_cache()
        log("stereo 16M workload")
        out["stereo_16m"] = bench_stereo(max(3, args.steps // 5))
        torch.cuda.empty_cache()
        log("differentiable compositor")
        out["differentiable"] = bench_backward(max(3, args.steps // 5))
        torch.cuda.empty_cache()
    if not args.no_cpu:
        log("cpu baseline")
        keys_h = keys.cpu().numpy().view(np.uint32)
        out["cpu_baseline"] = cpu_baseline(keys_h, n, planes, rig)
        out["cpu_oracle_projection"] = oracle_projection_rate(planes, rig, 262144)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

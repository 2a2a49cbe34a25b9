"""The batched form against the plain path: each view's segment of a batch
equals rdx_project_sort of that view alone (words, keys, count, indices and
draw arguments, bit for bit), and the plain path is itself held to the
oracle by test_project_gpu. Graph replays are checked against the eager
run with other GPU work between them."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import Batch, DeviceTable, Params, Projector, workloads as W
from tests.projection_check import _decode_rgb, _fields

pytestmark = pytest.mark.gpu


def plain(table, rig, prm):
    out = Projector().project_sort(table, rig, prm)
    torch.cuda.synchronize()
    n = table.count
    c = int(out["count"].cpu().item())
    return {
        "words": out["words"][0].cpu().numpy().view(np.uint32)[:n],
        "keys": out["keys"].cpu().numpy().view(np.uint32)[:n],
        "count": c,
        "indices": out["indices"].cpu().numpy().view(np.uint32)[:c],
        "args": out["args"].cpu().numpy().view(np.uint32)[:5],
    }


def check_views(b: Batch, tables, prm, views=None):
    torch.cuda.synchronize()
    keys = b.keys.cpu().numpy().view(np.uint32)
    idx = b.indices.cpu().numpy().view(np.uint32)
    counts = b.counts.cpu().numpy()
    args = b.args.cpu().numpy().view(np.uint32).reshape(-1, 5)
    for v, (t, rig) in enumerate(views or b.views):
        off, n = b.segment(v)
        ref = plain(tables[t], rig, prm)
        print(f"view {v}: table {t}, {n} splats, {ref['count']} visible", flush=True)
        assert np.array_equal(keys[off:off + n], ref["keys"]), v
        assert counts[v] == ref["count"], v
        assert np.array_equal(idx[off:off + ref["count"]], ref["indices"]), v
        assert np.array_equal(args[v], ref["args"]), v
        if b.words is not None:
            check_words(b.words[v].cpu().numpy().view(np.uint32)[:n], ref["words"], v)


def check_words(got, ref, v):
    """The batched kernel inlines the same float code as the plain one in
    another kernel; contraction may still differ in the last bit, so the
    parity bar is the word codecs' one code (angle modulo its wrap, the rgb
    word by its decoded colour), and the visible set is exact."""
    same = np.all(got == ref, axis=1)
    print(f"view {v}: {np.count_nonzero(~same)} of {len(same)} words differ", flush=True)
    assert np.array_equal(got[:, 3] != 0, ref[:, 3] != 0), v
    gf, rf = _fields(got), _fields(ref)
    for k in gf:
        d = np.abs(gf[k] - rf[k])
        if k == "angle":
            d = np.minimum(d, 4096 - d)
        assert d.max(initial=0) <= 1, (v, k)
    diff = np.nonzero(got[:, 2] != ref[:, 2])[0][:2000]
    if diff.size:
        gc, rc = _decode_rgb(got[diff]), _decode_rgb(ref[diff])
        # one step of the shared-exponent code: at most max|c| / 127
        step = np.maximum(np.abs(gc).max(axis=1), np.abs(rc).max(axis=1))[:, None] / 127.0
        assert np.all(np.abs(gc - rc) <= step + 1e-12), v


def small_batch(seed=3, collections=3, views=7, scale=0.004):
    wl = W.batched(seed=seed, collections=collections, views=views, scale=scale, width=640, height=360)
    tables = []
    for spec in wl.specs:
        planes = W.generate(spec)
        tables.append(DeviceTable.from_f32_planes(planes, spec.count, 3))
    return wl, tables


def test_batch_equals_plain_per_view(cuda):
    wl, tables = small_batch()
    prm = Params(sort_bits=32)
    views = list(zip(wl.view_table, wl.cameras))
    b = Batch(tables, views, prm)
    print("counts", [t.count for t in tables], "views", wl.view_table, flush=True)
    b.run(graph=False)
    check_views(b, tables, prm)


def test_batch_graph_replays_match_eager(cuda):
    wl, tables = small_batch(seed=5, collections=4, views=9)
    prm = Params(sort_bits=24)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm)
    b.run(graph=False)
    torch.cuda.synchronize()
    eager = (b.keys.clone(), b.indices.clone(), b.counts.clone(), b.args.clone())
    x = torch.randn(1 << 20, device="cuda")
    for r in range(3):
        b.indices.zero_()
        b.counts.zero_()
        b.keys.zero_()
        _ = (x * 2.0).sum()  # other GPU work between replays
        torch.cuda.synchronize()
        b.run(graph=True)
        torch.cuda.synchronize()
        print(f"replay {r}", flush=True)
        assert torch.equal(b.keys, eager[0])
        assert torch.equal(b.counts, eager[2])
        assert torch.equal(b.args, eager[3])
        for v in range(len(b.views)):
            off, _ = b.segment(v)
            c = int(eager[2][v])
            assert torch.equal(b.indices[off:off + c], eager[1][off:off + c])


def test_batch_new_cameras_through_the_graph(cuda):
    wl, tables = small_batch(seed=6, collections=2, views=5)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm)
    b.run(graph=True)
    moved = [W.orbit_camera(900 + v, 5.0 + v, 640, 360) for v in range(5)]
    b.set_cameras(moved)
    b.run(graph=True)
    check_views(b, tables, prm, views=list(zip(wl.view_table, moved)))


def test_batch_mixed_layouts_and_many_views_of_one_table(cuda):
    # f32 SH3, f32 SH0 and packed SH3 tables: three projection launches; 11
    # views of one table: two groups of that table.
    specs = [W.plain_4m_spec(seed=41, count=9000), W.plain_small_spec(seed=42), W.stereo_16m_spec(seed=43, count=5000)]
    specs[1].count = 7000
    p0 = W.generate(specs[0])
    p1 = W.generate(specs[1])
    p2 = W.generate(specs[2])
    words, chans = W.pack(p2, 3)
    tables = [DeviceTable.from_f32_planes(p0, specs[0].count, 3),
              DeviceTable.from_f32_planes(p1, specs[1].count, 0),
              DeviceTable.from_packed(words, specs[2].count, (0.0, 1.0, -10.0, 2.0), chans, 3)]
    views = [(0, W.orbit_camera(100 + v, 6.0 + 0.3 * v, 320, 240)) for v in range(11)]
    views += [(1, W.plain_small_camera()), (2, W.orbit_camera(7, 7.0, 400, 300)), (0, W.orbit_camera(8, 9.0, 400, 300))]
    prm = Params(sort_bits=32)
    b = Batch(tables, views, prm)
    b.run(graph=True)
    check_views(b, tables, prm)


def test_batch_with_cov_and_ext_tables(cuda):
    # covSplats (the cov form) and extSplats tables in one batch, ext and packed SH bundles
    from tests.test_project_gpu import cov_table, ext_table

    ca, cb, cch = cov_table(6000, 44, 3, lod=True)
    ea, eb, ech = ext_table(5000, 45, 2)
    tables = [DeviceTable.from_cov(ca, cb, 6000, cch, sh_codec=3, sh_degree=3),
              DeviceTable.from_ext(ea, eb, 5000, ech, sh_codec=2, sh_degree=3)]
    views = [(0, W.orbit_camera(60 + v, 6.5, 320, 240)) for v in range(3)] + [(1, W.orbit_camera(70, 7.0, 400, 300))]
    prm = Params(sort_bits=32)
    b = Batch(tables, views, prm)
    b.run(graph=False)
    check_views(b, tables, prm)


def test_batch_tiny_and_empty_segments(cuda):
    spec = W.plain_small_spec(seed=44)
    spec.count = 1
    p = W.generate(spec)
    t1 = DeviceTable.from_f32_planes(p, 1, 0)
    spec2 = W.plain_small_spec(seed=45)
    spec2.count = 4097
    t2 = DeviceTable.from_f32_planes(W.generate(spec2), 4097, 0)
    behind = W.plain_small_camera()
    views = [(0, behind), (1, behind), (0, W.orbit_camera(3, 50.0, 64, 64))]
    prm = Params(sort_bits=16)
    b = Batch([t1, t2], views, prm)
    b.run(graph=False)
    check_views(b, [t1, t2], prm)


def test_batch_segments_sort_like_the_oracle(cuda):
    # Each segment's order is SortKeys.sort of its keys (the oracle per segment).
    wl, tables = small_batch(seed=8, collections=2, views=4, scale=0.01)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm, words=False)
    b.run(graph=True)
    torch.cuda.synchronize()
    keys = b.keys.cpu().numpy().view(np.uint32)
    idx = b.indices.cpu().numpy().view(np.uint32)
    for v in range(4):
        off, n = b.segment(v)
        ordering, visible = O.sort(keys[off:off + n], 32)
        assert int(b.counts[v]) == visible
        assert np.array_equal(idx[off:off + visible], ordering[:visible])


@pytest.mark.parametrize("tuning", [(6, 16, 256), (4, 8, 512), (8, 12, 256), (8, 20, 256)])
def test_batch_under_other_tunings_equals_plain(cuda, tuning):
    # every tuning the autotuner may pick gives the plain path's order
    wl, tables = small_batch(seed=9, collections=3, views=6)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm, tuning=tuning)
    assert b.tuning == tuning and b.tuning_source == "given"
    b.run(graph=True)
    check_views(b, tables, prm)


def test_batched_workload_at_its_own_shape_autotuned(cuda):
    # 8 collections (log-uniform 1e4..6e6 splats, SH3), 32 views at 1280x720:
    # the tuning from the autotuner (then from its cache), graph replays; every
    # segment's indices, count and draw arguments equal SortKeys.sort of its keys
    wl = W.batched()
    tables = []
    for spec in wl.specs:
        print(f"generating {spec.count} splats", flush=True)
        tables.append(DeviceTable.from_f32_planes(W.generate(spec), spec.count, 3))
    prm = Params(sort_bits=32)
    views = list(zip(wl.view_table, wl.cameras))
    b = Batch(tables, views, prm, words=False, tuning="auto", tune_reps=3)
    print(f"tuning {b.tuning} ({b.tuning_source}, {b.tuning_ms:.3f} ms per run)", flush=True)
    assert b.tuning_source == "autotuned" and b.tuning_ms > 0
    again = Batch(tables, views, prm, words=False, tuning="auto")
    assert again.tuning_source == "cache" and again.tuning == b.tuning
    del again
    assert b.total == sum(tables[t].count for t in wl.view_table) > 10_000_000
    for r in range(2):
        b.indices.fill_(-1)
        b.counts.zero_()
        b.run(graph=True)
        torch.cuda.synchronize()
        keys = b.keys.cpu().numpy().view(np.uint32)
        idx = b.indices.cpu().numpy().view(np.uint32)
        counts = b.counts.cpu().numpy()
        args = b.args.cpu().numpy().view(np.uint32).reshape(-1, 5)
        for v in range(len(views)):
            off, n = b.segment(v)
            ordering, visible = O.sort(keys[off:off + n], 32)
            assert counts[v] == visible, (r, v)
            assert list(args[v]) == [6144, -(-visible // 1024), 0, 0, 0], (r, v)
            assert np.array_equal(idx[off:off + visible], ordering[:visible]), (r, v)
        print(f"replay {r}: {len(views)} segments, {int(counts.sum())} visible of {b.total}", flush=True)

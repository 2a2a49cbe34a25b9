"""The batched form's persistent launch (RDX_BATCH_PERSISTENT): the whole
batch, projection and every radix pass of every segment, in one launch with
grid-wide barriers between the phases. Its output must equal the launch
sequence's bit for bit (keys, indices, counts, draw arguments), each view
must equal the plain path on that view alone, and each segment's order must
be the oracle's SortKeys.sort of its keys. Graph replays are checked against
the eager run with other GPU work between them, and every persistent run is
followed by the batch's status check (a barrier wait that ran out raises)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import Batch, DeviceTable, Params, workloads as W
from radix_amd import camera as cam
from radix_amd._native import RadixError
from tests.test_batch_gpu import check_views, small_batch
from tests.test_ranged_sort_gpu import INVALID, SPAN, Rng, depth_table

pytestmark = pytest.mark.gpu

PERSIST = (8, 12, 512)


def outputs(b: Batch):
    torch.cuda.synchronize()
    return (b.keys.cpu().numpy().view(np.uint32).copy(), b.indices.cpu().numpy().view(np.uint32).copy(),
            b.counts.cpu().numpy().copy(), b.args.cpu().numpy().view(np.uint32).copy())


def assert_same(b: Batch, got, ref, tag):
    assert np.array_equal(got[0], ref[0]), f"{tag}: keys"
    assert np.array_equal(got[2], ref[2]), f"{tag}: counts"
    assert np.array_equal(got[3], ref[3]), f"{tag}: draw args"
    for v in range(len(b.views)):
        off, _ = b.segment(v)
        c = int(ref[2][v])
        assert np.array_equal(got[1][off:off + c], ref[1][off:off + c]), f"{tag}: view {v} indices"


def clear(b: Batch):
    b.keys.zero_()
    b.indices.fill_(-1)
    b.counts.zero_()
    b.args.zero_()


def test_persistent_equals_sequence_and_plain(cuda):
    wl, tables = small_batch(seed=11, collections=3, views=7)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm, tuning=PERSIST)
    print(f"persistent grid {b.persistent_grid}, total {b.total}", flush=True)
    assert b.persistent_grid >= 8 and b.persistent_grid % 8 == 0
    b.run(graph=False, persistent=False)
    seq = outputs(b)
    clear(b)
    b.run(graph=False, persistent=True)
    b.status()
    assert_same(b, outputs(b), seq, "eager persistent")
    check_views(b, tables, prm)


def test_persistent_graph_replays(cuda):
    wl, tables = small_batch(seed=12, collections=4, views=9)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm, tuning=PERSIST)
    b.run(graph=False, persistent=False)
    seq = outputs(b)
    x = torch.randn(1 << 20, device="cuda")
    for r in range(3):
        clear(b)
        _ = (x * 2.0).sum()  # other GPU work between replays
        torch.cuda.synchronize()
        b.run(graph=True, persistent=True)
        b.status()
        print(f"replay {r}", flush=True)
        assert_same(b, outputs(b), seq, f"replay {r}")
    # the sequence's graph and the persistent graph alternate on one batch
    clear(b)
    b.run(graph=True, persistent=False)
    assert_same(b, outputs(b), seq, "sequence graph")
    clear(b)
    b.run(graph=True, persistent=True)
    b.status()
    assert_same(b, outputs(b), seq, "persistent graph again")


def test_persistent_new_cameras(cuda):
    wl, tables = small_batch(seed=13, collections=2, views=5)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm, tuning=PERSIST)
    b.run(graph=True, persistent=True)
    b.status()
    moved = [W.orbit_camera(910 + v, 5.0 + v, 640, 360) for v in range(5)]
    b.set_cameras(moved)
    b.run(graph=True, persistent=True)
    b.status()
    check_views(b, tables, prm, views=list(zip(wl.view_table, moved)))


def test_persistent_three_and_four_pass_segments(cuda):
    # one segment within 2^27 of its base (three passes), one beyond (the gated fourth)
    rng = Rng(35)
    narrow = depth_table(1.0 + 49.0 * rng.random(40_000), rng)
    wide = depth_table(np.exp(np.log(0.012) + (np.log(990.0) - np.log(0.012)) * rng.random(60_000)), rng)
    rig = cam.mono(np.eye(4), 60.0, 640, 360)
    prm = Params(sort_bits=32)
    b = Batch([narrow, wide], [(0, rig), (1, rig), (0, rig)], prm, tuning=PERSIST)
    b.run(graph=False, persistent=False)
    seq = outputs(b)
    clear(b)
    b.run(graph=False, persistent=True)
    b.status()
    got = outputs(b)
    assert_same(b, got, seq, "persistent")
    spans = []
    for v in range(3):
        off, n = b.segment(v)
        k = got[0][off:off + n]
        ordering, visible = O.sort(k, 32)
        valid = k[k != INVALID]
        spans.append(int(valid.max()) - (int(valid.min()) & ~511))
        print(f"segment {v}: {n} keys, {visible} kept, span 2^{np.log2(spans[-1]):.2f}", flush=True)
        assert got[2][v] == visible
        assert np.array_equal(got[1][off:off + visible], ordering[:visible]), v
    assert spans[0] < SPAN and spans[1] >= SPAN and spans[2] < SPAN


def test_persistent_tiny_and_empty_segments(cuda):
    spec = W.plain_small_spec(seed=46)
    spec.count = 1
    t1 = DeviceTable.from_f32_planes(W.generate(spec), 1, 0)
    spec2 = W.plain_small_spec(seed=47)
    spec2.count = 6145
    t2 = DeviceTable.from_f32_planes(W.generate(spec2), 6145, 0)
    behind = W.plain_small_camera()
    views = [(0, behind), (1, behind), (0, W.orbit_camera(3, 50.0, 64, 64))]
    prm = Params(sort_bits=32)
    b = Batch([t1, t2], views, prm, tuning=PERSIST)
    b.run(graph=False, persistent=True)
    b.status()
    check_views(b, [t1, t2], prm)


def test_persistent_word_tables(cuda):
    # packed and covSplats tables (their codecs in the persistent projection)
    from tests.test_project_gpu import cov_table

    spec = W.stereo_16m_spec(seed=48, count=9000)
    words, chans = W.pack(W.generate(spec), 3)
    packed = DeviceTable.from_packed(words, spec.count, (0.0, 1.0, -10.0, 2.0), chans, 3)
    ca, cb, cch = cov_table(7000, 49, 3, lod=True)
    cov = DeviceTable.from_cov(ca, cb, 7000, cch, sh_codec=3, sh_degree=3)
    prm = Params(sort_bits=32)
    for tables, views in (([packed], [(0, W.orbit_camera(5 + v, 6.0 + v, 320, 240)) for v in range(3)]),
                          ([cov], [(0, W.orbit_camera(15 + v, 6.5, 320, 240)) for v in range(2)])):
        b = Batch(tables, views, prm, tuning=PERSIST)
        assert b.persistent_grid > 0
        b.run(graph=False, persistent=False)
        seq = outputs(b)
        clear(b)
        b.run(graph=True, persistent=True)
        b.status()
        assert_same(b, outputs(b), seq, "word table")
        check_views(b, tables, prm)


def test_persistent_form_not_offered(cuda):
    # other tunings, key widths and mixed kinds keep the launch sequence: the flag is refused
    wl, tables = small_batch(seed=14, collections=2, views=3)
    prm = Params(sort_bits=32)
    b = Batch(tables, list(zip(wl.view_table, wl.cameras)), prm, tuning=(8, 16, 512))
    assert b.persistent_grid == 0
    with pytest.raises(RadixError, match="persistent form"):
        b.run(graph=False, persistent=True)
    b24 = Batch(tables, list(zip(wl.view_table, wl.cameras)), Params(sort_bits=24), tuning=PERSIST)
    assert b24.persistent_grid == 0
    spec = W.stereo_16m_spec(seed=48, count=3000)
    words, chans = W.pack(W.generate(spec), 3)
    packed = DeviceTable.from_packed(words, spec.count, (0.0, 1.0, -10.0, 2.0), chans, 3)
    bm = Batch([tables[0], packed], [(0, wl.cameras[0]), (1, W.orbit_camera(5, 7.0, 320, 240))], prm, tuning=PERSIST)
    assert bm.persistent_grid == 0
    with pytest.raises(RadixError, match="persistent form"):
        bm.run(graph=False, persistent=True)


def test_batched_workload_persistent_at_its_own_shape(cuda):
    # 8 collections (log-uniform 1e4..6e6 splats, SH3), 32 views at 1280x720
    # in one persistent launch, replayed from a graph: each segment equals
    # SortKeys.sort of its keys, and the whole output equals the sequence's
    wl = W.batched()
    tables = []
    for spec in wl.specs:
        print(f"generating {spec.count} splats", flush=True)
        tables.append(DeviceTable.from_f32_planes(W.generate(spec), spec.count, 3))
    prm = Params(sort_bits=32)
    views = list(zip(wl.view_table, wl.cameras))
    b = Batch(tables, views, prm, words=False, tuning=PERSIST)
    print(f"persistent grid {b.persistent_grid}, {b.total} splat-views", flush=True)
    b.run(graph=True, persistent=False)
    seq = outputs(b)
    for r in range(2):
        clear(b)
        b.run(graph=True, persistent=True)
        b.status()
        got = outputs(b)
        assert_same(b, got, seq, f"replay {r}")
        for v in range(len(views)):
            off, n = b.segment(v)
            ordering, visible = O.sort(got[0][off:off + n], 32)
            assert got[2][v] == visible, (r, v)
            assert np.array_equal(got[1][off:off + visible], ordering[:visible]), (r, v)
        print(f"replay {r}: {len(views)} segments, {int(got[2].sum())} visible of {b.total}", flush=True)
    form, ms_seq, ms_persist = b.choose_form(reps=3)
    print(f"sequence {ms_seq:.3f} ms, persistent {ms_persist:.3f} ms: persistent={form}", flush=True)
    assert ms_seq > 0 and ms_persist > 0

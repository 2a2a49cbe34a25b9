"""Compacting 32-bit sorts under the default tuning sort key - base by 9-bit
digits: three passes when the kept keys lie within 2^27 of the base (the
smallest kept key rounded down to a multiple of 512), a fourth pass for
segments that reach further. Every case is held bit-exact to the oracle's
SortKeys.sort (src/core/sort.ts:64-163) of the same keys: both sides of the
2^27 boundary, narrow and wide random keys, and one batch whose two segments
take different numbers of passes."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import Batch, DeviceTable, Params, Projector, SplatSorter
from radix_amd import camera as cam
from tests.helpers import Rng, first_mismatch

pytestmark = pytest.mark.gpu
INVALID = 0xFFFFFFFF
SPAN = 1 << 27


def dev_u32(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def host_u32(t):
    return t.cpu().numpy().view(np.uint32)


def sort_and_check(keys):
    n = keys.size
    sorter = SplatSorter(n)
    sorter.sort(dev_u32(keys), n, 32)
    torch.cuda.synchronize()
    ordering, visible = O.sort(keys, 32)
    assert sorter.read_count() == visible
    assert list(host_u32(sorter.indirect)) == [6 * 1024, -(-visible // 1024), 0, 0, 0]
    assert first_mismatch(host_u32(sorter.values)[:visible], ordering[:visible]) is None


def with_invalid(keys, rng, frac=0.25):
    keys = keys.astype(np.uint32)
    keys[rng.random(keys.size) < frac] = INVALID
    return keys


@pytest.mark.parametrize("n", [100_003, 1 << 20])
def test_distance_keys_three_passes(cuda, n):
    # distances within a factor 20: 2^27 above the base is never reached
    rng = Rng(31)
    keys = O.keys_compute(2.0 + 38.0 * rng.random(n), 32)
    keys = with_invalid(keys, rng)
    valid = keys[keys != INVALID]
    assert int(valid.max()) - (int(valid.min()) & ~511) < SPAN
    sort_and_check(keys)


@pytest.mark.parametrize("frac", [0.97, 0.9995])
def test_mostly_culled_keys(cuda, frac):
    # A view that culls most splats: after the dropping pass the kept keys
    # fill only the first tiles, which the later passes' scatter spreads over
    # every XCD (and the tiles past them return). Three and four passes.
    rng = Rng(37)
    n = 3_000_017
    for spread in (38.0, 1e6):
        keys = with_invalid(O.keys_compute(2.0 + spread * rng.random(n), 32), rng, frac)
        sort_and_check(keys)


@pytest.mark.parametrize("reach", [SPAN - 1, SPAN])
def test_range_boundary(cuda, reach):
    # the smallest kept key 0x3F800123 (base 0x3F800000); the largest at
    # base + 2^27 - 1 (three passes) or base + 2^27 (the fourth pass)
    rng = Rng(32)
    lo, base = 0x3F800123, 0x3F800000
    n = 70_001
    keys = (lo + np.floor(rng.random(n) * (base + reach - lo))).astype(np.uint64)
    keys[5] = lo
    keys[n // 2] = base + reach
    keys = with_invalid(keys.astype(np.uint32), rng, 0.1)
    keys[5], keys[n // 2] = lo, base + reach
    sort_and_check(keys)


def test_wide_random_keys(cuda):
    # below 0xFFFF0000: WgpuCompact drops only the invalid key, SortKeys.sort
    # also hides keys whose top 16 bits are set (sort.ts:154)
    rng = Rng(33)
    keys = np.floor(rng.random(300_000) * float(0xFFFF0000)).astype(np.uint64).astype(np.uint32)
    sort_and_check(with_invalid(keys, rng, 0.05))


def test_all_kept_keys_equal(cuda):
    keys = np.full(50_000, 0x40123456, dtype=np.uint32)
    keys[::7] = INVALID
    sort_and_check(keys)


def depth_table(distances, rng):
    """SH0 f32 splats straight ahead of a camera at the origin looking down -z,
    sized to stay visible (about 5 pixels) at every distance."""
    n = distances.size
    planes = np.zeros((14, n), dtype=np.float32)
    spread = 0.3 * distances
    planes[0] = (rng.random(n) - 0.5) * spread
    planes[1] = (rng.random(n) - 0.5) * spread
    planes[2] = -distances
    planes[3:6] = np.log(0.01 * distances)
    planes[9] = 1.0  # quaternion (0, 0, 0, 1)
    planes[10] = 0.8
    planes[11:14] = 0.5
    return DeviceTable.from_f32_planes(planes, n, 0)


def test_batch_segments_take_three_and_four_passes(cuda):
    rng = Rng(34)
    narrow = depth_table(1.0 + 49.0 * rng.random(40_000), rng)
    # 0.012 .. 990: a factor above 2^16, so its keys reach beyond 2^27
    wide = depth_table(np.exp(np.log(0.012) + (np.log(990.0) - np.log(0.012)) * rng.random(60_000)), rng)
    rig = cam.mono(np.eye(4), 60.0, 640, 360)
    prm = Params(sort_bits=32)
    b = Batch([narrow, wide], [(0, rig), (1, rig), (0, rig)], prm)
    b.run(graph=False)
    torch.cuda.synchronize()
    keys = host_u32(b.keys)
    idx = host_u32(b.indices)
    counts = b.counts.cpu().numpy()
    spans = []
    for v in range(3):
        off, n = b.segment(v)
        k = keys[off:off + n]
        ordering, visible = O.sort(k, 32)
        valid = k[k != INVALID]
        spans.append(int(valid.max()) - (int(valid.min()) & ~511))
        print(f"segment {v}: {n} keys, {visible} kept, span above the base 2^{np.log2(spans[-1]):.2f}", flush=True)
        assert counts[v] == visible, v
        assert first_mismatch(idx[off:off + visible], ordering[:visible]) is None, v
    assert spans[0] < SPAN and spans[1] >= SPAN and spans[2] < SPAN
    # the single-segment path on the wide table
    out = Projector().project_sort(wide, rig, prm)
    torch.cuda.synchronize()
    k = host_u32(out["keys"])[:wide.count]
    ordering, visible = O.sort(k, 32)
    assert int(out["count"].cpu().item()) == visible
    assert first_mismatch(host_u32(out["indices"])[:visible], ordering[:visible]) is None

"""GPU radix sort parity: bit-exact against the oracle's SortKeys.sort
(the reference's test shapes, renderTests.ts:31-89) and size-independent
properties at full size."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import RadixSort, SortKeys, SplatSorter
from radix_amd.sort import Context
from tests.helpers import Rng, first_mismatch

pytestmark = pytest.mark.gpu
BITS = [16, 24, 32]
COUNTS = [0, 1, 4095, 4096, 4097, 70000]
TUNINGS = [(8, 16, 512), (8, 16, 256), (8, 8, 256), (8, 12, 512), (6, 16, 256), (4, 8, 256), (8, 20, 256), (6, 8, 512)]


def dev_u32(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def host_u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("tuning", TUNINGS)
def test_radix_sort_equals_sortkeys(cuda, tuning):
    random = Rng(7)
    ctx = Context(tuning)
    for bits in BITS:
        for count in COUNTS:
            print(f"sort_pairs tuning={tuning} bits={bits} count={count}", flush=True)
            keys = np.floor(random.random(max(count, 1)) * 2.0 ** bits).astype(np.uint64).astype(np.uint32)
            kd = dev_u32(keys)
            vd = dev_u32(np.arange(max(count, 1), dtype=np.uint32))
            RadixSort(kd, vd, context=ctx, digit_bits=tuning[0], keys_per_thread=tuning[1],
                      block_threads=tuning[2]).sort(count, bits=bits)
            torch.cuda.synchronize()
            expected, _ = O.sort(keys[:count], bits)
            got = host_u32(vd)[:count]
            assert first_mismatch(got, expected) is None, (tuning, bits, count)
            assert first_mismatch(host_u32(kd)[:count], keys[:count][expected]) is None


def test_sort_indirect_reads_count_on_device(cuda):
    random = Rng(3)
    keys = np.floor(random.random(50000) * 2.0 ** 24).astype(np.uint32)
    for count in (0, 1, 777, 50000):
        kd, vd = dev_u32(keys), dev_u32(np.arange(keys.size, dtype=np.uint32))
        cnt = dev_u32(np.array([count], dtype=np.uint32))
        RadixSort(kd, vd).sort_indirect(cnt, bits=24)
        torch.cuda.synchronize()
        expected, _ = O.sort(keys[:count], 24)
        assert first_mismatch(host_u32(vd)[:count], expected) is None, count
        # pairs past the count are untouched
        assert np.array_equal(host_u32(vd)[count:], np.arange(count, keys.size, dtype=np.uint32))


@pytest.mark.parametrize("bits", BITS)
def test_splat_sorter_compacts_and_sorts(cuda, bits):
    # renderTests.ts:59-89: 30% invalid keys, the rest below 0xff * 2^(bits-8)
    random = Rng(11)
    count = 20000
    r = random.random(2 * count)
    keys = np.where(r[0::2] < 0.3, O.key_invalid(bits), np.floor(r[1::2] * (0xFF * 2.0 ** (bits - 8)))).astype(np.uint32)
    sorter = SplatSorter(count)
    sorter.sort(dev_u32(keys), count, bits)
    torch.cuda.synchronize()
    ordering, visible = O.sort(keys, bits)
    assert sorter.read_count() == visible
    assert list(host_u32(sorter.indirect)) == [6 * 1024, -(-visible // 1024), 0, 0, 0]
    assert first_mismatch(host_u32(sorter.values)[:visible], ordering[:visible]) is None


def test_splat_sorter_all_invalid_and_empty(cuda):
    sorter = SplatSorter(100)
    sorter.sort(dev_u32(np.full(100, 0xFFFFFF, dtype=np.uint32)), 100, 24)
    torch.cuda.synchronize()
    assert sorter.read_count() == 0
    assert list(host_u32(sorter.indirect)) == [6144, 0, 0, 0, 0]
    sorter.sort(dev_u32(np.zeros(1, dtype=np.uint32)), 0, 24)
    torch.cuda.synchronize()
    assert sorter.read_count() == 0


@pytest.mark.parametrize("bits", BITS)
def test_sortkeys_sort_host_arrays(cuda, bits):
    random = Rng(9)
    keys = np.where(random.random(50000) < 0.1, O.key_invalid(bits),
                    np.floor(random.random(50000) * (0xFF * 2.0 ** (bits - 8) - 1))).astype(np.uint32)
    ordering, visible = SortKeys.sort(keys, bits=bits)
    exp, exp_visible = O.sort(keys, bits)
    assert ordering.size == 53248  # alignCapacity(50000)
    assert visible == exp_visible
    assert first_mismatch(ordering[:keys.size], exp) is None


def test_sortkeys_sort_visible_rule_on_arbitrary_keys(cuda):
    random = Rng(21)
    for bits in BITS:
        keys = np.floor(random.random(30000) * 2.0 ** bits).astype(np.uint64).astype(np.uint32)
        keys[::50] = O.key_invalid(bits) - np.arange(keys[::50].size, dtype=np.uint32) % 7
        ordering, visible = SortKeys.sort(keys, bits=bits)
        exp, exp_visible = O.sort(keys, bits)
        assert visible == exp_visible, bits
        assert first_mismatch(ordering[:keys.size], exp) is None


def test_full_size_properties_4m(cuda):
    # 4,194,304 keys: a permutation, keys non-decreasing, ties in index order,
    # the same as numpy's stable argsort, and identical on a repeat.
    n = 1 << 22
    random = Rng(5)
    keys = np.floor(random.random(n) * 2.0 ** 32).astype(np.uint64).astype(np.uint32)
    keys[: n // 3] &= np.uint32(0xFFF000FF)  # many ties
    results = []
    for _ in range(2):
        kd, vd = dev_u32(keys), dev_u32(np.arange(n, dtype=np.uint32))
        RadixSort(kd, vd).sort(n, bits=32)
        torch.cuda.synchronize()
        results.append(host_u32(vd).copy())
    assert np.array_equal(results[0], results[1])
    assert first_mismatch(results[0], np.argsort(keys, kind="stable").astype(np.uint32)) is None


@pytest.mark.parametrize("tuning,n", [((8, 16, 256), 1 << 22),          # 1024 tiles: two scan rounds of 512
                                      ((8, 8, 256), (1 << 22) + 4097),  # 2051 tiles: 32 per thread, two rounds
                                      ((4, 8, 256), 1 << 22)])          # 2048 tiles, 8 passes
def test_many_tiles_equal_stable_argsort(cuda, tuning, n):
    # the tile-count scan over more tiles than one round covers
    random = Rng(11)
    keys = np.floor(random.random(n) * 2.0 ** 32).astype(np.uint64).astype(np.uint32)
    keys[::3] &= np.uint32(0xFF00FF00)
    ctx = Context(tuning)
    kd, vd = dev_u32(keys), dev_u32(np.arange(n, dtype=np.uint32))
    RadixSort(kd, vd, context=ctx).sort(n, bits=32)
    torch.cuda.synchronize()
    assert first_mismatch(host_u32(vd), np.argsort(keys, kind="stable").astype(np.uint32)) is None, tuning


def test_compute_keys_device_matches_host(cuda):
    f = np.arange(0, 0xFFFFFFFF, 65537, dtype=np.uint64).astype(np.uint32)
    d = torch.from_numpy(f.view(np.float32).copy()).cuda()
    for bits in BITS:
        k = host_u32(SortKeys.compute_device(d, bits))
        exp = np.array([O.key_compute(float(x), bits) for x in f.view(np.float32)], dtype=np.uint32)
        assert first_mismatch(k, exp) is None, bits


@pytest.mark.parametrize("tuning", [(8, 16, 256), (6, 12, 512), (4, 8, 256)])
def test_skewed_digits_equal_sortkeys(cuda, tuning):
    # Rounds whose keys share one, two or a few digits (the ballot-ranked
    # digits) next to rounds of many digits (the LDS-mask ranked ones).
    random = Rng(11)
    n = 150000
    dist = 1.0 + 15.0 * random.random(n)
    cases = {
        "distances": (~dist.astype(np.float32).view(np.uint32)).astype(np.uint32),
        "equal": np.full(n, 0x3C00_1234, dtype=np.uint32),
        "two": np.where(random.random(n) < 0.5, 7, 0xBEEF_0000).astype(np.uint32),
        "three_runs": np.repeat(np.array([5, 9, 5, 300, 9], dtype=np.uint32), n // 5),
    }
    ctx = Context(tuning)
    for name, keys in cases.items():
        print(f"skewed {name} tuning={tuning}", flush=True)
        kd = dev_u32(keys)
        vd = dev_u32(np.arange(keys.size, dtype=np.uint32))
        RadixSort(kd, vd, context=ctx, digit_bits=tuning[0], keys_per_thread=tuning[1],
                  block_threads=tuning[2]).sort(keys.size, bits=32)
        torch.cuda.synchronize()
        expected, _ = O.sort(keys, 32)
        assert first_mismatch(host_u32(vd), expected) is None, name


def test_autotune_picks_a_supported_tuning(cuda):
    ctx = Context()
    chosen, ms = ctx.autotune(1 << 18, bits=24, reps=2)
    assert chosen in Context.candidates()
    assert ms > 0.0
    assert ctx.tuning() == chosen
    random = Rng(5)
    keys = np.floor(random.random(1 << 18) * 2.0 ** 24).astype(np.uint32)
    kd, vd = dev_u32(keys), dev_u32(np.arange(keys.size, dtype=np.uint32))
    RadixSort(kd, vd, context=ctx).sort(keys.size, bits=24)
    torch.cuda.synchronize()
    expected, _ = O.sort(keys, 24)
    assert first_mismatch(host_u32(vd), expected) is None

"""The oracle's sort keys and CPU sort, pinned (no GPU).

Pins: the reference's known answers (scripts/check-core.ts:1194-1232), the
sort.ts header's closed forms, brute force (a stable sort is the unique
permutation that orders keys and keeps ties in index order: numpy's stable
argsort), and the visible-count rule of sort.ts:104/128/154.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import Rng, first_mismatch

BITS = [16, 24, 32]


def mulberry32_scalar(seed):
    """test/gpu/renderUtils.ts:23-30, one number at a time (Math.imul semantics)."""
    state = seed & 0xFFFFFFFF

    def imul(a, b):
        return (a * b) & 0xFFFFFFFF

    def nxt():
        nonlocal state
        state = (state + 0x6D2B79F5) & 0xFFFFFFFF
        t = imul(state ^ (state >> 15), 1 | state)
        t = ((t + imul(t ^ (t >> 7), 61 | t)) & 0xFFFFFFFF) ^ t
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296

    return nxt


def test_rng_matches_scalar_mulberry32():
    for seed in (7, 11, 12345, 0xFFFFFFFF):
        r = Rng(seed)
        s = mulberry32_scalar(seed)
        a = np.concatenate([r.random(5), r.random(100)])
        b = np.array([s() for _ in range(105)])
        assert np.array_equal(a, b), seed


def test_key_of_one_at_32_bits():
    # check-core.ts:1208: compute(1, 32) == (0xffffffff ^ 0x3f800000) >>> 0
    assert O.key_compute(1.0, 32) == 0xC07FFFFF


def test_invalid_and_valid_max():
    assert [O.key_invalid(b) for b in BITS] == [0xFFFF, 0xFFFFFF, 0xFFFFFFFF]
    # 0xff * 2^(bits - 8) - 1: the top digit all ones is reserved
    assert [O.key_valid_max(b) for b in BITS] == [0xFEFF, 0xFEFFFF, 0xFEFFFFFF]


def test_keys_monotone_back_to_front_and_nonfinite_invalid():
    # check-core.ts:1194-1208
    distances = [0, 1e-12, 1e-6, 0.001, 0.5, 1, 1.0001, 2, 100, 1e6, 1e12, 3e38]
    for bits in BITS:
        keys = [O.key_compute(d, bits) for d in distances]
        assert all(keys[i] <= keys[i - 1] for i in range(1, len(keys))), bits
        assert all(k <= O.key_valid_max(bits) for k in keys)
        for d in (math.inf, math.nan, -math.inf):
            assert O.key_compute(d, bits) == O.key_invalid(bits)


def test_keys_closed_forms_over_float_patterns():
    # sort.ts:5-11: 32-bit q = f, 24-bit q = f >>> 7, 16-bit keeps 6 exponent
    # bits (2^-32..2^32, clamped) and 10 mantissa bits; key = min(mask ^ q, validMax).
    f = np.arange(0, 0x7F800000, 4093, dtype=np.uint32)
    d = f.view(np.float32).astype(np.float64)
    for bits in BITS:
        k = O.keys_compute(d, bits)
        assert np.all(np.diff(k.astype(np.int64)) <= 0), bits
        assert k.max() <= O.key_valid_max(bits)
    k32 = O.keys_compute(d, 32)
    assert np.array_equal(k32, np.minimum(~f, np.uint32(0xFEFFFFFF)))
    k24 = O.keys_compute(d, 24)
    assert np.array_equal(k24, np.minimum(np.uint32(0xFFFFFF) ^ (f >> np.uint32(7)), np.uint32(0xFEFFFF)))
    # 16 bits: distinct codes only inside [2^-32, 2^32)
    k16 = O.keys_compute(np.array([2.0 ** -40, 2.0 ** -32, 2.0 ** 32, 2.0 ** 40]), 16)
    assert k16[0] == k16[1] and k16[2] == k16[3] and k16[1] > k16[2]
    assert O.key_compute(2.0 ** -32, 16) == 0xFEFF  # q = 0: mask ^ 0 clamped to validMax
    assert O.key_compute(2.0 ** 32, 16) == 0xFFFF ^ 0xFFFF  # q = 0xffff


def test_negative_distance_is_invalid():
    # sort.ts:50-52: the sign bit puts f above the infinity pattern
    for bits in BITS:
        assert O.key_compute(-1.0, bits) == O.key_invalid(bits)
        assert O.key_compute(-0.0, bits) == O.key_invalid(bits)


def _brute(keys, bits):
    mask = np.uint32(O.key_invalid(bits))
    k = keys & mask if bits == 16 else keys
    return np.argsort(k, kind="stable").astype(np.uint32)


def _visible_rule(keys, bits):
    if bits == 16:
        return int(keys.size - np.count_nonzero((keys & 0xFFFF) == 0xFFFF))
    if bits == 24:
        return int(keys.size - np.count_nonzero((keys >> 11) == 0x1FFF))
    return int(keys.size - np.count_nonzero((keys >> 16) == 0xFFFF))


@pytest.mark.parametrize("bits", BITS)
def test_sort_equals_stable_brute_force(bits):
    # renderTests.ts:31-57 shapes and generator: rng(7), floor(random() * 2^bits)
    random = Rng(7)
    for count in (0, 1, 4095, 4096, 4097, 70000):
        keys = np.floor(random.random(max(count, 1)) * 2.0 ** bits).astype(np.uint64).astype(np.uint32)[:count]
        ordering, visible = O.sort(keys, bits)
        assert first_mismatch(ordering, _brute(keys, bits)) is None, (bits, count)
        assert visible == _visible_rule(keys, bits)


def test_sort_16_bits_ignores_high_bits():
    keys = np.array([0x10005, 0x00003, 0x20003, 0xFFFF0001], dtype=np.uint32)
    ordering, _ = O.sort(keys, 16)
    assert list(ordering) == [3, 1, 2, 0]


def test_sort_24_bits_rejects_keys_above_contract():
    with pytest.raises(ValueError):
        O.sort(np.array([1 << 24], dtype=np.uint32), 24)


def test_stable_same_order_every_width_invalid_last():
    # check-core.ts:1210-1232 (own stream: the reference's global seed
    # position depends on every earlier check)
    count = 20000
    r = Rng(12345).random(count)
    distances = np.round(r * 1000) / 10
    distances[::97] = np.inf
    finite = np.isfinite(distances)
    idx = np.arange(count)
    order = sorted(idx[finite], key=lambda i: (-distances[i], i))
    for bits in BITS:
        keys = O.keys_compute(distances, bits)
        ordering, visible = O.sort(keys, bits)
        assert visible == int(finite.sum())
        assert first_mismatch(ordering[:visible], np.array(order, dtype=np.uint32)) is None, bits
        assert not np.any(finite[ordering[visible:]])


def test_sort_planted_mistake_is_caught():
    # An unstable tie order (reversed within equal keys) must fail the brute-force check.
    keys = np.array([5, 3, 5, 3, 5], dtype=np.uint32)
    good, _ = O.sort(keys, 32)
    bad = np.array([3, 1, 4, 2, 0], dtype=np.uint32)
    assert first_mismatch(good, _brute(keys, 32)) is None
    assert first_mismatch(bad, _brute(keys, 32)) is not None

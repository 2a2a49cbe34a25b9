"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU calls)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from radix_amd import _native as N
from radix_amd import sort as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(rdx_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = N.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_binding():
    bound = {name for name, _, _ in N._SIGNATURES}
    assert set(declared_functions()) <= bound


def test_host_key_compute_matches_oracle():
    from oracle import oracle as O

    f = np.arange(0, 0x7F800000, 1_000_003, dtype=np.uint32).view(np.float32)
    for bits in (16, 24, 32):
        for d in list(f[:400]) + [np.inf, np.nan, -1.0]:
            assert S.SortKeys.compute(float(d), bits) == O.key_compute(float(d), bits)
        assert S.SortKeys.invalid(bits) == O.key_invalid(bits)
        assert S.SortKeys.valid_max(bits) == O.key_valid_max(bits)


def test_sort_errors_carry_the_reference_messages():
    keys = np.arange(10, dtype=np.uint32)
    with pytest.raises(N.RadixError, match=r"SortKeys.sort: ordering holds 3 of 10 indices"):
        S.SortKeys.sort(keys, bits=24, ordering=np.zeros(3, dtype=np.uint32))
    with pytest.raises(N.RadixError, match=r"bits must be 16, 24 or 32 \(got 12\)"):
        S.SortKeys.sort(keys, bits=12)


def test_tuning_validation():
    ctx = S.Context()
    assert ctx.tuning() == (8, 16, 256)
    ctx.set_tuning(6, 12, 512)
    assert ctx.tuning() == (6, 12, 512)
    with pytest.raises(N.RadixError):
        ctx.set_tuning(5, 16, 256)


def test_align_capacity():
    # TextureLayout.alignCapacity: whole rows of 4096, at least one
    assert [S.align_capacity(c) for c in (0, 1, 4096, 4097)] == [4096, 4096, 4096, 8192]

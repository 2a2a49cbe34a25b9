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
    assert ctx.tuning() == (8, 16, 512)
    ctx.set_tuning(6, 12, 512)
    assert ctx.tuning() == (6, 12, 512)
    with pytest.raises(N.RadixError):
        ctx.set_tuning(5, 16, 256)


def test_align_capacity():
    # TextureLayout.alignCapacity: whole rows of 4096, at least one
    assert [S.align_capacity(c) for c in (0, 1, 4096, 4097)] == [4096, 4096, 4096, 8192]


def _fake_f32_table(n=10):
    # A descriptor whose pointers are never dereferenced (argument checks only).
    t = N.SplatTable()
    t.layout = N.LAYOUT_F32_SOA
    t.count = n
    t.plane_stride = n
    for name in ("center", "log_scale", "quat", "opacity", "rgb"):
        setattr(t, name, 0x1000)
    t.sh_codec = N.SH_NONE
    t.object_to_world[:] = [1.0 if i % 5 == 0 else 0.0 for i in range(16)]
    return t


def test_batch_argument_errors_before_any_device_work():
    from radix_amd import Params
    from radix_amd import camera as cam

    ctx = S.Context()
    tables = (N.SplatTable * 1)(_fake_f32_table())
    views = (N.BatchView * 2)()
    views[0].table = 0
    views[0].camera = cam.mono(np.eye(4), 60, 64, 64).desc()
    views[1].table = 3
    views[1].camera = cam.mono(np.eye(4), 60, 64, 64).desc()
    h = C.c_void_p()
    lib = N.lib()
    rc = lib.rdx_batch_create(ctx.handle, tables, 1, views, 2, C.byref(Params(sort_bits=32).desc()), None, C.byref(h))
    assert rc == N.RDX_E_INVALID and b"table 3 of 1" in lib.rdx_last_error()
    views[1].table = 0
    views[1].camera = cam.stereo(np.eye(4), 0.06, 60, 64, 64).desc()
    rc = lib.rdx_batch_create(ctx.handle, tables, 1, views, 2, C.byref(Params(sort_bits=32).desc()), None, C.byref(h))
    assert rc == N.RDX_E_INVALID and b"one eye per view" in lib.rdx_last_error()
    rc = lib.rdx_batch_create(ctx.handle, tables, 1, views, 1, C.byref(Params(sort_bits=12).desc()), None, C.byref(h))
    assert rc == N.RDX_E_INVALID and b"bits must be 16, 24 or 32 (got 12)" in lib.rdx_last_error()
    assert not h.value

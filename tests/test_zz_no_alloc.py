"""rdx_sort_splats_static, the splat sorter that allocates nothing, against
rdx_sort_splats (the context's scratch) through the C ABI: equal indices,
counts and draw arguments on consecutive calls, the second smaller, at each
key width, and a clear error past its static scratch."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import _native as N
from radix_amd.sort import Context
from tests.helpers import Rng, first_mismatch

pytestmark = pytest.mark.gpu


def dev_u32(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def host_u32(t):
    return t.cpu().numpy().view(np.uint32)


def keys_for(bits, n, seed, wide):
    """Kept keys below validMax with a share of invalid ones; `wide` spreads
    32-bit keys over more than 2^27 above their smallest (the fourth pass)."""
    r = Rng(seed).random(2 * n)
    hi = 0xFF * 2.0 ** (bits - 8) - 1
    if bits == 32 and not wide:
        kept = 0x3F000000 + np.floor(r[1::2] * 2.0 ** 26)
    else:
        kept = np.floor(r[1::2] * hi)
    return np.where(r[0::2] < 0.2, O.key_invalid(bits), kept).astype(np.uint64).astype(np.uint32)


def run_static(keys_d, n, bits, idx, cnt, args):
    s = torch.cuda.current_stream().cuda_stream
    return N.lib().rdx_sort_splats_static(C.c_void_p(s), C.c_void_p(keys_d.data_ptr()), n, bits,
                                          C.c_void_p(idx.data_ptr()), C.c_void_p(cnt.data_ptr()),
                                          C.c_void_p(args.data_ptr()))


def run_context(ctx, keys_d, n, bits, idx, cnt, args):
    s = torch.cuda.current_stream().cuda_stream
    return N.lib().rdx_sort_splats(ctx.handle, C.c_void_p(s), C.c_void_p(keys_d.data_ptr()), n, bits,
                                   C.c_void_p(idx.data_ptr()), C.c_void_p(cnt.data_ptr()),
                                   C.c_void_p(args.data_ptr()))


@pytest.mark.parametrize("bits", [32, 24, 16])
def test_static_sorter_equals_context_sorter_twice(cuda, bits):
    ctx = Context()
    cap = N.STATIC_SORT_MAX_KEYS
    for call, (n, wide) in enumerate(((cap, False), (70_001, True))):
        print(f"static sorter bits={bits} call={call} n={n} wide={wide}", flush=True)
        keys = keys_for(bits, n, 100 * bits + call, wide)
        kd = dev_u32(keys)
        got = [torch.full((cap,), -1, dtype=torch.int32, device="cuda"), torch.zeros(1, dtype=torch.int32, device="cuda"),
               torch.zeros(5, dtype=torch.int32, device="cuda")]
        exp = [torch.full((cap,), -1, dtype=torch.int32, device="cuda"), torch.zeros(1, dtype=torch.int32, device="cuda"),
               torch.zeros(5, dtype=torch.int32, device="cuda")]
        N.check(run_static(kd, n, bits, *got))
        N.check(run_context(ctx, kd, n, bits, *exp))
        torch.cuda.synchronize()
        visible = int(host_u32(exp[1])[0])
        assert int(host_u32(got[1])[0]) == visible
        assert np.array_equal(host_u32(got[2]), host_u32(exp[2]))
        assert first_mismatch(host_u32(got[0])[:visible], host_u32(exp[0])[:visible]) is None
        if call == 1:
            ordering, ovis = O.sort(keys, bits)
            assert ovis == visible
            assert first_mismatch(host_u32(got[0])[:visible], ordering[:visible]) is None


def test_static_sorter_empty_and_limit(cuda):
    idx = torch.zeros(4, dtype=torch.int32, device="cuda")
    cnt = torch.full((1,), 7, dtype=torch.int32, device="cuda")
    args = torch.zeros(5, dtype=torch.int32, device="cuda")
    kd = dev_u32(np.zeros(4, dtype=np.uint32))
    N.check(run_static(kd, 0, 32, idx, cnt, args))
    torch.cuda.synchronize()
    assert int(host_u32(cnt)[0]) == 0 and list(host_u32(args)) == [6144, 0, 0, 0, 0]
    rc = run_static(kd, N.STATIC_SORT_MAX_KEYS + 1, 32, idx, cnt, args)
    assert rc == N.RDX_E_CAPACITY
    assert "static scratch of 1048576 keys" in N.lib().rdx_last_error().decode()

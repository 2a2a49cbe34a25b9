"""The HIP decoders read the known-answer words of tests/test_layout_kat.py
(packedSplats, extSplats, packedSh, extSh, derived by hand from the
reference's bit layouts) as the layouts define them: the projection's float
outputs of tables of those literal words equal the hand-decoded values and
the oracle's projection of the same words, to the parity bar."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector, camera as cam
from tests import test_layout_kat as K
from tests.projection_check import compare, to_host

pytestmark = pytest.mark.gpu
N_COPIES = 257  # one partial wave past four full ones


def rig():
    return cam.mono(np.eye(4), 60, 256, 128)


def run(table, params):
    out = Projector().project(table, rig(), params, side=True)
    torch.cuda.synchronize()
    return to_host(out, table.count, 1)


def tile(words, n=N_COPIES):
    return np.tile(np.asarray(words, dtype=np.uint32), (n, 1))


@pytest.mark.parametrize("sh_degree", [0, 3])
def test_packed_known_answer_words_on_the_device(cuda, sh_degree):
    words = tile(K.PACKED_WORDS)
    chans = []
    if sh_degree:
        w = K.packed_sh_words()
        chans = [tile(w[4 * t:4 * t + 4]) for t in range(3)]
    table = DeviceTable.from_packed(words, N_COPIES, tuple(K.PACKED_ENC), chans, sh_degree)
    prm = Params(sort_bits=32)
    g = run(table, prm)
    # hand-decoded: rgb (0, 1, 51 / 255) without SH; the key distance |(1, -2, -4)|
    if sh_degree == 0:
        assert np.allclose(g["side"]["rgb"].reshape(3, -1)[:, :N_COPIES].T, [0.0, 1.0, 51 / 255], rtol=0, atol=1e-7)
    dist = g["side"]["distance"].reshape(-1)[:N_COPIES]
    assert np.allclose(dist, math.sqrt(21.0), rtol=1e-7)
    ref = O.project_packed(words, K.PACKED_ENC, chans, sh_degree, O.view_from_rig(rig()), O.params(sort_bits=32))
    rep = compare(g, ref, 1)
    assert rep["eye0_compared"] == N_COPIES


@pytest.mark.parametrize("sh_codec", [0, 2, 3])
def test_ext_known_answer_words_on_the_device(cuda, sh_codec):
    a, b = tile(K.EXT_A), tile(K.EXT_B)
    chans = []
    if sh_codec == 2:
        w = np.zeros(16, dtype=np.uint32)
        O.lib().orc_sh_encode(2, O.dptr(K.ext_sh_coeffs()), 3, O.uptr(w))
        for k, v in K.EXT_SH_KNOWN.items():
            assert w[k] == v
        chans = [tile(w[4 * t:4 * t + 4]) for t in range(4)]
    elif sh_codec == 3:
        w = K.packed_sh_words()
        chans = [tile(w[4 * t:4 * t + 4]) for t in range(3)]
    deg = 3 if sh_codec else 0
    table = DeviceTable.from_ext(a, b, N_COPIES, chans, sh_codec=sh_codec or 2, sh_degree=deg)
    g = run(table, Params(sort_bits=32))
    if sh_codec == 0:
        # extRgb (64, 255 negative, 128) / 255 at base 15 (no SH: not clamped); LoD opacity 192 / 128 = 1.5
        assert np.allclose(g["side"]["rgb"].reshape(3, -1)[:, :N_COPIES].T, [64 / 255, -1.0, 128 / 255], rtol=0,
                           atol=1e-7)
    ref = O.project_ext(a, b, chans, sh_codec or 2, deg, O.view_from_rig(rig()), O.params(sort_bits=32))
    rep = compare(g, ref, 1)
    assert rep["eye0_compared"] == N_COPIES

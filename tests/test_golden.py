"""The oracle against the committed golden fixtures (tests/golden/,
written by tests/golden/make_golden.py): the seeded generator, the FP64
projection, the keys and SortKeys.sort reproduce them exactly."""
import os

import numpy as np

from oracle import oracle as O
from radix_amd import workloads as W

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name + ".npz"))


def test_plain_small_golden():
    g = load("plain_small")
    spec = W.plain_small_spec()
    planes = W.generate(spec)
    ref = O.project_f32(planes, 0, O.view_from_rig(W.plain_small_camera()), O.params(sort_bits=32))
    assert np.array_equal(ref["keys"], g["keys"])
    assert np.array_equal(ref["words"][0][:4096], g["words"])
    ordering, visible = O.sort(g["keys"], 32)
    assert visible == int(g["visible"][0])
    assert np.array_equal(ordering[:spec.count], g["ordering"])
    # the order is the stable one: keys ascending, ties by index
    assert np.array_equal(ordering[:spec.count], np.argsort(g["keys"], kind="stable"))


def test_plain_4m_sample_golden():
    g = load("plain_4m_4096")
    planes = np.ascontiguousarray(W.generate(W.plain_4m_spec())[:, :4096])
    ref = O.project_f32(planes, 3, O.view_from_rig(W.plain_4m_camera()), O.params(sort_bits=32))
    for k in ("keys", "rgb", "margin"):
        assert np.array_equal(ref[k], g[k]), k
    assert np.array_equal(ref["words"][0], g["words"])
    assert np.array_equal(ref["conic"][0], g["conic"])
    assert np.array_equal(ref["opacity"][0], g["opacity"])


def test_stereo_sample_golden():
    g = load("stereo_4096")
    planes = W.generate(W.stereo_16m_spec(count=4096))
    words, chans = W.pack(planes, 3)
    ref = O.project_packed(words, (0.0, 1.0, -10.0, 2.0), chans, 3, O.view_from_rig(W.stereo_camera()),
                           O.params(sort_bits=32, key_per_eye=True))
    assert np.array_equal(ref["keys"], g["keys"])
    assert np.array_equal(ref["words"][0], g["words0"])
    assert np.array_equal(ref["words"][1], g["words1"])

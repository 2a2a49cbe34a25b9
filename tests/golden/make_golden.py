"""Writes the golden fixtures of tests/golden/ from the oracle (FP64 C
restatement of the reference's path) on the seeded synthetic workloads:

  plain_small.npz  65,536-splat SH0 cloud, 1280x720 camera, 32-bit keys:
                   all keys, the SortKeys.sort ordering and visible count,
                   the splats whose cull decision sits within 1e-5 of its
                   threshold, and the ewa words of the first 4096 splats
  plain_4m_4096.npz  the first 4096 splats of the 4M SH3 cloud under its
                   camera: words, keys, rgb (SH3 colour), conic, opacity
                   (the inputs are regenerated from the seeded generator)
  stereo_4096.npz  4096 packedSplats + packedSh SH3 splats under the
                   2064x2208 stereo rig, per-eye keys: words and keys per eye

The reference holds no fixture files (its tests generate inputs with the
same mulberry32 generator at run time), so these pin the oracle against
itself across changes and give the GPU tests committed vectors at small
sizes. Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from radix_amd import workloads as W  # noqa: E402


def plain_small():
    spec = W.plain_small_spec()
    planes = W.generate(spec)
    rig = W.plain_small_camera()
    ref = O.project_f32(planes, 0, O.view_from_rig(rig), O.params(sort_bits=32))
    ordering, visible = O.sort(ref["keys"], 32)
    return dict(keys=ref["keys"], ordering=ordering[:spec.count], visible=np.array([visible]),
                words=ref["words"][0][:4096], near=ref["margin"] < 1e-5)


def plain_4m_sample():
    spec = W.plain_4m_spec()
    planes = np.ascontiguousarray(W.generate(spec)[:, :4096])
    rig = W.plain_4m_camera()
    ref = O.project_f32(planes, 3, O.view_from_rig(rig), O.params(sort_bits=32))
    return dict(keys=ref["keys"], words=ref["words"][0], rgb=ref["rgb"], conic=ref["conic"][0],
                opacity=ref["opacity"][0], margin=ref["margin"])


def stereo_sample():
    spec = W.stereo_16m_spec(count=4096)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = (0.0, 1.0, -10.0, 2.0)
    rig = W.stereo_camera()
    ref = O.project_packed(words, enc, chans, 3, O.view_from_rig(rig), O.params(sort_bits=32, key_per_eye=True))
    return dict(keys=ref["keys"], words0=ref["words"][0], words1=ref["words"][1], margin=ref["margin"])


def main():
    for name, fn in (("plain_small", plain_small), ("plain_4m_4096", plain_4m_sample), ("stereo_4096", stereo_sample)):
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()

"""SPZ decode (rdx_spz_decode, radix_amd.spz) against the oracle's restatement
of readSpz (oracle/spz.py, src/core/spz.ts:297-347). CPU: the oracle pinned
by closed forms and round trips; GPU: the device planes against the oracle
(bit-exact, or one float32 ulp where a double map is rounded),
the projection of an SPZ table against the FP64 projection oracle, and the
errors."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import spz as S
from tests.helpers import Rng
from tests.spz_format import encode_smallest_three, encode_spz


def scene(n, seed, sh_degree=3):
    r = Rng(seed)
    u = r.random(n * 60).reshape(60, n)
    centers = np.stack([(u[0] - 0.5) * 6, (u[1] - 0.5) * 4, -2 - u[2] * 8], 1)
    log_scales = np.stack([-4.5 + u[3] * 2, -4.5 + u[4] * 2, -4.5 + u[5] * 2], 1)
    quats = Rotation.random(n, random_state=seed).as_quat()
    opacity = 0.05 + 0.9 * u[6]
    rgb = np.stack([u[7], u[8], u[9]], 1)
    sh = (u[10:10 + 45].T - 0.5) * 0.4
    return centers, log_scales, quats, opacity, rgb, sh


def test_smallest_three_round_trip_and_known_answers():
    assert S.decode_smallest_three(encode_smallest_three([0, 0, 0, 1])) == [0.0, 0.0, 0.0, 1.0]
    step = math.sqrt(0.5) / 511
    for q in Rotation.random(200, random_state=3).as_quat():
        d = np.array(S.decode_smallest_three(encode_smallest_three(list(q))))
        # the same rotation: q or -q, each kept component within half a step
        e = min(np.abs(d - q).max(), np.abs(d + q).max())
        assert e < 2.5 * step
        assert abs(np.linalg.norm(d) - 1.0) < 1e-12


def test_decode_closed_forms():
    # int24 sign extension, scales, colours, opacity and SH at the end points
    body = bytearray()
    body += bytes([0x00, 0x00, 0x80, 0xFF, 0xFF, 0x7F, 0x01, 0x00, 0x00])  # -2^23, 2^23 - 1, 1
    body += bytes([255])  # alpha
    body += bytes([0, 255, 128])  # colours
    body += bytes([0, 160, 255])  # scales
    body += (0xC0000000).to_bytes(4, "little")  # largest w, the rest 0: identity
    body += bytes([0, 128, 255] + [128] * 6)  # SH degree 1: 9 values
    out = S.decode(bytes(body), 1, 3, 1, 12)
    assert list(out[0:3, 0]) == [-(2 ** 23) / 4096, (2 ** 23 - 1) / 4096, 1 / 4096]
    assert list(out[3:6, 0]) == [-10.0, 0.0, 255 / 16 - 10]
    assert list(out[6:10, 0]) == [0.0, 0.0, 0.0, 1.0]
    assert out[10, 0] == 1.0
    C = S.SH_C0 / 0.15
    assert out[11, 0] == pytest.approx(0.5 - 0.5 * C) and out[12, 0] == pytest.approx(0.5 + 0.5 * C)
    assert list(out[14:17, 0]) == [-1.0, 0.0, 127 / 128]
    # version 1: half-float centres; version 2: xyz with w >= 0
    v1 = np.array([1.5, -2.25, 0.0009765625], dtype=np.float16).tobytes() + bytes([0, 0, 0, 0, 0, 0, 0]) + bytes([255, 127, 0])
    o1 = S.decode(v1, 1, 1, 0, 12)
    assert list(o1[0:3, 0]) == [1.5, -2.25, 0.0009765625]
    assert list(o1[6:9, 0]) == [1.0, 127 / 127.5 - 1.0, -1.0] and o1[9, 0] == 0.0


@pytest.mark.parametrize("version", [1, 2, 3])
def test_writer_round_trip(version):
    centers, log_scales, quats, opacity, rgb, sh = scene(300, 5 + version)
    f = encode_spz(centers, log_scales, quats, opacity, rgb, sh, 3, version=version)
    import gzip

    raw = gzip.decompress(f)
    out = S.decode(raw[16:], 300, version, 3, 12)
    tol_c = 1e-3 * np.abs(centers).max() if version == 1 else 0.5 / 4096
    assert np.abs(out[0:3].T - centers).max() <= tol_c + 1e-12
    assert np.abs(out[3:6].T - log_scales).max() <= 1 / 32 + 1e-12
    assert np.abs(out[10] - opacity).max() <= 0.5 / 255 + 1e-12
    assert np.abs(out[14:].T - sh).max() <= 0.5 / 128 + 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("version,degree,n", [(1, 3, 5000), (2, 3, 5000), (3, 3, 5000), (4, 3, 5000), (3, 0, 4999),
                                              (2, 1, 257), (3, 2, 1)])
def test_device_planes_equal_the_oracle(cuda, version, degree, n):
    import gzip

    from radix_amd import spz

    centers, log_scales, quats, opacity, rgb, sh = scene(n, 11 + version)
    f = encode_spz(centers, log_scales, quats, opacity, rgb, sh, degree, version=min(version, 3))
    if version == 4:  # version 4 inside gzip: v3's layout
        raw = bytearray(gzip.decompress(f))
        raw[4] = 4
        f = gzip.compress(bytes(raw))
    table = spz.read_spz(f)
    got = table.tensors["planes"].cpu().numpy()[:, :n].astype(np.float64)
    head, body = spz.unpack(f)
    ref = S.decode(body, n, head["version"], degree, head["fractional_bits"]).astype(np.float32).astype(np.float64)
    assert got.shape[0] == 14 + 3 * S.SH_VECTORS[degree]
    exact = list(range(0, 6)) + [10] + list(range(14, got.shape[0]))
    assert np.array_equal(got[exact], ref[exact])
    # the quaternion and colour, evaluated in double and rounded once: within
    # one float32 ulp of the oracle's rounded double value
    rows = list(range(6, 10)) + list(range(11, 14))
    ulp = np.spacing(np.abs(ref[rows]).astype(np.float32)).astype(np.float64)
    assert (np.abs(got[rows] - ref[rows]) <= ulp).all()


@pytest.mark.gpu
@pytest.mark.parametrize("version,degree,n", [(3, 3, 1001), (1, 2, 301)])  # 1001 x 65 B: a partial last word
def test_misaligned_body_takes_the_byte_path(cuda, version, degree, n):
    # A body at an odd address is decoded by the byte path; a 4-byte-aligned
    # one by the LDS-staged path. Both must write the same planes.
    import ctypes as C
    import gzip

    import torch

    from radix_amd import _native as N

    centers, log_scales, quats, opacity, rgb, sh = scene(n, 31 + version)
    raw = gzip.decompress(encode_spz(centers, log_scales, quats, opacity, rgb, sh, degree, version=version))
    body = raw[16:]
    k = 3 * S.SH_VECTORS[degree]
    outs = []
    for pad in (0, 1):
        dev = torch.zeros(len(body) + 4, dtype=torch.uint8, device="cuda")
        dev[pad:pad + len(body)] = torch.frombuffer(bytearray(body), dtype=torch.uint8).to("cuda")
        planes = torch.full((14 + k, n), float("nan"), dtype=torch.float32, device="cuda")
        N.check(N.lib().rdx_spz_decode(None, C.c_void_p(dev.data_ptr() + pad), len(body), n, version, degree, 12,
                                       C.c_void_p(planes.data_ptr()), n))
        torch.cuda.synchronize()
        outs.append(planes.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    ref = S.decode(body, n, version, degree, 12).astype(np.float32)
    assert np.array_equal(outs[0][:6], ref[:6]) and np.array_equal(outs[0][14:], ref[14:])


@pytest.mark.gpu
def test_decode_on_a_caller_stream(cuda):
    # The upload and the decode run on the caller's stream (default priority).
    import torch

    from radix_amd import spz

    n = 3000
    f = encode_spz(*scene(n, 51), 3, version=3)
    a = spz.read_spz(f).tensors["planes"]
    b = spz.read_spz(f, stream=torch.cuda.Stream()).tensors["planes"]
    assert torch.equal(a, b)


@pytest.mark.gpu
def test_empty_file(cuda):
    import gzip
    import struct

    from radix_amd import spz

    f = gzip.compress(struct.pack("<IIIBBBB", 0x5053474E, 3, 0, 2, 12, 1, 0))
    table = spz.read_spz(f)
    assert table.count == 0 and table.meta["sh_degree"] == 2
    assert table.tensors["planes"].shape[0] == 14 + 3 * 8


@pytest.mark.gpu
def test_spz_table_projects_like_the_oracle(cuda):
    import torch

    from oracle import oracle as O
    from radix_amd import Params, Projector, spz
    from radix_amd import workloads as W
    from tests.projection_check import compare, to_host

    n = 20000
    centers, log_scales, quats, opacity, rgb, sh = scene(n, 21)
    f = spz.read_spz(encode_spz(centers, log_scales, quats, opacity, rgb, sh, 3, version=3))
    planes = f.tensors["planes"].cpu().numpy()[:, :n]
    rig = W.plain_small_camera()
    out = Projector().project(f, rig, Params(sort_bits=32), side=True)
    torch.cuda.synchronize()
    g = to_host(out, n, 1)
    ref = O.project_f32(np.ascontiguousarray(planes), 3, O.view_from_rig(rig), O.params(sort_bits=32))
    rep = compare(g, ref, 1)
    assert rep["eye0_compared"] > 5000


def test_errors():
    import gzip

    from radix_amd import spz

    with pytest.raises(ValueError, match="Invalid SPZ file"):
        spz.unpack(gzip.compress(b"\0" * 32))
    bad = encode_spz([[0, 0, -1]], [[-4, -4, -4]], [[0, 0, 0, 1]], [0.5], [[0.5, 0.5, 0.5]], version=3)
    raw = bytearray(gzip.decompress(bad))
    raw[4] = 9
    with pytest.raises(ValueError, match="Unsupported SPZ version 9 or SH degree 0"):
        spz.unpack(gzip.compress(bytes(raw)))
    with pytest.raises(ValueError, match="NGSP"):
        spz.unpack(b"NGSP" + b"\0" * 28)

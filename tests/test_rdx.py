"""RDX files decoded onto the device (radix_amd.rdx, rdx_untranspose_bytes),
against the format of src/core/rdx.ts and compress.ts: the byte transpose's
closed form and inverse (transposeBytes restated byte by byte), the
containers' fields and the reference's error texts (CPU); on the GPU the
untranspose kernel against numpy's transpose for aligned and ragged sizes,
and whole files (several chunks, a ragged last one, chunks decoded out of
order) whose device channels equal the source arrays bit for bit and whose
splat tables project exactly as the same words uploaded directly, as the
reference's own round trips check (check-core.ts:201-236)."""
import numpy as np
import pytest

from radix_amd import rdx
from tests.helpers import Rng
from tests.rdx_format import CHUNK_MAGIC, RDX_MAGIC, container, encode_rdx, transpose_bytes


def test_transpose_closed_form_and_inverse():
    data = bytes(range(8))
    assert transpose_bytes(data, 4) == bytes([0, 4, 1, 5, 2, 6, 3, 7])
    r = Rng(1)
    for n, eb in ((5, 16), (7, 3), (12, 8), (1, 4)):
        x = bytes(np.floor(r.random(n * eb) * 256).astype(np.uint8))
        t = transpose_bytes(x, eb)
        # decompressTransposed: a transpose with the element count as the stride
        assert transpose_bytes(t, n) == x
        # numpy's view of the same: the (n, eb) byte matrix transposed
        assert t == np.frombuffer(x, np.uint8).reshape(n, eb).T.tobytes()


def test_containers_and_errors():
    count = 5000
    a = np.arange(count * 4, dtype=np.uint32).reshape(count, 4)
    f = encode_rdx({"extSplats": {"channels": {"extA": "uint32x4", "extB": "uint32x4"}, "data": {"extA": a, "extB": a + 7}}},
                   count, chunk_size=4096, meta={"name": "x"})
    header, length = rdx.decode_header(f)
    assert header["count"] == count and header["chunkSize"] == 4096 and len(header["chunks"]) == 2
    assert header["meta"] == {"name": "x"}
    meta, payloads, n = rdx.decode_container(f[length:], CHUNK_MAGIC)
    assert meta == {"start": 0, "count": 4096} and len(payloads) == 2 and n == header["chunks"][0]
    with pytest.raises(ValueError, match="Not an RDX piece of the expected type"):
        rdx.decode_header(f[length:])
    with pytest.raises(ValueError, match="RDX version 2 is not 1"):
        rdx.decode_header(container(RDX_MAGIC, {"version": 2}, []))
    assert rdx.format_bytes("uint32x4") == 16 and rdx.format_bytes("uint16") == 2 and rdx.format_bytes("float32x2") == 8


@pytest.mark.gpu
@pytest.mark.parametrize("count,eb", [(4096, 16), (4097, 16), (1200, 16), (1000, 8), (999, 4), (777, 3), (5, 1), (65536, 16), (3, 16)])
def test_untranspose_kernel(cuda, count, eb):
    import ctypes as C

    import torch

    from radix_amd import _native as N

    r = Rng(count + eb)
    x = np.floor(r.random(count * eb) * 256).astype(np.uint8).reshape(count, eb)
    planes = torch.from_numpy(np.ascontiguousarray(x.T).reshape(-1)).cuda()
    out = torch.zeros(count * eb + 16, dtype=torch.uint8, device="cuda")
    N.check(N.lib().rdx_untranspose_bytes(None, C.c_void_p(planes.data_ptr()), count, eb, C.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.array_equal(got[:count * eb], x.reshape(-1))
    assert not got[count * eb:].any()  # nothing past the elements


def ext_source(count, seed):
    from tests.test_project_gpu import ext_table

    a, b, chans = ext_table(count, seed, 2)
    return a, b, chans


@pytest.mark.gpu
def test_ext_file_decodes_to_the_same_table(cuda):
    import torch

    from radix_amd import DeviceTable, Params, Projector
    from radix_amd import workloads as W

    count = 13000
    a, b, chans = ext_source(count, 31)
    bundles = {"extSplats": {"channels": {"extA": "uint32x4", "extB": "uint32x4"}, "data": {"extA": a, "extB": b}},
               "extSh": {"channels": {f"extSh{c}": "uint32x4" for c in "ABCD"}, "meta": {"codec": "ext", "degree": 3},
                         "data": {f"extSh{c}": chans[k] for k, c in enumerate("ABCD")}}}
    f = encode_rdx(bundles, count, chunk_size=4096)
    data = rdx.read_rdx(f)
    assert np.array_equal(data.words("extSplats", "extA").cpu().numpy().view(np.uint32), a)
    assert np.array_equal(data.words("extSh", "extShD").cpu().numpy().view(np.uint32), chans[3])
    table = rdx.splat_table(data)
    direct = DeviceTable.from_ext(a, b, count, chans, sh_codec=2, sh_degree=3)
    rig = W.orbit_camera(4, 7.0, 1024, 768)
    p = Projector()
    x, y = p.project(table, rig, Params(sort_bits=32)), p.project(direct, rig, Params(sort_bits=32))
    torch.cuda.synchronize()
    assert torch.equal(x["words"][0], y["words"][0]) and torch.equal(x["keys"], y["keys"])
    assert int((x["words"][0][:, 3] & 0xFF).ne(0).sum()) > count // 4
    # chunks decoded in reverse order give the same channels
    header, off = rdx.decode_header(f)
    rev = rdx.RdxDeviceData(header, "cuda")
    pieces = []
    for size in header["chunks"]:
        pieces.append(memoryview(f)[off:off + size])
        off += size
    for piece in reversed(pieces):
        rdx.decode_chunk(header, piece, rev)
    assert torch.equal(rev.words("extSplats", "extB"), data.words("extSplats", "extB"))


@pytest.mark.gpu
def test_packed_file_with_encoding_meta(cuda):
    import torch

    from radix_amd import DeviceTable, Params, Projector
    from radix_amd import workloads as W

    spec = W.stereo_16m_spec(count=9000, seed=2)
    planes = W.generate(spec)
    words, chans = W.pack(planes, 3)
    enc = {"rgbMin": 0.0, "rgbMax": 1.0, "lnScaleMin": -10.0, "lnScaleMax": 2.0}
    bundles = {"packedSplats": {"channels": {"packed": "uint32x4"}, "meta": {"encoding": enc}, "data": {"packed": words}},
               "packedSh": {"channels": {f"packedSh{c}": "uint32x4" for c in "ABC"}, "meta": {"codec": "packed", "degree": 3},
                            "data": {f"packedSh{c}": chans[k] for k, c in enumerate("ABC")}}}
    data = rdx.read_rdx(encode_rdx(bundles, spec.count, chunk_size=8192))
    table = rdx.splat_table(data)
    direct = DeviceTable.from_packed(words, spec.count, (0.0, 1.0, -10.0, 2.0), chans, 3)
    rig = W.stereo_camera(width=512, height=544)
    prm = Params(sort_bits=32, key_per_eye=True)
    x, y = Projector().project(table, rig, prm), Projector().project(direct, rig, prm)
    torch.cuda.synchronize()
    for e in range(2):
        assert torch.equal(x["words"][e], y["words"][e])
    assert torch.equal(x["keys"], y["keys"])


@pytest.mark.gpu
def test_decode_on_a_caller_stream(cuda):
    # Uploads and transposes run on the caller's stream: the channels equal
    # the sources (a stream of default priority).
    import torch

    count = 9001
    a, b, _ = ext_source(count, 41)
    f = encode_rdx({"extSplats": {"channels": {"extA": "uint32x4", "extB": "uint32x4"}, "data": {"extA": a, "extB": b}}},
                   count, chunk_size=4096)
    data = rdx.read_rdx(f, stream=torch.cuda.Stream())
    assert np.array_equal(data.words("extSplats", "extA").cpu().numpy().view(np.uint32), a)
    assert np.array_equal(data.words("extSplats", "extB").cpu().numpy().view(np.uint32), b)

"""The production draw on the GPU against the FP64 draw oracle (oracle/draw.c,
src/render/splatDraw.ts:98-181) on the same words and order: a projected and
sorted scene (the per-frame project, sort and draw), crafted words with LoD
opacities at several falloffs, linear output, the early stop, out-of-range
indices and counts, and the entry overflow."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector
from radix_amd import workloads as W
from radix_amd import _native as N
from radix_amd.draw import Drawer, draw_projected
from tests.helpers import Rng

pytestmark = pytest.mark.gpu


def dev_i32(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def compare(img, ref, frac=0.005, tol=1e-4):
    """fp32 against fp64: a cutoff or minAlpha decision may flip at a few
    pixels, everything else agrees to float rounding."""
    d = np.abs(img - ref)
    print(f"max |diff| {d.max():.3e}, median {np.median(d):.3e}, pixels above {tol}: {np.mean(d.max(axis=0) > tol):.5f}",
          flush=True)
    assert np.mean(d.max(axis=0) > tol) < frac
    assert np.median(d) < 1e-6


def projected_scene(width, height, seed=1, count=20000):
    spec = W.plain_small_spec(seed=seed, count=count)
    planes = W.generate(spec)
    table = DeviceTable.from_f32_planes(planes, count, 0)
    rig = W.orbit_camera(seed, 4.0, width, height)
    prm = Params(sort_bits=32)
    out = Projector().project_sort(table, rig, prm)
    torch.cuda.synchronize()
    return out, prm


@pytest.mark.parametrize("size", [(320, 240), (333, 197)])
def test_projected_scene_matches_oracle(cuda, size):
    Wd, Ht = size
    out, prm = projected_scene(Wd, Ht)
    drawer = Drawer()
    img = draw_projected(out, Wd, Ht, prm, drawer=drawer)
    torch.cuda.synchronize()
    st = drawer.stats()
    words = out["words"][0].cpu().numpy().view(np.uint32)
    order = out["indices"].cpu().numpy().view(np.uint32)
    count = int(out["count"].cpu().item())
    print("stats", st, "count", count, flush=True)
    assert st["drawn"] == count and not st["overflowed"] and st["entries"] > count
    ref = O.draw(words, order, count, Wd, Ht, prm.max_std_dev, prm.min_alpha, prm.falloff)
    assert ref[3].max() > 0.5
    compare(img.cpu().numpy(), ref)


def crafted(n, W_, H_, seed):
    """Words of n splats spread over the image with LoD opacity bytes among them."""
    r = Rng(seed)
    u = r.random(8 * n).reshape(8, n)
    fx = np.floor(((u[0] * 2.2 - 1.1) + 2.0) / 4.0 * 65535.0).astype(np.uint32)
    fy = np.floor(((u[1] * 2.2 - 1.1) + 2.0) / 4.0 * 65535.0).astype(np.uint32)
    angle = np.floor(u[2] * 4096).astype(np.uint32)
    sx = np.floor(u[3] * 800 + 1500).astype(np.uint32)  # log2 scale ~ 1.9 .. 5
    sy = np.floor(u[4] * 800 + 1300).astype(np.uint32)
    byte = np.floor(u[5] * 255 + 1).astype(np.uint32)  # 1 .. 255: above 128 are LoD opacities
    rgb = np.array([O.lib().orc_encode_ext_rgb(float(a), float(b), float(c))
                    for a, b, c in zip(u[6], u[7], 1.0 - u[6])], dtype=np.uint32)
    w = np.stack([fx | (fy << 16), np.uint32(5000) | (angle << 20), rgb, byte | (sx << 8) | (sy << 20)], axis=1)
    return w.astype(np.uint32)


@pytest.mark.parametrize("falloff,linear", [(1.0, False), (0.5, False), (0.0, False), (1.0, True)])
def test_crafted_words_lod_falloff_linear(cuda, falloff, linear):
    Wd, Ht, n = 160, 128, 3000
    words = crafted(n, Wd, Ht, 5)
    order = np.arange(n, dtype=np.uint32)[::-1].copy()
    prm = Params(falloff=falloff)
    img = Drawer(linear_output=linear).draw(dev_i32(words), dev_i32(order), dev_i32([n]), Wd, Ht, prm)
    ref = O.draw(words, order, n, Wd, Ht, prm.max_std_dev, prm.min_alpha, falloff, linear)
    assert (words[:, 3] & 0xFF).max() > 128
    compare(img.cpu().numpy(), ref)


def test_indices_past_n_and_count_past_order_are_skipped(cuda):
    Wd, Ht, n = 96, 64, 500
    words = crafted(n, Wd, Ht, 6)
    order = np.arange(n, dtype=np.uint32)
    order[::7] = n + 12345  # not drawn (the vertex stage's index guard)
    prm = Params()
    d = Drawer()
    img = d.draw(dev_i32(words), dev_i32(order), dev_i32([10 * n]), Wd, Ht, prm)  # count clamped to the order
    torch.cuda.synchronize()
    assert d.stats()["drawn"] == n
    ref = O.draw(words, order, n, Wd, Ht, prm.max_std_dev, prm.min_alpha, prm.falloff)
    compare(img.cpu().numpy(), ref)


def test_early_stop_within_its_bound(cuda):
    Wd, Ht = 320, 240
    out, prm = projected_scene(Wd, Ht, seed=3)
    exact = draw_projected(out, Wd, Ht, prm, drawer=Drawer()).cpu().numpy()
    stop = 1e-3
    early = draw_projected(out, Wd, Ht, prm, drawer=Drawer(t_stop=stop)).cpu().numpy()
    # a stopped pixel leaves out fragments worth at most its transmittance
    # (below t_stop) times the largest colour
    words = out["words"][0].cpu().numpy().view(np.uint32)
    count = int(out["count"].cpu().item())
    drawn = words[out["indices"].cpu().numpy().view(np.uint32)[:count]]
    buf = (C.c_double * 3)()
    cmax = 0.0
    for w in drawn[:, 2]:
        O.lib().orc_decode_ext_rgb(int(w), buf)
        cmax = max(cmax, *buf[:])
    d = np.abs(early - exact)
    bound = stop * max(cmax, 1.0)  # the alpha plane: at most the transmittance left
    print(f"early stop: max |diff| {d.max():.3e}, bound {bound:.3e}", flush=True)
    assert d.max() <= bound + 1e-5
    assert d.max() > 0.0


def test_dense_tile_split_into_chunks(cuda):
    # 6,000 splats over one 16x16 tile: its entries span several chunks of
    # 1,024, composited apart and folded front to back
    Wd, Ht, n = 64, 64, 6000
    r = Rng(8)
    u = r.random(6 * n).reshape(6, n)
    ndc = lambda px, size: (px + 0.5) * 2.0 / size - 1.0
    fx = np.floor((ndc(16 + 16 * u[0], Wd) + 2.0) / 4.0 * 65535.0).astype(np.uint32)
    fy = np.floor((-ndc(16 + 16 * u[1], Ht) + 2.0) / 4.0 * 65535.0).astype(np.uint32)
    sx = np.floor(u[2] * 300 + 1000).astype(np.uint32)  # 1.5 .. 3 px
    byte = np.floor(u[3] * 40 + 4).astype(np.uint32)
    rgb = np.array([O.lib().orc_encode_ext_rgb(float(a), float(b), 0.3) for a, b in zip(u[4], u[5])], dtype=np.uint32)
    words = np.stack([fx | (fy << 16), np.full(n, 7 | (100 << 20), dtype=np.uint32), rgb, byte | (sx << 8) | (sx << 20)],
                     axis=1).astype(np.uint32)
    order = np.arange(n, dtype=np.uint32)
    prm = Params()
    d = Drawer()
    img = d.draw(dev_i32(words), dev_i32(order), dev_i32([n]), Wd, Ht, prm)
    torch.cuda.synchronize()
    print("stats", d.stats(), flush=True)
    ref = O.draw(words, order, n, Wd, Ht, prm.max_std_dev, prm.min_alpha, prm.falloff)
    assert ref[3].max() > 0.99
    compare(img.cpu().numpy(), ref, frac=0.01)
    stop = 1e-3
    early = Drawer(t_stop=stop).draw(dev_i32(words), dev_i32(order), dev_i32([n]), Wd, Ht, prm).cpu().numpy()
    assert np.abs(early - img.cpu().numpy()).max() <= 1.01 * stop + 1e-5


def test_entry_limit_raises_and_no_limit_sizes_to_the_frame(cuda):
    Wd, Ht, n = 128, 128, 200
    words = crafted(n, Wd, Ht, 7)
    order = np.arange(n, dtype=np.uint32)
    d = Drawer(entries_per_splat=1)
    with pytest.raises(N.RadixError, match=r"entries exceed the limit of 200") as err:
        d.draw(dev_i32(words), dev_i32(order), dev_i32([n]), Wd, Ht, Params())
    assert err.value.code == N.RDX_E_CAPACITY
    free = Drawer()
    img = free.draw(dev_i32(words), dev_i32(order), dev_i32([n]), Wd, Ht, Params())
    torch.cuda.synchronize()
    st = free.stats()
    assert st["entries"] > n and st["capacity"] >= st["entries"] and not st["overflowed"]
    ref = O.draw(words, order, n, Wd, Ht)
    compare(img.cpu().numpy(), ref, frac=0.01)

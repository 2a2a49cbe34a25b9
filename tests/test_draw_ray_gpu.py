"""The draw of the ray-evaluated methods (rdx_draw_ray) on the GPU against the
FP64 oracle (orc_draw_ray, pinned by tests/test_oracle_draw_ray.py) on the same
words and order: each method's words of a projected cloud, mono and stereo
(either eye), an orthographic camera, LoD opacities above one, linear output,
out-of-range indices and a count past the order; a zero count, and an eye
the camera does not have."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import DeviceTable, Params, Projector
from radix_amd import _native as N
from radix_amd import camera as cam
from radix_amd import workloads as W
from radix_amd.draw import Drawer
from tests.test_draw_gpu import compare, dev_i32

pytestmark = pytest.mark.gpu

METHODS = ["3dgut", "htgs", "raygs"]


def cloud(seed, count, lod=False):
    spec = W.plain_small_spec(seed=seed, count=count)
    planes = W.generate(spec)
    if lod:
        # a tenth of the splats with LoD opacities above one (stored (op + 3) / 4)
        planes[10, ::10] = 1.0 + 2.0 * (np.arange(planes[10, ::10].size) % 7) / 7.0
    return DeviceTable.from_f32_planes(planes, count, 0)


def order_of(keys):
    """SortKeys.sort of the words' keys: the back-to-front order of the visible splats."""
    ordering, visible = O.sort(keys, 32)
    return ordering[:visible].astype(np.uint32), visible


def run(method, table, rig, prm, eye=0, linear=False, order=None, count=None):
    out = Projector().project_ray(table, rig, prm, method)
    torch.cuda.synchronize()
    keys = out["keys"].cpu().numpy().view(np.uint32)
    if order is None:
        order, count = order_of(keys)
    drawer = Drawer(linear_output=linear)
    img = drawer.draw_ray(method, out["channels"], dev_i32(order), dev_i32(np.array([count])), rig, prm, eye=eye)
    torch.cuda.synchronize()
    words = out["words"].cpu().numpy().view(np.uint32)
    ref = O.draw_ray(method, words, order, count, rig, eye=eye, max_std_dev=prm.max_std_dev,
                     min_alpha=prm.min_alpha, falloff=prm.falloff, linear=linear)
    st = drawer.stats()
    print(f"{method}: {count} drawn, stats {st}, alpha max {ref[3].max():.3f}", flush=True)
    assert st["drawn"] == min(count, len(order))
    return img.cpu().numpy(), ref


@pytest.mark.parametrize("method", METHODS)
def test_projected_cloud_matches_oracle(cuda, method):
    table = cloud(61, 6000)
    rig = W.orbit_camera(61, 4.0, 320, 240)
    prm = Params(sort_bits=32)
    img, ref = run(method, table, rig, prm)
    assert ref[3].max() > 0.5
    compare(img, ref)


@pytest.mark.parametrize("method", METHODS)
def test_stereo_eyes(cuda, method):
    table = cloud(62, 5000)
    rig = cam.stereo(cam.orbit_world(cam.random_direction(62), 3.5), 0.064, 80.0, 256, 200)
    prm = Params(sort_bits=32)
    for eye in (0, 1):
        img, ref = run(method, table, rig, prm, eye=eye)
        assert ref[3].max() > 0.3
        compare(img, ref)


@pytest.mark.parametrize("method", METHODS)
def test_orthographic(cuda, method):
    table = cloud(63, 4000)
    world = cam.orbit_world(cam.random_direction(63), 4.0)
    rig = cam.mono_ortho(world, 1.6, 240, 180)
    prm = Params(sort_bits=32)
    img, ref = run(method, table, rig, prm)
    assert ref[3].max() > 0.3
    compare(img, ref)


@pytest.mark.parametrize("method", METHODS)
def test_lod_opacities_and_linear_output(cuda, method):
    table = cloud(64, 5000, lod=True)
    rig = W.orbit_camera(64, 4.0, 256, 192)
    prm = Params(sort_bits=32, falloff=0.6)
    img, ref = run(method, table, rig, prm, linear=True)
    compare(img, ref)


def test_skipped_indices_and_clamped_count(cuda):
    table = cloud(65, 3000)
    rig = W.orbit_camera(65, 4.0, 200, 160)
    prm = Params(sort_bits=32)
    out = Projector().project_ray(table, rig, prm, "htgs")
    torch.cuda.synchronize()
    order, visible = order_of(out["keys"].cpu().numpy().view(np.uint32))
    # indices past n are skipped; a count above the order's length draws it all
    order = order.copy()
    order[::5] = 3000 + np.arange(order[::5].size, dtype=np.uint32)
    img, ref = run("htgs", table, rig, prm, order=order, count=len(order) + 100)
    compare(img, ref)


def test_zero_count_draws_a_transparent_frame(cuda):
    table = cloud(66, 2000)
    rig = W.orbit_camera(66, 4.0, 96, 64)
    prm = Params(sort_bits=32)
    img, ref = run("raygs", table, rig, prm, order=np.arange(2000, dtype=np.uint32), count=0)
    assert not img.any() and not ref.any()


def test_an_eye_the_camera_lacks_is_refused(cuda):
    table = cloud(67, 1000)
    rig = W.orbit_camera(67, 4.0, 96, 64)  # one eye
    prm = Params(sort_bits=32)
    out = Projector().project_ray(table, rig, prm, "htgs")
    order, count = dev_i32(np.arange(1000, dtype=np.uint32)), dev_i32(np.array([1000]))
    with pytest.raises(N.RadixError, match=r"eye 1 of a camera of 1 eyes") as err:
        Drawer().draw_ray("htgs", out["channels"], order, count, rig, prm, eye=1)
    assert err.value.code == N.RDX_E_INVALID
    with pytest.raises(ValueError, match=r"is not one of"):
        Drawer().draw_ray("ewa", out["channels"], order, count, rig, prm)

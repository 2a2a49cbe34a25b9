"""The Node N-API binding: tests/node/check_napi.js run by node (CPU checks
here, GPU checks marked gpu)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "tests", "node", "check_napi.js")
ADDON = os.path.join(ROOT, "radix_amd", "napi", "radix_napi.node")


def run_node(mode: str):
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    assert os.path.exists(ADDON), "radix_napi.node is not built"
    r = subprocess.run([node, SCRIPT, mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout


def test_napi_cpu_checks():
    run_node("cpu")


@pytest.mark.gpu
def test_napi_gpu_checks():
    run_node("gpu")


@pytest.mark.gpu
def test_napi_adjoint_and_draw_match_the_fp64_oracle(tmp_path):
    """The Node binding's compositor (Rasterizer forward + adjoint, on a
    GpuStream), draw and ray draw (SplatDrawer) on fixtures written here: the adjoint
    of the 400-splat scene's L2 loss within 1e-3 of the FP64 host gradient
    per parameter kind, the image and the drawn frame to the compositor's
    and draw's parity bar."""
    import json

    import numpy as np

    from oracle import oracle as O
    from radix_amd import camera as cam
    from radix_amd import workloads as W
    from tests.test_oracle_raster import small_scene

    theta, r, target = small_scene()
    n = theta.shape[0]
    rig = cam.mono(np.eye(4), 60.0, 64, 64)
    np.ascontiguousarray(theta.T, dtype=np.float32).tofile(tmp_path / "theta.f32")
    np.ascontiguousarray(target, dtype=np.float32).tofile(tmp_path / "target.f32")
    camera = {"head": list(cam.cm(rig.head)), "size": [64, 64], "ndcZMin": 0,
              "eyes": [{"view": list(cam.cm(rig.eyes[0].view)), "projection": list(cam.cm(rig.eyes[0].projection))}]}
    (tmp_path / "scene.json").write_text(json.dumps({"n": n, "camera": camera}))
    # the draw: a 20,000-splat frame projected and sorted by the oracle
    spec = W.plain_small_spec(seed=1, count=20000)
    planes = W.generate(spec)
    drig = W.orbit_camera(1, 4.0, 320, 240)
    proj = O.project_f32(planes, 0, O.view_from_rig(drig), O.params(sort_bits=32))
    ordering, visible = O.sort(proj["keys"], 32)
    words = np.ascontiguousarray(proj["words"][0], dtype=np.uint32)
    words.tofile(tmp_path / "words.u32")
    np.ascontiguousarray(ordering[:visible], dtype=np.uint32).tofile(tmp_path / "order.u32")
    (tmp_path / "draw.json").write_text(json.dumps({"n": spec.count, "count": int(visible), "capacity": int(visible),
                                                    "width": 320, "height": 240}))
    # a covSplats table with packedSh bundles, projected and sorted by the binding
    from radix_amd import DeviceTable, Params, Projector
    from tests.test_project_gpu import cov_table

    nc = 5000
    ca, cb, cch = cov_table(nc, 46, 3, lod=True)
    ca.tofile(tmp_path / "covA.u32")
    cb.tofile(tmp_path / "covB.u32")
    for t, ch in enumerate(cch):
        ch.tofile(tmp_path / f"sh{t}.u32")
    crig = W.orbit_camera(46, 6.0, 640, 480)
    ccam = {"head": list(cam.cm(crig.head)), "size": [640, 480], "ndcZMin": 0,
            "eyes": [{"view": list(cam.cm(crig.eyes[0].view)), "projection": list(cam.cm(crig.eyes[0].projection))}]}
    (tmp_path / "cov.json").write_text(json.dumps({"n": nc, "camera": ccam}))
    # the ray draw: a stereo raygs frame (channels w0, w1, w2) projected and sorted by the oracle, eye 1
    rrig = cam.stereo(cam.orbit_world(cam.random_direction(47), 4.0), 0.064, 70.0, 256, 192)
    rproj = O.project_ray_f32("raygs", planes[:, :8000].copy(), 0, O.view_from_rig(rrig), O.params(sort_bits=32))
    rorder, rvisible = O.sort(rproj["keys"], 32)
    for k, ch in enumerate(O.ray_channels(rproj["words"])):
        ch.tofile(tmp_path / f"ray_w{k}.u32")
    np.ascontiguousarray(rorder[:rvisible], dtype=np.uint32).tofile(tmp_path / "ray_order.u32")
    rcam = {"head": list(cam.cm(rrig.head)), "size": [256, 192], "ndcZMin": 0,
            "eyes": [{"view": list(cam.cm(e.view)), "projection": list(cam.cm(e.projection))} for e in rrig.eyes]}
    (tmp_path / "ray.json").write_text(json.dumps({"n": 8000, "count": int(rvisible), "method": "raygs", "eye": 1,
                                                   "camera": rcam}))
    node = shutil.which("node")
    if node is None:
        pytest.skip("node is not installed")
    run = subprocess.run([node, SCRIPT, "fixture", str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=240)
    print(run.stdout)
    assert run.returncode == 0, run.stdout

    loss_ref, g_ref = O.diff_grad(theta, r, target)
    g = np.fromfile(tmp_path / "grad.f32", dtype=np.float32).reshape(59, n).T.astype(np.float64)
    kinds = {"mean": slice(0, 3), "log_scale": slice(3, 6), "quat": slice(6, 10), "opacity_logit": slice(10, 11),
             "rgb": slice(11, 14), "sh": slice(14, 59)}
    for name, sl in kinds.items():
        rel = np.linalg.norm(g[:, sl] - g_ref[:, sl]) / max(np.linalg.norm(g_ref[:, sl]), 1e-30)
        print(f"{name}: relative error {rel:.2e}")
        assert rel < 1e-3, name
    img = np.fromfile(tmp_path / "image.f32", dtype=np.float32).reshape(3, 64, 64).astype(np.float64)
    ref, _, _ = O.diff_render(theta, r)
    d = np.abs(img - ref)
    assert np.mean(d > 1e-4) < 0.005 and np.median(d) < 1e-6
    rgba = np.fromfile(tmp_path / "rgba.f32", dtype=np.float32).reshape(4, 240, 320).astype(np.float64)
    dref = O.draw(words, ordering[:visible], int(visible), 320, 240)
    dd = np.abs(rgba - dref)
    assert np.mean(dd.max(axis=0) > 1e-4) < 0.005 and np.median(dd) < 1e-6
    rgba = np.fromfile(tmp_path / "ray_rgba.f32", dtype=np.float32).reshape(4, 192, 256).astype(np.float64)
    rref = O.draw_ray("raygs", rproj["words"], rorder[:rvisible], int(rvisible), rrig, eye=1)
    assert rref[3].max() > 0.3
    dd = np.abs(rgba - rref)
    assert np.mean(dd.max(axis=0) > 1e-4) < 0.005 and np.median(dd) < 1e-6
    # the covSplats table through the binding equals the Python host's call of the same kernels,
    # and its order is SortKeys.sort of its keys
    ctable = DeviceTable.from_cov(ca, cb, nc, cch, sh_codec=3, sh_degree=3)
    out = Projector().project_sort(ctable, crig, Params(sort_bits=32))
    import torch

    torch.cuda.synchronize()
    keys = np.fromfile(tmp_path / "cov_keys.u32", dtype=np.uint32)
    assert np.array_equal(keys, out["keys"].cpu().numpy().view(np.uint32)[:nc])
    assert np.array_equal(np.fromfile(tmp_path / "cov_words.u32", dtype=np.uint32).reshape(nc, 4),
                          out["words"][0].cpu().numpy().view(np.uint32)[:nc])
    ordering, visible = O.sort(keys, 32)
    assert np.array_equal(np.fromfile(tmp_path / "cov_indices.u32", dtype=np.uint32), ordering[:visible])

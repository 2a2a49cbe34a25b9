"""The differentiable compositor on the GPU against the FP64 oracle
(oracle/raster.c): image, loss and the gradient of all 59 parameters per
splat; the torch.autograd.Function harness; the 262,144-splat 512x512
workload through size-independent checks (directional derivatives of the
loss along the base colours and SH coefficients, where the image is linear,
against central differences of the GPU's own forward)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from radix_amd import Params
from radix_amd import _native as N
from radix_amd import camera as cam
from radix_amd import diff as D
from radix_amd import workloads as W
from tests.test_oracle_raster import small_scene

pytestmark = pytest.mark.gpu

KINDS = {"mean": slice(0, 3), "log_scale": slice(3, 6), "quat": slice(6, 10), "opacity_logit": slice(10, 11),
         "rgb": slice(11, 14), "sh": slice(14, 59)}


def gpu_loss_grad(theta, rig, target, rast=None):
    rast = rast or D.Rasterizer(sh_degree=3)
    p = D.params_tensor(theta).requires_grad_(True)
    img = D.composite(p, rig, Params(), rast)
    t = torch.from_numpy(np.ascontiguousarray(target, dtype=np.float32)).cuda()
    loss = ((img - t) ** 2).mean()
    loss.backward()
    torch.cuda.synchronize()
    return float(loss.item()), p.grad.cpu().numpy().T.astype(np.float64), img.detach().cpu().numpy(), rast


def scene_rig(size=64):
    return cam.mono(np.eye(4), 60.0, size, size)


def test_forward_image_matches_oracle(cuda):
    theta, r, target = small_scene()
    rig = scene_rig()
    rast = D.Rasterizer(sh_degree=3)
    img, alpha = rast.forward(D.params_tensor(theta), rig, Params())
    torch.cuda.synchronize()
    ref, ref_alpha, drawn = O.diff_render(theta, r)
    st = rast.stats()
    print("stats", st, "drawn by the oracle", drawn, flush=True)
    assert st["visible"] == drawn and not st["overflowed"]
    d = np.abs(img.cpu().numpy() - ref)
    da = np.abs(alpha.cpu().numpy() - ref_alpha)
    print("max |image diff|", d.max(), "median", np.median(d), flush=True)
    # fp32 against fp64: a cutoff or stop decision may flip at a few pixels
    assert np.mean(d > 1e-4) < 0.005
    assert np.mean(da > 1e-4) < 0.005
    assert np.median(d) < 1e-6


def test_gradient_matches_fp64_reference(cuda):
    theta, r, target = small_scene()
    loss_ref, g_ref = O.diff_grad(theta, r, target)
    loss, g, _, _ = gpu_loss_grad(theta, scene_rig(), target)
    assert loss == pytest.approx(loss_ref, rel=1e-3)
    for name, sl in KINDS.items():
        a, b = g[:, sl], g_ref[:, sl]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{name}: |g| {np.linalg.norm(b):.3e}, relative error {rel:.2e}", flush=True)
        assert rel < 1e-3, name


def test_gradient_across_chunks_matches_fp64_reference(cuda):
    # The adjoint takes each tile's list in chunks of 256 entries, each from
    # the forward's record of its transmittance and of the later chunks'
    # colour: 4,000 large, faint splats over 3 x 3 tiles put 1,500 to 4,000
    # entries on every tile (6 to 16 chunks); the centre's pixels stop early.
    spec = W.CloudSpec(seed=71, count=4000, sh_degree=3, centers=0, offset=(0.0, 0.0, -3.0), log_scale_mean=-1.2,
                       log_scale_std=0.2)
    theta = D.theta_from_planes(W.generate(spec))
    theta[:, 10] = -3.3 + 0.5 * np.sin(np.arange(theta.shape[0]))  # opacity 0.023 to 0.060
    size = 48
    rig = scene_rig(size)
    r = O.raster(size, size, rig.eyes[0].view, rig.eyes[0].projection, sh_degree=3)
    target, _, _ = O.diff_render(D.perturbed(theta, 72), r)
    loss_ref, g_ref = O.diff_grad(theta, r, target)
    loss, g, img, rast = gpu_loss_grad(theta, rig, target)
    st = rast.stats()
    print("stats", st, flush=True)
    assert st["entries"] > 9 * 6 * 256
    _, alpha, _ = O.diff_render(theta, r)
    assert 0.5 < alpha.mean() < 0.99 and alpha.max() > 0.9998  # most pixels reach deep, some stop
    assert loss == pytest.approx(loss_ref, rel=1e-3)
    for name, sl in KINDS.items():
        a, b = g[:, sl], g_ref[:, sl]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{name}: |g| {np.linalg.norm(b):.3e}, relative error {rel:.2e}", flush=True)
        assert rel < 1e-3, name


def test_sh_degree_and_untouched_splats(cuda):
    # SH degree 1 uses bands 1 only: gradients of bands 2-3 are zero; splats
    # behind the camera get zero gradients.
    theta, r, target = small_scene(n=200, seed=40)
    theta[:10, 2] = 5.0  # behind the camera (it looks down -z from the origin)
    r.sh_degree = 1
    target1, _, _ = O.diff_render(theta, r)
    target1 = target1 * 0.9
    _, g_ref = O.diff_grad(theta, r, target1)
    _, g, _, _ = gpu_loss_grad(theta, scene_rig(), target1, D.Rasterizer(sh_degree=1))
    assert np.all(g[:10] == 0.0)
    assert np.all(g[:, 14 + 9:] == 0.0)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel < 1e-3


def test_entry_limit_raises_and_no_limit_sizes_to_the_frame(cuda):
    # large splats: many more (splat, tile) entries than splats
    theta, r, target = small_scene(n=100, seed=50, log_scale=-1.5)
    rast = D.Rasterizer(sh_degree=3, entries_per_splat=1)
    with pytest.raises(N.RadixError, match=r"entries exceed the limit of 100") as err:
        rast.forward(D.params_tensor(theta), scene_rig(), Params())
    assert err.value.code == N.RDX_E_CAPACITY
    # the failed forward left nothing to differentiate
    with pytest.raises(N.RadixError, match=r"no forward to differentiate") as err:
        rast.backward(torch.zeros((3, 64, 64), dtype=torch.float32, device="cuda"))
    assert err.value.code == N.RDX_E_INVALID
    free = D.Rasterizer(sh_degree=3)
    img, alpha = free.forward(D.params_tensor(theta), scene_rig(), Params())
    torch.cuda.synchronize()
    st = free.stats()
    assert st["entries"] > 100 and st["capacity"] >= st["entries"] and not st["overflowed"]
    ref, ref_alpha, _ = O.diff_render(theta, r)
    d = np.abs(img.cpu().numpy().astype(np.float64) - ref)
    assert np.mean(d > 1e-4) < 0.005 and np.median(d) < 1e-6


def test_full_size_directional_derivatives(cuda):
    theta, target_theta, rig = D.workload()
    rast = D.Rasterizer(sh_degree=3)
    target, _ = rast.forward(D.params_tensor(target_theta), rig, Params())
    target = target.clone()
    print("target entries", rast.stats(), flush=True)
    p = D.params_tensor(theta)
    img, _ = rast.forward(p, rig, Params())
    st = rast.stats()
    print("stats", st, flush=True)
    assert not st["overflowed"] and st["visible"] > 200000
    gi = 2.0 * (img - target) / img.numel()
    g = rast.backward(gi)
    torch.cuda.synchronize()
    assert torch.isfinite(g).all()

    def loss(q):
        im, _ = rast.forward(q, rig, Params())
        return float(((im.double() - target.double()) ** 2).mean().item())

    # base colours: the image is linear in them away from the clamp at 0
    # (splats whose base colours all exceed 0.5 stay clear of it: their SH
    # terms are a few hundredths), the loss quadratic, so central
    # differences are exact up to rounding
    gen = torch.Generator(device="cpu").manual_seed(3)
    v = torch.zeros_like(p)
    bright = torch.from_numpy((theta[:, 11:14] > 0.5).all(axis=1)).cuda()
    v[11:14] = torch.randn((3, p.shape[1]), generator=gen).cuda() * bright
    assert int(bright.sum()) > 10000
    h = 1e-2
    fd = (loss(p + h * v) - loss(p - h * v)) / (2 * h)
    an = float((g * v).sum().item())
    print("rgb direction: fd", fd, "analytic", an, flush=True)
    assert an == pytest.approx(fd, rel=2e-3)
    # SH coefficients of the same splats: linear too, through the basis. The
    # float32 forwards' rounding leaves about 3e-10 of noise in a central
    # difference of this loss at h = 1e-2, so the direction is scaled to keep
    # its derivative three orders above that (its colour changes, about
    # 0.02, stay clear of the clamp at 0 for these bright splats).
    v = torch.zeros_like(p)
    v[14:59] = torch.randn((45, p.shape[1]), generator=gen).cuda() * bright * 0.5
    fd = (loss(p + h * v) - loss(p - h * v)) / (2 * h)
    an = float((g * v).sum().item())
    print("sh direction: fd", fd, "analytic", an, flush=True)
    assert an == pytest.approx(fd, rel=2e-3)
    # Opacity, scales and means move the cutoff ellipse, where alpha drops
    # from minAlpha to 0: a jump the adjoint (the derivative with the drawn
    # pixels fixed, held to the FP64 reference above) does not see, so
    # finite differences of the whole image differ by that boundary term.
    v = torch.zeros_like(p)
    v[10] = torch.randn(p.shape[1], generator=gen).cuda()
    h = 1e-3
    fd = (loss(p + h * v) - loss(p - h * v)) / (2 * h)
    an = float((g * v).sum().item())
    print("opacity direction: fd", fd, "analytic", an, flush=True)
    assert an == pytest.approx(fd, rel=0.1)


def test_full_size_gradient_matches_the_fp64_adjoint(cuda):
    # 262,144 splats at 512 x 512: the GPU adjoint of the L2 loss against the
    # FP64 host adjoint (oracle/raster.c, binned) of the same loss, per kind
    # of parameter: over 512 random parameters of each kind and over all of them
    theta, target_theta, rig = D.workload()
    r = O.raster(rig.width, rig.height, rig.eyes[0].view, rig.eyes[0].projection, sh_degree=3)
    target, _, _ = O.diff_render(target_theta, r)
    loss_ref, g_ref = O.diff_grad(theta, r, target)
    loss, g, _, rast = gpu_loss_grad(theta, rig, target)
    print(f"loss {loss:.6e} vs FP64 {loss_ref:.6e}; stats {rast.stats()}", flush=True)
    assert loss == pytest.approx(loss_ref, rel=1e-4)
    rng = np.random.default_rng(512)
    for name, sl in KINDS.items():
        a, b = g[:, sl], g_ref[:, sl]
        flat = np.nonzero(b.reshape(-1))[0]
        pick = rng.choice(flat, size=512, replace=False)
        ra, rb = a.reshape(-1)[pick], b.reshape(-1)[pick]
        rel512 = np.linalg.norm(ra - rb) / np.linalg.norm(rb)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{name}: 512 random parameters {rel512:.2e}, all {rel:.2e}", flush=True)
        assert rel512 < 1e-3 and rel < 1e-3, name

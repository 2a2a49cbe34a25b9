"""The oracle's differentiable compositor and adjoint, pinned (no GPU).

Pins: closed forms of the compositing (one splat at its centre, two in
order, the transmittance stop), the analytic gradient against central
finite differences in FP64 (h = 1e-4) on 512 parameters, with the
contributor lists frozen (the loss is then smooth, so the two must agree
to 1e-3 relative) and without (agreement wherever no discrete decision
moves), and the SH basis gradient against finite differences.
"""
import numpy as np
import pytest

from oracle import oracle as O
from radix_amd import camera as cam
from radix_amd import diff as D
from radix_amd import workloads as W


def small_scene(n=400, size=64, seed=21, log_scale=-3.2, sh=3):
    spec = W.CloudSpec(seed=seed, count=n, sh_degree=3, centers=0, offset=(0.0, 0.0, -3.0),
                       log_scale_mean=log_scale, log_scale_std=0.3)
    theta = D.theta_from_planes(W.generate(spec))
    target_theta = D.perturbed(theta, seed + 5)
    rig = cam.mono(np.eye(4), 60.0, size, size)
    r = O.raster(size, size, rig.eyes[0].view, rig.eyes[0].projection, sh_degree=sh)
    target, _, _ = O.diff_render(target_theta, r)
    return theta, r, target


def one_splat(center, scale=0.05, logit=2.0, rgb=(0.8, 0.4, 0.2)):
    t = np.zeros((1, 59))
    t[0, 0:3] = center
    t[0, 3:6] = np.log(scale)
    t[0, 9] = 1.0
    t[0, 10] = logit
    t[0, 11:14] = rgb
    return t


def test_one_splat_at_its_centre():
    rig = cam.mono(np.eye(4), 60.0, 33, 33)
    r = O.raster(33, 33, rig.eyes[0].view, rig.eyes[0].projection, sh_degree=0)
    t = one_splat([0.0, 0.0, -3.0])
    img, alpha, drawn = O.diff_render(t, r)
    assert drawn == 1
    # the centre projects to pixel (16.5, 16.5): the middle pixel's centre
    fy = 0.5 * 33 * rig.eyes[0].projection[1, 1]
    var = (0.05 * fy / 3.0) ** 2 + 0.3
    op = 1.0 / (1.0 + np.exp(-2.0))
    opc = op * np.sqrt((var - 0.3) ** 2 / var ** 2)
    assert alpha[16, 16] == pytest.approx(opc, rel=1e-12)
    assert img[:, 16, 16] == pytest.approx(np.array([0.8, 0.4, 0.2]) * opc, rel=1e-12)
    # one pixel right: exp(-z^2 / 2) with z^2 = 1 / var
    assert alpha[16, 17] == pytest.approx(opc * np.exp(-0.5 / var), rel=1e-12)


def test_two_splats_over_and_the_stop():
    rig = cam.mono(np.eye(4), 60.0, 9, 9)
    r = O.raster(9, 9, rig.eyes[0].view, rig.eyes[0].projection, sh_degree=0)
    front = one_splat([0.0, 0.0, -2.0], scale=0.3, logit=1.0, rgb=(1.0, 0.0, 0.0))
    back = one_splat([0.0, 0.0, -4.0], scale=0.6, logit=0.5, rgb=(0.0, 0.0, 1.0))
    both = np.concatenate([back, front])  # input order does not matter: depth does
    img, alpha, _ = O.diff_render(both, r)
    i1, a1, _ = O.diff_render(front, r)
    i2, a2, _ = O.diff_render(back, r)
    assert alpha[4, 4] == pytest.approx(a1[4, 4] + (1 - a1[4, 4]) * a2[4, 4], rel=1e-12)
    assert img[:, 4, 4] == pytest.approx(i1[:, 4, 4] + (1 - a1[4, 4]) * i2[:, 4, 4], rel=1e-12)
    # 40 opaque layers: compositing stops before T would drop below 1e-4
    layers = np.concatenate([one_splat([0.0, 0.0, -2.0 - 0.1 * k], scale=0.5, logit=4.0) for k in range(40)])
    _, al, _ = O.diff_render(layers, r)
    a = 1.0 / (1.0 + np.exp(-4.0))
    assert 1.0 - al[4, 4] >= 1e-4
    assert 1.0 - al[4, 4] < 1e-4 / (1.0 - a * 0.98)


def test_sh_basis_gradient():
    rng = np.random.default_rng(3)
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        g = np.zeros((24, 3))
        O.lib().orc_sh_basis_grad(O.dptr(d), O.dptr(g))
        for e in range(3):
            dp, dm = d.copy(), d.copy()
            dp[e] += 1e-6
            dm[e] -= 1e-6
            bp, bm = np.zeros(24), np.zeros(24)
            O.lib().orc_sh_basis(O.dptr(dp), O.dptr(bp))
            O.lib().orc_sh_basis(O.dptr(dm), O.dptr(bm))
            assert np.allclose((bp - bm) / 2e-6, g[:, e], rtol=1e-6, atol=1e-8)


def test_gradient_equals_frozen_finite_differences_on_512_params():
    theta, r, target = small_scene()
    loss, g = O.diff_grad(theta, r, target)
    assert loss > 0
    fr = O.Frozen(theta, r, target)
    assert fr.loss(theta) == pytest.approx(loss, rel=1e-14)
    rng = np.random.default_rng(12)
    drawn = np.nonzero(np.abs(g).sum(axis=1) > 0)[0]
    assert drawn.size > 200
    # every parameter kind, then random ones: 512 in all
    picks = [(rng.choice(drawn), k) for k in range(59)]
    picks += [(rng.choice(drawn), int(rng.integers(0, 59))) for _ in range(512 - len(picks))]
    h = 1e-4
    scale = np.abs(g).max()
    worst = 0.0
    for i, k in picks:
        tp, tm = theta.copy(), theta.copy()
        tp[i, k] += h
        tm[i, k] -= h
        fd = (fr.loss(tp) - fr.loss(tm)) / (2 * h)
        err = abs(fd - g[i, k]) / max(abs(fd), abs(g[i, k]), 1e-6 * scale)
        worst = max(worst, err)
        assert err <= 1e-3, (i, k, fd, g[i, k])
    print("worst relative error", worst)


def test_gradient_matches_unfrozen_differences_away_from_decisions():
    theta, r, target = small_scene(n=300, seed=33)
    _, g = O.diff_grad(theta, r, target)
    rng = np.random.default_rng(4)
    drawn = np.nonzero(np.abs(g).sum(axis=1) > 0)[0]
    h = 1e-6
    agree = 0
    trials = 120
    scale = np.abs(g).max()
    for _ in range(trials):
        i, k = rng.choice(drawn), int(rng.integers(0, 59))
        tp, tm = theta.copy(), theta.copy()
        tp[i, k] += h
        tm[i, k] -= h
        fd = (O.diff_loss(tp, r, target) - O.diff_loss(tm, r, target)) / (2 * h)
        if abs(fd - g[i, k]) <= 1e-3 * max(abs(fd), abs(g[i, k]), 1e-6 * scale):
            agree += 1
    assert agree >= 0.9 * trials, agree


def test_binning_changes_nothing():
    # the tile lists hold every splat that can reach a pixel, in draw order:
    # image, loss and gradient equal the walk over every splat bit for bit
    theta, r, target = small_scene(n=300, seed=71, log_scale=-2.0)
    L = O.lib()
    try:
        L.orc_raster_binning(0)
        img0, a0, d0 = O.diff_render(theta, r)
        loss0, g0 = O.diff_grad(theta, r, target)
    finally:
        L.orc_raster_binning(1)
    img1, a1, d1 = O.diff_render(theta, r)
    loss1, g1 = O.diff_grad(theta, r, target)
    assert d0 == d1 and np.array_equal(img0, img1) and np.array_equal(a0, a1)
    assert loss0 == loss1 and np.array_equal(g0, g1)
    assert np.count_nonzero(img1) > 1000

"""The autograd harness's guards (no GPU): composite() refuses the backward of
a forward whose state a later forward on the same Rasterizer replaced, and
the parameters it saves catch in-place changes before backward."""
import pytest
import torch

from radix_amd import diff as D


class FakeRasterizer:
    """The Rasterizer's interface on the CPU: the image is the first three parameter planes."""

    def __init__(self):
        self.generation = 0
        self.calls = []

    def forward(self, p, rig, prm):
        self.generation += 1
        return p.detach()[:3].clone(), None

    def backward(self, grad_image, generation=None):
        D.check_generation(self, generation)
        g = torch.zeros(59, grad_image.shape[-1])
        g[:3] = grad_image
        self.calls.append(generation)
        return g


def test_one_forward_one_backward():
    r = FakeRasterizer()
    p = torch.ones(59, 5, requires_grad=True)
    D.composite(p, None, None, r).sum().backward()
    assert r.calls == [1] and float(p.grad[:3].sum()) == 15.0


def test_second_forward_before_backward_raises():
    r = FakeRasterizer()
    p = torch.ones(59, 5, requires_grad=True)
    a = D.composite(p, None, None, r)
    b = D.composite(p * 2, None, None, r)
    with pytest.raises(RuntimeError, match="replaced its state"):
        (a.sum() + b.sum()).backward()


def test_in_place_change_before_backward_raises():
    r = FakeRasterizer()
    p = torch.ones(59, 5, requires_grad=True)
    q = p * 1.0
    img = D.composite(q, None, None, r)
    with torch.no_grad():
        q.add_(1.0)
    with pytest.raises(RuntimeError, match="modified by an inplace operation"):
        img.sum().backward()

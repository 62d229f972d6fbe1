"""Host logic of api_diff.rollout on a stub scene (CPU tensors, no device):
a backward whose scene ran a later forward raises instead of returning that
later rollout's gradients; one forward per backward works."""
import numpy as np
import pytest
import torch

from diff_fem_amd import api_diff


class StubScene:
    """Counts tapes like engine.Scene; the 'rollout' is x_T = 2 x0."""

    def __init__(self):
        self.tape_id = 0
        self._E_cached = np.ones(3)
        self.last = None

    def set_material(self, E):
        self.tape_id += 1

    def forward_device(self, x0, v0, steps, act=None, record=False):
        self.tape_id += 1
        self.last = x0.detach().clone()
        return 2 * x0.detach(), v0.detach().clone()

    def backward_device(self, gx, gv):
        return dict(x0=2 * gx, v0=gv, E=torch.zeros(3, dtype=torch.float64),
                    act=torch.zeros((gx.shape[0], 1, 1), dtype=torch.float64))


def _inputs():
    x0 = torch.ones((2, 6), dtype=torch.float64, requires_grad=True)
    v0 = torch.zeros((2, 6), dtype=torch.float64, requires_grad=True)
    return x0, v0, torch.ones(3, dtype=torch.float64), torch.zeros((2, 1, 1), dtype=torch.float64)


def test_single_rollout_backward():
    sc = StubScene()
    x0, v0, E, act = _inputs()
    xT, vT = api_diff.rollout(x0, v0, E, act, sc, 1)
    xT.sum().backward()
    assert torch.equal(x0.grad, torch.full_like(x0, 2.0))


def test_backward_after_a_later_forward_raises():
    sc = StubScene()
    x0, v0, E, act = _inputs()
    xa, _ = api_diff.rollout(x0, v0, E, act, sc, 1)
    xb, _ = api_diff.rollout(x0, v0, E, act, sc, 1)
    with pytest.raises(RuntimeError, match="tape is gone"):
        (xa.sum() + xb.sum()).backward()

"""Host wiring of the torch.autograd wrapper (no GPU): forward returns the engine's E_SIC and hands it the FOD
positions, backward scales the engine's adjoint gradient by the incoming gradient."""
import numpy as np
import torch

from publicrelease_2020_amd.autograd import fod_energy


class _Engine:
    def __init__(self):
        self.fods = {}

    def set_fods(self, s, f):
        self.fods[s] = np.array(f)

    def evaluate(self):
        return float((self.fods[1] ** 2).sum() + 3.0 * self.fods[2].sum())

    def fod_gradient(self, s):
        return 2.0 * self.fods[1] if s == 1 else np.full_like(self.fods[2], 3.0)


def test_forward_backward_wiring():
    up = torch.tensor([[0.1, 0.2, 0.3], [1.0, -1.0, 0.5]], dtype=torch.float64, requires_grad=True)
    dn = torch.tensor([[0.4, 0.0, -0.2]], dtype=torch.float64, requires_grad=True)
    eng = _Engine()
    e = fod_energy(eng, up, dn)
    assert abs(e.item() - ((up.detach() ** 2).sum() + 3.0 * dn.detach().sum()).item()) < 1e-15
    (2.5 * e).backward()
    np.testing.assert_allclose(up.grad.numpy(), 2.5 * 2.0 * up.detach().numpy())
    np.testing.assert_allclose(dn.grad.numpy(), np.full((1, 3), 7.5))

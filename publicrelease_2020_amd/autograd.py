"""E_SIC as a differentiable function of the FOD positions (torch.autograd), for harness code.

`fod_energy(ctx, fod_up, fod_dn)` returns E_SIC (a 0-d float64 tensor) evaluated by the HIP library for the
given FOD positions (n_up x 3 and n_dn x 3 float64 CPU tensors); its backward pass returns the library's
adjoint gradient dE_SIC/da (flosic_fod_gradient: dE/dTMAT = 2 SIC^T TMAT back-propagated through the Lowdin
step and the Fermi-orbital construction, with SIC the exact mesh derivative of the discrete energy), which for
LSDA equals the FRMORB2 FOD force DEBDAX (frmorb2.F:589-729); for PBE DEBDAX uses the Hessian-form GGA potential
and differs from the derivative of the energy.
The evaluation runs on the GPU inside the library; the tensors exchanged with torch are small host arrays,
so the function composes with any torch optimiser over FOD positions (the role of FOD_OPT/FOD_LBFGS,
electronic_geometry_lbfgs.f:143-146).
"""
from __future__ import annotations

import numpy as np
import torch

from .flosic import Flosic


class FodEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fod_up: torch.Tensor, fod_dn: torch.Tensor, engine: Flosic) -> torch.Tensor:
        engine.set_fods(1, fod_up.detach().cpu().numpy().astype(np.float64))
        engine.set_fods(2, fod_dn.detach().cpu().numpy().astype(np.float64))
        e = engine.evaluate()
        g_up = torch.from_numpy(engine.fod_gradient(1)).to(fod_up.dtype)
        g_dn = torch.from_numpy(engine.fod_gradient(2)).to(fod_dn.dtype)
        ctx.save_for_backward(g_up, g_dn)
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        g_up, g_dn = ctx.saved_tensors
        return grad_out * g_up, grad_out * g_dn, None


def fod_energy(engine: Flosic, fod_up: torch.Tensor, fod_dn: torch.Tensor) -> torch.Tensor:
    """E_SIC(fod_up, fod_dn) with the adjoint FOD gradient as its backward pass."""
    return FodEnergy.apply(fod_up, fod_dn, engine)

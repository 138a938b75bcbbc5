"""E_SIC(FOD positions) through torch.autograd: the backward pass is the library's FOD gradient.

Central differences of the GPU energy (h = 1e-4 bohr, FP64) agree with it to 1e-5 Ha/bohr for LSDA and PBE
(the residual is the mesh quadrature of the Coulomb matrix elements; the oracle shows the same 2e-6..3e-6 on this
80x302 mesh).  For PBE the gradient takes the exact mesh derivative of the discrete E_xc (e_rho, e_g of the
energy density), not the FRMORB2 force with the Hessian-form GGA potential, which misses the same differences by
up to 8e-3 Ha/bohr here; the gradient is also pinned to the oracle's to 1e-9."""
import numpy as np
import pytest
import torch

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("xc", [S.XC_LSDA, S.XC_PBE], ids=["LSDA", "PBE"])
def test_autograd_gradient_matches_fd(xc):
    from publicrelease_2020_amd import flosic as F
    from publicrelease_2020_amd.autograd import fod_energy

    sy = S.make_synthetic(3, 3, 7, xc=xc, nrad=80, nang=302, box=4.5)
    eng = F.Flosic(0)
    eng.load_system(sy)
    up = torch.tensor(sy.fod[0], dtype=torch.float64, requires_grad=True)
    dn = torch.tensor(sy.fod[1], dtype=torch.float64, requires_grad=True)
    e = fod_energy(eng, up, dn)
    e.backward()
    np.testing.assert_allclose(up.grad.numpy(), eng.fod_gradient(1), rtol=0, atol=0)
    np.testing.assert_allclose(dn.grad.numpy(), eng.fod_gradient(2), rtol=0, atol=0)
    if xc == S.XC_LSDA:  # LSDA: the adjoint and the FRMORB2 formula are one number
        np.testing.assert_allclose(up.grad.numpy(), eng.fod_forces(1), atol=1e-10)
    h = 1e-4
    if xc == S.XC_PBE:
        from oracle import oracle as O
        ref = O.evaluate(sy, want_vab=False)
        np.testing.assert_allclose(up.grad.numpy(), ref["spins"][0]["grad"], atol=1e-9)
        np.testing.assert_allclose(dn.grad.numpy(), ref["spins"][1]["grad"], atol=1e-9)
    for (s, i, x) in [(0, 0, 0), (1, 2, 1), (0, 1, 2)]:
        fods = [up.detach().clone(), dn.detach().clone()]
        fods[s][i, x] += h
        ep = float(fod_energy(eng, *fods))
        fods[s][i, x] -= 2 * h
        em = float(fod_energy(eng, *fods))
        fd = (ep - em) / (2 * h)
        g = (up.grad if s == 0 else dn.grad)[i, x].item()
        print(f"xc={xc} coord {(s, i, x)}: fd {fd:.9f} adjoint {g:.9f}", flush=True)
        assert abs(fd - g) < 1e-5

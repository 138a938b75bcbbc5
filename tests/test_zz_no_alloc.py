"""flosic_apotnl_dft_noalloc (the allocation-free variant of flosic_apotnl_dft) against flosic_apotnl_dft on the same
evaluation, through the C ABI: identical potentials and energies bit for bit (same kernels, same ordered sums), on
two calls in a row with the second on a smaller mesh (the shared scratch keeps the first call's contents), LSDA and
PBE; and its error paths (no current evaluation, mesh-block mode)."""
import numpy as np
import pytest

from publicrelease_2020_amd import system as S

pytestmark = pytest.mark.gpu


def _check(sysm):
    from publicrelease_2020_amd import flosic as F
    ctx = F.Flosic(0)
    try:
        ctx.load_system(sysm)
        with pytest.raises(F.FlosicError) as e:
            ctx.apotnl_dft_noalloc()
        assert e.value.status == F.ESTATE
        ctx.evaluate()
        pot, en = ctx.apotnl_dft()
        pot2, en2 = ctx.apotnl_dft_noalloc()
        print(f"{sysm.name}: M={sysm.mesh.npts} xc={sysm.xc} E_x {en['ex']:.12f} (no-alloc {en2['ex']:.12f})",
              flush=True)
        assert np.array_equal(pot, pot2)
        assert en == en2
    finally:
        ctx.close()


def test_noalloc_matches_allocating_entry_point_twice():
    _check(S.make_synthetic(4, 4, 5, S.XC_LSDA, nrad=40, nang=110, box=5.0))  # 17 600 points
    _check(S.make_synthetic(3, 3, 7, S.XC_PBE, nrad=24, nang=50, box=5.0))    # 3 600 points, then smaller


def test_noalloc_rejects_mesh_block_mode():
    from publicrelease_2020_amd import flosic as F
    sysm = S.make_synthetic(3, 3, 7, S.XC_LSDA, nrad=24, nang=50, box=5.0)
    ctx = F.Flosic(0)
    try:
        ctx.load_system(sysm)
        ctx.set_mesh_block(1024)
        ctx.evaluate()
        with pytest.raises(F.FlosicError) as e:
            ctx.apotnl_dft_noalloc()
        assert e.value.status == F.EINVAL
    finally:
        ctx.close()

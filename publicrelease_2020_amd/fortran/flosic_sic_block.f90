! Fortran host side of the SIC block of the SCF iteration (the shape of main.F:616-680) over the C ABI:
! per spin the KS coefficients and FODs are handed over, one evaluation replaces the FRMORB2(0,0) /
! FRMIORB / APOTNL_SIC loop / SICLAGM / FRMORB2(-ISPN,-1) sequence, and RESULTS, SIC and DEBDAX come back
! in the reference's layouts.  A nonzero library status is returned to the caller (the reference would
! call STOPIT there).
module flosic_sic_block
  use iso_c_binding
  use flosic_hip
  implicit none
contains
  subroutine flosic_sic_evaluate(ctx, nbas, nwf, nocc, coef1, coef2, bfrm1, bfrm2, results1, results2, &
                                 sic1, sic2, debdax1, debdax2, zsic, status)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: nbas, nwf(2), nocc(2)
    real(c_double), intent(in) :: coef1(nbas, nwf(1)), coef2(nbas, nwf(2))
    real(c_double), intent(in) :: bfrm1(3, nocc(1)), bfrm2(3, nocc(2))
    real(c_double), intent(out) :: results1(13, nocc(1)), results2(13, nocc(2))
    real(c_double), intent(out) :: sic1(nwf(1), nocc(1)), sic2(nwf(2), nocc(2))
    real(c_double), intent(out) :: debdax1(3, nocc(1)), debdax2(3, nocc(2))
    real(c_double), intent(out) :: zsic
    integer(c_int32_t), intent(out) :: status
    real(c_double) :: z
    status = flosic_set_orbitals(ctx, 1_c_int32_t, nbas, nwf(1), nocc(1), coef1)
    if (status /= FLOSIC_OK) return
    status = flosic_set_orbitals(ctx, 2_c_int32_t, nbas, nwf(2), nocc(2), coef2)
    if (status /= FLOSIC_OK) return
    status = flosic_set_fods(ctx, 1_c_int32_t, nocc(1), bfrm1)
    if (status /= FLOSIC_OK) return
    status = flosic_set_fods(ctx, 2_c_int32_t, nocc(2), bfrm2)
    if (status /= FLOSIC_OK) return
    status = flosic_evaluate(ctx, zsic)
    if (status /= FLOSIC_OK) return
    status = flosic_apotnl_sic(ctx, 1_c_int32_t, results1, z)
    if (status == FLOSIC_OK) status = flosic_apotnl_sic(ctx, 2_c_int32_t, results2, z)
    if (status == FLOSIC_OK) status = flosic_siclagm(ctx, 1_c_int32_t, sic1)
    if (status == FLOSIC_OK) status = flosic_siclagm(ctx, 2_c_int32_t, sic2)
    if (status == FLOSIC_OK) status = flosic_fod_forces(ctx, 1_c_int32_t, debdax1)
    if (status == FLOSIC_OK) status = flosic_fod_forces(ctx, 2_c_int32_t, debdax2)
  end subroutine flosic_sic_evaluate
end module flosic_sic_block

! flosic_sic_drive INFILE OUTFILE: reads one system (basis, mesh, options, KS orbitals and FODs of both
! spins) from an unformatted stream file, runs flosic_sic_evaluate on HIP device 0 and writes E_SIC,
! RESULTS, SIC and DEBDAX of both spins, then the force step's DERSIC (SICLAG_DER) and, for NSPN = 2, the DFT-side
! POTOUT(NMSH,2) and its energies (APOTNL_DFT).  Exit status 0 on success, 10 + library status otherwise.
program flosic_sic_drive
  use iso_c_binding
  use flosic_hip
  use flosic_sic_block
  implicit none
  character(len=1024) :: fin, fout
  integer(c_int32_t) :: natom, nbare, ncoef, nspn, xc, prec, nbas, nwf(2), nocc(2), status, s
  integer(c_int64_t) :: nmsh
  real(c_double), allocatable :: xyz(:), alpha(:), con(:), rmsh(:), wmsh(:), c1(:, :), c2(:, :), f1(:, :), f2(:, :)
  real(c_double), allocatable :: r1(:, :), r2(:, :), s1(:, :), s2(:, :), d1(:, :), d2(:, :), dpk(:, :, :), potdft(:)
  ! SICMAT's DERSIC(3,MAX_OCC,MX_CNT) layout (modules_sic.F90:44, MAX_OCC = 1500): the packed library result is
  ! copied into its leading columns, as a binding into the reference's modules must do
  integer, parameter :: MAX_OCC = 1500
  real(c_double), allocatable :: dersic(:, :, :)
  integer :: nf
  real(c_double) :: edft(5)
  integer(c_int32_t), allocatable :: bare_off(:), ncon(:), con_off(:)
  real(c_double) :: zsic
  type(c_ptr) :: ctx
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) natom, nbare, ncoef, nmsh, nspn, xc, prec, nbas, nwf, nocc
  allocate(xyz(3 * natom), bare_off(natom + 1), alpha(nbare), ncon(3 * natom), con_off(3 * natom), con(ncoef))
  allocate(rmsh(3 * nmsh), wmsh(nmsh), c1(nbas, nwf(1)), c2(nbas, nwf(2)), f1(3, nocc(1)), f2(3, nocc(2)))
  read(10) xyz, bare_off, alpha, ncon, con_off, con, rmsh, wmsh, c1, f1, c2, f2
  close(10)
  allocate(r1(13, nocc(1)), r2(13, nocc(2)), s1(nwf(1), nocc(1)), s2(nwf(2), nocc(2)), d1(3, nocc(1)), d2(3, nocc(2)))
  status = flosic_create(0_c_int32_t, ctx)
  if (status /= FLOSIC_OK) call fail(status)
  status = flosic_set_basis(ctx, natom, xyz, bare_off, alpha, ncon, con_off, con)
  if (status /= FLOSIC_OK) call fail(status)
  status = flosic_set_mesh(ctx, nmsh, rmsh, wmsh)
  if (status /= FLOSIC_OK) call fail(status)
  status = flosic_set_options(ctx, nspn, xc, prec)
  if (status /= FLOSIC_OK) call fail(status)
  call flosic_sic_evaluate(ctx, nbas, nwf, nocc, c1, c2, f1, f2, r1, r2, s1, s2, d1, d2, zsic, status)
  if (status /= FLOSIC_OK) call fail(status)
  nf = nocc(1) + nocc(2)
  allocate(dpk(3, nf, natom), dersic(3, MAX_OCC, natom), potdft(2 * nmsh))
  dersic = 0.0d0
  status = flosic_siclag_der(ctx, dpk)
  if (status /= FLOSIC_OK) call fail(status)
  dersic(:, 1:nf, 1:natom) = dpk
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) zsic, r1, r2, s1, s2, d1, d2, dersic(:, 1:nf, 1:natom)
  if (nspn == 2) then
    status = flosic_apotnl_dft(ctx, potdft, edft)
    if (status /= FLOSIC_OK) call fail(status)
    write(11) potdft, edft
  end if
  close(11)
  s = flosic_destroy(ctx)
contains
  subroutine fail(st)
    integer(c_int32_t), intent(in) :: st
    write(0, '(a,i0)') 'flosic status ', st
    call exit(10 + st)
  end subroutine fail
end program flosic_sic_drive

! Drop-in replacements for the reference's per-Fermi-orbital SIC routines, with the reference's own names and
! argument lists, backed by the HIP library through the ISO_C_BINDING module flosic_hip:
!
!   FRMORB2(LSXX,LFM)           frmorb2.F:10    LSXX = 0: Fermi orbitals, Lowdin step and FLOs of both spins (one
!                                               library evaluation when the orbitals or FODs changed); FMAT into
!                                               SICMAT.  LSXX = -ISPN, LFM = -1: DEBDAX(:,1:NFRM,ISPN) into FRM.
!   FRMIORB(NXBG,NXED)          frmiorb.f:5     TMAT (LOCORB) and NFRM (MOCORB) of both spins from the evaluation
!   APOTNL_SIC(TOTQNUM,LSPX,IORBX) apotnl_sic.F:31  RESULTS(1:13,IORBX,LSPX) into FRM; TOTQNUM = the FLO's charge
!   SICLAGM(LSPN)               frmorb_mpi.F:2  SIC(1:NWFS(LSPN),1:NFRM(LSPN),LSPN) into SICMAT
!   SICLMSLV(...)               frmorb_mpi.F:184 the MPI worker half of SICLAGM; never reached (SICLAGM above does
!                               the whole product), stops if called
!   DIAGV_FO(IPREONLY)          diagv_fo.F:5    SCFSICW's occupied-virtual rotation (PHIRES of DIAGV1 from HAM of
!                                               FOR_DIAG1); DIAGV_FO_UNI, diagv_fo.F:157, its unified-Hamiltonian form
!
! They read what the replaced routines read from the reference's modules -- COMMON2 (RIDT, IFUIDT, NIDENT, NCNT,
! BFALP, BFCON, N_BARE, N_CON, NSPN, IGGA), MESH1 (NMSH, RMSH, WMSH), COMMON5 (PSI_COEF, NWFS), COMMON8 (N_REP,
! NS_TOT), FRM (BFRM, LFRM) -- and write what the reference's consumers read back (FRM RESULTS and DEBDAX, SICMAT
! SIC and FMAT, LOCORB TMAT, MOCORB NFRM).  A nonzero library status prints the library's message and calls the
! reference's STOPIT.  Point group NONE only (every identity member one atom, one representation), as FLO-SIC runs;
! LDA-PW91 or GGA-PBE.  The HIP device is $FLOSIC_DEVICE (default 0), the precision $FLOSIC_PRECISION (0 FP64,
! 1 FP32-MFMA, 2 split-BF16x3; default 0).
!
! Compiled against the reference's module files and linked in place of apotnl_sic.o, frmorb2.o, frmiorb.o,
! frmorb_mpi.o and diagv_fo.o (oracle/build_nrlmol.py --dropin).
module flosic_dropin
  use iso_c_binding
  use flosic_hip
  implicit none
  type(c_ptr), save :: fctx = c_null_ptr
  logical, save :: have_eval = .false.
  integer, save :: nbas_s = 0, nwf_s(2) = 0, nfod_s(2) = 0
  real(c_double), allocatable, save :: coef_s(:, :, :), fod_s(:, :, :)
  real(c_double), save :: zsic_s = 0.0d0

contains

  subroutine dropin_check(st, what)
    integer(c_int32_t), intent(in) :: st
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (st == FLOSIC_OK) return
    write(6, '(a,a,a,i0)') ' FLOSIC_HIP: ', what, ' failed with status ', st
    if (c_associated(fctx)) then
      call c_f_pointer(flosic_last_error(fctx), msg, [512])
      k = 1
      do while (k <= 512)
        if (msg(k) == c_null_char) exit
        k = k + 1
      end do
      write(6, '(a,512a)') ' FLOSIC_HIP: ', msg(1:k - 1)
    end if
    call stopit
  end subroutine dropin_check

  ! The reference's basis (COMMON2) as flosic_set_basis takes it: atoms in identity-member order; per atom its bare
  ! exponents BFALP(1:N_BARE), per l = 0, 1, 2 the N_CON contractions, each a coefficient vector over the bare
  ! exponents (BFCON(1:N_BARE, ICON, l+1, IFNCT) is contiguous, so the (contraction, bare) block is copied as is).
  subroutine dropin_gather_basis(natom, xyz, bare_off, alpha, ncon, con_off, con, nbas)
    use common2, only : RIDT, IFUIDT, NIDENT, NCNT, BFALP, BFCON, N_BARE, N_CON
    integer, intent(out) :: natom, nbas
    real(c_double), allocatable, intent(out) :: xyz(:), alpha(:), con(:)
    integer(c_int32_t), allocatable, intent(out) :: bare_off(:), ncon(:), con_off(:)
    integer :: a, ifn, l, c, nb, nalpha, ncoef, k
    integer, parameter :: ncart(0:2) = [1, 3, 6]
    if (NCNT /= NIDENT) then
      write(6, *) 'FLOSIC_HIP: point group NONE only (NCNT, NIDENT =', NCNT, NIDENT, ')'
      call stopit
    end if
    natom = NIDENT
    nalpha = 0
    ncoef = 0
    nbas = 0
    do a = 1, natom
      ifn = IFUIDT(a)
      nalpha = nalpha + N_BARE(ifn)
      do l = 0, 2
        ncoef = ncoef + N_CON(l + 1, ifn) * N_BARE(ifn)
        nbas = nbas + ncart(l) * N_CON(l + 1, ifn)
      end do
    end do
    allocate(xyz(3 * natom), alpha(nalpha), con(max(ncoef, 1)), bare_off(natom + 1), ncon(3 * natom), con_off(3 * natom))
    bare_off(1) = 0
    k = 0
    do a = 1, natom
      ifn = IFUIDT(a)
      nb = N_BARE(ifn)
      xyz(3 * a - 2:3 * a) = RIDT(1:3, a)
      alpha(bare_off(a) + 1:bare_off(a) + nb) = BFALP(1:nb, ifn)
      bare_off(a + 1) = bare_off(a) + nb
      do l = 0, 2
        ncon(3 * (a - 1) + l + 1) = N_CON(l + 1, ifn)
        con_off(3 * (a - 1) + l + 1) = k
        do c = 1, N_CON(l + 1, ifn)
          con(k + 1:k + nb) = BFCON(1:nb, c, l + 1, ifn)
          k = k + nb
        end do
      end do
    end do
  end subroutine dropin_gather_basis

  ! PSI_COEF rows (symmetry-adapted order: per site l, contraction, component -- UNRAVEL, unravel.f:51-89, with the
  ! identity U_MAT of point group NONE) to the local order of GORBDRV (per site l, component, contraction).
  subroutine dropin_salc_to_local(nbas, ncols, src, dst)
    use common2, only : IFUIDT, NIDENT, N_CON
    integer, intent(in) :: nbas, ncols
    real(c_double), intent(in) :: src(:, :)
    real(c_double), intent(out) :: dst(nbas, ncols)
    integer :: a, ifn, l, c, m, ks, kl, base
    integer, parameter :: ncart(0:2) = [1, 3, 6]
    ks = 0
    base = 0
    do a = 1, NIDENT
      ifn = IFUIDT(a)
      do l = 0, 2
        do c = 1, N_CON(l + 1, ifn)
          do m = 1, ncart(l)
            ks = ks + 1
            kl = base + (m - 1) * N_CON(l + 1, ifn) + c
            dst(kl, 1:ncols) = src(ks, 1:ncols)
          end do
        end do
        base = base + ncart(l) * N_CON(l + 1, ifn)
      end do
    end do
  end subroutine dropin_salc_to_local

  ! Create the context and hand over the basis, mesh and options (once: the SCF keeps them fixed).
  subroutine dropin_setup()
    use common2, only : NSPN, IGGA
    use mesh1, only : NMSH, RMSH, WMSH
    integer :: natom, nbas, dev, prec, ios
    real(c_double), allocatable :: xyz(:), alpha(:), con(:)
    integer(c_int32_t), allocatable :: bare_off(:), ncon(:), con_off(:)
    character(len=32) :: env
    integer(c_int32_t) :: xc
    if (c_associated(fctx)) return
    dev = 0
    prec = FLOSIC_PREC_FP64
    call get_environment_variable('FLOSIC_DEVICE', env, status=ios)
    if (ios == 0) read(env, *) dev
    call get_environment_variable('FLOSIC_PRECISION', env, status=ios)
    if (ios == 0) read(env, *) prec
    call dropin_gather_basis(natom, xyz, bare_off, alpha, ncon, con_off, con, nbas)
    call dropin_check(flosic_create(int(dev, c_int32_t), fctx), 'flosic_create')
    call dropin_check(flosic_set_basis(fctx, int(natom, c_int32_t), xyz, bare_off, alpha, ncon, con_off, con), &
                      'flosic_set_basis')
    call dropin_check(flosic_set_mesh(fctx, int(NMSH, c_int64_t), RMSH, WMSH), 'flosic_set_mesh')
    xc = FLOSIC_XC_LSDA
    if (IGGA(1) > 0 .or. IGGA(2) > 0) xc = FLOSIC_XC_PBE
    call dropin_check(flosic_set_options(fctx, int(NSPN, c_int32_t), xc, int(prec, c_int32_t)), 'flosic_set_options')
    nbas_s = nbas
    write(6, '(a,i0,a,i0,a,i0,a,i0)') ' FLOSIC_HIP: context on device ', dev, ': ', natom, ' atoms, ', nbas, &
        ' basis functions, mesh ', NMSH
  end subroutine dropin_setup

  ! With $FLOSIC_DROPIN_DUMP set, the arrays handed to the library (basis as flosic_set_basis takes it, the mesh size,
  ! per spin the local-order orbital coefficients and the FODs) are written to that file, unformatted stream, before
  ! any library call: a check of the gathering against a reader of the same run directory, on any machine.
  subroutine dropin_dump()
    use common2, only : NSPN
    use mesh1, only : NMSH
    use common5, only : PSI_COEF, NWFS
    use FRM, only : BFRM, LFRM
    character(len=1024) :: path
    integer :: ios, natom, nbas, s, sp, nwf, u
    real(c_double), allocatable :: xyz(:), alpha(:), con(:), c(:, :)
    integer(c_int32_t), allocatable :: bare_off(:), ncon(:), con_off(:)
    call get_environment_variable('FLOSIC_DROPIN_DUMP', path, status=ios)
    if (ios /= 0) return
    call dropin_gather_basis(natom, xyz, bare_off, alpha, ncon, con_off, con, nbas)
    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
    write(u) int(natom, c_int32_t), int(nbas, c_int32_t), int(size(alpha), c_int32_t), int(size(con), c_int32_t), &
             int(NMSH, c_int32_t), int(NSPN, c_int32_t)
    write(u) xyz, bare_off, alpha, ncon, con_off, con
    do s = 1, NSPN
      sp = min(s, NSPN)
      nwf = max(NWFS(sp), LFRM(s))
      allocate(c(nbas, nwf))
      call dropin_salc_to_local(nbas, nwf, PSI_COEF(1:nbas, 1:nwf, 1, sp), c)
      write(u) int(nwf, c_int32_t), int(LFRM(s), c_int32_t)
      write(u) c, BFRM(1:3, 1:LFRM(s), s)
      deallocate(c)
    end do
    close(u)
  end subroutine dropin_dump

  ! Hand over orbitals and FODs of both spins and evaluate, unless they equal those of the last evaluation.
  subroutine dropin_evaluate()
    use common2, only : NSPN
    use common5, only : PSI_COEF, NWFS
    use common8, only : N_REP, NS_TOT
    use FRM, only : BFRM, LFRM
    integer :: s, sp, nwf(2), nf(2)
    logical :: same
    real(c_double), allocatable :: c(:, :)
    if (.not. c_associated(fctx)) call dropin_dump()
    call dropin_setup()
    if (N_REP /= 1 .or. NS_TOT(1) /= nbas_s) then
      write(6, *) 'FLOSIC_HIP: point group NONE only (N_REP, NS_TOT, basis =', N_REP, NS_TOT(1), nbas_s, ')'
      call stopit
    end if
    nf = 0
    nwf = 0
    do s = 1, NSPN
      nf(s) = LFRM(s)
      nwf(s) = max(NWFS(s), nf(s))
    end do
    same = have_eval .and. allocated(coef_s)
    if (same) same = all(nwf == nwf_s) .and. all(nf == nfod_s)
    if (same) then
      do s = 1, NSPN
        sp = s
        same = same .and. all(coef_s(:, 1:nwf(s), s) == PSI_COEF(1:nbas_s, 1:nwf(s), 1, sp)) .and. &
               all(fod_s(:, 1:nf(s), s) == BFRM(:, 1:nf(s), s))
      end do
    end if
    if (same) return
    if (allocated(coef_s)) deallocate(coef_s, fod_s)
    allocate(coef_s(nbas_s, maxval(nwf), 2), fod_s(3, max(1, maxval(nf)), 2))
    do s = 1, NSPN
      sp = min(s, NSPN)
      coef_s(:, 1:nwf(s), s) = PSI_COEF(1:nbas_s, 1:nwf(s), 1, sp)
      fod_s(:, 1:nf(s), s) = BFRM(:, 1:nf(s), s)
      allocate(c(nbas_s, nwf(s)))
      call dropin_salc_to_local(nbas_s, nwf(s), PSI_COEF(1:nbas_s, 1:nwf(s), 1, sp), c)
      call dropin_check(flosic_set_orbitals(fctx, int(s, c_int32_t), int(nbas_s, c_int32_t), int(nwf(s), c_int32_t), &
                                            int(nf(s), c_int32_t), c), 'flosic_set_orbitals')
      deallocate(c)
      call dropin_check(flosic_set_fods(fctx, int(s, c_int32_t), int(nf(s), c_int32_t), BFRM(:, :, s)), &
                        'flosic_set_fods')
    end do
    nwf_s = nwf
    nfod_s = nf
    call dropin_check(flosic_evaluate(fctx, zsic_s), 'flosic_evaluate')
    have_eval = .true.
    write(6, '(a,es24.16)') ' FLOSIC_HIP: evaluation done, E_SIC = ', zsic_s
  end subroutine dropin_evaluate
end module flosic_dropin


subroutine FRMORB2(LSXX, LFM)
  use flosic_dropin
  use SICMAT, only : FMAT
  use FRM, only : DEBDAX
  implicit none
  integer, intent(in) :: LSXX, LFM
  integer :: s, n
  real(c_double), allocatable :: f(:, :, :), o(:, :), q(:), u(:, :), t(:, :)
  if (LSXX == 0) then
    call dropin_evaluate()
    do s = 1, 2
      n = nfod_s(s)
      if (n == 0) cycle
      allocate(f(n, n, 4), o(n, n), q(n), u(n, n), t(n, n))
      call dropin_check(flosic_get_fo(fctx, int(s, c_int32_t), f, o, q, u, t), 'flosic_get_fo')
      FMAT(1:n, 1:n, 1:4, s) = f
      deallocate(f, o, q, u, t)
    end do
  else if (LSXX < 0 .and. LFM == -1) then
    s = -LSXX
    n = nfod_s(s)
    if (n > 0) call dropin_check(flosic_fod_forces(fctx, int(s, c_int32_t), DEBDAX(:, :, s)), 'flosic_fod_forces')
  else
    write(6, *) 'FLOSIC_HIP: FRMORB2 task not provided by the library:', LSXX, LFM
    call stopit
  end if
end subroutine FRMORB2


subroutine FRMIORB(NXBG, NXED)
  use flosic_dropin
  use LOCORB, only : TMAT
  use MOCORB, only : NFRM
  implicit none
  integer, intent(in) :: NXBG, NXED
  integer :: s, n
  real(c_double), allocatable :: f(:, :, :), o(:, :), q(:), u(:, :), t(:, :)
  call dropin_evaluate()
  do s = 1, 2
    n = nfod_s(s)
    NFRM(s) = n
    if (n == 0) cycle
    allocate(f(n, n, 4), o(n, n), q(n), u(n, n), t(n, n))
    call dropin_check(flosic_get_fo(fctx, int(s, c_int32_t), f, o, q, u, t), 'flosic_get_fo')
    TMAT(1:n, 1:n, s) = t
    deallocate(f, o, q, u, t)
  end do
end subroutine FRMIORB


subroutine APOTNL_SIC(TOTQNUM, LSPX, IORBX)
  use flosic_dropin
  use FRM, only : RESULTS
  implicit none
  real(8), intent(out) :: TOTQNUM
  integer, intent(in) :: LSPX, IORBX
  real(c_double) :: z
  call dropin_evaluate()
  if (IORBX == 1) call dropin_check(flosic_apotnl_sic(fctx, int(LSPX, c_int32_t), RESULTS(:, :, LSPX), z), &
                                    'flosic_apotnl_sic')
  TOTQNUM = RESULTS(2, IORBX, LSPX)
end subroutine APOTNL_SIC


subroutine SICLAGM(LSPN)
  use flosic_dropin
  use SICMAT, only : SIC, MAX_OCC
  implicit none
  integer, intent(in) :: LSPN
  real(c_double), allocatable :: s(:, :)
  integer :: n, nwf
  call dropin_evaluate()
  n = nfod_s(LSPN)
  nwf = nwf_s(LSPN)
  if (n == 0) return
  if (.not. allocated(SIC)) allocate(SIC(MAX_OCC, MAX_OCC, 2))
  ! SCFSICW adds SIC(IWF,JWF) over all orbitals to the Hamiltonian (scfsicw.f:400-404): columns past the FLOs are zero
  SIC(:, :, LSPN) = 0.0d0
  allocate(s(nwf, n))
  call dropin_check(flosic_siclagm(fctx, int(LSPN, c_int32_t), s), 'flosic_siclagm')
  SIC(1:nwf, 1:n, LSPN) = s
  deallocate(s)
  write(6, *) 'SIC MATRIX FOR SPIN LSPN:', LSPN
  do n = 1, nfod_s(LSPN)
    write(6, '(1x,15f12.5)') SIC(1:nfod_s(LSPN), n, LSPN)
  end do
end subroutine SICLAGM


subroutine SICLMSLV(LSPN, LFX, MPTS, WMSA, SICP, RPTS)
  implicit none
  integer :: LSPN, LFX, MPTS
  real(8) :: WMSA(*), SICP(*), RPTS(*)
  write(6, *) 'FLOSIC_HIP: SICLMSLV is not used with the library (SICLAGM computes the whole SIC matrix)'
  call stopit
end subroutine SICLMSLV


! DIAGV_FO(IPREONLY) (diagv_fo.F:5-156), SCFSICW's orbital rotation of the FLO-SIC SCF (scfsicw.f:556): HAM of
! FOR_DIAG1 (leading dimension its own first extent, NDH), m = INBAS orbitals of which NORB are occupied, PHIRES of
! DIAGV1 on return.  IPREONLY = 1 returns the pre-diagonalisation's eigenvectors as columns, as the reference does
! before its transpose (diagv_fo.F:64-71).  The eigenvectors' signs follow the library's eigensolver, so orbitals can
! differ from the reference's in sign only.
subroutine DIAGV_FO(IPREONLY)
  use flosic_dropin, only : dropin_check
  use flosic_hip, only : flosic_diagv_fo
  use iso_c_binding
  use for_diag1, only : HAM
  use diagv1, only : NORB, PHIRES
  use global_inputs, only : INBAS
  implicit none
  integer, intent(in) :: IPREONLY
  integer(c_int32_t) :: sweeps
  integer :: m, n
  m = INBAS
  n = NORB
  if (IPREONLY == 1) n = 0
  call dropin_check(flosic_diagv_fo(int(m, c_int32_t), int(n, c_int32_t), HAM, int(size(HAM, 1), c_int32_t), &
                                    PHIRES, sweeps), 'flosic_diagv_fo')
  if (IPREONLY == 1) then
    PHIRES(1:m, 1:m) = transpose(PHIRES(1:m, 1:m))
  else if (sweeps > 0) then
    write(6, *) 'FLOSIC_HIP: DIAGV_FO converged in sweep', sweeps
  else
    write(6, *) 'WARNING: DIAGV not converged!'
  end if
end subroutine DIAGV_FO


! DIAGV_FO_UNI(matlen,N,M,FOHAM,PHIRES,IPREONLY) (diagv_fo.F:157-220): the unified Hamiltonian's eigenvectors, PHIRES(i,j)
! = component j of the eigenvector of the i-th eigenvalue (ascending); the same call with no occupied orbitals
subroutine DIAGV_FO_UNI(matlen, N, M, FOHAM, PHIRES, IPREONLY)
  use flosic_dropin, only : dropin_check
  use flosic_hip, only : flosic_diagv_fo
  use iso_c_binding
  implicit none
  integer, intent(in) :: matlen, N, M, IPREONLY
  real(8) :: FOHAM(matlen, matlen), PHIRES(matlen, matlen)
  real(c_double), allocatable :: p(:, :)
  integer(c_int32_t) :: sweeps
  allocate(p(M, M))
  call dropin_check(flosic_diagv_fo(int(M, c_int32_t), 0_c_int32_t, FOHAM, int(matlen, c_int32_t), p, sweeps), &
                    'flosic_diagv_fo')
  PHIRES = 0.0d0
  PHIRES(1:M, 1:M) = p
  deallocate(p)
end subroutine DIAGV_FO_UNI

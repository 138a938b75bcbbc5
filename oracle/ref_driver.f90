! ref_driver.f90 -- C-callable driver around routines of the FLOSIC reference, compiled from the
! reference sources where they lie.  TEST INFRASTRUCTURE ONLY: it lets the CPU tests check the C oracle
! against the reference's own implementation of
!   FFFMT3/FFFMT4/FFFMTC (fffmt3.F, fffmt4.F, fffmtc.F)     Boys functions
!   POISSON2/POISSON1 (poisson2.f, poisson1.f) in DEBUG mode  exact Gaussian-pair potential at every point
!   GORBDRV (gorbdrv.f)                                       basis values and derivatives
!   PW91EX/PW91LC/PW91NC/PBEEX/PBECOR                         exchange-correlation kernels
!   FRMORB2(-ISPN,-1) (frmorb2.F:140-729)                     Lowdin step and FOD force from FMAT and SIC
! The module variables these routines read are set here from the arguments.
module flosic_ref_driver
  use iso_c_binding
  implicit none
contains

  subroutine ref_boys(top, n, t, f) bind(C, name="ref_boys")
    integer(c_int), value :: top, n
    real(c_double) :: t(n), f(n, 5)
    real(8), allocatable :: chgd(:)
    allocate(chgd(n))
    chgd = exp(-t)
    f = 0.0d0
    if (top == 2) then
      call FFFMT3(n, n, t, chgd, f)
    else if (top == 3) then
      call FFFMT4(n, n, t, chgd, f)
    else
      call FFFMTC(n, n, t, chgd, f)
    end if
    deallocate(chgd)
  end subroutine ref_boys

  ! exact potential of npair Cartesian pair densities RHO(10,10,npair) at npts points
  subroutine ref_poisson(npair, alpha, a, beta, b, rho, npts, pts, v) bind(C, name="ref_poisson")
    use mesh1, only : rmsh, nmsh
    use pot_dens, only : coulomb, rhog
    use common2, only : nspn, igga
    use common7, only : modden
    use debug1, only : debug
    integer(c_int), value :: npair, npts
    real(c_double) :: alpha(npair), a(3, npair), beta(npair), b(3, npair), rho(10, 10, npair)
    real(c_double) :: pts(3, npts), v(npts)
    real(8) :: al(100), aa(3, 100), be(100), bb(3, 100), rr(10, 10, 100)
    integer :: i0, k, np
    nmsh = npts
    rmsh(:, 1:npts) = pts
    if (allocated(coulomb)) deallocate(coulomb)
    if (allocated(rhog)) deallocate(rhog)
    allocate(coulomb(npts), rhog(npts, 10, 2))
    coulomb = 0.0d0
    rhog = 0.0d0
    nspn = 2
    igga = 0
    modden = 0
    debug = .true.
    do i0 = 1, npair, 50
      np = min(50, npair - i0 + 1)
      rr = 0.0d0
      do k = 1, np
        al(2*k-1) = alpha(i0+k-1)
        al(2*k) = alpha(i0+k-1)
        be(2*k-1) = beta(i0+k-1)
        be(2*k) = beta(i0+k-1)
        aa(:, 2*k-1) = a(:, i0+k-1)
        aa(:, 2*k) = a(:, i0+k-1)
        bb(:, 2*k-1) = b(:, i0+k-1)
        bb(:, 2*k) = b(:, i0+k-1)
        rr(:, :, 2*k-1) = rho(:, :, i0+k-1)
      end do
      call POISSON2(.false., 2*np, 10, 10, al, aa, be, bb, rr)
    end do
    v = coulomb(1:npts)
  end subroutine ref_poisson

  ! one atom type: nbare exponents, ncon(3) contractions, con(nbare, maxcon, 3); npts <= 25 points
  subroutine ref_gorbdrv(nderv, nbare, alpha, ncon, maxcon, con, lmax, npts, pts, grad) bind(C, name="ref_gorbdrv")
    use common2, only : bfcon, bfalp, n_bare, n_con, lsymmax
    integer(c_int), value :: nderv, nbare, maxcon, lmax, npts
    integer(c_int) :: ncon(3)
    real(c_double) :: alpha(nbare), con(nbare, maxcon, 3), pts(3, npts)
    real(c_double) :: grad(25, 10, 6, 15, 3)
    real(8) :: p25(25, 3)
    logical :: iupdat, icount(15, 3)
    integer :: l
    n_bare(1) = nbare
    bfalp(1:nbare, 1) = alpha
    lsymmax(1) = lmax
    n_con(1:3, 1) = ncon
    bfcon(:, :, :, 1) = 0.0d0
    do l = 1, 3
      bfcon(1:nbare, 1:maxcon, l, 1) = con(:, :, l)
    end do
    p25 = 0.0d0
    p25(1:npts, 1) = pts(1, :)
    p25(1:npts, 2) = pts(2, :)
    p25(1:npts, 3) = pts(3, :)
    grad = 0.0d0
    call GORBDRV(nderv, iupdat, icount, npts, p25, 1, grad)
  end subroutine ref_gorbdrv

  ! SUBVLXC with NSPN = 2 and no atom centres (no core density, MODE = 1: no local potential): gga = 0 takes
  ! IGGA = 0 and IDFTYP = (6, 6) (PW91EX at s = u = v = 0, PW91LC + PW91NC), gga = 1 IGGA = 1 and IDFTYP = (7, 7)
  ! (PBEEX, PBECOR).  rgu, rgd: value, gradient and Hessian of the spin densities; pot(:, 1) spin up, pot(:, 2)
  ! spin down; e = EXCVEC summed over the 100-point blocks SUBVLXC takes at a time.
  subroutine ref_subvlxc(gga, npts, rgu, rgd, w, pot, e) bind(C, name="ref_subvlxc")
    use mesh1, only : rmsh, wmsh, nmsh
    use common2, only : nspn, igga, idftyp, ncnt
    integer(c_int), value :: gga, npts
    real(c_double) :: rgu(10, npts), rgd(10, npts), w(npts), pot(npts, 2), e(4)
    real(8) :: rhov(10 * 2 * 100), vxcs(2 * 100), vlos(100), excvec(4)
    integer :: i0, m, k, ng
    nmsh = npts
    rmsh(:, 1:npts) = 0.0d0
    wmsh(1:npts) = w
    nspn = 2
    ncnt = 0
    if (gga /= 0) then
      igga = 1
      idftyp = 7
      ng = 10
    else
      igga = 0
      idftyp = 6
      ng = 1
    end if
    e = 0.0d0
    do i0 = 1, npts, 100
      m = min(100, npts - i0 + 1)
      rhov = 0.0d0
      do k = 1, m
        rhov(ng * 2 * (k - 1) + 1 : ng * 2 * (k - 1) + ng) = rgu(1:ng, i0 + k - 1) + rgd(1:ng, i0 + k - 1)
        rhov(ng * (2 * (k - 1) + 1) + 1 : ng * (2 * (k - 1) + 1) + ng) = rgd(1:ng, i0 + k - 1)
      end do
      call SUBVLXC(1, i0 - 1, m, rhov, vxcs, vlos, excvec)
      pot(i0 : i0 + m - 1, 1) = vxcs(1:m)
      pot(i0 : i0 + m - 1, 2) = vxcs(m + 1 : 2 * m)
      e = e + excvec
    end do
  end subroutine ref_subvlxc

  subroutine ref_pw91ex(dkf, s, u, v, exl, exn, vx) bind(C, name="ref_pw91ex")
    real(c_double), value :: dkf, s, u, v
    real(c_double) :: exl, exn, vx
    call PW91EX(dkf, s, u, v, exl, exn, vx)
  end subroutine ref_pw91ex

  subroutine ref_pw91lc(rs, zet, ec, vcup, vcdn) bind(C, name="ref_pw91lc")
    real(c_double), value :: rs, zet
    real(c_double) :: ec, vcup, vcdn
    real(8) :: ecrs, eczet, alfc
    call PW91LC(rs, zet, ec, vcup, vcdn, ecrs, eczet, alfc)
  end subroutine ref_pw91lc

  subroutine ref_pbeex(rho, s, u, v, exl, exn, vx) bind(C, name="ref_pbeex")
    real(c_double), value :: rho, s, u, v
    real(c_double) :: exl, exn, vx
    call PBEEX(1, rho, s, u, v, 1, 1, exl, exn, vx)
  end subroutine ref_pbeex

  subroutine ref_pbecor(rs, zet, t, uu, vv, ww, ec, vcup, vcdn, h, dvcup, dvcdn) bind(C, name="ref_pbecor")
    real(c_double), value :: rs, zet, t, uu, vv, ww
    real(c_double) :: ec, vcup, vcdn, h, dvcup, dvcdn
    call PBECOR(rs, zet, t, uu, vv, ww, 1, 1, ec, vcup, vcdn, h, dvcup, dvcdn)
  end subroutine ref_pbecor

  ! FOD forces of spin ispn from FMAT(j,i,K,spin) (n(spin) x n(spin) x 4) and SIC(k,l) of that spin.
  ! Outputs DEBDAX(3,n) and SLAT(n,n) = S^-1/2 T of that spin.
  subroutine ref_frmorb2_force(ispn, n1, n2, fmat1, fmat2, sic, debdax_out, slat_out) bind(C, name="ref_frmorb2_force")
    use common2, only : nspn
    use common5, only : nwfs, nwf
    use FRM, only : lfrm, debdax
    use SICMAT, only : SIC_M => SIC, FMAT
    use MOCORB, only : slat
    integer(c_int), value :: ispn, n1, n2
    real(c_double) :: fmat1(n1, n1, 4), fmat2(n2, n2, 4), sic(*), debdax_out(3, *), slat_out(*)
    integer :: n, k, i, j
    nspn = 2
    lfrm(1) = n1
    lfrm(2) = n2
    nwfs(1) = n1
    nwfs(2) = n2
    nwf = n1 + n2
    FMAT = 0.0d0
    FMAT(1:n1, 1:n1, 1:4, 1) = fmat1
    FMAT(1:n2, 1:n2, 1:4, 2) = fmat2
    if (.not. allocated(SIC_M)) allocate(SIC_M(1500, 1500, 2))
    SIC_M = 0.0d0
    n = n1
    if (ispn == 2) n = n2
    k = 0
    do i = 1, n
      do j = 1, n
        k = k + 1
        SIC_M(j, i, ispn) = sic(k)
      end do
    end do
    call FRMORB2(-ispn, -1)
    debdax_out(1:3, 1:n) = debdax(1:3, 1:n, ispn)
    k = 0
    do j = 1, n
      do i = 1, n
        k = k + 1
        slat_out(k) = slat(i, j, ispn)
      end do
    end do
  end subroutine ref_frmorb2_force

end module flosic_ref_driver

! flosic_hip.f90 -- ISO_C_BINDING interfaces to libflosic_hip.so (see flosic_hip.h for the contract of each
! entry point).  A Fortran host such as the FLOSIC SCF driver USEs this module; arrays are passed as plain
! column-major REAL*8 / INTEGER*4 buffers, the context as an opaque TYPE(C_PTR), and every function returns
! the library status (0 = success).
module flosic_hip
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: FLOSIC_OK = 0, FLOSIC_EINVAL = 1, FLOSIC_ELOWDIN = 2, FLOSIC_ECOUPOT = 3, &
                                   FLOSIC_EDEVICE = 4, FLOSIC_ESTATE = 5, FLOSIC_ENOMEM = 6
  integer(c_int32_t), parameter :: FLOSIC_XC_LSDA = 0, FLOSIC_XC_PBE = 1
  integer(c_int32_t), parameter :: FLOSIC_PREC_FP64 = 0, FLOSIC_PREC_MIXED = 1, FLOSIC_PREC_BF16X3 = 2, &
                                   FLOSIC_PREC_BF16X2 = 3

  interface
    integer(c_int32_t) function flosic_create(device, ctx) bind(C, name="flosic_create")
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: device
      type(c_ptr) :: ctx
    end function
    integer(c_int32_t) function flosic_destroy(ctx) bind(C, name="flosic_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function flosic_last_error(ctx) bind(C, name="flosic_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function flosic_set_basis(ctx, natom, xyz, bare_off, alpha, ncon, con_off, con) &
        bind(C, name="flosic_set_basis")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: natom
      real(c_double) :: xyz(*), alpha(*), con(*)
      integer(c_int32_t) :: bare_off(*), ncon(*), con_off(*)
    end function
    integer(c_int32_t) function flosic_set_mesh(ctx, nmsh, rmsh, wmsh) bind(C, name="flosic_set_mesh")
      import :: c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nmsh
      real(c_double) :: rmsh(3, *), wmsh(*)
    end function
    integer(c_int32_t) function flosic_set_options(ctx, nspn, xc, precision) bind(C, name="flosic_set_options")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nspn, xc, precision
    end function
    integer(c_int32_t) function flosic_set_orbitals(ctx, ispn, nbas, nwf, nocc, coef) bind(C, name="flosic_set_orbitals")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn, nbas, nwf, nocc
      real(c_double) :: coef(nbas, *)
    end function
    integer(c_int32_t) function flosic_set_fods(ctx, ispn, nfrm, bfrm) bind(C, name="flosic_set_fods")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn, nfrm
      real(c_double) :: bfrm(3, *)
    end function
    integer(c_int32_t) function flosic_frmorb(ctx) bind(C, name="flosic_frmorb")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int32_t) function flosic_get_fo(ctx, ispn, fmat, over, eval, filo, tmat) bind(C, name="flosic_get_fo")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: fmat(*), over(*), eval(*), filo(*), tmat(*)
    end function
    integer(c_int32_t) function flosic_evaluate(ctx, zsic) bind(C, name="flosic_evaluate")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: zsic
    end function
    integer(c_int32_t) function flosic_apotnl_sic(ctx, ispn, results, zsic) bind(C, name="flosic_apotnl_sic")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: results(13, *), zsic
    end function
    integer(c_int32_t) function flosic_apotnl_dft(ctx, pot, energies) bind(C, name="flosic_apotnl_dft")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: pot(*), energies(5)
    end function
    integer(c_int32_t) function flosic_apotnl_dft_noalloc(ctx, pot, energies) bind(C, name="flosic_apotnl_dft_noalloc")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: pot(*), energies(5)
    end function
    integer(c_int32_t) function flosic_siclag_der(ctx, dersic) bind(C, name="flosic_siclag_der")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: dersic(*)
    end function
    integer(c_int32_t) function flosic_siclagm(ctx, ispn, sic) bind(C, name="flosic_siclagm")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: sic(*)
    end function
    integer(c_int32_t) function flosic_sic_vab(ctx, ispn, vab) bind(C, name="flosic_sic_vab")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: vab(*)
    end function
    integer(c_int32_t) function flosic_fod_forces(ctx, ispn, debdax) bind(C, name="flosic_fod_forces")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: debdax(3, *)
    end function
    integer(c_int32_t) function flosic_fod_gradient(ctx, ispn, grad) bind(C, name="flosic_fod_gradient")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: grad(3, *)
    end function
    integer(c_int32_t) function flosic_get_potential(ctx, ispn, iorb, pot, vh) bind(C, name="flosic_get_potential")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn, iorb
      real(c_double) :: pot(*), vh(*)
    end function
    integer(c_int32_t) function flosic_lagmult_slv(ctx, ispn, pot, ldsic, sic) bind(C, name="flosic_lagmult_slv")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn, ldsic
      real(c_double) :: pot(*), sic(ldsic, *)
    end function
    integer(c_int32_t) function flosic_sic_hamiltonian(ctx, ispn, hao) bind(C, name="flosic_sic_hamiltonian")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn
      real(c_double) :: hao(*)
    end function
    integer(c_int32_t) function flosic_get_pairs(ctx, npair, list) bind(C, name="flosic_get_pairs")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t) :: npair, list(*)
    end function
    integer(c_int32_t) function flosic_get_timings(ctx, n, ms) bind(C, name="flosic_get_timings")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      real(c_double) :: ms(*)
    end function
    integer(c_int32_t) function flosic_get_counts(ctx, n, counts) bind(C, name="flosic_get_counts")
      import :: c_int32_t, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      integer(c_int64_t) :: counts(*)
    end function
    integer(c_int32_t) function flosic_set_coulomb_cols(ctx, ncols) bind(C, name="flosic_set_coulomb_cols")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ncols
    end function
    integer(c_int32_t) function flosic_diagv_fo(m, n, ham, ldh, phires, sweeps) bind(C, name="flosic_diagv_fo")
      import :: c_int32_t, c_double
      integer(c_int32_t), value :: m, n, ldh
      real(c_double) :: ham(*), phires(*)
      integer(c_int32_t) :: sweeps
    end function
    integer(c_int32_t) function flosic_scf_sic_orbitals(ctx, ispn, ham_ao, ldham, coef, eval, sweeps) &
        bind(C, name="flosic_scf_sic_orbitals")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ispn, ldham
      real(c_double) :: ham_ao(*), coef(*), eval(*)
      integer(c_int32_t) :: sweeps
    end function
    integer(c_int32_t) function flosic_set_far_threshold(ctx, tfar) bind(C, name="flosic_set_far_threshold")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tfar
    end function
    integer(c_int32_t) function flosic_set_mesh_block(ctx, block_points) bind(C, name="flosic_set_mesh_block")
      import :: c_int32_t, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: block_points
    end function
    integer(c_int32_t) function flosic_batch_create(n, ctxs, use_graph, batch) bind(C, name="flosic_batch_create")
      import :: c_int32_t, c_ptr
      integer(c_int32_t), value :: n
      type(c_ptr) :: ctxs(*)
      integer(c_int32_t), value :: use_graph
      type(c_ptr) :: batch
    end function
    integer(c_int32_t) function flosic_batch_evaluate(batch, zsic) bind(C, name="flosic_batch_evaluate")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: batch
      real(c_double) :: zsic(*)
    end function
    type(c_ptr) function flosic_batch_last_error(batch) bind(C, name="flosic_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: batch
    end function
    integer(c_int32_t) function flosic_batch_get_timings(batch, n, ms) bind(C, name="flosic_batch_get_timings")
      import :: c_int32_t, c_ptr, c_double
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: n
      real(c_double) :: ms(*)
    end function
    integer(c_int32_t) function flosic_batch_get_counts(batch, n, counts) bind(C, name="flosic_batch_get_counts")
      import :: c_int32_t, c_ptr, c_int64_t
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: n
      integer(c_int64_t) :: counts(*)
    end function
    integer(c_int32_t) function flosic_batch_destroy(batch) bind(C, name="flosic_batch_destroy")
      import :: c_int32_t, c_ptr
      type(c_ptr), value :: batch
    end function
  end interface
end module flosic_hip

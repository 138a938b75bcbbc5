/*
 * flosic_hip.h -- C ABI of the MI355X FLO-SIC per-Fermi-orbital evaluation library (libflosic_hip.so).
 *
 * The reference (FLOSIC/PublicRelease_2020, Fortran) passes the hot path's data through module variables
 * and files.  Each entry point below replaces one of its subroutines; the module data they read is handed
 * over explicitly by the set-up calls.  Arrays are column-major with 1-based meaning in the comments and
 * 0-based storage, REAL*8 -> double, INTEGER -> int32_t.  Every call returns a status (0 = success) and
 * never aborts the process; flosic_last_error() gives the message.  Calls on one context are not
 * re-entrant; distinct contexts are independent (one HIP stream each).
 *
 * Typical sequence (mirrors main.F:577-775):
 *   flosic_create -> flosic_set_basis -> flosic_set_mesh -> flosic_set_options
 *   per SCF/FOD step: flosic_set_orbitals(1|2) -> flosic_set_fods(1|2) -> flosic_evaluate
 *   then flosic_apotnl_sic / flosic_siclagm / flosic_sic_vab / flosic_fod_forces / flosic_fod_gradient
 */
#ifndef FLOSIC_HIP_H
#define FLOSIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct flosic_ctx flosic_ctx;
typedef struct flosic_batch flosic_batch;

/* status codes */
#define FLOSIC_OK 0
#define FLOSIC_EINVAL 1   /* bad argument or inconsistent sizes */
#define FLOSIC_ELOWDIN 2  /* Fermi-orbital overlap eigenvalue < 1e-8: LOWSIC STOPIT, lowden_files.f:71-99 */
#define FLOSIC_ECOUPOT 3  /* "COUPOT IS ZERO": sum |V_H| < 0.001, apotnl_sic.F:277-283 */
#define FLOSIC_EDEVICE 4  /* HIP / rocBLAS runtime failure */
#define FLOSIC_ESTATE 5   /* a required set-up call is missing */
#define FLOSIC_ENOMEM 6   /* device allocation failed */

/* exchange-correlation: the two functionals of the SIC path (setdftyp.F:5-110) */
#define FLOSIC_XC_LSDA 0 /* LDA-PW91*LDA-PW91 : PW91EX at s=u=v=0 + PW91LC */
#define FLOSIC_XC_PBE 1  /* GGA-PBE*GGA-PBE   : PBEEX + PBECOR */

/* arithmetic of the Coulomb contraction (the dominant mesh contraction); Lowdin, Boys functions, the force
 * algebra and every energy reduction stay FP64 in all modes */
#define FLOSIC_PREC_FP64 0   /* FP64 MFMA */
#define FLOSIC_PREC_MIXED 1  /* FP32 MFMA, FP32 accumulation per chunk of <=64 Hermite rows, FP64 sums */
#define FLOSIC_PREC_BF16X3 2 /* operands split into 3 bf16 pieces, 6 cross products on bf16 MFMA, FP64 sums */
#define FLOSIC_PREC_BF16X2 3 /* operands split into 2 bf16 pieces, 3 cross products ("3-term BF16 split") */

/* Context on HIP device `device` (a hipSetDevice index). */
int32_t flosic_create(int32_t device, flosic_ctx **ctx);
int32_t flosic_destroy(flosic_ctx *ctx);
const char *flosic_last_error(const flosic_ctx *ctx);

/* Basis: module COMMON2 (BFALP, BFCON, N_BARE, N_CON, RIDT) as consumed by GORBDRV (gorbdrv.f:18) and
 * COUPOT_SIC (coupot_sic.F:228-482).  Atom a has bare exponents alpha[bare_off[a] .. bare_off[a+1]); for
 * l = 0,1,2 it has ncon[3a+l] contractions whose coefficients over its bare exponents are the row-major
 * block con[con_off[3a+l] ...] of shape ncon x nbare (normalisation folded in).  Basis functions are
 * ordered atom, l, Cartesian component (s; x,y,z; xx,yy,zz,xy,xz,yz), contraction (frmorb2.F:262-266).
 * Builds the primitive pair list and the FOD-independent basis planes on the mesh (if a mesh is set). */
int32_t flosic_set_basis(flosic_ctx *ctx, int32_t natom, const double *xyz, const int32_t *bare_off,
                         const double *alpha, const int32_t *ncon, const int32_t *con_off, const double *con);

/* Mesh: module MESH1 (RMSH(3,NMSH), WMSH(NMSH)). */
int32_t flosic_set_mesh(flosic_ctx *ctx, int64_t nmsh, const double *rmsh, const double *wmsh);

/* NSPN, functional and precision.  NSPN = 1 (spin-restricted): one spin channel whose SIC energies count twice
 * (RESULTS(9) = XENG * 2/NSPN, apotnl_sic.F:581); spin 2 is empty and flosic_set_orbitals / flosic_set_fods reject
 * it.  Switching back to NSPN = 2 requires spin 2 to be set again. */
int32_t flosic_set_options(flosic_ctx *ctx, int32_t nspn, int32_t xc, int32_t precision);

/* KS orbitals of spin ispn (1 or 2): COMMON5 PSI_COEF as unravelled by UNRAVEL2 (unravel2.f:90-225);
 * coef is nbas x nwf column-major; the first nocc columns are the occupied orbitals (occupation 1). */
int32_t flosic_set_orbitals(flosic_ctx *ctx, int32_t ispn, int32_t nbas, int32_t nwf, int32_t nocc,
                            const double *coef);

/* FOD positions of spin ispn: module FRM BFRM(3,nfrm); nfrm must equal that spin's nocc (frmiorb.f:443). */
int32_t flosic_set_fods(flosic_ctx *ctx, int32_t ispn, int32_t nfrm, const double *bfrm);

/* FRMORB2(0,0) + FRMIORB(1,1) (frmorb2.F:140-525, frmiorb.f:336-482): Fermi orbitals at the FODs,
 * FO overlap, Lowdin S^-1/2 and TMAT for both spins. */
int32_t flosic_frmorb(flosic_ctx *ctx);

/* Fermi-orbital data of spin ispn after flosic_frmorb: FMAT(j,i,K) nfrm x nfrm x 4, OVER nfrm x nfrm,
 * EVAL (q ascending) nfrm, FILO (eigenvectors) nfrm x nfrm, TMAT nfrm x nfrm.  Any pointer may be NULL. */
int32_t flosic_get_fo(flosic_ctx *ctx, int32_t ispn, double *fmat, double *over, double *eval, double *filo,
                      double *tmat);

/* The whole SIC evaluation for both spins: APOTNL_SIC over every FLO (apotnl_sic.F:31-668, with
 * COUPOT_SIC/POISSON exact Coulomb), SICLAGM (frmorb_mpi.F:2-178), V^i_ab, and FRMORB2(-ISPN,-1) FOD
 * forces (frmorb2.F:589-729).  zsic (may be NULL) receives E_SIC = sum of RESULTS(9,:,:) (main.F:756). */
int32_t flosic_evaluate(flosic_ctx *ctx, double *zsic);

/* APOTNL_SIC for all FLOs: RESULTS(13,nfrm) of spin ispn (apotnl_sic.F:595-623) and ZSIC. */
int32_t flosic_apotnl_sic(flosic_ctx *ctx, int32_t ispn, double *results, double *zsic);

/* The DFT-side potential of the last evaluation's densities (APOTNL_DFT with NSPN = 2: getvlxc.F -> SUBVLXC,
 * subvlxc_libxc.F90:197-470, the SUBVLXC the program links): rho_s = sum of the FLO densities of spin s (an
 * orthonormal rotation of the occupied orbitals, occupations 1), V_H = the exact electronic Hartree potential of
 * rho_up + rho_dn (the sum of the FLO Hartree potentials); the terms APOTNL_DFT adds beyond it (nuclear and external
 * potentials, EXTPOT and the field term, apotnl_dft.F) are not included: the context holds no nuclear charges.
 * LSDA: exchange per spin at 2 rho_s (PW91EX, s = u = v = 0) and PW92 correlation (PW91LC) at the point's zeta.  PBE: exchange per spin at 2 rho_s (PBEEX with the s, u, v of 2 rho_s)
 * and PBECOR at the point's zeta with t, uu, vv of the total density and the zeta-gradient term ww.  pot
 * (nmsh x 2, may be NULL): V_H + V_xc of spin up, then spin down (the POTOUT layout); energies (may be NULL):
 * {E_x, E_c, E_H = 1/2 int rho V_H, N_up, N_dn} on the mesh.  NSPN = 2 only (otherwise FLOSIC_EINVAL); in
 * mesh-block mode the FLO fields are rebuilt block by block. */
int32_t flosic_apotnl_dft(flosic_ctx *ctx, double *pot, double *energies);

/* No counterpart in the reference: flosic_apotnl_dft for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same kernels, same outputs bit for
 * bit, but it allocates nothing: the staged potential planes and the block partials live in a fixed per-GPU
 * scratch array of the library (16.8 MB, held until the process ends), so the mesh may have at most 1 048 576
 * points (FLOSIC_EINVAL beyond).  It runs no evaluation (FLOSIC_ESTATE unless the last flosic_evaluate is current)
 * and needs the resident FLO fields (FLOSIC_EINVAL in mesh-block mode).  All calls on one GPU share the scratch:
 * calls that overlap on one GPU (several host threads, several contexts) race. */
int32_t flosic_apotnl_dft_noalloc(flosic_ctx *ctx, double *pot, double *energies);

/* SICLAG_DER (siclag_der.F:2-181, SICLM_DERSLV siclag_der.F:186-640, without point-group symmetry and without the
 * libxc/meta-GGA terms): DERSIC(x, i, A) = sum_k w_k dphi_i^A/dR_A,x V_i phi_i for the FLOs of both spins (spin 1
 * first, nF = nfrm_up + nfrm_dn) and every atom A, phi_i^A the part of FLO i on atom A's basis functions and V_i the
 * ZPOT potential of the last evaluation; element (x, i, A) at dersic[x + 3*(i + nF*A)], 3 x nF x natom.  NSPN = 1
 * doubles it (siclag_der.F:158-166).  The SIC part of the atomic Pulay force: FPULAY subtracts the sum over i from
 * FRC1 (mfpulay.F:88-96). */
int32_t flosic_siclag_der(flosic_ctx *ctx, double *dersic);

/* SICLAGM: SIC(j,i) = -sum_k w_k psi~_j V_i phi_i, j over all nwf orbitals (FLOs first), i over FLOs;
 * nwf x nfrm column-major (frmorb_mpi.F:456-517). */
int32_t flosic_siclagm(flosic_ctx *ctx, int32_t ispn, double *sic);

/* V^i_ab = sum_k w_k phi_a V_i phi_b over the FLOs of spin ispn: nfrm^3, element (a,b,i) at
 * a + nfrm*(b + nfrm*i) (the FLO-space restriction of AHAMMAT, sicmat.F:380-411).  Computed by every
 * evaluation; copied to the host on the first call after it (both spins), so evaluations that do not ask for it
 * skip the n^3 transfer. */
int32_t flosic_sic_vab(flosic_ctx *ctx, int32_t ispn, double *vab);

/* FRMORB2(-ISPN,-1): DEBDAX(3,nfrm) = +dE_SIC/da (frmorb2.F:589-729; consumed as the gradient by
 * FOD_OPT, electronic_geometry_lbfgs.f:143-146). */
int32_t flosic_fod_forces(flosic_ctx *ctx, int32_t ispn, double *debdax);

/* Adjoint FOD gradient dE_SIC/da(3,nfrm): dE/dTMAT = 2 SIC^T TMAT back-propagated through the Lowdin
 * step and the Fermi-orbital construction, with SIC the exact mesh derivative of the discrete energy (for PBE
 * the e_rho / e_g form; LSDA's SIC is already that, and there the gradient equals the FRMORB2 formula); the
 * Hartree term keeps the continuum symmetry <psi|V_H|phi>.  Per spin channel and without the 2/NSPN factor,
 * as DEBDAX.  No counterpart entry point in the reference. */
int32_t flosic_fod_gradient(flosic_ctx *ctx, int32_t ispn, double *grad);

/* ZPOTnnnn analogue (apotnl_sic.F:497-522): V_i = V_xc + V_H of FLO iorb (1-based) of spin ispn on the
 * mesh; vh (may be NULL) receives V_H alone. */
int32_t flosic_get_potential(flosic_ctx *ctx, int32_t ispn, int32_t iorb, double *pot, double *vh);

/* LAGMULT_SLV (lagmult_slv.F:2-187, caller NUMFORCE numforce.F:83-90): adds the DFT part of the Lagrange
 * multipliers to the SIC block of spin ispn,
 *   sic[j + ldsic*k] += sum_mesh w phi_j (-1/2 lap phi_k + pot phi_k),   j, k = 1..nfrm (FLOs of ispn),
 * with pot the DFT potential of that spin on the mesh (POTOUT(NMSH*(ISPN-1)+1 ..), nmsh values).  The FLOs are
 * those of the last evaluation; sic is typically the array flosic_siclagm returned (ldsic = nwf). */
int32_t flosic_lagmult_slv(flosic_ctx *ctx, int32_t ispn, const double *pot, int32_t ldsic, double *sic);

/* SIC Hamiltonian in the AO basis as SICHAM/APATCH assemble it (sicham.F:920-945, sicmat.F:146-262 with MM=2):
 *   H_ab = sum_i 1/2 ( <chi_a|V_i|phi_i><phi_i|chi_b> + <chi_b|V_i|phi_i><phi_i|chi_a> )
 * over the FLOs i of spin ispn, both brackets integrated on the mesh, V_i = V_H + V_xc of rho_i.  hao is
 * nbas x nbas column-major (full, symmetric). */
int32_t flosic_sic_hamiltonian(flosic_ctx *ctx, int32_t ispn, double *hao);

/* DIAGV_FO (diagv_fo.F:5-156; caller SCFSICW, scfsicw.f:548-552): the orbital rotation of the FLO-SIC SCF.  ham is
 * the non-symmetric m x m orbital Hamiltonian, ham[J + ldh*I] = HAM(J,I), of m orbitals whose first n are occupied
 * (the FLOs).  Its symmetric part is diagonalised (the eigenvectors, by ascending eigenvalue, start the orbitals;
 * diagv_fo.F:48-71), then sweeps of occupied-virtual Jacobi rotations, t = atan(2 h_ij / (h_ii - h_jj)) / 2,
 * continue until max |phi_occ^T HAM phi_virt| < 1e-6 (diagv_fo.F:74-140), at most 600.  phires[inew + m*iold] =
 * PHIRES(INEW,IOLD) on return (the new orbitals in the old ones, after the reference's final transpose); *sweeps
 * (may be NULL) = the converged sweep, or 0 when 600 sweeps did not converge (the reference warns and keeps the
 * last orbitals, diagv_fo.F:148-151; FLOSIC_OK either way).  The eigenvectors' signs follow this library's
 * eigensolver (largest component positive), so orbitals may differ from the reference's in sign only.  A host
 * computation; no context. */
int32_t flosic_diagv_fo(int32_t m, int32_t n, const double *ham, int32_t ldh, double *phires, int32_t *sweeps);

/* SCFSICW's FLO-SIC orbital update of spin ispn (scfsicw.f:330-614, NITER > 1 without the unified Hamiltonian),
 * from the last evaluation: C = [C_occ TMAT^T, C_virt] (the FLOs, then the virtual orbitals; scfsicw.f:330-349),
 * HAM = C^T ham_ao C (scfsicw.f:353-371), HAM(J,I) += SIC(I,J) for every FLO J (scfsicw.f:389-395, SIC as
 * flosic_siclagm returns it), DIAGV_FO (flosic_diagv_fo), then EVAL(i) = sum_kl PHIRES(i,k) PHIRES(i,l) HAM(l,k)
 * (scfsicw.f:584-597) and the new orbitals C_new(:,i) = sum_k PHIRES(i,k) C(:,k) (scfsicw.f:601-614).  ham_ao: the
 * spin's Kohn-Sham Hamiltonian in the basis (nbas x nbas, column-major, ldham >= nbas).  coef receives C_new
 * (nbas x nwf, leading dimension nbas), eval (may be NULL) nwf values, *sweeps (may be NULL) as flosic_diagv_fo.
 * The reference uses the full basis as orbital set (nwf = nbas); with nwf < nbas the update stays in the orbitals'
 * span.  The transforms run on rocBLAS, DIAGV_FO on the host. */
int32_t flosic_scf_sic_orbitals(flosic_ctx *ctx, int32_t ispn, const double *ham_ao, int32_t ldham, double *coef,
                                double *eval, int32_t *sweeps);

/* Primitive pair list built by flosic_set_basis (coupot_sic.F:228-310): npair rows of
 * (atomA, lA, bareA, atomB, lB, bareB), 0-based, in canonical order.  Pass list = NULL to query npair. */
int32_t flosic_get_pairs(flosic_ctx *ctx, int32_t *npair, int32_t *list);

/* Phase timings of the last flosic_evaluate in milliseconds (HIP events on the context's stream):
 * [0] FO+Lowdin  [1] orbitals on mesh (mesh-block mode: the whole block loop, [4] and [6] then hold only what
 * follows it)  [2] Hermite pair coefficients  [3] Coulomb operand + Coulomb  [4] XC+energies  [5] SIC matrix
 * [6] V^i_ab  [7] results and forces  [8] total  [9] the Coulomb kernel alone (inside [3]; in a batch, from this
 * system's operand being ready to the end of the batched launch).  After a graph-replayed batch evaluation only
 * [8] is set.  n <= 10 values are written. */
int32_t flosic_get_timings(flosic_ctx *ctx, int32_t n, double *ms);

/* Counters of the last evaluation: [0] screened primitive shell pairs (the flosic_get_pairs list) [1] Hermite
 * rows (shell pairs that share both primitive exponents and centres are one group with one row block)
 * [2] nmsh [3] nbas [4] nflo (both spins) [5] exponent-pair groups [6] Hermite rows as the Coulomb kernel of
 * the current precision streams them (chunks zero-extended to its MFMA K step) [7] peak device bytes the
 * context has held since its creation [8] FLO columns per Coulomb block in use (0 for the VALU form) [9] (64-point
 * tile, Hermite row) pairs of the split-BF16 stream contracted with two pieces (far chunks, flosic_set_far_threshold;
 * 0 for the other forms).  n <= 10 values are written. */
int32_t flosic_get_counts(flosic_ctx *ctx, int32_t n, int64_t *counts);

/* Tile shape of the Coulomb kernels: FLO columns per block (each block contracts 64 mesh points with this many
 * columns; a narrower block repeats the Boys/Hermite R generation for every column block but gives more blocks).
 * 0 (default): the widest the precision allows; 64, 128, 192, 256: fixed (must divide the padded FLO count:
 * 64, 128, 192 or a multiple of 256; the split-BF16 forms take at most 192); -1: autotuned at the next
 * evaluation, every legal width timed on a strided sample of <= 32768 mesh points and the fastest kept (cached
 * per process by precision, FLO count, Hermite rows and mesh size).  No reference counterpart. */
int32_t flosic_set_coulomb_cols(flosic_ctx *ctx, int32_t ncols);

/* Far chunks of the three-piece split-BF16 form (FLOSIC_PREC_BF16X3).  A (64-point tile, chunk of exponent-pair
 * groups) pair whose every (point, group) pair has T = p|r - P|^2 > tfar (p the summed exponent and P the centre of
 * the pair's Gaussian product, bounded from below through the tile's centroid and radius) is contracted with two
 * bf16 pieces of R and H (3 products) instead of three (6 products); every other pair keeps three.  tfar <= 0 turns
 * this off; the default is 30, POISSON1's THIRTYSOMETHING (poisson1.f:101), whose RCUT = sqrt(30/DELTA_MIN) + DMAX
 * (poisson1.f:528) is where the reference's default mode hands a pair batch to its multipole far field.  The bits
 * are built once per mesh, basis and threshold.  No reference counterpart. */
int32_t flosic_set_far_threshold(flosic_ctx *ctx, double tfar);

/* Mesh-block recomputation.  block_points = 0 (default): the FOD-independent basis planes (values, and for PBE
 * gradients and Hessians: 10 planes of nbas x nmsh) stay resident between evaluations and the FLO fields of the
 * whole mesh are stored during one.  block_points > 0 (rounded up to a multiple of 256): nothing of size
 * nbas x nmsh or nflo x nmsh is kept except V_H; each evaluation walks the mesh in blocks of block_points points
 * and builds, uses and drops the basis planes, FLO fields, V_tot and U of each block, as the reference's
 * SICLMSLV does over its 25-point chunks (frmorb_mpi.F:184-572).  Energies are identical in both modes (the
 * same 256-point partials); the SIC matrix and V^i_ab differ by FP64 summation order only.  No reference
 * counterpart as an entry point. */
int32_t flosic_set_mesh_block(flosic_ctx *ctx, int64_t block_points);

/* Batches of independent systems (one context each, all on one device; no counterpart in the reference, whose
 * GROUP mode spreads one system's orbitals over ranks, pasapotnl_grp.F90).  flosic_batch_evaluate is
 * flosic_evaluate of every context, with the Coulomb contractions of all systems of one precision in one launch
 * (each system at its own column-block width); zsic[i] (may be null) gets ZSIC of ctxs[i], and every per-context
 * getter then reads that context's results.  With use_graph the device part of the evaluation (Hermite
 * coefficients, batched Coulomb, fields, SIC matrix, V^i_ab of every system) is captured into one hipGraph on
 * the second evaluation and replayed on later ones while no context's settings or buffers change (moving FODs
 * keeps it valid); the FO/Lowdin algebra and the forces run on the host around it, the systems shared out over
 * up to 8 host threads.  The contexts must not be used elsewhere while the batch evaluates, and must outlive the
 * batch. */
int32_t flosic_batch_create(int32_t n, flosic_ctx *const *ctxs, int32_t use_graph, flosic_batch **batch);
int32_t flosic_batch_evaluate(flosic_batch *batch, double *zsic);
const char *flosic_batch_last_error(const flosic_batch *batch);
/* [0] host wall time of the last batch evaluation [1] its device part (batch stream events) [2] the batched
 * Coulomb launches alone (0 when the device part was a graph) [3] 0 eager, 1 captured, 2 replayed; ms */
int32_t flosic_batch_get_timings(flosic_batch *batch, int32_t n, double *ms);
/* [0] evaluations [1] graph captures [2] graph replays [3] batched Coulomb launches (one per precision form)
 * [4] Coulomb work items */
int32_t flosic_batch_get_counts(flosic_batch *batch, int32_t n, int64_t *counts);
int32_t flosic_batch_destroy(flosic_batch *batch);

#ifdef __cplusplus
}
#endif
#endif

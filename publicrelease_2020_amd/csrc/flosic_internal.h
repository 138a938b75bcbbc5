// Internal interface between the C-ABI host code (flosic_api.hip) and the device kernels
// (flosic_kernels.hip).  All pointers are device pointers unless named h_*.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flosic {

constexpr double kCutExp = 40.0;     // PARAMA2:31
constexpr int kBoysRows = 1601;      // fffmt3.F:24 table at dT = 0.01
constexpr int kBoysCols = 8;         // F_0..F_7 per row (top + 3 <= 7)

// basis as consumed by the mesh kernels: contracted shells (atom, l, contraction) with the bare
// primitives that carry a nonzero coefficient
struct BasisDev {
    int natom;
    int nbas;
    int ncs;
    const double *atom_xyz;   // 3*natom
    const double *atom_r2cut; // natom: max_i CUTEXP/alpha_i (beyond it no primitive survives)
    const int *atom_cs;       // natom+1 offsets into the shell arrays
    const int *cs_l;          // ncs
    const int *cs_f0;         // ncs: function index of Cartesian component 0
    const int *cs_stride;     // ncs: distance between Cartesian components (= ncon of the (atom,l))
    const int *cs_p0;         // ncs+1 offsets into prim arrays
    const double *prim_alpha;
    const double *prim_coef;
};

// Exponent-pair group: every primitive shell pair whose two primitives sit on primitive centres (atom, bare
// exponent) cA <= cB.  All of its shell pairs share p = a + b and P = (aA + bB)/p, hence one set of Hermite
// integrals R_tuv(p, P - r); their Hermite coefficients add into one row block of ntuv(lA + lB) rows, lA and
// lB the highest l present on each centre.  Groups are sorted by L = lA + lB.
struct PairDev {
    double A[3], B[3];
    double a, b;
    int psA[3], psB[3];  // primitive shell index of l = 0,1,2 on each centre, -1 where absent
    int lA, lB;
    int row;             // first Hermite row (rows of consecutive groups are contiguous, no padding)
    double sym;          // 2 for cA != cB, else 1
};

// chunk of consecutive groups of one order L: Hermite rows [hrow0, hrow0 + (pe-pb)*ntuv(L)) staged in LDS and
// zero-extended to kr rows (a multiple of the MFMA K step of the kernel that uses the chunk list)
constexpr int kChunkRows = 64;      // FP64 / FP32 MFMA kernels: K step 4
constexpr int kChunkRowsBf16 = 64;  // split-BF16 kernel: K step 32
constexpr int kFarWords = 1000;     // far-chunk bits per tile the split-BF16 kernel holds in LDS (32000 chunks)
struct ChunkDev {
    // groups [pb, pe); L: their Hermite order, or in the split-BF16 list 5 / 6 for the first 32 / last 3 rows of
    // L = 4 groups; kr: rows after zero extension; hrow0: first Hermite row of group pb
    int pb, pe, L, kr, hrow0;
    int prow0;  // first row of the chunk in the zero-extended row space of its list (split-BF16 H layout)
};

// One system's Coulomb contraction as the kernels see it.  H is in the operand order of the form (double or
// float fragments over K steps of 4 rows, or bf16 pieces over K steps of 32 rows: nsteps steps); ncb column
// blocks of NFB columns; V[f * ldv + k].
template <typename T>
struct CoulJob {
    const ChunkDev *chunks;
    int nchunks;
    const double4 *pP;
    const void *H;
    int nct;  // ldh / 16
    const double *X, *Y, *Z;
    int64_t M;
    int nF;
    double *V;
    int64_t ldv;
    int64_t nsteps;
    int ncb;
    int nfb;  // FLO columns per block (the single-launch batch kernels switch on it)
    // three-piece split-BF16 form: far-chunk bits of every 64-point tile (farw words per tile), or null
    const unsigned *far;
    int farw;
};

// Column-block widths the Coulomb kernels are built for: nfb must divide ldh; the split-BF16 form takes at most 192
bool coulomb_cols_legal(int nfb, int ldh, bool split);

// FLO columns handled by one Coulomb block: all of them up to 256; the H leading dimension ldh is nF rounded
// up to a multiple of this
int coulomb_cols(int nF);

// per-phase launch helpers (stream-ordered)
// host: fills kBoysRows x kBoysCols doubles followed by the same table in FP32 (kBoysTableDoubles in all)
constexpr int kBoysTableDoubles = kBoysRows * kBoysCols * 3 / 2;
void boys_table(double *h_tab);

// nd = 1 (values), 4 (+ gradient), 10 (+ Hessian xx,yy,zz,xy,xz,yz) planes, -1 (one Laplacian plane) or -3 (the
// three gradient planes alone)
template <typename T>
void launch_basis_planes(const BasisDev &B, const double *x, const double *y, const double *z, int64_t M, int nd,
                         T *phi, int64_t ldphi, hipStream_t s);

// T[f][k] = w_k (a * lap[f][k] + pot[k] * phi[f][k]); lap may be null, pot null means 1
void launch_weighted_op(int n, int64_t M, const double *w, const double *pot, const double *phi, const double *lap,
                        double a, double *T, hipStream_t s);

void launch_prim_coef(int nps, const int *ps_f0, const int *ps_stride, const int *ps_ncart, const int *ps_c0,
                      const int *ps_nc, const double *ps_coef, const double *cflo, int nF, int ldc, double *dps,
                      hipStream_t s);

// H[row][f] for every group; columns nF <= f < ldh are written as zero
void launch_hermite(int ngroup, const PairDev *pairs, const double *dps, int nF, double *H, int ldh,
                    hipStream_t s);

// V[f*M + k] = sum_pairs sum_tuv R_tuv(P - r_k) H[row+tuv][f]
void launch_coulomb(const int *segL /*6 offsets*/, const double4 *pairP, const int *pairRow, const double *H,
                    int ldh, const double *x, const double *y, const double *z, int64_t M, const double *btab,
                    int nF, double *V, int64_t ldv, hipStream_t s);

// MFMA forms.  H is first rearranged into MFMA operand order over the chunk list's zero-extended rows
// (nsteps K steps: rows/4 for the FP64/FP32 list, rows/32 for the bf16 list); the Coulomb kernels then generate
// R_tuv of every (point, group) once and stream H.  T = double (FP64 MFMA) or float (FP32 MFMA, FP64 sums per
// chunk).
template <typename T>
void launch_frag_h(const ChunkDev *chunks, int nchunks, const double *H, int ldh, int nF, T *Hf, hipStream_t s);
template <typename T>
void launch_coulomb_mfma(const ChunkDev *chunks, int nchunks, int64_t nsteps, const double4 *pairP, const T *Hf,
                         int ldh, const double *x, const double *y, const double *z, int64_t M, const double *btab,
                         int nF, double *V, int64_t ldv, hipStream_t s, int nfb = 0);

// Batched form: every job of one Coulomb form in one launch over the (system, point tile, column block) items; the
// job records and the item offsets in device memory (off[j] the first item of job j, which has ceil(M/64) * ncb items,
// tile-major; off[njobs] the total, < 2^31), each job with its own column-block width (jobs[j].nfb).
template <typename T>
void launch_coulomb_batch_all(const CoulJob<T> *jobs, const int64_t *off, int njobs, int64_t items,
                              const double *btab, hipStream_t s);
void launch_coulomb_bf16_batch_all(int npieces, const CoulJob<__bf16> *jobs, const int64_t *off, int njobs,
                                   int64_t items, const double *btab, hipStream_t s);
int coulomb_cols_split(int nF);

// split-BF16 form: R and H split into NP = 2 or 3 bf16 pieces, 3 or 6 cross products per 16x16x32 MFMA step,
// FP32 accumulation inside a chunk, FP64 sums across chunks.  H is split once per evaluation into
// Hs[piece][nsteps][ldh/16][64 lanes][8].
void launch_split_h(int npieces, const ChunkDev *chunks, int nchunks, const double *H, int ldh, int nF,
                    int64_t nsteps, __bf16 *Hs, hipStream_t s);
void launch_coulomb_bf16(int npieces, const ChunkDev *chunks, int nchunks, int64_t nsteps, const double4 *pairP,
                         const __bf16 *Hs, int ldh, const double *x, const double *y, const double *z, int64_t M,
                         const double *btab, int nF, double *V, int64_t ldv, hipStream_t s, int nfb = 0,
                         const unsigned *far = nullptr, int farw = 0);
// Far-chunk bits for the three-piece split-BF16 form: bit c of word far[t * farw + c / 32] is set when every (point,
// group) pair of 64-point tile t and chunk c of the split-BF16 chunk list has T = p|r - P|^2 > tfar (a lower bound
// from the tile's centroid and radius); farrows[0] (zeroed here) accumulates the rows of the far (tile, chunk) pairs
void launch_far_bits(const double *x, const double *y, const double *z, int64_t M, const ChunkDev *chunks,
                     int nchunks, const double4 *pairP, double tfar, int farw, unsigned *far, double *farrows,
                     hipStream_t s);

// XC + total potential + energy partials over M points: phi nd planes of [nF][M] (plane stride nF*M), Vtot and U
// [nF][M], VH[f * ldvh + k], w[k]; 256-point block b writes partial[(f * nblk + blk0 + b) * 9 + q].  For PBE,
// Ux (4 planes of [nF][M], or null) receives the exact mesh derivative of the orbital's E_xc + Hartree term:
// plane 0 w[(e_rho + V_H) phi + 2 q phi |grad phi|^2], planes 1..3 w 2 q phi^2 d_x,y,z phi with q = e_g / g
void launch_xc(int xc, int nd, const double *phi, int64_t M, int nF, const double *VH, int64_t ldvh,
               const double *w, double *Vtot, double *U, double *partial /*[nF][nblk][9]*/, int nblk, int blk0,
               hipStream_t s, double *Ux = nullptr);

// p[0..n) = 0 (a kernel: usable inside captured graphs); dst[i] = src[i * stride], i < n
void launch_fill_zero(double *p, size_t n, hipStream_t s);
void launch_sum_partials_ordered(const double *part, int nblk, int nq, double *out, hipStream_t s);
// DFT-side LSDA potential (NSPN = 2) over M points: phi values [nF][M] (FLOs 0..nup-1 spin up), VH[f * ldvh + k],
// w[k]; pot[k] spin up, pot[ldpot + k] spin down; partial[b * 5 + q] per 256-point block b of these M points
void launch_dft_lsda(const double *phi, int64_t M, int nup, int nF, const double *VH, int64_t ldvh, const double *w,
                     double *pot, int64_t ldpot, double *partial, hipStream_t s);
// DFT-side PBE potential (NSPN = 2): phi 10 planes of [nF][M] (value, gradient, Hessian); otherwise as launch_dft_lsda
void launch_dft_gga(const double *phi, int64_t M, int nup, int nF, const double *VH, int64_t ldvh, const double *w,
                    double *pot, int64_t ldpot, double *partial, hipStream_t s);
void launch_gather(const double *src, int64_t stride, int n, double *dst, hipStream_t s);

// out[f][q] (+)= sum over b of partial[f][b][q], compensated, in block order
void launch_reduce_partials(const double *partial, int nF, int nblk, int nq, double *out /*[nF][nq]*/,
                            hipStream_t s, bool accumulate = false);

// V^i_ab for one spin: phi0 (n FLOs of this spin, [n][M]), Vtot ([n][M]), w -> partial sums [slice][i][b][a]
int vab_slices(int64_t M);
void launch_vab(const double *phi0, const double *Vtot, const double *w, int64_t M, int n, int nslice,
                double *partial /*[nslice][n^3]*/, hipStream_t s, bool fp32 = false);

}  // namespace flosic

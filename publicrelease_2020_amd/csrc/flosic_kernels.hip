// Device kernels of the FLO-SIC per-Fermi-orbital evaluation, written for gfx950 (64-wide wavefronts).
//
//   k_basis_planes   contracted Cartesian Gaussians and their derivatives on the mesh (gorbdrv.f:18-316);
//                    FOD independent, computed once per system and kept resident in HBM
//   k_prim_coef      FLO coefficients folded onto primitive shells (the ADD/DMAT contraction of
//                    coupot_sic.F:398-482)
//   k_hermite        Hermite expansion of every primitive pair density for every FLO (the COEFICIENT
//                    polynomial set of poisson1.f:295-496, McMurchie-Davidson form), 2 pi/p folded in
//   k_coulomb        exact Gaussian-pair potential at every mesh point for all FLOs at once: Boys
//                    functions by the table/Taylor/asymptote scheme of fffmt3.F/fffmt4.F/fffmtc.F,
//                    Hermite recursion in registers, contraction with the pair coefficients
//   k_xc             spin-polarised LSDA/PBE for rho_up = rho_i, rho_dn = 0 (subvlxc_libxc.F90), total SIC
//                    potential and per-block partial sums of the RESULTS energies (apotnl_sic.F:536-581)
//   k_vab_mfma       V^i_ab = sum_k w phi_a V_i phi_b as an FP64 MFMA GEMM over point slices
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "flosic_internal.h"

namespace flosic {

// ---------------------------------------------------------------------------------------------------
// Cartesian component tables: index 0..9 = s, x, y, z, xx, yy, zz, xy, xz, yz (gorbdrv.f:78-98)
// ---------------------------------------------------------------------------------------------------
__host__ __device__ constexpr int cpow(int idx, int ax) {
    return idx == 0 ? 0
           : idx <= 3 ? (idx - 1 == ax ? 1 : 0)
           : idx <= 6 ? (idx - 4 == ax ? 2 : 0)
           : idx == 7 ? (ax < 2 ? 1 : 0)
           : idx == 8 ? (ax != 1 ? 1 : 0)
                      : (ax > 0 ? 1 : 0);
}
__host__ __device__ constexpr int ncart(int l) { return l == 0 ? 1 : (l == 1 ? 3 : 6); }
__host__ __device__ constexpr int loff(int l) { return l == 0 ? 0 : (l == 1 ? 1 : 4); }

// value (+ gradient, + Hessian xx,yy,zz,xy,xz,yz) of component IDX times the radial sums
// S0 = sum c e^{-a r^2}, S1 = sum c a e^{-a r^2}, S2 = sum c a^2 e^{-a r^2}
template <int IDX, int ND>
__device__ __forceinline__ void cart_eval(double x, double y, double z, double S0, double S1, double S2,
                                          double *o) {
    constexpr int a = cpow(IDX, 0), b = cpow(IDX, 1), c = cpow(IDX, 2);
    const double px[3] = {1.0, x, x * x}, py[3] = {1.0, y, y * y}, pz[3] = {1.0, z, z * z};
    const double P = px[a] * py[b] * pz[c];
    o[0] = P * S0;
    if constexpr (ND >= 4) {
        const double dPx = a ? a * px[a - 1] * py[b] * pz[c] : 0.0;
        const double dPy = b ? b * px[a] * py[b - 1] * pz[c] : 0.0;
        const double dPz = c ? c * px[a] * py[b] * pz[c - 1] : 0.0;
        o[1] = dPx * S0 - 2.0 * x * P * S1;
        o[2] = dPy * S0 - 2.0 * y * P * S1;
        o[3] = dPz * S0 - 2.0 * z * P * S1;
        if constexpr (ND == 10) {
            const double dPxx = a >= 2 ? a * (a - 1) * py[b] * pz[c] : 0.0;
            const double dPyy = b >= 2 ? b * (b - 1) * px[a] * pz[c] : 0.0;
            const double dPzz = c >= 2 ? c * (c - 1) * px[a] * py[b] : 0.0;
            const double dPxy = (a && b) ? a * b * px[a - 1] * py[b - 1] * pz[c] : 0.0;
            const double dPxz = (a && c) ? a * c * px[a - 1] * py[b] * pz[c - 1] : 0.0;
            const double dPyz = (b && c) ? b * c * px[a] * py[b - 1] * pz[c - 1] : 0.0;
            o[4] = dPxx * S0 - 4.0 * x * dPx * S1 + P * (4.0 * x * x * S2 - 2.0 * S1);
            o[5] = dPyy * S0 - 4.0 * y * dPy * S1 + P * (4.0 * y * y * S2 - 2.0 * S1);
            o[6] = dPzz * S0 - 4.0 * z * dPz * S1 + P * (4.0 * z * z * S2 - 2.0 * S1);
            o[7] = dPxy * S0 - 2.0 * x * dPy * S1 - 2.0 * y * dPx * S1 + 4.0 * x * y * P * S2;
            o[8] = dPxz * S0 - 2.0 * x * dPz * S1 - 2.0 * z * dPx * S1 + 4.0 * x * z * P * S2;
            o[9] = dPyz * S0 - 2.0 * y * dPz * S1 - 2.0 * z * dPy * S1 + 4.0 * y * z * P * S2;
        }
    }
}

// ND = 1, 4 or 10 planes (value, gradient, Hessian); ND = -1: one plane holding the Laplacian
template <typename T, int ND, int L, int M0>
__device__ __forceinline__ void store_comps(double x, double y, double z, double S0, double S1, double S2, T *phi,
                                            int64_t ldphi, int64_t plane, int f0, int stride, int64_t k) {
    if constexpr (M0 < ncart(L)) {
        const int64_t col = (int64_t)(f0 + M0 * stride) * ldphi + k;
        if constexpr (ND == -1) {
            double o[10];
            cart_eval<loff(L) + M0, 10>(x, y, z, S0, S1, S2, o);
            phi[col] = (T)(o[4] + o[5] + o[6]);
        } else if constexpr (ND == -3) {
            double o[4];
            cart_eval<loff(L) + M0, 4>(x, y, z, S0, S1, S2, o);
#pragma unroll
            for (int d = 0; d < 3; ++d) phi[d * plane + col] = (T)o[1 + d];
        } else {
            double o[ND];
            cart_eval<loff(L) + M0, ND>(x, y, z, S0, S1, S2, o);
#pragma unroll
            for (int d = 0; d < ND; ++d) phi[d * plane + col] = (T)o[d];
        }
        store_comps<T, ND, L, M0 + 1>(x, y, z, S0, S1, S2, phi, ldphi, plane, f0, stride, k);
    }
}

// one thread per mesh point; planes phi[d][mu][k] (leading dimension ldphi >= M, plane = nbas*ldphi);
// the planes must be zeroed beforehand (screened functions are not written)
template <typename T, int ND>
__global__ void __launch_bounds__(256) k_basis_planes(BasisDev B, const double *__restrict__ X,
                                                      const double *__restrict__ Y, const double *__restrict__ Z,
                                                      int64_t M, T *__restrict__ phi, int64_t ldphi) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const double px = X[k], py = Y[k], pz = Z[k];
    const int64_t plane = (int64_t)B.nbas * ldphi;
    const double geps = exp(-kCutExp);
    for (int a = 0; a < B.natom; ++a) {
        const double x = px - B.atom_xyz[3 * a], y = py - B.atom_xyz[3 * a + 1], z = pz - B.atom_xyz[3 * a + 2];
        const double r2 = x * x + y * y + z * z;
        if (!(r2 < B.atom_r2cut[a])) continue;
        for (int s = B.atom_cs[a]; s < B.atom_cs[a + 1]; ++s) {
            double S0 = 0.0, S1 = 0.0, S2 = 0.0;
            for (int p = B.cs_p0[s]; p < B.cs_p0[s + 1]; ++p) {
                const double al = B.prim_alpha[p];
                const double arg = al * r2;
                if (!(arg < kCutExp)) continue;
                const double g = exp(-arg);
                const double cc = B.prim_coef[p];
                if (!(fabs(cc) * g > geps)) continue;
                const double gc = g * cc;
                S0 += gc;
                S1 += gc * al;
                S2 += gc * al * al;
            }
            const int l = B.cs_l[s];
            if (l == 0) store_comps<T, ND, 0, 0>(x, y, z, S0, S1, S2, phi, ldphi, plane, B.cs_f0[s], B.cs_stride[s], k);
            else if (l == 1) store_comps<T, ND, 1, 0>(x, y, z, S0, S1, S2, phi, ldphi, plane, B.cs_f0[s], B.cs_stride[s], k);
            else store_comps<T, ND, 2, 0>(x, y, z, S0, S1, S2, phi, ldphi, plane, B.cs_f0[s], B.cs_stride[s], k);
        }
    }
}

template <typename T>
void launch_basis_planes(const BasisDev &B, const double *x, const double *y, const double *z, int64_t M, int nd,
                         T *phi, int64_t ldphi, hipStream_t s) {
    const dim3 grid((unsigned)((M + 255) / 256)), blk(256);
    if (nd == 1) hipLaunchKernelGGL((k_basis_planes<T, 1>), grid, blk, 0, s, B, x, y, z, M, phi, ldphi);
    else if (nd == 4) hipLaunchKernelGGL((k_basis_planes<T, 4>), grid, blk, 0, s, B, x, y, z, M, phi, ldphi);
    else if (nd == -1) hipLaunchKernelGGL((k_basis_planes<T, -1>), grid, blk, 0, s, B, x, y, z, M, phi, ldphi);
    else if (nd == -3) hipLaunchKernelGGL((k_basis_planes<T, -3>), grid, blk, 0, s, B, x, y, z, M, phi, ldphi);
    else hipLaunchKernelGGL((k_basis_planes<T, 10>), grid, blk, 0, s, B, x, y, z, M, phi, ldphi);
}
template void launch_basis_planes<double>(const BasisDev &, const double *, const double *, const double *, int64_t,
                                          int, double *, int64_t, hipStream_t);
template void launch_basis_planes<float>(const BasisDev &, const double *, const double *, const double *, int64_t,
                                         int, float *, int64_t, hipStream_t);

// T[f][k] = w_k (a * lap[f][k] + pot[k] * phi[f][k]) over n FLOs (lap may be null: T = w * pot * phi, or with
// pot null: T = w * phi)
__global__ void k_weighted_op(int n, int64_t M, const double *__restrict__ w, const double *__restrict__ pot,
                              const double *__restrict__ phi, const double *__restrict__ lap, double a,
                              double *__restrict__ T) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * M) return;
    const int64_t k = e % M;
    double v = pot ? pot[k] * phi[e] : phi[e];
    if (lap) v += a * lap[e];
    T[e] = w[k] * v;
}
void launch_weighted_op(int n, int64_t M, const double *w, const double *pot, const double *phi, const double *lap,
                        double a, double *T, hipStream_t s) {
    const int64_t cnt = (int64_t)n * M;
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_weighted_op, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, n, M, w, pot, phi, lap, a,
                       T);
}

// ---------------------------------------------------------------------------------------------------
// FLO coefficients on primitive shells: dps[(ps*6 + m)*nF + f] = sum_c con_c * cflo[f*ldc + f0 + m*stride + c]
// ---------------------------------------------------------------------------------------------------
__global__ void k_prim_coef(int nps, const int *__restrict__ ps_f0, const int *__restrict__ ps_stride,
                            const int *__restrict__ ps_ncart, const int *__restrict__ ps_c0,
                            const int *__restrict__ ps_nc, const double *__restrict__ ps_coef,
                            const double *__restrict__ cflo, int nF, int ldc, double *__restrict__ dps) {
    const int ps = blockIdx.x;
    const int f0 = ps_f0[ps], st = ps_stride[ps], nc = ps_nc[ps], c0 = ps_c0[ps], nm = ps_ncart[ps];
    for (int f = threadIdx.x; f < nF; f += blockDim.x) {
        for (int m = 0; m < 6; ++m) {
            double s = 0.0;
            if (m < nm)
                for (int c = 0; c < nc; ++c) s += ps_coef[c0 + c] * cflo[(int64_t)f * ldc + f0 + m * st + c];
            dps[((int64_t)ps * 6 + m) * nF + f] = s;
        }
    }
}

void launch_prim_coef(int nps, const int *ps_f0, const int *ps_stride, const int *ps_ncart, const int *ps_c0,
                      const int *ps_nc, const double *ps_coef, const double *cflo, int nF, int ldc, double *dps,
                      hipStream_t s) {
    if (nps <= 0) return;
    hipLaunchKernelGGL(k_prim_coef, dim3(nps), dim3(128), 0, s, nps, ps_f0, ps_stride, ps_ncart, ps_c0, ps_nc,
                       ps_coef, cflo, nF, ldc, dps);
}

// ---------------------------------------------------------------------------------------------------
// Hermite index tables: entries grouped by total order L' = 0..4; within an order t descending, then u
// descending (t+u+v = L')
// ---------------------------------------------------------------------------------------------------
__host__ __device__ constexpr int ntuv(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// Chunk types of the split-BF16 chunk list: L = 0..4 whole groups of that L; L = 5 the first 32 Hermite rows of L = 4
// groups, L = 6 their last 3 rows (a 35-row group would otherwise fill a 64-row chunk with 29 zero rows).  Rows per
// group in the chunk, offset of those rows inside the group's Hermite block, and the group's Hermite order:
__host__ __device__ constexpr int chunk_rpg(int L) { return L == 5 ? 32 : L == 6 ? 3 : ntuv(L); }
__host__ __device__ constexpr int chunk_roff(int L) { return L == 6 ? 32 : 0; }
__host__ __device__ constexpr int chunk_herm(int L) { return L > 4 ? 4 : L; }
__host__ __device__ constexpr int tuv_idx(int t, int u, int v) {
    // position inside order block: enumerate t from L' down, u from L'-t down
    int L = t + u + v;
    int base = L == 0 ? 0 : ntuv(L - 1);
    int pos = 0;
    for (int tt = L; tt > t; --tt) pos += L - tt + 1;
    pos += (L - t) - u;
    return base + pos;
}
__host__ __device__ constexpr int tuv_comp(int i, int which) {
    for (int L = 0; L <= 4; ++L)
        for (int t = L; t >= 0; --t)
            for (int u = L - t; u >= 0; --u) {
                int v = L - t - u;
                if (tuv_idx(t, u, v) == i) return which == 0 ? t : (which == 1 ? u : v);
            }
    return -1;
}

// ---------------------------------------------------------------------------------------------------
// Hermite coefficients of every pair for every FLO.  E^{ij}_t along one axis by the McMurchie-Davidson
// recursion; H[row+tuv][f] = sym * 2pi/p * sum_{mu,nu} dA[mu][f] dB[nu][f] Ex Ey Ez
// ---------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(128) k_hermite(const PairDev *__restrict__ pairs, const double *__restrict__ dps,
                                                  int nF, double *__restrict__ H, int ldh) {
    __shared__ double E[3][3][3][5];
    const PairDev pr = pairs[blockIdx.x];
    const double p = pr.a + pr.b;
    if (threadIdx.x < 3) {
        const int d = threadIdx.x;
        const double P = (pr.a * pr.A[d] + pr.b * pr.B[d]) / p;
        const double xpa = P - pr.A[d], xpb = P - pr.B[d];
        const double dd = pr.A[d] - pr.B[d];
        const double K = exp(-pr.a * pr.b / p * dd * dd);
        const double inv2p = 0.5 / p;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int t = 0; t < 5; ++t) E[d][i][j][t] = 0.0;
        E[d][0][0][0] = K;
        for (int i = 0; i <= 2; ++i)
            for (int j = 0; j <= 2; ++j) {
                if (i == 0 && j == 0) continue;
                const int pi = i > 0 ? i - 1 : i, pj = i > 0 ? j : j - 1;
                const double shift = i > 0 ? xpa : xpb;
                const int np = i + j - 1;
                for (int t = 0; t <= i + j; ++t) {
                    double v = 0.0;
                    if (t >= 1 && t - 1 <= np) v += inv2p * E[d][pi][pj][t - 1];
                    if (t <= np) v += shift * E[d][pi][pj][t];
                    if (t + 1 <= np) v += (t + 1) * E[d][pi][pj][t + 1];
                    E[d][i][j][t] = v;
                }
            }
    }
    __syncthreads();
    const int L = pr.lA + pr.lB;
    const int nt = ntuv(L);
    const double fac = pr.sym * 2.0 * M_PI / p;
    for (int f = threadIdx.x; f < ldh; f += blockDim.x) {
        if (f >= nF) {
            for (int t = 0; t < nt; ++t) H[(int64_t)(pr.row + t) * ldh + f] = 0.0;
            continue;
        }
        double h[35];
        for (int t = 0; t < 35; ++t) h[t] = 0.0;
        for (int la = 0; la <= pr.lA; ++la) {
            if (pr.psA[la] < 0) continue;
            for (int m = 0; m < ncart(la); ++m) {
                const double da = dps[((int64_t)pr.psA[la] * 6 + m) * nF + f];
                if (da == 0.0) continue;
                const int ax = cpow(loff(la) + m, 0), ay = cpow(loff(la) + m, 1), az = cpow(loff(la) + m, 2);
                for (int lb = 0; lb <= pr.lB; ++lb) {
                    if (pr.psB[lb] < 0) continue;
                    for (int n = 0; n < ncart(lb); ++n) {
                        const double dab = da * dps[((int64_t)pr.psB[lb] * 6 + n) * nF + f];
                        if (dab == 0.0) continue;
                        const int bx = cpow(loff(lb) + n, 0), by = cpow(loff(lb) + n, 1), bz = cpow(loff(lb) + n, 2);
                        for (int t = 0; t <= ax + bx; ++t) {
                            const double et = dab * E[0][ax][bx][t];
                            for (int u = 0; u <= ay + by; ++u) {
                                const double etu = et * E[1][ay][by][u];
                                for (int v = 0; v <= az + bz; ++v) h[tuv_idx(t, u, v)] += etu * E[2][az][bz][v];
                            }
                        }
                    }
                }
            }
        }
        for (int t = 0; t < nt; ++t) H[(int64_t)(pr.row + t) * ldh + f] = fac * h[t];
    }
}

void launch_hermite(int ngroup, const PairDev *pairs, const double *dps, int nF, double *H, int ldh, hipStream_t s) {
    if (ngroup <= 0) return;
    hipLaunchKernelGGL(k_hermite, dim3(ngroup), dim3(128), 0, s, pairs, dps, nF, H, ldh);
}

// ---------------------------------------------------------------------------------------------------
// Boys functions F_0..F_TOP (fffmt3.F:66-88): T > 16 one-term asymptote + upward recursion; otherwise
// cubic Taylor of F_TOP about the nearest 0.01 node and downward recursion
// ---------------------------------------------------------------------------------------------------
// Boys F_0..F_TOP of the reference scheme (fffmt3.F): T > 16 one-term asymptote and upward recursion, else the
// cubic Taylor expansion of F_TOP about the nearest table node and downward recursion.  C = double for the FP64
// forms; C = float (table rounded to FP32) where R feeds FP32 or split-BF16 MFMA operands anyway.
template <int TOP, typename C>
__device__ __forceinline__ void boys(C T, C eT, const C *__restrict__ tab, C *F) {
    if (T > C(16.0)) {
        // FP32: the hardware reciprocal and square root (1 ulp) instead of correctly rounded ones, well inside the
        // FP32 R's own rounding; FP64 keeps the reference scheme's operations
        C rect1, rt;
        if constexpr (sizeof(C) == 4) {
            rect1 = 0.5f * __builtin_amdgcn_rcpf(T);
            rt = __builtin_amdgcn_sqrtf(T);
        } else {
            rect1 = C(0.5) / T;
            rt = sqrt(T);
        }
        F[0] = rect1 * (rt * C(1.77245385090551602728) - eT);
#pragma unroll
        for (int m = 1; m <= TOP; ++m) F[m] = rect1 * ((2 * m - 1) * F[m - 1] - eT);
    } else {
        const int it = (int)(C(100.0) * T + C(1.5));
        const C dt = C(1.0e-2) * (it - 1) - T;
        const C *f = tab + (it - 1) * kBoysCols;
        if constexpr (sizeof(C) == 4)  // FP32: times the rounded 1/6 (a correctly rounded division costs 10 VALU)
            F[TOP] = f[TOP] + dt * (f[TOP + 1] + dt * (f[TOP + 2] * 0.5f + dt * (f[TOP + 3] * (1.0f / 6.0f))));
        else
            F[TOP] = f[TOP] + dt * (f[TOP + 1] + dt * (f[TOP + 2] / C(2.0) + dt * (f[TOP + 3] / C(6.0))));
        const C rect1 = C(2.0) * T;
#pragma unroll
        for (int m = TOP - 1; m >= 0; --m) {
            const C r = F[m + 1] * rect1 + eT;
            F[m] = (m == 0) ? r : r * (C(1.0) / (2 * m + 1));
        }
    }
}

// one Hermite step R^{(n)}_I from R^{(n+1)} in place (entries of lower order still hold level n+1)
template <int I, typename C>
__device__ __forceinline__ void rstep(C *R, C X, C Y, C Z) {
    constexpr int t = tuv_comp(I, 0), u = tuv_comp(I, 1), v = tuv_comp(I, 2);
    if constexpr (t > 0) {
        C r = X * R[tuv_idx(t - 1, u, v)];
        if constexpr (t > 1) r += (t - 1) * R[tuv_idx(t - 2, u, v)];
        R[I] = r;
    } else if constexpr (u > 0) {
        C r = Y * R[tuv_idx(t, u - 1, v)];
        if constexpr (u > 1) r += (u - 1) * R[tuv_idx(t, u - 2, v)];
        R[I] = r;
    } else {
        C r = Z * R[tuv_idx(t, u, v - 1)];
        if constexpr (v > 1) r += (v - 1) * R[tuv_idx(t, u, v - 2)];
        R[I] = r;
    }
}
template <int I, typename C>
__device__ __forceinline__ void rdown(C *R, C X, C Y, C Z) {
    if constexpr (I >= 1) {
        rstep<I, C>(R, X, Y, Z);
        rdown<I - 1, C>(R, X, Y, Z);
    }
}
template <int L, int N, typename C>
__device__ __forceinline__ void rlevels(C *R, C X, C Y, C Z, const C *F, const C *pw) {
    if constexpr (N >= 0) {
        rdown<ntuv(L - N) - 1, C>(R, X, Y, Z);
        R[0] = pw[N] * F[N];
        rlevels<L, N - 1, C>(R, X, Y, Z, F, pw);
    }
}

// R_tuv(p, P - C) for all t+u+v <= L, TUV order
template <int L, typename C = double>
__device__ __forceinline__ void hermite_R(C p, C X, C Y, C Z, const C *__restrict__ tab, C *R) {
    constexpr int TOP = L < 2 ? 2 : L;
    const C T = p * (X * X + Y * Y + Z * Z);
    C F[TOP + 1];
    C eT;
    if constexpr (sizeof(C) == 8)
        // beyond T = 60, e^-T (< 9e-27) is below half an ulp of every term it enters (F_4(60) ~ 6e-8), so
        // dropping it leaves each F_m bitwise unchanged and saves the exponential for most (point, group) pairs
        eT = T > 60.0 ? 0.0 : exp(-T);
    else  // FP32: the hardware exponential of -T log2(e) (relative error < 1e-6 at T = 16, where e^-T adds 1e-13)
        eT = __expf(-T);
    boys<TOP, C>(T, eT, tab, F);
    C pw[L + 1];
    pw[0] = C(1.0);
#pragma unroll
    for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (C(-2.0) * p);
    R[0] = pw[L] * F[L];
    rlevels<L, L - 1, C>(R, X, Y, Z, F, pw);
}

template <int L, int CH>
__device__ __forceinline__ void coulomb_segment(int beg, int end, const double4 *__restrict__ pP,
                                                const int *__restrict__ prow, const double *__restrict__ H, int ldh,
                                                int f0, double x, double y, double z, const double *__restrict__ tab,
                                                double *acc) {
    for (int ip = beg; ip < end; ++ip) {
        const double4 P = pP[ip];
        const int row = prow[ip];
        double R[ntuv(L)];
        hermite_R<L>(P.w, P.x - x, P.y - y, P.z - z, tab, R);
        const double *h = H + (int64_t)row * ldh + f0;
#pragma unroll
        for (int t = 0; t < ntuv(L); ++t) {
#pragma unroll
            for (int c = 0; c < CH; ++c) acc[c] = fma(R[t], h[(int64_t)t * ldh + c], acc[c]);
        }
    }
}

// block = 4 wavefronts; wave w of block (bx, by) handles points bx*64 .. +63 and FLO chunk (by*4 + w)
template <int CH>
__global__ void __launch_bounds__(256) k_coulomb(int s0, int s1, int s2, int s3, int s4, int s5,
                                                  const double4 *__restrict__ pP, const int *__restrict__ prow,
                                                  const double *__restrict__ H, int ldh, const double *__restrict__ X,
                                                  const double *__restrict__ Y, const double *__restrict__ Z,
                                                  int64_t M, const double *__restrict__ tab, int nF,
                                                  double *__restrict__ V, int64_t ldv) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int f0 = (blockIdx.y * 4 + wave) * CH;
    if (f0 >= nF) return;
    const int64_t k = (int64_t)blockIdx.x * 64 + lane;
    const bool valid = k < M;
    const double x = valid ? X[k] : 0.0, y = valid ? Y[k] : 0.0, z = valid ? Z[k] : 0.0;
    double acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0;
    coulomb_segment<0, CH>(s0, s1, pP, prow, H, ldh, f0, x, y, z, tab, acc);
    coulomb_segment<1, CH>(s1, s2, pP, prow, H, ldh, f0, x, y, z, tab, acc);
    coulomb_segment<2, CH>(s2, s3, pP, prow, H, ldh, f0, x, y, z, tab, acc);
    coulomb_segment<3, CH>(s3, s4, pP, prow, H, ldh, f0, x, y, z, tab, acc);
    coulomb_segment<4, CH>(s4, s5, pP, prow, H, ldh, f0, x, y, z, tab, acc);
    if (valid) {
#pragma unroll
        for (int c = 0; c < CH; ++c)
            if (f0 + c < nF) V[(int64_t)(f0 + c) * ldv + k] = acc[c];
    }
}

void launch_coulomb(const int *segL, const double4 *pairP, const int *pairRow, const double *H, int ldh,
                    const double *x, const double *y, const double *z, int64_t M, const double *btab, int nF, double *V,
                    int64_t ldv, hipStream_t s) {
    constexpr int CH = 16;
    const int nchunk = (nF + CH - 1) / CH;
    const dim3 grid((unsigned)((M + 63) / 64), (unsigned)((nchunk + 3) / 4)), blk(256);
    hipLaunchKernelGGL((k_coulomb<CH>), grid, blk, 0, s, segL[0], segL[1], segL[2], segL[3], segL[4], segL[5], pairP,
                       pairRow, H, ldh, x, y, z, M, btab, nF, V, ldv);
}

// ---------------------------------------------------------------------------------------------------
// MFMA Coulomb.  V[f][k] = sum_rows R[k][row] H[row][f]: R_tuv(p, P - r_k) is generated on the fly (FP64 Boys
// and Hermite recursion), H is the Hermite-coefficient matrix of every exponent-pair group for every FLO, stored
// in MFMA operand order.  Block = 16 wavefronts on a tile of TP = 64 points x NFB FLO columns (NFB covers all
// FLOs up to 256, so R is computed once per (point, group)).  A chunk of Hermite rows is consumed in sub-steps
// of SR rows; per sub-step, separated by one barrier:
//   * R waves generate their share of R of the next chunk into the other of two LDS tiles,
//   * copy waves move the H rows of the sub-step two ahead into a 3-slot LDS ring with asynchronous
//     global->LDS loads (global_load_lds_dwordx4: lane l of a wave lands at base + 16 l bytes), then wait until
//     the rows of the next sub-step have landed (vmcnt counts only their own copies),
//   * 8 consumer waves (2 point halves x 4 column quarters) run the MFMAs of the current sub-step from LDS.
// FP32 and split-BF16 accumulators are added into FP64 sums after every chunk.
// ---------------------------------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kCoulThreads = 1024;  // 16 wavefronts: R waves, copy waves, then 8 consumer waves
constexpr int kConsumerWave0 = 8;
constexpr int kSlots = 3;     // FP operand stream: ring of LDS slots filled by asynchronous global->LDS loads
// split-BF16: the consumers' FP32 accumulators go into their FP64 sums once at least this many Hermite rows have
// been accumulated (at chunk ends, and after the last chunk).  On the K=64 workload (same box): every 64 rows
// 1109 ms, 128 rows 1083 ms, 256 rows 1060-1070 ms, 512 rows 1053 ms, with |dE| against FP64 of 4.9e-7, 5.3e-7,
// 4.4e-7 and 6.2e-7 Ha; the flush (24 FP64 conversions and adds per consumer lane) is VALU work that the kernel's
// sub-steps wait for.  (The FP32 MFMA form keeps its flush per chunk: every 256 rows saved it under 1% there.)
constexpr int kFlushRowsBf16 = 256;
constexpr int kSlotsReg = 2;  // split-BF16: H staged through the copy waves' registers, the sub-step in use and the next

typedef __attribute__((address_space(3))) void lds_void;

// s_waitcnt with vmcnt <= n and no wait on the other counters (gfx9 encoding: vmcnt bits 3:0 and 15:14)
template <int N>
__device__ __forceinline__ void wait_vm_le() {
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
__device__ __forceinline__ void wait_vm(int n) {
    switch (n) {
#define FLOSIC_W(k) \
    case k: wait_vm_le<k>(); break;
        FLOSIC_W(0) FLOSIC_W(1) FLOSIC_W(2) FLOSIC_W(3) FLOSIC_W(4) FLOSIC_W(5) FLOSIC_W(6) FLOSIC_W(7)
        FLOSIC_W(8) FLOSIC_W(9) FLOSIC_W(10) FLOSIC_W(11) FLOSIC_W(12) FLOSIC_W(13) FLOSIC_W(14) FLOSIC_W(15)
        FLOSIC_W(16) FLOSIC_W(17) FLOSIC_W(18) FLOSIC_W(19) FLOSIC_W(20) FLOSIC_W(21) FLOSIC_W(22) FLOSIC_W(23)
        FLOSIC_W(24) FLOSIC_W(25) FLOSIC_W(26) FLOSIC_W(27) FLOSIC_W(28) FLOSIC_W(29) FLOSIC_W(30) FLOSIC_W(31)
#undef FLOSIC_W
        default: wait_vm_le<0>(); break;
    }
}

// KiB units [0, units) of a contiguous 16-byte aligned global run into LDS; unit u goes to copy wave
// u % ncw; returns the number of loads this wave issued
__device__ __forceinline__ int lds_copy_async(const char *src, char *dst, int units, int cw, int ncw, int lane,
                                              int first) {
    int n = 0;
    for (int u = (cw - first % ncw + ncw) % ncw; u < units; u += ncw, ++n)
        __builtin_amdgcn_global_load_lds((const void *)(src + (size_t)u * 1024 + lane * 16),
                                         (lds_void *)(dst + (size_t)u * 1024), 16, 0, 0);
    return n;
}

// Loads and stores through a pointer typed as global memory.  The batch kernels read their arrays' addresses from a
// job table in device memory, and to the compiler such an address is generic: accesses through it become flat
// instructions, which also count against the LDS counter, so a copy wave's wait for its LDS stores would drain its
// global loads in flight as well.  The Coulomb tiles load and store through these (with scalar or ext_vector
// types: a class type such as double4 would be copied through its generic-address copy constructor).
template <typename T>
__device__ __forceinline__ T gload(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
template <typename T>
__device__ __forceinline__ void gstore(T *p, T v) {
    *(__attribute__((address_space(1))) T *)p = v;
}
// Load of read-only data at a wave-uniform address through the constant address space: a scalar load (lgkm
// counter), which leaves the wave's vector loads in flight.  A vector load there would need a vmcnt wait before
// its value could be made uniform, draining the copy waves' H loads with it.
template <typename T>
__device__ __forceinline__ T cload(const T *p) {
    return *(const __attribute__((address_space(4))) T *)p;
}

// chunk record c with every field read into scalar registers.  The tile code loops and branches on these fields; read
// through a pointer that the compiler cannot prove unwritten during the kernel (the batch kernels take it from a job
// record in device memory) they would otherwise be per-lane values, and those loops divergent ones.
__device__ __forceinline__ ChunkDev chunk_at(const ChunkDev *__restrict__ chunks, int c) {
    static_assert(sizeof(ChunkDev) % 4 == 0, "chunk record in 32-bit words");
    constexpr int NW = sizeof(ChunkDev) / 4;
    const unsigned *src = reinterpret_cast<const unsigned *>(chunks + c);
    unsigned w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = __builtin_amdgcn_readfirstlane(cload(src + i));
    ChunkDev ch;
    __builtin_memcpy(&ch, w, sizeof(ch));
    return ch;
}
// position in the chunk / sub-step sequence
struct SubStep {
    int c, q;
};
__device__ __forceinline__ void advance(SubStep &p, const ChunkDev *chunks, int sr) {
    if (++p.q * sr >= chunk_at(chunks, p.c).kr) {
        p.q = 0;
        ++p.c;
    }
}

// one wave64 fragment of H in operand order: lane l holds row 4s + l/16 and column 16*ct + l%16 (FP64/FP32)
template <typename T>
__global__ void k_frag_h(const ChunkDev *__restrict__ chunks, const double *__restrict__ H, int ldh, int nF,
                         T *__restrict__ Hf) {
    const ChunkDev ch = chunks[blockIdx.x];
    const int nreal = (ch.pe - ch.pb) * ntuv(ch.L);
    const int nct = ldh / 16;
    for (int e = threadIdx.x; e < ch.kr * ldh; e += blockDim.x) {
        const int r = e / ldh, f = e - r * ldh;
        const double v = (r < nreal && f < nF) ? H[(int64_t)(ch.hrow0 + r) * ldh + f] : 0.0;
        const int64_t s = (ch.prow0 + r) / 4;
        Hf[(s * nct + f / 16) * 64 + (r % 4) * 16 + f % 16] = (T)v;
    }
}

template <typename T> struct MfmaTraits;
// Operands of both 16x16x4 forms: lane l holds A[l%16][l/16] and B[l/16][l%16].  The accumulators differ:
// FP32 register r of lane l holds D[4*(l/16) + r][l%16], FP64 register r holds D[l/16 + 4*r][l%16].
template <> struct MfmaTraits<double> {
    using acc = v4d;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lr, int r) { return lr + 4 * r; }
};
template <> struct MfmaTraits<float> {
    using acc = v4f;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lr, int r) { return 4 * lr + r; }
};

// R of the groups [np*q/nq, np*(q+1)/nq) of chunk `ch` (share q of nq) for the block's 64 points: R wave rw of
// nrw takes every nrw-th group (wave-uniform, so the group's centre and exponent are scalar loads) and lane l
// the point k0 + l, whose coordinates (px, py, pz) stay in registers.  LAYOUT 0: Rl[row * LD + point] in T;
// LAYOUT 2: Rl[point * LD + row] in T.  Share 0 also zeroes the rows beyond the groups.
// R is computed in FP64 for the FP64 form and in FP32 for the FP32 and split-BF16 operands, with the Boys table of
// that precision (the FP32 copy follows the FP64 table in device memory; the split-BF16 kernel reads it from LDS)
template <typename T>
using RType = typename std::conditional<sizeof(T) == 8, double, float>::type;

// RPG rows per group from Hermite row ROFF on (the whole group by default; the split-BF16 list's L = 4 head / tail
// chunks take rows [0, 32) / [32, 35) of each group, see chunk_rpg)
template <int L, typename T, int LD, int LAYOUT, int NP, int RPG = ntuv(L), int ROFF = 0>
__device__ __forceinline__ void chunk_R(int rw, int nrw, int lane, double px, double py, double pz, bool valid, int q,
                                        int nq, const ChunkDev &ch, const double4 *__restrict__ pP,
                                        const RType<T> *__restrict__ ctab, T *Rl) {
    constexpr int nt = ntuv(L);
    static_assert(ROFF + RPG <= nt, "chunk rows beyond the group's Hermite block");
    using C = RType<T>;
    const int np = ch.pe - ch.pb;
    const int glo = np * q / nq, ghi = np * (q + 1) / nq;
    for (int g = glo + rw; g < ghi; g += nrw) {
        const v4d P = gload(reinterpret_cast<const v4d *>(pP + ch.pb + g));  // x, y, z, p
        // lanes past the mesh end (valid false, coordinates 0) compute finite values too: an A row feeds only its own
        // point's output row, which is never stored, so no lane needs its R zeroed (that took nt moves per group)
        C R[nt];
        hermite_R<L, C>((C)P.w, (C)(P.x - px), (C)(P.y - py), (C)(P.z - pz), ctab, R);
#pragma unroll
        for (int t = 0; t < RPG; ++t) {
            if constexpr (LAYOUT == 0)
                Rl[(g * RPG + t) * LD + lane] = (T)R[ROFF + t];
            else
                Rl[lane * LD + g * RPG + t] = (T)R[ROFF + t];
        }
    }
    if (q == 0)
        for (int r = np * RPG + rw; r < ch.kr; r += nrw) {
            if constexpr (LAYOUT == 0)
                Rl[r * LD + lane] = (T)0;
            else
                Rl[lane * LD + r] = (T)0;
        }
}

template <typename T, int LD, int LAYOUT, int NP>
__device__ __forceinline__ void produce_R(int rw, int nrw, int lane, double px, double py, double pz, bool valid,
                                          int q, int nq, const ChunkDev &ch, const double4 *__restrict__ pP,
                                          const RType<T> *__restrict__ tab, T *Rl) {
    switch (ch.L) {
        case 0: chunk_R<0, T, LD, LAYOUT, NP>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP, tab, Rl); break;
        case 1: chunk_R<1, T, LD, LAYOUT, NP>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP, tab, Rl); break;
        case 2: chunk_R<2, T, LD, LAYOUT, NP>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP, tab, Rl); break;
        case 3: chunk_R<3, T, LD, LAYOUT, NP>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP, tab, Rl); break;
        case 4: chunk_R<4, T, LD, LAYOUT, NP>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP, tab, Rl); break;
        case 5:
            chunk_R<4, T, LD, LAYOUT, NP, chunk_rpg(5), chunk_roff(5)>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP,
                                                                      tab, Rl);
            break;
        default:
            chunk_R<4, T, LD, LAYOUT, NP, chunk_rpg(6), chunk_roff(6)>(rw, nrw, lane, px, py, pz, valid, q, nq, ch, pP,
                                                                      tab, Rl);
            break;
    }
}

// H rows of sub-step p (FP operand stream, rows multiple of 4) for block columns [fb, fb + NFB) into an LDS slot
// laid out [step][NFB/16 column tiles][64 lanes]; returns this copy wave's load count
template <typename T, int NFB, int SR>
__device__ __forceinline__ int stage_h_fp(const ChunkDev *__restrict__ chunks, SubStep p, const T *__restrict__ Hf,
                                          int nct, int fb, T *slot, int cw, int ncw, int lane) {
    constexpr int UPS = NFB * 4 * (int)sizeof(T) / 1024;  // KiB per K step
    const ChunkDev ch = chunk_at(chunks, p.c);
    const int64_t r0 = ch.prow0 + (int64_t)p.q * SR;
    const int nstep = min(SR, ch.kr - p.q * SR) / 4;
    int n = 0;
    for (int s = 0; s < nstep; ++s) {
        const char *src = (const char *)(Hf + ((r0 / 4 + s) * (int64_t)nct + fb / 16) * 64);
        n += lds_copy_async(src, (char *)(slot + (size_t)s * NFB * 4), UPS, cw, ncw, lane, s * UPS);
    }
    return n;
}

// s_waitcnt lgkmcnt(0) alone (vmcnt 63, expcnt 7)
__device__ __forceinline__ void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xF | (7 << 4) | (0 << 8) | (3 << 14)); }

// One (point tile, column block) of one system's Coulomb contraction, executed by a whole block.  All waves
// return from here together; the persistent batch kernel calls it once per work item.
// LDS of one FP Coulomb tile: the two R tiles and the ring of staged H sub-steps
template <typename T, int NFB>
struct FpTileLds {
    static constexpr int SR = sizeof(T) == 4 ? 32 : (NFB <= 192 ? 16 : 8);  // Hermite rows per sub-step
    static constexpr int LDR = 64 + (sizeof(T) == 8 ? 8 : 16);
    T Rl[2][kChunkRows * LDR];
    __attribute__((aligned(16))) T Hl[kSlots][SR * NFB];
};

template <typename T, int NFB>
__device__ __forceinline__ void coulomb_tile_fp(FpTileLds<T, NFB> &L, const ChunkDev *__restrict__ chunks, int nchunks,
                                                const double4 *__restrict__ pP, const T *__restrict__ Hf, int nct,
                                                const double *__restrict__ X, const double *__restrict__ Y,
                                                const double *__restrict__ Z, int64_t M, int nF,
                                                double *__restrict__ V, int64_t ldv, const double *__restrict__ tab,
                                                int64_t tile, int cb) {
    constexpr int TP = 64, NRT = 2, NCW = NFB / 64;  // consumer wave: 32 points x NFB/4 columns (TP = wave size)
    constexpr int SR = FpTileLds<T, NFB>::SR;        // Hermite rows per sub-step
    constexpr int NRW = 7, NCPW = 1;                 // R waves, copy waves
    constexpr int LDR = FpTileLds<T, NFB>::LDR;
    T(&Rl)[2][kChunkRows * LDR] = L.Rl;
    T(&Hl)[kSlots][SR * NFB] = L.Hl;
    using Tr = MfmaTraits<T>;
    using acc_t = typename Tr::acc;
    const int64_t k0 = tile * TP;
    const int fb = cb * NFB;
    // the wave index through readfirstlane: the compiler then knows it (and every role, group, unit and tile index
    // derived from it) to be wave-uniform and keeps that arithmetic and the load addresses built on it scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (wave < NRW) {  // R waves: lane l works on point k0 + l throughout
        const int64_t k = k0 + lane;
        const bool valid = k < M;
        const double px = valid ? gload(X + k) : 0.0, py = valid ? gload(Y + k) : 0.0, pz = valid ? gload(Z + k) : 0.0;
        const RType<T> *ctab = sizeof(T) == 8 ? (const RType<T> *)tab : (const RType<T> *)(tab + kBoysRows * kBoysCols);
        produce_R<T, LDR, 0, 1>(wave, NRW, lane, px, py, pz, valid, 0, 1, chunk_at(chunks, 0), pP, ctab, Rl[0]);
        __syncthreads();
        for (int c = 0; c < nchunks; ++c) {
            const int nsub = (chunk_at(chunks, c).kr + SR - 1) / SR;
            for (int q = 0; q < nsub; ++q) {
                if (c + 1 < nchunks)
                    produce_R<T, LDR, 0, 1>(wave, NRW, lane, px, py, pz, valid, q, nsub, chunk_at(chunks, c + 1), pP,
                                            ctab, Rl[(c + 1) & 1]);
                __syncthreads();
            }
        }
    } else if (wave < kConsumerWave0) {  // copy waves: keep the sub-step two ahead in flight
        const int cw = wave - NRW;
        SubStep nxt{0, 0};
        int gs_ahead = 0, pend = 0;
        pend = stage_h_fp<T, NFB, SR>(chunks, nxt, Hf, nct, fb, Hl[0], cw, NCPW, lane);
        advance(nxt, chunks, SR);
        ++gs_ahead;
        if (nxt.c < nchunks) {
            pend = stage_h_fp<T, NFB, SR>(chunks, nxt, Hf, nct, fb, Hl[1], cw, NCPW, lane);
            advance(nxt, chunks, SR);
            ++gs_ahead;
        } else {
            pend = 0;
        }
        wait_vm(pend);  // sub-step 0 has landed
        // raw barrier: a __syncthreads() fence would also wait for the copies still in flight
        __builtin_amdgcn_s_barrier();
        for (int c = 0; c < nchunks; ++c) {
            const int nsub = (chunk_at(chunks, c).kr + SR - 1) / SR;
            for (int q = 0; q < nsub; ++q) {
                int issued = 0;
                if (nxt.c < nchunks) {
                    issued = stage_h_fp<T, NFB, SR>(chunks, nxt, Hf, nct, fb, Hl[gs_ahead % kSlots], cw, NCPW, lane);
                    advance(nxt, chunks, SR);
                    ++gs_ahead;
                }
                wait_vm(issued);  // everything but the newest sub-step has landed
                __builtin_amdgcn_s_barrier();
            }
        }
    } else {
    const int cwv = wave - kConsumerWave0;
    const int lr = lane >> 4, lc = lane & 15;
    const int pw = (cwv >> 2) * NRT * 16;  // first point of this wave
    const int cl0 = (cwv & 3) * NCW;        // first column tile of this wave inside the block
    acc_t acc[NRT][NCW];
    double sum[NRT][NCW][4];
#pragma unroll
    for (int i = 0; i < NRT; ++i)
#pragma unroll
        for (int j = 0; j < NCW; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[i][j][r] = (T)0;
                sum[i][j][r] = 0.0;
            }
    __syncthreads();
    int gs = 0;
    for (int c = 0; c < nchunks; ++c) {
        const int kr = chunk_at(chunks, c).kr;
        const int nsub = (kr + SR - 1) / SR;
        const T *R = Rl[c & 1];
        for (int q = 0; q < nsub; ++q, ++gs) {
            const T *Hs = Hl[gs % kSlots];
            const int ns = min(SR, kr - q * SR) / 4;
            for (int s = 0; s < ns; ++s) {
                T a[NRT], b[NCW];
#pragma unroll
                for (int j = 0; j < NCW; ++j) b[j] = Hs[(s * (NFB / 16) + cl0 + j) * 64 + lane];
#pragma unroll
                for (int i = 0; i < NRT; ++i) a[i] = R[(q * SR + 4 * s + lr) * LDR + pw + i * 16 + lc];
#pragma unroll
                for (int i = 0; i < NRT; ++i)
#pragma unroll
                    for (int j = 0; j < NCW; ++j) acc[i][j] = Tr::mma(a[i], b[j], acc[i][j]);
            }
            if (q + 1 == nsub) {
                if constexpr (sizeof(T) == 4) {
#pragma unroll
                    for (int i = 0; i < NRT; ++i)
#pragma unroll
                        for (int j = 0; j < NCW; ++j)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                sum[i][j][r] += (double)acc[i][j][r];
                                acc[i][j][r] = 0.0f;
                            }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < NRT; ++i)
#pragma unroll
        for (int j = 0; j < NCW; ++j) {
            const int f = fb + (cl0 + j) * 16 + lc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t k = k0 + pw + i * 16 + Tr::row(lr, r);
                const double v = sizeof(T) == 4 ? sum[i][j][r] : (double)acc[i][j][r];
                if (f < nF && k < M) gstore(V + (int64_t)f * ldv + k, v);
            }
        }
    }
}

template <typename T, int NFB>
__global__ void __launch_bounds__(kCoulThreads) k_coulomb_ws(CoulJob<T> J, const double *__restrict__ tab) {
    __shared__ FpTileLds<T, NFB> lds;
    coulomb_tile_fp<T, NFB>(lds, J.chunks, J.nchunks, J.pP, (const T *)J.H, J.nct, J.X, J.Y, J.Z, J.M, J.nF, J.V, J.ldv, tab,
                            blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------------------------------
// Split-BF16 Coulomb.  x = x1 + x2 (+ x3) in bf16 pieces, so the product x*y is carried by x1y1 + x1y2 + x2y1
// (+ x1y3 + x2y2 + x3y1) at ~2^-16 (2^-24) relative error, every term on the 16x16x32 bf16 MFMA (lane l holds
// A[l%16][8*(l/16) .. +7], B[8*(l/16) .. +7][l%16]; FP32 register r holds D[4*(l/16) + r][l%16]).  Same block
// organisation as k_coulomb_ws with one 32-row K step per sub-step and consumer waves of 16 points x half the
// columns.  R is kept in LDS as FP32 [point][row]; each consumer splits its 8-row A operand when it loads it; H
// pieces come from k_split_h in operand order.  The copy waves stage H through their registers (ordinary 16-byte
// loads one sub-step ahead, 16-byte LDS stores into two slots): on the 36 KiB split-BF16 sub-steps this moves H
// faster than the asynchronous global->LDS loads of the FP kernels, which remain faster for their 24 KiB steps.
// ---------------------------------------------------------------------------------------------------
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int NP>
__device__ __forceinline__ void split_bf16(double x, __bf16 *out, int64_t stride) {
    double r = x;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        const __bf16 b = (__bf16)(float)r;
        out[p * stride] = b;
        r -= (double)(float)b;
    }
}

// Hs[p][g][ct][lane][e]: piece p of H row 32g + 8*(lane/16) + e, column 16ct + lane%16 (zero-extended rows and
// columns); one block per chunk of the bf16 chunk list
template <int NP>
__global__ void k_split_h(const ChunkDev *__restrict__ chunks, const double *__restrict__ H, int ldh, int nF,
                          int64_t nsteps, __bf16 *__restrict__ Hs) {
    const ChunkDev ch = chunks[blockIdx.x];
    const int rpg = chunk_rpg(ch.L), nt = ntuv(chunk_herm(ch.L)), roff = chunk_roff(ch.L);
    const int nreal = (ch.pe - ch.pb) * rpg;
    const int nct = ldh / 16;
    for (int e = threadIdx.x; e < ldh * ch.kr; e += blockDim.x) {
        const int f = e / ch.kr, r = e - f * ch.kr;
        // chunk row r: row roff + r % rpg of group r / rpg, whose Hermite block starts nt rows after the previous one
        const int64_t hr = ch.hrow0 + (int64_t)(r / rpg) * nt + roff + r % rpg;
        const double v = (r < nreal && f < nF) ? H[hr * ldh + f] : 0.0;
        const int64_t R = ch.prow0 + r;
        const int64_t idx = ((R / 32) * nct + f / 16) * 512 + (16 * ((R % 32) / 8) + f % 16) * 8 + R % 8;
        split_bf16<NP>(v, Hs + idx, nsteps * nct * 512);
    }
}

// Far-chunk bits, one block of 256 threads per 64-point tile.  The tile's centroid o and radius rho (largest distance
// of its points from o) bound every point's distance to a group centre P from below by |P - o| - rho, so
// p (|P - o| - rho)^2 bounds T = p|r - P|^2 of every (point, group) pair; a chunk is far when the bound of each of its
// groups exceeds tfar.  Each wave assembles 64 chunk bits with one ballot.
__global__ void __launch_bounds__(256) k_far_bits(const double *__restrict__ X, const double *__restrict__ Y,
                                                  const double *__restrict__ Z, int64_t M,
                                                  const ChunkDev *__restrict__ chunks, int nchunks,
                                                  const double4 *__restrict__ pP, double tfar, int farw,
                                                  unsigned *__restrict__ far, double *__restrict__ farrows) {
    __shared__ double geo[4];
    const int64_t t = blockIdx.x, k0 = t * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (wave == 0) {
        const int64_t k = k0 + lane;
        const bool valid = k < M;
        const double x = valid ? X[k] : 0.0, y = valid ? Y[k] : 0.0, z = valid ? Z[k] : 0.0;
        double sx = x, sy = y, sz = z, n = valid ? 1.0 : 0.0;
        for (int o = 32; o > 0; o >>= 1) {
            sx += __shfl_xor(sx, o);
            sy += __shfl_xor(sy, o);
            sz += __shfl_xor(sz, o);
            n += __shfl_xor(n, o);
        }
        const double cx = sx / n, cy = sy / n, cz = sz / n;
        double d = valid ? sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz)) : 0.0;
        for (int o = 32; o > 0; o >>= 1) d = fmax(d, __shfl_xor(d, o));
        if (lane == 0) {
            geo[0] = cx;
            geo[1] = cy;
            geo[2] = cz;
            geo[3] = d;
        }
    }
    __syncthreads();
    const double cx = geo[0], cy = geo[1], cz = geo[2], rho = geo[3];
    double rows = 0.0;
    for (int base = wave * 64; base < farw * 32; base += 256) {
        const int c = base + lane;
        bool is_far = false;
        if (c < nchunks) {
            const ChunkDev ch = chunks[c];
            double tmin = INFINITY;
            for (int g = ch.pb; g < ch.pe; ++g) {
                const double4 q = pP[g];
                const double dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
                const double d = fmax(sqrt(dx * dx + dy * dy + dz * dz) - rho, 0.0);
                tmin = fmin(tmin, q.w * d * d);
            }
            is_far = tmin > tfar;
            if (is_far) rows += ch.kr;
        }
        const unsigned long long m = __ballot(is_far);
        const int w = base / 32;
        if (lane == 0) far[t * farw + w] = (unsigned)m;
        if (lane == 1 && w + 1 < farw) far[t * farw + w + 1] = (unsigned)(m >> 32);
    }
    for (int o = 32; o > 0; o >>= 1) rows += __shfl_xor(rows, o);
    if (lane == 0 && rows > 0.0) atomicAdd(farrows, rows);
}

void launch_far_bits(const double *x, const double *y, const double *z, int64_t M, const ChunkDev *chunks,
                     int nchunks, const double4 *pairP, double tfar, int farw, unsigned *far, double *farrows,
                     hipStream_t s) {
    launch_fill_zero(farrows, 1, s);
    if (M <= 0 || farw <= 0) return;
    hipLaunchKernelGGL(k_far_bits, dim3((unsigned)((M + 63) / 64)), dim3(256), 0, s, x, y, z, M, chunks, nchunks,
                       pairP, tfar, farw, far, farrows);
}

// A operand of one K step: the 8 FP32 rows of this lane's point split into NP bf16 pieces, each the rounded
// remainder of the previous ones, so that the pieces shrink by 2^-8 each and the last remainder is below 2^-8 of
// the last piece.  (An exact split by truncation, keeping the upper 16 bits of each remainder, is 1% faster here
// but its pieces shrink only by 2^-7, which triples the worst dropped cross product.)
// Rows go two at a time through one packed conversion (v_cvt_pk_bf16_f32, round to nearest even like the scalar
// cast); a piece's bf16 value is back in FP32 by a shift (low half) or a mask (high half) of the packed word, so a
// piece of two rows costs 5 VALU operations and the split needs no repacking into the operand vector.
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef unsigned v4w __attribute__((ext_vector_type(4)));
template <int NP>
__device__ __forceinline__ void split_a(const float *src, v8bf (&a)[NP]) {
    const v4f x0 = *(const v4f *)src, x1 = *(const v4f *)(src + 4);
    float r[8] = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        v4w w;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const unsigned u = __builtin_bit_cast(unsigned, __builtin_convertvector((v2f){r[2 * e], r[2 * e + 1]}, v2bf));
            w[e] = u;
            if (p + 1 < NP) {
                r[2 * e] -= __builtin_bit_cast(float, u << 16);
                r[2 * e + 1] -= __builtin_bit_cast(float, u & 0xffff0000u);
            }
        }
        a[p] = __builtin_bit_cast(v8bf, w);
    }
}

// The 1 KiB units (64 lanes x 16 bytes) of one staged K step that copy wave cw of NCPW moves, through registers:
// unit u = pc * NFB/16 + v is KiB v of piece pc's run for the block's columns, at LDS slot offset u KiB.
template <int NP, int NFB, int NCPW>
struct HUnits {
    static constexpr int UPP = NFB / 16, NU = NP * UPP, UPW = (NU + NCPW - 1) / NCPW;
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    v4u r[UPW];
    // the units of the sub-step whose piece-0 run for the block's columns starts at sub_base, the pieces pstride bytes
    // apart; npe: pieces read from H (NP, or 2 for a far chunk, whose third piece is staged as zeros)
    __device__ __forceinline__ void load_at(const char *sub_base, int64_t pstride, int cw, int lane, int npe) {
#pragma unroll
        for (int k = 0; k < UPW; ++k) {
            const int u = cw + k * NCPW;
            if (u < npe * UPP) {
                const int pc = u / UPP, v = u - pc * UPP;
                r[k] = gload((const v4u *)(sub_base + pc * pstride + v * 1024 + lane * 16));
            } else if (u < NU) {
                r[k] = v4u{0u, 0u, 0u, 0u};
            }
        }
    }
    __device__ __forceinline__ void store(__bf16 *slot, int cw, int lane) const {
#pragma unroll
        for (int k = 0; k < UPW; ++k) {
            const int u = cw + k * NCPW;
            if (u < NU) *(v4u *)((char *)slot + u * 1024 + lane * 16) = r[k];
        }
    }
};

// LDS of one split-BF16 Coulomb tile: the two FP32 R tiles, the ring of staged H sub-steps, the FP32 Boys table
template <int NP, int NFB>
struct Bf16TileLds {
    static constexpr int LDK = kChunkRowsBf16 + 4;  // FP32 point rows of 272 bytes: conflict-free 16-byte reads
    static constexpr int SLOT = NP * NFB * 32;      // bf16 elements of one staged K step
    __attribute__((aligned(16))) float Rf[2][64 * LDK];
    __attribute__((aligned(16))) __bf16 Hl[kSlotsReg][SLOT];
    __attribute__((aligned(16))) float Tb[kBoysRows * kBoysCols];
    unsigned Fl[NP == 3 ? kFarWords : 1];  // far-chunk bits of this tile (three-piece form)
};

// One 32-row K step of a consumer wave (16 points x NFB/2 columns): the A operand split into NPE pieces and the
// products of the pieces down to order 2^-8(NPE-1): 3 (a0b0, a0b1, a1b0) for NPE = 2, plus a0b2, a1b1, a2b0 for
// NPE = 3 (zero for a0b2 in a far chunk, whose staged third H piece is zero)
template <int NPE, int NFB>
__device__ __forceinline__ void bf16_substep(const float *Rrow, const __bf16 *Hs_, int cl0, int lane,
                                             v4f (&acc)[NFB / 32]) {
    v8bf a[NPE];
    split_a<NPE>(Rrow, a);
#pragma unroll
    for (int j = 0; j < NFB / 32; ++j) {
        v8bf b[NPE];
#pragma unroll
        for (int p = 0; p < NPE; ++p) b[p] = *(const v8bf *)(Hs_ + ((p * (NFB / 16) + cl0 + j) * 64 + lane) * 8);
        v4f cc = acc[j];
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], cc, 0, 0, 0);
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], cc, 0, 0, 0);
        cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], cc, 0, 0, 0);
        if constexpr (NPE == 3) {
            cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], cc, 0, 0, 0);
            cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], cc, 0, 0, 0);
            cc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], cc, 0, 0, 0);
        }
        acc[j] = cc;
    }
}

// far: null, or the far-chunk bits of every tile (farw 32-bit words per tile, bit c of a tile's row set where every
// (point, group) pair of tile and chunk c has T = p|r - P|^2 above the far threshold): the three-piece form reads
// two H pieces of those chunks from memory and stages the third as zeros (H to ~2^-17 there, R still in three)
template <int NP, int NFB>
__device__ __forceinline__ void coulomb_tile_bf16(Bf16TileLds<NP, NFB> &L, const ChunkDev *__restrict__ chunks,
                                                  int nchunks, int64_t nsteps,
                                                  const double4 *__restrict__ pP, const __bf16 *__restrict__ Hs,
                                                  int nct, const double *__restrict__ X, const double *__restrict__ Y,
                                                  const double *__restrict__ Z, int64_t M, int nF,
                                                  double *__restrict__ V, int64_t ldv,
                                                  const double *__restrict__ tab, int64_t tile, int cb,
                                                  const unsigned *__restrict__ far, int farw) {
    constexpr int TP = 64, NCW = NFB / 32, SR = 32;  // consumer wave: 16 points x NFB/2 columns
    constexpr int NCPW = NP == 3 ? 2 : 1, NRW = kConsumerWave0 - NCPW;
    constexpr int LDK = Bf16TileLds<NP, NFB>::LDK;
    constexpr int SLOT = Bf16TileLds<NP, NFB>::SLOT;
    float(&Rf)[2][TP * LDK] = L.Rf;
    __bf16(&Hl)[kSlotsReg][SLOT] = L.Hl;
    float(&Tb)[kBoysRows * kBoysCols] = L.Tb;  // FP32 Boys table for the R waves
    const int64_t k0 = tile * TP;
    const int fb = cb * NFB;
    // the wave index through readfirstlane: the compiler then knows it (and every role, group, unit and tile index
    // derived from it) to be wave-uniform and keeps that arithmetic and the load addresses built on it scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const bool use_far = NP == 3 && far != nullptr;
    {  // the table in LDS keeps the R waves' per-lane table lookups off the load path of the H copies
        const v4f *src = (const v4f *)(tab + kBoysRows * kBoysCols);
        for (int i = threadIdx.x; i < kBoysRows * kBoysCols / 4; i += kCoulThreads) ((v4f *)Tb)[i] = src[i];
        if (use_far)
            for (int i = threadIdx.x; i < farw; i += kCoulThreads) L.Fl[i] = gload(far + tile * farw + i);
        __syncthreads();
    }
    // H pieces read from memory for chunk c (uniform)
    auto pieces = [&](int c) -> int {
        if (!use_far) return NP;
        const unsigned w = __builtin_amdgcn_readfirstlane(L.Fl[c >> 5]);
        return (w >> (c & 31)) & 1u ? 2 : NP;
    };
    if (wave < NRW) {  // R waves: lane l works on point k0 + l throughout
        const int64_t k = k0 + lane;
        const bool valid = k < M;
        const double px = valid ? gload(X + k) : 0.0, py = valid ? gload(Y + k) : 0.0, pz = valid ? gload(Z + k) : 0.0;
        produce_R<float, LDK, 2, 1>(wave, NRW, lane, px, py, pz, valid, 0, 1, chunk_at(chunks, 0), pP, Tb, Rf[0]);
        __syncthreads();
        for (int c = 0; c < nchunks; ++c) {
            const int nsub = chunk_at(chunks, c).kr / SR;
            for (int q = 0; q < nsub; ++q) {
                if (c + 1 < nchunks)
                    produce_R<float, LDK, 2, 1>(wave, NRW, lane, px, py, pz, valid, q, nsub, chunk_at(chunks, c + 1),
                                                pP, Tb, Rf[(c + 1) & 1]);
                __syncthreads();
            }
        }
    } else if (wave < kConsumerWave0) {  // copy waves: H of sub-step gs + 2 is loaded in iteration gs, stored in gs + 1
        const int cw = wave - NRW;
        HUnits<NP, NFB, NCPW> h;
        SubStep nxt{0, 0};
        // H of sub-step g (chunk c, step q: g = prow0(c) / 32 + q, the sub-steps consecutive in the padded row space)
        // for the block's columns starts g * nct KiB after hb; the pieces' runs are pstride bytes apart
        const int64_t pstride = nsteps * nct * 1024;
        const char *hb = (const char *)(Hs + (fb / 16) * 512);
        int64_t gnext = 0;  // sub-step index of nxt
        h.load_at(hb, pstride, cw, lane, pieces(0));
        h.store(Hl[0], cw, lane);
        advance(nxt, chunks, SR);
        ++gnext;
        bool held = nxt.c < nchunks;  // h holds the next sub-step
        if (held) {
            h.load_at(hb + gnext * nct * 1024, pstride, cw, lane, pieces(nxt.c));
            advance(nxt, chunks, SR);
            ++gnext;
        }
        wait_lgkm0();
        __builtin_amdgcn_s_barrier();
        int gs = 0;
        for (int c = 0; c < nchunks; ++c) {
            const int nsub = chunk_at(chunks, c).kr / SR;
            for (int q = 0; q < nsub; ++q, ++gs) {
                if (held) h.store(Hl[(gs + 1) % kSlotsReg], cw, lane);
                held = nxt.c < nchunks;
                if (held) {
                    h.load_at(hb + gnext * nct * 1024, pstride, cw, lane, pieces(nxt.c));
                    advance(nxt, chunks, SR);
                    ++gnext;
                }
                // the stores must have landed; the loads just issued stay in flight across the barrier
                wait_lgkm0();
                __builtin_amdgcn_s_barrier();
            }
        }
    } else {
    const int cwv = wave - kConsumerWave0;
    const int lr = lane >> 4, lc = lane & 15;
    const int pw = (cwv >> 1) * 16;   // point tile of this wave
    const int cl0 = (cwv & 1) * NCW;  // first column tile of this wave inside the block
    v4f acc[NCW];
    double sum[NCW][4];
#pragma unroll
    for (int j = 0; j < NCW; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            acc[j][r] = 0.0f;
            sum[j][r] = 0.0;
        }
    __syncthreads();
    int gs = 0, since = 0;  // since: rows accumulated in FP32 since the last flush into the FP64 sums
    for (int c = 0; c < nchunks; ++c) {
        const int kr = chunk_at(chunks, c).kr, nsub = kr / SR;
        const float *R = Rf[c & 1];
        since += kr;
        const bool flush = since >= kFlushRowsBf16 || c + 1 == nchunks;
        if (flush) since = 0;
        for (int q = 0; q < nsub; ++q, ++gs) {
            bf16_substep<NP, NFB>(R + (pw + lc) * LDK + SR * q + 8 * lr, Hl[gs % kSlotsReg], cl0, lane, acc);
            if (flush && q + 1 == nsub) {
#pragma unroll
                for (int j = 0; j < NCW; ++j)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        sum[j][r] += (double)acc[j][r];
                        acc[j][r] = 0.0f;
                    }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int j = 0; j < NCW; ++j) {
        const int f = fb + (cl0 + j) * 16 + lc;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t k = k0 + pw + 4 * lr + r;
            if (f < nF && k < M) gstore(V + (int64_t)f * ldv + k, sum[j][r]);
        }
    }
    }
}

template <int NP, int NFB>
__global__ void __launch_bounds__(kCoulThreads) k_coulomb_ws_bf16(CoulJob<__bf16> J, const double *__restrict__ tab) {
    __shared__ Bf16TileLds<NP, NFB> lds;
    coulomb_tile_bf16<NP, NFB>(lds, J.chunks, J.nchunks, J.nsteps, J.pP, (const __bf16 *)J.H, J.nct, J.X, J.Y, J.Z, J.M,
                               J.nF, J.V, J.ldv, tab, blockIdx.x, blockIdx.y, J.far, J.farw);
}

// ---------------------------------------------------------------------------------------------------
// Batched Coulomb: one launch over the (system, point tile, column block) items of every system of one form.
// Block b runs item b with the single-system tile code; the host orders the systems (heaviest items first) so
// that the dispatcher, which hands blocks to CUs in index order as CUs free up, balances them like a work queue.
// One block fills a CU (16 waves at 128 VGPRs), so the grid is the queue and no block loops.
// ---------------------------------------------------------------------------------------------------
// Every job of a batch form in one launch.  The job records sit in device memory (a batch of 48 systems does not fit
// the kernel argument space); each block copies its job through readfirstlane word by word, so every field is a
// scalar (uniform) value in the tile code, as when it comes from the kernel arguments, and the tile's loop bounds
// stay scalar branches.  Jobs of different column-block widths share the launch: the block switches on its job's
// width, and the LDS of the widest tile is shared by all widths.
template <typename T>
__device__ __forceinline__ CoulJob<T> uniform_job(const CoulJob<T> *__restrict__ jobs, int j) {
    static_assert(sizeof(CoulJob<T>) % 4 == 0, "job record in 32-bit words");
    constexpr int NW = sizeof(CoulJob<T>) / 4;
    const unsigned *src = reinterpret_cast<const unsigned *>(jobs + j);
    unsigned w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = __builtin_amdgcn_readfirstlane(cload(src + i));
    CoulJob<T> J;
    __builtin_memcpy(&J, w, sizeof(J));
    return J;
}
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(uint64_t)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
// block -> (job, local item): off[] ascending, off[0] = 0
__device__ __forceinline__ int job_of_block(const int64_t *__restrict__ off, int njobs, int64_t it, int64_t &local) {
    int lo = 0, hi = njobs - 1;  // largest j with off[j] <= it
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (uniform_i64(cload(off + mid)) <= it) lo = mid;
        else hi = mid - 1;
    }
    const int j = __builtin_amdgcn_readfirstlane(lo);
    local = it - uniform_i64(cload(off + j));
    return j;
}

template <typename T>
constexpr size_t fp_lds_bytes() {
    size_t b = sizeof(FpTileLds<T, 64>);
    b = b > sizeof(FpTileLds<T, 128>) ? b : sizeof(FpTileLds<T, 128>);
    b = b > sizeof(FpTileLds<T, 192>) ? b : sizeof(FpTileLds<T, 192>);
    b = b > sizeof(FpTileLds<T, 256>) ? b : sizeof(FpTileLds<T, 256>);
    return b;
}
template <typename T>
__global__ void __launch_bounds__(kCoulThreads) k_coulomb_ws_batch_all(const CoulJob<T> *__restrict__ jobs,
                                                                       const int64_t *__restrict__ off, int njobs,
                                                                       const double *__restrict__ tab) {
    __shared__ __attribute__((aligned(16))) unsigned char raw[fp_lds_bytes<T>()];
    int64_t local = 0;
    const int j = job_of_block(off, njobs, (int64_t)blockIdx.x, local);
    const CoulJob<T> J = uniform_job(jobs, j);
    const int64_t tile = local / J.ncb;
    const int cb = (int)(local - tile * J.ncb);
#define FLOSIC_TILE_FP(NFB)                                                                                          \
    coulomb_tile_fp<T, NFB>(*reinterpret_cast<FpTileLds<T, NFB> *>(raw), J.chunks, J.nchunks, J.pP, (const T *)J.H,   \
                            J.nct, J.X, J.Y, J.Z, J.M, J.nF, J.V, J.ldv, tab, tile, cb)
    switch (J.nfb) {
        case 64: FLOSIC_TILE_FP(64); break;
        case 128: FLOSIC_TILE_FP(128); break;
        case 192: FLOSIC_TILE_FP(192); break;
        default: FLOSIC_TILE_FP(256); break;
    }
#undef FLOSIC_TILE_FP
}

template <int NP>
constexpr size_t bf16_lds_bytes() {
    size_t b = sizeof(Bf16TileLds<NP, 64>);
    b = b > sizeof(Bf16TileLds<NP, 128>) ? b : sizeof(Bf16TileLds<NP, 128>);
    b = b > sizeof(Bf16TileLds<NP, 192>) ? b : sizeof(Bf16TileLds<NP, 192>);
    return b;
}
template <int NP>
__global__ void __launch_bounds__(kCoulThreads) k_coulomb_ws_bf16_batch_all(const CoulJob<__bf16> *__restrict__ jobs,
                                                                            const int64_t *__restrict__ off,
                                                                            int njobs, const double *__restrict__ tab) {
    __shared__ __attribute__((aligned(16))) unsigned char raw[bf16_lds_bytes<NP>()];
    int64_t local = 0;
    const int j = job_of_block(off, njobs, (int64_t)blockIdx.x, local);
    const CoulJob<__bf16> J = uniform_job(jobs, j);
    const int64_t tile = local / J.ncb;
    const int cb = (int)(local - tile * J.ncb);
#define FLOSIC_TILE_BF(NFB)                                                                                         \
    coulomb_tile_bf16<NP, NFB>(*reinterpret_cast<Bf16TileLds<NP, NFB> *>(raw), J.chunks, J.nchunks, J.nsteps, J.pP, \
                               (const __bf16 *)J.H, J.nct, J.X, J.Y, J.Z, J.M, J.nF, J.V, J.ldv, tab, tile, cb,  \
                               J.far, J.farw)
    switch (J.nfb) {
        case 64: FLOSIC_TILE_BF(64); break;
        case 128: FLOSIC_TILE_BF(128); break;
        default: FLOSIC_TILE_BF(192); break;
    }
#undef FLOSIC_TILE_BF
}

// FLO columns per Coulomb block: all of them up to 256 (ldh is nF rounded up to a multiple of this); the
// split-BF16 kernel takes at most 192 (its register budget), 128 beyond that
int coulomb_cols(int nF) { return nF <= 64 ? 64 : nF <= 128 ? 128 : nF <= 192 ? 192 : 256; }
bool coulomb_cols_legal(int nfb, int ldh, bool split) {
    if (nfb != 64 && nfb != 128 && nfb != 192 && nfb != 256) return false;
    if (split && nfb > 192) return false;
    return ldh % nfb == 0;
}
int coulomb_cols_split(int nF) { return nF <= 64 ? 64 : nF <= 128 ? 128 : nF <= 192 ? 192 : 128; }

template <typename T>
void launch_frag_h(const ChunkDev *chunks, int nchunks, const double *H, int ldh, int nF, T *Hf, hipStream_t s) {
    if (nchunks <= 0) return;
    hipLaunchKernelGGL((k_frag_h<T>), dim3(nchunks), dim3(256), 0, s, chunks, H, ldh, nF, Hf);
}
template void launch_frag_h<double>(const ChunkDev *, int, const double *, int, int, double *, hipStream_t);
template void launch_frag_h<float>(const ChunkDev *, int, const double *, int, int, float *, hipStream_t);

template <typename T>
void launch_coulomb_mfma(const ChunkDev *chunks, int nchunks, int64_t nsteps, const double4 *pairP, const T *Hf,
                         int ldh, const double *x, const double *y, const double *z, int64_t M, const double *btab,
                         int nF, double *V, int64_t ldv, hipStream_t s, int nfb) {
    if (nchunks <= 0 || M <= 0) return;
    if (nfb <= 0) nfb = coulomb_cols(nF);
    const CoulJob<T> J{chunks, nchunks, pairP, Hf, ldh / 16, x, y, z, M, nF, V, ldv, nsteps, ldh / nfb};
    const dim3 grid((unsigned)((M + 63) / 64), (unsigned)(ldh / nfb)), blk(kCoulThreads);
    switch (nfb) {
        case 64: hipLaunchKernelGGL((k_coulomb_ws<T, 64>), grid, blk, 0, s, J, btab); break;
        case 128: hipLaunchKernelGGL((k_coulomb_ws<T, 128>), grid, blk, 0, s, J, btab); break;
        case 192: hipLaunchKernelGGL((k_coulomb_ws<T, 192>), grid, blk, 0, s, J, btab); break;
        default: hipLaunchKernelGGL((k_coulomb_ws<T, 256>), grid, blk, 0, s, J, btab); break;
    }
}
template void launch_coulomb_mfma<double>(const ChunkDev *, int, int64_t, const double4 *, const double *, int,
                                          const double *, const double *, const double *, int64_t, const double *, int,
                                          double *, int64_t, hipStream_t, int);
template void launch_coulomb_mfma<float>(const ChunkDev *, int, int64_t, const double4 *, const float *, int,
                                         const double *, const double *, const double *, int64_t, const double *, int,
                                         double *, int64_t, hipStream_t, int);

void launch_split_h(int npieces, const ChunkDev *chunks, int nchunks, const double *H, int ldh, int nF,
                    int64_t nsteps, __bf16 *Hs, hipStream_t s) {
    if (nchunks <= 0) return;
    if (npieces == 2)
        hipLaunchKernelGGL((k_split_h<2>), dim3(nchunks), dim3(256), 0, s, chunks, H, ldh, nF, nsteps, Hs);
    else
        hipLaunchKernelGGL((k_split_h<3>), dim3(nchunks), dim3(256), 0, s, chunks, H, ldh, nF, nsteps, Hs);
}

void launch_coulomb_bf16(int npieces, const ChunkDev *chunks, int nchunks, int64_t nsteps, const double4 *pairP,
                         const __bf16 *Hs, int ldh, const double *x, const double *y, const double *z, int64_t M,
                         const double *btab, int nF, double *V, int64_t ldv, hipStream_t s, int nfb,
                         const unsigned *far, int farw) {
    if (nchunks <= 0 || M <= 0) return;
    if (nfb <= 0) nfb = coulomb_cols_split(nF);
    const CoulJob<__bf16> J{chunks, nchunks, pairP, Hs, ldh / 16, x, y, z, M, nF, V, ldv, nsteps, ldh / nfb, nfb,
                            npieces == 3 ? far : nullptr, npieces == 3 ? farw : 0};
    const dim3 grid((unsigned)((M + 63) / 64), (unsigned)(ldh / nfb)), blk(kCoulThreads);
#define FLOSIC_WSB(NP, NFB) hipLaunchKernelGGL((k_coulomb_ws_bf16<NP, NFB>), grid, blk, 0, s, J, btab)
#define FLOSIC_WSB_NFB(NP)                        \
    switch (nfb) {                                \
        case 64: FLOSIC_WSB(NP, 64); break;       \
        case 128: FLOSIC_WSB(NP, 128); break;     \
        default: FLOSIC_WSB(NP, 192); break;      \
    }
    if (npieces == 2) {
        FLOSIC_WSB_NFB(2)
    } else {
        FLOSIC_WSB_NFB(3)
    }
#undef FLOSIC_WSB_NFB
#undef FLOSIC_WSB
}


template <typename T>
void launch_coulomb_batch_all(const CoulJob<T> *jobs, const int64_t *off, int njobs, int64_t items,
                              const double *btab, hipStream_t s) {
    if (njobs <= 0 || items <= 0) return;
    hipLaunchKernelGGL((k_coulomb_ws_batch_all<T>), dim3((unsigned)items), dim3(kCoulThreads), 0, s, jobs, off, njobs,
                       btab);
}
template void launch_coulomb_batch_all<double>(const CoulJob<double> *, const int64_t *, int, int64_t,
                                               const double *, hipStream_t);
template void launch_coulomb_batch_all<float>(const CoulJob<float> *, const int64_t *, int, int64_t, const double *,
                                              hipStream_t);
void launch_coulomb_bf16_batch_all(int npieces, const CoulJob<__bf16> *jobs, const int64_t *off, int njobs,
                                   int64_t items, const double *btab, hipStream_t s) {
    if (njobs <= 0 || items <= 0) return;
    if (npieces == 2)
        hipLaunchKernelGGL((k_coulomb_ws_bf16_batch_all<2>), dim3((unsigned)items), dim3(kCoulThreads), 0, s, jobs, off,
                           njobs, btab);
    else
        hipLaunchKernelGGL((k_coulomb_ws_bf16_batch_all<3>), dim3((unsigned)items), dim3(kCoulThreads), 0, s, jobs, off,
                           njobs, btab);
}


// ---------------------------------------------------------------------------------------------------
// exchange-correlation for rho_up = rho_i, rho_dn = 0, NSPN = 2 (subvlxc_libxc.F90:197-470)
// ---------------------------------------------------------------------------------------------------
__device__ __forceinline__ void gcor(double A, double A1, double B1, double B2, double B3, double B4, double RS,
                                     double &GG, double &GGRS) {
    const double Q0 = -2 * A * (1.0 + A1 * RS);
    const double RS12 = sqrt(RS);
    const double RS32 = RS12 * RS12 * RS12;
    const double Q1 = 2 * A * (B1 * RS12 + B2 * RS + B3 * RS32 + B4 * RS * RS);
    const double Q2 = log(1.0 + 1.0 / Q1);
    GG = Q0 * Q2;
    const double Q3 = A * (B1 / RS12 + 2 * B2 + 3 * B3 * RS12 + 4 * B4 * RS);
    GGRS = -2 * A * A1 * Q2 - Q0 * Q3 / (Q1 * Q1 + Q1);
}
__device__ __forceinline__ void gcor2(double A, double A1, double B1, double B2, double B3, double B4, double RTRS,
                                      double &GG, double &GGRS) {
    const double Q0 = -2.0 * A * (1.0 + A1 * RTRS * RTRS);
    const double Q1 = 2.0 * A * RTRS * (B1 + RTRS * (B2 + RTRS * (B3 + B4 * RTRS)));
    const double Q2 = log(1.0 + 1.0 / Q1);
    GG = Q0 * Q2;
    const double Q3 = A * (B1 / RTRS + 2.0 * B2 + RTRS * (3.0 * B3 + 4.0 * B4 * RTRS));
    GGRS = -2.0 * A * A1 * Q2 - Q0 * Q3 / (Q1 * (1.0 + Q1));
}

// PW91 exchange (pw91ex.f:7-50)
__device__ __forceinline__ void pw91ex(double DKF, double S, double U, double V, double &EXL, double &EXN,
                                       double &VX) {
    const double A1 = 0.19645, A2 = 0.27430, A3 = 0.15084, A4 = 100.0, AX = -0.238732414, A = 7.7956, B1 = 0.004;
    const double THRD4 = 1.3333333333333333;
    const double FAC = AX * DKF;
    const double S2 = S * S, S3 = S2 * S, S4 = S2 * S2;
    const double AS = A * S;
    const double PA = sqrt(1.0 + AS * AS);
    const double P0 = AS / PA;
    const double P1 = A1 * log(AS + PA);
    const double P2 = A3 * exp(-A4 * S2);
    const double P3 = 1.0 / (1.0 + S * P1 + B1 * S4);
    const double P4 = 1.0 + S * P1 + (A2 - P2) * S2;
    const double F = P3 * P4;
    EXL = FAC;
    EXN = FAC * (F - 1.0);
    const double P5 = B1 * S2 - (A2 - P2);
    const double P6 = S * (P1 + A1 * P0);
    const double P7 = 2 * ((A2 - P2) + A4 * S2 * P2 - 2 * B1 * S2 * F);
    const double FS = P3 * (P3 * P5 * P6 + P7);
    const double P8 = 2 * S * (B1 - A4 * P2);
    const double P9 = P1 + A1 * P0 * (3.0 - P0 * P0);
    const double P10 = 4 * (A4 * S * P2 * (2.0 - A4 * S2) - 2 * B1 * S * F - B1 * S3 * FS);
    const double P11 = P3 * P3 * (P1 + A1 * P0 + 4 * B1 * S3);
    const double FSS = P3 * P3 * (P5 * P9 + P6 * P8) - 2 * P3 * P5 * P6 * P11 + P3 * P10 - P7 * P11;
    VX = FAC * (THRD4 * F - (U - THRD4 * S3) * FSS - V * FS);
}

// PW92 local correlation at zeta = 1 (pw91lc.f:4-43)
__device__ __forceinline__ void pw91lc_pol(double RS, double &EC, double &VCUP) {
    const double GAMR = 1.92366105, FZZR = 0.58482234, THRD = 0.333333333333, THRD4 = 1.333333333333;
    const double ZETPR = pow(2.0, THRD);
    const double F = (ZETPR * ZETPR * ZETPR * ZETPR - 2.0) * GAMR;
    double EU, EURS, EP, EPRS, ALFM, ALFRSM;
    gcor(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, RS, EU, EURS);
    gcor(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, RS, EP, EPRS);
    gcor(0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671, RS, ALFM, ALFRSM);
    // Z3 = Z4 = 1
    EC = EU * (1.0 - F) + EP * F - ALFM * F * (1.0 - 1.0) * FZZR;
    const double ECRS = EURS * (1.0 - F) + EPRS * F - ALFRSM * F * (1.0 - 1.0) * FZZR;
    const double FZ = THRD4 * ZETPR * GAMR;
    const double ECZET = 4.0 * F * (EP - EU + ALFM * FZZR) + FZ * (EP - EU - (1.0 - 1.0) * ALFM * FZZR);
    const double COMM = EC - RS * ECRS * THRD - ECZET;
    VCUP = COMM + ECZET;
}

// PW92 local correlation at general zeta (pw91lc.f:4-43): EC and both spin potentials
__device__ __forceinline__ void pw91lc_gen(double RS, double ZET, double &EC, double &VCUP, double &VCDN) {
    const double GAMR = 1.92366105, FZZR = 0.58482234, THRD = 0.333333333333, THRD4 = 1.333333333333;
    const double ZETPR = pow(1.0 + ZET, THRD), ZETMR = pow(1.0 - ZET, THRD);
    const double F = (ZETPR * ZETPR * ZETPR * ZETPR + ZETMR * ZETMR * ZETMR * ZETMR - 2.0) * GAMR;
    double EU, EURS, EP, EPRS, ALFM, ALFRSM;
    gcor(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, RS, EU, EURS);
    gcor(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, RS, EP, EPRS);
    gcor(0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671, RS, ALFM, ALFRSM);
    const double Z3 = ZET * ZET * ZET, Z4 = Z3 * ZET;
    EC = EU * (1.0 - F * Z4) + EP * F * Z4 - ALFM * F * (1.0 - Z4) * FZZR;
    const double ECRS = EURS * (1.0 - F * Z4) + EPRS * F * Z4 - ALFRSM * F * (1.0 - Z4) * FZZR;
    const double FZ = THRD4 * (ZETPR - ZETMR) * GAMR;
    const double ECZET = 4.0 * Z3 * F * (EP - EU + ALFM * FZZR) + FZ * (Z4 * EP - Z4 * EU - (1.0 - Z4) * ALFM * FZZR);
    const double COMM = EC - RS * ECRS * THRD - ZET * ECZET;
    VCUP = COMM + ECZET;
    VCDN = COMM - ECZET;
}

// PBE exchange (pbeex.f, MODE=1)
__device__ __forceinline__ void pbeex(double RHO, double S, double U, double V, double &EXL, double &EXN,
                                      double &VX) {
    const double THRD = 1.0 / 3.0, THRD4 = 4.0 / 3.0;
    const double AX = -0.738558766382022405884230032680836, UM = 0.2195149727645171, UK = 0.8040, UL = UM / UK;
    const double EXUNIF = AX * pow(RHO, THRD);
    const double S2 = S * S;
    const double P0 = 1.0 + UL * S2;
    const double FXPBE = 1.0 + UK - UK / P0;
    EXL = EXUNIF;
    EXN = EXUNIF * (FXPBE - 1.0);
    const double FS = 2.0 * UK * UL / (P0 * P0);
    const double FSS = -4.0 * UL * S * FS / P0;
    VX = EXUNIF * (THRD4 * FXPBE - (U - THRD4 * S2 * S) * FSS - V * FS);
}

// PBE correlation at zeta = 1, WW = 0 (pbecor.f)
__device__ __forceinline__ void pbecor_pol(double RS, double T, double UU, double VV, double &EC, double &VCUP,
                                           double &H, double &DVCUP) {
    const double THRD = 1.0 / 3.0, THRD2 = 2.0 * THRD, THRD4 = 4.0 * THRD, SIXTHM = -THRD / 2.0;
    const double GAM = 0.5198420997897463295344212145565, FZZ = 8.0 / (9.0 * GAM);
    const double GAMMA = 0.03109069086965489503494086371273, BET = 0.06672455060314922, DELT = BET / GAMMA;
    const double ETA = 1.0e-12;
    const double ZET = 1.0;
    const double RTRS = sqrt(RS);
    double EU, EURS, EP, EPRS, ALFM, ALFRSM;
    gcor2(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, RTRS, EU, EURS);
    gcor2(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, RTRS, EP, EPRS);
    gcor2(0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671, RTRS, ALFM, ALFRSM);
    const double Z4 = 1.0;
    const double F = (pow(1.0 + ZET, THRD4) + 0.0 - 2.0) / GAM;
    const double ec = EU * (1.0 - F * Z4) + EP * F * Z4 - ALFM * F * (1.0 - Z4) / FZZ;
    const double ECRS = EURS * (1.0 - F * Z4) + EPRS * F * Z4 - ALFRSM * F * (1.0 - Z4) / FZZ;
    const double FZ = THRD4 * (pow(1.0 + ZET, THRD) - 0.0) / GAM;
    const double ECZET = 4.0 * F * (EP - EU + ALFM / FZZ) + FZ * (Z4 * EP - Z4 * EU - (1.0 - Z4) * ALFM / FZZ);
    const double COMM0 = ec - RS * ECRS / 3.0 - ZET * ECZET;
    EC = ec;
    VCUP = COMM0 + ECZET;
    const double G = (pow(1.0 + ZET, THRD2) + 0.0) / 2.0;
    const double G3 = G * G * G;
    const double PON = -ec / (G3 * GAMMA);
    const double B = DELT / (exp(PON) - 1.0);
    const double B2 = B * B, T2 = T * T, T4 = T2 * T2;
    const double Q4 = 1.0 + B * T2;
    const double Q5 = 1.0 + B * T2 + B2 * T4;
    const double h = G3 * (BET / DELT) * log(1.0 + DELT * Q4 * T2 / Q5);
    const double G4 = G3 * G, T6 = T4 * T2, RSTHRD = RS / 3.0;
    const double GZ = (pow((1.0 + ZET) * (1.0 + ZET) + ETA, SIXTHM) - pow(ETA, SIXTHM)) / 3.0;
    const double FAC = DELT / B + 1.0;
    const double BG = -3.0 * B2 * ec * FAC / (BET * G4);
    const double BEC = B2 * FAC / (BET * G3);
    const double Q8 = Q5 * Q5 + DELT * Q4 * Q5 * T2;
    const double Q9 = 1.0 + 2.0 * B * T2;
    const double HB = -BET * G3 * B * T6 * (2.0 + B * T2) / Q8;
    const double HRS = -RSTHRD * HB * BEC * ECRS;
    const double FACT0 = 2.0 * DELT - 6.0 * B;
    const double FACT1 = Q5 * Q9 + Q4 * Q9 * Q9;
    const double HBT = 2.0 * BET * G3 * T4 * ((Q4 * Q5 * FACT0 - DELT * FACT1) / Q8) / Q8;
    const double HRST = RSTHRD * T2 * HBT * BEC * ECRS;
    const double HZ = 3.0 * GZ * h / G + HB * (BG * GZ + BEC * ECZET);
    const double HT = 2.0 * BET * G3 * Q9 / Q8;
    const double FACT2 = Q4 * Q5 + B * T2 * (Q4 * Q9 + Q5);
    const double FACT3 = 2.0 * B * Q5 * Q9 + DELT * FACT2;
    const double HTT = 4.0 * BET * G3 * T * (2.0 * B / Q8 - (Q9 * FACT3 / Q8) / Q8);
    double comm = h + HRS + HRST + T2 * HT / 6.0 + 7.0 * T2 * T * HTT / 6.0;
    const double PREF = HZ - GZ * T2 * HT / G;
    comm = comm - PREF * ZET - UU * HTT - VV * HT;
    H = h;
    DVCUP = comm + PREF;
}

// ---------------------------------------------------------------------------------------------------
// Exact partial derivatives of the PBE energy density.  The energy that xc_point integrates depends on the
// orbital density only through rho and g = |grad rho| (U, V, UU, VV enter the potential alone), so
// e(rho, g) = EXL + EXN + ECL + ECN is restated below as one expression and differentiated in forward mode
// (value + derivative pairs), once along rho and once along g.
// ---------------------------------------------------------------------------------------------------
struct Fwd {
    double v, d;
};
__device__ __forceinline__ Fwd operator+(Fwd a, Fwd b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Fwd operator-(Fwd a, Fwd b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Fwd operator*(Fwd a, Fwd b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Fwd operator/(Fwd a, Fwd b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ __forceinline__ Fwd operator+(double a, Fwd b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Fwd operator-(double a, Fwd b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Fwd operator*(double a, Fwd b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Fwd operator/(double a, Fwd b) { return {a / b.v, -a * b.d / (b.v * b.v)}; }
__device__ __forceinline__ Fwd operator*(Fwd a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Fwd fpow(Fwd a, double e) {
    const double p = pow(a.v, e);
    return {p, e * p / a.v * a.d};
}
__device__ __forceinline__ Fwd fsqrt(Fwd a) {
    const double r = sqrt(a.v);
    return {r, 0.5 / r * a.d};
}
__device__ __forceinline__ Fwd flog(Fwd a) { return {log(a.v), a.d / a.v}; }
__device__ __forceinline__ Fwd fexp(Fwd a) {
    const double e = exp(a.v);
    return {e, e * a.d};
}

// the value part of gcor2 (pbecor.f) in forward mode
__device__ __forceinline__ Fwd gcor2_fwd(double A, double A1, double B1, double B2, double B3, double B4, Fwd RTRS) {
    const Fwd Q0 = -2.0 * A * (1.0 + A1 * (RTRS * RTRS));
    const Fwd Q1 = 2.0 * A * (RTRS * (B1 + RTRS * (B2 + RTRS * (B3 + B4 * RTRS))));
    return Q0 * flog(1.0 + 1.0 / Q1);
}

// e(rho, g): PBE exchange of the up spin at D = 2 rho (pbeex.f) + PBE correlation at zeta = 1 (pbecor.f), per
// unit volume, with the same constants as xc_point
__device__ __forceinline__ Fwd pbe_edens(Fwd rho, Fwd g) {
    const double THIRD = 0.3333333333333333, TRPI2 = 29.6088132032680759, FRDPI = 1.2732395447351627;
    const double DKF2RS = 1.9191582926775128;
    const double AX = -0.738558766382022405884230032680836, UM = 0.2195149727645171, UK = 0.8040, UL = UM / UK;
    const double GAM = 0.5198420997897463295344212145565;
    const double GAMMA = 0.03109069086965489503494086371273, BET = 0.06672455060314922, DELT = BET / GAMMA;
    // exchange: S = 2g / (2 kF D)
    const Fwd D = 2.0 * rho;
    const Fwd DKF = fpow(TRPI2 * D, THIRD);
    const Fwd S = (2.0 * g) * (0.5 / (DKF * D));
    const Fwd EXUNIF = AX * fpow(D, THIRD);
    const Fwd FX = (1.0 + UK) - UK / (1.0 + UL * (S * S));
    const Fwd ex = 0.5 * ((EXUNIF * FX) * D);
    // correlation: T = g / (2 G ks rho)
    const Fwd DKFc = fpow(TRPI2 * rho, THIRD);
    const Fwd RS = DKF2RS / DKFc;
    const Fwd RTRS = fsqrt(RS);
    const Fwd EU = gcor2_fwd(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, RTRS);
    const Fwd EP = gcor2_fwd(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, RTRS);
    const double F = (pow(2.0, 4.0 / 3.0) - 2.0) / GAM;
    const Fwd ec = EU * Fwd{1.0 - F, 0.0} + EP * Fwd{F, 0.0};
    const double G = 0.5 * pow(2.0, 0.6666666666666667), G3 = G * G * G;
    const Fwd DKS = fsqrt(FRDPI * DKFc);
    const Fwd T = g * (0.5 / (DKS * G)) / rho;
    const Fwd B = DELT / (fexp((-1.0 / (G3 * GAMMA)) * ec) - Fwd{1.0, 0.0});
    const Fwd T2 = T * T;
    const Fwd Q4 = 1.0 + B * T2;
    const Fwd Q5 = 1.0 + B * T2 + (B * B) * (T2 * T2);
    const Fwd H = (G3 * (BET / DELT)) * flog(1.0 + DELT * (Q4 * T2) / Q5);
    return ex + rho * (ec + H);
}

// rg: rho, grad(3), Hessian xx,yy,zz,xy,xz,yz of the orbital density; returns v_xc and EXL,EXN,ECL,ECN
// densities per unit weight
template <int GGA>
__device__ __forceinline__ void xc_point(const double *rg, double &vxc, double *e4) {
    const double THIRD = 0.3333333333333333, TRPI2 = 29.6088132032680759, FRDPI = 1.2732395447351627;
    const double DKF2RS = 1.9191582926775128, EPS = 1.0e-20;
    const double rho = rg[0] < 0.0 ? 0.0 : rg[0];
    e4[0] = e4[1] = e4[2] = e4[3] = 0.0;
    vxc = 0.0;
    if (rho < EPS) return;
    double DG2 = 0.0, DGG = 0.0, DLP = 0.0;
    if (GGA) {
        DG2 = rg[1] * rg[1] + rg[2] * rg[2] + rg[3] * rg[3];
        DGG = rg[4] * rg[1] * rg[1] + rg[5] * rg[2] * rg[2] + rg[6] * rg[3] * rg[3] +
              2 * (rg[7] * rg[1] * rg[2] + rg[8] * rg[1] * rg[3] + rg[9] * rg[2] * rg[3]);
        DLP = rg[4] + rg[5] + rg[6];
    }
    {   // exchange of the up spin at density 2 rho
        const double D = 2.0 * rho;
        if (D > EPS) {
            const double DKF = pow(TRPI2 * D, THIRD);
            double S = 0.0, U = 0.0, V = 0.0;
            if (GGA) {
                const double DGRAD = 2.0 * sqrt(DG2);
                const double DLAP = 2.0 * DLP;
                const double REC2 = 0.5 / (DKF * D);
                const double REC1 = REC2 * D;
                const double REC22 = REC2 * REC2;
                S = DGRAD * REC2;
                V = DLAP * REC1 * REC2;
                U = DGRAD > EPS ? 8.0 * DGG * REC1 * REC22 / DGRAD : 0.0;
            }
            double EX, DEX, VX;
            if (GGA) pbeex(D, S, U, V, EX, DEX, VX);
            else pw91ex(DKF, S, U, V, EX, DEX, VX);
            vxc += VX;
            e4[0] += 0.5 * EX * D;
            e4[1] += 0.5 * DEX * D;
        }
    }
    {   // correlation, zeta = 1
        const double D = rho;
        const double DKF = pow(TRPI2 * D, THIRD);
        const double RS = DKF2RS / DKF;
        double EC, VCUP, DEC = 0.0, DVCUP = 0.0;
        if (!GGA) {
            pw91lc_pol(RS, EC, VCUP);
        } else {
            const double DKS = sqrt(FRDPI * DKF);
            const double G = 0.5 * pow(2.0, 0.6666666666666667);
            const double DGRAD = sqrt(DG2);
            const double REC3 = 1.0 / D;
            const double REC1 = 0.5 / (DKS * G);
            const double REC2 = REC1 * REC3;
            const double REC22 = REC2 * REC2;
            const double T = DGRAD * REC2;
            const double VV = DLP * REC1 * REC2;
            const double UU = DGRAD > EPS ? DGG * REC1 * REC22 / DGRAD : 0.0;
            pbecor_pol(RS, T, UU, VV, EC, VCUP, DEC, DVCUP);
        }
        e4[2] += EC * D;
        e4[3] += DEC * D;
        vxc += VCUP + DVCUP;
    }
}

// thread (k, f); per-block partial sums q: 0 XCH1 1 XCOU 2 XEXX 3 -sum w rho vxc 4 EXL 5 EXN 6 ECL 7 ECN
// 8 sum|V_H|
template <int GGA>
__global__ void __launch_bounds__(256) k_xc(const double *__restrict__ phi, int64_t M, int nF,
                                             const double *__restrict__ VH, int64_t ldvh,
                                             const double *__restrict__ w, double *__restrict__ Vtot,
                                             double *__restrict__ U, double *__restrict__ partial, int nblk,
                                             int blk0, double *__restrict__ Ux) {
    __shared__ double red[9][256];
    const int f = blockIdx.y;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double q[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) q[i] = 0.0;
    if (k < M) {
        const int64_t plane = (int64_t)nF * M;
        const int64_t off = (int64_t)f * M + k;
        const double ph = phi[off];
        double rg[10];
        rg[0] = ph * ph;
        if (GGA) {
            const double gx = phi[plane + off], gy = phi[2 * plane + off], gz = phi[3 * plane + off];
            rg[1] = 2 * ph * gx;
            rg[2] = 2 * ph * gy;
            rg[3] = 2 * ph * gz;
            rg[4] = 2 * (gx * gx + phi[4 * plane + off] * ph);
            rg[5] = 2 * (gy * gy + phi[5 * plane + off] * ph);
            rg[6] = 2 * (gz * gz + phi[6 * plane + off] * ph);
            rg[7] = 2 * (gx * gy + phi[7 * plane + off] * ph);
            rg[8] = 2 * (gx * gz + phi[8 * plane + off] * ph);
            rg[9] = 2 * (gy * gz + phi[9 * plane + off] * ph);
        }
        double vxc, e4[4];
        xc_point<GGA>(rg, vxc, e4);
        const double wk = w[k];
        const double vh = VH[(int64_t)f * ldvh + k];
        const double vt = vxc + vh;
        Vtot[off] = vt;
        U[off] = wk * vt * ph;
        if (GGA && Ux) {
            // exact mesh derivative of E_xc + the Hartree term: dE/dphi = w[(e_rho + vh) phi + 2 q phi |grad phi|^2],
            // dE/d(grad phi) = w 2 q phi^2 grad phi, q = e_g / g
            double er = 0.0, qg = 0.0, g2phi = 0.0;
            const double rho = rg[0];
            if (rho >= 1.0e-20) {
                const double g = sqrt(rg[1] * rg[1] + rg[2] * rg[2] + rg[3] * rg[3]);
                er = pbe_edens(Fwd{rho, 1.0}, Fwd{g, 0.0}).d;
                if (g > 1.0e-30) qg = pbe_edens(Fwd{rho, 0.0}, Fwd{g, 1.0}).d / g;
            }
            const int64_t plane = (int64_t)nF * M;
            double gp[3];
            for (int d = 0; d < 3; ++d) {
                gp[d] = phi[(1 + d) * plane + off];
                g2phi += gp[d] * gp[d];
            }
            Ux[off] = wk * ((er + vh) * ph + 2.0 * qg * ph * g2phi);
            for (int d = 0; d < 3; ++d) Ux[(1 + d) * plane + off] = wk * 2.0 * qg * ph * ph * gp[d];
        }
        q[0] = wk * rg[0];
        q[1] = wk * rg[0] * vh;
        q[2] = wk * pow(fabs(rg[0]), (double)(4.0f / 3.0f));
        q[3] = -wk * rg[0] * vxc;
        q[4] = e4[0] * wk;
        q[5] = e4[1] * wk;
        q[6] = e4[2] * wk;
        q[7] = e4[3] * wk;
        q[8] = fabs(vh);
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) red[i][threadIdx.x] = q[i];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int i = 0; i < 9; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x < 9)
        partial[((int64_t)f * nblk + blk0 + blockIdx.x) * 9 + threadIdx.x] = red[threadIdx.x][0];
}

void launch_xc(int xc, int nd, const double *phi, int64_t M, int nF, const double *VH, int64_t ldvh,
               const double *w, double *Vtot, double *U, double *partial, int nblk, int blk0, hipStream_t s,
               double *Ux) {
    const dim3 grid((unsigned)((M + 255) / 256), (unsigned)nF), blk(256);
    if (xc == 1)
        hipLaunchKernelGGL((k_xc<1>), grid, blk, 0, s, phi, M, nF, VH, ldvh, w, Vtot, U, partial, nblk, blk0, Ux);
    else
        hipLaunchKernelGGL((k_xc<0>), grid, blk, 0, s, phi, M, nF, VH, ldvh, w, Vtot, U, partial, nblk, blk0,
                           (double *)nullptr);
    (void)nd;
}

// DFT-side LSDA potential of the total spin densities (APOTNL_DFT with NSPN = 2, getvlxc.F ->
// subvlxc_libxc.F90:290-470 on the LDA path): rho_s = sum of the FLO densities of spin s (the FLOs of a spin are an
// orthonormal rotation of its occupied orbitals), V_H = sum of the FLO Hartree potentials (the potential of the total
// density), exchange per spin at 2 rho_s (PW91EX with s = u = v = 0), PW92 correlation at the point's zeta.
// pot[0, M) spin up, pot[ldpot, ldpot + M) spin down = V_H + V_xc,s; partial[block * 5 + q] = block sums of
// {E_x, E_c, E_H = 1/2 w rho V_H, N_up, N_dn} in a fixed order.
__global__ void __launch_bounds__(256) k_dft_lsda(const double *__restrict__ phi, int64_t M, int nup, int nF,
                                                  const double *__restrict__ VH, int64_t ldvh,
                                                  const double *__restrict__ w, double *__restrict__ pot, int64_t ldpot,
                                                  double *__restrict__ partial) {
    const double THIRD = 0.3333333333333333, TRPI2 = 29.6088132032680759, DKF2RS = 1.9191582926775128;
    const double EPS = 1.0e-20;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double q[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (k < M) {
        double ru = 0.0, rd = 0.0, vh = 0.0;
        for (int f = 0; f < nF; ++f) {
            const double p = phi[(int64_t)f * M + k];
            if (f < nup)
                ru += p * p;
            else
                rd += p * p;
            vh += VH[(int64_t)f * ldvh + k];
        }
        const double D = ru + rd;
        double vxu = 0.0, vxd = 0.0, vcu = 0.0, vcd = 0.0, ex = 0.0, ec = 0.0;
        if (D >= EPS) {
            for (int sp = 0; sp < 2; ++sp) {
                const double Ds = 2.0 * (sp == 0 ? ru : rd);
                if (Ds > EPS) {
                    double EX, DEX, VX;
                    pw91ex(pow(TRPI2 * Ds, THIRD), 0.0, 0.0, 0.0, EX, DEX, VX);
                    (sp == 0 ? vxu : vxd) = VX;
                    ex += 0.5 * EX * Ds;
                }
            }
            double EC;
            pw91lc_gen(DKF2RS / pow(TRPI2 * D, THIRD), (ru - rd) / D, EC, vcu, vcd);
            ec = EC * D;
        }
        pot[k] = vh + vxu + vcu;
        pot[ldpot + k] = vh + vxd + vcd;
        const double wk = w[k];
        q[0] = wk * ex;
        q[1] = wk * ec;
        q[2] = 0.5 * wk * D * vh;
        q[3] = wk * ru;
        q[4] = wk * rd;
    }
    __shared__ double red[5][256];
#pragma unroll
    for (int i = 0; i < 5; ++i) red[i][threadIdx.x] = q[i];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int i = 0; i < 5; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[(int64_t)blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

void launch_dft_lsda(const double *phi, int64_t M, int nup, int nF, const double *VH, int64_t ldvh, const double *w,
                     double *pot, int64_t ldpot, double *partial, hipStream_t s) {
    if (M <= 0) return;
    hipLaunchKernelGGL(k_dft_lsda, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, phi, M, nup, nF, VH, ldvh, w,
                       pot, ldpot, partial);
}

// PBE correlation at general zeta (pbecor.f, LGGA = LPOT = 1): EC, VCUP, VCDN, H, DVCUP, DVCDN
__device__ __forceinline__ void pbecor_gen(double RS, double ZET, double T, double UU, double VV, double WW, double &EC,
                                           double &VCUP, double &VCDN, double &H, double &DVCUP, double &DVCDN) {
    const double THRD = 1.0 / 3.0, THRD2 = 2.0 * THRD, THRD4 = 4.0 * THRD, SIXTHM = -THRD / 2.0;
    const double GAM = 0.5198420997897463295344212145565, FZZ = 8.0 / (9.0 * GAM);
    const double GAMMA = 0.03109069086965489503494086371273, BET = 0.06672455060314922, DELT = BET / GAMMA;
    const double ETA = 1.0e-12;
    const double RTRS = sqrt(RS);
    double EU, EURS, EP, EPRS, ALFM, ALFRSM;
    gcor2(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294, RTRS, EU, EURS);
    gcor2(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517, RTRS, EP, EPRS);
    gcor2(0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671, RTRS, ALFM, ALFRSM);
    const double Z3 = ZET * ZET * ZET, Z4 = Z3 * ZET;
    const double F = (pow(1.0 + ZET, THRD4) + pow(1.0 - ZET, THRD4) - 2.0) / GAM;
    const double ec = EU * (1.0 - F * Z4) + EP * F * Z4 - ALFM * F * (1.0 - Z4) / FZZ;
    const double ECRS = EURS * (1.0 - F * Z4) + EPRS * F * Z4 - ALFRSM * F * (1.0 - Z4) / FZZ;
    const double FZ = THRD4 * (pow(1.0 + ZET, THRD) - pow(1.0 - ZET, THRD)) / GAM;
    const double ECZET = 4.0 * Z3 * F * (EP - EU + ALFM / FZZ) + FZ * (Z4 * EP - Z4 * EU - (1.0 - Z4) * ALFM / FZZ);
    const double COMM0 = ec - RS * ECRS / 3.0 - ZET * ECZET;
    EC = ec;
    VCUP = COMM0 + ECZET;
    VCDN = COMM0 - ECZET;
    const double G = (pow(1.0 + ZET, THRD2) + pow(1.0 - ZET, THRD2)) / 2.0;
    const double G3 = G * G * G;
    const double PON = -ec / (G3 * GAMMA);
    const double B = DELT / (exp(PON) - 1.0);
    const double B2 = B * B, T2 = T * T, T4 = T2 * T2;
    const double Q4 = 1.0 + B * T2;
    const double Q5 = 1.0 + B * T2 + B2 * T4;
    const double h = G3 * (BET / DELT) * log(1.0 + DELT * Q4 * T2 / Q5);
    const double G4 = G3 * G, T6 = T4 * T2, RSTHRD = RS / 3.0;
    const double GZ = (pow((1.0 + ZET) * (1.0 + ZET) + ETA, SIXTHM) - pow((1.0 - ZET) * (1.0 - ZET) + ETA, SIXTHM)) / 3.0;
    const double FAC = DELT / B + 1.0;
    const double BG = -3.0 * B2 * ec * FAC / (BET * G4);
    const double BEC = B2 * FAC / (BET * G3);
    const double Q8 = Q5 * Q5 + DELT * Q4 * Q5 * T2;
    const double Q9 = 1.0 + 2.0 * B * T2;
    const double HB = -BET * G3 * B * T6 * (2.0 + B * T2) / Q8;
    const double HRS = -RSTHRD * HB * BEC * ECRS;
    const double FACT0 = 2.0 * DELT - 6.0 * B;
    const double FACT1 = Q5 * Q9 + Q4 * Q9 * Q9;
    const double HBT = 2.0 * BET * G3 * T4 * ((Q4 * Q5 * FACT0 - DELT * FACT1) / Q8) / Q8;
    const double HRST = RSTHRD * T2 * HBT * BEC * ECRS;
    const double HZ = 3.0 * GZ * h / G + HB * (BG * GZ + BEC * ECZET);
    const double HT = 2.0 * BET * G3 * Q9 / Q8;
    const double HZT = 3.0 * GZ * HT / G + HBT * (BG * GZ + BEC * ECZET);
    const double FACT2 = Q4 * Q5 + B * T2 * (Q4 * Q9 + Q5);
    const double FACT3 = 2.0 * B * Q5 * Q9 + DELT * FACT2;
    const double HTT = 4.0 * BET * G3 * T * (2.0 * B / Q8 - (Q9 * FACT3 / Q8) / Q8);
    double comm = h + HRS + HRST + T2 * HT / 6.0 + 7.0 * T2 * T * HTT / 6.0;
    const double PREF = HZ - GZ * T2 * HT / G;
    const double FACT5 = GZ * (2.0 * HT + T * HTT) / G;
    comm = comm - PREF * ZET - UU * HTT - VV * HT - WW * (HZT - FACT5);
    H = h;
    DVCUP = comm + PREF;
    DVCDN = comm - PREF;
}

// |grad|^2, grad.H.grad and the Laplacian of a 10-component density (value, gradient, xx,yy,zz,xy,xz,yz)
__device__ __forceinline__ void gga_invariants(const double *r, double &dg2, double &dgg, double &dlp) {
    dg2 = r[1] * r[1] + r[2] * r[2] + r[3] * r[3];
    dgg = r[4] * r[1] * r[1] + r[5] * r[2] * r[2] + r[6] * r[3] * r[3] +
          2 * (r[7] * r[1] * r[2] + r[8] * r[1] * r[3] + r[9] * r[2] * r[3]);
    dlp = r[4] + r[5] + r[6];
}

// DFT-side PBE potential of the total spin densities (APOTNL_DFT with NSPN = 2 and IGGA = 1: SUBVLXC as the program links it, subvlxc_libxc.F90:197-470; subvlxc.F:138-323 has the same non-libxc branch): per
// point the value, gradient and Hessian of rho_up and rho_dn from the FLO planes (rho = sum phi^2,
// grad rho = 2 sum phi grad phi, H = 2 sum (grad phi grad phi^T + phi H phi)); the total and down densities clipped
// as there; exchange per spin at D = 2 rho_s (PBEEX with s, u, v of 2 rho_s), PBE correlation at the point's zeta with
// t, uu, vv of the total density and the zeta-gradient term ww.  phi: 10 planes of [nF][M]; pot and partial as in
// k_dft_lsda (E_x = EXL + EXN, E_c = ECL + ECN).
__global__ void __launch_bounds__(256) k_dft_gga(const double *__restrict__ phi, int64_t M, int nup, int nF,
                                                 const double *__restrict__ VH, int64_t ldvh,
                                                 const double *__restrict__ w, double *__restrict__ pot, int64_t ldpot,
                                                 double *__restrict__ partial) {
    const double THIRD = 0.3333333333333333, THIRD2 = 0.6666666666666667;
    const double TRPI2 = 29.6088132032680759, FRDPI = 1.2732395447351627, DKF2RS = 1.9191582926775128;
    const double EPS = 1.0e-20;
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double q[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (k < M) {
        const int64_t plane = (int64_t)nF * M;
        double ru[10], rd[10], vh = 0.0;
#pragma unroll
        for (int i = 0; i < 10; ++i) ru[i] = rd[i] = 0.0;
        for (int f = 0; f < nF; ++f) {
            const int64_t off = (int64_t)f * M + k;
            const double ph = phi[off];
            const double gx = phi[plane + off], gy = phi[2 * plane + off], gz = phi[3 * plane + off];
            double *r = f < nup ? ru : rd;
            r[0] += ph * ph;
            r[1] += 2 * ph * gx;
            r[2] += 2 * ph * gy;
            r[3] += 2 * ph * gz;
            r[4] += 2 * (gx * gx + phi[4 * plane + off] * ph);
            r[5] += 2 * (gy * gy + phi[5 * plane + off] * ph);
            r[6] += 2 * (gz * gz + phi[6 * plane + off] * ph);
            r[7] += 2 * (gx * gy + phi[7 * plane + off] * ph);
            r[8] += 2 * (gx * gz + phi[8 * plane + off] * ph);
            r[9] += 2 * (gy * gz + phi[9 * plane + off] * ph);
            vh += VH[(int64_t)f * ldvh + k];
        }
        // RHT(:,1) = total, RHT(:,2) = down, values clipped to [0, total]; the up components are their difference
        double R1[10], R2[10], RU[10];
#pragma unroll
        for (int i = 0; i < 10; ++i) {
            R1[i] = ru[i] + rd[i];
            R2[i] = rd[i];
        }
        if (R1[0] < 0.0) R1[0] = 0.0;
        if (R2[0] < 0.0) R2[0] = 0.0;
        if (R2[0] > R1[0]) R2[0] = R1[0];
#pragma unroll
        for (int i = 0; i < 10; ++i) RU[i] = R1[i] - R2[i];
        const double DEN1 = R1[0], DEN3 = R2[0], DEN2 = DEN1 - DEN3;
        double DG2[3], DGG[3], DLP[3];
        gga_invariants(R1, DG2[0], DGG[0], DLP[0]);
        gga_invariants(RU, DG2[1], DGG[1], DLP[1]);
        gga_invariants(R2, DG2[2], DGG[2], DLP[2]);
        double v[2] = {0.0, 0.0}, ex = 0.0, ec = 0.0;
        if (DEN1 >= EPS) {
            for (int sp = 0; sp < 2; ++sp) {
                const double D = 2.0 * (sp == 0 ? DEN2 : DEN3);
                if (!(D > EPS)) continue;
                const double DKF = pow(TRPI2 * D, THIRD);
                const double DGRAD = 2.0 * sqrt(DG2[sp + 1]), DLAP = 2.0 * DLP[sp + 1];
                const double REC2 = 0.5 / (DKF * D), REC1 = REC2 * D, REC22 = REC2 * REC2;
                const double S = DGRAD * REC2, V = DLAP * REC1 * REC2;
                const double U = DGRAD > EPS ? 8.0 * DGG[sp + 1] * REC1 * REC22 / DGRAD : 0.0;
                double EX, DEX, VX;
                pbeex(D, S, U, V, EX, DEX, VX);
                v[sp] += VX;
                ex += 0.5 * (EX + DEX) * D;
            }
            const double D = DEN1, DREC = 1.0 / D, ZET = (DEN2 - DEN3) * DREC;
            const double DKF = pow(TRPI2 * D, THIRD), DKS = sqrt(FRDPI * DKF), RS = DKF2RS / DKF;
            double G = 0.5 * (pow(1.0 + ZET, THIRD2) + pow(1.0 - ZET, THIRD2));
            if (ZET == 1.0 || ZET == -1.0) G = 0.5 * pow(2.0, THIRD2);
            const double DGRAD = sqrt(DG2[0]);
            const double REC3 = DREC, REC1 = 0.5 / (DKS * G), REC2 = REC1 * REC3, REC22 = REC2 * REC2;
            const double T = DGRAD * REC2, VV = DLP[0] * REC1 * REC2;
            const double WW = (D * (DG2[1] - DG2[2]) - (DEN2 - DEN3) * DG2[0]) * REC3 * REC22;
            const double UU = DGRAD > EPS ? DGG[0] * REC1 * REC22 / DGRAD : 0.0;
            double EC, VCUP, VCDN, DEC, DVCUP, DVCDN;
            pbecor_gen(RS, ZET, T, UU, VV, WW, EC, VCUP, VCDN, DEC, DVCUP, DVCDN);
            ec = (EC + DEC) * D;
            v[0] += VCUP + DVCUP;
            v[1] += VCDN + DVCDN;
        }
        pot[k] = vh + v[0];
        pot[ldpot + k] = vh + v[1];
        const double wk = w[k];
        q[0] = wk * ex;
        q[1] = wk * ec;
        q[2] = 0.5 * wk * DEN1 * vh;
        q[3] = wk * DEN2;
        q[4] = wk * DEN3;
    }
    __shared__ double red[5][256];
#pragma unroll
    for (int i = 0; i < 5; ++i) red[i][threadIdx.x] = q[i];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int i = 0; i < 5; ++i) red[i][threadIdx.x] += red[i][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[(int64_t)blockIdx.x * 5 + threadIdx.x] = red[threadIdx.x][0];
}

void launch_dft_gga(const double *phi, int64_t M, int nup, int nF, const double *VH, int64_t ldvh, const double *w,
                    double *pot, int64_t ldpot, double *partial, hipStream_t s) {
    if (M <= 0) return;
    hipLaunchKernelGGL(k_dft_gga, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, s, phi, M, nup, nF, VH, ldvh, w,
                       pot, ldpot, partial);
}

// out[q] = sum over b in increasing order of part[b * nq + q]: the host's block-order sum, on one thread (the same
// additions in the same order, so the same bits).
__global__ void k_sum_partials_ordered(const double *__restrict__ part, int nblk, int nq, double *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int q = 0; q < nq; ++q) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * nq + q];
        out[q] = s;
    }
}
void launch_sum_partials_ordered(const double *part, int nblk, int nq, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_partials_ordered, dim3(1), dim3(64), 0, s, part, nblk, nq, out);
}

__global__ void k_fill_zero(double *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}
void launch_fill_zero(double *p, size_t n, hipStream_t s) {
    if (n == 0) return;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, s, p, (int64_t)n);
}

__global__ void k_gather(const double *__restrict__ src, int64_t stride, int n, double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[(int64_t)i * stride];
}
void launch_gather(const double *src, int64_t stride, int n, double *dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, src, stride, n, dst);
}

// deterministic compensated (Neumaier) reduction of per-block partials: thread per (f, q); with accumulate the
// sum starts from out
__global__ void k_reduce_partials(const double *__restrict__ partial, int nF, int nblk, int nq,
                                  double *__restrict__ out, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nF * nq) return;
    const int f = i / nq, q = i % nq;
    double s = accumulate ? out[i] : 0.0, c = 0.0;
    for (int b = 0; b < nblk; ++b) {
        const double x = partial[((int64_t)f * nblk + b) * nq + q];
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    out[i] = s + c;
}

void launch_reduce_partials(const double *partial, int nF, int nblk, int nq, double *out, hipStream_t s,
                            bool accumulate) {
    const int n = nF * nq;
    hipLaunchKernelGGL(k_reduce_partials, dim3((n + 127) / 128), dim3(128), 0, s, partial, nF, nblk, nq, out,
                       accumulate ? 1 : 0);
}

// ---------------------------------------------------------------------------------------------------
// V^i_ab = sum_k Vtot_i(k) [w phi_a phi_b](k): a GEMM over the points with the pair-product operand generated in
// registers.  Block (pair block, point slice, i block): pair block = 8 a x 8 b with a-block <= b-block (V is
// symmetric in a,b; the mirror is stored too), 4 wavefronts, wave w takes a = 8A + 2w + {0,1} and all 8 b, i.e.
// 16 pair columns, against up to 16*NIT i rows.  Points are staged 32 at a time; 16x16x4 FP64 MFMAs, or FP32
// MFMAs (operands rounded to FP32, FP32 accumulation over each 32 staged points, FP64 sums).  Partial sums per
// slice: part[slice][i][b][a].
// ---------------------------------------------------------------------------------------------------
constexpr int kVabKS = 32;

template <typename T, int NIT>
__global__ void __launch_bounds__(256) k_vab_mfma(const double *__restrict__ phi, const double *__restrict__ Vt,
                                                  const double *__restrict__ w, int64_t M, int n, int nb8,
                                                  int64_t kslice, double *__restrict__ part) {
    constexpr int NI = 16 * NIT, LDV = kVabKS + 1;
    using Tr = MfmaTraits<T>;
    __shared__ T Vl[NI * LDV];
    __shared__ double Pa[8 * LDV], Pb[8 * LDV], Wl[kVabKS];
    // pair block index -> (A, B), A <= B, row-major over the upper triangle
    int A = 0, rem = blockIdx.x;
    while (rem >= nb8 - A) {
        rem -= nb8 - A;
        ++A;
    }
    const int B = A + rem;
    const int s = blockIdx.y, i0 = blockIdx.z * NI;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane >> 4, lc = lane & 15;
    const int la = 2 * wave + (lc >> 3), lb = lc & 7;  // this lane's pair inside the 8x8 block
    const int64_t kb = (int64_t)s * kslice, ke = min(M, kb + kslice);
    typename Tr::acc acc[NIT];
    double sum[NIT][4];  // FP32 form: the FP32 accumulators of every 32 staged points are added here
#pragma unroll
    for (int it = 0; it < NIT; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            acc[it][r] = (T)0;
            sum[it][r] = 0.0;
        }
    // the 32 points of the next step are loaded into registers while the MFMAs consume the current step from LDS
    constexpr int NV = NI * kVabKS / 256, NP = 16 * kVabKS / 256;
    double rv[NV], rp[NP], rw = 0.0;
    auto load = [&](int64_t t0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int e = threadIdx.x + 256 * j, i = e / kVabKS, kk = e - i * kVabKS;
            const int64_t k = t0 + kk;
            rv[j] = (i0 + i < n && k < ke) ? Vt[(int64_t)(i0 + i) * M + k] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const int e = threadIdx.x + 256 * j, r = e / kVabKS, kk = e - r * kVabKS;
            const int64_t k = t0 + kk;
            const int f = r < 8 ? 8 * A + r : 8 * B + r - 8;
            rp[j] = (f < n && k < ke) ? phi[(int64_t)f * M + k] : 0.0;
        }
        if (threadIdx.x < kVabKS) rw = t0 + threadIdx.x < ke ? w[t0 + threadIdx.x] : 0.0;
    };
    auto store = [&]() {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            const int e = threadIdx.x + 256 * j, i = e / kVabKS, kk = e - i * kVabKS;
            Vl[i * LDV + kk] = (T)rv[j];
        }
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const int e = threadIdx.x + 256 * j, r = e / kVabKS, kk = e - r * kVabKS;
            if (r < 8)
                Pa[r * LDV + kk] = rp[j];
            else
                Pb[(r - 8) * LDV + kk] = rp[j];
        }
        if (threadIdx.x < kVabKS) Wl[threadIdx.x] = rw;
    };
    load(kb);
    store();
    __syncthreads();
    for (int64_t t0 = kb; t0 < ke; t0 += kVabKS) {
        const bool more = t0 + kVabKS < ke;  // uniform over the block
        if (more) load(t0 + kVabKS);
#pragma unroll
        for (int ks = 0; ks < kVabKS; ks += 4) {
            const int kk = ks + lr;
            const T bv = (T)(Wl[kk] * Pa[la * LDV + kk] * Pb[lb * LDV + kk]);
#pragma unroll
            for (int it = 0; it < NIT; ++it) acc[it] = Tr::mma(Vl[(it * 16 + lc) * LDV + kk], bv, acc[it]);
        }
        if (sizeof(T) == 4) {
#pragma unroll
            for (int it = 0; it < NIT; ++it)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    sum[it][r] += (double)acc[it][r];
                    acc[it][r] = (T)0;
                }
        }
        __syncthreads();
        if (more) store();
        __syncthreads();
    }
    const int a = 8 * A + 2 * wave + (lc >> 3), b = 8 * B + (lc & 7);
    if (a >= n || b >= n) return;
    const int64_t nn = (int64_t)n * n, base = (int64_t)s * n * nn;
#pragma unroll
    for (int it = 0; it < NIT; ++it)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = i0 + it * 16 + Tr::row(lr, r);
            if (i >= n) continue;
            const double v = sizeof(T) == 4 ? sum[it][r] : (double)acc[it][r];
            part[base + i * nn + (int64_t)b * n + a] = v;
            if (A != B) part[base + i * nn + (int64_t)a * n + b] = v;
        }
}

int vab_slices(int64_t M) {
    // point slices per spin: enough blocks to fill the chip, slices of at least 1024 points
    int s = 64;
    while (s > 1 && M / s < 1024) --s;
    return s;
}

void launch_vab(const double *phi0, const double *Vtot, const double *w, int64_t M, int n, int nslice, double *part,
                hipStream_t s, bool fp32) {
    if (n <= 0) return;
    const int nb8 = (n + 7) / 8;
    const int64_t kslice = ((M + nslice - 1) / nslice + kVabKS - 1) / kVabKS * kVabKS;
    const int nit = std::min(8, (n + 15) / 16);
    const dim3 grid((unsigned)(nb8 * (nb8 + 1) / 2), (unsigned)nslice, (unsigned)((n + 16 * nit - 1) / (16 * nit)));
#define FLOSIC_VAB(NIT)                                                                                         \
    if (fp32)                                                                                                   \
        hipLaunchKernelGGL((k_vab_mfma<float, NIT>), grid, dim3(256), 0, s, phi0, Vtot, w, M, n, nb8, kslice, part); \
    else                                                                                                        \
        hipLaunchKernelGGL((k_vab_mfma<double, NIT>), grid, dim3(256), 0, s, phi0, Vtot, w, M, n, nb8, kslice, part)
    switch (nit) {
        case 1: FLOSIC_VAB(1); break;
        case 2: FLOSIC_VAB(2); break;
        case 3: FLOSIC_VAB(3); break;
        case 4: FLOSIC_VAB(4); break;
        case 5: FLOSIC_VAB(5); break;
        case 6: FLOSIC_VAB(6); break;
        case 7: FLOSIC_VAB(7); break;
        default: FLOSIC_VAB(8); break;
    }
#undef FLOSIC_VAB
}

// ---------------------------------------------------------------------------------------------------
// host: Boys table (fffmt3.F:24-65): F_19 by its series, downward recursion, F_0..F_7 kept
// ---------------------------------------------------------------------------------------------------
void boys_table(double *tab) {
    // rows 0 .. kBoysRows-1 in FP64, followed by the same table rounded to FP32
    float *tab32 = (float *)(tab + kBoysRows * kBoysCols);
    for (int i = 0; i < kBoysCols; ++i) tab[i] = 1.0 / (2 * (i + 1) - 1);
    for (int L = 2; L <= kBoysRows; ++L) {
        const double T = 1.0e-2 * (L - 1);
        const double X = 2 * T;
        double den = 39.0, term = 1.0 / den, sum = term;
        for (int i = 2; i <= 100; ++i) {
            den += 2.0;
            term = term * X / den;
            sum = sum + term;
            if (term / sum - 1.0e-12 <= 0) break;
        }
        const double ex = exp(-T);
        double G[21];
        G[20] = ex * sum;
        for (int i = 1; i <= 19; ++i) {
            const int kk = 20 - i;
            G[kk] = (X * G[kk + 1] + ex) / (2 * kk - 1);
        }
        for (int i = 0; i < kBoysCols; ++i) tab[(L - 1) * kBoysCols + i] = G[i + 1];
    }
    for (int i = 0; i < kBoysRows * kBoysCols; ++i) tab32[i] = (float)tab[i];
}

}  // namespace flosic

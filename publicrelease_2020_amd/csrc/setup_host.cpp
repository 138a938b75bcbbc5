// Host-side set-up helpers for the synthetic systems (CPU, OpenMP): Becke fuzzy-cell partition weights of
// an atom-centred mesh and the analytic overlap matrix of a general-contraction Cartesian Gaussian basis.
// These replace the reference's mesh construction (vmesh.F) and overlap integrals (ovmxsf.f) for the
// synthetic inputs; they are not on the evaluation hot path.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
constexpr int kNcart[3] = {1, 3, 6};
constexpr int kCpow[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0},
                              {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
constexpr int kLoff[3] = {0, 1, 4};

// E^{ij}_0 for i, j <= 2 along one axis (McMurchie-Davidson), K folded in
void e0_table(double xpa, double xpb, double p, double K, double e0[3][3]) {
    double E[3][3][5] = {};
    const double inv2p = 0.5 / p;
    E[0][0][0] = K;
    for (int i = 0; i <= 2; ++i)
        for (int j = 0; j <= 2; ++j) {
            if (i == 0 && j == 0) continue;
            const int pi = i > 0 ? i - 1 : i, pj = i > 0 ? j : j - 1;
            const double sh = i > 0 ? xpa : xpb;
            const int np = i + j - 1;
            for (int t = 0; t <= i + j; ++t) {
                double v = 0.0;
                if (t >= 1 && t - 1 <= np) v += inv2p * E[pi][pj][t - 1];
                if (t <= np) v += sh * E[pi][pj][t];
                if (t + 1 <= np) v += (t + 1) * E[pi][pj][t + 1];
                E[i][j][t] = v;
            }
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) e0[i][j] = E[i][j][0];
}
}  // namespace

extern "C" {

// w[k] *= P_owner(k)(r_k): Becke (1988) cell functions, three iterations of p(x) = 1.5x - 0.5x^3
void setup_becke(int natom, const double *xyz, int64_t npts, const double *pts, const int32_t *owner, double *w) {
    std::vector<double> inv_rab((size_t)natom * natom, 0.0);
    for (int a = 0; a < natom; ++a)
        for (int b = 0; b < natom; ++b)
            if (a != b) {
                const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1], dz = xyz[3 * a + 2] - xyz[3 * b + 2];
                inv_rab[(size_t)a * natom + b] = 1.0 / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
#pragma omp parallel
    {
        std::vector<double> d(natom), cell(natom);
#pragma omp for schedule(static)
        for (int64_t k = 0; k < npts; ++k) {
            for (int a = 0; a < natom; ++a) {
                const double dx = pts[3 * k] - xyz[3 * a], dy = pts[3 * k + 1] - xyz[3 * a + 1], dz = pts[3 * k + 2] - xyz[3 * a + 2];
                d[a] = std::sqrt(dx * dx + dy * dy + dz * dz);
            }
            double tot = 0.0;
            for (int a = 0; a < natom; ++a) {
                double c = 1.0;
                for (int b = 0; b < natom && c != 0.0; ++b) {
                    if (b == a) continue;
                    double f = (d[a] - d[b]) * inv_rab[(size_t)a * natom + b];
                    for (int it = 0; it < 3; ++it) f = 1.5 * f - 0.5 * f * f * f;
                    c *= 0.5 * (1.0 - f);
                }
                cell[a] = c;
                tot += c;
            }
            w[k] *= tot > 0.0 ? cell[owner[k]] / tot : 0.0;
        }
    }
}

// S (nbas x nbas, symmetric) over contracted Cartesian functions, ordering atom, l, component, contraction
void setup_overlap(int natom, const double *xyz, const int32_t *bare_off, const double *alpha, const int32_t *ncon,
                   const int32_t *con_off, const double *con, int nbas, double *S) {
    std::vector<int> start(3 * natom);
    int k = 0;
    for (int a = 0; a < natom; ++a)
        for (int l = 0; l < 3; ++l) {
            start[3 * a + l] = k;
            k += kNcart[l] * ncon[3 * a + l];
        }
    std::memset(S, 0, sizeof(double) * (size_t)nbas * nbas);
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int A = 0; A < natom; ++A)
        for (int B = 0; B < natom; ++B) {
            const int na = bare_off[A + 1] - bare_off[A], nb = bare_off[B + 1] - bare_off[B];
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const double ai = alpha[bare_off[A] + i], bj = alpha[bare_off[B] + j];
                    const double p = ai + bj;
                    double e0[3][3][3];
                    for (int d = 0; d < 3; ++d) {
                        const double P = (ai * xyz[3 * A + d] + bj * xyz[3 * B + d]) / p;
                        const double dd = xyz[3 * A + d] - xyz[3 * B + d];
                        e0_table(P - xyz[3 * A + d], P - xyz[3 * B + d], p, std::exp(-ai * bj / p * dd * dd), e0[d]);
                    }
                    const double pref = std::pow(M_PI / p, 1.5);
                    for (int la = 0; la < 3; ++la) {
                        const int nca = ncon[3 * A + la];
                        for (int lb = 0; lb < 3; ++lb) {
                            const int ncb = ncon[3 * B + lb];
                            if (nca == 0 || ncb == 0) continue;
                            for (int ma = 0; ma < kNcart[la]; ++ma)
                                for (int mb = 0; mb < kNcart[lb]; ++mb) {
                                    const int *pa = kCpow[kLoff[la] + ma], *pb = kCpow[kLoff[lb] + mb];
                                    const double prim = pref * e0[0][pa[0]][pb[0]] * e0[1][pa[1]][pb[1]] * e0[2][pa[2]][pb[2]];
                                    if (prim == 0.0) continue;
                                    for (int ca = 0; ca < nca; ++ca) {
                                        const double cfa = con[con_off[3 * A + la] + ca * na + i];
                                        if (cfa == 0.0) continue;
                                        const int mu = start[3 * A + la] + ma * nca + ca;
                                        for (int cb = 0; cb < ncb; ++cb) {
                                            const double cfb = con[con_off[3 * B + lb] + cb * nb + j];
                                            if (cfb == 0.0) continue;
                                            const int nu = start[3 * B + lb] + mb * ncb + cb;
                                            S[(size_t)mu * nbas + nu] += cfa * cfb * prim;
                                        }
                                    }
                                }
                        }
                    }
                }
        }
}

}  // extern "C"

// Host side of libflosic_hip.so: context, set-up (basis, mesh, pair list, resident basis planes), the
// per-evaluation pipeline on one HIP stream, the FP64 Fermi-orbital / Lowdin / force algebra, and the
// C ABI declared in include/flosic_hip.h.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flosic_hip.h"
#include "flosic_internal.h"

using namespace flosic;

namespace {

double wall_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// f(i0, i1) over [0, n) split into contiguous ranges on up to kHostThreads host threads (the calling thread takes the
// first range); with worth false (work too small to pay for starting threads, some tens of microseconds each) it all
// runs on the calling thread.  Work started from inside a range runs on the calling thread alone.  Results do not
// depend on the thread count: every range writes its own outputs and nothing is reduced across ranges.
constexpr int kHostThreads = 8;
constexpr int kThreadMinOrbitals = 32;  // orbitals per spin below which both spins' n^3 algebra shares one thread
thread_local bool t_in_range = false;
template <typename F>
void host_parallel(int n, F &&f, bool worth = true) {
    int nt = (t_in_range || !worth) ? 1 : std::min(n, kHostThreads);
    const unsigned hw = std::thread::hardware_concurrency();
    if (hw > 0) nt = std::min(nt, (int)hw);
    if (nt <= 1) {
        if (n > 0) f(0, n);
        return;
    }
    std::vector<std::thread> th;
    int t = 1;
    try {
        for (; t < nt; ++t)
            th.emplace_back([&f, n, nt, t] {
                t_in_range = true;
                f((int)((int64_t)n * t / nt), (int)((int64_t)n * (t + 1) / nt));
            });
    } catch (...) {  // no thread to be had: the remaining ranges run here
    }
    const bool outer = t_in_range;
    t_in_range = true;
    f(0, (int)((int64_t)n / nt));
    for (int u = t; u < nt; ++u) f((int)((int64_t)n * u / nt), (int)((int64_t)n * (u + 1) / nt));
    t_in_range = outer;
    for (auto &x : th) x.join();
}

constexpr int kNcart[3] = {1, 3, 6};
constexpr double kZed = 1.0e-30;  // coupot_sic.F:61
// three-piece split-BF16 Coulomb: (tile, chunk) pairs whose every T = p|r - P|^2 exceeds this run with two pieces.
// 30 is POISSON1's THIRTYSOMETHING (poisson1.f:101): beyond RCUT = sqrt(30/DELTA_MIN) + DMAX (poisson1.f:528) the
// reference replaces the exact pair integrals by its multipole far field.
constexpr double kFarTDefault = 30.0;

struct Status {
    int code = 0;
    std::string msg;
};

// device bytes held by one context: current and high-water mark
struct MemTally {
    size_t cur = 0, peak = 0;
    uint64_t allocs = 0;  // allocations so far: a captured graph is valid while this is unchanged
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    MemTally *tally = nullptr;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            if (tally) tally->cur -= n * sizeof(T);
        }
        p = nullptr;
        n = 0;
    }
    // (re)allocate to at least cnt elements; returns false on failure
    bool ensure(size_t cnt, MemTally *t = nullptr) {
        if (cnt <= n && p) return true;
        release();
        if (cnt == 0) cnt = 1;
        if (hipMalloc(&p, cnt * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = cnt;
        tally = t;
        if (t) {
            ++t->allocs;
            t->cur += n * sizeof(T);
            t->peak = std::max(t->peak, t->cur);
        }
        return true;
    }
};

struct Spin {
    int nwf = 0, nocc = 0;
    std::vector<double> coef;  // nbas x nwf col-major
    std::vector<double> fod;   // 3 x nocc
    bool have_coef = false, have_fod = false;
    // FO / Lowdin (host, FP64)
    std::vector<double> fmat;  // FMAT(j,i,K): n x n x 4
    std::vector<double> over, q, U, sinvh, tmat;
    // results
    std::vector<double> results;  // 13 x n
    std::vector<double> sic;      // nwf x n
    std::vector<double> sic_ex;   // n x n: exact mesh derivative form of SIC(1:n,1:n) (PBE; LSDA uses sic)
    std::vector<double> vab;      // n^3
    std::vector<double> debdax, grad;  // 3 x n
    DevBuf<double> d_ctil;        // nbas x nwf: FLOs then virtual KS orbitals
    DevBuf<double> d_cocc;        // nbas x nocc: occupied KS orbitals
    DevBuf<double> d_psi, d_tmat;  // orbitals at the FODs (4 x n x n), TMAT
};

}  // namespace

struct flosic_ctx {
    MemTally mem;  // first member: outlives every buffer below
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    hipEvent_t ev[11] = {};
    Status st;
    // options
    int nspn = 2, xc = FLOSIC_XC_LSDA, prec = FLOSIC_PREC_FP64;
    // basis (host)
    int natom = 0, nbas = 0;
    std::vector<double> xyz, alpha, con;
    std::vector<int> bare_off, ncon, con_off;
    std::vector<int> fstart;  // 3*natom: first function of (atom,l)
    // contracted shells and primitives for the mesh kernels
    DevBuf<double> d_atom_xyz, d_atom_r2cut, d_prim_alpha, d_prim_coef;
    DevBuf<int> d_atom_cs, d_cs_l, d_cs_f0, d_cs_stride, d_cs_p0;
    int ncs = 0;
    // primitive shells and pairs
    std::vector<int> ps_atom, ps_l, ps_bare;
    std::vector<int32_t> pair_list;  // 6 per shell pair, canonical order
    int npair = 0;                   // screened primitive shell pairs
    int ngroup = 0, nrows = 0;       // exponent-pair groups and their Hermite rows
    int segL[6] = {0, 0, 0, 0, 0, 0};
    DevBuf<int> d_ps_f0, d_ps_stride, d_ps_ncart, d_ps_c0, d_ps_nc, d_pair_row;
    DevBuf<double> d_ps_coef;
    DevBuf<PairDev> d_pairs;
    DevBuf<ChunkDev> d_chunks, d_chunks_bf16;
    int nchunks = 0, nchunks_bf16 = 0;
    int64_t rows_padded = 0, rows_padded_bf16 = 0;
    DevBuf<double> d_lapplane, d_lap, d_pot, d_wop, d_small, d_B, d_hao;  // LAGMULT_SLV / AO SIC Hamiltonian
    DevBuf<double> d_scf;  // SCFSICW orbital update: ham_ao, ham_ao C, C^T ham_ao C
    bool lapplane_ok = false;
    DevBuf<float> d_H32;   // H in FP32 MFMA operand order
    DevBuf<double> d_H64;  // H in FP64 MFMA operand order
    DevBuf<__bf16> d_Hs;
    // far-chunk bits of the three-piece split-BF16 form (64-point tiles x far_words), FOD-independent: built once per
    // mesh, basis and threshold; far_rows = streamed (tile, row) pairs of the far chunks
    DevBuf<unsigned> d_far;
    DevBuf<double> d_farrows;
    int far_words = 0;
    double far_t = kFarTDefault;
    bool far_ok = false;
    DevBuf<double4> d_pairP;
    DevBuf<double> d_btab;
    // mesh
    int64_t M = 0;
    DevBuf<double> d_x, d_y, d_z, d_w;
    int64_t kmax = 0;
    double rmax = 0.0;
    // resident basis planes (FOD independent), or, in mesh-block mode (block_pts > 0), the planes of one block of
    // block_pts points, evaluated again for every block of every evaluation
    DevBuf<double> d_planes;
    int planes_nd = 0;
    bool planes_ok = false;
    int64_t block_pts = 0;
    uint64_t cfg_gen = 0;  // bumped by every setter that changes the launch sequence
    // Coulomb column-block width: requested (0 widest, -1 autotune, else fixed) and in use
    int coul_req = 0, coul_nfb = 0;
    std::string tuned_key;
    DevBuf<double> d_tx, d_ty, d_tz, d_tv;  // autotuner point sample and its potentials
    bool timed = true;     // per-phase events (off while a batch graph is captured or replayed)
    // spins
    Spin sp[2];
    // evaluation buffers
    int nF = 0, ldh = 0;
    DevBuf<double> d_fodxyz, d_fodchi;  // FOD coordinates of both spins (x, y, z arrays), basis planes at them
    DevBuf<double> d_cflo, d_dps, d_H, d_phi, d_VH, d_Vtot, d_U, d_part, d_energy, d_G, d_Gpart, d_vabpart, d_vab,
        d_lambda, d_asy, d_Ux, d_lamex, d_dft, d_dftpart, d_dplanes, d_dphi, d_dvt, d_du, d_dG;
    std::vector<double> h_energy, h_asy;
    double zsic = 0.0;
    bool evaluated = false, fo_done = false;
    bool vab_on_host = false;  // V^i_ab of the last evaluation copied to sp[].vab (on the first request)
    // the host results and forces of a single evaluation run while the V^i_ab kernels do: their inputs are copied to the
    // host before those kernels are queued, and ev[EV_HOST] marks the copies done
    bool results_queued = false;
    double timings[10] = {};
    int64_t counts[10] = {};
};

// Many independent systems evaluated together: the Coulomb contractions of all of them run as one launch per
// Coulomb form (the systems keep their own column-block widths), optionally with the whole device part captured
// once in a hipGraph and replayed while no context's buffers or settings change.
struct flosic_batch {
    MemTally mem;
    int device = 0;
    std::vector<flosic_ctx *> ctxs;
    bool use_graph = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_coul = nullptr, ev_end = nullptr, ev_t0 = nullptr, ev_c0 = nullptr,
               ev_c1 = nullptr;
    // launch plan: per class of equal (form, pieces) one launch over the jobs [job0, job0 + njobs) of d_jobs, whose
    // item offsets are d_off[off0 .. off0 + njobs] (each class's own, starting at 0); jobs keep their column widths
    struct Class {
        int form = 0, npieces = 0;
        std::vector<int> members;
        int64_t job0 = 0, off0 = 0, nitems = 0;
        int njobs = 0;
    };
    std::vector<Class> classes;
    DevBuf<CoulJob<double>> d_jobs;  // the job layout does not depend on the operand type
    DevBuf<int64_t> d_off;
    std::vector<uint64_t> plan_sig;  // context signatures the plan was built for
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graph_sig;
    int64_t evals = 0, captures = 0, replays = 0;
    double timings[4] = {};
    Status st;
};

// Scratch of flosic_apotnl_dft_noalloc, one copy per GPU for the life of the process: the two potential planes of
// at most kNoAllocMaxPts points, their 256-point block partials (5 per block) and the 5 ordered sums.  Every call on
// one GPU shares it; calls that overlap on one GPU race.  Host code never passes `scratch` itself to a kernel or a
// copy (on the host the name is a stand-in, not a device address): it uses the address hipGetSymbolAddress returns
// for the current device, looked up once per device when the first context on that device is created.
constexpr int64_t kNoAllocMaxPts = int64_t(1) << 20;
constexpr size_t kNoAllocScratchDoubles = 2 * (size_t)kNoAllocMaxPts + 5 * (size_t)(kNoAllocMaxPts / 256) + 8;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current); nullptr if the lookup fails.
double *scratch_on_device(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    return g_scratch_ptr[device];
}

int fail(flosic_ctx *c, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) {
        c->st.code = code;
        c->st.msg = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                           \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return fail(ctx, FLOSIC_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)
#define BLASCHK(ctx, call)                                                                       \
    do {                                                                                         \
        rocblas_status s_ = (call);                                                              \
        if (s_ != rocblas_status_success) return fail(ctx, FLOSIC_EDEVICE, "%s: rocblas status %d", #call, (int)s_); \
    } while (0)
#define ALLOC(ctx, buf, cnt)                                                                     \
    do {                                                                                         \
        if (!(buf).ensure(cnt, &(ctx)->mem)) return fail(ctx, FLOSIC_ENOMEM, "device allocation of %zu elements failed", (size_t)(cnt)); \
    } while (0)

template <typename T>
int upload(flosic_ctx *c, DevBuf<T> &b, const T *h, size_t cnt) {
    // synchronous: the host buffer may be released as soon as this returns
    ALLOC(c, b, cnt);
    if (cnt) {
        HIPCHK(c, hipMemcpyAsync(b.p, h, cnt * sizeof(T), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}
template <typename T>
int upload(flosic_ctx *c, DevBuf<T> &b, const std::vector<T> &h) {
    return upload(c, b, h.data(), h.size());
}

// symmetric eigen-decomposition: Householder reduction to tridiagonal form, then implicit QL with Wilkinson
// shifts (Wilkinson & Reinsch); q ascending, U column-major with the largest |component| of each vector positive
void eigh_sym(int n, const double *Ain, double *q, double *U) {
    // z: row-major working copy, z[i*n + j]
    std::vector<double> z(Ain, Ain + (size_t)n * n), d(n), e(n);
    auto Z = [&](int i, int j) -> double & { return z[(size_t)i * n + j]; };
    for (int i = n - 1; i > 0; --i) {
        const int l = i - 1;
        double h = 0.0;
        if (l > 0) {
            double scale = 0.0;
            for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
            if (scale == 0.0) {
                e[i] = Z(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    Z(i, k) /= scale;
                    h += Z(i, k) * Z(i, k);
                }
                double f = Z(i, l);
                const double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                Z(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; ++j) {
                    Z(j, i) = Z(i, j) / h;
                    double gg = 0.0;
                    for (int k = 0; k <= j; ++k) gg += Z(j, k) * Z(i, k);
                    for (int k = j + 1; k <= l; ++k) gg += Z(k, j) * Z(i, k);
                    e[j] = gg / h;
                    f += e[j] * Z(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; ++j) {
                    const double ff = Z(i, j);
                    const double gg = e[j] - hh * ff;
                    e[j] = gg;
                    for (int k = 0; k <= j; ++k) Z(j, k) -= ff * e[k] + gg * Z(i, k);
                }
            }
        } else {
            e[i] = Z(i, l);
        }
        d[i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; ++i) {  // accumulate the transformations
        if (d[i] != 0.0) {
            for (int j = 0; j < i; ++j) {
                double g = 0.0;
                for (int k = 0; k < i; ++k) g += Z(i, k) * Z(k, j);
                for (int k = 0; k < i; ++k) Z(k, j) -= g * Z(k, i);
            }
        }
        d[i] = Z(i, i);
        Z(i, i) = 1.0;
        for (int j = 0; j < i; ++j) Z(j, i) = Z(i, j) = 0.0;
    }
    // implicit QL on the tridiagonal (d, e), rotating the rows of v = z^T (row j = eigenvector j)
    std::vector<double> v((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) v[(size_t)j * n + i] = Z(i, j);
    for (int i = 1; i < n; ++i) e[i - 1] = e[i];
    if (n > 0) e[n - 1] = 0.0;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 200) break;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = std::hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    double f = s * e[i];
                    const double b = c * e[i];
                    e[i + 1] = r = std::hypot(f, g);
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    double *vi = &v[(size_t)i * n], *vi1 = &v[(size_t)(i + 1) * n];
                    for (int k = 0; k < n; ++k) {
                        const double t = vi1[k];
                        vi1[k] = s * vi[k] + c * t;
                        vi[k] = c * vi[k] - s * t;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int c = 0; c < n; ++c) {
        const int j = idx[c];
        q[c] = d[j];
        const double *vj = &v[(size_t)j * n];
        int big = 0;
        for (int k = 0; k < n; ++k)
            if (std::fabs(vj[k]) > std::fabs(vj[big])) big = k;
        const double sg = vj[big] < 0 ? -1.0 : 1.0;
        for (int k = 0; k < n; ++k) U[k + (size_t)n * c] = sg * vj[k];
    }
}

// FMAT, OVER, the Lowdin eigen-decomposition, S^-1/2 and TMAT of one spin on the host (frmorb2.F:140-290,
// lowden_files.F) from Psi_K(j, i) = d_K psi_j at FOD i (K = 0 value, 1..3 gradient; psiK[K][i][j]); returns 0, or
// the 1-based index of a Lowdin eigenvalue below 1e-8
int lowdin_host(Spin &s, const std::vector<double> &psiK) {
    const int n = s.nocc;
    const size_t nn = (size_t)n * n;
    s.fmat.assign(nn * 4, 0.0);
    for (int i = 0; i < n; ++i) {
        auto psi = [&](int K, int j) { return psiK[(size_t)K * nn + j + (size_t)n * i]; };
        double rhi[4] = {0, 0, 0, 0};
        for (int j = 0; j < n; ++j)
            for (int K = 0; K < 4; ++K) rhi[K] += psi(0, j) * psi(K, j);
        const double srho = std::sqrt(rhi[0]);
        for (int j = 0; j < n; ++j) s.fmat[j + (size_t)n * i] = psi(0, j) / srho;
        for (int K = 1; K < 4; ++K)
            for (int j = 0; j < n; ++j)
                s.fmat[j + (size_t)n * (i + (size_t)n * K)] =
                    psi(K, j) / srho - s.fmat[j + (size_t)n * i] * rhi[K] / rhi[0];
    }
    s.over.assign(nn, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += s.fmat[k + (size_t)n * i] * s.fmat[k + (size_t)n * j];
            s.over[j + (size_t)n * i] = acc;
        }
    s.q.assign(n, 0.0);
    s.U.assign(nn, 0.0);
    eigh_sym(n, s.over.data(), s.q.data(), s.U.data());
    for (int i = 0; i < n; ++i)
        if (s.q[i] < 1.0e-8) return i + 1;
    s.sinvh.assign(nn, 0.0);
    {
        std::vector<double> Uq(nn);  // U_ak / sqrt(q_k)
        for (int k = 0; k < n; ++k) {
            const double f = 1.0 / std::sqrt(s.q[k]);
            for (int a = 0; a < n; ++a) Uq[a + (size_t)n * k] = s.U[a + (size_t)n * k] * f;
        }
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += Uq[a + (size_t)n * k] * s.U[b + (size_t)n * k];
                s.sinvh[a + (size_t)n * b] = acc;
            }
    }
    s.tmat.assign(nn, 0.0);  // TMAT(i,j) = sum_k S^-1/2(i,k) T(k,j), T(k,j) = FMAT(j,k,1)
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += s.sinvh[i + (size_t)n * k] * s.fmat[j + (size_t)n * k];
            s.tmat[i + (size_t)n * j] = acc;
        }
    return 0;
}

// FRMORB2(-ISPN,-1) force formula (frmorb2.F:612-717) from FMAT, U, q and SIC(1:n,1:n)
void force_formula(const Spin &s, double *deb) {
    const int n = s.nocc, ld = s.nwf;
    auto F = [&](int j, int i, int K) { return s.fmat[j + (size_t)n * (i + (size_t)n * K)]; };
    auto U = [&](int k, int a) { return s.U[k + (size_t)n * a]; };
    auto L = [&](int k, int l) { return s.sic[k + (size_t)ld * l]; };
    std::vector<double> H((size_t)n * n), O((size_t)n * n), sc(n);
    std::fill(deb, deb + 3 * n, 0.0);
    // H(b,a) = sum_kl (U_ka U_lb - U_la U_kb) SIC(k,l) / sqrt(q_a q_b): antisymmetrised SIC in the eigenbasis
    std::vector<double> LU((size_t)n * n), ULU((size_t)n * n);
    for (int k = 0; k < n; ++k)
        for (int b = 0; b < n; ++b) {
            double acc = 0.0;
            for (int l = 0; l < n; ++l) acc += L(k, l) * U(l, b);
            LU[k + (size_t)n * b] = acc;
        }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += U(k, a) * LU[k + (size_t)n * b];
            ULU[a + (size_t)n * b] = acc;  // (U^T SIC U)_ab
        }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
            H[b + (size_t)n * a] = (ULU[a + (size_t)n * b] - ULU[b + (size_t)n * a]) / (std::sqrt(s.q[b]) * std::sqrt(s.q[a]));
    // D_x(f,m) = <F_f | dF_m/da_mx>
    std::vector<double> D(3 * (size_t)n * n);
    for (int x = 0; x < 3; ++x)
        for (int f = 0; f < n; ++f)
            for (int m = 0; m < n; ++m) {
                double acc = 0.0;
                for (int k = 0; k < n; ++k) acc += F(k, f, 0) * F(k, m, x + 1);
                D[(size_t)x * n * n + f + (size_t)n * m] = acc;
            }
    // first term: sum_a U_ma sum_f (sum_b H(b,a) U_fb) D(f,m)
    for (int a = 0; a < n; ++a) {
        for (int f = 0; f < n; ++f) {
            double acc = 0.0;
            for (int b = 0; b < n; ++b) acc += H[b + (size_t)n * a] * U(f, b);
            sc[f] = acc;
        }
        for (int x = 0; x < 3; ++x)
            for (int m = 0; m < n; ++m) {
                double acc = 0.0;
                for (int f = 0; f < n; ++f) acc += sc[f] * D[(size_t)x * n * n + f + (size_t)n * m];
                deb[x + 3 * m] -= U(m, a) * acc;
            }
    }
    // second term with H(b,a) (sqrt q_b - sqrt q_a)/(sqrt q_b + sqrt q_a) and the doubled overlap derivative
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            const double qa = std::sqrt(s.q[a]), qb = std::sqrt(s.q[b]);
            H[b + (size_t)n * a] = H[b + (size_t)n * a] * (qb - qa) / (qb + qa);
        }
    // -0.5 sum_ab (-0.5 h_ba) 2 sum_f (U_fb U_ma + U_mb U_fa) D(f,m) = 0.5 sum_a U_ma [(H^T Y) + (H Y)]_am with
    // Y = U^T D: three n^3 products per direction instead of the n^4 double loop
    std::vector<double> Y((size_t)n * n), Z((size_t)n * n);
    for (int x = 0; x < 3; ++x) {
        const double *Dx = D.data() + (size_t)x * n * n;
        for (int m = 0; m < n; ++m)
            for (int b = 0; b < n; ++b) {
                double acc = 0.0;
                for (int f = 0; f < n; ++f) acc += U(f, b) * Dx[f + (size_t)n * m];
                Y[b + (size_t)n * m] = acc;
            }
        for (int m = 0; m < n; ++m)
            for (int a = 0; a < n; ++a) {
                double acc = 0.0;
                for (int b = 0; b < n; ++b)
                    acc += H[b + (size_t)n * a] * Y[b + (size_t)n * m] + H[a + (size_t)n * b] * Y[b + (size_t)n * m];
                Z[a + (size_t)n * m] = acc;
            }
        for (int m = 0; m < n; ++m) {
            double acc = 0.0;
            for (int a = 0; a < n; ++a) acc += U(m, a) * Z[a + (size_t)n * m];
            deb[x + 3 * m] += 0.5 * acc;
        }
    }
}

// adjoint: Gbar = 2 SIC^T TMAT (= dE/dTMAT), back-propagated through TMAT = S^-1/2 T and S = T T^T
void force_adjoint(const Spin &s, const double *sic, int ld, double *grad) {
    const int n = s.nocc;
    const size_t nn = (size_t)n * n;
    auto F = [&](int j, int i, int K) { return s.fmat[j + (size_t)n * (i + (size_t)n * K)]; };
    std::vector<double> G(nn), B(nn), Bt(nn), W(nn), tmp(nn), Tb(nn);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += sic[k + (size_t)ld * i] * s.tmat[k + (size_t)n * j];
            G[i + (size_t)n * j] = 2.0 * acc;
        }
    for (int i = 0; i < n; ++i)  // B = G T^T
        for (int k = 0; k < n; ++k) {
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc += G[i + (size_t)n * j] * F(j, k, 0);
            B[i + (size_t)n * k] = acc;
        }
    for (int a = 0; a < n; ++a)
        for (int k = 0; k < n; ++k) {
            double acc = 0.0;
            for (int i = 0; i < n; ++i) acc += s.U[i + (size_t)n * a] * B[i + (size_t)n * k];
            tmp[a + (size_t)n * k] = acc;
        }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            double acc = 0.0;
            for (int k = 0; k < n; ++k) acc += tmp[a + (size_t)n * k] * s.U[k + (size_t)n * b];
            const double dq = s.q[a] - s.q[b];
            const double f = std::fabs(dq) > 1e-10 * (s.q[a] + s.q[b])
                                 ? (1.0 / std::sqrt(s.q[a]) - 1.0 / std::sqrt(s.q[b])) / dq
                                 : -0.5 / (s.q[a] * std::sqrt(s.q[a]));
            Bt[a + (size_t)n * b] = acc * f;  // Daleckii-Krein divided difference of q^-1/2
        }
    for (int i = 0; i < n; ++i)
        for (int b = 0; b < n; ++b) {
            double acc = 0.0;
            for (int a = 0; a < n; ++a) acc += s.U[i + (size_t)n * a] * Bt[a + (size_t)n * b];
            tmp[i + (size_t)n * b] = acc;
        }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            double acc = 0.0;
            for (int b = 0; b < n; ++b) acc += tmp[i + (size_t)n * b] * s.U[k + (size_t)n * b];
            W[i + (size_t)n * k] = acc;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double a1 = 0.0, a2 = 0.0;
            for (int k = 0; k < n; ++k) {
                a1 += (W[i + (size_t)n * k] + W[k + (size_t)n * i]) * F(j, k, 0);
                a2 += s.sinvh[i + (size_t)n * k] * G[k + (size_t)n * j];
            }
            Tb[i + (size_t)n * j] = a1 + a2;
        }
    for (int m = 0; m < n; ++m)
        for (int x = 0; x < 3; ++x) {
            double acc = 0.0;
            for (int j = 0; j < n; ++j) acc += Tb[m + (size_t)n * j] * F(j, m, x + 1);
            grad[x + 3 * m] = acc;
        }
}

// ------------------------------------------------------------------------------------------------
// set-up helpers
// ------------------------------------------------------------------------------------------------
int build_basis_device(flosic_ctx *c) {
    std::vector<double> axyz = c->xyz, r2cut(c->natom, 0.0), palpha, pcoef;
    std::vector<int> atom_cs(c->natom + 1, 0), cs_l, cs_f0, cs_stride, cs_p0(1, 0);
    for (int a = 0; a < c->natom; ++a) {
        const int nb = c->bare_off[a + 1] - c->bare_off[a];
        for (int l = 0; l < 3; ++l) {
            const int nc = c->ncon[3 * a + l];
            for (int cc = 0; cc < nc; ++cc) {
                for (int i = 0; i < nb; ++i) {
                    const double cf = c->con[c->con_off[3 * a + l] + cc * nb + i];
                    if (cf == 0.0) continue;
                    const double al = c->alpha[c->bare_off[a] + i];
                    palpha.push_back(al);
                    pcoef.push_back(cf);
                    r2cut[a] = std::max(r2cut[a], kCutExp / al);
                }
                cs_l.push_back(l);
                cs_f0.push_back(c->fstart[3 * a + l] + cc);
                cs_stride.push_back(nc);
                cs_p0.push_back((int)palpha.size());
            }
        }
        atom_cs[a + 1] = (int)cs_l.size();
    }
    c->ncs = (int)cs_l.size();
    int r = 0;
    if ((r = upload(c, c->d_atom_xyz, axyz)) || (r = upload(c, c->d_atom_r2cut, r2cut)) ||
        (r = upload(c, c->d_prim_alpha, palpha)) || (r = upload(c, c->d_prim_coef, pcoef)) ||
        (r = upload(c, c->d_atom_cs, atom_cs)) || (r = upload(c, c->d_cs_l, cs_l)) ||
        (r = upload(c, c->d_cs_f0, cs_f0)) || (r = upload(c, c->d_cs_stride, cs_stride)) ||
        (r = upload(c, c->d_cs_p0, cs_p0)))
        return r;
    return 0;
}

int build_pairs(flosic_ctx *c) {
    // primitive shells (atom, l, bare) carrying a nonzero contraction coefficient (ZED test, coupot_sic.F:384-397)
    c->ps_atom.clear();
    c->ps_l.clear();
    c->ps_bare.clear();
    std::vector<int> ps_f0, ps_stride, ps_ncart, ps_c0, ps_nc;
    std::vector<double> ps_coef;
    for (int a = 0; a < c->natom; ++a) {
        const int nb = c->bare_off[a + 1] - c->bare_off[a];
        for (int l = 0; l < 3; ++l) {
            const int nc = c->ncon[3 * a + l];
            for (int i = 0; i < nb; ++i) {
                bool used = false;
                for (int cc = 0; cc < nc; ++cc)
                    if (std::fabs(c->con[c->con_off[3 * a + l] + cc * nb + i]) > kZed) used = true;
                if (!used) continue;
                c->ps_atom.push_back(a);
                c->ps_l.push_back(l);
                c->ps_bare.push_back(i);
                ps_f0.push_back(c->fstart[3 * a + l]);
                ps_stride.push_back(nc);
                ps_ncart.push_back(kNcart[l]);
                ps_c0.push_back((int)ps_coef.size());
                ps_nc.push_back(nc);
                for (int cc = 0; cc < nc; ++cc) ps_coef.push_back(c->con[c->con_off[3 * a + l] + cc * nb + i]);
            }
        }
    }
    const int nps = (int)c->ps_atom.size();
    // canonical shell pairs ps1 <= ps2, screened by a*b/(a+b)*|A-B|^2 > CUTEXP (coupot_sic.F:293-310)
    c->pair_list.clear();
    auto screened = [&](int A, int B, double ai, double aj) {
        const double dx = c->xyz[3 * A] - c->xyz[3 * B], dy = c->xyz[3 * A + 1] - c->xyz[3 * B + 1],
                     dz = c->xyz[3 * A + 2] - c->xyz[3 * B + 2];
        const double rarg = dx * dx + dy * dy + dz * dz;
        const double rcalp = 1.0 / (ai + aj);
        return ai * aj * rcalp * rarg > kCutExp;
    };
    for (int i = 0; i < nps; ++i)
        for (int j = i; j < nps; ++j) {
            const int A = c->ps_atom[i], B = c->ps_atom[j];
            if (screened(A, B, c->alpha[c->bare_off[A] + c->ps_bare[i]], c->alpha[c->bare_off[B] + c->ps_bare[j]]))
                continue;
            const int32_t row[6] = {A, c->ps_l[i], c->ps_bare[i], B, c->ps_l[j], c->ps_bare[j]};
            c->pair_list.insert(c->pair_list.end(), row, row + 6);
        }
    c->npair = (int)(c->pair_list.size() / 6);
    // primitive centres (atom, bare) in atom-major order with their shells; the screening test depends on the
    // exponents and centres only, so a group survives exactly when each of its shell pairs does
    std::vector<int> pc_atom, pc_bare;
    std::vector<std::array<int, 3>> pc_ps;
    {
        for (int a = 0; a < c->natom; ++a) {
            const int nb = c->bare_off[a + 1] - c->bare_off[a];
            for (int i = 0; i < nb; ++i) {
                std::array<int, 3> ls = {-1, -1, -1};
                bool any = false;
                for (int k = 0; k < nps; ++k)
                    if (c->ps_atom[k] == a && c->ps_bare[k] == i) {
                        ls[c->ps_l[k]] = k;
                        any = true;
                    }
                if (!any) continue;
                pc_atom.push_back(a);
                pc_bare.push_back(i);
                pc_ps.push_back(ls);
            }
        }
    }
    const int npc = (int)pc_atom.size();
    std::vector<PairDev> pairs;
    for (int i = 0; i < npc; ++i)
        for (int j = i; j < npc; ++j) {
            const int A = pc_atom[i], B = pc_atom[j];
            const double ai = c->alpha[c->bare_off[A] + pc_bare[i]], aj = c->alpha[c->bare_off[B] + pc_bare[j]];
            if (screened(A, B, ai, aj)) continue;
            PairDev p{};
            for (int d = 0; d < 3; ++d) {
                p.A[d] = c->xyz[3 * A + d];
                p.B[d] = c->xyz[3 * B + d];
            }
            p.a = ai;
            p.b = aj;
            p.lA = p.lB = 0;
            for (int l = 0; l < 3; ++l) {
                p.psA[l] = pc_ps[i][l];
                p.psB[l] = pc_ps[j][l];
                if (p.psA[l] >= 0) p.lA = l;
                if (p.psB[l] >= 0) p.lB = l;
            }
            p.sym = (i == j) ? 1.0 : 2.0;
            pairs.push_back(p);
        }
    // order by L for the Coulomb kernels (stable: canonical order inside each L); rows contiguous, unpadded
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const PairDev &x, const PairDev &y) { return x.lA + x.lB < y.lA + y.lB; });
    c->ngroup = (int)pairs.size();
    std::vector<double4> pP(c->ngroup);
    std::vector<int> prow(c->ngroup);
    int row = 0;
    for (int k = 0; k < c->ngroup; ++k) {
        PairDev &p = pairs[k];
        const int L = p.lA + p.lB;
        p.row = prow[k] = row;
        row += (L + 1) * (L + 2) * (L + 3) / 6;
        const double pe = p.a + p.b;
        pP[k] = make_double4((p.a * p.A[0] + p.b * p.B[0]) / pe, (p.a * p.A[1] + p.b * p.B[1]) / pe,
                             (p.a * p.A[2] + p.b * p.B[2]) / pe, pe);
    }
    c->nrows = row;
    // chunk lists: consecutive groups of one L, at most `cap_rows` rows, zero-extended to a multiple of `quant`.  With
    // split_l4 (the split-BF16 list, 32-row steps) the L = 4 groups (35 rows) go into two chunk types instead: chunk
    // type 5 holds the first 32 rows of two groups, type 6 the last 3 rows of 21 groups; a whole 35-row group would
    // pad its 64-row chunk with 29 zero rows.
    auto make_chunks = [&](int cap_rows, int quant, int64_t &padded, bool split_l4) {
        std::vector<ChunkDev> out;
        padded = 0;
        auto emit = [&](int k, int e, int type, int rpg) {
            const int kr = ((e - k) * rpg + quant - 1) / quant * quant;
            out.push_back(ChunkDev{k, e, type, kr, pairs[k].row, (int)padded});
            padded += kr;
        };
        for (int k = 0; k < c->ngroup;) {
            const int L = pairs[k].lA + pairs[k].lB;
            const int nt = (L + 1) * (L + 2) * (L + 3) / 6;
            if (split_l4 && L == 4) {
                int end = k;
                while (end < c->ngroup && pairs[end].lA + pairs[end].lB == 4) ++end;
                for (int type = 5; type <= 6; ++type) {
                    const int rpg = type == 5 ? 32 : 3, cap = std::max(1, cap_rows / rpg);
                    for (int b = k; b < end; b += cap) emit(b, std::min(end, b + cap), type, rpg);
                }
                k = end;
                continue;
            }
            const int cap = std::max(1, cap_rows / nt);
            int e = k;
            while (e < c->ngroup && e - k < cap && pairs[e].lA + pairs[e].lB == L) ++e;
            emit(k, e, L, nt);
            k = e;
        }
        return out;
    };
    const std::vector<ChunkDev> chunks = make_chunks(kChunkRows, 4, c->rows_padded, false);
    const std::vector<ChunkDev> chunks16 = make_chunks(kChunkRowsBf16, 32, c->rows_padded_bf16, true);
    c->nchunks = (int)chunks.size();
    c->nchunks_bf16 = (int)chunks16.size();
    // segment offsets: segL[L] = first group with lA+lB >= L
    for (int L = 0; L <= 5; ++L) {
        int k = 0;
        while (k < c->ngroup && pairs[k].lA + pairs[k].lB < L) ++k;
        c->segL[L] = k;
    }
    int r = 0;
    if ((r = upload(c, c->d_pairs, pairs)) || (r = upload(c, c->d_pairP, pP)) || (r = upload(c, c->d_pair_row, prow)) ||
        (r = upload(c, c->d_chunks, chunks)) || (r = upload(c, c->d_chunks_bf16, chunks16)) ||
        (r = upload(c, c->d_ps_f0, ps_f0)) || (r = upload(c, c->d_ps_stride, ps_stride)) ||
        (r = upload(c, c->d_ps_ncart, ps_ncart)) || (r = upload(c, c->d_ps_c0, ps_c0)) ||
        (r = upload(c, c->d_ps_nc, ps_nc)) || (r = upload(c, c->d_ps_coef, ps_coef)))
        return r;
    return 0;
}

BasisDev basis_view(const flosic_ctx *c) {
    BasisDev B;
    B.natom = c->natom;
    B.nbas = c->nbas;
    B.ncs = c->ncs;
    B.atom_xyz = c->d_atom_xyz.p;
    B.atom_r2cut = c->d_atom_r2cut.p;
    B.atom_cs = c->d_atom_cs.p;
    B.cs_l = c->d_cs_l.p;
    B.cs_f0 = c->d_cs_f0.p;
    B.cs_stride = c->d_cs_stride.p;
    B.cs_p0 = c->d_cs_p0.p;
    B.prim_alpha = c->d_prim_alpha.p;
    B.prim_coef = c->d_prim_coef.p;
    return B;
}

int xc_planes(const flosic_ctx *c) { return c->xc == FLOSIC_XC_PBE ? 10 : 1; }

// FOD-independent basis planes on the whole mesh, kept resident (mesh-block mode keeps none)
int ensure_planes(flosic_ctx *c) {
    const int nd = xc_planes(c);
    if (c->block_pts > 0 || (c->planes_ok && c->planes_nd == nd)) return 0;
    const size_t cnt = (size_t)nd * c->nbas * (size_t)c->M;
    ALLOC(c, c->d_planes, cnt);
    HIPCHK(c, hipMemsetAsync(c->d_planes.p, 0, cnt * sizeof(double), c->stream));
    launch_basis_planes<double>(basis_view(c), c->d_x.p, c->d_y.p, c->d_z.p, c->M, nd, c->d_planes.p, c->M, c->stream);
    HIPCHK(c, hipGetLastError());
    c->planes_nd = nd;
    c->planes_ok = true;
    return 0;
}

// Basis planes of the mesh block [k0, k0 + m) into dst (leading dimension m); nd as launch_basis_planes
int block_planes(flosic_ctx *c, int64_t k0, int64_t m, int nd, double *dst) {
    const size_t cnt = (size_t)(nd == -1 ? 1 : nd == -3 ? 3 : nd) * c->nbas * (size_t)m;
    launch_fill_zero(dst, cnt, c->stream);  // a kernel, not a memset node, so that the block loop can be captured
    launch_basis_planes<double>(basis_view(c), c->d_x.p + k0, c->d_y.p + k0, c->d_z.p + k0, m, nd, dst, m, c->stream);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// The mesh is walked in blocks: one block of M points with resident planes, else blocks of block_pts points
// (a multiple of 256, so the 256-point energy partials fall on the same points in both modes)
int64_t mesh_block(const flosic_ctx *c) { return c->block_pts > 0 ? std::min<int64_t>(c->block_pts, c->M) : c->M; }

int check_ready(flosic_ctx *c) {
    if (c->nbas <= 0) return fail(c, FLOSIC_ESTATE, "basis not set");
    if (c->M <= 0) return fail(c, FLOSIC_ESTATE, "mesh not set");
    for (int s = 0; s < 2; ++s) {
        if (!c->sp[s].have_coef) return fail(c, FLOSIC_ESTATE, "orbitals of spin %d not set", s + 1);
        if (!c->sp[s].have_fod) return fail(c, FLOSIC_ESTATE, "FODs of spin %d not set", s + 1);
        if ((int)c->sp[s].fod.size() != 3 * c->sp[s].nocc)
            return fail(c, FLOSIC_EINVAL, "spin %d: number of FODs differs from occupied orbitals", s + 1);
    }
    if (c->sp[0].nocc + c->sp[1].nocc <= 0) return fail(c, FLOSIC_EINVAL, "no occupied orbitals");
    return 0;
}

// FRMORB2(0,0) + FRMIORB(1,1) for both spins: the basis and its gradient at every FOD (the mesh basis kernel with the
// FODs of both spins as its points), Psi_K = C^T chi_K at the FODs (4 products of nocc x nbas x nocc per spin,
// rocBLAS), the n x n Lowdin algebra of the two spins on two host threads, then cflo = C_occ TMAT^T (rocBLAS); the FLO
// coefficients of all FLOs land in d_cflo (nbas x nF, spin 1 first)
int run_frmorb(flosic_ctx *c) {
    const int N = c->nbas;
    const int n0 = c->sp[0].nocc, nfod = n0 + c->sp[1].nocc;
    ALLOC(c, c->d_cflo, (size_t)N * nfod);
    int r = 0;
    {   // FOD coordinates of both spins as three point arrays; chi planes [K][mu][FOD], K = value, d/dx, d/dy, d/dz
        std::vector<double> xyz(3 * (size_t)nfod);
        for (int t = 0; t < nfod; ++t) {
            const double *f = t < n0 ? &c->sp[0].fod[3 * t] : &c->sp[1].fod[3 * (t - n0)];
            for (int d = 0; d < 3; ++d) xyz[(size_t)d * nfod + t] = f[d];
        }
        if ((r = upload(c, c->d_fodxyz, xyz))) return r;
        ALLOC(c, c->d_fodchi, 4 * (size_t)N * nfod);
        launch_fill_zero(c->d_fodchi.p, 4 * (size_t)N * nfod, c->stream);
        if (nfod > 0)
            launch_basis_planes<double>(basis_view(c), c->d_fodxyz.p, c->d_fodxyz.p + nfod, c->d_fodxyz.p + 2 * nfod,
                                        nfod, 4, c->d_fodchi.p, nfod, c->stream);
        HIPCHK(c, hipGetLastError());
    }
    std::vector<double> psiK[2];
    for (int s = 0, off = 0; s < 2; off += c->sp[s].nocc, ++s) {
        Spin &sp = c->sp[s];
        const int n = sp.nocc;
        if (n == 0) continue;
        const size_t nn = (size_t)n * n;
        if ((r = upload(c, sp.d_cocc, sp.coef.data(), (size_t)N * n))) return r;
        ALLOC(c, sp.d_psi, 4 * nn);
        // Psi_K(j, i) = sum_mu C(mu, j) chi_K(mu, a_i): C^T times the transpose of the [mu][FOD] plane of this spin's FODs
        const double one = 1.0, zero = 0.0;
        BLASCHK(c, rocblas_dgemm_strided_batched(c->blas, rocblas_operation_transpose, rocblas_operation_transpose, n, n,
                                                 N, &one, sp.d_cocc.p, N, 0, c->d_fodchi.p + off, nfod,
                                                 (rocblas_stride)N * nfod, &zero, sp.d_psi.p, n, (rocblas_stride)nn, 4));
        psiK[s].resize(4 * nn);
        HIPCHK(c, hipMemcpyAsync(psiK[s].data(), sp.d_psi.p, sizeof(double) * 4 * nn, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int bad[2] = {0, 0};
    host_parallel(2, [&](int s0, int s1) {
        for (int s = s0; s < s1; ++s)
            if (c->sp[s].nocc > 0) bad[s] = lowdin_host(c->sp[s], psiK[s]);
    }, std::min(n0, nfod - n0) >= kThreadMinOrbitals);
    for (int s = 0; s < 2; ++s)
        if (bad[s])
            return fail(c, FLOSIC_ELOWDIN, "Lowdin eigenvalue %d = %.3e < 1e-8 (coincident FODs?)", bad[s],
                        c->sp[s].q[bad[s] - 1]);
    int off = 0;
    for (int s = 0; s < 2; ++s) {
        Spin &sp = c->sp[s];
        const int n = sp.nocc;
        if (n > 0) {
            if ((r = upload(c, sp.d_tmat, sp.tmat))) return r;
            // cflo(mu, i) = sum_j C(mu, j) TMAT(i, j)
            const double one = 1.0, zero = 0.0;
            BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_transpose, N, n, n, &one,
                                     sp.d_cocc.p, N, sp.d_tmat.p, n, &zero, c->d_cflo.p + (size_t)off * N, N));
        }
        off += n;
    }
    c->fo_done = true;
    return 0;
}

// out (rows x n, column-major) (+)= A^T B with A (M x rows) and B (M x n) column-major, leading dimension M: the
// point dimension is split into slices (the products are skinny with K = M) and the slice partials are summed
// in a fixed order, onto out when accumulate is set
int gemm_tn_points(flosic_ctx *c, int rows, int n, const double *A, const double *B, int64_t M, double *out,
                   bool accumulate = false, int64_t lda = 0, int64_t ldb = 0) {
    if (rows <= 0 || n <= 0) return 0;
    if (lda <= 0) lda = M;
    if (ldb <= 0) ldb = M;
    const double one = 1.0, zero = 0.0;
    int nsl = 64;
    while (nsl > 1 && M / nsl < 4096) --nsl;
    const int64_t ms = M / nsl, rest = M - ms * nsl;  // nsl slices of ms points, then the remainder
    ALLOC(c, c->d_Gpart, (size_t)(nsl + 1) * rows * n);
    BLASCHK(c, rocblas_dgemm_strided_batched(c->blas, rocblas_operation_transpose, rocblas_operation_none, rows, n,
                                             (rocblas_int)ms, &one, A, (rocblas_int)lda, ms, B, (rocblas_int)ldb, ms,
                                             &zero, c->d_Gpart.p, rows, (rocblas_stride)rows * n, nsl));
    int nparts = nsl;
    if (rest > 0) {
        BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, rows, n,
                                 (rocblas_int)rest, &one, A + ms * nsl, (rocblas_int)lda, B + ms * nsl,
                                 (rocblas_int)ldb, &zero, c->d_Gpart.p + (size_t)nsl * rows * n, rows));
        nparts = nsl + 1;
    }
    launch_reduce_partials(c->d_Gpart.p, 1, nparts, rows * n, out, c->stream, accumulate);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// phi (nd planes of [ncol][m], plane stride ncol*m) = planes (nd planes of [nbas][m]) . cf (nbas x ncol)
int orbitals_on_points(flosic_ctx *c, const double *planes, int64_t m, int nd, const double *cf, int ncol,
                       double *phi) {
    const double one = 1.0, zero = 0.0;
    const int N = c->nbas;
    for (int d = 0; d < nd; ++d)
        BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, ncol, N,
                                 &one, planes + (size_t)d * N * m, (rocblas_int)m, cf, N, &zero,
                                 phi + (size_t)d * ncol * m, (rocblas_int)m));
    return 0;
}

// Event slots of one evaluation
enum { EV_START, EV_FO, EV_HERM, EV_COUL0, EV_COUL1, EV_PHI, EV_XC, EV_SIC, EV_VAB, EV_END, EV_HOST };

int mark(flosic_ctx *c, int slot) {
    if (c->timed) HIPCHK(c, hipEventRecord(c->ev[slot], c->stream));
    return 0;
}
#define MARK(c, slot)                         \
    do {                                      \
        const int m_ = mark((c), (slot));     \
        if (m_) return m_;                    \
    } while (0)

// Coulomb form of a context
enum CoulForm { COUL_VALU, COUL_FP64, COUL_FP32, COUL_BF16 };
CoulForm coulomb_form(const flosic_ctx *c) {
    const char *alg = getenv("FLOSIC_COULOMB");
    if (alg && !strcmp(alg, "valu")) return COUL_VALU;  // per-thread VALU form, FP64 (kept for A/B measurements)
    if (c->prec == FLOSIC_PREC_BF16X3 || c->prec == FLOSIC_PREC_BF16X2) return COUL_BF16;
    return c->prec == FLOSIC_PREC_MIXED ? COUL_FP32 : COUL_FP64;
}
int coulomb_nfb(const flosic_ctx *c) {
    if (c->coul_nfb > 0) return c->coul_nfb;
    return coulomb_form(c) == COUL_BF16 ? coulomb_cols_split(c->nF) : coulomb_cols(c->nF);
}

// One single-system Coulomb launch of the context's form over M points (x, y, z) into V (ldv), nfb columns
void coulomb_launch(flosic_ctx *c, const double *x, const double *y, const double *z, int64_t M, double *V,
                    int64_t ldv, int nfb) {
    const int nF = c->nF;
    switch (coulomb_form(c)) {
        case COUL_VALU:
            launch_coulomb(c->segL, c->d_pairP.p, c->d_pair_row.p, c->d_H.p, c->ldh, x, y, z, M, c->d_btab.p, nF, V,
                           ldv, c->stream);
            break;
        case COUL_BF16: {
            // the far bits describe the tiles of the whole mesh (not the autotuner's point sample)
            const bool mesh = x == c->d_x.p && M == c->M && c->far_ok && c->far_words > 0;
            launch_coulomb_bf16(c->prec == FLOSIC_PREC_BF16X3 ? 3 : 2, c->d_chunks_bf16.p, c->nchunks_bf16,
                                c->rows_padded_bf16 / 32, c->d_pairP.p, c->d_Hs.p, c->ldh, x, y, z, M, c->d_btab.p, nF,
                                V, ldv, c->stream, nfb, mesh ? c->d_far.p : nullptr, mesh ? c->far_words : 0);
            break;
        }
        case COUL_FP32:
            launch_coulomb_mfma<float>(c->d_chunks.p, c->nchunks, c->rows_padded / 4, c->d_pairP.p, c->d_H32.p, c->ldh,
                                       x, y, z, M, c->d_btab.p, nF, V, ldv, c->stream, nfb);
            break;
        case COUL_FP64:
            launch_coulomb_mfma<double>(c->d_chunks.p, c->nchunks, c->rows_padded / 4, c->d_pairP.p, c->d_H64.p,
                                        c->ldh, x, y, z, M, c->d_btab.p, nF, V, ldv, c->stream, nfb);
            break;
    }
}

// The autotuner: every legal column-block width of the context's Coulomb form is timed on a strided sample of at
// most 32768 mesh points (every point of a small mesh) and the fastest kept.  Results are cached per process by
// form, FLO count, Hermite rows and mesh size, so equal shapes are tuned once.
std::mutex g_tune_mu;
std::map<std::string, int> g_tune_cache;

int autotune_coulomb(flosic_ctx *c) {
    const CoulForm form = coulomb_form(c);
    char key[160];
    snprintf(key, sizeof key, "f%d p%d n%d r%lld m%lld", (int)form, c->prec, c->nF,
             (long long)(form == COUL_BF16 ? c->rows_padded_bf16 : c->rows_padded), (long long)c->M);
    if (c->tuned_key == key && c->coul_nfb > 0) return 0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            c->coul_nfb = it->second;
            c->tuned_key = key;
            return 0;
        }
    }
    c->coul_nfb = 0;
    if (form == COUL_VALU) {
        c->tuned_key = key;
        return 0;
    }
    const int64_t S = std::min<int64_t>(c->M, 32768);
    const int64_t stride = c->M / S;
    ALLOC(c, c->d_tx, (size_t)S);
    ALLOC(c, c->d_ty, (size_t)S);
    ALLOC(c, c->d_tz, (size_t)S);
    ALLOC(c, c->d_tv, (size_t)c->nF * S);
    launch_gather(c->d_x.p, stride, (int)S, c->d_tx.p, c->stream);
    launch_gather(c->d_y.p, stride, (int)S, c->d_ty.p, c->stream);
    launch_gather(c->d_z.p, stride, (int)S, c->d_tz.p, c->stream);
    HIPCHK(c, hipGetLastError());
    int best = 0;
    float best_ms = 0.f;
    for (int nfb : {256, 192, 128, 64}) {
        if (!coulomb_cols_legal(nfb, c->ldh, form == COUL_BF16)) continue;
        float ms = 0.f;
        for (int rep = 0; rep < 2; ++rep) {  // the second launch is timed
            HIPCHK(c, hipEventRecord(c->ev[EV_COUL0], c->stream));
            coulomb_launch(c, c->d_tx.p, c->d_ty.p, c->d_tz.p, S, c->d_tv.p, S, nfb);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipEventRecord(c->ev[EV_COUL1], c->stream));
            HIPCHK(c, hipEventSynchronize(c->ev[EV_COUL1]));
            HIPCHK(c, hipEventElapsedTime(&ms, c->ev[EV_COUL0], c->ev[EV_COUL1]));
        }
        if (best == 0 || ms < best_ms) {
            best = nfb;
            best_ms = ms;
        }
    }
    c->coul_nfb = best;
    c->tuned_key = key;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = best;
    return 0;
}

// Column-block width of this evaluation: fixed, autotuned, or the widest the form allows
int resolve_coulomb_cols(flosic_ctx *c) {
    if (c->coul_req == -1) return autotune_coulomb(c);
    c->coul_nfb = 0;
    if (c->coul_req > 0) {
        if (coulomb_form(c) == COUL_VALU) return 0;
        if (!coulomb_cols_legal(c->coul_req, c->ldh, coulomb_form(c) == COUL_BF16))
            return fail(c, FLOSIC_EINVAL, "Coulomb column block %d does not divide ldh %d for this precision",
                        c->coul_req, c->ldh);
        c->coul_nfb = c->coul_req;
    }
    return 0;
}

// Phase A1: Fermi orbitals, Lowdin, FLO coefficients (FP64; n x n algebra on the host), resident planes
int eval_fo(flosic_ctx *c) {
    int r = check_ready(c);
    if (r) return r;
    c->evaluated = false;
    const int N = c->nbas;
    const int n0 = c->sp[0].nocc, n1 = c->sp[1].nocc;
    const int nF = n0 + n1;
    c->nF = nF;
    {
        const int nfb = coulomb_cols(nF);
        c->ldh = (nF + nfb - 1) / nfb * nfb;
    }
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(c->ev[EV_START], c->stream));
    if ((r = run_frmorb(c))) return r;
    for (int s = 0; s < 2; ++s) {
        Spin &sp = c->sp[s];
        ALLOC(c, sp.d_ctil, (size_t)N * sp.nwf);
        const int off = s == 0 ? 0 : n0;
        if (sp.nocc > 0)
            HIPCHK(c, hipMemcpyAsync(sp.d_ctil.p, c->d_cflo.p + (size_t)off * N, sizeof(double) * N * sp.nocc,
                                     hipMemcpyDeviceToDevice, c->stream));
        if (sp.nwf > sp.nocc)
            HIPCHK(c, hipMemcpyAsync(sp.d_ctil.p + (size_t)sp.nocc * N, sp.coef.data() + (size_t)sp.nocc * N,
                                     sizeof(double) * N * (sp.nwf - sp.nocc), hipMemcpyHostToDevice, c->stream));
    }
    if ((r = ensure_planes(c))) return r;
    MARK(c, EV_FO);
    return 0;
}

// Far-chunk bits of the three-piece split-BF16 form for the whole mesh: built once per mesh, basis and threshold
// (they depend on neither the FODs nor the orbitals).  Threshold <= 0, or more chunks than the kernel's LDS row
// holds, leaves the form at three pieces everywhere.
int ensure_far(flosic_ctx *c) {
    if (c->far_ok) return 0;
    c->far_words = 0;
    if (c->far_t > 0.0 && c->nchunks_bf16 <= 32 * kFarWords && c->M > 0) {
        const int w = (c->nchunks_bf16 + 31) / 32;
        ALLOC(c, c->d_far, (size_t)((c->M + 63) / 64) * w);
        ALLOC(c, c->d_farrows, 1);
        launch_far_bits(c->d_x.p, c->d_y.p, c->d_z.p, c->M, c->d_chunks_bf16.p, c->nchunks_bf16, c->d_pairP.p,
                        c->far_t, w, c->d_far.p, c->d_farrows.p, c->stream);
        HIPCHK(c, hipGetLastError());
        c->far_words = w;
    }
    c->far_ok = true;
    return 0;
}

// Phase A2 (device work only): Hermite pair coefficients H[row][f] of all FLOs and the Coulomb operand of the
// context's form
int eval_hermite(flosic_ctx *c) {
    const int N = c->nbas, nF = c->nF;
    const int nps = (int)c->ps_atom.size();
    ALLOC(c, c->d_dps, (size_t)nps * 6 * nF);
    launch_prim_coef(nps, c->d_ps_f0.p, c->d_ps_stride.p, c->d_ps_ncart.p, c->d_ps_c0.p, c->d_ps_nc.p, c->d_ps_coef.p,
                     c->d_cflo.p, nF, N, c->d_dps.p, c->stream);
    ALLOC(c, c->d_H, (size_t)c->nrows * c->ldh);
    launch_hermite(c->ngroup, c->d_pairs.p, c->d_dps.p, nF, c->d_H.p, c->ldh, c->stream);
    HIPCHK(c, hipGetLastError());
    MARK(c, EV_HERM);
    // the Coulomb operand in MFMA order
    switch (coulomb_form(c)) {
        case COUL_BF16: {
            const int np = c->prec == FLOSIC_PREC_BF16X3 ? 3 : 2;
            ALLOC(c, c->d_Hs, (size_t)np * c->ldh * c->rows_padded_bf16);
            launch_split_h(np, c->d_chunks_bf16.p, c->nchunks_bf16, c->d_H.p, c->ldh, nF, c->rows_padded_bf16 / 32,
                           c->d_Hs.p, c->stream);
            if (np == 3) {
                const int r = ensure_far(c);
                if (r) return r;
            }
            break;
        }
        case COUL_FP32:
            ALLOC(c, c->d_H32, (size_t)c->rows_padded * c->ldh);
            launch_frag_h<float>(c->d_chunks.p, c->nchunks, c->d_H.p, c->ldh, nF, c->d_H32.p, c->stream);
            break;
        case COUL_FP64:
            ALLOC(c, c->d_H64, (size_t)c->rows_padded * c->ldh);
            launch_frag_h<double>(c->d_chunks.p, c->nchunks, c->d_H.p, c->ldh, nF, c->d_H64.p, c->stream);
            break;
        default: break;
    }
    HIPCHK(c, hipGetLastError());
    ALLOC(c, c->d_VH, (size_t)nF * c->M);
    return resolve_coulomb_cols(c);
}

// Phase B (one system): Coulomb potentials of all FLOs, V_H[f][k]
int eval_coulomb(flosic_ctx *c) {
    MARK(c, EV_COUL0);
    coulomb_launch(c, c->d_x.p, c->d_y.p, c->d_z.p, c->M, c->d_VH.p, c->M, coulomb_nfb(c));
    HIPCHK(c, hipGetLastError());
    MARK(c, EV_COUL1);
    return 0;
}

// Phase C: FLOs on the mesh, XC, total potential, energies, SIC matrix and V^i_ab.  With resident planes the mesh
// is one block; in mesh-block mode the planes, FLO fields, V_tot and U of each block are built, used and dropped,
// and G, V^i_ab accumulate over the blocks in block order.
// Results of the evaluation that the host needs (energy partial sums, V_H at the outermost point, SIC(j,i), and for
// PBE SIC_ex) queued for copying to the host on the context stream
int queue_results(flosic_ctx *c) {
    const int nF = c->nF;
    c->h_energy.resize((size_t)nF * 9);
    c->h_asy.resize(nF);
    HIPCHK(c, hipMemcpyAsync(c->h_energy.data(), c->d_energy.p, sizeof(double) * nF * 9, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_asy.data(), c->d_asy.p, sizeof(double) * nF, hipMemcpyDeviceToHost, c->stream));
    {
        size_t lo = 0;
        for (int s = 0; s < 2; ++s) {
            Spin &sp = c->sp[s];
            const int n = sp.nocc;
            sp.sic.resize((size_t)sp.nwf * n);
            if (n > 0)
                HIPCHK(c, hipMemcpyAsync(sp.sic.data(), c->d_lambda.p + lo, sizeof(double) * sp.nwf * n,
                                         hipMemcpyDeviceToHost, c->stream));
            lo += (size_t)sp.nwf * n;
        }
        c->vab_on_host = false;  // V^i_ab (n^3 per spin) stays on the device until it is asked for
    }
    if (c->xc == FLOSIC_XC_PBE) {
        size_t lo = 0;
        for (int s = 0; s < 2; ++s) {
            Spin &sp = c->sp[s];
            const size_t nn = (size_t)sp.nocc * sp.nocc;
            sp.sic_ex.resize(nn);
            if (nn) HIPCHK(c, hipMemcpyAsync(sp.sic_ex.data(), c->d_lamex.p + lo, sizeof(double) * nn,
                                             hipMemcpyDeviceToHost, c->stream));
            lo += nn;
        }
    }
    return 0;
}

// overlap_host: with the fields resident (no mesh blocks), queue the host's results before the V^i_ab kernels so that
// the host works on them while those kernels run (single evaluations; a batch keeps its captured region free of
// copies to host memory)
int eval_fields(flosic_ctx *c, bool overlap_host = false) {
    int r = 0;
    c->results_queued = false;
    const int N = c->nbas, nF = c->nF;
    const int n0 = c->sp[0].nocc;
    const int64_t M = c->M;
    const int nd = xc_planes(c);
    const bool blocked = c->block_pts > 0;
    const int64_t mb = mesh_block(c);
    const int nblk = (int)((M + 255) / 256);
    ALLOC(c, c->d_part, (size_t)nF * nblk * 9);
    ALLOC(c, c->d_energy, (size_t)nF * 9);
    ALLOC(c, c->d_asy, (size_t)nF);
    ALLOC(c, c->d_G, (size_t)N * nF);
    ALLOC(c, c->d_phi, (size_t)nd * nF * mb);
    ALLOC(c, c->d_Vtot, (size_t)nF * mb);
    ALLOC(c, c->d_U, (size_t)nF * mb);
    if (blocked) ALLOC(c, c->d_planes, (size_t)nd * N * mb);
    const bool exact = c->xc == FLOSIC_XC_PBE;  // exact mesh derivative for the FOD gradient
    size_t lamex_cnt = 0;
    for (int s = 0; s < 2; ++s) lamex_cnt += (size_t)c->sp[s].nocc * c->sp[s].nocc;
    if (exact) {
        ALLOC(c, c->d_Ux, (size_t)4 * nF * mb);
        ALLOC(c, c->d_lamex, lamex_cnt);
    }
    size_t vab_cnt = 0, vpart_cnt = 0;
    for (int s = 0; s < 2; ++s) {
        const size_t n = (size_t)c->sp[s].nocc;
        vab_cnt += n * n * n;
        vpart_cnt = std::max(vpart_cnt, (size_t)vab_slices(mb) * n * n * n);
    }
    ALLOC(c, c->d_vab, vab_cnt);
    ALLOC(c, c->d_vabpart, vpart_cnt);
    const bool defer_vab = overlap_host && !blocked;
    auto vab_all = [&](int64_t k0, int64_t m) -> int {  // V^i_ab partials of points [k0, k0 + m), both spins
        size_t vo = 0;
        for (int s = 0; s < 2; ++s) {
            const int n = c->sp[s].nocc;
            if (n == 0) continue;
            const int off = s == 0 ? 0 : n0;
            const int nsl = vab_slices(m);
            launch_vab(c->d_phi.p + (size_t)off * m, c->d_Vtot.p + (size_t)off * m, c->d_w.p + k0, m, n, nsl,
                       c->d_vabpart.p, c->stream, c->prec != FLOSIC_PREC_FP64);
            HIPCHK(c, hipGetLastError());
            launch_reduce_partials(c->d_vabpart.p, 1, nsl, n * n * n, c->d_vab.p + vo, c->stream, k0 > 0);
            vo += (size_t)n * n * n;
        }
        return 0;
    };
    for (int64_t k0 = 0; k0 < M; k0 += mb) {
        const int64_t m = std::min(mb, M - k0);
        const double *planes = c->d_planes.p;
        if (blocked && (r = block_planes(c, k0, m, nd, c->d_planes.p))) return r;
        if ((r = orbitals_on_points(c, planes, m, nd, c->d_cflo.p, nF, c->d_phi.p))) return r;
        if (!blocked) MARK(c, EV_PHI);
        launch_xc(c->xc, nd, c->d_phi.p, m, nF, c->d_VH.p + k0, M, c->d_w.p + k0, c->d_Vtot.p, c->d_U.p, c->d_part.p,
                  nblk, (int)(k0 / 256), c->stream, exact ? c->d_Ux.p : nullptr);
        HIPCHK(c, hipGetLastError());
        if (!blocked) MARK(c, EV_XC);
        if (exact) {  // -SIC_ex(j,i) = sum_k [Ux0_i phi_j + sum_d Ux_d,i d_d phi_j], FLOs of one spin
            size_t lo = 0;
            for (int s = 0; s < 2; ++s) {
                const int n = c->sp[s].nocc;
                const size_t off = s == 0 ? 0 : (size_t)n0;
                for (int d = 0; d < 4 && n > 0; ++d)
                    if ((r = gemm_tn_points(c, n, n, c->d_phi.p + ((size_t)d * nF + off) * m,
                                            c->d_Ux.p + ((size_t)d * nF + off) * m, m, c->d_lamex.p + lo,
                                            k0 > 0 || d > 0)))
                        return r;
                lo += (size_t)n * n;
            }
        }
        // G = Phi_0^T U over the points (split K, fixed-order partial sums)
        if ((r = gemm_tn_points(c, N, nF, planes, c->d_U.p, m, c->d_G.p, k0 > 0))) return r;
        if (!blocked) MARK(c, EV_SIC);
        if (!defer_vab && (r = vab_all(k0, m))) return r;
    }
    if (blocked) MARK(c, EV_PHI);
    launch_reduce_partials(c->d_part.p, nF, nblk, 9, c->d_energy.p, c->stream);
    launch_gather(c->d_VH.p + c->kmax, M, nF, c->d_asy.p, c->stream);  // V_H of every FLO at the outermost point
    if (blocked) MARK(c, EV_XC);
    // SIC_s = -Ctil_s^T G_s
    size_t lam_cnt = 0;
    for (int s = 0; s < 2; ++s) lam_cnt += (size_t)c->sp[s].nwf * c->sp[s].nocc;
    ALLOC(c, c->d_lambda, lam_cnt);
    {
        const double zero = 0.0, mone = -1.0;
        size_t lo = 0;
        for (int s = 0; s < 2; ++s) {
            Spin &sp = c->sp[s];
            const int off = s == 0 ? 0 : n0;
            if (sp.nocc > 0)
                BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, sp.nwf, sp.nocc,
                                         N, &mone, sp.d_ctil.p, N, c->d_G.p + (size_t)off * N, N, &zero,
                                         c->d_lambda.p + lo, sp.nwf));
            lo += (size_t)sp.nwf * sp.nocc;
        }
    }
    if (blocked) MARK(c, EV_SIC);
    if (defer_vab) {
        if ((r = queue_results(c))) return r;
        HIPCHK(c, hipEventRecord(c->ev[EV_HOST], c->stream));
        c->results_queued = true;
        if ((r = vab_all(0, M))) return r;
    }
    MARK(c, EV_VAB);
    return 0;
}

// Phase D: results to the host, RESULTS(1:13), forces, timings
int eval_finish(flosic_ctx *c) {
    const int nF = c->nF;
    int r = 0;
    if (c->results_queued) {  // copies queued ahead of the V^i_ab kernels, which may still be running
        c->results_queued = false;
        HIPCHK(c, hipEventSynchronize(c->ev[EV_HOST]));
    } else {
        if ((r = queue_results(c))) return r;
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    for (int s = 0; s < 2; ++s)
        for (double &x : c->sp[s].sic_ex) x = -x;
    // RESULTS(1:13) per FLO (apotnl_sic.F:564-623)
    c->zsic = 0.0;
    const double e13 = (double)(1.0f / 3.0f);
    for (int s = 0, f = 0; s < 2; ++s) {
        Spin &sp = c->sp[s];
        sp.results.assign((size_t)13 * sp.nocc, 0.0);
        for (int i = 0; i < sp.nocc; ++i, ++f) {
            const double *q = &c->h_energy[(size_t)f * 9];
            if (q[8] < 0.001) return fail(c, FLOSIC_ECOUPOT, "COUPOT IS ZERO for spin %d orbital %d", s + 1, i + 1);
            const double XCOU = q[1];
            const double EXC = q[4] + q[5] + q[6] + q[7];
            const double XENG = -EXC - XCOU / 2.0;
            double *R = &sp.results[(size_t)13 * i];
            R[0] = i + 1;
            R[1] = q[0];
            R[2] = 0.0;
            R[3] = c->h_asy[f] * c->rmax;
            R[4] = q[2] * 0.75 * std::pow(6.0 / M_PI, e13);
            R[5] = -XCOU / 2.0;
            R[6] = q[3] - XCOU;
            R[7] = XENG;
            R[8] = XENG * (2.0 / c->nspn);
            R[9] = q[4];
            R[10] = q[6];
            R[11] = q[5];
            R[12] = q[7];
            c->zsic += R[8];
        }
    }
    host_parallel(2, [&](int s0, int s1) {  // the two spins' forces on two host threads (n^3 work each)
        for (int s = s0; s < s1; ++s) {
            Spin &sp = c->sp[s];
            sp.debdax.assign(3 * (size_t)sp.nocc, 0.0);
            sp.grad.assign(3 * (size_t)sp.nocc, 0.0);
            if (sp.nocc > 0) {
                force_formula(sp, sp.debdax.data());
                if (c->xc == FLOSIC_XC_PBE) force_adjoint(sp, sp.sic_ex.data(), sp.nocc, sp.grad.data());
                else force_adjoint(sp, sp.sic.data(), sp.nwf, sp.grad.data());
            }
        }
    }, std::min(c->sp[0].nocc, c->sp[1].nocc) >= kThreadMinOrbitals);
    HIPCHK(c, hipEventRecord(c->ev[EV_END], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[EV_END]));
    // timing slots: [0] FO+Lowdin [1] orbitals on the mesh (mesh-block mode: the whole block loop) [2] Hermite
    // [3] operand + Coulomb [4] XC + energies [5] SIC matrix [6] V^i_ab [7] host results and forces [8] total
    // [9] Coulomb kernel
    const int from[10] = {EV_START, EV_COUL1, EV_FO, EV_HERM, EV_PHI, EV_XC, EV_SIC, EV_VAB, EV_START, EV_COUL0};
    const int to[10] = {EV_FO, EV_PHI, EV_HERM, EV_COUL1, EV_XC, EV_SIC, EV_VAB, EV_END, EV_END, EV_COUL1};
    for (int i = 0; i < 10; ++i) {
        float ms = 0.f;
        if (c->timed) HIPCHK(c, hipEventElapsedTime(&ms, c->ev[from[i]], c->ev[to[i]]));
        c->timings[i] = ms;
    }
    if (!c->timed) {  // only the whole evaluation is timed
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev[EV_START], c->ev[EV_END]));
        c->timings[8] = ms;
    }
    c->counts[0] = c->npair;
    c->counts[1] = c->nrows;
    c->counts[2] = c->M;
    c->counts[3] = c->nbas;
    c->counts[4] = nF;
    c->counts[5] = c->ngroup;
    c->counts[6] = coulomb_form(c) == COUL_BF16 ? c->rows_padded_bf16 : c->rows_padded;
    c->counts[7] = (int64_t)c->mem.peak;
    c->counts[8] = coulomb_form(c) == COUL_VALU ? 0 : coulomb_nfb(c);
    c->counts[9] = -1;  // read from the device on request
    c->evaluated = true;
    return 0;
}

int evaluate(flosic_ctx *c) {
    int r = 0;
    if ((r = eval_fo(c)) || (r = eval_hermite(c)) || (r = eval_coulomb(c)) || (r = eval_fields(c, true)) ||
        (r = eval_finish(c)))
        return r;
    return 0;
}

// ---------------------------------------------------------------------------------------------------
// Batches of systems
// ---------------------------------------------------------------------------------------------------
int bfail(flosic_batch *b, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    b->st.code = code;
    b->st.msg = buf;
    return code;
}
int bfail_ctx(flosic_batch *b, int i, int code) {
    return bfail(b, code, "system %d: %s", i, b->ctxs[i]->st.msg.c_str());
}
#define BHIPCHK(b, call)                                                                            \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) return bfail(b, FLOSIC_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// what a captured graph or an uploaded plan depends on: every context's allocations and settings
std::vector<uint64_t> batch_sig(const flosic_batch *b) {
    std::vector<uint64_t> v;
    for (const flosic_ctx *c : b->ctxs) {
        v.push_back(c->mem.allocs);
        v.push_back(c->cfg_gen);
        v.push_back((uint64_t)coulomb_form(c));
    }
    return v;
}

// The Coulomb job of one context in its form
CoulJob<double> coulomb_job(const flosic_ctx *c) {
    CoulJob<double> J{};
    const CoulForm form = coulomb_form(c);
    const int nfb = coulomb_nfb(c);
    J.chunks = form == COUL_BF16 ? c->d_chunks_bf16.p : c->d_chunks.p;
    J.nchunks = form == COUL_BF16 ? c->nchunks_bf16 : c->nchunks;
    J.pP = c->d_pairP.p;
    J.H = form == COUL_BF16 ? (const void *)c->d_Hs.p : form == COUL_FP32 ? (const void *)c->d_H32.p
                                                                           : (const void *)c->d_H64.p;
    J.nct = c->ldh / 16;
    J.X = c->d_x.p;
    J.Y = c->d_y.p;
    J.Z = c->d_z.p;
    J.M = c->M;
    J.nF = c->nF;
    J.V = c->d_VH.p;
    J.ldv = c->M;
    J.nsteps = form == COUL_BF16 ? c->rows_padded_bf16 / 32 : c->rows_padded / 4;
    J.ncb = c->ldh / nfb;
    J.nfb = nfb;
    const bool far = form == COUL_BF16 && c->prec == FLOSIC_PREC_BF16X3 && c->far_ok && c->far_words > 0;
    J.far = far ? c->d_far.p : nullptr;
    J.farw = far ? c->far_words : 0;
    return J;
}

// Classes of equal (form, pieces), each one launch whatever the members' column-block widths; jobs heaviest first
// (work per item ~ K steps x columns), so that the longest items are dispatched first and the hardware dispatcher
// balances the tail.  The job records and item offsets go to device memory once per plan.
int build_plan(flosic_batch *b) {
    static_assert(sizeof(CoulJob<double>) == sizeof(CoulJob<float>) && sizeof(CoulJob<double>) == sizeof(CoulJob<__bf16>),
                  "job layout is independent of the operand type");
    b->classes.clear();
    for (int i = 0; i < (int)b->ctxs.size(); ++i) {
        const flosic_ctx *c = b->ctxs[i];
        const CoulForm form = coulomb_form(c);
        if (form == COUL_VALU || c->nchunks <= 0) continue;
        const int np = form == COUL_BF16 ? (c->prec == FLOSIC_PREC_BF16X3 ? 3 : 2) : 0;
        auto it = std::find_if(b->classes.begin(), b->classes.end(), [&](const flosic_batch::Class &k) {
            return k.form == form && k.npieces == np;
        });
        if (it == b->classes.end()) {
            b->classes.push_back({});
            it = b->classes.end() - 1;
            it->form = form;
            it->npieces = np;
        }
        it->members.push_back(i);
    }
    std::vector<CoulJob<double>> all;
    std::vector<int64_t> offs;
    for (auto &k : b->classes) {
        std::vector<CoulJob<double>> jobs;
        for (int i : k.members) jobs.push_back(coulomb_job(b->ctxs[i]));
        std::stable_sort(jobs.begin(), jobs.end(), [](const CoulJob<double> &x, const CoulJob<double> &y) {
            return x.nsteps * x.nfb > y.nsteps * y.nfb;
        });
        k.job0 = (int64_t)all.size();
        k.off0 = (int64_t)offs.size();
        k.njobs = (int)jobs.size();
        int64_t items = 0;
        for (const auto &J : jobs) {
            all.push_back(J);
            offs.push_back(items);
            items += (J.M + 63) / 64 * J.ncb;
        }
        offs.push_back(items);
        if (items >= (int64_t)INT32_MAX)
            return bfail(b, FLOSIC_EINVAL, "batch too large: %lld Coulomb items in one launch", (long long)items);
        k.nitems = items;
    }
    if (!all.empty()) {
        if (!b->d_jobs.ensure(all.size(), &b->mem) || !b->d_off.ensure(offs.size(), &b->mem))
            return bfail(b, FLOSIC_ENOMEM, "device allocation of the batch job table failed");
        BHIPCHK(b, hipMemcpy(b->d_jobs.p, all.data(), sizeof(CoulJob<double>) * all.size(), hipMemcpyHostToDevice));
        BHIPCHK(b, hipMemcpy(b->d_off.p, offs.data(), sizeof(int64_t) * offs.size(), hipMemcpyHostToDevice));
    }
    b->plan_sig = batch_sig(b);
    return 0;
}

int launch_classes(flosic_batch *b) {
    for (const auto &k : b->classes) {
        const double *btab = b->ctxs[k.members[0]]->d_btab.p;
        const CoulJob<double> *jobs = b->d_jobs.p + k.job0;
        const int64_t *off = b->d_off.p + k.off0;
        if (k.form == COUL_FP64)
            launch_coulomb_batch_all<double>(jobs, off, k.njobs, k.nitems, btab, b->stream);
        else if (k.form == COUL_FP32)
            launch_coulomb_batch_all<float>(reinterpret_cast<const CoulJob<float> *>(jobs), off, k.njobs, k.nitems,
                                            btab, b->stream);
        else
            launch_coulomb_bf16_batch_all(k.npieces, reinterpret_cast<const CoulJob<__bf16> *>(jobs), off, k.njobs,
                                          k.nitems, btab, b->stream);
        BHIPCHK(b, hipGetLastError());
    }
    return 0;
}

// The device part of a batch evaluation, issued on the batch stream and forked to the context streams:
// Hermite coefficients and operands of every system, the batched Coulomb launches, then every system's fields.
int issue_region(flosic_batch *b, bool eager) {
    int r = 0;
    const int n = (int)b->ctxs.size();
    BHIPCHK(b, hipEventRecord(b->ev_fork, b->stream));
    for (int i = 0; i < n; ++i) {
        flosic_ctx *c = b->ctxs[i];
        BHIPCHK(b, hipStreamWaitEvent(c->stream, b->ev_fork, 0));
        if ((r = eval_hermite(c))) return bfail_ctx(b, i, r);
        BHIPCHK(b, hipEventRecord(c->ev[EV_COUL0], c->stream));
        BHIPCHK(b, hipStreamWaitEvent(b->stream, c->ev[EV_COUL0], 0));
    }
    if (eager && b->plan_sig != batch_sig(b) && (r = build_plan(b))) return r;
    if (eager) BHIPCHK(b, hipEventRecord(b->ev_c0, b->stream));
    if ((r = launch_classes(b))) return r;
    if (eager) BHIPCHK(b, hipEventRecord(b->ev_c1, b->stream));
    BHIPCHK(b, hipEventRecord(b->ev_coul, b->stream));
    for (int i = 0; i < n; ++i) {
        flosic_ctx *c = b->ctxs[i];
        BHIPCHK(b, hipStreamWaitEvent(c->stream, b->ev_coul, 0));
        if (coulomb_form(c) == COUL_VALU) {
            if ((r = eval_coulomb(c))) return bfail_ctx(b, i, r);
        } else {
            if ((r = mark(c, EV_COUL1))) return bfail_ctx(b, i, r);
        }
        if ((r = eval_fields(c))) return bfail_ctx(b, i, r);
        BHIPCHK(b, hipEventRecord(c->ev[EV_VAB], c->stream));
        BHIPCHK(b, hipStreamWaitEvent(b->stream, c->ev[EV_VAB], 0));
    }
    return 0;
}

void drop_graph(flosic_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_sig.clear();
}

int batch_evaluate(flosic_batch *b, double *zsic) {
    const double t_host0 = wall_ms();
    int r = 0;
    const int n = (int)b->ctxs.size();
    b->st = Status();
    BHIPCHK(b, hipSetDevice(b->device));
    // host work of the systems (FO/Lowdin before the device region, results and forces after it) on host threads, one
    // system per range step; a failure is reported for the first failing system in batch order
    std::vector<int> rc(n, 0);
    for (int i = 0; i < n; ++i) {
        b->ctxs[i]->st = Status();
        b->ctxs[i]->timed = true;
    }
    host_parallel(n, [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) rc[i] = eval_fo(b->ctxs[i]);
    });
    BHIPCHK(b, hipSetDevice(b->device));
    for (int i = 0; i < n; ++i)
        if (rc[i]) return bfail_ctx(b, i, rc[i]);
    for (int i = 0; i < n; ++i) {
        flosic_ctx *c = b->ctxs[i];
        BHIPCHK(b, hipEventRecord(c->ev[EV_FO], c->stream));
        BHIPCHK(b, hipStreamWaitEvent(b->stream, c->ev[EV_FO], 0));
    }
    const std::vector<uint64_t> sig = batch_sig(b);
    const bool replay = b->use_graph && b->exec && sig == b->graph_sig;
    const bool capture = b->use_graph && !replay && b->evals > 0 && sig == b->plan_sig;
    if (!replay) drop_graph(b);
    for (flosic_ctx *c : b->ctxs) c->timed = !(replay || capture);
    BHIPCHK(b, hipEventRecord(b->ev_t0, b->stream));
    if (replay) {
        BHIPCHK(b, hipGraphLaunch(b->exec, b->stream));
        ++b->replays;
    } else if (capture) {
        BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        r = issue_region(b, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (r || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            for (flosic_ctx *c : b->ctxs) c->timed = true;
            return r ? r : bfail(b, FLOSIC_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
        }
        b->graph = g;
        BHIPCHK(b, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_sig = sig;
        ++b->captures;
        BHIPCHK(b, hipGraphLaunch(b->exec, b->stream));
    } else {
        if ((r = issue_region(b, true))) return r;
    }
    BHIPCHK(b, hipEventRecord(b->ev_end, b->stream));
    for (flosic_ctx *c : b->ctxs) BHIPCHK(b, hipStreamWaitEvent(c->stream, b->ev_end, 0));
    host_parallel(n, [&](int i0, int i1) {
        for (int i = i0; i < i1; ++i) {
            flosic_ctx *c = b->ctxs[i];
            (void)hipSetDevice(c->device);
            rc[i] = eval_finish(c);
            c->timed = true;
        }
    });
    BHIPCHK(b, hipSetDevice(b->device));
    for (int i = 0; i < n; ++i) {
        if (rc[i]) return bfail_ctx(b, i, rc[i]);
        if (zsic) zsic[i] = b->ctxs[i]->zsic;
    }
    if (!replay && !capture) b->plan_sig = batch_sig(b);
    float ms = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->ev_t0, b->ev_end));
    b->timings[1] = ms;
    b->timings[2] = 0.0;
    if (!replay && !capture) {
        BHIPCHK(b, hipEventElapsedTime(&ms, b->ev_c0, b->ev_c1));
        b->timings[2] = ms;
    }
    b->timings[3] = replay ? 2 : capture ? 1 : 0;
    ++b->evals;
    b->timings[0] = wall_ms() - t_host0;
    return 0;
}

int need_eval(flosic_ctx *c) {
    if (c->evaluated) return 0;
    return evaluate(c);
}

int spin_index(flosic_ctx *c, int ispn) {
    if (ispn != 1 && ispn != 2) return fail(c, FLOSIC_EINVAL, "ispn must be 1 or 2 (got %d)", ispn), -1;
    return ispn - 1;
}

}  // namespace

// ====================================================================================================
// C ABI
// ====================================================================================================
extern "C" {

int32_t flosic_create(int32_t device, flosic_ctx **out) {
    if (!out) return FLOSIC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FLOSIC_EDEVICE;
    if (device < 0 || device >= ndev) return FLOSIC_EINVAL;
    flosic_ctx *c = new (std::nothrow) flosic_ctx();
    if (!c) return FLOSIC_ENOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FLOSIC_EDEVICE;
    }
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return FLOSIC_EDEVICE;
        }
    if (rocblas_create_handle(&c->blas) != rocblas_status_success ||
        rocblas_set_stream(c->blas, c->stream) != rocblas_status_success) {
        delete c;
        return FLOSIC_EDEVICE;
    }
    if (!scratch_on_device(device)) {  // the no-allocation entry point's scratch, looked up with the device current
        delete c;
        return FLOSIC_EDEVICE;
    }
    std::vector<double> tab((size_t)kBoysTableDoubles);
    boys_table(tab.data());
    if (upload(c, c->d_btab, tab) || hipStreamSynchronize(c->stream) != hipSuccess) {
        delete c;
        return FLOSIC_EDEVICE;
    }
    *out = c;
    return FLOSIC_OK;
}

int32_t flosic_destroy(flosic_ctx *c) {
    if (!c) return FLOSIC_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->blas) rocblas_destroy_handle(c->blas);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FLOSIC_OK;
}

const char *flosic_last_error(const flosic_ctx *c) { return c ? c->st.msg.c_str() : "null context"; }

int32_t flosic_set_basis(flosic_ctx *c, int32_t natom, const double *xyz, const int32_t *bare_off, const double *alpha,
                         const int32_t *ncon, const int32_t *con_off, const double *con) {
    if (!c) return FLOSIC_EINVAL;
    ++c->cfg_gen;
    if (natom <= 0 || !xyz || !bare_off || !alpha || !ncon || !con_off || !con)
        return fail(c, FLOSIC_EINVAL, "flosic_set_basis: bad arguments");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "hipSetDevice");
    c->natom = natom;
    c->xyz.assign(xyz, xyz + 3 * natom);
    c->bare_off.assign(bare_off, bare_off + natom + 1);
    if (c->bare_off[0] != 0) return fail(c, FLOSIC_EINVAL, "bare_off[0] must be 0");
    c->alpha.assign(alpha, alpha + c->bare_off[natom]);
    for (double a : c->alpha)
        if (!(a > 0.0)) return fail(c, FLOSIC_EINVAL, "exponents must be positive");
    c->ncon.assign(ncon, ncon + 3 * natom);
    c->con_off.assign(con_off, con_off + 3 * natom);
    size_t ncoef = 0;
    c->fstart.assign(3 * natom, 0);
    int nbas = 0;
    for (int a = 0; a < natom; ++a)
        for (int l = 0; l < 3; ++l) {
            if (c->ncon[3 * a + l] < 0) return fail(c, FLOSIC_EINVAL, "negative ncon");
            const size_t end = (size_t)c->con_off[3 * a + l] +
                               (size_t)c->ncon[3 * a + l] * (c->bare_off[a + 1] - c->bare_off[a]);
            ncoef = std::max(ncoef, end);
            c->fstart[3 * a + l] = nbas;
            nbas += kNcart[l] * c->ncon[3 * a + l];
        }
    c->con.assign(con, con + ncoef);
    c->nbas = nbas;
    int r = 0;
    if ((r = build_basis_device(c)) || (r = build_pairs(c))) return r;
    c->planes_ok = false;
    c->far_ok = false;
    c->lapplane_ok = false;
    c->evaluated = false;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "set_basis sync");
    return FLOSIC_OK;
}

int32_t flosic_set_mesh(flosic_ctx *c, int64_t nmsh, const double *rmsh, const double *wmsh) {
    if (!c) return FLOSIC_EINVAL;
    ++c->cfg_gen;
    if (nmsh <= 0 || !rmsh || !wmsh) return fail(c, FLOSIC_EINVAL, "flosic_set_mesh: bad arguments");
    if (nmsh > (int64_t)INT32_MAX / 2) return fail(c, FLOSIC_EINVAL, "mesh too large");
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "hipSetDevice");
    std::vector<double> x(nmsh), y(nmsh), z(nmsh), w(wmsh, wmsh + nmsh);
    c->rmax = 0.0;
    c->kmax = 0;
    for (int64_t k = 0; k < nmsh; ++k) {
        x[k] = rmsh[3 * k];
        y[k] = rmsh[3 * k + 1];
        z[k] = rmsh[3 * k + 2];
        const double r = std::sqrt(x[k] * x[k] + y[k] * y[k] + z[k] * z[k]);
        if (r > c->rmax) {  // first point of largest radius (apotnl_sic.F:409-418)
            c->rmax = r;
            c->kmax = k;
        }
    }
    c->M = nmsh;
    int r = 0;
    if ((r = upload(c, c->d_x, x)) || (r = upload(c, c->d_y, y)) || (r = upload(c, c->d_z, z)) ||
        (r = upload(c, c->d_w, w)))
        return r;
    c->planes_ok = false;
    c->far_ok = false;
    c->lapplane_ok = false;
    c->evaluated = false;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "set_mesh sync");
    return FLOSIC_OK;
}

int32_t flosic_set_options(flosic_ctx *c, int32_t nspn, int32_t xc, int32_t precision) {
    if (!c) return FLOSIC_EINVAL;
    ++c->cfg_gen;
    if (nspn != 1 && nspn != 2) return fail(c, FLOSIC_EINVAL, "nspn must be 1 or 2");
    if (xc != FLOSIC_XC_LSDA && xc != FLOSIC_XC_PBE) return fail(c, FLOSIC_EINVAL, "unknown functional %d", xc);
    if (precision < FLOSIC_PREC_FP64 || precision > FLOSIC_PREC_BF16X2)
        return fail(c, FLOSIC_EINVAL, "unknown precision %d", precision);
    if (xc != c->xc) c->planes_ok = false;
    if (nspn != c->nspn) {
        // NSPN = 1 (spin-restricted): one spin channel, its SIC energies counted twice (RESULTS(9) = XENG * 2 / NSPN,
        // apotnl_sic.F:581); spin 2 is empty.  Back to NSPN = 2, spin 2 must be set again.
        Spin &s2 = c->sp[1];
        s2.nwf = s2.nocc = 0;
        s2.coef.clear();
        s2.fod.clear();
        s2.have_coef = s2.have_fod = nspn == 1;
        c->fo_done = false;
    }
    c->nspn = nspn;
    c->xc = xc;
    c->prec = precision;
    c->evaluated = false;
    return FLOSIC_OK;
}

int32_t flosic_set_mesh_block(flosic_ctx *c, int64_t block_points) {
    if (!c) return FLOSIC_EINVAL;
    if (block_points < 0) return fail(c, FLOSIC_EINVAL, "block_points must be >= 0");
    const int64_t b = block_points == 0 ? 0 : (block_points + 255) / 256 * 256;
    if (b != c->block_pts) {
        if (hipSetDevice(c->device) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "hipSetDevice");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "set_mesh_block sync");
        // mesh-sized buffers of the other mode are released
        c->d_planes.release();
        c->d_lapplane.release();
        c->d_phi.release();
        c->d_Vtot.release();
        c->d_U.release();
        c->d_lap.release();
        c->d_wop.release();
        c->planes_ok = false;
        c->lapplane_ok = false;
        c->evaluated = false;
    }
    c->block_pts = b;
    ++c->cfg_gen;
    return FLOSIC_OK;
}

int32_t flosic_set_coulomb_cols(flosic_ctx *c, int32_t ncols) {
    if (!c) return FLOSIC_EINVAL;
    if (ncols != -1 && ncols != 0 && ncols != 64 && ncols != 128 && ncols != 192 && ncols != 256)
        return fail(c, FLOSIC_EINVAL, "Coulomb column block must be -1, 0, 64, 128, 192 or 256 (got %d)", ncols);
    c->coul_req = ncols;
    c->coul_nfb = 0;
    c->tuned_key.clear();
    c->evaluated = false;
    ++c->cfg_gen;
    return FLOSIC_OK;
}

int32_t flosic_set_far_threshold(flosic_ctx *c, double tfar) {
    if (!c) return FLOSIC_EINVAL;
    if (!(tfar == tfar)) return fail(c, FLOSIC_EINVAL, "far threshold is NaN");
    c->far_t = tfar;
    c->far_ok = false;
    c->evaluated = false;
    ++c->cfg_gen;
    return FLOSIC_OK;
}

int32_t flosic_set_orbitals(flosic_ctx *c, int32_t ispn, int32_t nbas, int32_t nwf, int32_t nocc, const double *coef) {
    if (!c) return FLOSIC_EINVAL;
    ++c->cfg_gen;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    if (s == 1 && c->nspn == 1) return fail(c, FLOSIC_EINVAL, "NSPN = 1: spin 2 is not used");
    if (nbas != c->nbas) return fail(c, FLOSIC_EINVAL, "nbas %d differs from the basis (%d)", nbas, c->nbas);
    if (nocc < 0 || nwf < nocc || (nwf > 0 && !coef)) return fail(c, FLOSIC_EINVAL, "bad orbital sizes");
    Spin &sp = c->sp[s];
    sp.nwf = nwf;
    sp.nocc = nocc;
    sp.coef.assign(coef, coef + (size_t)nbas * nwf);
    sp.have_coef = true;
    c->evaluated = false;
    c->fo_done = false;
    return FLOSIC_OK;
}

int32_t flosic_set_fods(flosic_ctx *c, int32_t ispn, int32_t nfrm, const double *bfrm) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    if (s == 1 && c->nspn == 1) return fail(c, FLOSIC_EINVAL, "NSPN = 1: spin 2 is not used");
    if (nfrm < 0 || (nfrm > 0 && !bfrm)) return fail(c, FLOSIC_EINVAL, "bad FOD arguments");
    for (int i = 0; i < 3 * nfrm; ++i)
        if (!std::isfinite(bfrm[i])) return fail(c, FLOSIC_EINVAL, "non-finite FOD coordinate");
    Spin &sp = c->sp[s];
    if (sp.have_coef && nfrm != sp.nocc)
        return fail(c, FLOSIC_EINVAL, "MORB .NE. NFRM: %d FODs for %d occupied orbitals (frmiorb.f:443)", nfrm, sp.nocc);
    sp.fod.assign(bfrm, bfrm + 3 * nfrm);
    sp.have_fod = true;
    c->evaluated = false;
    c->fo_done = false;
    return FLOSIC_OK;
}

int32_t flosic_frmorb(flosic_ctx *c) {
    if (!c) return FLOSIC_EINVAL;
    int r = check_ready(c);
    if (r) return r;
    return run_frmorb(c);
}

int32_t flosic_get_fo(flosic_ctx *c, int32_t ispn, double *fmat, double *over, double *eval, double *filo, double *tmat) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    if (!c->fo_done) {
        int r = flosic_frmorb(c);
        if (r) return r;
    }
    const Spin &sp = c->sp[s];
    if (fmat) std::copy(sp.fmat.begin(), sp.fmat.end(), fmat);
    if (over) std::copy(sp.over.begin(), sp.over.end(), over);
    if (eval) std::copy(sp.q.begin(), sp.q.end(), eval);
    if (filo) std::copy(sp.U.begin(), sp.U.end(), filo);
    if (tmat) std::copy(sp.tmat.begin(), sp.tmat.end(), tmat);
    return FLOSIC_OK;
}

int32_t flosic_evaluate(flosic_ctx *c, double *zsic) {
    if (!c) return FLOSIC_EINVAL;
    c->st = Status();
    const int r = evaluate(c);
    if (r) return r;
    if (zsic) *zsic = c->zsic;
    return FLOSIC_OK;
}

int32_t flosic_apotnl_sic(flosic_ctx *c, int32_t ispn, double *results, double *zsic) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    const int r = need_eval(c);
    if (r) return r;
    if (results) std::copy(c->sp[s].results.begin(), c->sp[s].results.end(), results);
    if (zsic) *zsic = c->zsic;
    return FLOSIC_OK;
}

int32_t flosic_apotnl_dft(flosic_ctx *c, double *pot, double *energies) {
    if (!c) return FLOSIC_EINVAL;
    if (c->nspn != 2) return fail(c, FLOSIC_EINVAL, "flosic_apotnl_dft: NSPN = 2 only");
    int r = need_eval(c);
    if (r) return r;
    const int64_t M = c->M;
    const int nblk = (int)((M + 255) / 256);
    ALLOC(c, c->d_dft, (size_t)2 * M);
    ALLOC(c, c->d_dftpart, (size_t)5 * nblk);
    // resident FLO fields: one pass; mesh-block mode: the planes and fields of each block are built again (blocks are
    // multiples of 256 points, so the 256-point partials are those of the resident pass)
    const bool blocked = c->block_pts > 0;
    const int64_t mb = mesh_block(c);
    const int nd = xc_planes(c), nF = c->nF;  // PBE: the FLO planes hold value, gradient and Hessian
    for (int64_t k0 = 0; k0 < M; k0 += mb) {
        const int64_t m = std::min(mb, M - k0);
        if (blocked) {
            if ((r = block_planes(c, k0, m, nd, c->d_planes.p))) return r;
            if ((r = orbitals_on_points(c, c->d_planes.p, m, nd, c->d_cflo.p, nF, c->d_phi.p))) return r;
        }
        double *part = c->d_dftpart.p + (size_t)5 * (k0 / 256);
        if (c->xc == FLOSIC_XC_PBE)
            launch_dft_gga(c->d_phi.p, m, c->sp[0].nocc, nF, c->d_VH.p + k0, M, c->d_w.p + k0, c->d_dft.p + k0, M, part,
                           c->stream);
        else
            launch_dft_lsda(c->d_phi.p, m, c->sp[0].nocc, nF, c->d_VH.p + k0, M, c->d_w.p + k0, c->d_dft.p + k0, M,
                            part, c->stream);
        HIPCHK(c, hipGetLastError());
    }
    std::vector<double> part((size_t)5 * nblk);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->d_dftpart.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                             c->stream));
    if (pot)
        HIPCHK(c, hipMemcpyAsync(pot, c->d_dft.p, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (energies) {  // block partials summed in block order (deterministic)
        for (int q = 0; q < 5; ++q) energies[q] = 0.0;
        for (int b = 0; b < nblk; ++b)
            for (int q = 0; q < 5; ++q) energies[q] += part[(size_t)5 * b + q];
    }
    return FLOSIC_OK;
}

int32_t flosic_apotnl_dft_noalloc(flosic_ctx *c, double *pot, double *energies) {
    if (!c) return FLOSIC_EINVAL;
    if (c->nspn != 2) return fail(c, FLOSIC_EINVAL, "flosic_apotnl_dft_noalloc: NSPN = 2 only");
    if (!c->evaluated)
        return fail(c, FLOSIC_ESTATE, "flosic_apotnl_dft_noalloc: no current evaluation (call flosic_evaluate first; "
                                      "this entry point runs no evaluation)");
    if (c->block_pts > 0)
        return fail(c, FLOSIC_EINVAL, "flosic_apotnl_dft_noalloc: needs resident FLO fields (mesh-block mode rebuilds "
                                      "them in allocated buffers)");
    const int64_t M = c->M;
    if (M > kNoAllocMaxPts)
        return fail(c, FLOSIC_EINVAL, "flosic_apotnl_dft_noalloc: %lld mesh points exceed the scratch limit of %lld",
                    (long long)M, (long long)kNoAllocMaxPts);
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, FLOSIC_EDEVICE, "hipSetDevice");
    double *base = scratch_on_device(c->device);
    if (!base) return fail(c, FLOSIC_EDEVICE, "flosic_apotnl_dft_noalloc: scratch address lookup failed");
    const int nblk = (int)((M + 255) / 256);
    double *d_pot = base, *d_part = base + 2 * M, *d_sum = d_part + (size_t)5 * nblk;
    // every block writes its 5 partials and every point its two potentials: nothing in the scratch needs clearing
    if (c->xc == FLOSIC_XC_PBE)
        launch_dft_gga(c->d_phi.p, M, c->sp[0].nocc, c->nF, c->d_VH.p, M, c->d_w.p, d_pot, M, d_part, c->stream);
    else
        launch_dft_lsda(c->d_phi.p, M, c->sp[0].nocc, c->nF, c->d_VH.p, M, c->d_w.p, d_pot, M, d_part, c->stream);
    HIPCHK(c, hipGetLastError());
    launch_sum_partials_ordered(d_part, nblk, 5, d_sum, c->stream);
    HIPCHK(c, hipGetLastError());
    if (energies) HIPCHK(c, hipMemcpyAsync(energies, d_sum, sizeof(double) * 5, hipMemcpyDeviceToHost, c->stream));
    if (pot) HIPCHK(c, hipMemcpyAsync(pot, d_pot, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FLOSIC_OK;
}

// SICLAG_DER: G_d = (d_d chi)^T U over the mesh for d = x, y, z with U = w V_i phi_i, then
// DERSIC(x, i, A) = -sum over atom A's functions s of c_si G_x(s, i), on the host.  With resident fields U is the
// evaluation's; PBE keeps the gradient planes resident too (one pass over the mesh), LSDA rebuilds them per block of
// points.  In mesh-block mode the planes, FLO fields and U of each block are rebuilt and dropped (the energy
// partials of that recomputation land in the evaluation's partial buffer, scratch once the evaluation has finished).
constexpr int64_t kDerBlockPts = 32768;  // points per block when the planes are resident (a multiple of 256)

int32_t flosic_siclag_der(flosic_ctx *c, double *dersic) {
    if (!c || !dersic) return FLOSIC_EINVAL;
    int r = need_eval(c);
    if (r) return r;
    const int N = c->nbas, nF = c->nF, nd = xc_planes(c), ng = std::max(nd, 4);
    const int64_t M = c->M;
    const bool blocked = c->block_pts > 0;
    ALLOC(c, c->d_dG, (size_t)3 * N * nF);
    if (!blocked && nd == 10) {  // resident PBE planes: value, gradient, Hessian of every function on the mesh
        for (int d = 0; d < 3; ++d)
            if ((r = gemm_tn_points(c, N, nF, c->d_planes.p + (size_t)(1 + d) * N * M, c->d_U.p, M,
                                    c->d_dG.p + (size_t)d * N * nF)))
                return r;
    } else {
        const int64_t mb = blocked ? mesh_block(c) : std::min<int64_t>(M, kDerBlockPts);
        const int nblk = (int)((M + 255) / 256);
        // mesh-block mode needs the value planes too (the FLO fields are rebuilt); resident LSDA only the gradients
        const int pk = blocked ? ng : -3, np = blocked ? ng : 3, g0 = blocked ? 1 : 0;
        ALLOC(c, c->d_dplanes, (size_t)np * N * mb);
        if (blocked) {
            ALLOC(c, c->d_dphi, (size_t)nd * nF * mb);
            ALLOC(c, c->d_dvt, (size_t)nF * mb);
            ALLOC(c, c->d_du, (size_t)nF * mb);
        }
        for (int64_t k0 = 0; k0 < M; k0 += mb) {
            const int64_t m = std::min(mb, M - k0);
            if ((r = block_planes(c, k0, m, pk, c->d_dplanes.p))) return r;
            const double *u = c->d_U.p + k0;  // resident U: [nF][M]
            int64_t ldu = M;
            if (blocked) {
                if ((r = orbitals_on_points(c, c->d_dplanes.p, m, nd, c->d_cflo.p, nF, c->d_dphi.p))) return r;
                launch_xc(c->xc, nd, c->d_dphi.p, m, nF, c->d_VH.p + k0, M, c->d_w.p + k0, c->d_dvt.p, c->d_du.p,
                          c->d_part.p, nblk, (int)(k0 / 256), c->stream);
                HIPCHK(c, hipGetLastError());
                u = c->d_du.p;
                ldu = m;
            }
            for (int d = 0; d < 3; ++d)
                if ((r = gemm_tn_points(c, N, nF, c->d_dplanes.p + (size_t)(g0 + d) * N * m, u, m,
                                        c->d_dG.p + (size_t)d * N * nF, k0 > 0, m, ldu)))
                    return r;
        }
    }
    std::vector<double> G((size_t)3 * N * nF), cf((size_t)N * nF);
    HIPCHK(c, hipMemcpyAsync(G.data(), c->d_dG.p, sizeof(double) * G.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cf.data(), c->d_cflo.p, sizeof(double) * cf.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double fac = c->nspn == 1 ? 2.0 : 1.0;  // siclag_der.F:158-166
    for (int a = 0; a < c->natom; ++a) {
        const int f0 = c->fstart[3 * a], f1 = a + 1 < c->natom ? c->fstart[3 * (a + 1)] : N;
        for (int i = 0; i < nF; ++i)
            for (int x = 0; x < 3; ++x) {
                const double *g = G.data() + (size_t)x * N * nF + (size_t)N * i;
                double acc = 0.0;
                for (int f = f0; f < f1; ++f) acc -= cf[f + (size_t)N * i] * g[f];
                dersic[x + 3 * ((size_t)i + (size_t)nF * a)] = fac * acc;
            }
    }
    return FLOSIC_OK;
}

int32_t flosic_siclagm(flosic_ctx *c, int32_t ispn, double *sic) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    const int r = need_eval(c);
    if (r) return r;
    if (sic) std::copy(c->sp[s].sic.begin(), c->sp[s].sic.end(), sic);
    return FLOSIC_OK;
}

int32_t flosic_sic_vab(flosic_ctx *c, int32_t ispn, double *vab) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    const int r = need_eval(c);
    if (r) return r;
    if (!c->vab_on_host) {
        size_t vo = 0;
        for (int t = 0; t < 2; ++t) {
            Spin &sp = c->sp[t];
            const size_t cnt = (size_t)sp.nocc * sp.nocc * sp.nocc;
            sp.vab.resize(cnt);
            if (cnt) HIPCHK(c, hipMemcpyAsync(sp.vab.data(), c->d_vab.p + vo, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                                              c->stream));
            vo += cnt;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->vab_on_host = true;
    }
    if (vab) std::copy(c->sp[s].vab.begin(), c->sp[s].vab.end(), vab);
    return FLOSIC_OK;
}

int32_t flosic_fod_forces(flosic_ctx *c, int32_t ispn, double *debdax) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    const int r = need_eval(c);
    if (r) return r;
    if (debdax) std::copy(c->sp[s].debdax.begin(), c->sp[s].debdax.end(), debdax);
    return FLOSIC_OK;
}

int32_t flosic_fod_gradient(flosic_ctx *c, int32_t ispn, double *grad) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    const int r = need_eval(c);
    if (r) return r;
    if (grad) std::copy(c->sp[s].grad.begin(), c->sp[s].grad.end(), grad);
    return FLOSIC_OK;
}

int32_t flosic_get_potential(flosic_ctx *c, int32_t ispn, int32_t iorb, double *pot, double *vh) {
    if (!c) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    int r = need_eval(c);
    if (r) return r;
    if (iorb < 1 || iorb > c->sp[s].nocc) return fail(c, FLOSIC_EINVAL, "orbital index %d out of range", iorb);
    const int f = (s == 0 ? 0 : c->sp[0].nocc) + iorb - 1;
    const int64_t M = c->M;
    if (vh) HIPCHK(c, hipMemcpy(vh, c->d_VH.p + (size_t)f * M, sizeof(double) * M, hipMemcpyDeviceToHost));
    if (!pot) return FLOSIC_OK;
    if (c->block_pts <= 0) {
        HIPCHK(c, hipMemcpy(pot, c->d_Vtot.p + (size_t)f * M, sizeof(double) * M, hipMemcpyDeviceToHost));
        return FLOSIC_OK;
    }
    // mesh-block mode keeps no V_tot: the blocks are evaluated again (energy partials land in scratch)
    const int nF = c->nF, nd = xc_planes(c);
    const int64_t mb = mesh_block(c);
    const int nblk = (int)((M + 255) / 256);
    for (int64_t k0 = 0; k0 < M; k0 += mb) {
        const int64_t m = std::min(mb, M - k0);
        if ((r = block_planes(c, k0, m, nd, c->d_planes.p))) return r;
        if ((r = orbitals_on_points(c, c->d_planes.p, m, nd, c->d_cflo.p, nF, c->d_phi.p))) return r;
        launch_xc(c->xc, nd, c->d_phi.p, m, nF, c->d_VH.p + k0, M, c->d_w.p + k0, c->d_Vtot.p, c->d_U.p, c->d_part.p,
                  nblk, (int)(k0 / 256), c->stream);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(pot + k0, c->d_Vtot.p + (size_t)f * m, sizeof(double) * m, hipMemcpyDeviceToHost,
                                 c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FLOSIC_OK;
}

int32_t flosic_lagmult_slv(flosic_ctx *c, int32_t ispn, const double *pot, int32_t ldsic, double *sic) {
    if (!c || !pot || !sic) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    int r = need_eval(c);
    if (r) return r;
    const int n = c->sp[s].nocc, N = c->nbas;
    const int64_t M = c->M;
    if (ldsic < n) return fail(c, FLOSIC_EINVAL, "ldsic %d < nfrm %d", ldsic, n);
    if (n == 0) return FLOSIC_OK;
    const int off = s == 0 ? 0 : c->sp[0].nocc;
    const double one = 1.0, zero = 0.0;
    const double *cf = c->d_cflo.p + (size_t)off * N;
    const bool blocked = c->block_pts > 0;
    const int64_t mb = mesh_block(c);
    const int nd = xc_planes(c);
    if ((r = upload(c, c->d_pot, pot, (size_t)M))) return r;
    ALLOC(c, c->d_lap, (size_t)n * mb);
    ALLOC(c, c->d_wop, (size_t)n * mb);
    ALLOC(c, c->d_small, (size_t)n * n);
    // Laplacians of the FLOs: from the Hessian planes (PBE) or from a Laplacian plane (kept resident unless in
    // mesh-block mode)
    if (nd != 10) {
        if (blocked) {
            ALLOC(c, c->d_lapplane, (size_t)N * mb);
        } else if (!c->lapplane_ok) {
            ALLOC(c, c->d_lapplane, (size_t)N * M);
            if ((r = block_planes(c, 0, M, -1, c->d_lapplane.p))) return r;
            c->lapplane_ok = true;
        }
    }
    for (int64_t k0 = 0; k0 < M; k0 += mb) {
        const int64_t m = std::min(mb, M - k0);
        const double *planes = c->d_planes.p;
        const double *phi = c->d_phi.p + (size_t)off * M;
        if (blocked) {
            if ((r = block_planes(c, k0, m, nd, c->d_planes.p))) return r;
            if ((r = orbitals_on_points(c, planes, m, 1, cf, n, c->d_phi.p))) return r;
            phi = c->d_phi.p;
        }
        if (nd == 10) {
            for (int d = 4; d < 7; ++d)
                BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, n, N,
                                         &one, planes + (size_t)d * N * m, (rocblas_int)m, cf, N,
                                         d == 4 ? &zero : &one, c->d_lap.p, (rocblas_int)m));
        } else {
            if (blocked && (r = block_planes(c, k0, m, -1, c->d_lapplane.p))) return r;
            BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, n, N, &one,
                                     c->d_lapplane.p, (rocblas_int)m, cf, N, &zero, c->d_lap.p, (rocblas_int)m));
        }
        // T_k = w (-1/2 lap phi_k + V phi_k), L = phi^T T
        launch_weighted_op(n, m, c->d_w.p + k0, c->d_pot.p + k0, phi, c->d_lap.p, -0.5, c->d_wop.p, c->stream);
        HIPCHK(c, hipGetLastError());
        if ((r = gemm_tn_points(c, n, n, phi, c->d_wop.p, m, c->d_small.p, k0 > 0))) return r;
    }
    std::vector<double> L((size_t)n * n);
    HIPCHK(c, hipMemcpyAsync(L.data(), c->d_small.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) sic[j + (size_t)ldsic * k] += L[j + (size_t)n * k];
    return FLOSIC_OK;
}

int32_t flosic_sic_hamiltonian(flosic_ctx *c, int32_t ispn, double *hao) {
    if (!c || !hao) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    int r = need_eval(c);
    if (r) return r;
    const int n = c->sp[s].nocc, N = c->nbas;
    const int64_t M = c->M;
    std::fill(hao, hao + (size_t)N * N, 0.0);
    if (n == 0) return FLOSIC_OK;
    const int off = s == 0 ? 0 : c->sp[0].nocc;
    const bool blocked = c->block_pts > 0;
    const int64_t mb = mesh_block(c);
    // B_ai = <chi_a|phi_i> and G_ai = <chi_a|V_i|phi_i> on the mesh (G from the evaluation)
    ALLOC(c, c->d_wop, (size_t)n * mb);
    ALLOC(c, c->d_B, (size_t)N * n);
    for (int64_t k0 = 0; k0 < M; k0 += mb) {
        const int64_t m = std::min(mb, M - k0);
        const double *phi = c->d_phi.p + (size_t)off * M;
        if (blocked) {
            if ((r = block_planes(c, k0, m, 1, c->d_planes.p))) return r;
            if ((r = orbitals_on_points(c, c->d_planes.p, m, 1, c->d_cflo.p + (size_t)off * N, n, c->d_phi.p))) return r;
            phi = c->d_phi.p;
        }
        launch_weighted_op(n, m, c->d_w.p + k0, nullptr, phi, nullptr, 0.0, c->d_wop.p, c->stream);
        HIPCHK(c, hipGetLastError());
        if ((r = gemm_tn_points(c, N, n, c->d_planes.p, c->d_wop.p, m, c->d_B.p, k0 > 0))) return r;
    }
    ALLOC(c, c->d_hao, (size_t)N * N);
    const double half = 0.5, zero = 0.0;
    BLASCHK(c, rocblas_dsyr2k(c->blas, rocblas_fill_upper, rocblas_operation_none, N, n, &half,
                              c->d_G.p + (size_t)off * N, N, c->d_B.p, N, &zero, c->d_hao.p, N));
    HIPCHK(c, hipMemcpyAsync(hao, c->d_hao.p, sizeof(double) * N * N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < N; ++b)
        for (int a = b + 1; a < N; ++a) hao[a + (size_t)N * b] = hao[b + (size_t)N * a];
    return FLOSIC_OK;
}

namespace {

// DIAGV_FO (diagv_fo.F:5-156) on the host.  h: HAM(J,I) at h[J + ldh*I] (m x m, not symmetric); on return phi holds
// PHIRES after its final transpose, phi[inew + m*iold].  The symmetric part is diagonalised first (DIAGGE with a unit
// overlap, diagv_fo.F:48-71: eigenvectors by ascending eigenvalue are the starting orbitals); then sweeps over the
// (occupied i <= n, virtual j > n) pairs rotate orbitals i and j by t = atan(2 h_ij / (h_ii - h_jj)) / 2 with
// h_ab = phi_a^T H phi_b (diagv_fo.F:74-112), until max |phi_occ^T H phi_virt| < 1e-6 after a sweep
// (diagv_fo.F:117-140), at most 600 sweeps.  The reference forms H phi_i and H phi_j anew for every pair; here
// G = H Phi is formed at the start of each sweep and its columns i, j rotate with those of Phi (H is linear, so
// this is the same arithmetic up to rounding at O(m) instead of O(m^2) per pair).  Returns the converged sweep, or
// 0 when 600 sweeps did not converge (the reference then warns and keeps the last orbitals, diagv_fo.F:148-151).
int diagv_fo_host(int m, int n, const double *h, int ldh, double *phi) {
    constexpr int kMaxSweeps = 600;
    constexpr double kConvg = 1.0e-6;
    std::vector<double> avg((size_t)m * m), q(m), P((size_t)m * m), G((size_t)m * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) avg[i + (size_t)m * j] = 0.5 * (h[i + (size_t)ldh * j] + h[j + (size_t)ldh * i]);
    eigh_sym(m, avg.data(), q.data(), P.data());  // P column k: orbital k in the old basis
    auto form_g = [&]() {  // G = H P
        std::fill(G.begin(), G.end(), 0.0);
        for (int k = 0; k < m; ++k)
            for (int l = 0; l < m; ++l) {
                const double plk = P[l + (size_t)m * k];
                const double *hl = h + (size_t)ldh * l;
                double *gk = &G[(size_t)m * k];
                for (int r = 0; r < m; ++r) gk[r] += hl[r] * plk;
            }
    };
    auto dot = [&](const double *a, const double *b) {
        double acc = 0.0;
        for (int r = 0; r < m; ++r) acc += a[r] * b[r];
        return acc;
    };
    int converged = 0;
    for (int sweep = 1; sweep <= kMaxSweeps && n > 0 && n < m; ++sweep) {
        form_g();
        for (int i = 0; i < n; ++i)
            for (int j = n; j < m; ++j) {
                double *pi = &P[(size_t)m * i], *pj = &P[(size_t)m * j];
                double *gi = &G[(size_t)m * i], *gj = &G[(size_t)m * j];
                const double hii = dot(pi, gi), hij = dot(pi, gj), hjj = dot(pj, gj);
                const double t = 0.5 * std::atan(2.0 * hij / (hii - hjj));
                const double ct = std::cos(t), st = std::sin(t);
                for (int r = 0; r < m; ++r) {
                    const double a = pi[r], b = pj[r];
                    pi[r] = ct * a + st * b;
                    pj[r] = -st * a + ct * b;
                    const double ga = gi[r], gb = gj[r];
                    gi[r] = ct * ga + st * gb;
                    gj[r] = -st * ga + ct * gb;
                }
            }
        form_g();
        double worst = 0.0;  // max |phi_i^T H phi_j|, i occupied, j virtual
        for (int j = n; j < m; ++j)
            for (int i = 0; i < n; ++i)
                worst = std::max(worst, std::fabs(dot(&P[(size_t)m * i], &G[(size_t)m * j])));
        if (worst < kConvg) {
            converged = sweep;
            break;
        }
    }
    if (n == 0 || n == m) converged = 1;  // nothing to rotate
    for (int k = 0; k < m; ++k)
        for (int r = 0; r < m; ++r) phi[k + (size_t)m * r] = P[r + (size_t)m * k];
    return converged;
}

}  // namespace

int32_t flosic_diagv_fo(int32_t m, int32_t n, const double *ham, int32_t ldh, double *phires, int32_t *sweeps) {
    if (m <= 0 || n < 0 || n > m || ldh < m || !ham || !phires) return FLOSIC_EINVAL;
    const int k = diagv_fo_host(m, n, ham, ldh, phires);
    if (sweeps) *sweeps = k;
    return FLOSIC_OK;
}

int32_t flosic_scf_sic_orbitals(flosic_ctx *c, int32_t ispn, const double *ham_ao, int32_t ldham, double *coef,
                                double *eval, int32_t *sweeps) {
    if (!c || !ham_ao || !coef) return FLOSIC_EINVAL;
    const int s = spin_index(c, ispn);
    if (s < 0) return FLOSIC_EINVAL;
    int r = need_eval(c);
    if (r) return r;
    const Spin &sp = c->sp[s];
    const int N = c->nbas, m = sp.nwf, n = sp.nocc;
    if (ldham < N) return fail(c, FLOSIC_EINVAL, "flosic_scf_sic_orbitals: ldham %d < nbas %d", ldham, N);
    if (m <= 0) return fail(c, FLOSIC_EINVAL, "flosic_scf_sic_orbitals: spin %d has no orbitals", ispn);
    // HAM = C^T ham_ao C over the orbitals of the evaluation (FLOs, then virtual orbitals: d_ctil)
    ALLOC(c, c->d_scf, (size_t)N * N + (size_t)N * m + (size_t)m * m);
    double *d_h = c->d_scf.p, *d_t = d_h + (size_t)N * N, *d_ham = d_t + (size_t)N * m;
    HIPCHK(c, hipMemcpy2DAsync(d_h, sizeof(double) * N, ham_ao, sizeof(double) * ldham, sizeof(double) * N, N,
                               hipMemcpyHostToDevice, c->stream));
    const double one = 1.0, zero = 0.0;
    BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, N, m, N, &one, d_h, N,
                             sp.d_ctil.p, N, &zero, d_t, N));
    BLASCHK(c, rocblas_dgemm(c->blas, rocblas_operation_transpose, rocblas_operation_none, m, m, N, &one,
                             sp.d_ctil.p, N, d_t, N, &zero, d_ham, m));
    std::vector<double> ham((size_t)m * m), cold((size_t)N * m), phi((size_t)m * m);
    HIPCHK(c, hipMemcpyAsync(ham.data(), d_ham, sizeof(double) * ham.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(cold.data(), sp.d_ctil.p, sizeof(double) * cold.size(), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // HAM(J,I) += SIC(I,J), J a FLO (scfsicw.f:389-395; SIC(I,J) = <psi_I|V_J|phi_J> from SICLAGM)
    for (int J = 0; J < n; ++J)
        for (int I = 0; I < m; ++I) ham[J + (size_t)m * I] += sp.sic[I + (size_t)m * J];
    const int k = diagv_fo_host(m, n, ham.data(), m, phi.data());
    if (sweeps) *sweeps = k;
    // EVAL(i) = sum_kl PHIRES(i,k) PHIRES(i,l) HAM(l,k) (scfsicw.f:584-597); C_new(:,i) = sum_k PHIRES(i,k) C(:,k)
    // (scfsicw.f:601-614)
    for (int i = 0; i < m; ++i) {
        if (eval) {
            double e = 0.0;
            for (int kk = 0; kk < m; ++kk) {
                double t = 0.0;
                for (int l = 0; l < m; ++l) t += phi[i + (size_t)m * l] * ham[l + (size_t)m * kk];
                e += phi[i + (size_t)m * kk] * t;
            }
            eval[i] = e;
        }
        double *ci = coef + (size_t)N * i;
        std::fill(ci, ci + N, 0.0);
        for (int kk = 0; kk < m; ++kk) {
            const double w = phi[i + (size_t)m * kk];
            const double *ck = &cold[(size_t)N * kk];
            for (int b = 0; b < N; ++b) ci[b] += w * ck[b];
        }
    }
    return FLOSIC_OK;
}

int32_t flosic_get_pairs(flosic_ctx *c, int32_t *npair, int32_t *list) {
    if (!c || !npair) return FLOSIC_EINVAL;
    *npair = c->npair;
    if (list) std::copy(c->pair_list.begin(), c->pair_list.end(), list);
    return FLOSIC_OK;
}

int32_t flosic_get_timings(flosic_ctx *c, int32_t n, double *ms) {
    if (!c || !ms || n < 0) return FLOSIC_EINVAL;
    for (int i = 0; i < n && i < 10; ++i) ms[i] = c->timings[i];
    return FLOSIC_OK;
}

int32_t flosic_get_counts(flosic_ctx *c, int32_t n, int64_t *counts) {
    if (!c || !counts || n < 0) return FLOSIC_EINVAL;
    if (n > 9) {
        double rows = 0.0;
        const bool far = coulomb_form(c) == COUL_BF16 && c->prec == FLOSIC_PREC_BF16X3 && c->far_ok &&
                         c->far_words > 0;
        if (far && hipMemcpy(&rows, c->d_farrows.p, sizeof rows, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(c, FLOSIC_EDEVICE, "flosic_get_counts: far-row count");
        c->counts[9] = (int64_t)rows;
    }
    for (int i = 0; i < n && i < 10; ++i) counts[i] = c->counts[i];
    return FLOSIC_OK;
}

int32_t flosic_batch_create(int32_t n, flosic_ctx *const *ctxs, int32_t use_graph, flosic_batch **out) {
    if (!out || n <= 0 || !ctxs) return FLOSIC_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i] || ctxs[i]->device != ctxs[0]->device) return FLOSIC_EINVAL;
        for (int j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return FLOSIC_EINVAL;
    }
    flosic_batch *b = new (std::nothrow) flosic_batch();
    if (!b) return FLOSIC_ENOMEM;
    b->device = ctxs[0]->device;
    b->ctxs.assign(ctxs, ctxs + n);
    b->use_graph = use_graph != 0;
    bool ok = hipSetDevice(b->device) == hipSuccess &&
              hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
    for (hipEvent_t *e : {&b->ev_fork, &b->ev_coul, &b->ev_end, &b->ev_t0, &b->ev_c0, &b->ev_c1})
        ok = ok && hipEventCreate(e) == hipSuccess;
    if (!ok) {
        flosic_batch_destroy(b);
        return FLOSIC_EDEVICE;
    }
    *out = b;
    return FLOSIC_OK;
}

int32_t flosic_batch_evaluate(flosic_batch *b, double *zsic) {
    if (!b) return FLOSIC_EINVAL;
    return batch_evaluate(b, zsic);
}

const char *flosic_batch_last_error(const flosic_batch *b) { return b ? b->st.msg.c_str() : "null batch"; }

int32_t flosic_batch_get_timings(flosic_batch *b, int32_t n, double *ms) {
    if (!b || !ms || n < 0) return FLOSIC_EINVAL;
    for (int i = 0; i < n && i < 4; ++i) ms[i] = b->timings[i];
    return FLOSIC_OK;
}

int32_t flosic_batch_get_counts(flosic_batch *b, int32_t n, int64_t *counts) {
    if (!b || !counts || n < 0) return FLOSIC_EINVAL;
    int64_t items = 0;
    for (const auto &k : b->classes) items += k.nitems;
    const int64_t v[5] = {b->evals, b->captures, b->replays, (int64_t)b->classes.size(), items};
    for (int i = 0; i < n && i < 5; ++i) counts[i] = v[i];
    return FLOSIC_OK;
}

int32_t flosic_batch_destroy(flosic_batch *b) {
    if (!b) return FLOSIC_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (hipEvent_t e : {b->ev_fork, b->ev_coul, b->ev_end, b->ev_t0, b->ev_c0, b->ev_c1})
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FLOSIC_OK;
}

}  // extern "C"

// scf.hip -- the in-process RHF driver: the RHF loop of scf.f90
// (src/scf/scf.f90:61-216,720-904) with the dens and RHFI2G program spawns
// replaced by device-resident D, G and the stored integrals, plus DIIS.
//
// Everything n x n lives on the device in the context's SCF buffers, which
// persist across calls of the same basis size.  Per iteration the host reads
// one small vector of scalars (energy terms, the new row of the DIIS matrix
// and the previous density change) produced by one fused reduction, and
// decides convergence from it; the density itself comes from the MFMA
// purification in sp2.hip.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "context.h"

namespace myqc {

int sp2_purify(myqc_ctx *ctx, const double *Fo, double *P, double *work, double *lohi, double *slots, double *tpart,
               int n, int nocc,
               int *ok, int *steps);
int gershgorin_bounds(myqc_ctx *ctx, const double *A, int n, double *lohi, double *slots);

static dim3 grid2(int n) { return dim3((n + 127) / 128, n); }

// --- elementwise kernels -----------------------------------------------------

// U(:, j) *= s_j^power
__global__ void k_scale_cols(double *__restrict__ U, const double *__restrict__ s, int n, double power) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    U[i + (size_t)n * j] *= pow(s[j], power);
}

__global__ void k_antisym(const double *__restrict__ T, double *__restrict__ E, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    E[i + (size_t)n * j] = T[i + (size_t)n * j] - T[j + (size_t)n * i];
}

// Z = a X + b Y + c I  (Y may be null when b = 0)
__global__ void k_axpbyc(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, double a,
                         double b, double c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const size_t e = i + (size_t)n * j;
    double v = (X ? a * X[e] : 0.0) + (Y ? b * Y[e] : 0.0);
    if (i == j) v += c;
    Z[e] = v;
}

struct LinComb {
    double c[8];
    int m;
};

// Fx = sum_i c_i hF_i (DIIS extrapolation over the history slots)
__global__ void k_lincomb(const double *__restrict__ hF, size_t nn, LinComb lc, double *__restrict__ Fx) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int i = 0; i < lc.m; ++i) v += lc.c[i] * hF[nn * i + e];
        Fx[e] = v;
    }
}

// --- fused reductions ---------------------------------------------------------
//
// Job d: mode 0 sum x*y, mode 1 max |x - y|, mode 2 max |x - I|.  Stage one
// writes one partial per (job, block); stage two folds them in a fixed order,
// so results are bitwise reproducible.
constexpr int kRedJobs = 24, kRedBlocks = 64;

struct RedJobs {
    const double *x[kRedJobs];
    const double *y[kRedJobs];
    int mode[kRedJobs];
    int m;
    int n;   // matrix order (mode 2)
};

__device__ __forceinline__ double wave_fold(double v, int mode) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = mode == 0 ? v + o : fmax(v, o);
    }
    return v;
}

__global__ __launch_bounds__(256) void k_reduce_part(RedJobs J, double *__restrict__ part) {
    __shared__ double s[4];
    const int d = blockIdx.y, mode = J.mode[d];
    const size_t nn = (size_t)J.n * J.n;
    const double *x = J.x[d], *y = J.y[d];
    double acc = 0.0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
        if (mode == 0)
            acc += x[e] * y[e];
        else if (mode == 1)
            acc = fmax(acc, fabs(x[e] - y[e]));
        else
            acc = fmax(acc, fabs(x[e] - ((e % J.n) == (e / J.n) ? 1.0 : 0.0)));
    }
    acc = wave_fold(acc, mode);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = s[0];
        for (int w = 1; w < 4; ++w) v = mode == 0 ? v + s[w] : fmax(v, s[w]);
        part[d * kRedBlocks + blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(64) void k_reduce_final(RedJobs J, const double *__restrict__ part,
                                                     double *__restrict__ out) {
    const int d = blockIdx.x;
    const double v = wave_fold(part[d * kRedBlocks + threadIdx.x], J.mode[d]);
    if (threadIdx.x == 0) out[d] = v;
}

// --- buffers -------------------------------------------------------------------

struct ScfBuffers {
    int n = -1;
    double *H = nullptr, *S = nullptr, *X = nullptr, *Xi = nullptr, *F = nullptr, *Fx = nullptr, *D = nullptr,
           *Dn = nullptr, *Do = nullptr, *Don = nullptr, *Dref = nullptr, *G = nullptr, *C = nullptr, *T1 = nullptr,
           *T2 = nullptr, *P2 = nullptr, *eig = nullptr, *work = nullptr, *dsc = nullptr, *lohi = nullptr,
           *part = nullptr, *red = nullptr, *sp2t = nullptr;
    double *hF = nullptr, *hE = nullptr;   // DIIS history, kDiis x n^2 each
    int *info = nullptr;
    static constexpr int kDiis = 8;
    void release() {
        double **ps[] = {&H,  &S,  &X,   &Xi,  &F,   &Fx,   &D,    &Dn,   &Do,   &Don, &Dref, &G, &C,
                         &T1, &T2, &P2, &eig, &work, &dsc, &lohi, &part, &red, &hF,  &hE, &sp2t};
        for (auto p : ps) dfree((void **)p);
        dfree((void **)&info);
        n = -1;
    }
    int ensure(myqc_ctx *ctx, int nb) {
        if (n == nb) return 0;
        release();
        const size_t nn = (size_t)nb * nb;
        double **mats[] = {&H, &S, &X, &Xi, &F, &Fx, &D, &Dn, &Do, &Don, &Dref, &G, &C, &T1, &T2, &P2};
        for (auto m : mats) MYQC_CHECK(dalloc(ctx, (void **)m, sizeof(double) * nn));
        MYQC_CHECK(dalloc(ctx, (void **)&eig, sizeof(double) * nb));
        MYQC_CHECK(dalloc(ctx, (void **)&work, sizeof(double) * nb));
        MYQC_CHECK(dalloc(ctx, (void **)&dsc, sizeof(double) * 16));
        MYQC_CHECK(dalloc(ctx, (void **)&lohi, sizeof(double) * 2 * nb));
        const size_t tiles = (size_t)(nb + 15) / 16;   // purification trace partials: 2 steps x tiles^2 x 2
        MYQC_CHECK(dalloc(ctx, (void **)&sp2t, sizeof(double) * 4 * tiles * tiles));
        MYQC_CHECK(dalloc(ctx, (void **)&part, sizeof(double) * kRedJobs * kRedBlocks));
        MYQC_CHECK(dalloc(ctx, (void **)&red, sizeof(double) * kRedJobs));
        MYQC_CHECK(dalloc(ctx, (void **)&hF, sizeof(double) * nn * kDiis));
        MYQC_CHECK(dalloc(ctx, (void **)&hE, sizeof(double) * nn * kDiis));
        MYQC_CHECK(dalloc(ctx, (void **)&info, sizeof(int)));
        n = nb;
        return 0;
    }
};

void free_scf(myqc_ctx *ctx) {
    for (ScfBuffers **p : {&ctx->scf, &ctx->scf_b}) {
        if (*p) {
            (*p)->release();
            delete *p;
            *p = nullptr;
        }
    }
}

// runs the jobs and copies their results to the host (one synchronisation)
static int reduce(myqc_ctx *ctx, ScfBuffers &b, RedJobs &J, double *out) {
    J.n = b.n;
    k_reduce_part<<<dim3(kRedBlocks, J.m), 256, 0, ctx->stream>>>(J, b.part);
    k_reduce_final<<<J.m, 64, 0, ctx->stream>>>(J, b.part, b.red);
    MYQC_HIP(ctx, hipGetLastError());
    MYQC_HIP(ctx, d2h(ctx, out, b.red, sizeof(double) * J.m));
    return 0;
}

static int add_job(myqc_ctx *ctx, RedJobs &J, int mode, const double *x, const double *y) {
    if (J.m >= kRedJobs) return set_error(ctx, MYQC_E_ARG, "too many reductions");
    J.x[J.m] = x;
    J.y[J.m] = y;
    J.mode[J.m] = mode;
    return J.m++;
}

static int gemm(myqc_ctx *ctx, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc) {
    // the first call of each shape sizes the workspace (no allocation inside rocBLAS)
    const std::array<int, 8> key{0, (int)ta, (int)tb, m, n, k, lda, ldb};
    if (!ctx->blas_shapes.count(key)) {
        MYQC_CHECK(blas_reserve(
            ctx, [&] { (void)rocblas_dgemm(ctx->blas, ta, tb, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc); }));
        ctx->blas_shapes.insert(key);
    }
    MYQC_BLAS(ctx, rocblas_dgemm(ctx->blas, ta, tb, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc));
    return 0;
}
static const rocblas_operation NN = rocblas_operation_none, TT = rocblas_operation_transpose;

// eigen-decomposition of the symmetric n x n A (overwritten by eigenvectors)
static int syevd(myqc_ctx *ctx, ScfBuffers &b, double *A, double *w, int n) {
    const std::array<int, 8> key{1, n, 0, 0, 0, 0, 0, 0};
    if (!ctx->blas_shapes.count(key)) {
        MYQC_CHECK(blas_reserve(ctx, [&] {
            (void)rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_lower, n, A, n, w, b.work, b.info);
        }));
        ctx->blas_shapes.insert(key);
    }
    MYQC_BLAS(ctx, rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_lower, n, A, n, w, b.work, b.info));
    int info = 0;
    MYQC_HIP(ctx, d2h(ctx, &info, b.info, sizeof(int)));
    if (info != 0) return set_error(ctx, MYQC_E_HIP, "rocsolver_dsyevd info=%d", info);
    return 0;
}

// orbitals: T2 = eigenvectors of Fo, eig, and C = X T2
static int diag_orth(myqc_ctx *ctx, ScfBuffers &b, const double *Fo, int n) {
    MYQC_HIP(ctx, hipMemcpyAsync(b.T2, Fo, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    MYQC_CHECK(syevd(ctx, b, b.T2, b.eig, n));
    MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, b.X, n, b.T2, n, 0.0, b.C, n));
    return 0;
}

// Fo = X^T F X
static int to_orth(myqc_ctx *ctx, ScfBuffers &b, const double *F, double *Fo, int n) {
    MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, F, n, b.X, n, 0.0, b.T1, n));
    MYQC_CHECK(gemm(ctx, TT, NN, n, n, n, 1.0, b.X, n, b.T1, n, 0.0, Fo, n));
    return 0;
}

// D = occ X Do X^T (occ = 2 for the RHF total density, 1 for one spin)
static int to_ao_density(myqc_ctx *ctx, ScfBuffers &b, const double *Do, double *D, int n, double occ = 2.0) {
    MYQC_CHECK(gemm(ctx, NN, TT, n, n, n, 1.0, Do, n, b.X, n, 0.0, b.T1, n));
    MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, occ, b.X, n, b.T1, n, 0.0, D, n));
    return 0;
}

// X = S^{-1/2}, Xi = S^{1/2} by the coupled Newton-Schulz iteration
// (Higham, Numer. Algorithms 15, 227 (1997)): with A = S/c, c >= lambda_max,
//     T = (3 I - Z Y)/2,  Y <- Y T,  Z <- T Z,   Y -> A^{1/2}, Z -> A^{-1/2};
// matrix products only.  Returns *ok = 0 when it has not converged in its
// step budget (an ill-conditioned S), for the eigen-decomposition path.
static int inv_sqrt_ns(myqc_ctx *ctx, ScfBuffers &b, int n, int *ok) {
    *ok = 0;
    MYQC_CHECK(gershgorin_bounds(ctx, b.S, n, b.lohi, b.dsc));
    double bnd[2];
    MYQC_HIP(ctx, d2h(ctx, bnd, b.dsc + 6, sizeof bnd));
    const double c = bnd[1];
    if (!(c > 0.0)) return 0;
    double *Y = b.Xi, *Z = b.X, *Yn = b.T2, *Zn = b.P2;
    k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(b.S, nullptr, Y, 1.0 / c, 0.0, 0.0, n);
    k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(nullptr, nullptr, Z, 0.0, 0.0, 1.0, n);
    int extra = -1;
    for (int k = 0; k < 40; ++k) {
        MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, Z, n, Y, n, 0.0, b.T1, n));
        if (extra < 0 && k >= 6 && (k & 1) == 0) {
            RedJobs J{};
            add_job(ctx, J, 2, b.T1, nullptr);
            double r = 0.0;
            MYQC_CHECK(reduce(ctx, b, J, &r));
            // quadratic convergence: one more step takes 1e-9 below round-off
            if (r < 1e-9) extra = 1;
        }
        if (extra == 0) {
            *ok = 1;
            break;
        }
        k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(b.T1, nullptr, b.T1, -0.5, 0.0, 1.5, n);
        MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, Y, n, b.T1, n, 0.0, Yn, n));
        MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, b.T1, n, Z, n, 0.0, Zn, n));
        std::swap(Y, Yn);
        std::swap(Z, Zn);
        if (extra > 0) --extra;
    }
    if (!*ok) return 0;
    const double rc = sqrt(c);
    k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(Z, nullptr, b.X, 1.0 / rc, 0.0, 0.0, n);
    k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(Y, nullptr, b.Xi, rc, 0.0, 0.0, n);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// X = S^{-1/2}, Xi = S^{1/2} through the eigen-decomposition of S; rejects a
// numerically singular S
static int inv_sqrt_eig(myqc_ctx *ctx, ScfBuffers &b, int n) {
    MYQC_HIP(ctx, hipMemcpyAsync(b.T2, b.S, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    MYQC_CHECK(syevd(ctx, b, b.T2, b.eig, n));
    double s0 = 0.0;
    MYQC_HIP(ctx, d2h(ctx, &s0, b.eig, sizeof(double)));
    if (s0 <= 1e-10) return set_error(ctx, MYQC_E_ARG, "rhf: overlap matrix singular (min eigenvalue %g)", s0);
    MYQC_HIP(ctx, hipMemcpyAsync(b.T1, b.T2, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    k_scale_cols<<<grid2(n), 128, 0, ctx->stream>>>(b.T1, b.eig, n, -0.5);
    MYQC_CHECK(gemm(ctx, NN, TT, n, n, n, 1.0, b.T1, n, b.T2, n, 0.0, b.X, n));
    MYQC_HIP(ctx, hipMemcpyAsync(b.T1, b.T2, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    k_scale_cols<<<grid2(n), 128, 0, ctx->stream>>>(b.T1, b.eig, n, 0.5);
    MYQC_CHECK(gemm(ctx, NN, TT, n, n, n, 1.0, b.T1, n, b.T2, n, 0.0, b.Xi, n));
    return 0;
}

// Orthonormal-basis density projector onto the nocc lowest levels of Fo by
// purification; falls back to diagonalisation if it does not converge.
static int purify(myqc_ctx *ctx, ScfBuffers &b, const double *Fo, double *P, int n, int nocc) {
    if (nocc == 0 || nocc == n) {
        // empty or full: P = 0 or I (a one-level spectrum gives purification
        // no interval to map)
        k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(nullptr, nullptr, P, 0.0, 0.0, nocc == n ? 1.0 : 0.0, n);
        return 0;
    }
    int ok = 0;
    MYQC_CHECK(sp2_purify(ctx, Fo, P, b.P2, b.lohi, b.dsc, b.sp2t, n, nocc, &ok, &ctx->scf_sp2_steps));
    ++ctx->scf_purify;
    if (!ok) {
        // no clean gap at this step: occupy the nocc lowest eigenvectors
        ++ctx->scf_diag;
        MYQC_CHECK(diag_orth(ctx, b, Fo, n));
        MYQC_CHECK(gemm(ctx, NN, TT, n, n, nocc, 1.0, b.T2, n, b.T2, n, 0.0, P, n));
    }
    return 0;
}

// D = occ C_occ C_occ^T: occ = 2 for RHF (src/dens/dens.f90:115-124), 1 for
// one UHF spin (densUHF, :213-228)
static int density(myqc_ctx *ctx, const double *C, int nocc, double *D, int n, double occ = 2.0) {
    if (nocc == 0) {
        k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(nullptr, nullptr, D, 0.0, 0.0, 0.0, n);
        return 0;
    }
    return gemm(ctx, NN, TT, n, n, nocc, occ, C, n, C, n, 0.0, D, n);
}

// solve the DIIS system on the host (m <= 8), Gaussian elimination with pivoting
static bool diis_coeffs(const std::vector<double> &B, int m, std::vector<double> &c) {
    const int N = m + 1;
    std::vector<double> M((size_t)N * (N + 1), 0.0);
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) M[i * (N + 1) + j] = B[i * m + j];
        M[i * (N + 1) + m] = -1.0;
        M[i * (N + 1) + N] = 0.0;
    }
    for (int j = 0; j < m; ++j) M[m * (N + 1) + j] = -1.0;
    M[m * (N + 1) + m] = 0.0;
    M[m * (N + 1) + N] = -1.0;
    for (int col = 0; col < N; ++col) {
        int piv = col;
        for (int r = col + 1; r < N; ++r)
            if (fabs(M[r * (N + 1) + col]) > fabs(M[piv * (N + 1) + col])) piv = r;
        if (fabs(M[piv * (N + 1) + col]) < 1e-300) return false;
        if (piv != col)
            for (int k = 0; k <= N; ++k) std::swap(M[col * (N + 1) + k], M[piv * (N + 1) + k]);
        for (int r = 0; r < N; ++r) {
            if (r == col) continue;
            const double f = M[r * (N + 1) + col] / M[col * (N + 1) + col];
            for (int k = col; k <= N; ++k) M[r * (N + 1) + k] -= f * M[col * (N + 1) + k];
        }
    }
    c.resize(m);
    for (int i = 0; i < m; ++i) c[i] = M[i * (N + 1) + N] / M[i * (N + 1) + i];
    return true;
}

// The loop.  Iteration k builds F[D_k] and E_k = (D_k.F_k + D_k.H)/2 + E_nuc,
// extrapolates the orthonormal-basis Fock matrix by DIIS and purifies it into
// D_{k+1}.  Iteration k has converged when k > 1, |E_k - E_{k-1}| < e_tol,
// max|D_{k+1} - D_k| < d_tol and the DIIS error norm is below 100 d_tol; the
// density change is read with iteration k+1's scalars, whose F[D_{k+1}] and
// E_{k+1} are then the reported plain Fock matrix and energy (the reference
// prints E of the density it diagonalises, scf.f90:775-848).  With
// precision FP32, F_k = F_{k-1} + G32(D_k - D_{k-1}) between FP64 builds every
// 8 iterations, and convergence is only accepted on two FP64 builds in a row.
// A run started from a given density or orbitals (a guess that is already
// near the solution but not a Roothaan solution) level-shifts the virtual
// space, F' + sigma (I - D'), while the DIIS error is above kShiftErr: that
// keeps DIIS from wandering off the guess in the first iterations.  From the
// core-Hamiltonian guess the shift slows convergence, so it is not used there.
constexpr double kShift = 0.5, kShiftErr = 1.0;

static int rhf_scf(myqc_ctx *ctx, const double *H, const double *S, int nocc, double e_nuc, const double *C_guess,
                   const double *D_guess, double e_tol, double d_tol, int max_iter, int precision, double *energy,
                   double *C_out, double *eig_out, int *iters_out) {
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    const int MD = ScfBuffers::kDiis;
    if (!ctx->scf) ctx->scf = new ScfBuffers();
    ScfBuffers &b = *ctx->scf;
    MYQC_CHECK(b.ensure(ctx, n));
    MYQC_HIP(ctx, hipMemcpyAsync(b.H, H, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(b.S, S, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    int ok = 0;
    MYQC_CHECK(inv_sqrt_ns(ctx, b, n, &ok));
    if (!ok) MYQC_CHECK(inv_sqrt_eig(ctx, b, n));
    // guess density
    if (C_guess || D_guess) {
        if (C_guess) {
            MYQC_HIP(ctx, hipMemcpyAsync(b.C, C_guess, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
            MYQC_CHECK(density(ctx, b.C, nocc, b.D, n));
        } else {
            MYQC_HIP(ctx, hipMemcpyAsync(b.D, D_guess, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
        }
        // Do = S^{1/2} D S^{1/2} / 2
        MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, b.D, n, b.Xi, n, 0.0, b.T1, n));
        MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 0.5, b.Xi, n, b.T1, n, 0.0, b.Do, n));
    } else {
        // core Hamiltonian guess (scf.f90:410-512)
        MYQC_CHECK(to_orth(ctx, b, b.H, b.F, n));
        MYQC_CHECK(purify(ctx, b, b.F, b.Do, n, nocc));
        MYQC_CHECK(to_ao_density(ctx, b, b.Do, b.D, n));
    }
    std::vector<double> Bm((size_t)MD * MD, 0.0);
    int hist = 0, head = 0;
    double E = 0.0, E_old = 0.0;
    // convergence data of the previous iteration
    double dE_prev = 1e300, err_prev = 1e300;
    bool prev64 = false;
    const bool incr = (precision == MYQC_FP32);
    int since_full = 0;
    bool final_fp64 = false, conv = false;
    double *fslot = nullptr;
    int it = 1;
    // per-iteration convergence data to stderr when MYQC_SCF_TRACE is set
    const bool trace = getenv("MYQC_SCF_TRACE") != nullptr;
    while (true) {
        const bool full = !incr || final_fp64 || since_full == 0 || since_full >= 8;
        if (full) {
            MYQC_CHECK(digest_G(ctx, b.D, b.G, MYQC_FP64));
            since_full = 1;
        } else {
            k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(b.D, b.Dref, b.T1, 1.0, -1.0, 0.0, n);
            MYQC_CHECK(digest_G(ctx, b.T1, b.G, MYQC_FP32));   // G += G32(dD)
            ++since_full;
        }
        if (incr) MYQC_HIP(ctx, hipMemcpyAsync(b.Dref, b.D, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
        k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(b.H, b.G, b.F, 1.0, 1.0, 0.0, n);
        // orthonormal basis: Fo = X^T F X, DIIS error Fo Do - Do Fo
        fslot = b.hF + nn * head;
        double *eslot = b.hE + nn * head;
        MYQC_CHECK(to_orth(ctx, b, b.F, fslot, n));
        MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, fslot, n, b.Do, n, 0.0, b.T2, n));
        k_antisym<<<grid2(n), 128, 0, ctx->stream>>>(b.T2, eslot, n);
        const int hn = std::min(hist + 1, MD);   // history after this step, slot `head` included
        RedJobs J{};
        add_job(ctx, J, 0, b.D, b.F);
        add_job(ctx, J, 0, b.D, b.H);
        const int jdd = it > 1 ? add_job(ctx, J, 1, b.D, b.Dn) : -1;   // Dn holds D_{k-1}
        const int jb0 = J.m;
        for (int j = 0; j < hn; ++j) add_job(ctx, J, 0, eslot, b.hE + nn * j);
        double r[kRedJobs];
        MYQC_CHECK(reduce(ctx, b, J, r));
        E = 0.5 * (r[0] + r[1]) + e_nuc;
        if (trace)
            fprintf(stderr, "[scf] it %2d %s E %.12f dE %.2e max|dD| %.2e |err| %.2e hist %d\n", it,
                    full ? "fp64" : "fp32", E, fabs(E - E_old), jdd >= 0 ? r[jdd] : -1.0, sqrt(fabs(r[jb0 + head])), hn);
        if (jdd >= 0 && it > 2 && dE_prev < e_tol && r[jdd] < d_tol && err_prev < 100 * d_tol) {
            if (!incr || (prev64 && full)) {
                conv = true;
                break;
            }
            if (!final_fp64) {
                // confirm with FP64 builds, starting with this iteration's
                final_fp64 = true;
                if (!full) continue;
            }
        }
        if (it > max_iter) break;
        for (int j = 0; j < hn; ++j) {
            Bm[head * MD + j] = r[jb0 + j];
            Bm[j * MD + head] = r[jb0 + j];
        }
        hist = hn;
        const double *Fuse = fslot;
        if (hist >= 2) {
            std::vector<double> Bs((size_t)hist * hist), c;
            for (int i = 0; i < hist; ++i)
                for (int j = 0; j < hist; ++j) Bs[i * hist + j] = Bm[i * MD + j];
            if (diis_coeffs(Bs, hist, c)) {
                LinComb lc{};
                lc.m = hist;
                for (int i = 0; i < hist; ++i) lc.c[i] = c[i];
                k_lincomb<<<256, 256, 0, ctx->stream>>>(b.hF, nn, lc, b.Fx);
                Fuse = b.Fx;
            }
        }
        err_prev = sqrt(fabs(Bm[head * MD + head]));
        dE_prev = fabs(E - E_old);
        E_old = E;
        prev64 = full;
        head = (head + 1) % MD;
        if ((C_guess || D_guess) && err_prev > kShiftErr) {
            k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(Fuse, b.Do, b.T1, 1.0, -kShift, kShift, n);
            Fuse = b.T1;
        }
        MYQC_CHECK(purify(ctx, b, Fuse, b.Don, n, nocc));
        std::swap(b.D, b.Dn);   // Dn keeps D_k for the change
        std::swap(b.Do, b.Don);
        MYQC_CHECK(to_ao_density(ctx, b, b.Do, b.D, n));
        ++it;
    }
    // orbitals of the plain Fock matrix of the reported density (fslot holds
    // its orthonormal-basis form)
    MYQC_CHECK(diag_orth(ctx, b, fslot, n));
    if (energy) *energy = E;
    if (C_out) MYQC_HIP(ctx, d2h(ctx, C_out, b.C, sizeof(double) * nn));
    if (eig_out) MYQC_HIP(ctx, d2h(ctx, eig_out, b.eig, sizeof(double) * n));
    if (iters_out) *iters_out = it - 1;
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!conv) return set_error(ctx, MYQC_E_NOCONV, "RHF did not converge in %d iterations", max_iter);
    return 0;
}

// UHF (src/scf/scf.f90:221-404 UHF, :518-622 initUHF, :909-1117 UHFiter).
// Per spin s: F_s = H + G_s[Da, Db] (the UHFI2G build), the density D_s =
// C_s,occ C_s,occ^T of the nocc_s lowest levels (densUHF), and
// E = (Da.(Fa + H) + Db.(Fb + H))/2 + E_nuc (UHFiter, :969-975).  The loop
// is the RHF one run on both spins together: one DIIS extrapolation with
// shared coefficients over B_ij = <ea_i, ea_j> + <eb_i, eb_j>, and
// convergence on |dE|, max|dD| over both spins and the DIIS error.  The alpha
// spin uses ctx->scf (which also holds H, S, X), the beta spin ctx->scf_b.
static int uhf_scf(myqc_ctx *ctx, const double *H, const double *S, int na, int nb_, double e_nuc,
                   const double *Ca_guess, const double *Cb_guess, double e_tol, double d_tol, int max_iter,
                   double *energy, double *Ca_out, double *Cb_out, double *eiga_out, double *eigb_out,
                   int *iters_out) {
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    const int MD = ScfBuffers::kDiis;
    if (!ctx->scf) ctx->scf = new ScfBuffers();
    if (!ctx->scf_b) ctx->scf_b = new ScfBuffers();
    ScfBuffers &A = *ctx->scf, &B = *ctx->scf_b;
    ScfBuffers *sp[2] = {&A, &B};
    const int nocc[2] = {na, nb_};
    MYQC_CHECK(A.ensure(ctx, n));
    MYQC_CHECK(B.ensure(ctx, n));
    MYQC_HIP(ctx, hipMemcpyAsync(A.H, H, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(A.S, S, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    int ok = 0;
    MYQC_CHECK(inv_sqrt_ns(ctx, A, n, &ok));
    if (!ok) MYQC_CHECK(inv_sqrt_eig(ctx, A, n));
    MYQC_HIP(ctx, hipMemcpyAsync(B.X, A.X, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(B.Xi, A.Xi, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
    const bool guessed = Ca_guess != nullptr;
    if (guessed) {
        // orbitals from a Cui restart file (scf.f90:314-320)
        const double *Cg[2] = {Ca_guess, Cb_guess};
        for (int s = 0; s < 2; ++s) {
            ScfBuffers &b = *sp[s];
            MYQC_HIP(ctx, hipMemcpyAsync(b.C, Cg[s], sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
            MYQC_CHECK(density(ctx, b.C, nocc[s], b.D, n, 1.0));
            // Do = S^{1/2} D S^{1/2}
            MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, b.D, n, b.Xi, n, 0.0, b.T1, n));
            MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, b.Xi, n, b.T1, n, 0.0, b.Do, n));
        }
    } else {
        // both spins from the core Hamiltonian's orbitals (initUHF, scf.f90:518-622)
        MYQC_CHECK(to_orth(ctx, A, A.H, A.F, n));
        for (int s = 0; s < 2; ++s) {
            MYQC_CHECK(purify(ctx, *sp[s], A.F, sp[s]->Do, n, nocc[s]));
            MYQC_CHECK(to_ao_density(ctx, *sp[s], sp[s]->Do, sp[s]->D, n, 1.0));
        }
    }
    std::vector<double> Bm((size_t)MD * MD, 0.0);
    int hist = 0, head = 0;
    double E = 0.0, E_old = 0.0, dE_prev = 1e300, err_prev = 1e300;
    bool conv = false;
    double *fslot[2] = {nullptr, nullptr};
    int it = 1;
    const bool trace = getenv("MYQC_SCF_TRACE") != nullptr;
    while (true) {
        MYQC_CHECK(digest_G_uhf(ctx, A.D, B.D, A.G, B.G));
        const int hn = std::min(hist + 1, MD);
        RedJobs J{};
        int jdd[2] = {-1, -1}, jb0[2] = {0, 0}, jf[2], jh[2];
        for (int s = 0; s < 2; ++s) {
            ScfBuffers &b = *sp[s];
            k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(A.H, b.G, b.F, 1.0, 1.0, 0.0, n);
            fslot[s] = b.hF + nn * head;
            double *eslot = b.hE + nn * head;
            MYQC_CHECK(to_orth(ctx, b, b.F, fslot[s], n));
            MYQC_CHECK(gemm(ctx, NN, NN, n, n, n, 1.0, fslot[s], n, b.Do, n, 0.0, b.T2, n));
            k_antisym<<<grid2(n), 128, 0, ctx->stream>>>(b.T2, eslot, n);
            jf[s] = add_job(ctx, J, 0, b.D, b.F);
            jh[s] = add_job(ctx, J, 0, b.D, A.H);
            if (it > 1) jdd[s] = add_job(ctx, J, 1, b.D, b.Dn);   // Dn holds D_{k-1}
        }
        for (int s = 0; s < 2; ++s) {
            jb0[s] = J.m;
            for (int j = 0; j < hn; ++j) add_job(ctx, J, 0, sp[s]->hE + nn * head, sp[s]->hE + nn * j);
        }
        double r[kRedJobs];
        MYQC_CHECK(reduce(ctx, A, J, r));
        E = 0.5 * (r[jf[0]] + r[jh[0]] + r[jf[1]] + r[jh[1]]) + e_nuc;
        const double dmax = it > 1 ? std::max(r[jdd[0]], r[jdd[1]]) : 1e300;
        if (trace)
            fprintf(stderr, "[uhf] it %2d E %.12f dE %.2e max|dD| %.2e |err| %.2e hist %d\n", it, E, fabs(E - E_old),
                    it > 1 ? dmax : -1.0, sqrt(fabs(r[jb0[0] + head] + r[jb0[1] + head])), hn);
        if (it > 2 && dE_prev < e_tol && dmax < d_tol && err_prev < 100 * d_tol) {
            conv = true;
            break;
        }
        if (it > max_iter) break;
        for (int j = 0; j < hn; ++j) {
            const double v = r[jb0[0] + j] + r[jb0[1] + j];
            Bm[head * MD + j] = v;
            Bm[j * MD + head] = v;
        }
        hist = hn;
        const double *Fuse[2] = {fslot[0], fslot[1]};
        if (hist >= 2) {
            std::vector<double> Bs((size_t)hist * hist), c;
            for (int i = 0; i < hist; ++i)
                for (int j = 0; j < hist; ++j) Bs[i * hist + j] = Bm[i * MD + j];
            if (diis_coeffs(Bs, hist, c)) {
                LinComb lc{};
                lc.m = hist;
                for (int i = 0; i < hist; ++i) lc.c[i] = c[i];
                for (int s = 0; s < 2; ++s) {
                    k_lincomb<<<256, 256, 0, ctx->stream>>>(sp[s]->hF, nn, lc, sp[s]->Fx);
                    Fuse[s] = sp[s]->Fx;
                }
            }
        }
        err_prev = sqrt(fabs(Bm[head * MD + head]));
        dE_prev = fabs(E - E_old);
        E_old = E;
        head = (head + 1) % MD;
        for (int s = 0; s < 2; ++s) {
            ScfBuffers &b = *sp[s];
            const double *Fs = Fuse[s];
            if (guessed && err_prev > kShiftErr) {
                k_axpbyc<<<grid2(n), 128, 0, ctx->stream>>>(Fs, b.Do, b.T1, 1.0, -kShift, kShift, n);
                Fs = b.T1;
            }
            MYQC_CHECK(purify(ctx, b, Fs, b.Don, n, nocc[s]));
            std::swap(b.D, b.Dn);
            std::swap(b.Do, b.Don);
            MYQC_CHECK(to_ao_density(ctx, b, b.Do, b.D, n, 1.0));
        }
        ++it;
    }
    // orbitals of each spin's plain Fock matrix of the reported densities
    for (int s = 0; s < 2; ++s) MYQC_CHECK(diag_orth(ctx, *sp[s], fslot[s], n));
    if (energy) *energy = E;
    if (Ca_out) MYQC_HIP(ctx, d2h(ctx, Ca_out, A.C, sizeof(double) * nn));
    if (Cb_out) MYQC_HIP(ctx, d2h(ctx, Cb_out, B.C, sizeof(double) * nn));
    if (eiga_out) MYQC_HIP(ctx, d2h(ctx, eiga_out, A.eig, sizeof(double) * n));
    if (eigb_out) MYQC_HIP(ctx, d2h(ctx, eigb_out, B.eig, sizeof(double) * n));
    if (iters_out) *iters_out = it - 1;
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (!conv) return set_error(ctx, MYQC_E_NOCONV, "UHF did not converge in %d iterations", max_iter);
    return 0;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_density_rhf(myqc_ctx *ctx, const double *C, int nocc, double *D) {
    if (!ctx || !C || !D) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_density_rhf: no system set");
    const int n = ctx->n;
    if (nocc < 0 || nocc > n) return set_error(ctx, MYQC_E_ARG, "myqc_density_rhf: nocc=%d", nocc);
    const size_t nn = (size_t)n * n;
    double *dC = nullptr, *dD = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&dC, sizeof(double) * nn));
    int rc = dalloc(ctx, (void **)&dD, sizeof(double) * nn);
    if (rc == 0 && hipMemcpyAsync(dC, C, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "upload failed");
    if (rc == 0) rc = density(ctx, dC, nocc, dD, n);
    if (rc == 0 && d2h(ctx, D, dD, sizeof(double) * nn) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "download failed");
    dfree((void **)&dC);
    dfree((void **)&dD);
    return rc;
}

extern "C" int myqc_scf_info(myqc_ctx *ctx, int *purifications, int *diag_fallbacks) {
    if (!ctx) return MYQC_E_ARG;
    if (purifications) *purifications = ctx->scf_purify;
    if (diag_fallbacks) *diag_fallbacks = ctx->scf_diag;
    return 0;
}

extern "C" int myqc_rhf_scf(myqc_ctx *ctx, const double *H, const double *S, int nocc, double e_nuc,
                            const double *C_guess, double e_tol, double d_tol, int max_iter, int precision,
                            double *energy, double *C, double *eig, int *iterations) {
    if (!ctx || !H || !S) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_rhf_scf: no system set");
    if (nocc < 0 || nocc > ctx->n || max_iter < 1 || e_tol <= 0 || d_tol <= 0)
        return set_error(ctx, MYQC_E_ARG, "myqc_rhf_scf: invalid arguments");
    if (precision != MYQC_FP64 && precision != MYQC_FP32)
        return set_error(ctx, MYQC_E_ARG, "myqc_rhf_scf: precision %d", precision);
    if (!ctx->have_eri || (precision == MYQC_FP32 && !ctx->have_eri32))
        return set_error(ctx, MYQC_E_STATE, "myqc_rhf_scf: call myqc_eri_build first");
    ctx->scf_purify = ctx->scf_diag = ctx->scf_sp2_steps = 0;
    return rhf_scf(ctx, H, S, nocc, e_nuc, C_guess, nullptr, e_tol, d_tol, max_iter, precision, energy, C, eig,
                   iterations);
}

extern "C" int myqc_rhf_scf_density(myqc_ctx *ctx, const double *H, const double *S, int nocc, double e_nuc,
                                    const double *D_guess, double e_tol, double d_tol, int max_iter, int precision,
                                    double *energy, double *C, double *eig, int *iterations) {
    if (!ctx || !H || !S || !D_guess) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_rhf_scf_density: no system set");
    if (nocc < 0 || nocc > ctx->n || max_iter < 1 || e_tol <= 0 || d_tol <= 0)
        return set_error(ctx, MYQC_E_ARG, "myqc_rhf_scf_density: invalid arguments");
    if (precision != MYQC_FP64 && precision != MYQC_FP32)
        return set_error(ctx, MYQC_E_ARG, "myqc_rhf_scf_density: precision %d", precision);
    if (!ctx->have_eri || (precision == MYQC_FP32 && !ctx->have_eri32))
        return set_error(ctx, MYQC_E_STATE, "myqc_rhf_scf_density: call myqc_eri_build first");
    ctx->scf_purify = ctx->scf_diag = ctx->scf_sp2_steps = 0;
    return rhf_scf(ctx, H, S, nocc, e_nuc, nullptr, D_guess, e_tol, d_tol, max_iter, precision, energy, C, eig,
                   iterations);
}

extern "C" int myqc_density_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int nalpha, int nbeta, double *Da,
                                double *Db) {
    if (!ctx || !Ca || !Cb || !Da || !Db) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_density_uhf: no system set");
    const int n = ctx->n;
    if (nalpha < 0 || nalpha > n || nbeta < 0 || nbeta > n)
        return set_error(ctx, MYQC_E_ARG, "myqc_density_uhf: nalpha=%d nbeta=%d", nalpha, nbeta);
    const size_t nn = (size_t)n * n;
    double *buf = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&buf, sizeof(double) * 2 * nn));
    const double *Cs[2] = {Ca, Cb};
    double *Ds[2] = {Da, Db};
    const int occ[2] = {nalpha, nbeta};
    int rc = 0;
    for (int s = 0; s < 2 && rc == 0; ++s) {
        if (hipMemcpyAsync(buf, Cs[s], sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "upload failed");
        if (rc == 0) rc = density(ctx, buf, occ[s], buf + nn, n, 1.0);
        if (rc == 0 && d2h(ctx, Ds[s], buf + nn, sizeof(double) * nn) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "download failed");
    }
    dfree((void **)&buf);
    return rc;
}

extern "C" int myqc_uhf_scf(myqc_ctx *ctx, const double *H, const double *S, int nalpha, int nbeta, double e_nuc,
                            const double *Ca_guess, const double *Cb_guess, double e_tol, double d_tol, int max_iter,
                            double *energy, double *Ca, double *Cb, double *eig_a, double *eig_b, int *iterations) {
    if (!ctx || !H || !S) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_uhf_scf: no system set");
    if (nalpha < 0 || nalpha > ctx->n || nbeta < 0 || nbeta > ctx->n || max_iter < 1 || e_tol <= 0 || d_tol <= 0)
        return set_error(ctx, MYQC_E_ARG, "myqc_uhf_scf: invalid arguments");
    if ((Ca_guess == nullptr) != (Cb_guess == nullptr))
        return set_error(ctx, MYQC_E_ARG, "myqc_uhf_scf: give both guess orbital sets or neither");
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "myqc_uhf_scf: call myqc_eri_build first");
    ctx->scf_purify = ctx->scf_diag = ctx->scf_sp2_steps = 0;
    return uhf_scf(ctx, H, S, nalpha, nbeta, e_nuc, Ca_guess, Cb_guess, e_tol, d_tol, max_iter, energy, Ca, Cb, eig_a,
                   eig_b, iterations);
}

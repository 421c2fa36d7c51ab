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
#include "many.h"

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

// DIIS history slots (the extrapolation's coefficient arrays and the host
// solver are sized by it)
constexpr int kDiisSlots = 8;

struct LinComb {
    double c[kDiisSlots];
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
           *Dn = nullptr, *Do = nullptr, *Don = nullptr, *G = nullptr, *C = nullptr, *T1 = nullptr,
           *T2 = nullptr, *P2 = nullptr, *eig = nullptr, *work = nullptr, *dsc = nullptr, *lohi = nullptr,
           *part = nullptr, *red = nullptr, *sp2t = nullptr;
    double *hF = nullptr, *hE = nullptr;   // DIIS history, kDiis x n^2 each
    int *info = nullptr;
    static constexpr int kDiis = kDiisSlots;
    void release() {
        double **ps[] = {&H,  &S,  &X,   &Xi,  &F,   &Fx,   &D,    &Dn,   &Do,   &Don, &G, &C,
                         &T1, &T2, &P2, &eig, &work, &dsc, &lohi, &part, &red, &hF,  &hE, &sp2t};
        for (auto p : ps) dfree((void **)p);
        dfree((void **)&info);
        n = -1;
    }
    int ensure(myqc_ctx *ctx, int nb) {
        if (n == nb) return 0;
        release();
        const size_t nn = (size_t)nb * nb;
        double **mats[] = {&H, &S, &X, &Xi, &F, &Fx, &D, &Dn, &Do, &Don, &G, &C, &T1, &T2, &P2};
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

// solve the DIIS system on the host (m <= kDiisSlots), Gaussian elimination with pivoting
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
    // redo: this iteration rebuilds in FP64 the density the incremental
    // passes converged on (E32: their energy for it)
    bool redo = false, redo_next = false;
    double E32 = 0.0;
    double *fslot = nullptr;
    int it = 1;
    // per-iteration convergence data to stderr when MYQC_SCF_TRACE is set
    const bool trace = getenv("MYQC_SCF_TRACE") != nullptr;
    while (true) {
        redo = redo_next;
        redo_next = false;
        const bool full = !incr || final_fp64 || since_full == 0 || since_full >= 8;
        if (full) {
            MYQC_CHECK(digest_G(ctx, b.D, b.G, MYQC_FP64));
            since_full = 1;
        } else {
            // G += G32(D - Dn): Dn holds the density the previous pass digested
            MYQC_CHECK(digest_G(ctx, b.D, b.G, MYQC_FP32, b.Dn));
            ++since_full;
        }
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
            // the FP64 rebuild of the density the incremental passes
            // converged on is accepted at once when the exact Fock matrix
            // confirms it: its own commutator error below the bound, and its
            // energy within e_tol of the incremental one
            if (redo && sqrt(fabs(r[jb0 + head])) < 100 * d_tol && fabs(E - E32) < e_tol) {
                conv = true;
                break;
            }
            if (!final_fp64) {
                // confirm with FP64 builds, starting with this iteration's
                final_fp64 = true;
                if (!full) {
                    E32 = E;
                    redo_next = true;
                    continue;
                }
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

// ---------------------------------------------------------------------------
// Many systems of one basis size in lockstep (rhf_scf for each, on the first
// context's stream): the per-system digestion passes run back to back, and
// every small step of the loop -- the basis changes, the DIIS error and
// extrapolation, the level shift, the fused reductions and each purification
// step -- is one launch for all systems, with one host synchronisation per
// iteration for all of them.  Each system converges on its own; converged
// ones drop out of the launches.
// ---------------------------------------------------------------------------

struct EltMany {
    const double *X[kMany];
    const double *Y[kMany];
    double *Z[kMany];
    double a[kMany], b[kMany], c[kMany];
};

// Z_k = a_k X_k + b_k Y_k + c_k I (X or Y may be null); blockIdx.z: system
__global__ void k_axpbyc_many(EltMany e, int n) {
    const int k = blockIdx.z, i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const size_t q = i + (size_t)n * j;
    double v = (e.X[k] ? e.a[k] * e.X[k][q] : 0.0) + (e.Y[k] ? e.b[k] * e.Y[k][q] : 0.0);
    if (i == j) v += e.c[k];
    e.Z[k][q] = v;
}

// Z_k = X_k - X_k^T
__global__ void k_antisym_many(EltMany e, int n) {
    const int k = blockIdx.z, i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    e.Z[k][i + (size_t)n * j] = e.X[k][i + (size_t)n * j] - e.X[k][j + (size_t)n * i];
}

struct LinMany {
    const double *hF[kMany];
    double *Fx[kMany];
    double c[kMany][ScfBuffers::kDiis];
    int m[kMany];
};

// Fx_k = sum_i c_{k,i} hF_k[i]; blockIdx.y: system
__global__ void k_lincomb_many(LinMany l, size_t nn) {
    const int k = blockIdx.y;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int i = 0; i < l.m[k]; ++i) v += l.c[k][i] * l.hF[k][nn * i + e];
        l.Fx[k][e] = v;
    }
}

struct RedJob {
    const double *x, *y;
    int mode, pad;
};

// the fused reductions of k_reduce_part / k_reduce_final over a job table
__global__ __launch_bounds__(256) void k_reduce_part_tab(const RedJob *__restrict__ jobs, int n,
                                                         double *__restrict__ part) {
    __shared__ double sh[4];
    const int d = blockIdx.y;
    const RedJob J = jobs[d];
    const size_t nn = (size_t)n * n;
    double acc = 0.0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x) {
        if (J.mode == 0)
            acc += J.x[e] * J.y[e];
        else if (J.mode == 1)
            acc = fmax(acc, fabs(J.x[e] - J.y[e]));
        else
            acc = fmax(acc, fabs(J.x[e] - ((e % n) == (e / n) ? 1.0 : 0.0)));
    }
    acc = wave_fold(acc, J.mode);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = sh[0];
        for (int w = 1; w < 4; ++w) v = J.mode == 0 ? v + sh[w] : fmax(v, sh[w]);
        part[(size_t)d * kRedBlocks + blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(64) void k_reduce_final_tab(const RedJob *__restrict__ jobs,
                                                         const double *__restrict__ part, double *__restrict__ out) {
    const int d = blockIdx.x;
    const double v = wave_fold(part[(size_t)d * kRedBlocks + threadIdx.x], jobs[d].mode);
    if (threadIdx.x == 0) out[d] = v;
}

// copies of the systems' orthonormal-basis Fock matrices into one strided
// buffer (for the batched eigen-decomposition) and back out of it
__global__ void k_gather_mats(GemmMany g, int n, double *__restrict__ dst) {
    const int k = blockIdx.z, i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    dst[(size_t)k * n * n + i + (size_t)n * j] = g.A[k][i + (size_t)n * j];
}

// device buffers of the many-system loop, kept by the first context
struct ManyBuffers {
    int cap_jobs = 0, cap_k = 0, n = -1;
    RedJob *jobs = nullptr;
    double *part = nullptr, *red = nullptr, *eigA = nullptr, *eigW = nullptr, *eigE = nullptr;
    int *info = nullptr, *status = nullptr;
    RedJob *h_jobs = nullptr;   // pinned
    double *h_red = nullptr;    // pinned
    // the Fock builds of one iteration fan out over kFan streams (the
    // contexts' own): one event to start them after the lead stream's work,
    // one per stream to join them back.  Eight streams over the process's
    // four hardware queues measured 0.7 % faster than four for 32 40-water
    // clusters (181.9 against 180.6 solves/s, two runs each)
    static constexpr int kFan = 8;
    hipEvent_t fork = nullptr, join[kFan] = {};
};

static int many_buffers(myqc_ctx *ctx, int K, int n, int njobs, ManyBuffers *&mb) {
    if (!ctx->many) ctx->many = new ManyBuffers();
    mb = ctx->many;
    if (njobs > mb->cap_jobs) {
        MYQC_CHECK(dalloc(ctx, (void **)&mb->jobs, sizeof(RedJob) * njobs));
        MYQC_CHECK(dalloc(ctx, (void **)&mb->part, sizeof(double) * kRedBlocks * njobs));
        MYQC_CHECK(dalloc(ctx, (void **)&mb->red, sizeof(double) * njobs));
        if (mb->h_jobs) hipHostFree(mb->h_jobs);
        if (mb->h_red) hipHostFree(mb->h_red);
        mb->h_jobs = nullptr;
        mb->h_red = nullptr;
        MYQC_HIP(ctx, hipHostMalloc((void **)&mb->h_jobs, sizeof(RedJob) * njobs, hipHostMallocDefault));
        MYQC_HIP(ctx, hipHostMalloc((void **)&mb->h_red, sizeof(double) * njobs, hipHostMallocDefault));
        mb->cap_jobs = njobs;
    }
    if (K > mb->cap_k || n != mb->n) {
        const size_t nn = (size_t)n * n;
        MYQC_CHECK(dalloc(ctx, (void **)&mb->eigA, sizeof(double) * nn * K));
        MYQC_CHECK(dalloc(ctx, (void **)&mb->eigW, sizeof(double) * (size_t)n * K));
        MYQC_CHECK(dalloc(ctx, (void **)&mb->eigE, sizeof(double) * (size_t)n * K));
        MYQC_CHECK(dalloc(ctx, (void **)&mb->info, sizeof(int) * K));
        MYQC_CHECK(dalloc(ctx, (void **)&mb->status, sizeof(int) * 2 * K));
        mb->cap_k = K;
        mb->n = n;
    }
    if (!mb->fork) {
        MYQC_HIP(ctx, hipEventCreateWithFlags(&mb->fork, hipEventDisableTiming));
        for (int s = 0; s < ManyBuffers::kFan; ++s)
            MYQC_HIP(ctx, hipEventCreateWithFlags(&mb->join[s], hipEventDisableTiming));
    }
    return 0;
}

void free_many(myqc_ctx *ctx) {
    if (!ctx->many) return;
    ManyBuffers *mb = ctx->many;
    void **ps[] = {(void **)&mb->jobs, (void **)&mb->part, (void **)&mb->red, (void **)&mb->eigA,
                   (void **)&mb->eigW, (void **)&mb->eigE, (void **)&mb->info, (void **)&mb->status};
    for (auto p : ps) dfree(p);
    if (mb->h_jobs) hipHostFree(mb->h_jobs);
    if (mb->h_red) hipHostFree(mb->h_red);
    if (mb->fork) hipEventDestroy(mb->fork);
    for (hipEvent_t e : mb->join)
        if (e) hipEventDestroy(e);
    delete mb;
    ctx->many = nullptr;
}

// runs a job table (jobs filled in mb->h_jobs[0..m)) and returns the results
// in mb->h_red (one synchronisation)
static int reduce_tab(myqc_ctx *ctx, ManyBuffers &mb, int m, int n) {
    MYQC_HIP(ctx, hipMemcpyAsync(mb.jobs, mb.h_jobs, sizeof(RedJob) * m, hipMemcpyHostToDevice, ctx->stream));
    k_reduce_part_tab<<<dim3(kRedBlocks, m), 256, 0, ctx->stream>>>(mb.jobs, n, mb.part);
    k_reduce_final_tab<<<m, 64, 0, ctx->stream>>>(mb.jobs, mb.part, mb.red);
    MYQC_HIP(ctx, hipGetLastError());
    MYQC_HIP(ctx, hipMemcpyAsync(mb.h_red, mb.red, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// C_k = alpha A_k B_k (n x n) over the listed systems
static int gemm_many(myqc_ctx *lead, const std::vector<int> &sys, int n, const std::vector<const double *> &A,
                     const std::vector<const double *> &B, const std::vector<double *> &C, double alpha) {
    GemmMany g{};
    for (size_t z = 0; z < sys.size(); ++z) {
        g.A[z] = A[sys[z]];
        g.B[z] = B[sys[z]];
        g.C[z] = C[sys[z]];
    }
    if (sys.empty()) return 0;
    return gemm_f64_many(lead, n, n, n, g, (int)sys.size(), alpha);
}

// per-system state of the loop (the locals of rhf_scf)
struct ManyState {
    myqc_ctx *ctx = nullptr;
    ScfBuffers *b = nullptr;
    int nocc = 0;
    double e_nuc = 0.0;
    bool guessed = false;
    std::vector<double> Bm;
    int hist = 0, head = 0;
    double E = 0.0, E_old = 0.0, dE_prev = 1e300, err_prev = 1e300;
    bool prev64 = false, final_fp64 = false, conv = false, done = false;
    bool redo = false;   // this iteration rebuilds the incrementally converged density in FP64
    double E32 = 0.0;    // the incremental energy of that density
    int since_full = 0, it = 1;
    double *fslot = nullptr;
    const double *fuse = nullptr;
    bool full = true;
    int jdd = -1, jb0 = 0, j0 = 0, hn = 0;
    std::vector<double> cdiis;
};

static int rhf_scf_many(myqc_ctx *const *ctxs, int K, const double *const *H, const double *const *S, const int *nocc,
                        const double *e_nuc, const double *const *D_guess, double e_tol, double d_tol, int max_iter,
                        int precision, double *energy, double *const *C_out, double *const *eig_out,
                        int *iters_out) {
    myqc_ctx *lead = ctxs[0];
    const int n = lead->n;
    const size_t nn = (size_t)n * n;
    const int MD = ScfBuffers::kDiis;
    const bool incr = (precision == MYQC_FP32);
    ManyBuffers *mbp = nullptr;
    MYQC_CHECK(many_buffers(lead, K, n, K * (3 + MD) + 8, mbp));
    ManyBuffers &mb = *mbp;
    std::vector<ManyState> st(K);
    // every context's work goes to the lead stream for the call (restored at return)
    std::vector<hipStream_t> own(K);
    for (int k = 0; k < K; ++k) {
        MYQC_HIP(lead, hipStreamSynchronize(ctxs[k]->stream));
        own[k] = ctxs[k]->stream;
    }
    struct Restore {
        myqc_ctx *const *c;
        std::vector<hipStream_t> &s;
        ~Restore() {
            for (size_t k = 0; k < s.size(); ++k) (void)hipStreamSynchronize(s[k]);
            for (size_t k = 0; k < s.size(); ++k) {
                c[k]->stream = s[k];
                rocblas_set_stream(c[k]->blas, s[k]);
            }
        }
    } restore{ctxs, own};
    for (int k = 0; k < K; ++k) {
        ManyState &q = st[k];
        q.ctx = ctxs[k];
        q.ctx->stream = lead->stream;
        MYQC_BLAS(lead, rocblas_set_stream(q.ctx->blas, lead->stream));
        if (!q.ctx->scf) q.ctx->scf = new ScfBuffers();
        MYQC_CHECK(q.ctx->scf->ensure(q.ctx, n));
        q.b = q.ctx->scf;
        q.nocc = nocc[k];
        q.e_nuc = e_nuc[k];
        q.guessed = D_guess && D_guess[k];
        q.Bm.assign((size_t)MD * MD, 0.0);
        q.ctx->scf_purify = q.ctx->scf_diag = q.ctx->scf_sp2_steps = 0;
        MYQC_HIP(lead, hipMemcpyAsync(q.b->H, H[k], sizeof(double) * nn, hipMemcpyHostToDevice, lead->stream));
        MYQC_HIP(lead, hipMemcpyAsync(q.b->S, S[k], sizeof(double) * nn, hipMemcpyHostToDevice, lead->stream));
        if (q.guessed)
            MYQC_HIP(lead, hipMemcpyAsync(q.b->D, D_guess[k], sizeof(double) * nn, hipMemcpyHostToDevice, lead->stream));
    }
    std::vector<int> all(K);
    for (int k = 0; k < K; ++k) all[k] = k;
    auto ptrs = [&](auto f) {
        std::vector<decltype(f(st[0]))> v(K);
        for (int k = 0; k < K; ++k) v[k] = f(st[k]);
        return v;
    };
    const dim3 g2((n + 127) / 128, n);
    auto elt = [&](const std::vector<int> &sys, auto fill, bool antisym) {
        EltMany e{};
        for (size_t z = 0; z < sys.size(); ++z) fill(e, (int)z, st[sys[z]]);
        if (sys.empty()) return;
        if (antisym) k_antisym_many<<<dim3(g2.x, g2.y, (unsigned)sys.size()), 128, 0, lead->stream>>>(e, n);
        else k_axpbyc_many<<<dim3(g2.x, g2.y, (unsigned)sys.size()), 128, 0, lead->stream>>>(e, n);
    };
    // X = S^-1/2, Xi = S^1/2 by the coupled Newton-Schulz iteration of
    // inv_sqrt_ns for all systems at once (the eigen-decomposition path for
    // any that does not converge in its step budget)
    {
        Sp2Many sp{};
        for (int k = 0; k < K; ++k) {
            sp.F[k] = st[k].b->S;
            sp.lohi[k] = st[k].b->lohi;
            sp.slots[k] = st[k].b->dsc;
        }
        MYQC_CHECK(gershgorin_many(lead, sp, K, n));
        for (int k = 0; k < K; ++k)
            MYQC_HIP(lead, hipMemcpyAsync(mb.red + k, st[k].b->dsc + 7, sizeof(double), hipMemcpyDeviceToDevice,
                                          lead->stream));
        MYQC_HIP(lead, hipMemcpyAsync(mb.h_red, mb.red, sizeof(double) * K, hipMemcpyDeviceToHost, lead->stream));
        MYQC_HIP(lead, hipStreamSynchronize(lead->stream));
        std::vector<double> c(mb.h_red, mb.h_red + K);
        for (int k = 0; k < K; ++k)
            if (!(c[k] > 0.0)) return set_error(lead, MYQC_E_ARG, "rhf_many: system %d has no positive overlap bound", k);
        // Y = S / c (in Xi), Z = I (in X); Yn, Zn in T2, P2
        std::vector<double *> Y(K), Z(K), Yn(K), Zn(K), T1(K);
        for (int k = 0; k < K; ++k) {
            Y[k] = st[k].b->Xi;
            Z[k] = st[k].b->X;
            Yn[k] = st[k].b->T2;
            Zn[k] = st[k].b->P2;
            T1[k] = st[k].b->T1;
        }
        elt(all, [&](EltMany &e, int z, ManyState &q) {
                const int k = (int)(&q - &st[0]);
                e.X[z] = q.b->S; e.Z[z] = Y[k]; e.a[z] = 1.0 / c[k]; }, false);
        elt(all, [&](EltMany &e, int z, ManyState &q) {
                const int k = (int)(&q - &st[0]);
                e.Z[z] = Z[k]; e.c[z] = 1.0; }, false);
        auto cv = [](const std::vector<double *> &v) { return std::vector<const double *>(v.begin(), v.end()); };
        std::vector<int> live = all, extra(K, -1), okk(K, 0);
        for (int step = 0; step < 40 && !live.empty(); ++step) {
            MYQC_CHECK(gemm_many(lead, live, n, cv(Z), cv(Y), T1, 1.0));
            if (step >= 6 && (step & 1) == 0) {
                int m = 0;
                std::vector<int> who;
                for (int k : live)
                    if (extra[k] < 0) {
                        mb.h_jobs[m++] = RedJob{T1[k], nullptr, 2, 0};
                        who.push_back(k);
                    }
                if (m) {
                    MYQC_CHECK(reduce_tab(lead, mb, m, n));
                    // quadratic convergence: one more step takes 1e-9 below round-off
                    for (int j = 0; j < m; ++j)
                        if (mb.h_red[j] < 1e-9) extra[who[j]] = 1;
                }
            }
            std::vector<int> next;
            for (int k : live) {
                if (extra[k] == 0) okk[k] = 1;
                else next.push_back(k);
            }
            live.swap(next);
            if (live.empty()) break;
            elt(live, [&](EltMany &e, int z, ManyState &q) {
                    const int k = (int)(&q - &st[0]);
                    e.X[z] = T1[k]; e.Z[z] = T1[k]; e.a[z] = -0.5; e.c[z] = 1.5; }, false);
            MYQC_CHECK(gemm_many(lead, live, n, cv(Y), cv(T1), Yn, 1.0));
            MYQC_CHECK(gemm_many(lead, live, n, cv(T1), cv(Z), Zn, 1.0));
            for (int k : live) {
                std::swap(Y[k], Yn[k]);
                std::swap(Z[k], Zn[k]);
                if (extra[k] > 0) --extra[k];
            }
        }
        // X = Z / sqrt(c), Xi = Y sqrt(c) (through T1 and Don, as Z and Y may
        // sit in X and Xi), or the eigen-decomposition path
        for (int k = 0; k < K; ++k) {
            ScfBuffers &b = *st[k].b;
            if (!okk[k]) {
                MYQC_CHECK(inv_sqrt_eig(st[k].ctx, b, n));
                continue;
            }
            const double rc = sqrt(c[k]);
            k_axpbyc<<<g2, 128, 0, lead->stream>>>(Z[k], nullptr, b.T1, 1.0 / rc, 0.0, 0.0, n);
            k_axpbyc<<<g2, 128, 0, lead->stream>>>(Y[k], nullptr, b.Don, rc, 0.0, 0.0, n);
            MYQC_HIP(lead, hipMemcpyAsync(b.X, b.T1, sizeof(double) * nn, hipMemcpyDeviceToDevice, lead->stream));
            MYQC_HIP(lead, hipMemcpyAsync(b.Xi, b.Don, sizeof(double) * nn, hipMemcpyDeviceToDevice, lead->stream));
        }
    }
    auto Xs = ptrs([](ManyState &q) { return (const double *)q.b->X; });
    auto Xis = ptrs([](ManyState &q) { return (const double *)q.b->Xi; });
    auto T1s = ptrs([](ManyState &q) { return q.b->T1; });
    auto T2s = ptrs([](ManyState &q) { return q.b->T2; });
    auto cT1s = ptrs([](ManyState &q) { return (const double *)q.b->T1; });
    // purification of the listed systems' Fock matrices Fin[k] into Pout[k]
    int sp2_steps = 0;
    auto purify_many = [&](const std::vector<int> &sys, const std::vector<const double *> &Fin,
                           const std::vector<double *> &Pout) -> int {
        if (sys.empty()) return 0;
        Sp2Many sp{};
        for (size_t z = 0; z < sys.size(); ++z) {
            ScfBuffers &b = *st[sys[z]].b;
            sp.F[z] = Fin[sys[z]];
            sp.P[z] = Pout[sys[z]];
            sp.W[z] = b.P2;
            sp.lohi[z] = b.lohi;
            sp.slots[z] = b.dsc;
            sp.tp[z] = b.sp2t;
            sp.nocc[z] = st[sys[z]].nocc;
        }
        std::vector<int> ok;
        MYQC_CHECK(sp2_purify_many(lead, sp, (int)sys.size(), n, mb.status, ok, &sp2_steps));
        for (size_t z = 0; z < sys.size(); ++z) {
            ManyState &q = st[sys[z]];
            ++q.ctx->scf_purify;
            if (!ok[z]) {
                ++q.ctx->scf_diag;
                MYQC_CHECK(diag_orth(q.ctx, *q.b, Fin[sys[z]], n));
                MYQC_CHECK(gemm(q.ctx, NN, TT, n, n, q.nocc, 1.0, q.b->T2, n, q.b->T2, n, 0.0, Pout[sys[z]], n));
            }
        }
        return 0;
    };
    // the starting densities
    {
        std::vector<int> g, core;
        for (int k = 0; k < K; ++k) (st[k].guessed ? g : core).push_back(k);
        // Do = S^1/2 D S^1/2 / 2
        auto Ds = ptrs([](ManyState &q) { return (const double *)q.b->D; });
        auto Dos = ptrs([](ManyState &q) { return q.b->Do; });
        MYQC_CHECK(gemm_many(lead, g, n, Ds, Xis, T1s, 1.0));
        MYQC_CHECK(gemm_many(lead, g, n, Xis, cT1s, Dos, 0.5));
        // core Hamiltonian guess: Fo = X H X, purified, D = 2 X Do X
        auto Hs = ptrs([](ManyState &q) { return (const double *)q.b->H; });
        auto Fs = ptrs([](ManyState &q) { return q.b->F; });
        auto cFs = ptrs([](ManyState &q) { return (const double *)q.b->F; });
        MYQC_CHECK(gemm_many(lead, core, n, Hs, Xs, T1s, 1.0));
        MYQC_CHECK(gemm_many(lead, core, n, Xs, cT1s, Fs, 1.0));
        MYQC_CHECK(purify_many(core, cFs, Dos));
        auto cDos = ptrs([](ManyState &q) { return (const double *)q.b->Do; });
        auto DsW = ptrs([](ManyState &q) { return q.b->D; });
        MYQC_CHECK(gemm_many(lead, core, n, cDos, Xs, T1s, 1.0));
        MYQC_CHECK(gemm_many(lead, core, n, Xs, cT1s, DsW, 2.0));
    }
    std::vector<int> active = all;
    // the fan-out streams of the Fock builds: the contexts' own (the first is
    // the lead stream)
    const int nfan = std::min(K, (int)ManyBuffers::kFan);
    while (!active.empty()) {
        // Fock builds: system i of the active list on fan-out stream i mod
        // nfan, so that one pass's tail overlaps the next pass's start
        MYQC_HIP(lead, hipEventRecord(mb.fork, lead->stream));
        for (int f = 1; f < nfan; ++f) MYQC_HIP(lead, hipStreamWaitEvent(own[f], mb.fork, 0));
        for (size_t i = 0; i < active.size(); ++i) {
            ManyState &q = st[active[i]];
            ScfBuffers &b = *q.b;
            hipStream_t fs = own[i % nfan];
            q.ctx->stream = fs;
            q.full = !incr || q.final_fp64 || q.since_full == 0 || q.since_full >= 8;
            int rc;
            if (q.full) {
                rc = digest_G(q.ctx, b.D, b.G, MYQC_FP64);
                q.since_full = 1;
            } else {
                rc = digest_G(q.ctx, b.D, b.G, MYQC_FP32, b.Dn);   // Dn: the density the previous pass digested
                ++q.since_full;
            }
            q.ctx->stream = lead->stream;
            if (rc) return set_error(lead, MYQC_E_HIP, "%s", q.ctx->err.c_str());
            q.fslot = b.hF + nn * q.head;
        }
        for (int f = 1; f < nfan; ++f) {
            MYQC_HIP(lead, hipEventRecord(mb.join[f], own[f]));
            MYQC_HIP(lead, hipStreamWaitEvent(lead->stream, mb.join[f], 0));
        }
        // F = H + G; Fo = X F X into the DIIS slot; error Fo Do - Do Fo
        elt(active, [&](EltMany &e, int z, ManyState &q) { e.X[z] = q.b->H; e.Y[z] = q.b->G; e.Z[z] = q.b->F; e.a[z] = 1.0; e.b[z] = 1.0; }, false);
        auto cF = ptrs([](ManyState &q) { return (const double *)q.b->F; });
        auto fsl = ptrs([](ManyState &q) { return q.fslot; });
        auto cfsl = ptrs([](ManyState &q) { return (const double *)q.fslot; });
        auto cDo = ptrs([](ManyState &q) { return (const double *)q.b->Do; });
        MYQC_CHECK(gemm_many(lead, active, n, cF, Xs, T1s, 1.0));
        MYQC_CHECK(gemm_many(lead, active, n, Xs, cT1s, fsl, 1.0));
        MYQC_CHECK(gemm_many(lead, active, n, cfsl, cDo, T2s, 1.0));
        elt(active, [&](EltMany &e, int z, ManyState &q) { e.X[z] = q.b->T2; e.Z[z] = q.b->hE + nn * q.head; }, true);
        // one reduction table for all: D.F, D.H, max|D - D_prev|, the DIIS row
        int m = 0;
        for (int k : active) {
            ManyState &q = st[k];
            ScfBuffers &b = *q.b;
            q.hn = std::min(q.hist + 1, MD);
            q.j0 = m;
            mb.h_jobs[m++] = RedJob{b.D, b.F, 0, 0};
            mb.h_jobs[m++] = RedJob{b.D, b.H, 0, 0};
            q.jdd = -1;
            if (q.it > 1) {
                q.jdd = m;
                mb.h_jobs[m++] = RedJob{b.D, b.Dn, 1, 0};
            }
            q.jb0 = m;
            const double *es = b.hE + nn * q.head;
            for (int j = 0; j < q.hn; ++j) mb.h_jobs[m++] = RedJob{es, b.hE + nn * j, 0, 0};
        }
        MYQC_CHECK(reduce_tab(lead, mb, m, n));
        std::vector<int> go, diis, shift;
        for (int k : active) {
            ManyState &q = st[k];
            const double *r = mb.h_red;
            q.E = 0.5 * (r[q.j0] + r[q.j0 + 1]) + q.e_nuc;
            bool stop = false;
            const bool redo = q.redo;
            q.redo = false;
            if (q.jdd >= 0 && q.it > 2 && q.dE_prev < e_tol && r[q.jdd] < d_tol && q.err_prev < 100 * d_tol) {
                if (!incr || (q.prev64 && q.full)) {
                    q.conv = true;
                    stop = true;
                } else if (redo && sqrt(fabs(r[q.jb0 + q.head])) < 100 * d_tol && fabs(q.E - q.E32) < e_tol) {
                    // the FP64 rebuild confirms the incrementally converged
                    // density (as in the single-system loop)
                    q.conv = true;
                    stop = true;
                } else if (!q.final_fp64) {
                    q.final_fp64 = true;
                    if (!q.full) {
                        // confirm with an FP64 build of this same density
                        q.since_full = 0;
                        q.redo = true;
                        q.E32 = q.E;
                        go.push_back(-1 - k);
                        continue;
                    }
                }
            }
            if (stop || q.it > max_iter) {
                q.done = true;
                continue;
            }
            for (int j = 0; j < q.hn; ++j) {
                q.Bm[q.head * MD + j] = r[q.jb0 + j];
                q.Bm[j * MD + q.head] = r[q.jb0 + j];
            }
            q.hist = q.hn;
            q.fuse = q.fslot;
            if (q.hist >= 2) {
                std::vector<double> Bs((size_t)q.hist * q.hist), cc;
                for (int i = 0; i < q.hist; ++i)
                    for (int j = 0; j < q.hist; ++j) Bs[i * q.hist + j] = q.Bm[i * MD + j];
                if (diis_coeffs(Bs, q.hist, cc)) {
                    diis.push_back(k);
                    q.fuse = q.b->Fx;
                    q.cdiis = cc;
                }
            }
            q.err_prev = sqrt(fabs(q.Bm[q.head * MD + q.head]));
            q.dE_prev = fabs(q.E - q.E_old);
            q.E_old = q.E;
            q.prev64 = q.full;
            q.head = (q.head + 1) % MD;
            if (q.guessed && q.err_prev > kShiftErr) shift.push_back(k);
            go.push_back(k);
        }
        // DIIS extrapolation and level shift, all systems at once
        if (!diis.empty()) {
            LinMany l{};
            for (size_t z = 0; z < diis.size(); ++z) {
                ManyState &q = st[diis[z]];
                l.hF[z] = q.b->hF;
                l.Fx[z] = q.b->Fx;
                l.m[z] = q.hist;
                for (int i = 0; i < q.hist; ++i) l.c[z][i] = q.cdiis[i];
            }
            k_lincomb_many<<<dim3(64, (unsigned)diis.size()), 256, 0, lead->stream>>>(l, nn);
        }
        elt(shift, [&](EltMany &e, int z, ManyState &q) {
                e.X[z] = q.fuse; e.Y[z] = q.b->Do; e.Z[z] = q.b->T1; e.a[z] = 1.0; e.b[z] = -kShift; e.c[z] = kShift; }, false);
        for (int k : shift) st[k].fuse = st[k].b->T1;
        std::vector<int> pur, redo;
        for (int x : go) (x >= 0 ? pur : redo).push_back(x >= 0 ? x : -1 - x);
        // purify, swap, D = 2 X Do X for the systems that go on
        {
            std::vector<const double *> Fin(K);
            std::vector<double *> Pout(K);
            for (int k : pur) {
                Fin[k] = st[k].fuse;
                Pout[k] = st[k].b->Don;
            }
            MYQC_CHECK(purify_many(pur, Fin, Pout));
            for (int k : pur) {
                ScfBuffers &b = *st[k].b;
                std::swap(b.D, b.Dn);
                std::swap(b.Do, b.Don);
            }
            auto cDo2 = ptrs([](ManyState &q) { return (const double *)q.b->Do; });
            auto Dw = ptrs([](ManyState &q) { return q.b->D; });
            MYQC_CHECK(gemm_many(lead, pur, n, cDo2, Xs, T1s, 1.0));
            MYQC_CHECK(gemm_many(lead, pur, n, Xs, cT1s, Dw, 2.0));
            for (int k : pur) ++st[k].it;
        }
        std::vector<int> next;
        for (int k : active)
            if (!st[k].done) next.push_back(k);
        active.swap(next);
        (void)redo;
    }
    // orbitals: eigenvectors of each system's final orthonormal-basis Fock
    // matrix, one strided batch
    {
        GemmMany g{};
        for (int k = 0; k < K; ++k) g.A[k] = st[k].fslot;
        k_gather_mats<<<dim3(g2.x, g2.y, K), 128, 0, lead->stream>>>(g, n, mb.eigA);
        MYQC_CHECK(blas_reserve(lead, [&] {
            (void)rocsolver_dsyevd_strided_batched(lead->blas, rocblas_evect_original, rocblas_fill_lower, n, mb.eigA, n,
                                                   (rocblas_stride)nn, mb.eigW, n, mb.eigE, n, mb.info, K);
        }));
        MYQC_BLAS(lead, rocsolver_dsyevd_strided_batched(lead->blas, rocblas_evect_original, rocblas_fill_lower, n,
                                                         mb.eigA, n, (rocblas_stride)nn, mb.eigW, n, mb.eigE, n,
                                                         mb.info, K));
        std::vector<const double *> V(K);
        for (int k = 0; k < K; ++k) V[k] = mb.eigA + nn * k;
        auto Cs = ptrs([](ManyState &q) { return q.b->C; });
        MYQC_CHECK(gemm_many(lead, all, n, Xs, V, Cs, 1.0));
        std::vector<int> info(K);
        MYQC_HIP(lead, hipMemcpyAsync(info.data(), mb.info, sizeof(int) * K, hipMemcpyDeviceToHost, lead->stream));
        for (int k = 0; k < K; ++k) {
            if (C_out && C_out[k]) MYQC_HIP(lead, hipMemcpyAsync(C_out[k], st[k].b->C, sizeof(double) * nn, hipMemcpyDeviceToHost, lead->stream));
            if (eig_out && eig_out[k]) MYQC_HIP(lead, hipMemcpyAsync(eig_out[k], mb.eigW + (size_t)n * k, sizeof(double) * n, hipMemcpyDeviceToHost, lead->stream));
        }
        MYQC_HIP(lead, hipStreamSynchronize(lead->stream));
        for (int k = 0; k < K; ++k)
            if (info[k] != 0) return set_error(lead, MYQC_E_HIP, "rocsolver_dsyevd_strided_batched info=%d (system %d)", info[k], k);
    }
    int nconv = 0;
    for (int k = 0; k < K; ++k) {
        if (energy) energy[k] = st[k].E;
        if (iters_out) iters_out[k] = st[k].it - 1;
        nconv += st[k].conv;
    }
    if (nconv < K) return set_error(lead, MYQC_E_NOCONV, "rhf_many: %d of %d systems did not converge in %d iterations", K - nconv, K, max_iter);
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

extern "C" int myqc_rhf_scf_many(myqc_ctx *const *ctxs, int nsys, const double *const *H, const double *const *S,
                                 const int *nocc, const double *e_nuc, const double *const *D_guess, double e_tol,
                                 double d_tol, int max_iter, int precision, double *energy, double *const *C,
                                 double *const *eig, int *iterations) {
    if (!ctxs || nsys < 1 || !H || !S || !nocc || !e_nuc) return MYQC_E_ARG;
    myqc_ctx *lead = ctxs[0];
    if (!lead) return MYQC_E_ARG;
    if (nsys > kMany) return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: %d systems (at most %d)", nsys, kMany);
    if (max_iter < 1 || e_tol <= 0 || d_tol <= 0) return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: invalid tolerances");
    if (precision != MYQC_FP64 && precision != MYQC_FP32)
        return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: precision %d", precision);
    for (int k = 0; k < nsys; ++k) {
        myqc_ctx *c = ctxs[k];
        if (!c || !H[k] || !S[k]) return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: system %d: null argument", k);
        for (int j = 0; j < k; ++j)
            if (ctxs[j] == c) return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: context of system %d repeated", k);
        if (!c->have_system || !c->have_eri || (precision == MYQC_FP32 && !c->have_eri32))
            return set_error(lead, MYQC_E_STATE, "myqc_rhf_scf_many: system %d: call myqc_eri_build first", k);
        if (c->n != lead->n || c->device != lead->device)
            return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: system %d: n=%d on device %d (system 0: n=%d, %d)", k,
                             c->n, c->device, lead->n, lead->device);
        if (nocc[k] < 1 || nocc[k] >= c->n)
            return set_error(lead, MYQC_E_ARG, "myqc_rhf_scf_many: system %d: nocc=%d", k, nocc[k]);
    }
    return rhf_scf_many(ctxs, nsys, H, S, nocc, e_nuc, D_guess, e_tol, d_tol, max_iter, precision, energy, C, eig,
                        iterations);
}

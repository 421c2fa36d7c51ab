// jk.hip -- J/K (Fock G) digestion of the stored unique integrals, the RHF
// density, and the in-process RHF driver.
//
// Digestion: every stored value is v' = (ij|kl) times its degeneracy weight
// (0 for repeats inside a block, 1/2 per index coincidence), so one pass over
// the unique integrals with
//     A_ij += 2 D_kl v'   A_kl += 2 D_ij v'
//     A_ik -= D_jl v'/2   A_jk -= D_il v'/2   A_il -= D_jk v'/2   A_jl -= D_ik v'/2
// and G = A + A^T equals the reference's
//     G(i,j) = sum_kl D(k,l)[XX(i,j,k,l) - XX(i,k,j,l)/4 - XX(i,l,j,k)/4]
// (src/I2G/RHFI2G.f90:80-90) for the symmetrised D (the reference's formula
// only sees D through D + D^T).  One thread per shell quartet accumulates its
// six J/K blocks in registers and adds them to A with hardware atomics; A and
// D (n^2 each) stay L2-resident while the integrals stream from HBM.
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "context.h"

namespace myqc {

template <int C> struct PL;
template <> struct PL<0> { static constexpr int a = 0, b = 0; };
template <> struct PL<1> { static constexpr int a = 1, b = 0; };
template <> struct PL<2> { static constexpr int a = 1, b = 1; };

template <typename TV>
struct DigArgs {
    const ShellPair *pairs;
    const int *sorted_b, *sorted_k;
    const long long *cum;
    int nbra;
    long long nq;
    int same_class;
    const TV *v;
    const TV *D;     // symmetric density (or density change), n x n
    TV *A;           // accumulator, n x n
    int n;
};

__device__ __forceinline__ void find_q(const long long *__restrict__ cum, int nbra, long long t, int same, int &rb,
                                       long long &rk) {
    int lo = 0, hi = nbra;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (cum[mid] <= t) lo = mid;
        else hi = mid;
    }
    rb = lo;
    rk = t - cum[lo] + (same ? lo : 0);
}

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

template <int CB, int CK, typename TV>
__global__ __launch_bounds__(256) void k_digest(DigArgs<TV> a) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq) return;
    constexpr int na = ncart(PL<CB>::a), nb = ncart(PL<CB>::b), nc = ncart(PL<CK>::a), nd = ncart(PL<CK>::b);
    constexpr int BLK = na * nb * nc * nd;
    int rb;
    long long rk;
    find_q(a.cum, a.nbra, t, a.same_class, rb, rk);
    const ShellPair bra = a.pairs[a.sorted_b[rb]], ket = a.pairs[a.sorted_k[rk]];
    const int i0 = bra.ao_a, j0 = bra.ao_b, k0 = ket.ao_a, l0 = ket.ao_b;
    const size_t n = a.n;
    const TV *__restrict__ D = a.D;
    TV x[BLK];
#pragma unroll
    for (int e = 0; e < BLK; ++e) x[e] = a.v[t * BLK + e];
    TV Jab[na][nb], Jcd[nc][nd], Kac[na][nc], Kbc[nb][nc], Kad[na][nd], Kbd[nb][nd];
#pragma unroll
    for (int p = 0; p < na; ++p)
#pragma unroll
        for (int q = 0; q < nb; ++q) Jab[p][q] = 0;
#pragma unroll
    for (int p = 0; p < nc; ++p)
#pragma unroll
        for (int q = 0; q < nd; ++q) Jcd[p][q] = 0;
#pragma unroll
    for (int p = 0; p < na; ++p)
#pragma unroll
        for (int q = 0; q < nc; ++q) Kac[p][q] = 0;
#pragma unroll
    for (int p = 0; p < nb; ++p)
#pragma unroll
        for (int q = 0; q < nc; ++q) Kbc[p][q] = 0;
#pragma unroll
    for (int p = 0; p < na; ++p)
#pragma unroll
        for (int q = 0; q < nd; ++q) Kad[p][q] = 0;
#pragma unroll
    for (int p = 0; p < nb; ++p)
#pragma unroll
        for (int q = 0; q < nd; ++q) Kbd[p][q] = 0;
#pragma unroll
    for (int ia = 0; ia < na; ++ia)
#pragma unroll
        for (int ib = 0; ib < nb; ++ib) {
            const TV dij = D[(i0 + ia) + n * (j0 + ib)];
#pragma unroll
            for (int ic = 0; ic < nc; ++ic) {
                const TV dik = D[(i0 + ia) + n * (k0 + ic)];
                const TV djk = D[(j0 + ib) + n * (k0 + ic)];
#pragma unroll
                for (int id = 0; id < nd; ++id) {
                    const TV v = x[((ia * nb + ib) * nc + ic) * nd + id];
                    const TV dkl = D[(k0 + ic) + n * (l0 + id)];
                    const TV djl = D[(j0 + ib) + n * (l0 + id)];
                    const TV dil = D[(i0 + ia) + n * (l0 + id)];
                    Jab[ia][ib] += dkl * v;
                    Jcd[ic][id] += dij * v;
                    Kac[ia][ic] += djl * v;
                    Kbc[ib][ic] += dil * v;
                    Kad[ia][id] += djk * v;
                    Kbd[ib][id] += dik * v;
                }
            }
        }
    TV *A = a.A;
#pragma unroll
    for (int p = 0; p < na; ++p)
#pragma unroll
        for (int q = 0; q < nb; ++q) atomic_add(&A[(i0 + p) + n * (j0 + q)], (TV)2 * Jab[p][q]);
#pragma unroll
    for (int p = 0; p < nc; ++p)
#pragma unroll
        for (int q = 0; q < nd; ++q) atomic_add(&A[(k0 + p) + n * (l0 + q)], (TV)2 * Jcd[p][q]);
#pragma unroll
    for (int p = 0; p < na; ++p)
#pragma unroll
        for (int q = 0; q < nc; ++q) atomic_add(&A[(i0 + p) + n * (k0 + q)], (TV)(-0.5) * Kac[p][q]);
#pragma unroll
    for (int p = 0; p < nb; ++p)
#pragma unroll
        for (int q = 0; q < nc; ++q) atomic_add(&A[(j0 + p) + n * (k0 + q)], (TV)(-0.5) * Kbc[p][q]);
#pragma unroll
    for (int p = 0; p < na; ++p)
#pragma unroll
        for (int q = 0; q < nd; ++q) atomic_add(&A[(i0 + p) + n * (l0 + q)], (TV)(-0.5) * Kad[p][q]);
#pragma unroll
    for (int p = 0; p < nb; ++p)
#pragma unroll
        for (int q = 0; q < nd; ++q) atomic_add(&A[(j0 + p) + n * (l0 + q)], (TV)(-0.5) * Kbd[p][q]);
}

// --- small elementwise kernels ------------------------------------------------

__global__ void k_symmetrize(const double *__restrict__ D, double *__restrict__ Ds, float *__restrict__ Ds32, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double v = 0.5 * (D[i + (size_t)n * j] + D[j + (size_t)n * i]);
    if (Ds) Ds[i + (size_t)n * j] = v;
    if (Ds32) Ds32[i + (size_t)n * j] = (float)v;
}

// Ds32 = (float) sym(D - Dref)
__global__ void k_symdiff32(const double *__restrict__ D, const double *__restrict__ Dref, float *__restrict__ Ds32,
                            int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const size_t a = i + (size_t)n * j, b = j + (size_t)n * i;
    Ds32[a] = (float)(0.5 * ((D[a] - Dref[a]) + (D[b] - Dref[b])));
}

// G (+)= A + A^T
__global__ void k_assemble(const double *__restrict__ A, double *__restrict__ G, int n, int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double v = A[i + (size_t)n * j] + A[j + (size_t)n * i];
    G[i + (size_t)n * j] = accumulate ? G[i + (size_t)n * j] + v : v;
}

__global__ void k_assemble32(const float *__restrict__ A, double *__restrict__ G, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    G[i + (size_t)n * j] += (double)A[i + (size_t)n * j] + (double)A[j + (size_t)n * i];
}

template <typename T>
__global__ void k_zero(T *__restrict__ p, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

static dim3 grid2(int n) { return dim3((n + 127) / 128, n); }

int ensure_work(myqc_ctx *ctx) {
    if (ctx->work_n == ctx->n) return 0;
    const size_t nn = (size_t)ctx->n * ctx->n;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_D, sizeof(double) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_A, sizeof(double) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_G, sizeof(double) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_D32, sizeof(float) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_A32, sizeof(float) * nn));
    ctx->work_n = ctx->n;
    return 0;
}

template <typename TV>
static int launch_digest(myqc_ctx *ctx, const TV *D, TV *A) {
    for (int c = 0; c < 6; ++c) {
        const QuartetClass &qc = ctx->cls[c];
        if (qc.nq == 0) continue;
        DigArgs<TV> a;
        a.pairs = ctx->d_pairs;
        a.sorted_b = ctx->d_sorted[qc.cb];
        a.sorted_k = ctx->d_sorted[qc.ck];
        a.cum = qc.d_cum;
        a.nbra = qc.nbra;
        a.nq = qc.nq;
        a.same_class = qc.cb == qc.ck;
        a.D = D;
        a.A = A;
        a.n = ctx->n;
        if constexpr (sizeof(TV) == 8) a.v = (const TV *)qc.d_v64;
        else a.v = (const TV *)qc.d_v32;
        const unsigned nblk = (unsigned)((qc.nq + 255) / 256);
        switch (c) {
            case 0: k_digest<0, 0, TV><<<nblk, 256, 0, ctx->stream>>>(a); break;
            case 1: k_digest<1, 0, TV><<<nblk, 256, 0, ctx->stream>>>(a); break;
            case 2: k_digest<1, 1, TV><<<nblk, 256, 0, ctx->stream>>>(a); break;
            case 3: k_digest<2, 0, TV><<<nblk, 256, 0, ctx->stream>>>(a); break;
            case 4: k_digest<2, 1, TV><<<nblk, 256, 0, ctx->stream>>>(a); break;
            case 5: k_digest<2, 2, TV><<<nblk, 256, 0, ctx->stream>>>(a); break;
        }
        MYQC_HIP(ctx, hipGetLastError());
    }
    return 0;
}

// G(D) into d_G (precision FP64), or G += G32(D) (precision FP32: D is then
// the density change and the result is added to d_G)
int digest_G(myqc_ctx *ctx, const double *d_Din, double *d_Gout, int precision) {
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "no stored integrals (call myqc_eri_build)");
    MYQC_CHECK(ensure_work(ctx));
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    hipEvent_t ev;
    MYQC_CHECK(timer_begin(ctx, T_JK, &ev));
    if (precision == MYQC_FP64) {
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, ctx->d_D, nullptr, n);
        k_zero<double><<<256, 256, 0, ctx->stream>>>(ctx->d_A, nn);
        MYQC_CHECK(launch_digest<double>(ctx, ctx->d_D, ctx->d_A));
        k_assemble<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, d_Gout, n, 0);
    } else {
        if (!ctx->have_eri32) return set_error(ctx, MYQC_E_STATE, "no FP32 integrals stored");
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, nullptr, ctx->d_D32, n);
        k_zero<float><<<256, 256, 0, ctx->stream>>>(ctx->d_A32, nn);
        MYQC_CHECK(launch_digest<float>(ctx, ctx->d_D32, ctx->d_A32));
        k_assemble32<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A32, d_Gout, n);
    }
    MYQC_HIP(ctx, hipGetLastError());
    const double bytes = (precision == MYQC_FP64 ? 8.0 : 4.0) * (double)ctx->n_stored + 16.0 * nn;
    MYQC_CHECK(timer_end(ctx, T_JK, ev, bytes));
    return 0;
}

// ---------------------------------------------------------------------------
// RHF driver
// ---------------------------------------------------------------------------

__global__ void k_scale_cols(double *__restrict__ U, const double *__restrict__ s, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    U[i + (size_t)n * j] *= 1.0 / sqrt(s[j]);
}

__global__ void k_antisym(const double *__restrict__ T, double *__restrict__ E, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    E[i + (size_t)n * j] = T[i + (size_t)n * j] - T[j + (size_t)n * i];
}

__global__ void k_add(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        Z[i] = X[i] + Y[i];
}

__global__ void k_sub(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        Z[i] = X[i] - Y[i];
}

struct ScfBuffers {
    double *H = nullptr, *S = nullptr, *X = nullptr, *F = nullptr, *Fx = nullptr, *D = nullptr, *Dn = nullptr,
           *Dref = nullptr, *G = nullptr, *C = nullptr, *T1 = nullptr, *T2 = nullptr, *eig = nullptr, *work = nullptr;
    double *hF = nullptr, *hE = nullptr;   // DIIS history, m x n^2 each
    int *info = nullptr;
    void release() {
        double **ps[] = {&H, &S, &X, &F, &Fx, &D, &Dn, &Dref, &G, &C, &T1, &T2, &eig, &work, &hF, &hE};
        for (auto p : ps) dfree((void **)p);
        dfree((void **)&info);
    }
};

static int gemm(myqc_ctx *ctx, rocblas_operation ta, rocblas_operation tb, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb, double beta, double *C, int ldc) {
    MYQC_BLAS(ctx, rocblas_dgemm(ctx->blas, ta, tb, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc));
    return 0;
}

// eigen-decomposition of the symmetric n x n A (overwritten by eigenvectors)
static int syevd(myqc_ctx *ctx, ScfBuffers &b, double *A, double *w, int n) {
    MYQC_BLAS(ctx, rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_lower, n, A, n, w, b.work, b.info));
    int info = 0;
    MYQC_HIP(ctx, hipMemcpyAsync(&info, b.info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (info != 0) return set_error(ctx, MYQC_E_HIP, "rocsolver_dsyevd info=%d", info);
    return 0;
}

// C = X C', eigenvalues into eig, from F (symmetric, AO basis)
static int diag_fock(myqc_ctx *ctx, ScfBuffers &b, const double *F, int n) {
    MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, F, n, b.X, n, 0.0, b.T1, n));
    MYQC_CHECK(gemm(ctx, rocblas_operation_transpose, rocblas_operation_none, n, n, n, 1.0, b.X, n, b.T1, n, 0.0, b.T2, n));
    MYQC_CHECK(syevd(ctx, b, b.T2, b.eig, n));
    MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, b.X, n, b.T2, n, 0.0, b.C, n));
    return 0;
}

static int density(myqc_ctx *ctx, const double *C, int nocc, double *D, int n) {
    if (nocc == 0) {
        MYQC_HIP(ctx, hipMemsetAsync(D, 0, sizeof(double) * n * n, ctx->stream));
        return 0;
    }
    return gemm(ctx, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, 2.0, C, n, C, n, 0.0, D, n);
}

static double ddot(myqc_ctx *ctx, int cnt, const double *x, const double *y, int *rc) {
    double r = 0.0;
    rocblas_status s = rocblas_ddot(ctx->blas, cnt, x, 1, y, 1, &r);
    if (s != rocblas_status_success) *rc = set_error(ctx, MYQC_E_HIP, "rocblas_ddot failed (%d)", (int)s);
    return r;
}

static double maxabs_diff(myqc_ctx *ctx, ScfBuffers &b, const double *X, const double *Y, size_t cnt, int *rc) {
    k_sub<<<256, 256, 0, ctx->stream>>>(X, Y, b.T1, cnt);
    rocblas_int idx = 0;
    rocblas_status s = rocblas_idamax(ctx->blas, (int)cnt, b.T1, 1, &idx);
    if (s != rocblas_status_success) {
        *rc = set_error(ctx, MYQC_E_HIP, "rocblas_idamax failed");
        return 0.0;
    }
    double v = 0.0;
    d2h(ctx, &v, b.T1 + (idx - 1), sizeof(double));
    return fabs(v);
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

static int rhf_scf(myqc_ctx *ctx, const double *H, const double *S, int nocc, double e_nuc, const double *C_guess,
                   double e_tol, double d_tol, int max_iter, int precision, double *energy, double *C_out,
                   double *eig_out, int *iters_out) {
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    const int MD = 8;
    ScfBuffers b;
    int rc = 0;
    auto fail = [&](int code) {
        b.release();
        return code;
    };
#define TRY(x)                 \
    do {                       \
        rc = (x);              \
        if (rc) return fail(rc); \
    } while (0)
    double **mats[] = {&b.H, &b.S, &b.X, &b.F, &b.Fx, &b.D, &b.Dn, &b.Dref, &b.G, &b.C, &b.T1, &b.T2};
    for (auto m : mats) TRY(dalloc(ctx, (void **)m, sizeof(double) * nn));
    TRY(dalloc(ctx, (void **)&b.eig, sizeof(double) * n));
    TRY(dalloc(ctx, (void **)&b.work, sizeof(double) * n));
    TRY(dalloc(ctx, (void **)&b.hF, sizeof(double) * nn * MD));
    TRY(dalloc(ctx, (void **)&b.hE, sizeof(double) * nn * MD));
    TRY(dalloc(ctx, (void **)&b.info, sizeof(int)));
    if (hipMemcpyAsync(b.H, H, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(b.S, S, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return fail(set_error(ctx, MYQC_E_HIP, "rhf: upload failed"));
    // X = S^{-1/2}
    if (hipMemcpyAsync(b.T2, b.S, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        return fail(set_error(ctx, MYQC_E_HIP, "rhf: copy failed"));
    TRY(syevd(ctx, b, b.T2, b.eig, n));
    {
        std::vector<double> s(n);
        d2h(ctx, s.data(), b.eig, sizeof(double) * n);
        if (s[0] <= 1e-10) return fail(set_error(ctx, MYQC_E_ARG, "rhf: overlap matrix singular (min eigenvalue %g)", s[0]));
    }
    hipMemcpyAsync(b.T1, b.T2, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
    k_scale_cols<<<grid2(n), 128, 0, ctx->stream>>>(b.T1, b.eig, n);
    TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_transpose, n, n, n, 1.0, b.T1, n, b.T2, n, 0.0, b.X, n));
    // guess
    if (C_guess) {
        if (hipMemcpyAsync(b.C, C_guess, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            return fail(set_error(ctx, MYQC_E_HIP, "rhf: upload failed"));
    } else {
        TRY(diag_fock(ctx, b, b.H, n));   // core Hamiltonian guess (scf.f90:410-512)
    }
    TRY(density(ctx, b.C, nocc, b.D, n));
    std::vector<double> Bm((size_t)MD * MD, 0.0);
    int hist = 0, head = 0;
    double E = 0.0, E_old = 0.0;
    int it = 0;
    bool conv = false;
    // mixed precision: incremental FP32 builds F_k = F_{k-1} + G32(D_k - D_{k-1})
    // from an FP64 reference build; FP64 rebuild every 8 steps and at the end
    bool incr = (precision == MYQC_FP32);
    int since_full = 0;
    bool final_fp64 = false;
    for (it = 1; it <= max_iter; ++it) {
        if (!incr || final_fp64 || since_full == 0 || since_full >= 8) {
            TRY(digest_G(ctx, b.D, b.G, MYQC_FP64));
            hipMemcpyAsync(b.Dref, b.D, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
            since_full = 1;
        } else {
            k_sub<<<256, 256, 0, ctx->stream>>>(b.D, b.Dref, b.T1, nn);
            TRY(digest_G(ctx, b.T1, b.G, MYQC_FP32));   // G += G32(dD)
            hipMemcpyAsync(b.Dref, b.D, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
            ++since_full;
        }
        k_add<<<256, 256, 0, ctx->stream>>>(b.H, b.G, b.F, nn);
        E = 0.5 * (ddot(ctx, (int)nn, b.D, b.F, &rc) + ddot(ctx, (int)nn, b.D, b.H, &rc)) + e_nuc;
        if (rc) return fail(rc);
        // DIIS error X^T (F D S - S D F) X
        TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, b.D, n, b.S, n, 0.0, b.T1, n));
        TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, b.F, n, b.T1, n, 0.0, b.T2, n));
        k_antisym<<<grid2(n), 128, 0, ctx->stream>>>(b.T2, b.T1, n);
        TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, b.T1, n, b.X, n, 0.0, b.T2, n));
        double *eslot = b.hE + nn * head, *fslot = b.hF + nn * head;
        TRY(gemm(ctx, rocblas_operation_transpose, rocblas_operation_none, n, n, n, 1.0, b.X, n, b.T2, n, 0.0, eslot, n));
        hipMemcpyAsync(fslot, b.F, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
        if (hist < MD) ++hist;
        for (int j = 0; j < hist; ++j) {
            const double v = ddot(ctx, (int)nn, eslot, b.hE + nn * j, &rc);
            Bm[head * MD + j] = v;
            Bm[j * MD + head] = v;
        }
        if (rc) return fail(rc);
        const double err = sqrt(fabs(Bm[head * MD + head]));
        // extrapolate
        const double *Fuse = b.F;
        if (hist >= 2) {
            std::vector<double> Bs((size_t)hist * hist), c;
            for (int i = 0; i < hist; ++i)
                for (int j = 0; j < hist; ++j) Bs[i * hist + j] = Bm[i * MD + j];
            if (diis_coeffs(Bs, hist, c)) {
                MYQC_HIP(ctx, hipMemsetAsync(b.Fx, 0, sizeof(double) * nn, ctx->stream));
                for (int i = 0; i < hist; ++i)
                    MYQC_BLAS(ctx, rocblas_daxpy(ctx->blas, (int)nn, &c[i], b.hF + nn * i, 1, b.Fx, 1));
                Fuse = b.Fx;
            }
        }
        head = (head + 1) % MD;
        const double dE = fabs(E - E_old);
        E_old = E;
        TRY(diag_fock(ctx, b, Fuse, n));
        TRY(density(ctx, b.C, nocc, b.Dn, n));
        const double dD = maxabs_diff(ctx, b, b.Dn, b.D, nn, &rc);
        if (rc) return fail(rc);
        hipMemcpyAsync(b.D, b.Dn, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
        if (it > 1 && dE < e_tol && dD < d_tol && err < 100 * d_tol) {
            if (incr && !final_fp64) {
                final_fp64 = true;   // confirm with FP64 builds
                continue;
            }
            conv = true;
            break;
        }
    }
    // report E[D] with the orbitals of the plain Fock matrix of the last step
    // (the reference prints E of the density it diagonalises, scf.f90:775-848)
    TRY(digest_G(ctx, b.D, b.G, MYQC_FP64));
    k_add<<<256, 256, 0, ctx->stream>>>(b.H, b.G, b.F, nn);
    E = 0.5 * (ddot(ctx, (int)nn, b.D, b.F, &rc) + ddot(ctx, (int)nn, b.D, b.H, &rc)) + e_nuc;
    if (rc) return fail(rc);
    TRY(diag_fock(ctx, b, b.F, n));
    if (energy) *energy = E;
    if (C_out) d2h(ctx, C_out, b.C, sizeof(double) * nn);
    if (eig_out) d2h(ctx, eig_out, b.eig, sizeof(double) * n);
    if (iters_out) *iters_out = it > max_iter ? max_iter : it;
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b.release();
#undef TRY
    if (!conv) return set_error(ctx, MYQC_E_NOCONV, "RHF did not converge in %d iterations", max_iter);
    return 0;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_fock_G(myqc_ctx *ctx, const double *D, double *G, int precision) {
    if (!ctx || !D || !G) return MYQC_E_ARG;
    if (precision != MYQC_FP64 && precision != MYQC_FP32)
        return set_error(ctx, MYQC_E_ARG, "myqc_fock_G: precision %d", precision);
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "myqc_fock_G: call myqc_eri_build first");
    MYQC_CHECK(ensure_work(ctx));
    const size_t nn = (size_t)ctx->n * ctx->n;
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_G, D, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    double *d_out = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_out, sizeof(double) * nn));
    int rc = 0;
    if (precision == MYQC_FP32 && hipMemsetAsync(d_out, 0, sizeof(double) * nn, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "memset failed");
    if (rc == 0) rc = digest_G(ctx, ctx->d_G, d_out, precision);
    if (rc == 0 && d2h(ctx, G, d_out, sizeof(double) * nn) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_fock_G: copy failed");
    dfree((void **)&d_out);
    return rc;
}

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
    return rhf_scf(ctx, H, S, nocc, e_nuc, C_guess, e_tol, d_tol, max_iter, precision, energy, C, eig, iterations);
}

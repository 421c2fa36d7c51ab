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

int sp2_purify(myqc_ctx *ctx, const double *Fo, double *P, double *work, double *lohi, double *slots, int n, int nocc,
               int *ok);

template <int C> struct PL;
template <> struct PL<0> { static constexpr int a = 0, b = 0; };
template <> struct PL<1> { static constexpr int a = 1, b = 0; };
template <> struct PL<2> { static constexpr int a = 1, b = 1; };

template <typename TV>
struct DigArgs {
    const int *meta0, *meta1, *meta2;         // ket AOs by class and Schwarz rank
    const TV *dk0, *dk1, *dk2;                // ket density blocks by class and rank (1, 3, 9 values)
    const int *wg;                            // [w] = {A, u0, u1}
    const int *unit;                          // [u] = {B, c0, c1, c2}
    const long long *unit_off;                // [u][3]
    const int *shell;                         // [s] = {ao, L, wg0, wg1}
    const TV *x;                              // expanded values
    const TV *D;                              // symmetric density (or change), n x n
    TV *partial;                              // [w][3][n]
    int n;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One ket of rank r in class CK against the bra (A, B): J into the thread's
// accumulator, K into the workgroup's LDS rows of A.  s_drow holds D's rows
// of shell B.
template <int LA, int LB, int CK, typename TV>
__device__ __forceinline__ void digest_ket(const DigArgs<TV> &a, TV *s_rows, const TV *s_drow, int r, long long off,
                                           TV (&J)[3][3]) {
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(PL<CK>::a), nd = ncart(PL<CK>::b);
    constexpr int BLK = na * nb * nc * nd;
    const int n = a.n;
    const int m = (CK == 0 ? a.meta0 : CK == 1 ? a.meta1 : a.meta2)[r];
    const int aoC = m & 0x7fff, aoD = (m >> 15) & 0x7fff, same = m >> 30;
    const TV *__restrict__ v = a.x + off + (long long)r * BLK;
    TV x[BLK];
#pragma unroll
    for (int e = 0; e < BLK; ++e) x[e] = v[e];
    const TV *dk = (CK == 0 ? a.dk0 : CK == 1 ? a.dk1 : a.dk2) + (size_t)r * (nc * nd);
    TV dq[nc][nd];
#pragma unroll
    for (int c = 0; c < nc; ++c)
#pragma unroll
        for (int d = 0; d < nd; ++d) dq[c][d] = dk[c * nd + d];
    const TV fj = same ? (TV)1 : (TV)2;
    TV dbd[nb][nd], dbc[nb][nc];
#pragma unroll
    for (int b = 0; b < nb; ++b) {
#pragma unroll
        for (int d = 0; d < nd; ++d) dbd[b][d] = s_drow[b * n + aoD + d];
#pragma unroll
        for (int c = 0; c < nc; ++c) dbc[b][c] = s_drow[b * n + aoC + c];
    }
#pragma unroll
    for (int ia = 0; ia < na; ++ia) {
#pragma unroll
        for (int ib = 0; ib < nb; ++ib) {
            TV js = 0;
#pragma unroll
            for (int c = 0; c < nc; ++c)
#pragma unroll
                for (int d = 0; d < nd; ++d) js += dq[c][d] * x[((ia * nb + ib) * nc + c) * nd + d];
            J[ia][ib] += fj * js;
        }
        // G(i, k) -= 1/2 sum_{b,d} D(b,d) (i b|k d)
#pragma unroll
        for (int c = 0; c < nc; ++c) {
            TV ks = 0;
#pragma unroll
            for (int ib = 0; ib < nb; ++ib)
#pragma unroll
                for (int d = 0; d < nd; ++d) ks += dbd[ib][d] * x[((ia * nb + ib) * nc + c) * nd + d];
            atomicAdd(&s_rows[ia * n + aoC + c], (TV)(-0.5) * ks);
        }
        // the ket's other orientation: G(i, l) -= 1/2 sum_{b,c} D(b,c) (i b|k l)
        if (!same) {
#pragma unroll
            for (int d = 0; d < nd; ++d) {
                TV ks = 0;
#pragma unroll
                for (int ib = 0; ib < nb; ++ib)
#pragma unroll
                    for (int c = 0; c < nc; ++c) ks += dbc[ib][c] * x[((ia * nb + ib) * nc + c) * nd + d];
                atomicAdd(&s_rows[ia * n + aoD + d], (TV)(-0.5) * ks);
            }
        }
    }
}

// All kets of the unit (A, B): the workgroup's threads stride over the
// concatenated ket-class segments.
template <int LA, int LB, typename TV>
__device__ __forceinline__ void digest_unit(const DigArgs<TV> &a, TV *s_rows, const TV *s_drow, int u, int aoB) {
    constexpr int na = ncart(LA), nb = ncart(LB);
    const int c0 = a.unit[4 * u + 1], c1 = a.unit[4 * u + 2], c2 = a.unit[4 * u + 3];
    const long long o0 = a.unit_off[3 * u], o1 = a.unit_off[3 * u + 1], o2 = a.unit_off[3 * u + 2];
    TV J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const int total = c0 + c1 + c2;
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        if (k < c0) digest_ket<LA, LB, 0>(a, s_rows, s_drow, k, o0, J);
        else if (k < c0 + c1) digest_ket<LA, LB, 1>(a, s_rows, s_drow, k - c0, o1, J);
        else digest_ket<LA, LB, 2>(a, s_rows, s_drow, k - c0 - c1, o2, J);
    }
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int ia = 0; ia < na; ++ia)
#pragma unroll
        for (int ib = 0; ib < nb; ++ib) {
            const TV t = wave_sum(J[ia][ib]);
            if (lane == 0) atomicAdd(&s_rows[ia * a.n + aoB + ib], t);
        }
}

// One workgroup per block of units of one row shell A; writes the rows of A
// it accumulated to its partial slot.
template <typename TV>
__global__ __launch_bounds__(256) void k_digest(DigArgs<TV> a) {
    extern __shared__ unsigned char smem[];
    const int n = a.n;
    TV *s_rows = reinterpret_cast<TV *>(smem);
    TV *s_drow = s_rows + 3 * n;
    const int w = blockIdx.x;
    const int A = a.wg[3 * w], u0 = a.wg[3 * w + 1], u1 = a.wg[3 * w + 2];
    const int LA = a.shell[4 * A + 1];
    const int na = LA ? 3 : 1;
    for (int i = threadIdx.x; i < na * n; i += blockDim.x) s_rows[i] = 0;
    for (int u = u0; u < u1; ++u) {
        const int B = a.unit[4 * u];
        const int aoB = a.shell[4 * B], LB = a.shell[4 * B + 1];
        const int nb = LB ? 3 : 1;
        __syncthreads();   // previous unit done with s_drow
        for (int i = threadIdx.x; i < nb * n; i += blockDim.x) {
            const int b = i / n, j = i - b * n;
            s_drow[i] = a.D[(aoB + b) + (size_t)n * j];
        }
        __syncthreads();
        if (LA == 0) {
            if (LB == 0) digest_unit<0, 0>(a, s_rows, s_drow, u, aoB);
            else digest_unit<0, 1>(a, s_rows, s_drow, u, aoB);
        } else {
            if (LB == 0) digest_unit<1, 0>(a, s_rows, s_drow, u, aoB);
            else digest_unit<1, 1>(a, s_rows, s_drow, u, aoB);
        }
    }
    __syncthreads();
    TV *dst = a.partial + (size_t)w * 3 * n;
    for (int i = threadIdx.x; i < na * n; i += blockDim.x) dst[i] = s_rows[i];
}

// G(i,j) = sum over the workgroups of shell(i) of their partial rows
template <typename TV>
__global__ void k_gather_rows(const TV *__restrict__ partial, const int *__restrict__ shell,
                              const int *__restrict__ ao_shell, int n, double *__restrict__ Graw) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= n) return;
    const int A = ao_shell[i];
    const int ia = i - shell[4 * A], w0 = shell[4 * A + 2], w1 = shell[4 * A + 3];
    double s = 0.0;
    for (int w = w0; w < w1; ++w) s += (double)partial[((size_t)w * 3 + ia) * n + j];
    Graw[i + (size_t)n * j] = s;
}

// the kets' density blocks in Schwarz-rank order: dk[r*(nc*nd) + c*nd + d] =
// D(ao_a + c, ao_b + d) of pair sorted[r]
template <typename TV>
__global__ void k_pack_kets(const ShellPair *__restrict__ pairs, const int *__restrict__ sorted, int np, int nc, int nd,
                            const TV *__restrict__ D, int n, TV *__restrict__ dk) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    const ShellPair sp = pairs[sorted[r]];
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < nd; ++d) dk[(size_t)r * nc * nd + c * nd + d] = D[(sp.ao_a + c) + (size_t)n * (sp.ao_b + d)];
}

// --- small elementwise kernels ------------------------------------------------

__global__ void k_add(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, size_t cnt);

__global__ void k_symmetrize(const double *__restrict__ D, double *__restrict__ Ds, float *__restrict__ Ds32, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n) return;
    const double v = 0.5 * (D[i + (size_t)n * j] + D[j + (size_t)n * i]);
    if (Ds) Ds[i + (size_t)n * j] = v;
    if (Ds32) Ds32[i + (size_t)n * j] = (float)v;
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
static int launch_digest(myqc_ctx *ctx, const TV *D, const TV *x, TV *dk, TV *partial, double *Graw) {
    const int np0 = (int)ctx->pair_ids[0].size(), np1 = (int)ctx->pair_ids[1].size(),
              np2 = (int)ctx->pair_ids[2].size();
    TV *dk0 = dk, *dk1 = dk + np0, *dk2 = dk + np0 + 3 * (size_t)np1;
    const int nps[3] = {np0, np1, np2};
    TV *dks[3] = {dk0, dk1, dk2};
    for (int c = 0; c < 3; ++c)
        if (nps[c] > 0)
            k_pack_kets<TV><<<(nps[c] + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_sorted[c], nps[c],
                                                                            c ? 3 : 1, c == 2 ? 3 : 1, D, ctx->n,
                                                                            dks[c]);
    if (ctx->n_wg > 0) {
        DigArgs<TV> a;
        a.meta0 = ctx->d_meta[0];
        a.meta1 = ctx->d_meta[1];
        a.meta2 = ctx->d_meta[2];
        a.dk0 = dk0;
        a.dk1 = dk1;
        a.dk2 = dk2;
        a.wg = ctx->d_wg;
        a.unit = ctx->d_unit;
        a.unit_off = ctx->d_unit_off;
        a.shell = ctx->d_shell;
        a.x = x;
        a.D = D;
        a.partial = partial;
        a.n = ctx->n;
        const size_t shm = sizeof(TV) * 6 * ctx->n;
        k_digest<TV><<<ctx->n_wg, 256, shm, ctx->stream>>>(a);
    }
    k_gather_rows<TV><<<dim3((ctx->n + 127) / 128, ctx->n), 128, 0, ctx->stream>>>(partial, ctx->d_shell,
                                                                                  ctx->d_ao_shell, ctx->n, Graw);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// G(D) into d_Gout (precision FP64), or d_Gout += G32(D) (precision FP32:
// D is then the density change)
int digest_G(myqc_ctx *ctx, const double *d_Din, double *d_Gout, int precision) {
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "no stored integrals (call myqc_eri_build)");
    MYQC_CHECK(ensure_work(ctx));
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    if ((size_t)6 * n * sizeof(double) > 64 * 1024)
        return set_error(ctx, MYQC_E_ARG, "digestion rows for n=%d exceed 64 KB of LDS", n);
    hipEvent_t ev;
    MYQC_CHECK(timer_begin(ctx, T_JK, &ev));
    if (precision == MYQC_FP64) {
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, ctx->d_D, nullptr, n);
        MYQC_CHECK(launch_digest<double>(ctx, ctx->d_D, ctx->d_x64, ctx->d_Dpair, ctx->d_partial, ctx->d_A));
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, d_Gout, nullptr, n);
    } else {
        if (!ctx->have_eri32) return set_error(ctx, MYQC_E_STATE, "no FP32 integrals stored");
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, nullptr, ctx->d_D32, n);
        MYQC_CHECK(launch_digest<float>(ctx, ctx->d_D32, ctx->d_x32, ctx->d_Dpair32, ctx->d_partial32, ctx->d_A));
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, ctx->d_G, nullptr, n);
        k_add<<<256, 256, 0, ctx->stream>>>(d_Gout, ctx->d_G, d_Gout, nn);
    }
    MYQC_HIP(ctx, hipGetLastError());
    const double bytes = (precision == MYQC_FP64 ? 8.0 : 4.0) * (double)ctx->n_expanded + 16.0 * nn;
    MYQC_CHECK(timer_end(ctx, T_JK, ev, bytes));
    return 0;
}

// ---------------------------------------------------------------------------
// RHF driver
// ---------------------------------------------------------------------------

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

__global__ void k_add(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        Z[i] = X[i] + Y[i];
}

__global__ void k_sub(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        Z[i] = X[i] - Y[i];
}

struct ScfBuffers {
    double *H = nullptr, *S = nullptr, *X = nullptr, *Xi = nullptr, *F = nullptr, *Fo = nullptr, *Fx = nullptr,
           *D = nullptr, *Dn = nullptr, *Do = nullptr, *Don = nullptr, *Dref = nullptr, *G = nullptr, *C = nullptr,
           *T1 = nullptr, *T2 = nullptr, *P2 = nullptr, *eig = nullptr, *work = nullptr, *dsc = nullptr,
           *lohi = nullptr;
    double *hF = nullptr, *hE = nullptr;   // DIIS history, m x n^2 each
    int *info = nullptr;
    void release() {
        double **ps[] = {&H, &S, &X, &Xi, &F, &Fo, &Fx, &D, &Dn, &Do, &Don, &Dref, &G, &C, &T1, &T2, &P2, &eig, &work,
                         &dsc, &lohi, &hF, &hE};
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

// C = X C' and orbital energies from the orthonormal-basis Fock matrix Fo
static int diag_orth(myqc_ctx *ctx, ScfBuffers &b, const double *Fo, int n) {
    MYQC_HIP(ctx, hipMemcpyAsync(b.T2, Fo, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    MYQC_CHECK(syevd(ctx, b, b.T2, b.eig, n));
    MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, b.X, n, b.T2, n, 0.0, b.C, n));
    return 0;
}

// Fo = X^T F X
static int to_orth(myqc_ctx *ctx, ScfBuffers &b, const double *F, double *Fo, int n) {
    MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, F, n, b.X, n, 0.0, b.T1, n));
    MYQC_CHECK(gemm(ctx, rocblas_operation_transpose, rocblas_operation_none, n, n, n, 1.0, b.X, n, b.T1, n, 0.0, Fo, n));
    return 0;
}

// D = 2 X Do X^T
static int to_ao_density(myqc_ctx *ctx, ScfBuffers &b, const double *Do, double *D, int n) {
    MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_transpose, n, n, n, 1.0, Do, n, b.X, n, 0.0, b.T1, n));
    MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 2.0, b.X, n, b.T1, n, 0.0, D, n));
    return 0;
}

// Orthonormal-basis density projector onto the nocc lowest levels of Fo by
// purification; falls back to diagonalisation if it does not converge.
static int purify(myqc_ctx *ctx, ScfBuffers &b, const double *Fo, double *P, int n, int nocc) {
    if (nocc == 0) {
        MYQC_HIP(ctx, hipMemsetAsync(P, 0, sizeof(double) * n * n, ctx->stream));
        return 0;
    }
    int ok = 0;
    MYQC_CHECK(sp2_purify(ctx, Fo, P, b.P2, b.lohi, b.dsc, n, nocc, &ok));
    ++ctx->scf_purify;
    if (!ok) {
        // no clean gap at this step: occupy the nocc lowest eigenvectors
        ++ctx->scf_diag;
        MYQC_CHECK(diag_orth(ctx, b, Fo, n));
        MYQC_CHECK(gemm(ctx, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, 1.0, b.T2, n, b.T2, n, 0.0,
                        P, n));
    }
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
#define TRY(x)                   \
    do {                         \
        rc = (x);                \
        if (rc) return fail(rc); \
    } while (0)
    double **mats[] = {&b.H, &b.S, &b.X, &b.Xi, &b.F, &b.Fo, &b.Fx, &b.D, &b.Dn, &b.Do, &b.Don, &b.Dref, &b.G, &b.C,
                       &b.T1, &b.T2, &b.P2};
    for (auto m : mats) TRY(dalloc(ctx, (void **)m, sizeof(double) * nn));
    TRY(dalloc(ctx, (void **)&b.eig, sizeof(double) * n));
    TRY(dalloc(ctx, (void **)&b.work, sizeof(double) * n));
    TRY(dalloc(ctx, (void **)&b.dsc, sizeof(double) * 8));
    TRY(dalloc(ctx, (void **)&b.lohi, sizeof(double) * 2 * n));
    TRY(dalloc(ctx, (void **)&b.hF, sizeof(double) * nn * MD));
    TRY(dalloc(ctx, (void **)&b.hE, sizeof(double) * nn * MD));
    TRY(dalloc(ctx, (void **)&b.info, sizeof(int)));
    if (hipMemcpyAsync(b.H, H, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(b.S, S, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return fail(set_error(ctx, MYQC_E_HIP, "rhf: upload failed"));
    // X = S^{-1/2}, Xi = S^{1/2} (symmetric orthogonalisation)
    if (hipMemcpyAsync(b.T2, b.S, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
        return fail(set_error(ctx, MYQC_E_HIP, "rhf: copy failed"));
    TRY(syevd(ctx, b, b.T2, b.eig, n));
    {
        std::vector<double> s(n);
        d2h(ctx, s.data(), b.eig, sizeof(double) * n);
        if (s[0] <= 1e-10) return fail(set_error(ctx, MYQC_E_ARG, "rhf: overlap matrix singular (min eigenvalue %g)", s[0]));
    }
    hipMemcpyAsync(b.T1, b.T2, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
    k_scale_cols<<<grid2(n), 128, 0, ctx->stream>>>(b.T1, b.eig, n, -0.5);
    TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_transpose, n, n, n, 1.0, b.T1, n, b.T2, n, 0.0, b.X, n));
    hipMemcpyAsync(b.T1, b.T2, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
    k_scale_cols<<<grid2(n), 128, 0, ctx->stream>>>(b.T1, b.eig, n, 0.5);
    TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_transpose, n, n, n, 1.0, b.T1, n, b.T2, n, 0.0, b.Xi, n));
    // guess density
    if (C_guess) {
        if (hipMemcpyAsync(b.C, C_guess, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            return fail(set_error(ctx, MYQC_E_HIP, "rhf: upload failed"));
        TRY(density(ctx, b.C, nocc, b.D, n));
        // Do = S^{1/2} D S^{1/2} / 2
        TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, b.D, n, b.Xi, n, 0.0, b.T1, n));
        TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 0.5, b.Xi, n, b.T1, n, 0.0, b.Do, n));
    } else {
        // core Hamiltonian guess (scf.f90:410-512)
        TRY(to_orth(ctx, b, b.H, b.Fo, n));
        TRY(purify(ctx, b, b.Fo, b.Do, n, nocc));
        TRY(to_ao_density(ctx, b, b.Do, b.D, n));
    }
    std::vector<double> Bm((size_t)MD * MD, 0.0);
    int hist = 0, head = 0;
    double E = 0.0, E_old = 0.0;
    int it = 0;
    bool conv = false;
    // mixed precision: incremental FP32 builds F_k = F_{k-1} + G32(D_k - D_{k-1})
    // from an FP64 reference build; FP64 rebuild every 8 steps and at the end
    const bool incr = (precision == MYQC_FP32);
    int since_full = 0;
    bool final_fp64 = false;
    for (it = 1; it <= max_iter; ++it) {
        if (!incr || final_fp64 || since_full == 0 || since_full >= 8) {
            TRY(digest_G(ctx, b.D, b.G, MYQC_FP64));
            since_full = 1;
        } else {
            k_sub<<<256, 256, 0, ctx->stream>>>(b.D, b.Dref, b.T1, nn);
            TRY(digest_G(ctx, b.T1, b.G, MYQC_FP32));   // G += G32(dD)
            ++since_full;
        }
        hipMemcpyAsync(b.Dref, b.D, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream);
        k_add<<<256, 256, 0, ctx->stream>>>(b.H, b.G, b.F, nn);
        E = 0.5 * (ddot(ctx, (int)nn, b.D, b.F, &rc) + ddot(ctx, (int)nn, b.D, b.H, &rc)) + e_nuc;
        if (rc) return fail(rc);
        // orthonormal basis: Fo = X^T F X, DIIS error Fo Do - Do Fo
        double *fslot = b.hF + nn * head, *eslot = b.hE + nn * head;
        TRY(to_orth(ctx, b, b.F, fslot, n));
        TRY(gemm(ctx, rocblas_operation_none, rocblas_operation_none, n, n, n, 1.0, fslot, n, b.Do, n, 0.0, b.T2, n));
        k_antisym<<<grid2(n), 128, 0, ctx->stream>>>(b.T2, eslot, n);
        if (hist < MD) ++hist;
        for (int j = 0; j < hist; ++j) {
            const double v = ddot(ctx, (int)nn, eslot, b.hE + nn * j, &rc);
            Bm[head * MD + j] = v;
            Bm[j * MD + head] = v;
        }
        if (rc) return fail(rc);
        const double err = sqrt(fabs(Bm[head * MD + head]));
        const double *Fuse = fslot;
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
        TRY(purify(ctx, b, Fuse, b.Don, n, nocc));
        TRY(to_ao_density(ctx, b, b.Don, b.Dn, n));
        const double dD = maxabs_diff(ctx, b, b.Dn, b.D, nn, &rc);
        if (rc) return fail(rc);
        std::swap(b.D, b.Dn);
        std::swap(b.Do, b.Don);
        if (it > 1 && dE < e_tol && dD < d_tol && err < 100 * d_tol) {
            if (incr && !final_fp64) {
                final_fp64 = true;   // confirm with FP64 builds
                continue;
            }
            conv = true;
            break;
        }
    }
    // report E[D] with the orbitals of its plain Fock matrix (the reference
    // prints E of the density it diagonalises, scf.f90:775-848)
    TRY(digest_G(ctx, b.D, b.G, MYQC_FP64));
    k_add<<<256, 256, 0, ctx->stream>>>(b.H, b.G, b.F, nn);
    E = 0.5 * (ddot(ctx, (int)nn, b.D, b.F, &rc) + ddot(ctx, (int)nn, b.D, b.H, &rc)) + e_nuc;
    if (rc) return fail(rc);
    TRY(to_orth(ctx, b, b.F, b.Fo, n));
    TRY(diag_orth(ctx, b, b.Fo, n));
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
    ctx->scf_purify = ctx->scf_diag = 0;
    return rhf_scf(ctx, H, S, nocc, e_nuc, C_guess, e_tol, d_tol, max_iter, precision, energy, C, eig, iterations);
}

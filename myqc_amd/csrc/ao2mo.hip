// ao2mo.hip -- the RHF (ia|jb) four-index transform and the MP2 energy.
//
// The transform keeps the reference's four quarter steps
// (idx1..idx4_trans, src/ao2mo/ao2mo.f90:1306-1439) but runs each as one
// dense GEMM (or a strided batch of GEMMs) over the device-resident dense AO
// integrals; the FP64 form uses rocBLAS DGEMM (plain library GEMMs), the
// split-precision form the hand-written MFMA kernels in mfma_gemm.hip, with
// steps 1 and 2 fused so that the o n^3 intermediate is never stored.
// The MP2 reduction restates mp2_rhf (src/mp2/mp2.f90:79-144) as one pass
// with a deterministic two-stage sum.
#include <cmath>
#include <vector>

#include "context.h"

namespace myqc {
int gemm_bf16x3(myqc_ctx *ctx, bool transA, int m, int n, int k, const double *A, int lda, long long strideA,
                const double *B, int ldb, long long strideB, double *C, int ldc, long long strideC, int batch);
int ao2mo12_bf16x3(myqc_ctx *ctx, int o, int n, const double *C, const double *XX, double *M);
}

namespace myqc {

// out[((i*o + j)*nv + b)*nv + a] = O[i + o*(a + nv*(j + o*b))]
__global__ void k_records(const double *__restrict__ O, double *__restrict__ out, int o, int nv) {
    const size_t total = (size_t)o * o * nv * nv;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        size_t x = r;
        const int a = (int)(x % nv); x /= nv;
        const int b = (int)(x % nv); x /= nv;
        const int j = (int)(x % o); x /= o;
        const int i = (int)x;
        out[r] = O[i + (size_t)o * (a + (size_t)nv * (j + (size_t)o * b))];
    }
}

// per-block partial sums of E(AB) (all i,j,a,b) and E(AA) (i<j, a<b)
__global__ __launch_bounds__(256) void k_mp2(const double *__restrict__ O, const double *__restrict__ eig, int o, int nv,
                                             double *__restrict__ part) {
    const size_t total = (size_t)o * nv * o * nv;
    double s1 = 0.0, s2 = 0.0;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        size_t x = r;
        const int i = (int)(x % o); x /= o;
        const int a = (int)(x % nv); x /= nv;
        const int j = (int)(x % o); x /= o;
        const int b = (int)x;
        const double t = O[r];
        const double den = eig[i] + eig[j] - eig[a + o] - eig[b + o];
        s2 += t * t / den;
        if (i < j && a < b) {
            const double u = t - O[i + (size_t)o * (b + (size_t)nv * (j + (size_t)o * a))];
            s1 += u * u / den;
        }
    }
    __shared__ double sh1[256], sh2[256];
    sh1[threadIdx.x] = s1;
    sh2[threadIdx.x] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh1[threadIdx.x] += sh1[threadIdx.x + w];
            sh2[threadIdx.x] += sh2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sh1[0];
        part[2 * blockIdx.x + 1] = sh2[0];
    }
}

static int dgemm(myqc_ctx *ctx, rocblas_operation ta, int m, int n, int k, const double *A, int lda, long long sA,
                 const double *B, int ldb, long long sB, double *C, int ldc, long long sC, int batch, int precision) {
    if (precision == MYQC_BF16X3)
        return gemm_bf16x3(ctx, ta == rocblas_operation_transpose, m, n, k, A, lda, sA, B, ldb, sB, C, ldc, sC, batch);
    const double one = 1.0, zero = 0.0;
    if (batch == 1)
        MYQC_BLAS(ctx, rocblas_dgemm(ctx->blas, ta, rocblas_operation_none, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc));
    else
        MYQC_BLAS(ctx, rocblas_dgemm_strided_batched(ctx->blas, ta, rocblas_operation_none, m, n, k, &one, A, lda, sA, B,
                                                     ldb, sB, &zero, C, ldc, sC, batch));
    return 0;
}

static int ao2mo(myqc_ctx *ctx, const double *C_host, int o, int precision) {
    const int n = ctx->n, nv = n - o;
    const size_t N = n;
    double *dXX = nullptr, *dC = nullptr, *dL = nullptr, *dM = nullptr, *dN = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        dfree((void **)&dXX);
        dfree((void **)&dC);
        dfree((void **)&dL);
        dfree((void **)&dM);
        dfree((void **)&dN);
    };
#define TRY(x)          \
    do {                \
        rc = (x);       \
        if (rc) {       \
            cleanup();  \
            return rc;  \
        }               \
    } while (0)
    TRY(dalloc(ctx, (void **)&dC, sizeof(double) * N * N));
    if (hipMemcpyAsync(dC, C_host, sizeof(double) * N * N, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        cleanup();
        return set_error(ctx, MYQC_E_HIP, "ao2mo: upload failed");
    }
    TRY(dalloc(ctx, (void **)&dXX, sizeof(double) * N * N * N * N));
    // dense AO integrals at the stored build's screening threshold
    TRY(eri_fill_dense(ctx, ctx->have_eri ? ctx->eri_tol : 0.0, dXX));
    hipEvent_t ev;
    TRY(timer_begin(ctx, T_AO2MO, &ev));
    if (precision == MYQC_BF16X3) {
        // 1 + 2 fused per (l,d) block: L never leaves the workgroup
        TRY(dalloc(ctx, (void **)&dM, sizeof(double) * (size_t)o * nv * N * N));
        TRY(ao2mo12_bf16x3(ctx, o, n, dC, dXX, dM));
        dfree((void **)&dXX);   // (hipFree waits for the kernel)
    } else {
        // 1: L(i,v,l,d) = sum_u C(u,i) XX(u,v,l,d)          (idx1_trans)
        TRY(dalloc(ctx, (void **)&dL, sizeof(double) * (size_t)o * N * N * N));
        TRY(dgemm(ctx, rocblas_operation_transpose, o, n * n * n, n, dC, n, 0, dXX, n, 0, dL, o, 0, 1, precision));
        dfree((void **)&dXX);
        // 2: M(i,a,l,d) = sum_v L(i,v,l,d) C(v,a)            (idx2_trans)
        TRY(dalloc(ctx, (void **)&dM, sizeof(double) * (size_t)o * nv * N * N));
        TRY(dgemm(ctx, rocblas_operation_none, o, nv, n, dL, o, (long long)o * n, dC + N * o, n, 0, dM, o,
                  (long long)o * nv, n * n, precision));
        dfree((void **)&dL);
    }
    // 3: N(i,a,j,d) = sum_l M(i,a,l,d) C(l,j)            (idx3_trans)
    TRY(dalloc(ctx, (void **)&dN, sizeof(double) * (size_t)o * nv * o * N));
    TRY(dgemm(ctx, rocblas_operation_none, o * nv, o, n, dM, o * nv, (long long)o * nv * n, dC, n, 0, dN, o * nv,
              (long long)o * nv * o, n, precision));
    dfree((void **)&dM);
    // 4: O(i,a,j,b) = sum_d N(i,a,j,d) C(d,b)            (idx4_trans)
    const size_t oe = (size_t)o * nv * o * nv;
    if (ctx->iajb_elems != oe) {
        TRY(dalloc(ctx, (void **)&ctx->d_iajb, sizeof(double) * (oe ? oe : 1)));
        ctx->iajb_elems = oe;
    }
    TRY(dgemm(ctx, rocblas_operation_none, o * nv * o, nv, n, dN, o * nv * o, 0, dC + N * o, n, 0, ctx->d_iajb,
              o * nv * o, 0, 1, precision));
    const double flops = 2.0 * ((double)o * N * N * N * N + (double)o * nv * N * N * N + (double)o * o * nv * N * N +
                                (double)o * o * nv * nv * N);
    TRY(timer_end(ctx, T_AO2MO, ev, flops));
    ctx->iajb_nocc = o;
    ctx->iajb_nv = nv;
    cleanup();
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
#undef TRY
    return 0;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_ao2mo_ovov(myqc_ctx *ctx, const double *C, int nocc, double *iajb, int precision) {
    if (!ctx || !C) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_ao2mo_ovov: no system set");
    if (nocc < 1 || nocc >= ctx->n) return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_ovov: nocc=%d", nocc);
    if (precision != MYQC_FP64 && precision != MYQC_BF16X3)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_ovov: precision %d", precision);
    const size_t n = ctx->n;
    if (n * n * n * n * 8 > (size_t)200 << 30)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_ovov: n=%zu needs more than 200 GB of dense AO integrals", n);
    MYQC_CHECK(ao2mo(ctx, C, nocc, precision));
    if (iajb) {
        const int o = nocc, nv = ctx->n - nocc;
        const size_t oe = (size_t)o * o * nv * nv;
        double *d_rec = nullptr;
        MYQC_CHECK(dalloc(ctx, (void **)&d_rec, sizeof(double) * (oe ? oe : 1)));
        k_records<<<1024, 256, 0, ctx->stream>>>(ctx->d_iajb, d_rec, o, nv);
        int rc = 0;
        if (d2h(ctx, iajb, d_rec, sizeof(double) * oe) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "myqc_ao2mo_ovov: download failed");
        dfree((void **)&d_rec);
        return rc;
    }
    return 0;
}

extern "C" int myqc_mp2_rhf(myqc_ctx *ctx, const double *eig, int nocc, double *e_aa, double *e_ab, double *e_tot) {
    if (!ctx || !eig) return MYQC_E_ARG;
    if (!ctx->d_iajb || ctx->iajb_nocc != nocc)
        return set_error(ctx, MYQC_E_STATE, "myqc_mp2_rhf: no (ia|jb) for nocc=%d (call myqc_ao2mo_ovov)", nocc);
    const int o = nocc, nv = ctx->iajb_nv, n = o + nv;
    double *d_eig = nullptr, *d_part = nullptr;
    const int nblk = 512;
    MYQC_CHECK(dalloc(ctx, (void **)&d_eig, sizeof(double) * n));
    int rc = dalloc(ctx, (void **)&d_part, sizeof(double) * 2 * nblk);
    if (rc == 0 && hipMemcpyAsync(d_eig, eig, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "upload failed");
    std::vector<double> part(2 * nblk);
    if (rc == 0) {
        hipEvent_t ev;
        rc = timer_begin(ctx, T_MP2, &ev);
        k_mp2<<<nblk, 256, 0, ctx->stream>>>(ctx->d_iajb, d_eig, o, nv, d_part);
        if (rc == 0) rc = timer_end(ctx, T_MP2, ev, 8.0 * (double)o * o * nv * nv);
        if (rc == 0 && d2h(ctx, part.data(), d_part, sizeof(double) * 2 * nblk) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "download failed");
    }
    dfree((void **)&d_eig);
    dfree((void **)&d_part);
    if (rc) return rc;
    double s1 = 0.0, s2 = 0.0;
    for (int b = 0; b < nblk; ++b) {
        s1 += part[2 * b];
        s2 += part[2 * b + 1];
    }
    if (e_aa) *e_aa = s1;
    if (e_ab) *e_ab = s2;
    if (e_tot) *e_tot = 2.0 * s1 + s2;
    return 0;
}

// (ia|jb) records of an ijab_AB file (record (i,j) i-major, (a,b) a fastest)
// to the device's [i][a][j][b] layout
__global__ void k_from_records(const double *__restrict__ rec, double *__restrict__ O, int o, int nv) {
    const size_t total = (size_t)o * o * nv * nv;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        size_t x = r;
        const int a = (int)(x % nv); x /= nv;
        const int b = (int)(x % nv); x /= nv;
        const int j = (int)(x % o); x /= o;
        const int i = (int)x;
        O[i + (size_t)o * (a + (size_t)nv * (j + (size_t)o * b))] = rec[r];
    }
}

extern "C" int myqc_mp2_rhf_records(myqc_ctx *ctx, const double *iajb_records, const double *eig, int nocc, int nvrt,
                                    double *e_aa, double *e_ab, double *e_tot) {
    if (!ctx || !iajb_records || !eig) return MYQC_E_ARG;
    if (nocc < 1 || nvrt < 1) return set_error(ctx, MYQC_E_ARG, "myqc_mp2_rhf_records: nocc=%d nvrt=%d", nocc, nvrt);
    const size_t oe = (size_t)nocc * nocc * nvrt * nvrt;
    double *d_rec = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_rec, sizeof(double) * oe));
    if (ctx->iajb_elems != oe) {
        int rc = dalloc(ctx, (void **)&ctx->d_iajb, sizeof(double) * oe);
        if (rc) {
            dfree((void **)&d_rec);
            ctx->iajb_elems = 0;
            return rc;
        }
        ctx->iajb_elems = oe;
    }
    if (hipMemcpyAsync(d_rec, iajb_records, sizeof(double) * oe, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        dfree((void **)&d_rec);
        return set_error(ctx, MYQC_E_HIP, "myqc_mp2_rhf_records: upload failed");
    }
    k_from_records<<<1024, 256, 0, ctx->stream>>>(d_rec, ctx->d_iajb, nocc, nvrt);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree((void **)&d_rec);
    if (e != hipSuccess) return set_error(ctx, MYQC_E_HIP, "myqc_mp2_rhf_records: %s", hipGetErrorString(e));
    ctx->iajb_nocc = nocc;
    ctx->iajb_nv = nvrt;
    return myqc_mp2_rhf(ctx, eig, nocc, e_aa, e_ab, e_tot);
}

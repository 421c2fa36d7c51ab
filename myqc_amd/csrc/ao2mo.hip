// ao2mo.hip -- the RHF (ia|jb) four-index transform and the MP2 energy.
//
// The transform keeps the reference's four quarter steps
// (idx1..idx4_trans, src/ao2mo/ao2mo.f90:1306-1439) but runs each as dense
// matrix-core products over the device-resident dense AO integrals, in
// hand-written kernels: FP64 MFMA (mfma_f64.hip) for the plain form, split
// bf16 MFMA with FP32 accumulation (mfma_gemm.hip) for the mixed one; in both
// steps 1 and 2 are fused so that the o n^3 intermediate is never stored.
// The MP2 reduction restates mp2_rhf (src/mp2/mp2.f90:79-144) as one pass
// with a deterministic two-stage sum.
#include <cmath>
#include <vector>

#include "context.h"

namespace myqc {
int gemm_bf16x3(myqc_ctx *ctx, bool transA, int m, int n, int k, const double *A, int lda, long long strideA,
                const double *B, int ldb, long long strideB, double *C, int ldc, long long strideC, int batch);
int ao2mo12_bf16x3(myqc_ctx *ctx, int np, const double *Cp, int nq, const double *Cq, int n, const double *XX,
                   double *M);
int ao2mo12_f64(myqc_ctx *ctx, int np, const double *Cp, int nq, const double *Cq, int n, const double *XX,
                double *M);
int gemm_f64(myqc_ctx *ctx, int m, int nb, int k, const double *A, long long lda, long long sA, const double *B,
             long long ldb, double *C, long long ldc, long long sC, int batch);
}

namespace myqc {

// Record streams of an ijab file: record (i,j) i-major, (a,b) a fastest
// (src/ao2mo/ao2mo.f90:566-582 RHF; :718-724, :804-810, :890-896 UHF).  With
// tri set only the records i < j are written (ijab_AA, ijab_BB), else all
// o1 x o2 of them (ijab_AB).  O is [i][a][j][b] with i fastest.
__device__ __forceinline__ size_t record_of(int i, int j, int o1, int o2, bool tri) {
    return tri ? (size_t)i * (2 * o1 - i - 1) / 2 + (size_t)(j - i - 1) : (size_t)i * o2 + j;
}

__global__ void k_records(const double *__restrict__ O, double *__restrict__ out, int o1, int nv1, int o2, int nv2,
                          int tri) {
    const size_t total = (size_t)o1 * o2 * nv1 * nv2;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        size_t x = r;
        const int a = (int)(x % nv1); x /= nv1;
        const int b = (int)(x % nv2); x /= nv2;
        const int j = (int)(x % o2); x /= o2;
        const int i = (int)x;
        if (tri && j <= i) continue;
        out[(record_of(i, j, o1, o2, tri) * nv2 + b) * nv1 + a] = O[i + (size_t)o1 * (a + (size_t)nv1 * (j + (size_t)o2 * b))];
    }
}

// the inverse: records to [i][a][j][b] (entries of absent records are left alone)
__global__ void k_from_records(const double *__restrict__ rec, double *__restrict__ O, int o1, int nv1, int o2, int nv2,
                               int tri) {
    const size_t total = (size_t)o1 * o2 * nv1 * nv2;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        size_t x = r;
        const int a = (int)(x % nv1); x /= nv1;
        const int b = (int)(x % nv2); x /= nv2;
        const int j = (int)(x % o2); x /= o2;
        const int i = (int)x;
        if (tri && j <= i) continue;
        O[i + (size_t)o1 * (a + (size_t)nv1 * (j + (size_t)o2 * b))] = rec[(record_of(i, j, o1, o2, tri) * nv2 + b) * nv1 + a];
    }
}

// Per-block partial sums over one (ia|jb) set, O = [i][a][j][b] (i fastest),
// D = e1[i] + e2[j] - e1[o1+a] - e2[o2+b]:
//   part[2k]   = sum_{i<j, a<b} ((ia|jb) - (ib|ja))^2 / D   (same-spin sets only: same != 0)
//   part[2k+1] = sum_{ijab} (ia|jb)^2 / D
// (mp2_rhf, src/mp2/mp2.f90:79-144; mp2_uhf, :154-232)
__global__ __launch_bounds__(256) void k_mp2(const double *__restrict__ O, const double *__restrict__ e1,
                                             const double *__restrict__ e2, int o1, int nv1, int o2, int nv2, int same,
                                             double *__restrict__ part) {
    const size_t total = (size_t)o1 * nv1 * o2 * nv2;
    double s1 = 0.0, s2 = 0.0;
    for (size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x; r < total; r += (size_t)gridDim.x * blockDim.x) {
        size_t x = r;
        const int i = (int)(x % o1); x /= o1;
        const int a = (int)(x % nv1); x /= nv1;
        const int j = (int)(x % o2); x /= o2;
        const int b = (int)x;
        const double t = O[r];
        const double den = e1[i] + e2[j] - e1[a + o1] - e2[b + o2];
        s2 += t * t / den;
        if (same && i < j && a < b) {
            const double u = t - O[i + (size_t)o1 * (b + (size_t)nv1 * (j + (size_t)o2 * a))];
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

// MP2 sums straight from the ijab_AB record stream (record (i,j) i-major,
// (a,b) a fastest; src/mp2/mp2.f90:99-135): one workgroup per record writes
// its two partial sums to part[2 rec], part[2 rec + 1]
//   sum_{a<b} ((ia|jb) - (ib|ja))^2 / D  (for i < j only),   sum_{ab} (ia|jb)^2 / D,
// D = e_i + e_j - e_a - e_b; every slot of [0, 2 o^2) is written.
__global__ __launch_bounds__(256) void k_mp2_records(const double *__restrict__ R, const double *__restrict__ e, int o,
                                                     int nv, double *__restrict__ part) {
    __shared__ double sh1[256], sh2[256];
    const int rec = blockIdx.x, i = rec / o, j = rec - i * o;
    const size_t nvv = (size_t)nv * nv;
    const double *Rr = R + (size_t)rec * nvv;
    const double eij = e[i] + e[j];
    double s1 = 0.0, s2 = 0.0;
    for (size_t q = threadIdx.x; q < nvv; q += blockDim.x) {
        const int b = (int)(q / nv), a = (int)(q - (size_t)b * nv);
        const double t = Rr[q];
        const double den = eij - e[o + a] - e[o + b];
        s2 += t * t / den;
        if (i < j && a < b) {
            const double u = t - Rr[(size_t)a * nv + b];
            s1 += u * u / den;
        }
    }
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
        part[2 * (size_t)rec] = sh1[0];
        part[2 * (size_t)rec + 1] = sh2[0];
    }
}

// out = {E(AA), E(AB), E(MBPT2) = 2 E(AA) + E(AB)} from the nrec record
// partials, summed in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void k_mp2_fold(const double *__restrict__ part, int nrec, double *__restrict__ out) {
    __shared__ double sh1[256], sh2[256];
    double s1 = 0.0, s2 = 0.0;
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        s1 += part[2 * (size_t)r];
        s2 += part[2 * (size_t)r + 1];
    }
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
        out[0] = sh1[0];
        out[1] = sh2[0];
        out[2] = 2.0 * sh1[0] + sh2[0];
    }
}

// Quarter steps 1 and 2 (idx1_trans, idx2_trans) over two column blocks of
// orbitals, Cp (n x np) and Cq (n x nq), leading dimension n, fused per (l,d)
// block so that L(p,v,l,d) = sum_u Cp(u,p) XX(u,v,l,d) stays on chip:
//   M(p,q,l,d) = sum_v [sum_u Cp(u,p) XX(u,v,l,d)] Cq(v,q)      (np x nq x n x n)
// With release_xx, *dXX is freed once nothing reads it any more.
static int quarter12(myqc_ctx *ctx, double **dXX, bool release_xx, const double *Cp, int np, const double *Cq, int nq,
                     int precision, double **dM) {
    const int n = ctx->n;
    const size_t N = n, msize = sizeof(double) * (size_t)np * nq * N * N;
    MYQC_CHECK(dalloc(ctx, (void **)dM, msize));   // (a kept buffer of that size is reused)
    if (precision == MYQC_BF16X3) MYQC_CHECK(ao2mo12_bf16x3(ctx, np, Cp, nq, Cq, n, *dXX, *dM));
    else MYQC_CHECK(ao2mo12_f64(ctx, np, Cp, nq, Cq, n, *dXX, *dM));
    if (release_xx) dfree((void **)dXX);   // (hipFree waits for the kernel)
    return 0;
}

// C = A B with A (m x k, lda, stride sA per problem) and the shared B (k x nb,
// leading dimension k) for `batch` problems, in the transform's precision
static int quarter_gemm(myqc_ctx *ctx, int m, int nb, int k, const double *A, long long lda, long long sA,
                        const double *B, double *C, long long ldc, long long sC, int batch, int precision) {
    if (precision == MYQC_BF16X3)
        return gemm_bf16x3(ctx, false, m, nb, k, A, (int)lda, sA, B, k, 0, C, (int)ldc, sC, batch);
    return gemm_f64(ctx, m, nb, k, A, lda, sA, B, k, C, ldc, sC, batch);
}

// Quarter steps 3 and 4 (idx3_trans, idx4_trans) with the second pair's
// orbital blocks Cr (n x nr) and Cs (n x ns):
//   O(pq,r,s) = sum_d [sum_l M(pq,l,d) Cr(l,r)] Cs(d,s)     (m1 x nr x ns, pq fastest)
static int quarter34(myqc_ctx *ctx, const double *dM, int m1, const double *Cr, int nr, const double *Cs, int ns,
                     int precision, double *dO, double **keep = nullptr) {
    const int n = ctx->n;
    const size_t N = n;
    double *dN = keep ? *keep : nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&dN, sizeof(double) * (size_t)m1 * nr * N));
    if (keep) *keep = dN;
    // 3: N(pq,r,d) = sum_l M(pq,l,d) Cr(l,r)
    int rc = quarter_gemm(ctx, m1, nr, n, dM, m1, (long long)m1 * n, Cr, dN, m1, (long long)m1 * nr, n, precision);
    // 4: O(pq,r,s) = sum_d N(pq,r,d) Cs(d,s)
    if (rc == 0)
        rc = quarter_gemm(ctx, m1 * nr, ns, n, dN, (long long)m1 * nr, 0, Cs, dO, (long long)m1 * nr, 0, 1, precision);
    if (!keep) dfree((void **)&dN);
    return rc;
}

static double transform_flops(int o1, int nv1, int o2, int nv2, int n) {
    const double N = n;
    return 2.0 * ((double)o1 * N * N * N * N + (double)o1 * nv1 * N * N * N + (double)o1 * nv1 * o2 * N * N +
                  (double)o1 * nv1 * o2 * nv2 * N);
}

static int upload_square(myqc_ctx *ctx, const double *host, double **dev) {
    const size_t nn = (size_t)ctx->n * ctx->n;
    MYQC_CHECK(dalloc(ctx, (void **)dev, sizeof(double) * nn));
    if (hipMemcpyAsync(*dev, host, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return set_error(ctx, MYQC_E_HIP, "ao2mo: upload failed");
    return 0;
}

// dense AO integrals at the stored build's screening threshold
static int dense_integrals(myqc_ctx *ctx, double **dXX) {
    const size_t N = ctx->n;
    MYQC_CHECK(dalloc(ctx, (void **)dXX, sizeof(double) * N * N * N * N));
    return eri_fill_dense(ctx, ctx->have_eri ? ctx->eri_tol : 0.0, *dXX);
}

// RHF: (ia|jb) for the one set of orbitals (slow_ao2mo_MP2_RHF,
// src/ao2mo/ao2mo.f90:465-602) into ctx->d_iajb
static int ao2mo(myqc_ctx *ctx, const double *C_host, int o, int precision) {
    const int n = ctx->n, nv = n - o;
    double *dXX = nullptr, *dC = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        dfree((void **)&dXX);
        dfree((void **)&dC);
    };
#define TRY(x)          \
    do {                \
        rc = (x);       \
        if (rc) {       \
            cleanup();  \
            return rc;  \
        }               \
    } while (0)
    TRY(upload_square(ctx, C_host, &dC));
    TRY(dense_integrals(ctx, &dXX));
    // the output and the quarter-step intermediates (kept by the context, so
    // a repeated transform of one size allocates nothing) before the timed
    // region
    const size_t oe = (size_t)o * nv * o * nv, N = n;
    if (ctx->iajb_elems != oe) {
        TRY(dalloc(ctx, (void **)&ctx->d_iajb, sizeof(double) * (oe ? oe : 1)));
        ctx->iajb_elems = oe;
    }
    TRY(dalloc(ctx, (void **)&ctx->d_q12, sizeof(double) * std::max<size_t>((size_t)o * nv * N * N, 1)));
    TRY(dalloc(ctx, (void **)&ctx->d_q3, sizeof(double) * std::max<size_t>((size_t)o * nv * o * N, 1)));
    hipEvent_t ev;
    TRY(timer_begin(ctx, T_AO2MO, &ev));
    TRY(quarter12(ctx, &dXX, false, dC, o, dC + (size_t)n * o, nv, precision, &ctx->d_q12));
    TRY(quarter34(ctx, ctx->d_q12, o * nv, dC, o, dC + (size_t)n * o, nv, precision, ctx->d_iajb, &ctx->d_q3));
    TRY(timer_end(ctx, T_AO2MO, ev, transform_flops(o, nv, o, nv, n)));
    ctx->iajb_nocc = o;
    ctx->iajb_nv = nv;
    cleanup();
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// UHF: the three spin sets of slow_ao2mo_MP2_UHF (src/ao2mo/ao2mo.f90:614-904)
// into ctx->d_uhf[0..2] = AA (Ca, Ca), BB (Cb, Cb), AB (i, a alpha; j, b beta).
// With vvoo, also the same-spin (ab|ij) of slow_ao2mo_CIS_UHF (:919-1227)
// into ctx->d_vvoo[0..1] (alpha, beta), [a][b][i][j] with a fastest, from the
// same dense integrals.
// The reference recomputes the first two quarter steps for each set; here
// M_A = steps 1-2 with Ca serves both AA and AB, so the dense integrals are
// contracted twice (Ca, Cb) instead of three times.  A set with no occupied
// or no virtual orbital is empty.
int ao2mo_uhf(myqc_ctx *ctx, const double *Ca_host, const double *Cb_host, int oa, int ob, int precision, bool vvoo) {
    const int n = ctx->n, va = n - oa, vb = n - ob;
    double *dXX = nullptr, *dCa = nullptr, *dCb = nullptr, *dMa = nullptr, *dMb = nullptr;
    int rc = 0;
    auto cleanup = [&]() {
        dfree((void **)&dXX);
        dfree((void **)&dCa);
        dfree((void **)&dCb);
        dfree((void **)&dMa);
        dfree((void **)&dMb);
    };
    ctx->have_uhf = ctx->have_vvoo = false;
    const int dims[3][4] = {{oa, va, oa, va}, {ob, vb, ob, vb}, {oa, va, ob, vb}};
    for (int s = 0; s < 3; ++s) {
        const size_t e = (size_t)dims[s][0] * dims[s][1] * dims[s][2] * dims[s][3];
        TRY(dalloc(ctx, (void **)&ctx->d_uhf[s], sizeof(double) * (e ? e : 1)));
        for (int k = 0; k < 4; ++k) ctx->uhf_dims[s][k] = dims[s][k];
    }
    const bool want_a = (size_t)oa * va > 0, want_b = (size_t)ob * vb > 0;
    TRY(upload_square(ctx, Ca_host, &dCa));
    TRY(upload_square(ctx, Cb_host, &dCb));
    if (want_a || want_b) TRY(dense_integrals(ctx, &dXX));
    hipEvent_t ev;
    TRY(timer_begin(ctx, T_AO2MO, &ev));
    double flops = 0.0;
    if (want_a) TRY(quarter12(ctx, &dXX, !want_b && !vvoo, dCa, oa, dCa + (size_t)n * oa, va, precision, &dMa));
    if (want_b) TRY(quarter12(ctx, &dXX, !vvoo, dCb, ob, dCb + (size_t)n * ob, vb, precision, &dMb));
    if (vvoo) {
        const double *Cs[2] = {dCa, dCb};
        const int os[2] = {oa, ob}, vs[2] = {va, vb};
        for (int s = 0; s < 2; ++s) {
            const size_t e = (size_t)vs[s] * vs[s] * os[s] * os[s];
            TRY(dalloc(ctx, (void **)&ctx->d_vvoo[s], sizeof(double) * (e ? e : 1)));
            ctx->vvoo_dims[s][0] = vs[s];
            ctx->vvoo_dims[s][1] = os[s];
            if (e == 0) continue;
            double *dV = nullptr;
            const double *Cv = Cs[s] + (size_t)n * os[s];
            rc = quarter12(ctx, &dXX, false, Cv, vs[s], Cv, vs[s], precision, &dV);
            if (rc == 0) rc = quarter34(ctx, dV, vs[s] * vs[s], Cs[s], os[s], Cs[s], os[s], precision, ctx->d_vvoo[s]);
            dfree((void **)&dV);
            TRY(rc);
            flops += transform_flops(vs[s], vs[s], os[s], os[s], n);
        }
    }
    dfree((void **)&dXX);
    if (want_a) {
        TRY(quarter34(ctx, dMa, oa * va, dCa, oa, dCa + (size_t)n * oa, va, precision, ctx->d_uhf[0]));
        flops += transform_flops(oa, va, oa, va, n);
    }
    if (want_b) {
        TRY(quarter34(ctx, dMb, ob * vb, dCb, ob, dCb + (size_t)n * ob, vb, precision, ctx->d_uhf[1]));
        flops += transform_flops(ob, vb, ob, vb, n);
    }
    if (want_a && want_b) {
        TRY(quarter34(ctx, dMa, oa * va, dCb, ob, dCb + (size_t)n * ob, vb, precision, ctx->d_uhf[2]));
        // steps 3-4 only: steps 1-2 are shared with the AA set
        flops += 2.0 * ((double)oa * va * ob * n * n + (double)oa * va * ob * vb * n);
    }
    TRY(timer_end(ctx, T_AO2MO, ev, flops));
    cleanup();
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_uhf = true;
    ctx->have_vvoo = vvoo;
#undef TRY
    return 0;
}

// the records of one set into a host buffer (nothing is written for an empty set)
static int download_records(myqc_ctx *ctx, const double *dO, const int *d, int tri, double *host) {
    const size_t nrec = tri ? (size_t)d[0] * (d[0] > 0 ? d[0] - 1 : 0) / 2 : (size_t)d[0] * d[2];
    const size_t cnt = nrec * d[1] * d[3];
    if (cnt == 0) return 0;
    double *d_rec = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_rec, sizeof(double) * cnt));
    k_records<<<1024, 256, 0, ctx->stream>>>(dO, d_rec, d[0], d[1], d[2], d[3], tri);
    int rc = 0;
    if (d2h(ctx, host, d_rec, sizeof(double) * cnt) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "ao2mo: download failed");
    dfree((void **)&d_rec);
    return rc;
}

// the two partial sums of k_mp2 over one set, folded on the host in a fixed order
static int mp2_sums(myqc_ctx *ctx, const double *dO, const int *d, const double *d_e1, const double *d_e2, int same,
                    double *s_same, double *s_all) {
    *s_same = *s_all = 0.0;
    if ((size_t)d[0] * d[1] * d[2] * d[3] == 0) return 0;
    const int nblk = 512;
    double *d_part = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_part, sizeof(double) * 2 * nblk));
    std::vector<double> part(2 * nblk);
    k_mp2<<<nblk, 256, 0, ctx->stream>>>(dO, d_e1, d_e2, d[0], d[1], d[2], d[3], same, d_part);
    int rc = 0;
    if (d2h(ctx, part.data(), d_part, sizeof(double) * 2 * nblk) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "mp2: download failed");
    dfree((void **)&d_part);
    if (rc) return rc;
    for (int b = 0; b < nblk; ++b) {
        *s_same += part[2 * b];
        *s_all += part[2 * b + 1];
    }
    return 0;
}

// E(AA), E(BB), E(AB) of mp2_uhf (src/mp2/mp2.f90:154-232) from ctx->d_uhf
static int mp2_uhf_sums(myqc_ctx *ctx, const double *eig_a, const double *eig_b, double *e_aa, double *e_bb,
                        double *e_ab, double *e_tot) {
    const int n = ctx->uhf_dims[0][0] + ctx->uhf_dims[0][1];
    double *d_e = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_e, sizeof(double) * 2 * n));
    double s[3] = {0.0, 0.0, 0.0}, unused = 0.0;
    int rc = 0;
    if (hipMemcpyAsync(d_e, eig_a, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(d_e + n, eig_b, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "mp2: upload failed");
    hipEvent_t ev;
    if (rc == 0) rc = timer_begin(ctx, T_MP2, &ev);
    if (rc == 0) rc = mp2_sums(ctx, ctx->d_uhf[0], ctx->uhf_dims[0], d_e, d_e, 1, &s[0], &unused);
    if (rc == 0) rc = mp2_sums(ctx, ctx->d_uhf[1], ctx->uhf_dims[1], d_e + n, d_e + n, 1, &s[1], &unused);
    if (rc == 0) rc = mp2_sums(ctx, ctx->d_uhf[2], ctx->uhf_dims[2], d_e, d_e + n, 0, &unused, &s[2]);
    if (rc == 0) {
        double work = 0.0;
        for (int k = 0; k < 3; ++k)
            work += 8.0 * ctx->uhf_dims[k][0] * ctx->uhf_dims[k][1] * (double)ctx->uhf_dims[k][2] * ctx->uhf_dims[k][3];
        rc = timer_end(ctx, T_MP2, ev, work);
    }
    dfree((void **)&d_e);
    if (rc) return rc;
    if (e_aa) *e_aa = s[0];
    if (e_bb) *e_bb = s[1];
    if (e_ab) *e_ab = s[2];
    if (e_tot) *e_tot = s[0] + s[1] + s[2];
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
        const int d[4] = {nocc, ctx->n - nocc, nocc, ctx->n - nocc};
        return download_records(ctx, ctx->d_iajb, d, 0, iajb);
    }
    return 0;
}

extern "C" int myqc_mp2_rhf(myqc_ctx *ctx, const double *eig, int nocc, double *e_aa, double *e_ab, double *e_tot) {
    if (!ctx || !eig) return MYQC_E_ARG;
    if (!ctx->d_iajb || ctx->iajb_nocc != nocc)
        return set_error(ctx, MYQC_E_STATE, "myqc_mp2_rhf: no (ia|jb) for nocc=%d (call myqc_ao2mo_ovov)", nocc);
    const int o = nocc, nv = ctx->iajb_nv, n = o + nv;
    double *d_eig = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_eig, sizeof(double) * n));
    int rc = 0;
    if (hipMemcpyAsync(d_eig, eig, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "upload failed");
    double s1 = 0.0, s2 = 0.0;
    if (rc == 0) {
        hipEvent_t ev;
        const int d[4] = {o, nv, o, nv};
        rc = timer_begin(ctx, T_MP2, &ev);
        if (rc == 0) rc = mp2_sums(ctx, ctx->d_iajb, d, d_eig, d_eig, 1, &s1, &s2);
        if (rc == 0) rc = timer_end(ctx, T_MP2, ev, 8.0 * (double)o * o * nv * nv);
    }
    dfree((void **)&d_eig);
    if (rc) return rc;
    if (e_aa) *e_aa = s1;
    if (e_ab) *e_ab = s2;
    if (e_tot) *e_tot = 2.0 * s1 + s2;
    return 0;
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
    k_from_records<<<1024, 256, 0, ctx->stream>>>(d_rec, ctx->d_iajb, nocc, nvrt, nocc, nvrt, 0);
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    dfree((void **)&d_rec);
    if (e != hipSuccess) return set_error(ctx, MYQC_E_HIP, "myqc_mp2_rhf_records: %s", hipGetErrorString(e));
    ctx->iajb_nocc = nocc;
    ctx->iajb_nv = nvrt;
    return myqc_mp2_rhf(ctx, eig, nocc, e_aa, e_ab, e_tot);
}

extern "C" int myqc_ao2mo_ovov_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int nalpha, int nbeta,
                                   double *ijab_aa, double *ijab_bb, double *ijab_ab, int precision) {
    if (!ctx || !Ca || !Cb) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_ao2mo_ovov_uhf: no system set");
    if (nalpha < 0 || nalpha > ctx->n || nbeta < 0 || nbeta > ctx->n)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_ovov_uhf: nalpha=%d nbeta=%d", nalpha, nbeta);
    if (precision != MYQC_FP64 && precision != MYQC_BF16X3)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_ovov_uhf: precision %d", precision);
    const size_t n = ctx->n;
    if (n * n * n * n * 8 > (size_t)200 << 30)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_ovov_uhf: n=%zu needs more than 200 GB of dense AO integrals", n);
    MYQC_CHECK(ao2mo_uhf(ctx, Ca, Cb, nalpha, nbeta, precision, false));
    if (ijab_aa) MYQC_CHECK(download_records(ctx, ctx->d_uhf[0], ctx->uhf_dims[0], 1, ijab_aa));
    if (ijab_bb) MYQC_CHECK(download_records(ctx, ctx->d_uhf[1], ctx->uhf_dims[1], 1, ijab_bb));
    if (ijab_ab) MYQC_CHECK(download_records(ctx, ctx->d_uhf[2], ctx->uhf_dims[2], 0, ijab_ab));
    return 0;
}

extern "C" int myqc_mp2_uhf(myqc_ctx *ctx, const double *eig_a, const double *eig_b, int nalpha, int nbeta,
                            double *e_aa, double *e_bb, double *e_ab, double *e_tot) {
    if (!ctx || !eig_a || !eig_b) return MYQC_E_ARG;
    if (!ctx->have_uhf || ctx->uhf_dims[0][0] != nalpha || ctx->uhf_dims[1][0] != nbeta)
        return set_error(ctx, MYQC_E_STATE, "myqc_mp2_uhf: no UHF (ia|jb) for nalpha=%d nbeta=%d (call myqc_ao2mo_ovov_uhf)",
                         nalpha, nbeta);
    return mp2_uhf_sums(ctx, eig_a, eig_b, e_aa, e_bb, e_ab, e_tot);
}

extern "C" int myqc_mp2_uhf_records(myqc_ctx *ctx, const double *ijab_aa, const double *ijab_bb,
                                    const double *ijab_ab, const double *eig_a, const double *eig_b, int nalpha,
                                    int nbeta, int nvrt_a, int nvrt_b, double *e_aa, double *e_bb, double *e_ab,
                                    double *e_tot) {
    if (!ctx || !eig_a || !eig_b) return MYQC_E_ARG;
    if (nalpha < 0 || nbeta < 0 || nvrt_a < 0 || nvrt_b < 0 || nalpha + nvrt_a != nbeta + nvrt_b)
        return set_error(ctx, MYQC_E_ARG, "myqc_mp2_uhf_records: nalpha=%d nbeta=%d nvrt_a=%d nvrt_b=%d", nalpha,
                         nbeta, nvrt_a, nvrt_b);
    const int dims[3][4] = {{nalpha, nvrt_a, nalpha, nvrt_a}, {nbeta, nvrt_b, nbeta, nvrt_b},
                            {nalpha, nvrt_a, nbeta, nvrt_b}};
    const double *recs[3] = {ijab_aa, ijab_bb, ijab_ab};
    ctx->have_uhf = false;
    for (int s = 0; s < 3; ++s) {
        const int *d = dims[s];
        const int tri = s < 2;
        const size_t e = (size_t)d[0] * d[1] * d[2] * d[3];
        const size_t nrec = tri ? (size_t)d[0] * (d[0] > 0 ? d[0] - 1 : 0) / 2 : (size_t)d[0] * d[2];
        const size_t cnt = nrec * d[1] * d[3];
        MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_uhf[s], sizeof(double) * (e ? e : 1)));
        for (int k = 0; k < 4; ++k) ctx->uhf_dims[s][k] = d[k];
        if (e == 0) continue;
        // same-spin sets hold only the records i < j: the rest stays zero and unread
        MYQC_HIP(ctx, hipMemsetAsync(ctx->d_uhf[s], 0, sizeof(double) * e, ctx->stream));
        if (cnt == 0) continue;
        if (!recs[s]) return set_error(ctx, MYQC_E_ARG, "myqc_mp2_uhf_records: records of set %d missing", s);
        double *d_rec = nullptr;
        MYQC_CHECK(dalloc(ctx, (void **)&d_rec, sizeof(double) * cnt));
        if (hipMemcpyAsync(d_rec, recs[s], sizeof(double) * cnt, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
            dfree((void **)&d_rec);
            return set_error(ctx, MYQC_E_HIP, "myqc_mp2_uhf_records: upload failed");
        }
        k_from_records<<<1024, 256, 0, ctx->stream>>>(d_rec, ctx->d_uhf[s], d[0], d[1], d[2], d[3], tri);
        const hipError_t err = hipStreamSynchronize(ctx->stream);
        dfree((void **)&d_rec);
        if (err != hipSuccess) return set_error(ctx, MYQC_E_HIP, "myqc_mp2_uhf_records: %s", hipGetErrorString(err));
    }
    ctx->have_uhf = true;
    return mp2_uhf_sums(ctx, eig_a, eig_b, e_aa, e_bb, e_ab, e_tot);
}

// myqc_mp2_rhf_records without any allocation: the records and orbital
// energies are already on the device, the per-record partial sums go to the
// device scratch of this GPU (ctx->d_scratch, kNoAllocRecords records), and
// the three energies to d_out.  Work is enqueued on the context stream and
// complete at return.  Calls that overlap on one GPU share the scratch and
// race.
extern "C" int myqc_mp2_rhf_records_noalloc(myqc_ctx *ctx, const double *d_records, const double *d_eig, int nocc,
                                            int nvrt, double *d_out) {
    if (!ctx || !d_records || !d_eig || !d_out) return MYQC_E_ARG;
    if (nocc < 1 || nvrt < 1)
        return set_error(ctx, MYQC_E_ARG, "myqc_mp2_rhf_records_noalloc: nocc=%d nvrt=%d", nocc, nvrt);
    const size_t nrec = (size_t)nocc * nocc;
    if (nrec > kNoAllocRecords)
        return set_error(ctx, MYQC_E_ARG,
                         "myqc_mp2_rhf_records_noalloc: nocc=%d gives %zu records, beyond the scratch's %zu "
                         "(nocc <= 256); use myqc_mp2_rhf_records",
                         nocc, nrec, kNoAllocRecords);
    if (!ctx->d_scratch) return set_error(ctx, MYQC_E_STATE, "myqc_mp2_rhf_records_noalloc: no device scratch");
    k_mp2_records<<<(unsigned)nrec, 256, 0, ctx->stream>>>(d_records, d_eig, nocc, nvrt, ctx->d_scratch);
    k_mp2_fold<<<1, 256, 0, ctx->stream>>>(ctx->d_scratch, (int)nrec, d_out);
    MYQC_HIP(ctx, hipGetLastError());
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// cis.hip -- UHF-CIS: the integral streams of slow_ao2mo_CIS_UHF
// (src/ao2mo/ao2mo.f90:919-1227) and the CIS eigenproblem of cis_scf_uhf
// (src/CIS/cis.f90:87-223).
//
// The five streams are the reference's files, record (j,b) j-major with b
// fastest, each record over (i,a) with a fastest:
//   ajib_AA, ajib_BB: (ai|jb) of one spin;  ajbi_AA, ajbi_BB: (ab|ji) of one
//   spin;  ajib_AB: (a_A i_A|j_B b_B).
// (ai|jb) is the (ia|jb) set of the UHF MP2 transform (real orbitals), so
// only the same-spin (ab|ij) sets are extra.  The CIS matrix, dimension
// upA + upB (upS = nocc_S nvrt_S), takes record c as its column c:
//   A(ia, jb) = (ai|jb) - (ab|ji) + delta (e_a - e_i)  within a spin,
//   A(ia_A, jb_B) = (ai|jb)_AB,
// and is diagonalised whole (rocSOLVER dsyevd on its upper triangle, where the
// reference calls DSYEV('V','U'), cis.f90:200-209).
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

#include "context.h"

namespace myqc {

int ao2mo_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int oa, int ob, int precision, bool vvoo);

// kind 0: (ai|jb) from an ovov set O = [i][a][j][b] (i fastest), o1 x v1 x o2 x v2
// kind 1: (ab|ji) from a vvoo set V = [a][b][r][s] (a fastest), v1 x v1 x o1 x o1
//         (ao2mo.f90 writes Om(a,b,j,i), :1051 and :1210)
__global__ void k_cis_records(const double *__restrict__ src, double *__restrict__ out, int kind, int o1, int v1,
                              int o2, int v2) {
    const size_t up1 = (size_t)o1 * v1, total = up1 * o2 * v2;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const size_t rec = x / up1, e = x % up1;
        const int j = (int)(rec / v2), b = (int)(rec % v2);
        const int i = (int)(e / v1), a = (int)(e % v1);
        out[x] = kind == 0 ? src[i + (size_t)o1 * (a + (size_t)v1 * (j + (size_t)o2 * b))]
                           : src[a + (size_t)v1 * (b + (size_t)v1 * (j + (size_t)o1 * i))];
    }
}

// the CIS matrix (column-major, dim x dim) from the five device streams
__global__ void k_cis_matrix(double *__restrict__ A, int dim, int upA, int upB, const double *__restrict__ aa1,
                             const double *__restrict__ aa2, const double *__restrict__ ab,
                             const double *__restrict__ bb1, const double *__restrict__ bb2,
                             const double *__restrict__ eA, const double *__restrict__ eB, int oA, int vA, int oB,
                             int vB) {
    const size_t total = (size_t)dim * dim;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(x % dim), c = (int)(x / dim);
        double v;
        if (r < upA && c < upA) {
            const size_t e = (size_t)c * upA + r;
            v = aa1[e] - aa2[e];
            if (r == c) v += eA[oA + r % vA] - eA[r / vA];
        } else if (r < upA) {
            v = ab[(size_t)(c - upA) * upA + r];
        } else if (c < upA) {
            v = ab[(size_t)(r - upA) * upA + c];
        } else {
            const int rr = r - upA, cc = c - upA;
            const size_t e = (size_t)cc * upB + rr;
            v = bb1[e] - bb2[e];
            if (rr == cc) v += eB[oB + rr % vB] - eB[rr / vB];
        }
        A[x] = v;
    }
}

struct CisDims {
    int oA, vA, oB, vB;
    size_t upA() const { return (size_t)oA * vA; }
    size_t upB() const { return (size_t)oB * vB; }
    // element counts of the five streams: ajib_AA, ajbi_AA, ajib_AB, ajib_BB, ajbi_BB
    size_t count(int s) const {
        const size_t a = upA(), b = upB();
        return s < 2 ? a * a : s == 2 ? a * b : b * b;
    }
};

// the five streams on the device from the sets of the last transform
static int device_streams(myqc_ctx *ctx, const CisDims &d, double *const *out) {
    const double *src[5] = {ctx->d_uhf[0], ctx->d_vvoo[0], ctx->d_uhf[2], ctx->d_uhf[1], ctx->d_vvoo[1]};
    const int kind[5] = {0, 1, 0, 0, 1};
    const int dims[5][4] = {{d.oA, d.vA, d.oA, d.vA}, {d.oA, d.vA, d.oA, d.vA}, {d.oA, d.vA, d.oB, d.vB},
                            {d.oB, d.vB, d.oB, d.vB}, {d.oB, d.vB, d.oB, d.vB}};
    for (int s = 0; s < 5; ++s) {
        if (d.count(s) == 0) continue;
        k_cis_records<<<1024, 256, 0, ctx->stream>>>(src[s], out[s], kind[s], dims[s][0], dims[s][1], dims[s][2],
                                                     dims[s][3]);
    }
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// assemble and diagonalise; energies[0..nroots), vectors dim x nroots
static int cis_solve(myqc_ctx *ctx, const CisDims &d, double *const *dstreams, const double *eig_a,
                     const double *eig_b, int nroots, double *energies, double *vectors) {
    const int n = d.oA + d.vA;
    const size_t dim = d.upA() + d.upB();
    if (dim == 0 || nroots <= 0) return 0;
    if (dim > 0x7fffffff) return set_error(ctx, MYQC_E_ARG, "cis: dimension %zu too large", dim);
    double *buf = nullptr;
    int *info = nullptr;
    // A (dim^2), eigenvalues (dim), off-diagonal work (dim), eigenvalues of both spins (2 n)
    MYQC_CHECK(dalloc(ctx, (void **)&buf, sizeof(double) * (dim * dim + 2 * dim + 2 * (size_t)n)));
    int rc = dalloc(ctx, (void **)&info, sizeof(int));
    double *A = buf, *W = A + dim * dim, *E = W + dim, *eA = E + dim, *eB = eA + n;
    if (rc == 0 && (hipMemcpyAsync(eA, eig_a, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
                    hipMemcpyAsync(eB, eig_b, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        rc = set_error(ctx, MYQC_E_HIP, "cis: upload failed");
    if (rc == 0) {
        k_cis_matrix<<<1024, 256, 0, ctx->stream>>>(A, (int)dim, (int)d.upA(), (int)d.upB(), dstreams[0], dstreams[1],
                                                    dstreams[2], dstreams[3], dstreams[4], eA, eB, d.oA, d.vA, d.oB,
                                                    d.vB);
        rc = blas_reserve(ctx, [&] {
            (void)rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, (int)dim, A, (int)dim, W, E,
                                   info);
        });
        if (rc == 0 && rocsolver_dsyevd(ctx->blas, rocblas_evect_original, rocblas_fill_upper, (int)dim, A, (int)dim, W,
                                        E, info) != rocblas_status_success)
            rc = set_error(ctx, MYQC_E_HIP, "cis: rocsolver_dsyevd failed");
    }
    int h_info = 0;
    if (rc == 0 && d2h(ctx, &h_info, info, sizeof(int)) != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "cis: download");
    if (rc == 0 && h_info != 0) rc = set_error(ctx, MYQC_E_HIP, "cis: rocsolver_dsyevd info=%d", h_info);
    const size_t k = std::min((size_t)nroots, dim);
    if (rc == 0 && energies && d2h(ctx, energies, W, sizeof(double) * k) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "cis: download failed");
    if (rc == 0 && vectors && d2h(ctx, vectors, A, sizeof(double) * dim * k) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "cis: download failed");
    dfree((void **)&buf);
    dfree((void **)&info);
    return rc;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_ao2mo_cis_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int nalpha, int nbeta,
                                  double *ajib_aa, double *ajbi_aa, double *ajib_ab, double *ajib_bb,
                                  double *ajbi_bb) {
    if (!ctx || !Ca || !Cb) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_ao2mo_cis_uhf: no system set");
    if (nalpha < 0 || nalpha > ctx->n || nbeta < 0 || nbeta > ctx->n)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_cis_uhf: nalpha=%d nbeta=%d", nalpha, nbeta);
    const size_t n = ctx->n;
    if (n * n * n * n * 8 > (size_t)200 << 30)
        return set_error(ctx, MYQC_E_ARG, "myqc_ao2mo_cis_uhf: n=%zu needs more than 200 GB of dense AO integrals", n);
    MYQC_CHECK(ao2mo_uhf(ctx, Ca, Cb, nalpha, nbeta, MYQC_FP64, true));
    double *outs[5] = {ajib_aa, ajbi_aa, ajib_ab, ajib_bb, ajbi_bb};
    const CisDims d{nalpha, ctx->n - nalpha, nbeta, ctx->n - nbeta};
    size_t total = 0;
    for (int s = 0; s < 5; ++s) total += outs[s] ? d.count(s) : 0;
    if (total == 0) return 0;
    double *buf = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&buf, sizeof(double) * (d.count(0) * 2 + d.count(2) + d.count(3) * 2)));
    double *streams[5];
    streams[0] = buf;
    for (int s = 1; s < 5; ++s) streams[s] = streams[s - 1] + d.count(s - 1);
    int rc = device_streams(ctx, d, streams);
    for (int s = 0; s < 5 && rc == 0; ++s)
        if (outs[s] && d.count(s) && d2h(ctx, outs[s], streams[s], sizeof(double) * d.count(s)) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "myqc_ao2mo_cis_uhf: download failed");
    dfree((void **)&buf);
    return rc;
}

extern "C" int myqc_cis_uhf(myqc_ctx *ctx, const double *ajib_aa, const double *ajbi_aa, const double *ajib_ab,
                            const double *ajib_bb, const double *ajbi_bb, const double *eig_a, const double *eig_b,
                            int nalpha, int nbeta, int nvrt_a, int nvrt_b, int nroots, double *energies,
                            double *vectors) {
    if (!ctx || !eig_a || !eig_b) return MYQC_E_ARG;
    if (nalpha < 0 || nbeta < 0 || nvrt_a < 0 || nvrt_b < 0 || nalpha + nvrt_a != nbeta + nvrt_b || nroots < 0)
        return set_error(ctx, MYQC_E_ARG, "myqc_cis_uhf: nalpha=%d nbeta=%d nvrt_a=%d nvrt_b=%d nroots=%d", nalpha,
                         nbeta, nvrt_a, nvrt_b, nroots);
    const CisDims d{nalpha, nvrt_a, nbeta, nvrt_b};
    const double *in[5] = {ajib_aa, ajbi_aa, ajib_ab, ajib_bb, ajbi_bb};
    const bool resident = !in[0] && !in[1] && !in[2] && !in[3] && !in[4];
    if (resident && (!ctx->have_vvoo || ctx->uhf_dims[0][0] != nalpha || ctx->uhf_dims[1][0] != nbeta ||
                     ctx->uhf_dims[0][1] != nvrt_a || ctx->uhf_dims[1][1] != nvrt_b))
        return set_error(ctx, MYQC_E_STATE, "myqc_cis_uhf: no CIS integrals for nalpha=%d nbeta=%d (call "
                                            "myqc_ao2mo_cis_uhf, or pass the record streams)", nalpha, nbeta);
    const size_t total = d.count(0) * 2 + d.count(2) + d.count(3) * 2;
    double *buf = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&buf, sizeof(double) * (total ? total : 1)));
    double *streams[5];
    streams[0] = buf;
    for (int s = 1; s < 5; ++s) streams[s] = streams[s - 1] + d.count(s - 1);
    int rc = 0;
    if (resident) {
        rc = device_streams(ctx, d, streams);
    } else {
        for (int s = 0; s < 5 && rc == 0; ++s) {
            if (d.count(s) == 0) continue;
            if (!in[s]) rc = set_error(ctx, MYQC_E_ARG, "myqc_cis_uhf: record stream %d missing", s);
            else if (hipMemcpyAsync(streams[s], in[s], sizeof(double) * d.count(s), hipMemcpyHostToDevice,
                                    ctx->stream) != hipSuccess)
                rc = set_error(ctx, MYQC_E_HIP, "myqc_cis_uhf: upload failed");
        }
    }
    if (rc == 0) rc = cis_solve(ctx, d, streams, eig_a, eig_b, nroots, energies, vectors);
    dfree((void **)&buf);
    return rc;
}

// one_e.hip -- overlap S and core Hamiltonian H = T + V over contracted s/p
// shell pairs, one thread per shell pair (int1e drop-in,
// src/integrals/int1e.f90:132-280; the per-pair math is in one_e_math.h),
// the dipole integrals of int1p (src/integrals/int1p.f90:140-355) and the
// SCF dipole moment of cphf_dipole_rhf (src/props/cphf.f90:106-121).
#include <cmath>

#include "context.h"
#include "one_e_math.h"

namespace myqc {

__global__ __launch_bounds__(256) void k_one_e(const ShellPair *__restrict__ pairs, const PrimPair *__restrict__ prims,
                                               const double *__restrict__ ftT, const double *__restrict__ atoms,
                                               int nnuc, int npairs, int n, double *__restrict__ S, double *__restrict__ H) {
    __shared__ double s_ft[FT_TOTAL];
    for (int i = threadIdx.x; i < FT_TOTAL; i += blockDim.x) s_ft[i] = ftT[i];
    __syncthreads();
    // one wavefront per shell pair: lane l takes the nuclei l, l + 64, ... of
    // the attraction integrals and the primitive pairs l, l + 64, ... of S and
    // T, then the lanes' sums are added across the wavefront
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (t >= npairs) return;   // whole wavefronts
    const ShellPair sp = pairs[t];
    double s[3][3], h[3][3];
    one_e_pair(sp, prims, s_ft, atoms, nnuc, s, h, lane, 64);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                h[a][b] += __shfl_xor(h[a][b], off, 64);
                s[a][b] += __shfl_xor(s[a][b], off, 64);
            }
    if (lane != 0) return;
    const int na = sp.la ? 3 : 1, nb = sp.lb ? 3 : 1;
    // constant indices into s and h (registers), the sizes as guards
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            if (a >= na || b >= nb) continue;
            const size_t i = sp.ao_a + a, j = sp.ao_b + b;
            S[i + n * j] = s[a][b];
            S[j + n * i] = s[a][b];
            H[i + n * j] = h[a][b];
            H[j + n * i] = h[a][b];
        }
}

// P_c (c = x, y, z) at P + c n^2, both triangles
__global__ __launch_bounds__(128) void k_dipole(const ShellPair *__restrict__ pairs, const PrimPair *__restrict__ prims,
                                                int npairs, int n, double *__restrict__ P) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const ShellPair sp = pairs[t];
    double dip[3][3][3];
    dipole_pair(sp, prims, dip);
    const int na = sp.la ? 3 : 1, nb = sp.lb ? 3 : 1;
    const size_t nn = (size_t)n * n;
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) {
                const size_t i = sp.ao_a + a, j = sp.ao_b + b;
                P[c * nn + i + n * j] = dip[c][a][b];
                P[c * nn + j + n * i] = dip[c][a][b];
            }
}

// mu[c] = -sum_uv D(u,v) P_c(v,u): one workgroup per component, a fixed-order fold
__global__ __launch_bounds__(256) void k_dipole_contract(const double *__restrict__ P, const double *__restrict__ D,
                                                         int n, double *__restrict__ mu) {
    __shared__ double sh[256];
    const int c = blockIdx.x;
    const size_t nn = (size_t)n * n;
    double acc = 0.0;
    for (size_t e = threadIdx.x; e < nn; e += blockDim.x) {
        const size_t u = e % n, v = e / n;
        acc += D[e] * P[c * nn + v + (size_t)n * u];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mu[c] = -sh[0];
}

static int dipole_integrals(myqc_ctx *ctx, double *dP) {
    const int n = ctx->n, np = (int)ctx->pairs.size();
    k_dipole<<<(np + 127) / 128, 128, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims, np, n, dP);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_int1p(myqc_ctx *ctx, double *P) {
    if (!ctx || !P) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_int1p: no system set");
    const size_t nn = (size_t)ctx->n * ctx->n;
    double *dP = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&dP, sizeof(double) * 3 * nn));
    int rc = dipole_integrals(ctx, dP);
    if (rc == 0 && d2h(ctx, P, dP, sizeof(double) * 3 * nn) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_int1p: download failed");
    dfree((void **)&dP);
    return rc;
}

extern "C" int myqc_dipole_rhf(myqc_ctx *ctx, const double *D, double *mu) {
    if (!ctx || !D || !mu) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_dipole_rhf: no system set");
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    double *buf = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&buf, sizeof(double) * (4 * nn + 3)));
    double *dP = buf, *dD = buf + 3 * nn, *dmu = buf + 4 * nn;
    int rc = 0;
    if (hipMemcpyAsync(dD, D, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_dipole_rhf: upload failed");
    if (rc == 0) rc = dipole_integrals(ctx, dP);
    if (rc == 0) {
        k_dipole_contract<<<3, 256, 0, ctx->stream>>>(dP, dD, n, dmu);
        if (hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "myqc_dipole_rhf: launch failed");
    }
    if (rc == 0 && d2h(ctx, mu, dmu, sizeof(double) * 3) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_dipole_rhf: download failed");
    dfree((void **)&buf);
    if (rc) return rc;
    // nuclear part (cphf.f90:118-122)
    for (int a = 0; a < ctx->nnuc; ++a)
        for (int c = 0; c < 3; ++c) mu[c] += ctx->Z[a] * ctx->xyz[3 * a + c];
    return 0;
}

extern "C" int myqc_int1e(myqc_ctx *ctx, double *S, double *H) {
    if (!ctx || !S || !H) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_int1e: no system set");
    const int n = ctx->n, np = (int)ctx->pairs.size();
    std::vector<double> atoms(4 * ctx->nnuc);
    for (int c = 0; c < ctx->nnuc; ++c) {
        atoms[4 * c] = ctx->Z[c];
        for (int d = 0; d < 3; ++d) atoms[4 * c + 1 + d] = ctx->xyz[3 * c + d];
    }
    double *dS = nullptr, *dH = nullptr, *dA = nullptr;
    const size_t nn = (size_t)n * n;
    int rc = dtmp(ctx, TMP_S, &dS, sizeof(double) * nn);
    if (rc == 0) rc = dtmp(ctx, TMP_H, &dH, sizeof(double) * nn);
    if (rc == 0) rc = dtmp(ctx, TMP_ATOMS, &dA, sizeof(double) * atoms.size());
    if (rc == 0 && hipMemcpyAsync(dA, atoms.data(), sizeof(double) * atoms.size(), hipMemcpyHostToDevice,
                                  ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "upload failed");
    if (rc == 0) {
        k_one_e<<<(np + 3) / 4, 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims, ctx->d_ftT, dA, ctx->nnuc, np, n,
                                                      dS, dH);
        if (hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "one-electron kernel launch failed");
    }
    if (rc == 0 && (d2h(ctx, S, dS, sizeof(double) * nn) != hipSuccess ||
                    d2h(ctx, H, dH, sizeof(double) * nn) != hipSuccess))
        rc = set_error(ctx, MYQC_E_HIP, "download failed");
    return rc;
}

// one_e.hip -- overlap S and core Hamiltonian H = T + V over contracted s/p
// shell pairs, one thread per shell pair (int1e drop-in,
// src/integrals/int1e.f90:132-280; the per-pair math is in one_e_math.h).
#include <cmath>

#include "context.h"
#include "one_e_math.h"

namespace myqc {

__global__ __launch_bounds__(128) void k_one_e(const ShellPair *__restrict__ pairs, const PrimPair *__restrict__ prims,
                                               const double *__restrict__ ftT, const double *__restrict__ atoms,
                                               int nnuc, int npairs, int n, double *__restrict__ S, double *__restrict__ H) {
    __shared__ double s_ft[FT_NT * FT_NM];
    for (int i = threadIdx.x; i < FT_NT * FT_NM; i += blockDim.x) s_ft[i] = ftT[i];
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const ShellPair sp = pairs[t];
    double s[3][3], h[3][3];
    one_e_pair(sp, prims, s_ft, atoms, nnuc, s, h);
    const int na = sp.la ? 3 : 1, nb = sp.lb ? 3 : 1;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            const size_t i = sp.ao_a + a, j = sp.ao_b + b;
            S[i + n * j] = s[a][b];
            S[j + n * i] = s[a][b];
            H[i + n * j] = h[a][b];
            H[j + n * i] = h[a][b];
        }
}

}  // namespace myqc

using namespace myqc;

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
    int rc = dalloc(ctx, (void **)&dS, sizeof(double) * nn);
    if (rc == 0) rc = dalloc(ctx, (void **)&dH, sizeof(double) * nn);
    if (rc == 0) rc = dalloc(ctx, (void **)&dA, sizeof(double) * atoms.size());
    if (rc == 0 && hipMemcpyAsync(dA, atoms.data(), sizeof(double) * atoms.size(), hipMemcpyHostToDevice,
                                  ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "upload failed");
    if (rc == 0) {
        k_one_e<<<(np + 127) / 128, 128, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims, ctx->d_ftT, dA, ctx->nnuc, np,
                                                          n, dS, dH);
        if (hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "one-electron kernel launch failed");
    }
    if (rc == 0 && (d2h(ctx, S, dS, sizeof(double) * nn) != hipSuccess ||
                    d2h(ctx, H, dH, sizeof(double) * nn) != hipSuccess))
        rc = set_error(ctx, MYQC_E_HIP, "download failed");
    dfree((void **)&dS);
    dfree((void **)&dH);
    dfree((void **)&dA);
    return rc;
}

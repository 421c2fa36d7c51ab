// sp2.hip -- density-matrix purification on gfx950 FP64 matrix cores.
//
// The RHF driver needs, each iteration, the projector onto the nocc lowest
// levels of the orthonormal-basis Fock matrix F.  Trace-correcting
// purification (Niklasson, Phys. Rev. B 66, 155115 (2002)) reaches it with
// matrix products only:
//     P0 = (emax I - F) / (emax - emin)           (Gershgorin bounds)
//     P <- P^2  or  2P - P^2, whichever trace lands nearer nocc.
// For symmetric P, tr(P^2) = ||P||_F^2, so both traces the choice needs are
// known before the product; one kernel per step computes P^2 with
// v_mfma_f64_16x16x4f64 (one wavefront per 16x16 tile), applies the choice
// in its epilogue, writes the next P and accumulates that P's trace and
// Frobenius norm for the following step.  Three rotating scalar slots carry
// the traces between launches without host synchronisation.
#include <cmath>

#include "context.h"

namespace myqc {

typedef double d4_t __attribute__((ext_vector_type(4)));

// slots: 3 x {tr P, ||P||_F^2}; bounds at slots[6..7]
__global__ __launch_bounds__(64) void k_gersh_rows(const double *__restrict__ F, int n, double *__restrict__ lohi) {
    const int i = blockIdx.x, lane = threadIdx.x;
    double r = 0.0;
    for (int j = lane; j < n; j += 64)
        if (j != i) r += fabs(F[j + (size_t)n * i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    if (lane == 0) {
        const double c = F[i + (size_t)n * i];
        lohi[2 * i] = c - r;
        lohi[2 * i + 1] = c + r;
    }
}

__global__ __launch_bounds__(1024) void k_gersh_reduce(const double *__restrict__ lohi, int n,
                                                       double *__restrict__ slots) {
    __shared__ double smin[1024], smax[1024];
    double lo = 1e300, hi = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        lo = fmin(lo, lohi[2 * i]);
        hi = fmax(hi, lohi[2 * i + 1]);
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        slots[6] = smin[0];
        slots[7] = smax[0];
        for (int s = 0; s < 6; ++s) slots[s] = 0.0;
    }
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// P0 = (emax I - F)/(emax - emin), with its trace and Frobenius norm into slot 0
__global__ __launch_bounds__(64) void k_sp2_start(const double *__restrict__ F, double *__restrict__ P, int n,
                                                 double *__restrict__ slots) {
    const size_t e = (size_t)blockIdx.x * 64 + threadIdx.x;
    const double lo = slots[6], hi = slots[7];
    const double s = 1.0 / (hi - lo);
    double t = 0.0, f = 0.0;
    if (e < (size_t)n * n) {
        const int i = (int)(e % n), j = (int)(e / n);
        const double v = ((i == j ? hi : 0.0) - F[e]) * s;
        P[e] = v;
        f = v * v;
        if (i == j) t = v;
    }
    t = wsum(t);
    f = wsum(f);
    if (threadIdx.x == 0) {
        unsafeAtomicAdd(&slots[0], t);
        unsafeAtomicAdd(&slots[1], f);
    }
}

// One purification step k: Pn = choice(P^2, 2P - P^2) from the traces in
// slot k%3; Pn's traces accumulate into slot (k+1)%3; slot (k+2)%3 (last read
// by step k-1) is cleared for step k+1.  Grid: one workgroup of KS wavefronts
// per 16x16 output tile, wavefront w summing the k range [w*kc, (w+1)*kc) so
// that each dependent MFMA chain is n/(4 KS) long; the partial tiles meet in
// LDS.  Lane l feeds A row l%16 and B column l%16 at k offset l/16 and
// receives D rows l/16 + 4v, column l%16, v = 0..3 (the layout the GPU test of
// the builtin checks).
constexpr int KS = 4;

__global__ __launch_bounds__(64 * KS) void k_sp2_step_mfma(const double *__restrict__ P, double *__restrict__ Pn,
                                                          int n, double nocc, double *__restrict__ slots, int step) {
    __shared__ double s_acc[KS - 1][4][64];
    const int tiles = (n + 15) >> 4;
    const int ti = blockIdx.x % tiles, tj = blockIdx.x / tiles;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, kq = lane >> 4;
    const int ia = ti * 16 + r;     // row of P fed as A
    const int jb = tj * 16 + r;     // column of P fed as B (read as a row: P is symmetric)
    const int kc = ((n + 4 * KS - 1) / (4 * KS)) * 4;
    const int kb = w * kc, ke = min(n, kb + kc);
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 3
    for (int k0 = kb; k0 < ke; k0 += 4) {
        const int k = k0 + kq;
        const double a = (ia < n && k < ke) ? P[ia + (size_t)n * k] : 0.0;
        const double b = (jb < n && k < ke) ? P[jb + (size_t)n * k] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (w > 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_acc[w - 1][v][lane] = acc[v];
    }
    __syncthreads();
    if (w > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] += s_acc[q][v][lane];
    const double *sl = slots + 2 * (step % 3);
    const double tr = sl[0], fro = sl[1];
    const bool square = fabs(fro - nocc) < fabs(2.0 * tr - fro - nocc);
    double t = 0.0, f = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = ti * 16 + kq + 4 * v, col = tj * 16 + r;
        if (row < n && col < n) {
            const size_t e = row + (size_t)n * col;
            const double p2 = acc[v];
            const double nw = square ? p2 : 2.0 * P[e] - p2;
            Pn[e] = nw;
            f += nw * nw;
            if (row == col) t += nw;
        }
    }
    t = wsum(t);
    f = wsum(f);
    if (lane == 0) {
        double *nx = slots + 2 * ((step + 1) % 3);
        unsafeAtomicAdd(&nx[0], t);
        unsafeAtomicAdd(&nx[1], f);
        if (blockIdx.x == 0) {
            double *cl = slots + 2 * ((step + 2) % 3);
            cl[0] = 0.0;
            cl[1] = 0.0;
        }
    }
}

// Gershgorin bounds of the symmetric A into slots[6..7]; clears slots[0..5]
int gershgorin_bounds(myqc_ctx *ctx, const double *A, int n, double *lohi, double *slots) {
    k_gersh_rows<<<n, 64, 0, ctx->stream>>>(A, n, lohi);
    k_gersh_reduce<<<1, 1024, 0, ctx->stream>>>(lohi, n, slots);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// Projector onto the nocc lowest levels of the symmetric Fo (n x n) into P.
// work: n x n scratch; lohi: 2n; slots: 8 doubles.  Returns 1 in *ok when
// purification converged (idempotency below 1e-8, then two more steps).
// *steps carries the step count of the previous call in and this call's out:
// successive SCF iterations need nearly the same count, so the host checks
// convergence (a synchronising read of two scalars) only from a few steps
// before it.
int sp2_purify(myqc_ctx *ctx, const double *Fo, double *P, double *work, double *lohi, double *slots, int n, int nocc,
               int *ok, int *steps) {
    *ok = 0;
    MYQC_CHECK(gershgorin_bounds(ctx, Fo, n, lohi, slots));
    const size_t nn = (size_t)n * n;
    k_sp2_start<<<(unsigned)((nn + 63) / 64), 64, 0, ctx->stream>>>(Fo, P, n, slots);
    const int tiles = (n + 15) / 16;
    double *cur = P, *nxt = work;
    int extra = -1;
    const int first = *steps > 16 ? *steps - 4 : 12;
    *steps = 0;
    for (int step = 0; step < 120; ++step) {
        if (extra < 0 && step >= first && ((step - first) & 1) == 0) {
            double sl[2];
            MYQC_HIP(ctx, d2h(ctx, sl, slots + 2 * (step % 3), sizeof sl));
            // tr P - ||P||^2 = sum l(1-l) over P's eigenvalues l; near
            // convergence each step squares it, so two more steps finish
            if (fabs(sl[0] - sl[1]) < 1e-8 && fabs(sl[0] - nocc) < 1e-6) extra = 2;
        }
        if (extra == 0) {
            *ok = 1;
            *steps = step;
            break;
        }
        k_sp2_step_mfma<<<tiles * tiles, 64 * KS, 0, ctx->stream>>>(cur, nxt, n, (double)nocc, slots, step);
        double *t = cur;
        cur = nxt;
        nxt = t;
        if (extra > 0) --extra;
    }
    if (cur != P) MYQC_HIP(ctx, hipMemcpyAsync(P, cur, sizeof(double) * nn, hipMemcpyDeviceToDevice, ctx->stream));
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

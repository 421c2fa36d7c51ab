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
// in its epilogue, writes the next P and that P's per-tile trace and
// Frobenius-norm partials for the following step, so the traces pass between
// launches without host synchronisation.
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.h"
#include "many.h"


namespace myqc {

typedef double d4_t __attribute__((ext_vector_type(4)));

// slots: bounds at slots[6..7]; control words at slots + 8
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
        reinterpret_cast<int *>(slots + 8)[0] = -1;   // purification control word
    }
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// P0 = (emax I - F)/(emax - emin): one workgroup per 16x16 tile, writing the
// tile's trace and Frobenius-norm partials to tout (read by step 0)
__global__ __launch_bounds__(256) void k_sp2_start(const double *__restrict__ F, double *__restrict__ P, int n,
                                                  const double *__restrict__ slots, double *__restrict__ tout) {
    __shared__ double s_t[4], s_f[4];
    const int tiles = (n + 15) >> 4, tile = blockIdx.x;
    const int i = (tile % tiles) * 16 + (threadIdx.x & 15), j = (tile / tiles) * 16 + (threadIdx.x >> 4);
    const double lo = slots[6], hi = slots[7];
    const double s = 1.0 / (hi - lo);
    double t = 0.0, f = 0.0;
    if (i < n && j < n) {
        const size_t e = i + (size_t)n * j;
        const double v = ((i == j ? hi : 0.0) - F[e]) * s;
        P[e] = v;
        f = v * v;
        if (i == j) t = v;
    }
    t = wsum(t);
    f = wsum(f);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_t[w] = t;
        s_f[w] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        tout[2 * tile] = s_t[0] + s_t[1] + s_t[2] + s_t[3];
        tout[2 * tile + 1] = s_f[0] + s_f[1] + s_f[2] + s_f[3];
    }
}

// One purification step k: Pn = choice(P^2, 2P - P^2).  Grid: one workgroup
// of KS wavefronts per 16x16 output tile, wavefront w summing the k range
// [w*kc, (w+1)*kc) so that each dependent MFMA chain is n/(4 KS) long; the
// partial tiles meet in LDS.  Lane l feeds A row l%16 and B column l%16 at k
// offset l/16 and receives D rows l/16 + 4v, column l%16, v = 0..3 (the layout
// the GPU test of the builtin checks).
//
// The traces of P that the choice needs arrive as per-tile partials written
// by step k-1 (tin, nin of them; step 0 reads those of k_sp2_start), and this
// step writes Pn's per-tile partials to tout: no atomics,
// and every workgroup sums the same partials in the same order, so all of
// them make the same choice.  Every load a step needs (the partials, the
// control word, its MFMA operands and its own tile of P) is issued before
// any of them is used, so a step costs about one memory round trip.
//
// Convergence is decided on the device: step k meets the criterion when k >=
// first and P_k is idempotent to 1e-8 with trace nocc to 1e-6 (tr P -
// ||P||^2 = sum l(1-l) over P's eigenvalues; near convergence each step
// squares it).  The first such step is recorded in ctrl[0]; two steps later
// the projector is final and the remaining launched steps return at once.
constexpr int KS = 8, KU = 9;

__device__ __forceinline__ void sp2_step_tile(const double *__restrict__ P, double *__restrict__ Pn, int n,
                                              double nocc, const double *__restrict__ tin, int nin,
                                              double *__restrict__ tout, int step, int first,
                                              int *__restrict__ ctrl, int tile) {
    __shared__ double s_acc[KS - 1][4][64];
    __shared__ double s_t[KS], s_f[KS];
    const int tiles = (n + 15) >> 4;
    const int ti = tile % tiles, tj = tile / tiles;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, kq = lane >> 4;
    const int ia = ti * 16 + r;     // row of P fed as A
    const int jb = tj * 16 + r;     // column of P fed as B (read as a row: P is symmetric)
    const int kc = ((n + 4 * KS - 1) / (4 * KS)) * 4;
    const int kb = w * kc, ke = min(n, kb + kc);
    double t_in = 0.0, f_in = 0.0;
    for (int i = threadIdx.x; i < nin; i += blockDim.x) {
        t_in += tin[2 * i];
        f_in += tin[2 * i + 1];
    }
    const int seen = __hip_atomic_load(&ctrl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double own[4] = {0.0, 0.0, 0.0, 0.0};
    if (w == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = ti * 16 + kq + 4 * v, col = tj * 16 + r;
            if (row < n && col < n) own[v] = P[row + (size_t)n * col];
        }
    }
    d4_t acc = {0.0, 0.0, 0.0, 0.0};
    for (int k00 = kb; k00 < ke; k00 += 4 * KU) {
        double a[KU], b[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
            const int k = k00 + 4 * u + kq;
            a[u] = (ia < n && k < ke) ? P[ia + (size_t)n * k] : 0.0;
            b[u] = (jb < n && k < ke) ? P[jb + (size_t)n * k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < KU; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    t_in = wsum(t_in);
    f_in = wsum(f_in);
    if (lane == 0) {
        s_t[w] = t_in;
        s_f[w] = f_in;
    }
    if (w > 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_acc[w - 1][v][lane] = acc[v];
    }
    __syncthreads();
    double tr = 0.0, fro = 0.0;
#pragma unroll
    for (int q = 0; q < KS; ++q) {
        tr += s_t[q];
        fro += s_f[q];
    }
    const bool crit = step >= first && fabs(tr - fro) < 1e-8 && fabs(tr - nocc) < 1e-6;
    const int cs = (seen >= 0 && seen < step) ? seen : (crit ? step : -1);
    if (cs >= 0 && step >= cs + 2) return;
    if (tile == 0 && threadIdx.x == 0 && seen < 0 && crit) ctrl[0] = step;
    if (w != 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[v] += s_acc[q][v][lane];
    const bool square = fabs(fro - nocc) < fabs(2.0 * tr - fro - nocc);
    double t = 0.0, f = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = ti * 16 + kq + 4 * v, col = tj * 16 + r;
        if (row < n && col < n) {
            const double p2 = acc[v];
            const double nw = square ? p2 : 2.0 * own[v] - p2;
            Pn[row + (size_t)n * col] = nw;
            f += nw * nw;
            if (row == col) t += nw;
        }
    }
    t = wsum(t);
    f = wsum(f);
    if (lane == 0) {
        tout[2 * tile] = t;
        tout[2 * tile + 1] = f;
    }
}

__global__ __launch_bounds__(64 * KS) void k_sp2_step_mfma(const double *__restrict__ P, double *__restrict__ Pn,
                                                          int n, double nocc, const double *__restrict__ tin,
                                                          int nin, double *__restrict__ tout, int step, int first,
                                                          int *__restrict__ ctrl) {
    sp2_step_tile(P, Pn, n, nocc, tin, nin, tout, step, first, ctrl, blockIdx.x);
}

// The batched step for up to kMany systems of one size.  P is symmetric, so
// so is P^2: a workgroup computes one 64 x 64 tile (ti, tj) with ti <= tj and
// writes it and, off the diagonal, its mirror.  Four wavefronts, each a
// 32 x 32 quarter (2 x 2 MFMA blocks, every A and B fragment read from LDS
// feeding two MFMAs).  The k range goes through LDS in slices of 32 (A rows
// and B columns of the tile, both read as columns of P; a diagonal tile
// reads one side only), double-buffered: the next slice's global loads are
// in flight while the current slice's MFMAs run, with one barrier per slice.
// Workgroups are dealt to the eight XCDs in turn (workgroup b runs on XCD
// b mod 8), and the grid is laid out so that every tile of system k runs on
// XCD k mod 8, where that system's P stays in one L2.
// Step 0 reads the nin0 trace partials of the start kernel, later steps the
// nup tiles' own (an off-diagonal tile's Frobenius part counts twice); the
// choice and the convergence record are as in sp2_step_tile.
constexpr int SYT = 64, SYK = 32;

__global__ __launch_bounds__(256) void k_sp2_step_sym(Sp2Many a, int K, int n, int stride, int nin0, int step,
                                                     int first) {
    __shared__ double sA[2][SYK][SYT + 1], sB[2][SYK][SYT + 1];
    __shared__ double s_t[4], s_f[4];
    const int tiles = (n + SYT - 1) / SYT, nup = tiles * (tiles + 1) / 2;
    const int xcd = blockIdx.x & 7, j = blockIdx.x >> 3;
    const int k = xcd + 8 * (j / nup), tile = j % nup;
    if (k >= K) return;   // (uniform over the workgroup)
    int ti = 0, rem = tile;
    while (rem >= tiles - ti) {
        rem -= tiles - ti;
        ++ti;
    }
    const int tj = ti + rem;
    const bool diag = ti == tj;
    const double *P = (step & 1) ? a.W[k] : a.P[k];
    double *Pn = (step & 1) ? a.P[k] : a.W[k];
    const double *tin = a.tp[k] + 2 * (size_t)stride * ((step + 1) & 1);
    double *tout = a.tp[k] + 2 * (size_t)stride * (step & 1);
    int *ctrl = reinterpret_cast<int *>(a.slots[k] + 8);
    const int nin = step == 0 ? nin0 : nup;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, kq = lane >> 4;
    const int wr = w & 1, wc = w >> 1;
    // the traces of P from the previous step's partials
    double t_in = 0.0, f_in = 0.0;
    for (int i = threadIdx.x; i < nin; i += blockDim.x) {
        t_in += tin[2 * i];
        f_in += tin[2 * i + 1];
    }
    t_in = wsum(t_in);
    f_in = wsum(f_in);
    if (lane == 0) {
        s_t[w] = t_in;
        s_f[w] = f_in;
    }
    const int seen = __hip_atomic_load(&ctrl[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const double tr = (s_t[0] + s_t[1]) + (s_t[2] + s_t[3]), fro = (s_f[0] + s_f[1]) + (s_f[2] + s_f[3]);
    const double nocc = a.nocc[k];
    const bool crit = step >= first && fabs(tr - fro) < 1e-8 && fabs(tr - nocc) < 1e-6;
    const int cs = (seen >= 0 && seen < step) ? seen : (crit ? step : -1);
    if (cs >= 0 && step >= cs + 2) return;   // (uniform over the workgroup)
    if (tile == 0 && threadIdx.x == 0 && seen < 0 && crit) ctrl[0] = step;
    const int row0 = ti * SYT, col0 = tj * SYT;
    // element q of a thread's share of one slice: side (q >= 8: B), column
    // kk of the slice, row i of the tile; 2 x 32 x 64 values, 16 per thread
    double pre[16];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = threadIdx.x + 256 * q;
            const int side = e >> 11, kk = (e >> 6) & (SYK - 1), i = e & 63;
            const int kg = k0 + kk, rg = (side ? col0 : row0) + i;
            pre[q] = (kg < n && rg < n && !(side && diag)) ? P[rg + (size_t)n * kg] : 0.0;
        }
    };
    d4_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = d4_t{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    int buf = 0;
    for (int k0 = 0; k0 < n; k0 += SYK, buf ^= 1) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int e = threadIdx.x + 256 * q;
            const int side = e >> 11, kk = (e >> 6) & (SYK - 1), i = e & 63;
            if (!side) sA[buf][kk][i] = pre[q];
            else if (!diag) sB[buf][kk][i] = pre[q];
        }
        __syncthreads();
        if (k0 + SYK < n) fetch(k0 + SYK);
        const double(*As)[SYT + 1] = sA[buf];
        const double(*Bs)[SYT + 1] = diag ? sA[buf] : sB[buf];
#pragma unroll
        for (int ks = 0; ks < SYK; ks += 4) {
            double fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) fa[x] = As[ks + kq][wr * 32 + x * 16 + r];
#pragma unroll
            for (int y = 0; y < 2; ++y) fb[y] = Bs[ks + kq][wc * 32 + y * 16 + r];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[y], acc[x][y], 0, 0, 0);
        }
    }
    // the choice, the new tile (and its mirror) and its trace partials
    const bool square = fabs(fro - nocc) < fabs(2.0 * tr - fro - nocc);
    double t = 0.0, f = 0.0;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = row0 + wr * 32 + x * 16 + kq + 4 * v, col = col0 + wc * 32 + y * 16 + r;
                if (row < n && col < n) {
                    const double p2 = acc[x][y][v];
                    const double nw = square ? p2 : 2.0 * P[row + (size_t)n * col] - p2;
                    Pn[row + (size_t)n * col] = nw;
                    if (!diag) Pn[col + (size_t)n * row] = nw;
                    f += nw * nw;
                    if (row == col) t += nw;
                }
            }
    if (!diag) f *= 2.0;
    t = wsum(t);
    f = wsum(f);
    __syncthreads();
    if (lane == 0) {
        s_t[w] = t;
        s_f[w] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        tout[2 * tile] = (s_t[0] + s_t[1]) + (s_t[2] + s_t[3]);
        tout[2 * tile + 1] = (s_f[0] + s_f[1]) + (s_f[2] + s_f[3]);
    }
}

// After the launched steps: status = {converged, final step}; the final
// projector (the input of step ctrl[0] + 2) is moved into P when it sits in W.
__global__ void k_sp2_finish(double *__restrict__ P, const double *__restrict__ W, int n, int *__restrict__ ctrl) {
    const int cs = ctrl[0];
    const size_t nn = (size_t)n * n;
    if (cs >= 0 && ((cs + 2) & 1))
        for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x)
            P[e] = W[e];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl[1] = cs >= 0 ? 1 : 0;
        ctrl[2] = cs >= 0 ? cs + 2 : -1;
    }
}

// --- the same for up to kMany systems of one size (blockIdx.y: system) ---

__global__ __launch_bounds__(64) void k_gersh_rows_many(Sp2Many a, int n) {
    const int k = blockIdx.y, i = blockIdx.x, lane = threadIdx.x;
    const double *F = a.F[k];
    double r = 0.0;
    for (int j = lane; j < n; j += 64)
        if (j != i) r += fabs(F[j + (size_t)n * i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    if (lane == 0) {
        const double c = F[i + (size_t)n * i];
        a.lohi[k][2 * i] = c - r;
        a.lohi[k][2 * i + 1] = c + r;
    }
}

__global__ __launch_bounds__(1024) void k_gersh_reduce_many(Sp2Many a, int n) {
    __shared__ double smin[1024], smax[1024];
    const int k = blockIdx.x;
    const double *lohi = a.lohi[k];
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
        double *slots = a.slots[k];
        slots[6] = smin[0];
        slots[7] = smax[0];
        for (int s = 0; s < 6; ++s) slots[s] = 0.0;
        reinterpret_cast<int *>(slots + 8)[0] = -1;
    }
}

__global__ __launch_bounds__(256) void k_sp2_start_many(Sp2Many a, int n, int tt) {
    __shared__ double s_t[4], s_f[4];
    const int k = blockIdx.y, tiles = (n + 15) >> 4, tile = blockIdx.x;
    const int i = (tile % tiles) * 16 + (threadIdx.x & 15), j = (tile / tiles) * 16 + (threadIdx.x >> 4);
    const double lo = a.slots[k][6], hi = a.slots[k][7];
    const double s = 1.0 / (hi - lo);
    double t = 0.0, f = 0.0;
    if (i < n && j < n) {
        const size_t e = i + (size_t)n * j;
        const double v = ((i == j ? hi : 0.0) - a.F[k][e]) * s;
        a.P[k][e] = v;
        f = v * v;
        if (i == j) t = v;
    }
    t = wsum(t);
    f = wsum(f);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_t[w] = t;
        s_f[w] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double *tout = a.tp[k] + 2 * (size_t)tt;   // plays step -1
        tout[2 * tile] = s_t[0] + s_t[1] + s_t[2] + s_t[3];
        tout[2 * tile + 1] = s_f[0] + s_f[1] + s_f[2] + s_f[3];
    }
}

// status[2k] = converged, status[2k+1] = final step of system k
__global__ void k_sp2_finish_many(Sp2Many a, int n, int *__restrict__ status) {
    const int k = blockIdx.y;
    int *ctrl = reinterpret_cast<int *>(a.slots[k] + 8);
    const int cs = ctrl[0];
    const size_t nn = (size_t)n * n;
    if (cs >= 0 && ((cs + 2) & 1))
        for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nn; e += (size_t)gridDim.x * blockDim.x)
            a.P[k][e] = a.W[k][e];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctrl[1] = cs >= 0 ? 1 : 0;
        ctrl[2] = cs >= 0 ? cs + 2 : -1;
        status[2 * k] = ctrl[1];
        status[2 * k + 1] = ctrl[2];
    }
}

// Gershgorin bounds of K systems' a.F into their a.slots[6..7] (a.lohi: 2 n each)
int gershgorin_many(myqc_ctx *ctx, const Sp2Many &a, int K, int n) {
    if (K < 1 || K > kMany) return set_error(ctx, MYQC_E_ARG, "gershgorin_many: %d systems", K);
    k_gersh_rows_many<<<dim3(n, K), 64, 0, ctx->stream>>>(a, n);
    k_gersh_reduce_many<<<K, 1024, 0, ctx->stream>>>(a, n);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// Purification of K systems of order n at once (the projectors into a.P):
// one launch per step for all of them; status (device, 2 K ints) and host
// ok[k] say which converged; *steps carries the step count in and out as in
// sp2_purify.
int sp2_purify_many(myqc_ctx *ctx, const Sp2Many &a, int K, int n, int *status, std::vector<int> &ok, int *steps) {
    if (K < 1 || K > kMany) return set_error(ctx, MYQC_E_ARG, "sp2_purify_many: %d systems", K);
    const int tiles = (n + 15) / 16, tt = tiles * tiles;
    const int t64 = (n + SYT - 1) / SYT, nup = t64 * (t64 + 1) / 2;
    if (nup > tt) return set_error(ctx, MYQC_E_ARG, "sp2_purify_many: n=%d", n);
    const unsigned grid = 8u * (unsigned)((K + 7) / 8) * (unsigned)nup;
    k_gersh_rows_many<<<dim3(n, K), 64, 0, ctx->stream>>>(a, n);
    k_gersh_reduce_many<<<K, 1024, 0, ctx->stream>>>(a, n);
    k_sp2_start_many<<<dim3(tt, K), 256, 0, ctx->stream>>>(a, n, tt);
    const int first = 8, max_steps = 120;
    int launched = 0, target = *steps > 0 ? std::min(*steps + 6, max_steps) : 48;
    std::vector<int> st(2 * K);
    ok.assign(K, 0);
    while (true) {
        for (; launched < target; ++launched)
            k_sp2_step_sym<<<grid, 256, 0, ctx->stream>>>(a, K, n, tt, tt, launched, first);
        k_sp2_finish_many<<<dim3(64, K), 256, 0, ctx->stream>>>(a, n, status);
        MYQC_HIP(ctx, hipGetLastError());
        MYQC_HIP(ctx, d2h(ctx, st.data(), status, sizeof(int) * 2 * K));
        bool all = true;
        for (int k = 0; k < K; ++k) all = all && st[2 * k];
        if (all || launched >= max_steps) break;
        target = std::min(launched + 8, max_steps);
    }
    int mx = 0;
    for (int k = 0; k < K; ++k) {
        ok[k] = st[2 * k];
        if (ok[k]) mx = std::max(mx, st[2 * k + 1]);
    }
    *steps = mx;
    return 0;
}

// Gershgorin bounds of the symmetric A into slots[6..7]; clears slots[0..5]
int gershgorin_bounds(myqc_ctx *ctx, const double *A, int n, double *lohi, double *slots) {
    k_gersh_rows<<<n, 64, 0, ctx->stream>>>(A, n, lohi);
    k_gersh_reduce<<<1, 1024, 0, ctx->stream>>>(lohi, n, slots);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// Projector onto the nocc lowest levels of the symmetric Fo (n x n) into P.
// work: n x n scratch; lohi: 2n; slots: 16 doubles (traces, bounds, and the
// control words at slots + 8); tpart: 4 ceil(n/16)^2 doubles (per-tile trace
// partials of two consecutive steps).  Returns 1 in *ok when
// purification converged (idempotency below 1e-8, then two more steps).
// *steps carries the step count of the previous call in and this call's out:
// successive SCF iterations need nearly the same count, so the host checks
// convergence (a synchronising read of two scalars) only from a few steps
// before it.
int sp2_purify(myqc_ctx *ctx, const double *Fo, double *P, double *work, double *lohi, double *slots, double *tpart,
               int n, int nocc, int *ok, int *steps) {
    *ok = 0;
    MYQC_CHECK(gershgorin_bounds(ctx, Fo, n, lohi, slots));
    const int tiles = (n + 15) / 16;
    const size_t tt = (size_t)tiles * tiles;
    // partials of step k go to tpart + 2 tt (k & 1); k_sp2_start plays step -1
    k_sp2_start<<<(unsigned)tt, 256, 0, ctx->stream>>>(Fo, P, n, slots, tpart + 2 * tt);
    // launch about as many steps as the previous purification took (plus a
    // margin); steps past convergence cost an empty launch, and only the
    // status read at the end waits for the device
    int *ctrl = reinterpret_cast<int *>(slots + 8);
    const int first = 8, max_steps = 120;
    int launched = 0, target = *steps > 0 ? std::min(*steps + 6, max_steps) : 48;
    *steps = 0;
    while (true) {
        for (; launched < target; ++launched) {
            const double *cur = (launched & 1) ? work : P;
            double *nxt = (launched & 1) ? P : work;
            const double *tin = tpart + 2 * tt * ((launched + 1) & 1);   // step launched - 1
            double *tout = tpart + 2 * tt * (launched & 1);
            k_sp2_step_mfma<<<(unsigned)tt, 64 * KS, 0, ctx->stream>>>(cur, nxt, n, (double)nocc, tin, (int)tt, tout,
                                                                      launched, first, ctrl);
        }
        k_sp2_finish<<<64, 256, 0, ctx->stream>>>(P, work, n, ctrl);
        MYQC_HIP(ctx, hipGetLastError());
        int st[2] = {0, 0};
        MYQC_HIP(ctx, d2h(ctx, st, ctrl + 1, sizeof st));
        if (st[0]) {
            *ok = 1;
            *steps = st[1];
            break;
        }
        if (launched >= max_steps) break;
        target = std::min(launched + 8, max_steps);
    }
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

// mfma_f64.hip -- the FP64 quarter transforms of ao2mo on gfx950 matrix
// cores (v_mfma_f64_16x16x4f64), replacing library GEMMs.
//
// The four quarter steps of slow_ao2mo_MP2_RHF (idx1..idx4_trans,
// src/ao2mo/ao2mo.f90:1306-1439), with orbital blocks Cp, Cq, Cr, Cs:
//   1+2  M(p,q,l,d) = sum_v [sum_u Cp(u,p) XX(u,v,l,d)] Cq(v,q)   (k_ao2mo12_f64)
//   3    N(pq,r,d)  = sum_l M(pq,l,d) Cr(l,r)                      (k_gemm_f64)
//   4    O(pq,r,s)  = sum_d N(pq,r,d) Cs(d,s)                      (k_gemm_f64)
//
// MFMA lane layout (16x16x4, checked by the GPU layout test): lane l = 16 g +
// c feeds A[c][g] of the 16x4 A and B[g][c] of the 4x16 B, and receives
// D[g + 4 v][c], v = 0..3.  Both kernels compute transposed products so that
// every store is 16 consecutive doubles, and in steps 1+2 the accumulator of
// the first product is already the B operand of the second: with
// L^T(v, p) = sum_u X^T(v, u) Cp(u, p) held as D tiles, element v of a lane's
// accumulator is L^T(v0 + 4 kb + g, p0 + c) for kb = v, which is exactly the
// B fragment of k-step kb of M^T(q, p) = sum_v Cq^T(q, v) L^T(v, p).  The o n^3
// intermediate L never leaves the registers.
#include <cstdint>

#include "context.h"
#include "many.h"

namespace myqc {

typedef double d4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4_t mfma16x16x4(double a, double b, d4_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Quarter steps 1 and 2, fused.
//
// Workgroup: 8 wavefronts, wavefront w owning the 16 orbitals p of tile
// p0 = 128 blockIdx.y + 16 w, and the 16 NAT orbitals q from 16 NAT blockIdx.z.
// Its B fragments of Cp (k-steps of u) stay in registers for the whole launch.
// The workgroup walks (l,d) blocks blockIdx.x, + gridDim.x, ... and, within
// each, v tiles of 16 columns: the tile X(:, v0..v0+15 | l,d) is 16 n
// contiguous doubles, copied into LDS (double-buffered, the next tile's loads
// in flight during this tile's MFMAs) and shared by the 8 wavefronts.  Per v
// tile a wavefront runs NKS MFMAs of step 1 (two accumulators over even and
// odd k-steps) and 4 NAT of step 2; after the last v tile of an (l,d) block it
// stores M^T and clears the accumulators.
// ---------------------------------------------------------------------------
constexpr int Q12_WAVES = 8, Q12_THREADS = 64 * Q12_WAVES, Q12_NAT = 3;

template <int NKS>
__global__ __launch_bounds__(Q12_THREADS) void k_ao2mo12_f64(int np, int nq, int n, const double *__restrict__ Cp,
                                                             const double *__restrict__ Cq,
                                                             const double *__restrict__ XX, double *__restrict__ M,
                                                             long long nld) {
    constexpr int LDX = 4 * NKS + 2;                           // padded column of a tile in LDS
    constexpr int XPT = (16 * 4 * NKS + Q12_THREADS - 1) / Q12_THREADS;   // tile doubles per thread
    __shared__ double sX[2][16 * LDX];
    // Cq^T of this block of q, [q][v], in LDS when it fits beside the X tiles
    constexpr bool QLDS = NKS <= 48;
    __shared__ double sQ[QLDS ? 16 * Q12_NAT * LDX : 1];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, g = lane >> 4, c = lane & 15;
    const int p0 = blockIdx.y * (16 * Q12_WAVES) + 16 * w;
    const int q0 = blockIdx.z * (16 * Q12_NAT);
    const size_t nn = (size_t)n * n;
    const int nvt = (n + 15) >> 4;
    // rows u >= n of both buffers stay zero (Cp's fragments are zero there too)
    for (int e = t; e < 2 * 16 * LDX; e += Q12_THREADS) (&sX[0][0])[e] = 0.0;
    if constexpr (QLDS)
        for (int e = t; e < 16 * Q12_NAT * LDX; e += Q12_THREADS) {
            const int ql = e / LDX, v = e - ql * LDX, q = q0 + ql;
            sQ[e] = (v < n && q < nq) ? Cq[v + (size_t)n * q] : 0.0;
        }
    // Cp fragments: B[g][c] of k-step k = Cp(4 k + g, p0 + c)
    double cf[NKS];
#pragma unroll
    for (int k = 0; k < NKS; ++k) {
        const int u = 4 * k + g;
        cf[k] = (u < n && p0 + c < np) ? Cp[u + (size_t)n * (p0 + c)] : 0.0;
    }
    // the work items of this workgroup: (ld, vt) for ld = blockIdx.x + j gridDim.x
    const long long nitem_ld = (nld - blockIdx.x + gridDim.x - 1) / gridDim.x;
    const long long nitem = nitem_ld * nvt;
    double xr[XPT];
    auto load_tile = [&](long long it) {
        const long long ld = blockIdx.x + (it / nvt) * gridDim.x;
        const int v0 = (int)(it % nvt) * 16;
        const double *src = XX + ld * nn + (size_t)v0 * n;
        const int valid = min(16, n - v0) * n;   // doubles of the tile inside the block
#pragma unroll
        for (int r = 0; r < XPT; ++r) {
            const int e = t + r * Q12_THREADS;
            xr[r] = e < valid ? src[e] : 0.0;
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int r = 0; r < XPT; ++r) {
            const int e = t + r * Q12_THREADS;
            if (e < 16 * n) {
                const int col = e / n, u = e - col * n;
                sX[buf][col * LDX + u] = xr[r];
            }
        }
    };
    d4_t macc[Q12_NAT];
#pragma unroll
    for (int s = 0; s < Q12_NAT; ++s) macc[s] = (d4_t){0.0, 0.0, 0.0, 0.0};
    if (nitem > 0) load_tile(0);
    __syncthreads();   // the zero fill is complete before the first tile lands
    if (nitem > 0) store_tile(0);
    __syncthreads();
    int cur = 0;
    for (long long it = 0; it < nitem; ++it) {
        if (it + 1 < nitem) load_tile(it + 1);
        const long long ld = blockIdx.x + (it / nvt) * gridDim.x;
        const int vt = (int)(it % nvt), v0 = vt * 16;
        // step 1: L^T(v0 + ., p0 + .) over all u, four independent chains
        const double *xs = &sX[cur][c * LDX + g];
        d4_t ac[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) ac[k] = (d4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NKS; ++k) ac[k & 3] = mfma16x16x4(xs[4 * k], cf[k], ac[k & 3]);
        const d4_t lt = (ac[0] + ac[1]) + (ac[2] + ac[3]);
        // step 2: M^T(q, p) += sum_v Cq^T(q, v) L^T(v, p) over this tile's v
#pragma unroll
        for (int s = 0; s < Q12_NAT; ++s) {
            if constexpr (QLDS) {
                const double *qs = &sQ[(16 * s + c) * LDX + v0 + g];
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) macc[s] = mfma16x16x4(qs[4 * kb], lt[kb], macc[s]);
            } else {
                const int q = q0 + 16 * s + c;
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) {
                    const int v = v0 + 4 * kb + g;
                    const double qa = (v < n && q < nq) ? Cq[v + (size_t)n * q] : 0.0;
                    macc[s] = mfma16x16x4(qa, lt[kb], macc[s]);
                }
            }
        }
        if (vt == nvt - 1) {
            // M(p, q, l, d) at p + np (q + nq ld)
#pragma unroll
            for (int s = 0; s < Q12_NAT; ++s) {
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    const int q = q0 + 16 * s + g + 4 * vv, p = p0 + c;
                    if (p < np && q < nq) M[p + (size_t)np * (q + (size_t)nq * ld)] = macc[s][vv];
                }
                macc[s] = (d4_t){0.0, 0.0, 0.0, 0.0};
            }
        }
        if (it + 1 < nitem) store_tile(cur ^ 1);
        __syncthreads();
        cur ^= 1;
    }
}

// ---------------------------------------------------------------------------
// Steps 3 and 4: C (m x nb, ldc) = A (m x k, lda) B (k x nb, ldb) for `batch`
// problems (A += bat sA, C += bat sC; B shared), all column-major FP64.
// A workgroup of 4 wavefronts owns 128 rows of one problem (blockIdx.x) and
// 48 columns (blockIdx.y); each wavefront 2 row tiles x 3 column tiles, as
// C^T = B^T A^T so that lanes store 16 consecutive rows.  A's fragments are
// 16 consecutive rows at 4 k each (coalesced), B's come from L1/L2 (B is a
// small block of orbital coefficients); the next k-step's fragments are
// loaded while this one's MFMAs run.
// ---------------------------------------------------------------------------
constexpr int GF_MT = 2, GF_NT = 3, GF_THREADS = 256;

__global__ __launch_bounds__(GF_THREADS) void k_gemm_f64(int m, int nb, int k, const double *__restrict__ A,
                                                         long long lda, long long sA, const double *__restrict__ B,
                                                         long long ldb, double *__restrict__ C, long long ldc,
                                                         long long sC, int tiles_m) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
    const long long bat = blockIdx.x / tiles_m;
    const int r0 = (int)(blockIdx.x - bat * tiles_m) * (64 * GF_MT) + 16 * GF_MT * w;
    const int j0 = blockIdx.y * (16 * GF_NT);
    A += bat * sA;
    C += bat * sC;
    d4_t acc[GF_MT][GF_NT];
#pragma unroll
    for (int i = 0; i < GF_MT; ++i)
#pragma unroll
        for (int j = 0; j < GF_NT; ++j) acc[i][j] = (d4_t){0.0, 0.0, 0.0, 0.0};
    double fa[GF_MT], fb[GF_NT];
    auto load = [&](int k0) {
        const int kk = k0 + g;
#pragma unroll
        for (int i = 0; i < GF_MT; ++i) {
            const int r = r0 + 16 * i + c;
            fa[i] = (r < m && kk < k) ? A[r + lda * kk] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < GF_NT; ++j) {
            const int col = j0 + 16 * j + c;
            fb[j] = (col < nb && kk < k) ? B[kk + ldb * col] : 0.0;
        }
    };
    load(0);
    for (int k0 = 0; k0 < k; k0 += 4) {
        double a[GF_MT], b[GF_NT];
#pragma unroll
        for (int i = 0; i < GF_MT; ++i) a[i] = fa[i];
#pragma unroll
        for (int j = 0; j < GF_NT; ++j) b[j] = fb[j];
        if (k0 + 4 < k) load(k0 + 4);
#pragma unroll
        for (int i = 0; i < GF_MT; ++i)
#pragma unroll
            for (int j = 0; j < GF_NT; ++j) acc[i][j] = mfma16x16x4(b[j], a[i], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < GF_MT; ++i)
#pragma unroll
        for (int j = 0; j < GF_NT; ++j)
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int r = r0 + 16 * i + c, col = j0 + 16 * j + g + 4 * vv;
                if (r < m && col < nb) C[r + ldc * col] = acc[i][j][vv];
            }
}

template <int NKS>
static void launch12(dim3 grid, hipStream_t s, int np, int nq, int n, const double *Cp, const double *Cq,
                     const double *XX, double *M, long long nld) {
    k_ao2mo12_f64<NKS><<<grid, Q12_THREADS, 0, s>>>(np, nq, n, Cp, Cq, XX, M, nld);
}

// Steps 1+2 in FP64: M (np x nq x n x n) from Cp (n x np), Cq (n x nq), both
// with leading dimension n, and the dense integrals XX (n^4).
int ao2mo12_f64(myqc_ctx *ctx, int np, const double *Cp, int nq, const double *Cq, int n, const double *XX,
                double *M) {
    if (np <= 0 || nq <= 0 || n <= 0) return set_error(ctx, MYQC_E_ARG, "ao2mo12_f64: np=%d nq=%d n=%d", np, nq, n);
    const int nks = (n + 3) / 4;
    const long long nld = (long long)n * n;
    const int gy = (np + 16 * Q12_WAVES - 1) / (16 * Q12_WAVES), gz = (nq + 16 * Q12_NAT - 1) / (16 * Q12_NAT);
    if (gy > 65535 || gz > 65535) return set_error(ctx, MYQC_E_ARG, "ao2mo12_f64: too many orbitals");
    // one workgroup per CU and orbital block (256 CUs), each walking its (l,d) blocks
    const long long gx = std::min<long long>(nld, std::max(1LL, 256LL / ((long long)gy * gz)));
    const dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)gz);
    hipStream_t s = ctx->stream;
    if (nks <= 8) launch12<8>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 16) launch12<16>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 24) launch12<24>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 32) launch12<32>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 48) launch12<48>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 64) launch12<64>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 80) launch12<80>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else if (nks <= 100) launch12<100>(grid, s, np, nq, n, Cp, Cq, XX, M, nld);
    else return set_error(ctx, MYQC_E_ARG, "ao2mo12_f64: n=%d above 400", n);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// C = A B (FP64), `batch` problems of A (m x k, lda, stride sA) against the
// shared B (k x nb, ldb), into C (m x nb, ldc, stride sC).
int gemm_f64(myqc_ctx *ctx, int m, int nb, int k, const double *A, long long lda, long long sA, const double *B,
             long long ldb, double *C, long long ldc, long long sC, int batch) {
    if (m <= 0 || nb <= 0 || k <= 0 || batch <= 0) return set_error(ctx, MYQC_E_ARG, "gemm_f64: empty problem");
    const long long tiles_m = ((long long)m + 64 * GF_MT - 1) / (64 * GF_MT);
    const long long gx = tiles_m * batch;
    const int gy = (nb + 16 * GF_NT - 1) / (16 * GF_NT);
    if (gx > 0x7fffffffLL || tiles_m > 0x7fffffffLL || gy > 65535)
        return set_error(ctx, MYQC_E_ARG, "gemm_f64: grid too large");
    k_gemm_f64<<<dim3((unsigned)gx, (unsigned)gy), GF_THREADS, 0, ctx->stream>>>(m, nb, k, A, lda, sA, B, ldb, C, ldc,
                                                                                sC, (int)tiles_m);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// The same product for many problems in 64 x 64 output tiles: four
// wavefronts, each a 32 x 32 quarter (2 x 2 MFMA blocks); the k range goes
// through LDS in slices of 32, double-buffered (the next slice's global loads
// in flight during the current slice's MFMAs, one barrier per slice).  A is
// read down its columns (64 rows of one column contiguous), B down its
// columns too (32 k of one column contiguous), both stored [k][row or column]
// in LDS.  Workgroup b runs on XCD b mod 8; every tile of problem z runs on
// XCD z mod 8, so that its operands stay in one L2.
constexpr int GMT = 64, GMK = 32;

__global__ __launch_bounds__(256) void k_gemm_f64_many64(int m, int nb, int k, GemmMany g, double alpha, int nz,
                                                         int tm, int tn) {
    __shared__ double sA[2][GMK][GMT + 1], sB[2][GMK][GMT + 1];
    const int tiles = tm * tn;
    const int xcd = blockIdx.x & 7, jb = blockIdx.x >> 3;
    const int z = xcd + 8 * (jb / tiles), tile = jb % tiles;
    if (z >= nz) return;   // (uniform over the workgroup)
    const int ti = tile % tm, tj = tile / tm;
    const double *A = g.A[z], *B = g.B[z];
    double *C = g.C[z];
    const int row0 = ti * GMT, col0 = tj * GMT;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, r = lane & 15, kq = lane >> 4;
    const int wr = w & 1, wc = w >> 1;
    double pa[8], pb[8];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = t + 256 * q;
            const int ka = k0 + (e >> 6), ia = row0 + (e & 63);      // A: 64 rows of column ka
            pa[q] = (ia < m && ka < k) ? A[ia + (size_t)m * ka] : 0.0;
            const int kb = k0 + (e & 31), jc = col0 + (e >> 5);      // B: 32 k of column jc
            pb[q] = (jc < nb && kb < k) ? B[kb + (size_t)k * jc] : 0.0;
        }
    };
    d4_t acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (d4_t){0.0, 0.0, 0.0, 0.0};
    fetch(0);
    int buf = 0;
    for (int k0 = 0; k0 < k; k0 += GMK, buf ^= 1) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = t + 256 * q;
            sA[buf][e >> 6][e & 63] = pa[q];
            sB[buf][e & 31][e >> 5] = pb[q];
        }
        __syncthreads();
        if (k0 + GMK < k) fetch(k0 + GMK);
#pragma unroll
        for (int ks = 0; ks < GMK; ks += 4) {
            double fa[2], fb[2];
#pragma unroll
            for (int x = 0; x < 2; ++x) fa[x] = sA[buf][ks + kq][wr * 32 + x * 16 + r];
#pragma unroll
            for (int y = 0; y < 2; ++y) fb[y] = sB[buf][ks + kq][wc * 32 + y * 16 + r];
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) acc[x][y] = mfma16x16x4(fa[x], fb[y], acc[x][y]);
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = row0 + wr * 32 + x * 16 + kq + 4 * v, col = col0 + wc * 32 + y * 16 + r;
                if (row < m && col < nb) C[row + (size_t)m * col] = alpha * acc[x][y][v];
            }
}

// C_z = alpha A_z B_z for nz problems (nz <= kMany), square or not:
// A_z m x k, B_z k x nb, C_z m x nb, all with their row counts as leading
// dimensions.
int gemm_f64_many(myqc_ctx *ctx, int m, int nb, int k, const GemmMany &g, int nz, double alpha) {
    if (m <= 0 || nb <= 0 || k <= 0 || nz <= 0 || nz > kMany)
        return set_error(ctx, MYQC_E_ARG, "gemm_f64_many: m=%d nb=%d k=%d problems %d", m, nb, k, nz);
    const int tm = (m + GMT - 1) / GMT, tn = (nb + GMT - 1) / GMT;
    const long long grid = 8LL * ((nz + 7) / 8) * tm * tn;
    if (grid > 0x7fffffffLL) return set_error(ctx, MYQC_E_ARG, "gemm_f64_many: grid too large");
    k_gemm_f64_many64<<<(unsigned)grid, 256, 0, ctx->stream>>>(m, nb, k, g, alpha, nz, tm, tn);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

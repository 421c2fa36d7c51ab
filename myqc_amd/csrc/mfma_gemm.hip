// mfma_gemm.hip -- split-precision (bf16 hi/lo) GEMM on gfx950 MFMA with FP32
// accumulation, for the ao2mo quarter transforms.
//
//   C = op(A) B,   op(A) = A or A^T,   all column-major FP64 in HBM.
//
// Each FP64 operand element x is split on its way into LDS as
//   hi = bf16(x),   lo = bf16(x - hi)            (16 significant bits in all)
// and the product is accumulated as  hi*hi  +  (hi*lo + lo*hi)  with
// v_mfma_f32_16x16x32_bf16; the dropped lo*lo term is 2^-16 relative.  The
// main and correction sums live in separate FP32 accumulators and are added
// in FP64 at the store.  Relative error per output ~1e-5 (bf16 split) plus
// FP32 accumulation over k; the ao2mo/MP2 test states it against FP64.
//
// Tiling: a 256-thread workgroup owns a 128 x 64 tile of C; wavefront w
// computes rows 64 (w & 1) .. +63 and columns 32 (w >> 1) .. +31 as 4 x 2
// MFMA 16x16 sub-tiles; K advances 32 at a time through LDS.  MFMA operand
// layout (16x16x32 bf16): lane l supplies row (or column) l % 16 at
// k = 8 (l / 16) .. +7; the FP32 result of lane l, element v, is
// D[4 (l / 16) + v][l % 16] (checked by the GPU layout test).
#include <cstdint>

#include "context.h"

namespace myqc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GM = 128, GN = 64, GK = 32, GKP = GK + 8;   // LDS rows padded to 80 bytes

// round-to-nearest-even float -> bf16 bits (finite inputs)
__device__ __forceinline__ uint16_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void split(double x, uint16_t &hi, uint16_t &lo) {
    const float f = (float)x;
    hi = bf16_bits(f);
    lo = bf16_bits(f - bf16_value(hi));
}

template <bool TRANSA>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(int m, int n, int k, const double *__restrict__ A, long long lda,
                                                     long long sA, const double *__restrict__ B, long long ldb,
                                                     long long sB, double *__restrict__ C, long long ldc, long long sC,
                                                     long long tilesN) {
    __shared__ __attribute__((aligned(16))) uint16_t sAh[GM][GKP], sAl[GM][GKP], sBh[GN][GKP], sBl[GN][GKP];
    const long long bx = blockIdx.x;
    const long long bat = bx / tilesN, tn = bx - bat * tilesN;
    const int m0 = blockIdx.y * GM;
    const long long n0 = tn * GN;
    A += bat * sA;
    B += bat * sB;
    C += bat * sC;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 32;
    const int r16 = lane & 15, q = lane >> 4;
    f32x4 acc[4][2], cor[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            cor[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
    for (int k0 = 0; k0 < k; k0 += GK) {
        // A tile (GM x GK) -> sA[row][k]
        if (TRANSA) {
            // op(A)(i, kk) = A[kk + lda i]: 16 consecutive k of one row per thread
            const int i = t >> 1, kb = (t & 1) * 16;
            const int gi = m0 + i;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gk = k0 + kb + e;
                const double x = (gi < m && gk < k) ? A[gk + lda * gi] : 0.0;
                split(x, sAh[i][kb + e], sAl[i][kb + e]);
            }
        } else {
            // op(A)(i, kk) = A[i + lda kk]: 16 consecutive rows at one k per thread
            const int kk = t >> 3, ib = (t & 7) * 16;
            const int gk = k0 + kk;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gi = m0 + ib + e;
                const double x = (gi < m && gk < k) ? A[gi + lda * gk] : 0.0;
                split(x, sAh[ib + e][kk], sAl[ib + e][kk]);
            }
        }
        // B tile (GK x GN) -> sB[col][k]: 8 consecutive k of one column per thread
        {
            const int j = t >> 2, kb = (t & 3) * 8;
            const long long gj = n0 + j;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int gk = k0 + kb + e;
                const double x = (gj < n && gk < k) ? B[gk + ldb * gj] : 0.0;
                split(x, sBh[j][kb + e], sBl[j][kb + e]);
            }
        }
        __syncthreads();
        bf16x8 bh[2], bl[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            bh[j] = *reinterpret_cast<const bf16x8 *>(&sBh[wn + 16 * j + r16][8 * q]);
            bl[j] = *reinterpret_cast<const bf16x8 *>(&sBl[wn + 16 * j + r16][8 * q]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(&sAh[wm + 16 * i + r16][8 * q]);
            const bf16x8 al = *reinterpret_cast<const bf16x8 *>(&sAl[wm + 16 * i + r16][8 * q]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
                cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], cor[i][j], 0, 0, 0);
                cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], cor[i][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gi = m0 + wm + 16 * i + 4 * q + v;
                const long long gj = n0 + wn + 16 * j + r16;
                if (gi < m && gj < n) C[gi + ldc * gj] = (double)acc[i][j][v] + (double)cor[i][j][v];
            }
}

// C (m x n, ldc) = op(A) (m x k) * B (k x n, ldb) for `batch` strided
// problems; FP64 in and out, split-bf16 MFMA products inside.
int gemm_bf16x3(myqc_ctx *ctx, bool transA, int m, int n, int k, const double *A, int lda, long long strideA,
                const double *B, int ldb, long long strideB, double *C, int ldc, long long strideC, int batch) {
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return set_error(ctx, MYQC_E_ARG, "gemm_bf16x3: empty problem");
    const long long tilesN = ((long long)n + GN - 1) / GN;
    const long long gx = tilesN * batch;
    const int gy = (m + GM - 1) / GM;
    if (gx > 0x7fffffffLL || gy > 65535) return set_error(ctx, MYQC_E_ARG, "gemm_bf16x3: grid too large");
    const dim3 grid((unsigned)gx, (unsigned)gy);
    if (transA)
        k_gemm_bf16x3<true><<<grid, 256, 0, ctx->stream>>>(m, n, k, A, lda, strideA, B, ldb, strideB, C, ldc, strideC,
                                                            tilesN);
    else
        k_gemm_bf16x3<false><<<grid, 256, 0, ctx->stream>>>(m, n, k, A, lda, strideA, B, ldb, strideB, C, ldc, strideC,
                                                             tilesN);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

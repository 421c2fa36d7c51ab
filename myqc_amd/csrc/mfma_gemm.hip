// mfma_gemm.hip -- split-precision (bf16 hi/lo) GEMM on gfx950 MFMA with FP32
// accumulation, for the ao2mo quarter transforms.
//
//   C = op(A) B,   op(A) = A or A^T,   all column-major FP64 in HBM.
//
// Each FP64 operand element x is split as
//   hi = bf16(x),   lo = bf16(x - hi)            (16 significant bits in all)
// and the product is accumulated as  hi*hi  +  (hi*lo + lo*hi)  with
// v_mfma_f32_16x16x32_bf16; the dropped lo*lo term is 2^-16 relative.  The
// main and correction sums live in separate FP32 accumulators and are added
// in FP64 at the store.  Relative error per output ~1e-5 (bf16 split) plus
// FP32 accumulation over k; the ao2mo/MP2 test states it against FP64.
//
// The quarter transforms multiply one large operand (the integrals or a
// partly transformed tensor) by a small block of C: the small operand is split
// once into hi/lo planes (k contiguous per row or column, read with 16-byte
// loads), the large one is split on its way into LDS, its next k-step loaded
// into registers while the current one's MFMAs run.
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

// hi/lo planes of op(A) = A^T (A column-major, lda): rows i < m of length
// kp (k rounded up to GK), zero beyond k and beyond m (up to mp rows)
__global__ void k_split_rows_t(const double *__restrict__ A, long long lda, int m, int k, int mp, int kp,
                               uint16_t *__restrict__ hi, uint16_t *__restrict__ lo) {
    const long long cnt = (long long)mp * kp;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e / kp), kk = (int)(e - (long long)i * kp);
        const double x = (i < m && kk < k) ? A[kk + lda * i] : 0.0;
        split(x, hi[e], lo[e]);
    }
}

// C = A^T B with A's hi/lo planes pre-split (k_split_rows_t) and B (k x n,
// k-contiguous columns) split on its way into LDS with 16-byte vector loads;
// the next k-step's B values are loaded into registers while this step's
// MFMAs run.  Tiling and accumulation as described at the top of this file.
__global__ __launch_bounds__(256) void k_gemm_bf16x3_tn(int m, int n, int k, int kp, const uint16_t *__restrict__ Ahi,
                                                        const uint16_t *__restrict__ Alo, const double *__restrict__ B,
                                                        long long ldb, double *__restrict__ C, long long ldc,
                                                        int vec_ok) {
    __shared__ __attribute__((aligned(16))) uint16_t sAh[GM][GKP], sAl[GM][GKP], sBh[GN][GKP], sBl[GN][GKP];
    const int m0 = blockIdx.y * GM;
    const long long n0 = (long long)blockIdx.x * GN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 32;
    const int r16 = lane & 15, q = lane >> 4;
    // B loader: column j, 8 consecutive k
    const int bj = t >> 2, bk = (t & 3) * 8;
    const long long gj = n0 + bj;
    double nb[8];
    auto load_b = [&](int k0) {
        const int gk = k0 + bk;
        if (vec_ok && gj < n && gk + 8 <= k) {
            const double2 *p = reinterpret_cast<const double2 *>(B + gk + ldb * gj);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double2 d = p[v];
                nb[2 * v] = d.x;
                nb[2 * v + 1] = d.y;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) nb[e] = (gj < n && gk + e < k) ? B[gk + e + ldb * gj] : 0.0;
        }
    };
    f32x4 acc[4][2], cor[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            cor[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
    load_b(0);
    for (int k0 = 0; k0 < k; k0 += GK) {
        // A planes: 128 rows x 32 k, 16-byte chunks (two per thread per plane)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int id = t + 256 * c, row = id >> 2, ch = (id & 3) * 8;
            const size_t src = (size_t)(m0 + row) * kp + k0 + ch;
            *reinterpret_cast<uint4 *>(&sAh[row][ch]) = *reinterpret_cast<const uint4 *>(Ahi + src);
            *reinterpret_cast<uint4 *>(&sAl[row][ch]) = *reinterpret_cast<const uint4 *>(Alo + src);
        }
        // B: split the prefetched values, then start loading the next step's
        uint32_t h[4], l[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            uint16_t h0, l0, h1, l1;
            split(nb[2 * e], h0, l0);
            split(nb[2 * e + 1], h1, l1);
            h[e] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            l[e] = (uint32_t)l0 | ((uint32_t)l1 << 16);
        }
        *reinterpret_cast<uint4 *>(&sBh[bj][bk]) = make_uint4(h[0], h[1], h[2], h[3]);
        *reinterpret_cast<uint4 *>(&sBl[bj][bk]) = make_uint4(l[0], l[1], l[2], l[3]);
        if (k0 + GK < k) load_b(k0 + GK);
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
                const long long gjj = n0 + wn + 16 * j + r16;
                if (gi < m && gjj < n) C[gi + ldc * gjj] = (double)acc[i][j][v] + (double)cor[i][j][v];
            }
}

// C = A B for `batch` strided problems with A (m x k, column-major, lda,
// FP64) split on its way into LDS and B's hi/lo planes pre-split once
// (k_split_rows_t over B's columns; B is shared by the batch).  A wavefront
// loads 64 consecutive rows of A at one k (coalesced); the next k-step's A
// values are loaded into registers while this step's MFMAs run.  Tiling and
// accumulation as described at the top of this file.
__global__ __launch_bounds__(256) void k_gemm_bf16x3_nn(int m, int n, int k, int kp, const double *__restrict__ A,
                                                        long long lda, long long sA, const uint16_t *__restrict__ Bhi,
                                                        const uint16_t *__restrict__ Blo, double *__restrict__ C,
                                                        long long ldc, long long sC, long long tilesN) {
    __shared__ __attribute__((aligned(16))) uint16_t sAh[GM][GKP], sAl[GM][GKP], sBh[GN][GKP], sBl[GN][GKP];
    const long long bx = blockIdx.x;
    const long long bat = bx / tilesN, tn = bx - bat * tilesN;
    const int m0 = blockIdx.y * GM;
    const long long n0 = tn * GN;
    A += bat * sA;
    C += bat * sC;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = (w & 1) * 64, wn = (w >> 1) * 32;
    const int r16 = lane & 15, q = lane >> 4;
    // A loader: row ar of the tile, k = ah + 2 e (e < 16)
    const int ar = t & 127, ah = t >> 7;
    const int gi = m0 + ar;
    double na[16];
    auto load_a = [&](int k0) {
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int gk = k0 + ah + 2 * e;
            na[e] = (gi < m && gk < k) ? A[gi + lda * gk] : 0.0;
        }
    };
    // B loader: column bj of the tile, 8 consecutive k
    const int bj = t >> 2, bk = (t & 3) * 8;
    f32x4 acc[4][2], cor[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            acc[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            cor[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
    load_a(0);
    for (int k0 = 0; k0 < k; k0 += GK) {
        {
            const size_t src = (size_t)(n0 + bj) * kp + k0 + bk;
            *reinterpret_cast<uint4 *>(&sBh[bj][bk]) = *reinterpret_cast<const uint4 *>(Bhi + src);
            *reinterpret_cast<uint4 *>(&sBl[bj][bk]) = *reinterpret_cast<const uint4 *>(Blo + src);
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) split(na[e], sAh[ar][ah + 2 * e], sAl[ar][ah + 2 * e]);
        if (k0 + GK < k) load_a(k0 + GK);
        __syncthreads();
        bf16x8 bh[2], bl[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            bh[j] = *reinterpret_cast<const bf16x8 *>(&sBh[wn + 16 * j + r16][8 * q]);
            bl[j] = *reinterpret_cast<const bf16x8 *>(&sBl[wn + 16 * j + r16][8 * q]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const bf16x8 ahv = *reinterpret_cast<const bf16x8 *>(&sAh[wm + 16 * i + r16][8 * q]);
            const bf16x8 alv = *reinterpret_cast<const bf16x8 *>(&sAl[wm + 16 * i + r16][8 * q]);
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, bh[j], acc[i][j], 0, 0, 0);
                cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahv, bl[j], cor[i][j], 0, 0, 0);
                cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alv, bh[j], cor[i][j], 0, 0, 0);
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
                const int gii = m0 + wm + 16 * i + 4 * q + v;
                const long long gjj = n0 + wn + 16 * j + r16;
                if (gii < m && gjj < n) C[gii + ldc * gjj] = (double)acc[i][j][v] + (double)cor[i][j][v];
            }
}

// Quarter steps 1 and 2 of the transform fused, for one (l,d) block per
// workgroup (blockIdx.x = l + n d) and a 128-row tile of i (blockIdx.y) and a
// 64-column tile of a (blockIdx.z), with i over the first orbital block Cp and
// a over the second, Cq (o and nv below are their widths):
//   L(i,v) = sum_u Cp(u,i) XX(u,v,l,d),   M(i,a,l,d) = sum_v L(i,v) Cq(v,a)
// L is formed 32 columns of v at a time in LDS and never leaves the
// workgroup.  Each step keeps the split arithmetic of k_gemm_bf16x3_tn and
// k_gemm_bf16x3_nn exactly: L is rounded as the FP64 value acc + cor those
// kernels would store, then split again, and both k loops run in the same
// 32-wide order, so M is bitwise the unfused result.  Co: planes of Cp^T
// (k_split_rows_t, rows i, kp), Cv: planes of Cq^T (rows a, kp).
__global__ __launch_bounds__(256) void k_ao2mo12_bf16x3(int o, int nv, int n, int kp, const uint16_t *__restrict__ Coh,
                                                        const uint16_t *__restrict__ Col,
                                                        const uint16_t *__restrict__ Cvh,
                                                        const uint16_t *__restrict__ Cvl,
                                                        const double *__restrict__ XX, double *__restrict__ M) {
    __shared__ __attribute__((aligned(16))) uint16_t sAh[GM][GKP], sAl[GM][GKP];   // C_occ^T tile (i, u)
    __shared__ __attribute__((aligned(16))) uint16_t sBh[32][GKP], sBl[32][GKP];   // XX tile (v, u)
    __shared__ __attribute__((aligned(16))) uint16_t sLh[GM][GKP], sLl[GM][GKP];   // L slice (i, v)
    __shared__ __attribute__((aligned(16))) uint16_t sVh[64][GKP], sVl[64][GKP];   // C_virt^T tile (a, v)
    const long long ld = blockIdx.x;
    const int i0 = blockIdx.y * GM, a0 = blockIdx.z * 64;
    const double *X = XX + ld * (long long)n * n;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int r16 = lane & 15, q = lane >> 4;
    const int wr = 32 * w;   // the wavefront's 32 rows of i
    // XX loader: column v = t >> 3 of the slice, u = 4 (t & 7) .. +3 of the k-step
    const int bv = t >> 3, bu = (t & 7) * 4;
    f32x4 accM[2][4], corM[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            accM[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            corM[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
        }
    for (int v0 = 0; v0 < n; v0 += 32) {
        // C_virt^T planes for this slice of v (rows a0 .. a0 + 63)
        {
            const int row = t >> 2, ch = (t & 3) * 8;
            const size_t src = (size_t)(a0 + row) * kp + v0 + ch;
            *reinterpret_cast<uint4 *>(&sVh[row][ch]) = *reinterpret_cast<const uint4 *>(Cvh + src);
            *reinterpret_cast<uint4 *>(&sVl[row][ch]) = *reinterpret_cast<const uint4 *>(Cvl + src);
        }
        f32x4 accL[2][2], corL[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                accL[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
                corL[i][j] = (f32x4){0.f, 0.f, 0.f, 0.f};
            }
        for (int k0 = 0; k0 < n; k0 += GK) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const int id = t + 256 * c, row = id >> 2, ch = (id & 3) * 8;
                const size_t src = (size_t)(i0 + row) * kp + k0 + ch;
                *reinterpret_cast<uint4 *>(&sAh[row][ch]) = *reinterpret_cast<const uint4 *>(Coh + src);
                *reinterpret_cast<uint4 *>(&sAl[row][ch]) = *reinterpret_cast<const uint4 *>(Col + src);
            }
            {
                const int v = v0 + bv;
                uint16_t h[4], l[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int u = k0 + bu + e;
                    const double x = (v < n && u < n) ? X[u + (long long)n * v] : 0.0;
                    split(x, h[e], l[e]);
                }
                *reinterpret_cast<uint2 *>(&sBh[bv][bu]) =
                    make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
                *reinterpret_cast<uint2 *>(&sBl[bv][bu]) =
                    make_uint2((uint32_t)l[0] | ((uint32_t)l[1] << 16), (uint32_t)l[2] | ((uint32_t)l[3] << 16));
            }
            __syncthreads();
            bf16x8 bh[2], bl[2];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                bh[j] = *reinterpret_cast<const bf16x8 *>(&sBh[16 * j + r16][8 * q]);
                bl[j] = *reinterpret_cast<const bf16x8 *>(&sBl[16 * j + r16][8 * q]);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(&sAh[wr + 16 * i + r16][8 * q]);
                const bf16x8 al = *reinterpret_cast<const bf16x8 *>(&sAl[wr + 16 * i + r16][8 * q]);
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    accL[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], accL[i][j], 0, 0, 0);
                    corL[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], corL[i][j], 0, 0, 0);
                    corL[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], corL[i][j], 0, 0, 0);
                }
            }
            __syncthreads();
        }
        // the L slice, rounded as the stored FP64 value would be, split into LDS
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int vv = 0; vv < 4; ++vv) {
                    const int row = wr + 16 * i + 4 * q + vv, col = 16 * j + r16;
                    const double lv = (double)accL[i][j][vv] + (double)corL[i][j][vv];
                    split(lv, sLh[row][col], sLl[row][col]);
                }
        __syncthreads();
        bf16x8 vh[4], vl[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            vh[j] = *reinterpret_cast<const bf16x8 *>(&sVh[16 * j + r16][8 * q]);
            vl[j] = *reinterpret_cast<const bf16x8 *>(&sVl[16 * j + r16][8 * q]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const bf16x8 lh = *reinterpret_cast<const bf16x8 *>(&sLh[wr + 16 * i + r16][8 * q]);
            const bf16x8 ll = *reinterpret_cast<const bf16x8 *>(&sLl[wr + 16 * i + r16][8 * q]);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                accM[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lh, vh[j], accM[i][j], 0, 0, 0);
                corM[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lh, vl[j], corM[i][j], 0, 0, 0);
                corM[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ll, vh[j], corM[i][j], 0, 0, 0);
            }
        }
        __syncthreads();   // sL and sV are refilled by the next slice
    }
    // M(i, a, l, d) at i + o (a + nv (l + n d))
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int vv = 0; vv < 4; ++vv) {
                const int gi = i0 + wr + 16 * i + 4 * q + vv, ga = a0 + 16 * j + r16;
                if (gi < o && ga < nv)
                    M[gi + (long long)o * (ga + (long long)nv * ld)] = (double)accM[i][j][vv] + (double)corM[i][j][vv];
            }
}

// Quarter steps 1 and 2 of the transform in split-bf16 over two column
// blocks of orbitals: M (np x nq x n x n), M(p,q,l,d) = sum_uv Cp(u,p) Cq(v,q)
// XX(u,v,l,d), from Cp (n x np) and Cq (n x nq), both with leading dimension n,
// and the dense integrals XX (n^4).
int ao2mo12_bf16x3(myqc_ctx *ctx, int np, const double *Cp, int nq, const double *Cq, int n, const double *XX,
                   double *M) {
    if (np <= 0 || nq <= 0) return set_error(ctx, MYQC_E_ARG, "ao2mo12_bf16x3: np=%d nq=%d", np, nq);
    const int gy = (np + GM - 1) / GM, gz = (nq + 63) / 64;
    const long long nld = (long long)n * n;
    if (nld > 0x7fffffffLL || gy > 65535 || gz > 65535) return set_error(ctx, MYQC_E_ARG, "ao2mo12_bf16x3: too large");
    const int kp = ((n + GK - 1) / GK) * GK;
    const size_t po = (size_t)gy * GM * kp, pv = (size_t)gz * 64 * kp;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_planes, sizeof(uint16_t) * 2 * (po + pv)));
    uint16_t *coh = ctx->d_planes, *col = coh + po, *cvh = col + po, *cvl = cvh + pv;
    k_split_rows_t<<<256, 256, 0, ctx->stream>>>(Cp, n, np, n, gy * GM, kp, coh, col);
    k_split_rows_t<<<256, 256, 0, ctx->stream>>>(Cq, n, nq, n, gz * 64, kp, cvh, cvl);
    k_ao2mo12_bf16x3<<<dim3((unsigned)nld, (unsigned)gy, (unsigned)gz), 256, 0, ctx->stream>>>(np, nq, n, kp, coh, col,
                                                                                             cvh, cvl, XX, M);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// C (m x n, ldc) = op(A) (m x k) * B (k x n, ldb) for `batch` strided
// problems; FP64 in and out, split-bf16 MFMA products inside.  The transposed
// single problem (the first quarter transform: a small A against a large,
// k-contiguous B) pre-splits A once.
int gemm_bf16x3(myqc_ctx *ctx, bool transA, int m, int n, int k, const double *A, int lda, long long strideA,
                const double *B, int ldb, long long strideB, double *C, int ldc, long long strideC, int batch) {
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return set_error(ctx, MYQC_E_ARG, "gemm_bf16x3: empty problem");
    const long long tilesN = ((long long)n + GN - 1) / GN;
    const long long gx = tilesN * batch;
    const int gy = (m + GM - 1) / GM;
    if (gx > 0x7fffffffLL || gy > 65535) return set_error(ctx, MYQC_E_ARG, "gemm_bf16x3: grid too large");
    if (transA && batch != 1) return set_error(ctx, MYQC_E_ARG, "gemm_bf16x3: op(A) = A^T takes one problem");
    if (transA) {
        const int mp = gy * GM, kp = ((k + GK - 1) / GK) * GK;
        MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_planes, sizeof(uint16_t) * 2 * (size_t)mp * kp));
        uint16_t *hi = ctx->d_planes, *lo = ctx->d_planes + (size_t)mp * kp;
        k_split_rows_t<<<256, 256, 0, ctx->stream>>>(A, lda, m, k, mp, kp, hi, lo);
        const int vec_ok = (ldb % 2 == 0) && ((reinterpret_cast<uintptr_t>(B) & 15) == 0);
        k_gemm_bf16x3_tn<<<dim3((unsigned)tilesN, (unsigned)gy), 256, 0, ctx->stream>>>(m, n, k, kp, hi, lo, B, ldb, C,
                                                                                      ldc, vec_ok);
        MYQC_HIP(ctx, hipGetLastError());
        return 0;
    }
    // op(A) = A: B (shared by the batch) is split into planes once
    if (strideB != 0) return set_error(ctx, MYQC_E_ARG, "gemm_bf16x3: B must be shared by the batch");
    const int np_ = (int)(tilesN * GN), kp = ((k + GK - 1) / GK) * GK;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_planes, sizeof(uint16_t) * 2 * (size_t)np_ * kp));
    uint16_t *hi = ctx->d_planes, *lo = ctx->d_planes + (size_t)np_ * kp;
    k_split_rows_t<<<256, 256, 0, ctx->stream>>>(B, ldb, n, k, np_, kp, hi, lo);
    k_gemm_bf16x3_nn<<<dim3((unsigned)gx, (unsigned)gy), 256, 0, ctx->stream>>>(m, n, k, kp, A, lda, strideA, hi, lo, C,
                                                                                ldc, strideC, tilesN);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

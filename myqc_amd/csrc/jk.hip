// jk.hip -- J/K (Fock G) digestion of the stored integrals.
//
// The reference's two-electron part of the Fock matrix is
//     G(i,j) = sum_kl D(k,l)[XX(i,j,k,l) - XX(i,k,j,l)/4 - XX(i,l,j,k)/4]
// (src/I2G/RHFI2G.f90:80-90), which only sees D through D + D^T, so it is
// evaluated for the symmetrised D.  Each unique shell quartet (AB|CD), with
// (A,B) its bra pair (the higher class, or the lower Schwarz rank in one
// class), is stored three times, in segments of kets sorted by Schwarz rank:
// under the ordered pairs (A,B) and (B,A), and under (C,D) in the ket pair's
// stored orientation.  A copy under the ordered pair (X,Y) with ket (U,W)
// updates only rows of shell X:
//     J: A(i,j) += wJ f D(k,l) (ij|kl),           f = 2 unless U == W
//     K: A(i,k) -= wK D(j,l) (ij|kl) / 2,  and if U != W  A(i,l) -= wK D(j,k) (ij|kl) / 2
// for i in X, j in Y, and G = A + A^T.  The weights split the quartet's
// terms so that each lands once: the bra's forward copy adds J and its K rows,
// the bra's reverse copy its K rows only, the ket's copy J only (wK = 0);
// wJ halves for a same-shell row pair, wK for the diagonal quartet (P|P).
// (tests/hostcheck states this split on the host and checks it against
// RHFI2G.)  A workgroup owns the rows of one shell A for a run of about 2^15
// stored values: it stages in LDS the rows of D of the B shells it meets (at
// most dig_rows AOs), then each wavefront takes chunks of kets of one (A, B,
// ket class, role) range in turn, with no barrier between chunks.  A lane
// takes 1-8 kets of a chunk (more for the small classes) and issues all their
// loads of one bra row before their arithmetic.  J is reduced across the
// wavefront and K is added per ket, both into FP64 LDS rows of A (FP32
// values are accumulated in FP64 too); the rows go to the workgroup's own
// partial slot.  Workgroups launch longest first.  A gather sums the partial
// rows of each shell in a fixed order: no global atomics; the LDS atomics
// make the summation order, and so the last bits, vary from run to run.
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "context.h"

namespace myqc {

template <int C> struct PL;
template <> struct PL<0> { static constexpr int a = 0, b = 0; };
template <> struct PL<1> { static constexpr int a = 1, b = 0; };
template <> struct PL<2> { static constexpr int a = 1, b = 1; };

template <typename TV>
struct DigArgs {
    const int *meta0, *meta1, *meta2;         // ket AOs by class and Schwarz rank
    const TV *dk0, *dk1, *dk2;                // ket density blocks by class and rank (1, 3, 9 values)
    const int *wg;                            // [w] = {A, chunk0, chunk1, first D row, D rows, partial slot}
    const int *unit;                          // [u] = {B, c0, c1, c2}
    const int4 *chunk;                        // [c] = {offset lo, offset hi, aoB | LB << 15 | class << 16 | kets << 18, first rank}
    const long long *unit_off;                // [u][3]
    const int *shell;                         // [s] = {ao, L, wg0, wg1}
    const TV *x;                              // expanded values
    const TV *D;                              // symmetric density (or change), n x n
    double *partial;                          // [w][3][n]
    int n;
    double jscale;                            // 1, or 0 for the exchange part alone
    // density-weighted screen of the incremental (FP32) passes: a partner of
    // rank r in class c is skipped when q_row * q_c[r] * dmax < qtau (0: no
    // screen); q_c are the partners' Schwarz factors by rank (descending)
    const float *q0, *q1, *q2;
    const float4 *chunkq;                     // [c] = {q_row q_first, q_row q_last, q_row, 0}, rounded up
    const unsigned *dmax;                     // max |D| of this pass (device, float bits)
    double qtau;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Kets per lane of a chunk: enough that every lane has several loads of one
// bra row in flight (nb*nc*nd values per ket and row).
__host__ __device__ constexpr int kets_per_lane(int row) { return row == 1 ? 8 : row == 3 ? 4 : row == 9 ? 2 : 1; }

// One chunk of kets of class CK against the bra (A, B): lane l takes kets
// r0 + l + 64 j, j < KPL.  J is reduced across the wavefront into the LDS
// rows of A, K added per ket into them.  s_drow holds D's rows of shell B.
// The bra rows are streamed one at a time, with the loads of all the lane's
// kets issued before their arithmetic.
template <int LA, int LB, int CK, typename TV>
__device__ __forceinline__ void digest_chunk(const DigArgs<TV> &a, double *s_rows, const TV *s_drow, int r0, int nk,
                                             long long off, int aoB, double wJ, double wK, double thr) {
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(PL<CK>::a), nd = ncart(PL<CK>::b);
    constexpr int ROW = nb * nc * nd, BLK = na * ROW, KPL = kets_per_lane(ROW);
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int *meta = CK == 0 ? a.meta0 : CK == 1 ? a.meta1 : a.meta2;
    const TV *dkc = CK == 0 ? a.dk0 : CK == 1 ? a.dk1 : a.dk2;
    const float *qk = CK == 0 ? a.q0 : CK == 1 ? a.q1 : a.q2;
    bool act[KPL];
    int aoC[KPL], aoD[KPL], same[KPL];
    TV dq[KPL][nc * nd], dbd[KPL][nb * nd], dbc[KPL][nb * nc];
    const TV *v[KPL];
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        const int k = lane + 64 * j;
        // thr > 0: the screen of the incremental passes (partners whose
        // integrals with this row pair are all below it drop out)
        act[j] = k < nk && (thr <= 0.0 || (double)qk[r0 + k] >= thr);
        const int r = r0 + (act[j] ? k : 0);
        const int m = meta[r];
        aoC[j] = m & 0x7fff;
        aoD[j] = (m >> 15) & 0x7fff;
        same[j] = m >> 30;
        v[j] = a.x + off + (long long)r * BLK;
#pragma unroll
        for (int e = 0; e < nc * nd; ++e) dq[j][e] = act[j] ? dkc[(size_t)r * (nc * nd) + e] : (TV)0;
    }
#pragma unroll
    for (int j = 0; j < KPL; ++j)
#pragma unroll
        for (int b = 0; b < nb; ++b) {
#pragma unroll
            for (int d = 0; d < nd; ++d) dbd[j][b * nd + d] = s_drow[b * n + aoD[j] + d];
#pragma unroll
            for (int c = 0; c < nc; ++c) dbc[j][b * nc + c] = s_drow[b * n + aoC[j] + c];
        }
#pragma unroll 1
    for (int ia = 0; ia < na; ++ia) {
        TV x[KPL][ROW];
#pragma unroll
        for (int j = 0; j < KPL; ++j)
#pragma unroll
            for (int e = 0; e < ROW; ++e) x[j][e] = act[j] ? v[j][ia * ROW + e] : (TV)0;
        TV jr[nb];
#pragma unroll
        for (int ib = 0; ib < nb; ++ib) jr[ib] = 0;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const TV fj = same[j] ? (TV)1 : (TV)2;
#pragma unroll
            for (int ib = 0; ib < nb; ++ib) {
                TV js = 0;
#pragma unroll
                for (int c = 0; c < nc; ++c)
#pragma unroll
                    for (int d = 0; d < nd; ++d) js += dq[j][c * nd + d] * x[j][(ib * nc + c) * nd + d];
                jr[ib] += fj * js;
            }
            if (!act[j] || wK == 0.0) continue;   // (wK is uniform across the wavefront)
            // G(i, k) -= 1/2 sum_{b,d} D(b,d) (i b|k d)
#pragma unroll
            for (int c = 0; c < nc; ++c) {
                TV ks = 0;
#pragma unroll
                for (int ib = 0; ib < nb; ++ib)
#pragma unroll
                    for (int d = 0; d < nd; ++d) ks += dbd[j][ib * nd + d] * x[j][(ib * nc + c) * nd + d];
                atomicAdd(&s_rows[ia * n + aoC[j] + c], -0.5 * wK * (double)ks);
            }
            // the ket's other orientation: G(i, l) -= 1/2 sum_{b,c} D(b,c) (i b|k l)
            if (!same[j]) {
#pragma unroll
                for (int d = 0; d < nd; ++d) {
                    TV ks = 0;
#pragma unroll
                    for (int ib = 0; ib < nb; ++ib)
#pragma unroll
                        for (int c = 0; c < nc; ++c) ks += dbc[j][ib * nc + c] * x[j][(ib * nc + c) * nd + d];
                    atomicAdd(&s_rows[ia * n + aoD[j] + d], -0.5 * wK * (double)ks);
                }
            }
        }
#pragma unroll
        for (int ib = 0; ib < nb; ++ib) {
            const TV t = wave_sum(jr[ib]);
            if (lane == 0 && wJ != 0.0) atomicAdd(&s_rows[ia * n + aoB + ib], wJ * (double)t);
        }
    }
}

template <int LA, typename TV>
__device__ __forceinline__ void digest_chunk_la(const DigArgs<TV> &a, double *s_rows, const TV *s_drow, int LB, int ck,
                                                int r0, int nk, long long off, int aoB, double wJ, double wK,
                                                double thr) {
    if (LB == 0) {
        if (ck == 0) digest_chunk<LA, 0, 0>(a, s_rows, s_drow, r0, nk, off, aoB, wJ, wK, thr);
        else if (ck == 1) digest_chunk<LA, 0, 1>(a, s_rows, s_drow, r0, nk, off, aoB, wJ, wK, thr);
        else digest_chunk<LA, 0, 2>(a, s_rows, s_drow, r0, nk, off, aoB, wJ, wK, thr);
    } else {
        if (ck == 0) digest_chunk<LA, 1, 0>(a, s_rows, s_drow, r0, nk, off, aoB, wJ, wK, thr);
        else if (ck == 1) digest_chunk<LA, 1, 1>(a, s_rows, s_drow, r0, nk, off, aoB, wJ, wK, thr);
        else digest_chunk<LA, 1, 2>(a, s_rows, s_drow, r0, nk, off, aoB, wJ, wK, thr);
    }
}

// One workgroup per run of units of one row shell A: it stages the D rows of
// the units' B shells once, then each wavefront takes chunks in turn (no
// barrier between them; the next chunk's descriptor is loaded ahead), and
// the rows of A it accumulated go to its partial slot.
template <typename TV>
__global__ __launch_bounds__(256, 3) void k_digest(DigArgs<TV> a) {
    extern __shared__ unsigned char smem[];
    const int n = a.n;
    double *s_rows = reinterpret_cast<double *>(smem);
    TV *s_drow = reinterpret_cast<TV *>(s_rows + 3 * n);
    const int w = blockIdx.x;
    const int A = a.wg[6 * w], ch0 = a.wg[6 * w + 1], ch1 = a.wg[6 * w + 2];
    const int row0 = a.wg[6 * w + 3], nrows = a.wg[6 * w + 4], slot = a.wg[6 * w + 5];
    const int LA = a.shell[4 * A + 1], aoA = a.shell[4 * A];
    const int na = LA ? 3 : 1;
    for (int i = threadIdx.x; i < na * n; i += blockDim.x) s_rows[i] = 0;
    // D(row0 + b, j) = D(j, row0 + b): read columns, coalesced
    for (int i = threadIdx.x; i < nrows * n; i += blockDim.x) {
        const int b = i / n, j = i - b * n;
        s_drow[i] = a.D[j + (size_t)n * (row0 + b)];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const bool screen = a.qtau > 0.0;
    const double dmax = screen ? (double)__uint_as_float(*a.dmax) : 0.0;
    int c = ch0 + wave;
    int4 d = c < ch1 ? a.chunk[c] : make_int4(0, 0, 0, 0);
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 q = screen && c < ch1 ? a.chunkq[c] : z4;
    while (c < ch1) {
        const int cn = c + 4;
        const int4 dn = cn < ch1 ? a.chunk[cn] : make_int4(0, 0, 0, 0);
        const float4 qn = screen && cn < ch1 ? a.chunkq[cn] : z4;
        const long long off = (long long)(unsigned)d.x | ((long long)d.y << 32);
        const int aoB = d.z & 0x7fff, LB = (d.z >> 15) & 1, ck = (d.z >> 16) & 3, nk = (d.z >> 18) & 0x3ff;
        double thr = 0.0;
        if (screen) {
            // |integral| <= q_row q_partner, partners in descending order: a
            // chunk whose first partner is below the screen drops out whole,
            // one whose last is above it runs whole, and only a chunk that
            // straddles it tests its partners one by one (thr > 0)
            if ((double)q.x * dmax < a.qtau) {
                c = cn;
                d = dn;
                q = qn;
                continue;
            }
            if ((double)q.y * dmax < a.qtau) thr = a.qtau / ((double)q.z * dmax);
        }
        const int role = (d.z >> 28) & 7;   // 1: bra role, 2: forward orientation, 4: diagonal quartet
        // weights of the three-copy split (G = A + A^T): the bra pair's two
        // orientations share its J and each adds its K rows, the ket pair's
        // copy adds J only; halves for a same-shell row pair and the diagonal
        const double half_same = (aoB == aoA) ? 0.5 : 1.0;
        const double wJ = a.jscale * ((role & 1) ? ((role & 2) ? half_same : 0.0) : half_same);
        const double wK = (role & 1) ? ((role & 4) ? 0.5 : 1.0) : 0.0;
        const TV *drow = s_drow + (size_t)(aoB - row0) * n;
        if (LA == 0) digest_chunk_la<0>(a, s_rows, drow, LB, ck, d.w, nk, off, aoB, wJ, wK, thr);
        else digest_chunk_la<1>(a, s_rows, drow, LB, ck, d.w, nk, off, aoB, wJ, wK, thr);
        c = cn;
        d = dn;
        q = qn;
    }
    __syncthreads();
    double *dst = a.partial + (size_t)slot * 3 * n;
    for (int i = threadIdx.x; i < na * n; i += blockDim.x) dst[i] = s_rows[i];
}

// G(i,j) = sum over the workgroups of shell(i) of their partial rows
__global__ void k_gather_rows(const double *__restrict__ partial, const int *__restrict__ shell,
                              const int *__restrict__ ao_shell, int n, double *__restrict__ Graw) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= n) return;
    const int A = ao_shell[i];
    const int ia = i - shell[4 * A], w0 = shell[4 * A + 2], w1 = shell[4 * A + 3];
    // four independent sums so that the loads of consecutive slots overlap
    double s4[4] = {0.0, 0.0, 0.0, 0.0};
    int w = w0;
    for (; w + 4 <= w1; w += 4)
#pragma unroll
        for (int q = 0; q < 4; ++q) s4[q] += partial[((size_t)(w + q) * 3 + ia) * n + j];
    for (; w < w1; ++w) s4[0] += partial[((size_t)w * 3 + ia) * n + j];
    const double s = (s4[0] + s4[1]) + (s4[2] + s4[3]);
    Graw[i + (size_t)n * j] = s;
}

// the kets' density blocks in Schwarz-rank order: dk[r*(nc*nd) + c*nd + d] =
// D(ao_a + c, ao_b + d) of pair sorted[r]
template <typename TV>
__global__ void k_pack_kets(const ShellPair *__restrict__ pairs, const int *__restrict__ sorted, int np, int nc, int nd,
                            const TV *__restrict__ D, int n, TV *__restrict__ dk) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= np) return;
    const ShellPair sp = pairs[sorted[r]];
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < nd; ++d) dk[(size_t)r * nc * nd + c * nd + d] = D[(sp.ao_a + c) + (size_t)n * (sp.ao_b + d)];
}

// --- small elementwise kernels ------------------------------------------------

__global__ void k_add(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        Z[i] = X[i] + Y[i];
}

// Ds = scale (D + D^T); with amax, also max |Ds32| into *amax as the bit
// pattern of a non-negative float (ordered like the value; *amax starts at 0)
__global__ __launch_bounds__(128) void k_symmetrize(const double *__restrict__ D, double *__restrict__ Ds, float *__restrict__ Ds32, int n,
                             double scale, unsigned *__restrict__ amax = nullptr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    float m = 0.0f;
    if (i < n) {
        const double v = scale * (D[i + (size_t)n * j] + D[j + (size_t)n * i]);
        if (Ds) Ds[i + (size_t)n * j] = v;
        if (Ds32) {
            Ds32[i + (size_t)n * j] = (float)v;
            m = fabsf((float)v);
        }
    }
    if (amax) {
        // one atomic per workgroup at most, and none once the stored maximum
        // is already as large (the stored value only grows, so a stale read
        // can only cost an unneeded atomic)
        __shared__ float s_m[2];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned b = __float_as_uint(fmaxf(s_m[0], s_m[1]));
            if (b > __atomic_load_n(amax, __ATOMIC_RELAXED)) atomicMax(amax, b);
        }
    }
}





static dim3 grid2(int n) { return dim3((n + 127) / 128, n); }

int ensure_work(myqc_ctx *ctx) {
    if (ctx->work_n == ctx->n) return 0;
    const size_t nn = (size_t)ctx->n * ctx->n;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_D, sizeof(double) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_A, sizeof(double) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_G, sizeof(double) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_D32, sizeof(float) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_A32, sizeof(float) * nn));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_dmax, sizeof(double)));
    ctx->work_n = ctx->n;
    return 0;
}

template <typename TV>
static int launch_digest(myqc_ctx *ctx, const TV *D, const TV *x, TV *dk, double *partial, double *Graw,
                         double jscale = 1.0, double qtau = 0.0) {
    const int np0 = (int)ctx->pair_ids[0].size(), np1 = (int)ctx->pair_ids[1].size(),
              np2 = (int)ctx->pair_ids[2].size();
    TV *dk0 = dk, *dk1 = dk + np0, *dk2 = dk + np0 + 3 * (size_t)np1;
    const int nps[3] = {np0, np1, np2};
    TV *dks[3] = {dk0, dk1, dk2};
    for (int c = 0; c < 3; ++c)
        if (nps[c] > 0)
            k_pack_kets<TV><<<(nps[c] + 255) / 256, 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_sorted[c], nps[c],
                                                                            c ? 3 : 1, c == 2 ? 3 : 1, D, ctx->n,
                                                                            dks[c]);
    if (ctx->n_wg > 0) {
        DigArgs<TV> a;
        a.meta0 = ctx->d_meta[0];
        a.meta1 = ctx->d_meta[1];
        a.meta2 = ctx->d_meta[2];
        a.dk0 = dk0;
        a.dk1 = dk1;
        a.dk2 = dk2;
        a.wg = ctx->d_wg;
        a.unit = ctx->d_unit;
        a.chunk = reinterpret_cast<const int4 *>(ctx->d_chunk);
        a.unit_off = ctx->d_unit_off;
        a.shell = ctx->d_shell;
        a.x = x;
        a.D = D;
        a.partial = partial;
        a.n = ctx->n;
        a.jscale = jscale;
        a.q0 = ctx->d_qrank[0];
        a.q1 = ctx->d_qrank[1];
        a.q2 = ctx->d_qrank[2];
        a.chunkq = reinterpret_cast<const float4 *>(ctx->d_chunkq);
        a.dmax = reinterpret_cast<const unsigned *>(ctx->d_dmax);
        a.qtau = qtau;
        const size_t shm = (3 * sizeof(double) + ctx->dig_rows * sizeof(TV)) * (size_t)ctx->n;
        // timed alone: algorithmic bytes per pass are one read of every
        // stored unique integral plus the density in and G out (16 n^2)
        hipEvent_t ev;
        MYQC_CHECK(timer_begin(ctx, T_JK, &ev));
        k_digest<TV><<<ctx->n_wg, 256, shm, ctx->stream>>>(a);
        MYQC_CHECK(timer_end(ctx, T_JK, ev, (double)sizeof(TV) * (double)ctx->n_stored + 16.0 * ctx->n * ctx->n));
    }
    k_gather_rows<<<dim3((ctx->n + 127) / 128, ctx->n), 128, 0, ctx->stream>>>(partial, ctx->d_shell,
                                                                                  ctx->d_ao_shell, ctx->n, Graw);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// The incremental (FP32) passes skip a partner when every integral it forms
// with the row pair, times max |dD|, is below this (|integral| <= Q_row
// Q_partner); the full FP64 builds, which decide convergence and the final
// energy, never screen.
constexpr double kIncrementalScreen = 1e-12;

// G(D) into d_Gout (precision FP64), or d_Gout += G32(D) (precision FP32:
// D is then the density change; partners below the incremental screen drop
// out)
int digest_G(myqc_ctx *ctx, const double *d_Din, double *d_Gout, int precision) {
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "no stored integrals (call myqc_eri_build)");
    MYQC_CHECK(ensure_work(ctx));
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    if ((size_t)(3 + ctx->dig_rows) * n * sizeof(double) > 64 * 1024)
        return set_error(ctx, MYQC_E_ARG, "digestion rows for n=%d exceed 64 KB of LDS", n);
    if (precision == MYQC_FP64) {
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, ctx->d_D, nullptr, n, 0.5);
        MYQC_CHECK(launch_digest<double>(ctx, ctx->d_D, ctx->d_x64, ctx->d_Dpair, ctx->d_partial, ctx->d_A));
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, d_Gout, nullptr, n, 1.0);
    } else {
        if (!ctx->have_eri32) return set_error(ctx, MYQC_E_STATE, "no FP32 integrals stored");
        // max |dD| of the pass for the screen, folded into the symmetrisation
        MYQC_HIP(ctx, hipMemsetAsync(ctx->d_dmax, 0, sizeof(unsigned), ctx->stream));
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, nullptr, ctx->d_D32, n, 0.5,
                                                        reinterpret_cast<unsigned *>(ctx->d_dmax));
        MYQC_CHECK(launch_digest<float>(ctx, ctx->d_D32, ctx->d_x32, ctx->d_Dpair32, ctx->d_partial, ctx->d_A, 1.0,
                                        kIncrementalScreen));
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, ctx->d_G, nullptr, n, 1.0);
        k_add<<<256, 256, 0, ctx->stream>>>(d_Gout, ctx->d_G, d_Gout, nn);
    }
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// Z = a X + b Y
__global__ void k_axpby(const double *__restrict__ X, const double *__restrict__ Y, double *__restrict__ Z, double a,
                        double b, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        Z[i] = a * X[i] + b * Y[i];
}

// UHF: with G(D) = J[D] - K[D]/2 and the exchange part alone Gk(D) = -K[D]/2,
// G_a = J[Da + Db] - K[Da] = G(Da + Db) + Gk(Da - Db), G_b = G(Da + Db) - Gk(Da - Db)
int digest_G_uhf(myqc_ctx *ctx, const double *d_Da, const double *d_Db, double *d_Ga, double *d_Gb) {
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "no stored integrals (call myqc_eri_build)");
    MYQC_CHECK(ensure_work(ctx));
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    // total density into Ga, spin difference into Gb (both overwritten below)
    k_axpby<<<256, 256, 0, ctx->stream>>>(d_Da, d_Db, d_Ga, 1.0, 1.0, nn);
    k_axpby<<<256, 256, 0, ctx->stream>>>(d_Da, d_Db, d_Gb, 1.0, -1.0, nn);
    k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Gb, ctx->d_D, nullptr, n, 0.5);
    MYQC_CHECK(launch_digest<double>(ctx, ctx->d_D, ctx->d_x64, ctx->d_Dpair, ctx->d_partial, ctx->d_A, 0.0));
    k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, ctx->d_G, nullptr, n, 1.0);   // Gk(Da - Db)
    MYQC_CHECK(digest_G(ctx, d_Ga, d_Gb, MYQC_FP64));                                           // G(Da + Db)
    k_axpby<<<256, 256, 0, ctx->stream>>>(d_Gb, ctx->d_G, d_Ga, 1.0, 1.0, nn);
    k_axpby<<<256, 256, 0, ctx->stream>>>(d_Gb, ctx->d_G, d_Gb, 1.0, -1.0, nn);
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_fock_G_uhf(myqc_ctx *ctx, const double *Da, const double *Db, double *Ga, double *Gb) {
    if (!ctx || !Da || !Db || !Ga || !Gb) return MYQC_E_ARG;
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "myqc_fock_G_uhf: call myqc_eri_build first");
    const size_t nn = (size_t)ctx->n * ctx->n;
    double *buf = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&buf, sizeof(double) * 4 * nn));
    double *dA = buf, *dB = buf + nn, *gA = buf + 2 * nn, *gB = buf + 3 * nn;
    int rc = 0;
    if (hipMemcpyAsync(dA, Da, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(dB, Db, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_fock_G_uhf: upload failed");
    if (rc == 0) rc = digest_G_uhf(ctx, dA, dB, gA, gB);
    if (rc == 0 && (d2h(ctx, Ga, gA, sizeof(double) * nn) != hipSuccess ||
                    d2h(ctx, Gb, gB, sizeof(double) * nn) != hipSuccess))
        rc = set_error(ctx, MYQC_E_HIP, "myqc_fock_G_uhf: download failed");
    dfree((void **)&buf);
    return rc;
}

extern "C" int myqc_fock_G(myqc_ctx *ctx, const double *D, double *G, int precision) {
    if (!ctx || !D || !G) return MYQC_E_ARG;
    if (precision != MYQC_FP64 && precision != MYQC_FP32)
        return set_error(ctx, MYQC_E_ARG, "myqc_fock_G: precision %d", precision);
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "myqc_fock_G: call myqc_eri_build first");
    MYQC_CHECK(ensure_work(ctx));
    const size_t nn = (size_t)ctx->n * ctx->n;
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_G, D, sizeof(double) * nn, hipMemcpyHostToDevice, ctx->stream));
    double *d_out = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_out, sizeof(double) * nn));
    int rc = 0;
    if (precision == MYQC_FP32 && hipMemsetAsync(d_out, 0, sizeof(double) * nn, ctx->stream) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "memset failed");
    if (rc == 0) rc = digest_G(ctx, ctx->d_G, d_out, precision);
    if (rc == 0 && d2h(ctx, G, d_out, sizeof(double) * nn) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_fock_G: copy failed");
    dfree((void **)&d_out);
    return rc;
}

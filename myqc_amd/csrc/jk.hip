// jk.hip -- J/K (Fock G) digestion of the stored integrals.
//
// The reference's two-electron part of the Fock matrix is
//     G(i,j) = sum_kl D(k,l)[XX(i,j,k,l) - XX(i,k,j,l)/4 - XX(i,l,j,k)/4]
// (src/I2G/RHFI2G.f90:80-90), which only sees D through D + D^T, so it is
// evaluated for the symmetrised D.  Each unique shell quartet is stored once,
// pre-weighted, under its bra pair (digest_plan.h describes the layout, the
// schedule and the six updates each stored element makes).  A workgroup
// holds in LDS the rows of G of its row shell A and of its units' B shells,
// and the same rows of D; its wavefronts take its column chunks (one ket
// class, W consecutive Schwarz ranks, KPL kets per lane), longest first, as
// they become free, and per chunk walk the rows (unit, B function) of every
// unit that has kets there, the next row's loads in flight during the
// current row's arithmetic:
//   J, bra side:  reduced across the wavefront into row a, column b;
//   J, ket side:  summed per lane over the chunk's units, written once per
//                 chunk to the chunk's partial slot;
//   K:            four updates per element into rows a and b (LDS atomics).
// FP32 values (the incremental passes) are accumulated in FP64 and their
// partial rows and slots written in FP32.  A gather sums, per AO row, the
// partial rows of the workgroups holding it, and per ket, the slots of the
// chunks that cover it, in FP64; G = A + A^T.  No global atomics in the
// digestion; the LDS atomics make the summation order, and so the last
// bits, vary from run to run.
#include <algorithm>
#include <cmath>
#include <vector>

#include "context.h"
#include "digest_plan.h"
#include "wave.h"

namespace myqc {

template <int C> struct PL;
template <> struct PL<0> { static constexpr int a = 0, b = 0; };
template <> struct PL<1> { static constexpr int a = 1, b = 0; };
template <> struct PL<2> { static constexpr int a = 1, b = 1; };

template <typename TV>
struct DigArgs {
    const int *meta0, *meta1, *meta2;         // ket AOs by class and Schwarz rank: aoC | aoD << 15
    const TV *dk0, *dk1, *dk2;                // ket density blocks by class and rank (1, 3, 9 values)
    const int *wg;                            // kWgInts per workgroup (digest_plan.h)
    const int *unit;                          // kUnitInts per unit
    const long long *unit_off;                // [u][3] value offsets
    const int4 *chunk;                        // {class, first rank, kets, partial slot}
    const int *row_ao;                        // AO of each partial row
    const TV *x;                              // stored (pre-weighted) values
    const TV *D;                              // symmetric density (or change), n x n
    TV *partial;                              // [row][n] (FP32 for the incremental passes)
    TV *jk_part;                              // [slot][kJkSlot]
    int n;
    double jscale;                            // 1, or 0 for the exchange part alone
    // density-weighted screen of the incremental (FP32) passes: a unit's ket
    // of rank r in class c is skipped when q_unit * q_c[r] * dmax < qtau
    // (0: no screen); q_c are the kets' Schwarz factors by rank (descending)
    const float *q0, *q1, *q2;
    const unsigned *dmax;                     // max |D| of this pass (device, float bits)
    double qtau;
};

// The schedule's records are read-only for the whole launch: read through
// the constant address space, so that the wave-uniform reads of a unit,
// chunk or workgroup record become scalar loads and their values stay in
// scalar registers.
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T *p) {
    return (cptr<T>)p;
}

// One row of the walk over a column chunk for bra AO a = ia: row ib of
// unit u's block (B's function ib) over the lanes' kets.  off: value offset
// of its element 0 for rank 0 (element e of ket rank r at off + e len + r);
// bit j of act: the lane's ket j is in the unit's range and passes the
// screen, ri[j]: its rank (an inactive lane reads the unit's first ket, in range,
// and drops it).
template <int KPL>
struct DigRow {
    long long off;
    int u, len, nb, ib, aoB, sb;
    int ri[KPL];
    unsigned act;
};

// The first unit at or after u (of u1) whose rank range meets the chunk
// [r0, r0 + nk) and that the density screen keeps, as its row 0 for bra AO
// ia; false when there is none.  Every value but act / ri is wave-uniform;
// qv[j]: the Schwarz factor of the lane's ket j.  The screen's wave-uniform
// reads are scalar loads, so that no wait for them also waits for the
// stored values in flight.
template <int CK, typename TV, int KPL>
__device__ __forceinline__ bool dig_find(const DigArgs<TV> &a, int u, int u1, int r0, int nk, int ia, double dmax,
                                         bool screen, const float *qk, const int (&rk)[KPL], const float (&qv)[KPL],
                                         const bool (&valid)[KPL], DigRow<KPL> &row) {
    constexpr int NCD = dig_ncd(CK);
#pragma unroll 1
    for (; u < u1; ++u) {
        const cptr<int> ur = as_const(a.unit) + (size_t)kUnitInts * u;
        const int lo = ur[3 + 2 * CK], hi = ur[4 + 2 * CK];
        if (hi <= r0 || lo >= r0 + nk) continue;
        double thr = 0.0;
        if (screen) {
            // |x| <= q_unit q_ket, kets in descending order: the unit drops
            // out whole when its first ket in the chunk is below the screen,
            // runs whole when its last one is above, else tests ket by ket
            const double qu = (double)__int_as_float(ur[9]) * dmax;
            const int rf = max(lo, r0), rl = min(hi, r0 + nk) - 1;
            if ((double)as_const(qk)[rf] * qu < a.qtau) continue;
            if ((double)as_const(qk)[rl] * qu < a.qtau) thr = a.qtau / qu;
        }
        row.act = 0u;
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            const bool on = valid[j] && rk[j] >= lo && rk[j] < hi && (thr <= 0.0 || (double)qv[j] >= thr);
            row.act |= on ? 1u << j : 0u;
            row.ri[j] = on ? rk[j] : lo;
        }
        // (wave-uniform values, made visibly so: the walk's branches on
        // them then stay scalar)
        row.u = __builtin_amdgcn_readfirstlane(u);
        row.len = __builtin_amdgcn_readfirstlane(hi - lo);
        row.nb = __builtin_amdgcn_readfirstlane((ur[1] >> 15) ? 3 : 1);
        row.ib = 0;
        row.aoB = __builtin_amdgcn_readfirstlane(ur[1] & 0x7fff);
        row.sb = __builtin_amdgcn_readfirstlane(ur[2]);
        row.off = as_const(a.unit_off)[3 * (size_t)u + CK] + (long long)ia * row.nb * NCD * row.len;
        return true;
    }
    return false;
}

// The row after `cur`: the unit's next function, else the next unit's row 0.
template <int CK, typename TV, int KPL>
__device__ __forceinline__ bool dig_next(const DigArgs<TV> &a, const DigRow<KPL> &cur, int u1, int r0, int nk, int ia,
                                         double dmax, bool screen, const float *qk, const int (&rk)[KPL],
                                         const float (&qv)[KPL], const bool (&valid)[KPL], DigRow<KPL> &nxt) {
    const int ib = __builtin_amdgcn_readfirstlane(cur.ib), nb = __builtin_amdgcn_readfirstlane(cur.nb);
    if (ib + 1 < nb) {
        nxt = cur;
        nxt.ib = ib + 1;
        nxt.off = cur.off + (long long)dig_ncd(CK) * cur.len;
        return true;
    }
    return dig_find<CK>(a, cur.u + 1, u1, r0, nk, ia, dmax, screen, qk, rk, qv, valid, nxt);
}

template <int KPL, int NCD, typename TV>
__device__ __forceinline__ void dig_load(const TV *x, const DigRow<KPL> &row, TV (&v)[KPL][NCD]) {
#pragma unroll
    for (int e = 0; e < NCD; ++e)
#pragma unroll
        for (int j = 0; j < KPL; ++j) v[j][e] = x[row.off + (long long)e * row.len + row.ri[j]];
}

// The same loads from row `nx` when `more`, else again from row `cu`.
// Issued on every path, so that the wait counts the compiler places before
// the current row's arithmetic are the same on all paths and never cover
// these loads.
template <int KPL, int NCD, typename TV>
__device__ __forceinline__ void dig_load_sel(const TV *x, bool more, const DigRow<KPL> &nx, const DigRow<KPL> &cu,
                                             TV (&v)[KPL][NCD]) {
    const long long off = more ? nx.off : cu.off;
    const int len = more ? nx.len : cu.len;
#pragma unroll
    for (int e = 0; e < NCD; ++e)
#pragma unroll
        for (int j = 0; j < KPL; ++j) v[j][e] = x[off + (long long)e * len + (more ? nx.ri[j] : cu.ri[j])];
}

// The arithmetic of one row (bra AO a = ia, B's function ib) on its loaded
// values v.  s_acc / s_d: the workgroup's LDS rows of G and D (A's rows
// first, B's from row sb); dA[j]: D(a, c), D(a, d) at the lane's ket j's
// columns; jk / ka: the lane's ket-side J sums and its row-a K sums,
// carried over the chunk's units (every unit of a workgroup shares row
// shell A).
template <int CK, typename TV, int KPL = dig_kpl(CK), int NCD = dig_ncd(CK),
          int NK = ncart(PL<CK>::a) + ncart(PL<CK>::b)>
__device__ __forceinline__ void digest_row(const DigArgs<TV> &a, double *s_acc, const TV *s_d, int ia,
                                           const DigRow<KPL> &row, const TV (&v)[KPL][NCD], const int (&aoC)[KPL],
                                           const int (&aoD)[KPL], const TV (&dq)[KPL][NCD], const TV (&dA)[KPL][NK],
                                           double (&jk)[KPL][NCD], TV (&ka)[KPL][NK]) {
    constexpr int nc = ncart(PL<CK>::a), nd = ncart(PL<CK>::b);
    static_assert(nc * nd == NCD && nc + nd == NK, "ket block");
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int colB = row.aoB + row.ib, rowB = row.sb + row.ib;
    TV x[KPL][NCD];
#pragma unroll
    for (int j = 0; j < KPL; ++j)
#pragma unroll
        for (int e = 0; e < NCD; ++e) x[j][e] = (row.act >> j) & 1u ? v[j][e] : (TV)0;
    // J, bra side: A(a, b) += 2 sum_cd D(c, d) x
    TV t = 0;
#pragma unroll
    for (int j = 0; j < KPL; ++j)
#pragma unroll
        for (int e = 0; e < NCD; ++e) t += dq[j][e] * x[j][e];
    t = wave_total(t);
    if (lane == 0 && a.jscale != 0.0) atomicAdd(&s_acc[ia * n + colB], 2.0 * a.jscale * (double)t);
    const TV dab = s_d[ia * n + colB];   // D(a, b), wave-uniform
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        // J, ket side: A(c, d) += 2 D(a, b) x
#pragma unroll
        for (int e = 0; e < NCD; ++e) jk[j][e] += (double)(dab * x[j][e]);
        // K: A(a, c) -= D(b, d) x / 2, A(a, d) -= D(b, c) x / 2 (into ka),
        //    A(b, c) -= D(a, d) x / 2, A(b, d) -= D(a, c) x / 2 (LDS)
        TV dB[NK], kb[NK];
#pragma unroll
        for (int c = 0; c < nc; ++c) dB[c] = s_d[rowB * n + aoC[j] + c];
#pragma unroll
        for (int d = 0; d < nd; ++d) dB[nc + d] = s_d[rowB * n + aoD[j] + d];
#pragma unroll
        for (int k = 0; k < NK; ++k) kb[k] = 0;
#pragma unroll
        for (int c = 0; c < nc; ++c)
#pragma unroll
            for (int d = 0; d < nd; ++d) {
                const TV xv = x[j][c * nd + d];
                ka[j][c] += dB[nc + d] * xv;
                ka[j][nc + d] += dB[c] * xv;
                kb[c] += dA[j][nc + d] * xv;
                kb[nc + d] += dA[j][c] * xv;
            }
        if ((row.act >> j) & 1u) {
#pragma unroll
            for (int c = 0; c < nc; ++c) atomicAdd(&s_acc[rowB * n + aoC[j] + c], -0.5 * (double)kb[c]);
#pragma unroll
            for (int d = 0; d < nd; ++d) atomicAdd(&s_acc[rowB * n + aoD[j] + d], -0.5 * (double)kb[nc + d]);
        }
    }
}

// Half a step of the walk: find the row after `cur` into `nxt` and issue
// its loads into `vn`, then do `cur`'s arithmetic on `vc`.  The caller
// alternates two row records and two register buffers, so the loads of one
// row are in flight during the arithmetic of the one before.
template <int CK, typename TV, int KPL = dig_kpl(CK), int NCD = dig_ncd(CK),
          int NK = ncart(PL<CK>::a) + ncart(PL<CK>::b)>
__device__ __forceinline__ bool dig_half(const DigArgs<TV> &a, double *s_acc, const TV *s_d, int ia,
                                         const DigRow<KPL> &cur, DigRow<KPL> &nxt, const TV (&vc)[KPL][NCD],
                                         TV (&vn)[KPL][NCD], int u1, int r0, int nk, double dmax, bool screen,
                                         const float *qk, const int (&rk)[KPL], const float (&qv)[KPL],
                                         const bool (&valid)[KPL], const int (&aoC)[KPL], const int (&aoD)[KPL],
                                         const TV (&dq)[KPL][NCD], const TV (&dA)[KPL][NK], double (&jk)[KPL][NCD],
                                         TV (&ka)[KPL][NK]) {
    const bool more =
        __builtin_amdgcn_readfirstlane((int)dig_next<CK>(a, cur, u1, r0, nk, ia, dmax, screen, qk, rk, qv, valid, nxt));
    dig_load_sel(a.x, more, nxt, cur, vn);
    digest_row<CK>(a, s_acc, s_d, ia, cur, vc, aoC, aoD, dq, dA, jk, ka);
    return more;
}

// One column chunk: the lanes' kets; then for each bra row a of the row
// shell A, every unit of the workgroup whose rank range meets the chunk, and
// row a's K sums (LDS); then the chunk's ket-side partial slot.
template <int LA, int CK, typename TV>
__device__ __forceinline__ void digest_column(const DigArgs<TV> &a, double *s_acc, const TV *s_d, int4 ch, int u0,
                                              int u1, double dmax, bool screen) {
    constexpr int na = ncart(LA), nc = ncart(PL<CK>::a), nd = ncart(PL<CK>::b), NCD = nc * nd, NK = nc + nd;
    constexpr int KPL = dig_kpl(CK);
    const int n = a.n;
    const int lane = threadIdx.x & 63;
    const int r0 = ch.y, nk = ch.z;
    const int *meta = CK == 0 ? a.meta0 : CK == 1 ? a.meta1 : a.meta2;
    const TV *dkc = CK == 0 ? a.dk0 : CK == 1 ? a.dk1 : a.dk2;
    const float *qk = CK == 0 ? a.q0 : CK == 1 ? a.q1 : a.q2;
    int rk[KPL], aoC[KPL], aoD[KPL];
    bool valid[KPL];
    float qv[KPL];
    TV dq[KPL][NCD];
    double jk[KPL][NCD];
#pragma unroll
    for (int j = 0; j < KPL; ++j) {
        const int k = lane + 64 * j;
        valid[j] = k < nk;
        rk[j] = r0 + (valid[j] ? k : 0);
        const int m = meta[rk[j]];
        aoC[j] = m & 0x7fff;
        aoD[j] = (m >> 15) & 0x7fff;
        qv[j] = screen ? qk[rk[j]] : 0.0f;
#pragma unroll
        for (int e = 0; e < NCD; ++e) {
            dq[j][e] = valid[j] ? dkc[(size_t)rk[j] * NCD + e] : (TV)0;
            jk[j][e] = 0.0;
        }
    }
#pragma unroll 1
    for (int ia = 0; ia < na; ++ia) {
        TV dA[KPL][NK], ka[KPL][NK];
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
#pragma unroll
            for (int c = 0; c < nc; ++c) dA[j][c] = s_d[ia * n + aoC[j] + c];
#pragma unroll
            for (int d = 0; d < nd; ++d) dA[j][nc + d] = s_d[ia * n + aoD[j] + d];
#pragma unroll
            for (int k = 0; k < NK; ++k) ka[j][k] = 0;
        }
        DigRow<KPL> r0w, r1w;
        if (dig_find<CK>(a, u0, u1, r0, nk, ia, dmax, screen, qk, rk, qv, valid, r0w)) {
            TV v0[KPL][NCD], v1[KPL][NCD];
            dig_load(a.x, r0w, v0);
#pragma unroll 1
            for (;;) {
                if (!dig_half<CK>(a, s_acc, s_d, ia, r0w, r1w, v0, v1, u1, r0, nk, dmax, screen, qk, rk, qv, valid,
                                  aoC, aoD, dq, dA, jk, ka))
                    break;
                if (!dig_half<CK>(a, s_acc, s_d, ia, r1w, r0w, v1, v0, u1, r0, nk, dmax, screen, qk, rk, qv, valid,
                                  aoC, aoD, dq, dA, jk, ka))
                    break;
            }
        }
#pragma unroll
        for (int j = 0; j < KPL; ++j) {
            if (!valid[j]) continue;
#pragma unroll
            for (int c = 0; c < nc; ++c) atomicAdd(&s_acc[ia * n + aoC[j] + c], -0.5 * (double)ka[j][c]);
#pragma unroll
            for (int d = 0; d < nd; ++d) atomicAdd(&s_acc[ia * n + aoD[j] + d], -0.5 * (double)ka[j][nc + d]);
        }
    }
    TV *dst = a.jk_part + (size_t)ch.w * kJkSlot;
    const double js = 2.0 * a.jscale;
#pragma unroll
    for (int j = 0; j < KPL; ++j)
#pragma unroll
        for (int e = 0; e < NCD; ++e) dst[(lane + 64 * j) * NCD + e] = (TV)(js * jk[j][e]);
}

template <int LA, typename TV>
__device__ __forceinline__ void digest_column_la(const DigArgs<TV> &a, double *s_acc, const TV *s_d, int4 ch, int u0,
                                                 int u1, double dmax, bool screen) {
    if (ch.x == 0) digest_column<LA, 0>(a, s_acc, s_d, ch, u0, u1, dmax, screen);
    else if (ch.x == 1) digest_column<LA, 1>(a, s_acc, s_d, ch, u0, u1, dmax, screen);
    else digest_column<LA, 2>(a, s_acc, s_d, ch, u0, u1, dmax, screen);
}

// One workgroup: stage its D rows, run its column chunks (wavefront w takes
// chunk w first, then whichever chunk is next when it is free), write its
// rows of A to its partial rows.  Three
// workgroups per CU for the FP64 form (LDS: 16 B per row element); the FP32
// form's rows take 12 B, so it runs four (at 128 VGPRs, a few spilled: 5-7 %
// faster per pass than three).
template <typename TV>
__global__ __launch_bounds__(256, sizeof(TV) == 4 ? 4 : 3) void k_digest(DigArgs<TV> a) {
    extern __shared__ __align__(16) unsigned char smem[];   // after s_next: doubles need 8-byte alignment
    const int n = a.n;
    const cptr<int> wr = as_const(a.wg) + (size_t)kWgInts * blockIdx.x;
    const int LA = wr[0] >> 15, u0 = wr[1], u1 = wr[2], c0 = wr[3], c1 = wr[4], nrows = wr[5], row0 = wr[6];
    __shared__ int s_next;   // the next chunk not yet taken
    double *s_acc = reinterpret_cast<double *>(smem);
    TV *s_d = reinterpret_cast<TV *>(s_acc + (size_t)nrows * n);
    if (threadIdx.x == 0) s_next = c0 + 4;
#pragma unroll 1
    for (int i = threadIdx.x; i < nrows * n; i += blockDim.x) {
        s_acc[i] = 0.0;
        const int r = i / n, j = i - r * n;
        s_d[i] = a.D[j + (size_t)n * a.row_ao[row0 + r]];   // D(row, j) = D(j, row): coalesced
    }
    __syncthreads();
    const bool screen = a.qtau > 0.0;
    const double dmax = screen ? (double)__uint_as_float(*a.dmax) : 0.0;
    // the wavefront's index, made visibly uniform so that its chunk records
    // are scalar loads; wavefront w starts with chunk c0 + w, then takes the
    // next chunk not yet taken (the chunks are listed longest first)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
#pragma unroll 1
    for (int c = c0 + wave; c < c1;) {
        const cptr<int> cr = as_const(reinterpret_cast<const int *>(a.chunk)) + 4 * (size_t)c;
        const int4 ch = make_int4(cr[0], cr[1], cr[2], cr[3]);
        if (LA == 0) digest_column_la<0>(a, s_acc, s_d, ch, u0, u1, dmax, screen);
        else digest_column_la<1>(a, s_acc, s_d, ch, u0, u1, dmax, screen);
        int t = 0;
        if (lane == 0) t = atomicAdd(&s_next, 1);
        c = __builtin_amdgcn_readfirstlane(t);
    }
    __syncthreads();
    TV *dst = a.partial + (size_t)row0 * n;
#pragma unroll 1
    for (int i = threadIdx.x; i < nrows * n; i += blockDim.x) dst[i] = (TV)s_acc[i];
}

// The ket-side J of every pair: jk[base(c) + r ncd(c) + e] = sum of the
// partial slots of the chunks that cover rank r of class c.  One workgroup
// per group of at most kJkGroup slots of one block of W(c) ranks (positions
// of the block across the threads, coalesced; four independent sums), added
// into jk (zeroed before the pass) with one atomic per position.
template <typename PT>
__global__ __launch_bounds__(256) void k_gather_jk(const PT *__restrict__ jk_part, const int *__restrict__ list,
                                                   const int *__restrict__ group, int3 base, int3 block0, int4 nps,
                                                   double *__restrict__ jk) {
    const int g = blockIdx.x;
    const int blk = group[3 * g], k0 = group[3 * g + 1], k1 = group[3 * g + 2];
    const int c = blk >= block0.z ? 2 : blk >= block0.y ? 1 : 0;
    const int ncd = dig_ncd(c), W = dig_width(c), npos = W * ncd;
    const int r0 = (blk - (c == 0 ? block0.x : c == 1 ? block0.y : block0.z)) * W;
    const int np = c == 0 ? nps.x : c == 1 ? nps.y : nps.z;
    const int b0 = c == 0 ? base.x : c == 1 ? base.y : base.z;
    for (int pos = threadIdx.x; pos < npos; pos += blockDim.x) {
        if (r0 + pos / ncd >= np) break;
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
        int k = k0;
        for (; k + 4 <= k1; k += 4)
#pragma unroll
            for (int q = 0; q < 4; ++q) s4[q] += jk_part[(size_t)list[k + q] * kJkSlot + pos];
        for (; k < k1; ++k) s4[0] += jk_part[(size_t)list[k] * kJkSlot + pos];
        atomicAdd(&jk[b0 + (size_t)r0 * ncd + pos], (s4[0] + s4[1]) + (s4[2] + s4[3]));
    }
}

// A(i, j) = the sum of the partial rows holding AO i, plus the ket-side J of
// the pair (shell(i), shell(j)) when it is stored in that orientation.  One
// workgroup per (AO, 64 columns); its four thread rows take every fourth
// partial row (two independent sums each) and meet in LDS in a fixed order.
template <typename PT>
__global__ __launch_bounds__(256) void k_gather_rows(const PT *__restrict__ partial,
                                                     const int *__restrict__ list_off, const int *__restrict__ list,
                                                     const double *__restrict__ jk, const int *__restrict__ jkoff,
                                                     const int *__restrict__ ao_shell, const int *__restrict__ shell,
                                                     int ns, int n, double *__restrict__ Graw) {
    __shared__ double s_sum[4][64];
    const int j = blockIdx.x * 64 + threadIdx.x, i = blockIdx.y, q = threadIdx.y;
    double s = 0.0;
    if (j < n) {
        const int k0 = list_off[i], k1 = list_off[i + 1];
        double s2[2] = {0.0, 0.0};
        int k = k0 + q;
        for (; k + 4 < k1; k += 8) {
            s2[0] += partial[(size_t)list[k] * n + j];
            s2[1] += partial[(size_t)list[k + 4] * n + j];
        }
        if (k < k1) s2[0] += partial[(size_t)list[k] * n + j];
        s = s2[0] + s2[1];
    }
    s_sum[q][threadIdx.x] = s;
    __syncthreads();
    if (q != 0 || j >= n) return;
    s = (s_sum[0][threadIdx.x] + s_sum[1][threadIdx.x]) + (s_sum[2][threadIdx.x] + s_sum[3][threadIdx.x]);
    const int X = ao_shell[i], Y = ao_shell[j];
    const int o = jkoff[X * ns + Y];
    if (o >= 0) s += jk[o + (i - shell[4 * X]) * (shell[4 * Y + 1] ? 3 : 1) + (j - shell[4 * Y])];
    Graw[i + (size_t)n * j] = s;
}

__global__ void k_zero(double *__restrict__ p, size_t cnt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

// The kets' density blocks of the three classes in Schwarz-rank order,
// dk_c[r (nc nd) + c nd + d] = D(ao_a + c, ao_b + d) of pair sorted_c[r], with
// dk_c at dk + {0, np0, np0 + 3 np1}; and the zeroing of jk (njk values), in
// one launch
template <typename TV>
__global__ void k_pack_kets_all(const ShellPair *__restrict__ pairs, const int *__restrict__ s0,
                                const int *__restrict__ s1, const int *__restrict__ s2, int np0, int np1, int np2,
                                const TV *__restrict__ D, int n, TV *__restrict__ dk, double *__restrict__ jk,
                                size_t njk) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < njk) jk[id] = 0.0;
    if (id >= (size_t)np0 + np1 + np2) return;
    int r = (int)id, nc = 1, nd = 1;
    const int *sorted = s0;
    TV *out = dk;
    if (r >= np0 + np1) {
        r -= np0 + np1;
        sorted = s2;
        nc = nd = 3;
        out = dk + np0 + 3 * (size_t)np1;
    } else if (r >= np0) {
        r -= np0;
        sorted = s1;
        nc = 3;
        out = dk + np0;
    }
    const ShellPair sp = pairs[sorted[r]];
    for (int c = 0; c < nc; ++c)
        for (int d = 0; d < nd; ++d) out[(size_t)r * nc * nd + c * nd + d] = D[(sp.ao_a + c) + (size_t)n * (sp.ao_b + d)];
}

// --- small elementwise kernels ------------------------------------------------


// Ds = scale (E + E^T), E = D - Dm (Dm may be NULL: E = D); with amax, also
// max |Ds32| into *amax as the bit pattern of a non-negative float (ordered
// like the value; *amax starts at 0)
__global__ __launch_bounds__(128) void k_symmetrize(const double *__restrict__ D, double *__restrict__ Ds, float *__restrict__ Ds32, int n,
                             double scale, unsigned *__restrict__ amax = nullptr,
                             const double *__restrict__ Dm = nullptr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    float m = 0.0f;
    if (i < n) {
        const size_t ij = i + (size_t)n * j, ji = j + (size_t)n * i;
        const double v = Dm ? scale * ((D[ij] - Dm[ij]) + (D[ji] - Dm[ji])) : scale * (D[ij] + D[ji]);
        if (Ds) Ds[ij] = v;
        if (Ds32) {
            Ds32[ij] = (float)v;
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





// G += A + A^T (the incremental passes); the pass's max |dD| word is
// cleared for the next pass (the digestion has read it)
__global__ __launch_bounds__(128) void k_sym_add(const double *__restrict__ A, double *__restrict__ G, int n,
                                                 unsigned *__restrict__ amax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i < n) G[i + (size_t)n * j] += A[i + (size_t)n * j] + A[j + (size_t)n * i];
    if (i == 0 && j == 0) *amax = 0u;
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
    // the incremental passes' max |dD| word starts at 0 (k_sym_add clears it
    // after each pass)
    MYQC_HIP(ctx, hipMemsetAsync(ctx->d_dmax, 0, sizeof(double), ctx->stream));
    ctx->work_n = ctx->n;
    return 0;
}

template <typename TV>
static int launch_digest(myqc_ctx *ctx, const TV *D, const TV *x, TV *dk, double *partial, double *Graw,
                         double jscale = 1.0, double qtau = 0.0) {
    const DigestPlan &pl = ctx->plan;
    const int n = ctx->n;
    const int np0 = (int)ctx->pair_ids[0].size(), np1 = (int)ctx->pair_ids[1].size(),
              np2 = (int)ctx->pair_ids[2].size();
    TV *dk0 = dk, *dk1 = dk + np0, *dk2 = dk + np0 + 3 * (size_t)np1;
    // the kets' density blocks, and jk cleared (the ket-side J gather adds into it)
    const size_t njk = (size_t)std::max(pl.jk_base[3], 1);
    const size_t npack = std::max(njk, (size_t)np0 + np1 + np2);
    k_pack_kets_all<TV><<<(unsigned)((npack + 255) / 256), 256, 0, ctx->stream>>>(
        ctx->d_pairs, ctx->d_sorted[0], ctx->d_sorted[1], ctx->d_sorted[2], np0, np1, np2, D, n, dk, ctx->d_jk, njk);
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
        a.unit_off = ctx->d_unit_off;
        a.chunk = reinterpret_cast<const int4 *>(ctx->d_chunk);
        a.row_ao = ctx->d_row_ao;
        a.x = x;
        a.D = D;
        // the partial rows and ket-side slots in the pass's own precision
        // (sums of FP32 products, accumulated in FP64, rounded once to FP32
        // in the incremental passes), summed in FP64 by the gathers; the
        // FP64-sized buffers hold either
        a.partial = reinterpret_cast<TV *>(partial);
        a.jk_part = reinterpret_cast<TV *>(ctx->d_jk_part);
        a.n = n;
        a.jscale = jscale;
        a.q0 = ctx->d_qrank[0];
        a.q1 = ctx->d_qrank[1];
        a.q2 = ctx->d_qrank[2];
        a.dmax = reinterpret_cast<const unsigned *>(ctx->d_dmax);
        a.qtau = qtau;
        const size_t shm = (sizeof(double) + sizeof(TV)) * (size_t)pl.max_rows * n;
        // timed alone: algorithmic bytes per pass are one read of every
        // stored unique integral plus the density in and G out (16 n^2)
        hipEvent_t ev;
        MYQC_CHECK(timer_begin(ctx, T_JK, &ev));
        k_digest<TV><<<ctx->n_wg, 256, shm, ctx->stream>>>(a);
        MYQC_CHECK(timer_end(ctx, T_JK, ev, (double)sizeof(TV) * (double)ctx->n_stored + 16.0 * n * n));
        const int ngroups = (int)(pl.jk_group.size() / 3);
        if (ngroups > 0)
            k_gather_jk<TV><<<ngroups, 256, 0, ctx->stream>>>(
                reinterpret_cast<const TV *>(ctx->d_jk_part), ctx->d_jk_list, ctx->d_jk_group, make_int3(pl.jk_base[0], pl.jk_base[1], pl.jk_base[2]),
                make_int3(pl.jk_block0[0], pl.jk_block0[1], pl.jk_block0[2]),
                make_int4((int)ctx->pair_ids[0].size(), (int)ctx->pair_ids[1].size(), (int)ctx->pair_ids[2].size(), 0),
                ctx->d_jk);
    }
    k_gather_rows<TV><<<dim3((n + 63) / 64, n), dim3(64, 4), 0, ctx->stream>>>(
        reinterpret_cast<const TV *>(partial), ctx->d_row_list_off, ctx->d_row_list, ctx->d_jk, ctx->d_jkoff, ctx->d_ao_shell, ctx->d_shell,
        ctx->ns, n, Graw);
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
int digest_G(myqc_ctx *ctx, const double *d_Din, double *d_Gout, int precision, const double *d_Dprev) {
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "no stored integrals (call myqc_eri_build)");
    MYQC_CHECK(ensure_work(ctx));
    const int n = ctx->n;
    const size_t nn = (size_t)n * n;
    if (precision == MYQC_FP64) {
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, ctx->d_D, nullptr, n, 0.5);
        MYQC_CHECK(launch_digest<double>(ctx, ctx->d_D, ctx->d_x64, ctx->d_Dpair, ctx->d_partial, ctx->d_A));
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, d_Gout, nullptr, n, 1.0);
    } else {
        if (!ctx->have_eri32) return set_error(ctx, MYQC_E_STATE, "no FP32 integrals stored");
        // the density change (Din - Dprev when Dprev is given) symmetrised
        // into FP32, with the pass's max |dD| for the screen folded in; the
        // word is 0 here (k_sym_add clears it at the end of each pass)
        unsigned *amax = reinterpret_cast<unsigned *>(ctx->d_dmax);
        k_symmetrize<<<grid2(n), 128, 0, ctx->stream>>>(d_Din, nullptr, ctx->d_D32, n, 0.5, amax, d_Dprev);
        MYQC_CHECK(launch_digest<float>(ctx, ctx->d_D32, ctx->d_x32, ctx->d_Dpair32, ctx->d_partial, ctx->d_A, 1.0,
                                        kIncrementalScreen));
        k_sym_add<<<grid2(n), 128, 0, ctx->stream>>>(ctx->d_A, d_Gout, n, amax);
        (void)nn;
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

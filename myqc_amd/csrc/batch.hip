// batch.hip -- RHF of thousands of small molecules in one persistent launch.
//
// The reference solves one molecule per run of its programs; a batch of
// 2-8 atom molecules (2-40 functions each) is far too small per molecule to
// fill a GPU, so everything here is organised around the batch:
//
//   k_batch_one_e        one thread per (molecule, shell pair): S and H
//   k_batch_eri<CB,CK>   one thread per (molecule, unique shell quartet) of a
//                        pair-class combination: the same OS/HGP code and
//                        reference numerics as the single-molecule path,
//                        written into the molecule's packed 8-fold-unique
//                        array (ij|kl), i>=j, k>=l, ij>=kl
//   k_batch_scf          persistent: each workgroup takes molecules from an
//                        atomic queue (largest first) and runs the whole SCF
//                        on-chip -- G(D) by streaming the packed integrals
//                        from HBM into LDS accumulators, DIIS, and a cyclic
//                        Jacobi eigensolver in LDS -- with no host round trip
//
// The launch shapes (threads per ERI and SCF workgroup) are chosen by timing
// the candidates on the batch itself (batch_tune), and the whole sequence is
// captured once into a hipGraph that later calls replay.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "context.h"
#include "guess.h"
#include "ket_soa.h"
#include "one_e_math.h"

namespace myqc {

constexpr int kBatchDiis = 8;
constexpr int kBatchMaxN = 40;

struct BMol {
    int n, nocc, nnuc, atom0;
    int cls0[3], ncls[3];    // pairs of class c: cls_ids[c][cls0[c] .. cls0[c] + ncls[c])
    long long eri0;          // packed (ij|kl) offset
    long long mat0;          // n x n blocks: offset (sum of n^2 before)
    double e_nuc;
};

template <int C> struct BPL;
template <> struct BPL<0> { static constexpr int a = 0, b = 0; };
template <> struct BPL<1> { static constexpr int a = 1, b = 0; };
template <> struct BPL<2> { static constexpr int a = 1, b = 1; };

__host__ __device__ __forceinline__ long long tri(long long a) { return a * (a + 1) / 2; }

// ---------------------------------------------------------------------------
// one-electron integrals
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_batch_one_e(const ShellPair *__restrict__ pairs,
                                                     const PrimPair *__restrict__ prims, const double *__restrict__ ftT,
                                                     const double *__restrict__ atoms, const int *__restrict__ pair_mol,
                                                     const BMol *__restrict__ mols, int npairs, double *__restrict__ S,
                                                     double *__restrict__ H) {
    __shared__ double s_ft[FT_TOTAL];
    for (int i = threadIdx.x; i < FT_TOTAL; i += blockDim.x) s_ft[i] = ftT[i];
    __syncthreads();
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const ShellPair sp = pairs[t];
    const BMol mol = mols[pair_mol[t]];
    double s[3][3], h[3][3];
    one_e_pair(sp, prims, s_ft, atoms + 4 * mol.atom0, mol.nnuc, s, h);
    const int na = sp.la ? 3 : 1, nb = sp.lb ? 3 : 1;
    const long long n = mol.n;
    double *Sm = S + mol.mat0, *Hm = H + mol.mat0;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            const long long i = sp.ao_a + a, j = sp.ao_b + b;
            Sm[i + n * j] = s[a][b];
            Sm[j + n * i] = s[a][b];
            Hm[i + n * j] = h[a][b];
            Hm[j + n * i] = h[a][b];
        }
}

// ---------------------------------------------------------------------------
// two-electron integrals into packed per-molecule storage
// ---------------------------------------------------------------------------
struct BEriArgs {
    const ShellPair *pairs;
    const PrimPair *prims;
    const double *ftT;
    const BMol *mols;
    const int *cls0, *cls1, *cls2;
    const double *ksoa;      // primitive pairs of the ket class list (ket_soa.h)
    long long nket;          // entries of the ket class list
    const long long *qcum;   // nmol + 1 prefix counts of this class combination
    int nmol;
    long long nq;
    double *eri;
};

template <int CB, int CK>
__global__ __launch_bounds__(256) void k_batch_eri(BEriArgs a) {
    __shared__ double s_ft[FT_TOTAL];
    for (int i = threadIdx.x; i < FT_TOTAL; i += blockDim.x) s_ft[i] = a.ftT[i];
    __syncthreads();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq) return;
    int lo = 0, hi = a.nmol;   // last m with qcum[m] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.qcum[mid] <= t) lo = mid;
        else hi = mid;
    }
    const BMol mol = a.mols[lo];
    const long long q = t - a.qcum[lo];
    long long p, r;
    if (CB == CK) {
        p = (long long)floor((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
        while (tri(p) > q) --p;
        while (tri(p + 1) <= q) ++p;
        r = q - tri(p);
    } else {
        const int nk = mol.ncls[CK];
        p = q / nk;
        r = q - p * nk;
    }
    const int *idb = CB == 0 ? a.cls0 : CB == 1 ? a.cls1 : a.cls2;
    const int *idk = CK == 0 ? a.cls0 : CK == 1 ? a.cls1 : a.cls2;
    const ShellPair bra = a.pairs[idb[mol.cls0[CB] + p]], ket = a.pairs[idk[mol.cls0[CK] + r]];
    constexpr int LA = BPL<CB>::a, LB = BPL<CB>::b, LC = BPL<CK>::a, LD = BPL<CK>::b;
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);
    QuartetAcc<LA, LB, LC, LD> acc;
    quartet_prims_f<LA, LB, LC, LD>(bra, ket, a.prims, KetSoa{a.ksoa, a.nket, mol.cls0[CK] + r}, s_ft, acc);
    const double CD[3] = {ket.AB[0], ket.AB[1], ket.AB[2]};
    double out[na * nb * nc * nd];
    quartet_hrr<LA, LB, LC, LD>(acc, bra.AB, CD, out);
    double *E = a.eri + mol.eri0;
#pragma unroll
    for (int ia = 0; ia < na; ++ia)
#pragma unroll
        for (int ib = 0; ib < nb; ++ib)
#pragma unroll
            for (int ic = 0; ic < nc; ++ic)
#pragma unroll
                for (int id = 0; id < nd; ++id) {
                    const int i = bra.ao_a + ia, j = bra.ao_b + ib, k = ket.ao_a + ic, l = ket.ao_b + id;
                    const long long IJ = tri(max(i, j)) + min(i, j), KL = tri(max(k, l)) + min(k, l);
                    E[tri(max(IJ, KL)) + min(IJ, KL)] = out[((ia * nb + ib) * nc + ic) * nd + id];
                }
}

// ---------------------------------------------------------------------------
// the persistent SCF
// ---------------------------------------------------------------------------
struct BScfArgs {
    const BMol *mols;
    const int *order;          // molecules, largest first
    int nmol;
    int *counter;              // work queue head (zeroed before the launch)
    const double *eri, *S, *H, *guess;   // guess: NULL or n^2 blocks
    double *scratch;           // 4 n^2 blocks per molecule: X, F', Dp, W
    double *hist;              // 2 kBatchDiis n^2 blocks per molecule
    double *Dout;
    double *energy;
    int *iters, *conv;
    int *fallback;             // NULL, or 1 where a learned start restarted from the core guess
    double e_tol, d_tol;
    int max_iter;
    int learned_iter;          // iterations a start from `guess` gets before the restart
    int restart;               // 1: restart an unconverged start from `guess` from the core guess
    int check_it;              // a start from `guess` whose DIIS error after check_it
    double check_err;          // iterations is above check_err restarts at once
};

// Workgroup reductions: a wavefront butterfly, then the per-wave results
// summed in a fixed order (every thread returns the same value).
__device__ __forceinline__ double wave_red_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_red_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wg_sum(double v, double *s_red) {
    v = wave_red_sum(v);
    __syncthreads();   // s_red free
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += s_red[w];
    return r;
}

__device__ __forceinline__ double wg_max(double v, double *s_red) {
    v = wave_red_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = -1e300;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, s_red[w]);
    return r;
}

// C = alpha A B (or A B^T); all n x n column-major; caller synchronises.
// Operands may sit in HBM: the k loop is unrolled so that their loads issue
// together.
__device__ __forceinline__ void mm(double *C, const double *A, const double *B, int n, bool transB, double alpha) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e % n, j = e / n;
        double s = 0.0;
        if (transB) {
#pragma unroll 8
            for (int k = 0; k < n; ++k) s += A[i + n * k] * B[j + n * k];
        } else {
#pragma unroll 8
            for (int k = 0; k < n; ++k) s += A[i + n * k] * B[k + n * j];
        }
        C[e] = alpha * s;
    }
}

// Cyclic Jacobi on the symmetric A (n x n, LDS): eigenvalues on A's diagonal,
// eigenvectors in the columns of U.  Rounds pair the indices round-robin so
// the n/2 rotations of a round are disjoint; each thread applies J^T A J to
// whole 2x2 blocks in place, so one barrier separates the rounds.
__device__ void jacobi(double *A, double *U, int n, double *s_rot, double *s_red) {
    const int npad = n + (n & 1), half = npad >> 1;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) U[e] = (e % n == e / n) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const double v = A[e];
            if (e % n == e / n) dia += v * v;
            else off += v * v;
        }
        off = wg_sum(off, s_red);
        dia = wg_sum(dia, s_red);
        if (off <= 1e-28 * fmax(dia, 1e-300)) break;
        for (int r = 0; r < npad - 1; ++r) {
            // rotation of pair k: s_rot[4k] = p, [4k+1] = q, [4k+2] = c, [4k+3] = s
            for (int k = threadIdx.x; k < half; k += blockDim.x) {
                const int a = k == 0 ? 0 : ((k - 1 + r) % (npad - 1)) + 1;
                const int kb = npad - 1 - k;
                const int b = ((kb - 1 + r) % (npad - 1)) + 1;
                const int p = min(a, b), q = max(a, b);
                double c = 1.0, s = 0.0;
                if (q < n) {
                    const double apq = A[p + n * q];
                    if (fabs(apq) > 1e-300) {
                        const double th = (A[q + n * q] - A[p + n * p]) / (2.0 * apq);
                        const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        s = t * c;
                    }
                }
                s_rot[4 * k] = p;
                s_rot[4 * k + 1] = q;
                s_rot[4 * k + 2] = c;
                s_rot[4 * k + 3] = s;
            }
            __syncthreads();
            // A <- J^T A J on 2x2 blocks (pair k1 rows, pair k2 columns)
            for (int w = threadIdx.x; w < half * half; w += blockDim.x) {
                const int k1 = w % half, k2 = w / half;
                const int p1 = (int)s_rot[4 * k1], q1 = (int)s_rot[4 * k1 + 1];
                const int p2 = (int)s_rot[4 * k2], q2 = (int)s_rot[4 * k2 + 1];
                const double c1 = s_rot[4 * k1 + 2], s1 = s_rot[4 * k1 + 3];
                const double c2 = s_rot[4 * k2 + 2], s2 = s_rot[4 * k2 + 3];
                const bool q1ok = q1 < n, q2ok = q2 < n;
                const double M00 = A[p1 + n * p2];
                const double M01 = q2ok ? A[p1 + n * q2] : 0.0;
                const double M10 = q1ok ? A[q1 + n * p2] : 0.0;
                const double M11 = (q1ok && q2ok) ? A[q1 + n * q2] : 0.0;
                const double T00 = c1 * M00 - s1 * M10, T01 = c1 * M01 - s1 * M11;
                const double T10 = s1 * M00 + c1 * M10, T11 = s1 * M01 + c1 * M11;
                A[p1 + n * p2] = T00 * c2 - T01 * s2;
                if (q2ok) A[p1 + n * q2] = T00 * s2 + T01 * c2;
                if (q1ok) A[q1 + n * p2] = T10 * c2 - T11 * s2;
                if (q1ok && q2ok) A[q1 + n * q2] = T10 * s2 + T11 * c2;
            }
            // U <- U J
            for (int w = threadIdx.x; w < n * half; w += blockDim.x) {
                const int row = w % n, k = w / n;
                const int p = (int)s_rot[4 * k], q = (int)s_rot[4 * k + 1];
                if (q >= n) continue;
                const double c = s_rot[4 * k + 2], s = s_rot[4 * k + 3];
                const double up = U[row + n * p], uq = U[row + n * q];
                U[row + n * p] = c * up - s * uq;
                U[row + n * q] = s * up + c * uq;
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ void wg_sum2(double &x, double &y, double *s_red) {
    x = wave_red_sum(x);
    y = wave_red_sum(y);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        s_red[2 * (threadIdx.x >> 6)] = x;
        s_red[2 * (threadIdx.x >> 6) + 1] = y;
    }
    __syncthreads();
    x = y = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        x += s_red[2 * w];
        y += s_red[2 * w + 1];
    }
}

// Projector onto the nocc lowest levels of the symmetric A (n x n, LDS, left
// unchanged) into P by trace-correcting purification (Niklasson, PRB 66,
// 155115 (2002)): P0 = (hi I - A)/(hi - lo) from Gershgorin bounds, then
// P <- P^2 or 2P - P^2, whichever trace lands nearer nocc (tr P^2 = ||P||_F^2
// for symmetric P).  Two steps after P is idempotent to 1e-8 with trace nocc
// to 1e-6 it is final.  T: n x n LDS scratch.  False when that does not
// happen within 100 steps (near-degenerate frontier levels).
__device__ bool purify_lds(const double *A, double *P, double *T, int n, int nocc, double *s_red) {
    double lo = 1e300, hi = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double r = 0.0;
        for (int j = 0; j < n; ++j)
            if (j != i) r += fabs(A[i + n * j]);
        const double c = A[i * (n + 1)];
        lo = fmin(lo, c - r);
        hi = fmax(hi, c + r);
    }
    lo = -wg_max(-lo, s_red);
    hi = wg_max(hi, s_red);
    const double sc = 1.0 / fmax(hi - lo, 1e-300);
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) P[e] = ((e % n == e / n ? hi : 0.0) - A[e]) * sc;
    int cs = -1;
    for (int step = 0; step < 100; ++step) {
        double t = 0.0, f = 0.0;
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
            const double v = P[e];
            f += v * v;
            if (e % n == e / n) t += v;
        }
        wg_sum2(t, f, s_red);   // also orders the previous step's writes of P
        if (cs < 0 && step >= 8 && fabs(t - f) < 1e-8 && fabs(t - nocc) < 1e-6) cs = step;
        if (cs >= 0 && step >= cs + 2) return true;
        const bool square = fabs(f - nocc) < fabs(2.0 * t - f - nocc);
        mm(T, P, P, n, false, 1.0);
        __syncthreads();
        for (int e = threadIdx.x; e < n * n; e += blockDim.x) P[e] = square ? T[e] : 2.0 * P[e] - T[e];
    }
    __syncthreads();
    return false;
}

// indices of the nocc lowest eigenvalues (diagonal of A) into s_occ; thread 0
__device__ void select_occ(const double *A, int n, int nocc, int *s_occ, int *s_tmp) {
    if (threadIdx.x == 0) {
        for (int i = 0; i < n; ++i) s_tmp[i] = i;
        for (int a = 0; a < nocc; ++a) {
            int best = a;
            for (int b = a + 1; b < n; ++b)
                if (A[s_tmp[b] * (n + 1)] < A[s_tmp[best] * (n + 1)]) best = b;
            const int t = s_tmp[a];
            s_tmp[a] = s_tmp[best];
            s_tmp[best] = t;
            s_occ[a] = s_tmp[a];
        }
    }
    __syncthreads();
}

// P = U_occ U_occ^T
__device__ __forceinline__ void occ_projector(double *P, const double *U, const int *s_occ, int n, int nocc) {
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e % n, j = e / n;
        double s = 0.0;
        for (int a = 0; a < nocc; ++a) s += U[i + n * s_occ[a]] * U[j + n * s_occ[a]];
        P[e] = s;
    }
}

// DIIS coefficients from the (m+1)^2 system on thread 0, worked in the LDS
// array M ((kBatchDiis + 1) (kBatchDiis + 2) doubles); false if singular
__device__ bool diis_solve(const double *B, int MD, int m, double *M, double *c) {
    const int N = m + 1, W = N + 1;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < W; ++j) M[i * W + j] = 0.0;
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < m; ++j) M[i * W + j] = B[i * MD + j];
        M[i * W + m] = -1.0;
    }
    for (int j = 0; j < m; ++j) M[m * W + j] = -1.0;
    M[m * W + N] = -1.0;
    for (int col = 0; col < N; ++col) {
        int piv = col;
        for (int r = col + 1; r < N; ++r)
            if (fabs(M[r * W + col]) > fabs(M[piv * W + col])) piv = r;
        if (fabs(M[piv * W + col]) < 1e-300) return false;
        if (piv != col)
            for (int k = 0; k < W; ++k) {
                const double t = M[col * W + k];
                M[col * W + k] = M[piv * W + k];
                M[piv * W + k] = t;
            }
        for (int r = 0; r < N; ++r) {
            if (r == col) continue;
            const double f = M[r * W + col] / M[col * W + col];
            for (int k = col; k < W; ++k) M[r * W + k] -= f * M[col * W + k];
        }
    }
    for (int i = 0; i < m; ++i) c[i] = M[i * W + N] / M[i * W + i];
    return true;
}

// G(D) = A + A^T accumulated from the packed unique integrals (degeneracy
// weights 1/2 per index coincidence); F = H + G.  A wavefront takes one bra
// pair IJ at a time and each lane four consecutive kets KL <= IJ (a
// contiguous run of the packed array), loaded a step ahead.  The J(ij)
// term is reduced across the wavefront; the K terms on (i,k) and (j,k), whose
// k repeats along a run of kets, are summed in registers while k stays the
// same; the other three go to LDS per value.  Aacc and the copy Dl of D have
// an odd leading dimension, so the updates and reads of a column (same i,
// consecutive l) fall in distinct LDS banks.
__device__ void fock(const double *__restrict__ eri, const double *__restrict__ H, const double *D, double *Aacc,
                     double *Dl, double *F, int n, const unsigned char *s_pi, const unsigned char *s_pj) {
    const int la = n | 1;
    for (int e = threadIdx.x; e < la * n; e += blockDim.x) {
        const int i = e % la, j = e / la;
        Aacc[e] = 0.0;
        Dl[e] = i < n ? D[i + n * j] : 0.0;
    }
    __syncthreads();
    const int P = (int)tri(n);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    constexpr int U = 4, STEP = 64 * U;
    auto advance = [&](int &ij, int &kl) {
        kl += STEP;
        if (kl > ij) {
            ij += nwave;
            kl = 0;
        }
    };
    auto load = [&](int ij, int kl, double (&v)[U]) {
        const int x = kl + U * lane;
        const double *row = eri + tri(ij);
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = (ij < P && x + u <= ij) ? row[x + u] : 0.0;
    };
    int pij = wave, pkl = 0;
    double v0[U], v2[U];
    load(pij, pkl, v0);
    advance(pij, pkl);
    int IJ = wave, KL0 = 0;
    double jr = 0.0;
    while (IJ < P) {
#pragma unroll
        for (int u = 0; u < U; ++u) v2[u] = v0[u];
        load(pij, pkl, v0);
        advance(pij, pkl);
        const int i = s_pi[IJ], j = s_pj[IJ];
        const double dij = Dl[i + la * j];
        const double wij = (i == j) ? 0.5 : 1.0;
        int kc = -1;
        double aik = 0.0, ajk = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int KL = KL0 + U * lane + u;
            if (KL <= IJ) {
                const int k = s_pi[KL], l = s_pj[KL];
                double v = v2[u] * wij;
                if (k == l) v *= 0.5;
                if (IJ == KL) v *= 0.5;
                if (k != kc) {
                    if (kc >= 0) {
                        atomicAdd(&Aacc[i + la * kc], aik);
                        atomicAdd(&Aacc[j + la * kc], ajk);
                    }
                    kc = k;
                    aik = ajk = 0.0;
                }
                jr += 2.0 * Dl[k + la * l] * v;
                aik -= 0.5 * Dl[j + la * l] * v;
                ajk -= 0.5 * Dl[i + la * l] * v;
                atomicAdd(&Aacc[k + la * l], 2.0 * dij * v);
                atomicAdd(&Aacc[i + la * l], -0.5 * Dl[j + la * k] * v);
                atomicAdd(&Aacc[j + la * l], -0.5 * Dl[i + la * k] * v);
            }
        }
        if (kc >= 0) {
            atomicAdd(&Aacc[i + la * kc], aik);
            atomicAdd(&Aacc[j + la * kc], ajk);
        }
        KL0 += STEP;
        if (KL0 > IJ) {   // row done (uniform across the wavefront)
            jr = wave_red_sum(jr);
            if (lane == 0) atomicAdd(&Aacc[i + la * j], jr);
            jr = 0.0;
            IJ += nwave;
            KL0 = 0;
        }
    }
    __syncthreads();
    for (int x = threadIdx.x; x < n * n; x += blockDim.x) {
        const int i = x % n, j = x / n;
        F[x] = H[x] + Aacc[i + la * j] + Aacc[j + la * i];
    }
    __syncthreads();
}

// LDS: L0 (G accumulator / Jacobi matrix) and L2 (padded D in the Fock build
// / U / purification), (nmax | 1) nmax doubles each; L1 (D), L3 (temp), nmax^2
// doubles each; then reductions, rotations, DIIS matrix and index tables.
__host__ __device__ inline size_t scf_lds_bytes(int nmax, int threads) {
    const size_t n2 = (size_t)nmax * nmax;
    return sizeof(double) * (4 * n2 + 2 * nmax + threads + 2 * (nmax + 2) + kBatchDiis * kBatchDiis + kBatchDiis + 8 +
                             (kBatchDiis + 1) * (kBatchDiis + 2)) +
           sizeof(int) * 2 * nmax + 2 * (size_t)tri(nmax) + 16;
}

// (at most 128 VGPRs: two 512-thread workgroups per CU)
__global__ __launch_bounds__(512, 4) void k_batch_scf(BScfArgs a, int nmax) {
    extern __shared__ __attribute__((aligned(16))) double smem[];
    const size_t n2max = (size_t)nmax * nmax;
    // L0 and L2: (nmax | 1) x nmax (the Fock build's padded accumulator and D)
    double *L0 = smem, *L1 = L0 + n2max + nmax, *L2 = L1 + n2max, *L3 = L2 + n2max + nmax;
    double *s_red = L3 + n2max;
    double *s_rot = s_red + blockDim.x;                  // 2 (nmax + 2)
    double *s_B = s_rot + 2 * (nmax + 2);                // kBatchDiis^2
    double *s_c = s_B + kBatchDiis * kBatchDiis;         // kBatchDiis
    double *s_misc = s_c + kBatchDiis;                   // 8
    double *s_M = s_misc + 8;                            // DIIS system
    int *s_occ = reinterpret_cast<int *>(s_M + (kBatchDiis + 1) * (kBatchDiis + 2));
    int *s_tmp = s_occ + nmax;
    unsigned char *s_pi = reinterpret_cast<unsigned char *>(s_tmp + nmax);
    unsigned char *s_pj = s_pi + tri(nmax);
    int *s_job = reinterpret_cast<int *>(s_misc + 7);
    const int MD = kBatchDiis;
    while (true) {
        if (threadIdx.x == 0) *s_job = atomicAdd(a.counter, 1);
        __syncthreads();
        const int job = *s_job;
        __syncthreads();
        if (job >= a.nmol) break;
        const int m = a.order[job];
        const BMol mol = a.mols[m];
        const int n = mol.n, nocc = mol.nocc;
        const size_t nn = (size_t)n * n;
        const double *Sg = a.S + mol.mat0, *Hg = a.H + mol.mat0, *eri = a.eri + mol.eri0;
        double *X = a.scratch + 4 * mol.mat0, *Fp = X + nn, *Dp = Fp + nn, *W = Dp + nn;
        double *hF = a.hist + 2 * MD * mol.mat0, *hE = hF + MD * nn;
        for (int x = threadIdx.x; x < tri(n); x += blockDim.x) {
            int i = (int)floor((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
            while (tri(i) > x) --i;
            while (tri(i + 1) <= x) ++i;
            s_pi[x] = (unsigned char)i;
            s_pj[x] = (unsigned char)(x - tri(i));
        }
        // X = S^{-1/2} by the eigen-decomposition of S
        for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) L0[e] = Sg[e];
        __syncthreads();
        jacobi(L0, L2, n, s_rot, s_red);
        double smin = 1e300;
        for (int i = threadIdx.x; i < n; i += blockDim.x) smin = fmin(smin, L0[i * (n + 1)]);
        smin = -wg_max(-smin, s_red);
        if (smin <= 1e-10) {
            if (threadIdx.x == 0) {
                a.energy[m] = 0.0;
                a.iters[m] = 0;
                a.conv[m] = -1;   // overlap numerically singular
                if (a.fallback) a.fallback[m] = 0;
            }
            __syncthreads();
            continue;
        }
        for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) L3[e] = L2[e] / sqrt(L0[(e / n) * (n + 1)]);
        __syncthreads();
        mm(X, L3, L2, n, true, 1.0);
        __syncthreads();
        // a start from the given density gets learned_iter iterations; if it
        // does not converge, the molecule restarts from the core guess in the
        // same workgroup with the full budget
        int it_total = 0, conv = 0, it = 0;
        double E = 0.0;
        for (int attempt = a.guess ? 0 : 1; attempt < 2; ++attempt) {
        const bool from_guess = attempt == 0;
        const int maxit = from_guess ? a.learned_iter : a.max_iter;
        if (from_guess) {
            // Dp = S^{1/2} D S^{1/2} / 2
            for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) L3[e] = L2[e] * sqrt(L0[(e / n) * (n + 1)]);
            __syncthreads();
            mm(W, L3, L2, n, true, 1.0);              // S^{1/2}
            for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) L1[e] = a.guess[mol.mat0 + e];
            __syncthreads();
            mm(L3, W, L1, n, false, 1.0);
            __syncthreads();
            mm(Dp, L3, W, n, false, 0.5);
            __syncthreads();
        } else {
            // core Hamiltonian guess (scf.f90:410-512)
            mm(L3, Hg, X, n, false, 1.0);
            __syncthreads();
            mm(L0, X, L3, n, false, 1.0);
            __syncthreads();
            jacobi(L0, L2, n, s_rot, s_red);
            select_occ(L0, n, nocc, s_occ, s_tmp);
            occ_projector(Dp, L2, s_occ, n, nocc);
            __syncthreads();
            mm(L3, Dp, X, n, false, 1.0);
            __syncthreads();
            mm(L1, X, L3, n, false, 2.0);
            __syncthreads();
        }
        int hist = 0, head = 0;
        double E_old = 0.0, dE_prev = 1e300, err_prev = 1e300;
        bool abandoned = false;
        E = 0.0;
        for (it = 1; it <= maxit + 1; ++it) {
            // F = H + G(D) into W; E = (D.F + D.H)/2 + E_nuc
            fock(eri, Hg, L1, L0, L2, W, n, s_pi, s_pj);
            double ed = 0.0;
            for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) ed += L1[e] * (W[e] + Hg[e]);
            E = 0.5 * wg_sum(ed, s_red) + mol.e_nuc;
            // the previous iteration's density change, then its convergence
            if (it > 1) {
                double dd = 0.0;
                for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) dd = fmax(dd, fabs(L1[e] - L3[e]));
                dd = wg_max(dd, s_red);
                if (it > 2 && dE_prev < a.e_tol && dd < a.d_tol && err_prev < 100.0 * a.d_tol) {
                    conv = 1;
                    break;
                }
            }
            if (it > maxit) break;
            // F' = X F X into the history; error F' Dp - Dp F'
            double *fs = hF + head * nn, *es = hE + head * nn;
            mm(L3, W, X, n, false, 1.0);
            __syncthreads();
            mm(fs, X, L3, n, false, 1.0);
            __syncthreads();
            mm(L3, fs, Dp, n, false, 1.0);
            __syncthreads();
            const int hn = min(hist + 1, MD);
            double dots[kBatchDiis];
#pragma unroll
            for (int q = 0; q < kBatchDiis; ++q) dots[q] = 0.0;
            for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) {
                const int i = e % n, j = e / n;
                const double ev = L3[e] - L3[j + n * i];
                es[e] = ev;
#pragma unroll
                for (int q = 0; q < kBatchDiis; ++q) {
                    const double hv = q == head ? ev : (q < hn ? hE[q * nn + e] : 0.0);
                    dots[q] += ev * hv;
                }
            }
#pragma unroll
            for (int q = 0; q < kBatchDiis; ++q) {
                if (q < hn) {   // uniform across the workgroup
                    const double v = wg_sum(dots[q], s_red);
                    if (threadIdx.x == 0) {
                        s_B[head * MD + q] = v;
                        s_B[q * MD + head] = v;
                    }
                }
            }
            __syncthreads();
            hist = hn;
            // DIIS extrapolation of F' into L0
            if (threadIdx.x == 0) {
                const bool ok = hist >= 2 && diis_solve(s_B, MD, hist, s_M, s_c);
                if (!ok) {
                    for (int q = 0; q < MD; ++q) s_c[q] = 0.0;
                    s_c[head] = 1.0;
                }
                s_misc[0] = ok ? 1.0 : 0.0;
            }
            __syncthreads();
            for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) {
                double v = 0.0;
                for (int q = 0; q < hist; ++q) v += s_c[q] * hF[q * nn + e];
                L0[e] = v;
            }
            err_prev = sqrt(fabs(s_B[head * MD + head]));
            dE_prev = fabs(E - E_old);
            E_old = E;
            head = (head + 1) % MD;
            __syncthreads();
            // a start from the guess that is far from converging by
            // check_it iterations restarts now rather than at learned_iter
            if (from_guess && a.restart && it == a.check_it && err_prev > a.check_err) {
                abandoned = true;
                break;
            }
            // new orthonormal density: the projector onto the nocc lowest
            // levels of F' by purification, or by diagonalisation when the
            // frontier levels are too close for it
            if (purify_lds(L0, L2, L3, n, nocc, s_red)) {
                for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) Dp[e] = L2[e];
            } else {
                jacobi(L0, L2, n, s_rot, s_red);
                select_occ(L0, n, nocc, s_occ, s_tmp);
                occ_projector(Dp, L2, s_occ, n, nocc);
            }
            __syncthreads();
            // keep D_k in L3 for the change; D_{k+1} = 2 X Dp X into L1
            for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) L3[e] = L1[e];
            mm(W, Dp, X, n, false, 1.0);
            __syncthreads();
            mm(L1, X, W, n, false, 2.0);
            __syncthreads();
        }
        it_total += abandoned ? it : min(it - 1, maxit);
        if (conv || !from_guess || !a.restart) {
            if (threadIdx.x == 0 && a.fallback) a.fallback[m] = from_guess ? 0 : (a.guess ? 1 : 0);
            break;
        }
        }   // attempts
        for (int e = threadIdx.x; e < (int)nn; e += blockDim.x) a.Dout[mol.mat0 + e] = L1[e];
        if (threadIdx.x == 0) {
            a.energy[m] = E;
            a.iters[m] = it_total;
            a.conv[m] = conv;
        }
        __syncthreads();
    }
}

__global__ void k_zero_int(int *p, int cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) p[i] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct BatchState {
    int nmol = 0, nmax = 0, npairs = 0, natom = 0;
    long long total_eri = 0, total_mat = 0, nq[6] = {0, 0, 0, 0, 0, 0};
    std::vector<BMol> mols;
    std::vector<int> order;
    BMol *d_mols = nullptr;
    int *d_order = nullptr, *d_pair_mol = nullptr, *d_cls[3] = {nullptr, nullptr, nullptr};
    int ncls[3] = {0, 0, 0};
    double *d_ksoa[3] = {nullptr, nullptr, nullptr};   // per class list (ket_soa.h)
    ShellPair *d_pairs = nullptr;
    PrimPair *d_prims = nullptr;
    double *d_atoms = nullptr, *d_ftT = nullptr;
    long long *d_qcum[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *d_eri = nullptr, *d_S = nullptr, *d_H = nullptr, *d_D = nullptr, *d_guess = nullptr, *d_scratch = nullptr,
           *d_hist = nullptr, *d_energy = nullptr;
    int *d_iters = nullptr, *d_conv = nullptr, *d_counter = nullptr;
    // atom pairs of the learned guess
    bool guess_ok = false;
    int npair = 0;
    PairRec *d_pairrec = nullptr;
    int *d_mol_atoms = nullptr, *d_fail = nullptr;
    std::vector<int> pair_mol_of;   // molecule of each pair record
    hipGraphExec_t graph_learned = nullptr;
    double gl_etol = 0, gl_dtol = 0;
    int gl_maxit = 0, gl_gen = -1;
    // launch shapes (autotuned) and the captured graph
    int eri_threads[6] = {256, 256, 256, 256, 256, 256};
    int scf_threads = 256;
    bool tuned = false;
    hipGraphExec_t graph = nullptr;
    bool graph_guess = false;
    double g_etol = 0, g_dtol = 0;
    int g_maxit = 0;
    int replays = 0;
    void release() {
        void **ps[] = {(void **)&d_mols,   (void **)&d_order,  (void **)&d_pair_mol, (void **)&d_cls[0],
                       (void **)&d_cls[1], (void **)&d_cls[2], (void **)&d_pairs,    (void **)&d_prims,
                       (void **)&d_atoms,  (void **)&d_ftT,    (void **)&d_eri,      (void **)&d_S,
                       (void **)&d_H,      (void **)&d_D,      (void **)&d_guess,    (void **)&d_scratch,
                       (void **)&d_hist,   (void **)&d_energy, (void **)&d_iters,    (void **)&d_conv,
                       (void **)&d_counter, (void **)&d_pairrec, (void **)&d_mol_atoms, (void **)&d_fail};
        for (auto p : ps) dfree(p);
        if (graph_learned) (void)hipGraphExecDestroy(graph_learned);
        graph_learned = nullptr;
        for (auto &q : d_qcum) dfree((void **)&q);
        for (auto &q : d_ksoa) dfree((void **)&q);
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        nmol = 0;
        tuned = false;
    }
};

void free_batch(myqc_ctx *ctx) {
    if (ctx->batch) {
        ctx->batch->release();
        delete ctx->batch;
        ctx->batch = nullptr;
    }
}

static const int kCB[6] = {0, 1, 1, 2, 2, 2}, kCK[6] = {0, 0, 1, 0, 1, 2};

static void launch_batch_eri(myqc_ctx *ctx, BatchState &b, int c, int threads) {
    BEriArgs a;
    a.pairs = b.d_pairs;
    a.prims = b.d_prims;
    a.ftT = b.d_ftT;
    a.mols = b.d_mols;
    a.cls0 = b.d_cls[0];
    a.cls1 = b.d_cls[1];
    a.cls2 = b.d_cls[2];
    a.ksoa = b.d_ksoa[kCK[c]];
    a.nket = b.ncls[kCK[c]];
    a.qcum = b.d_qcum[c];
    a.nmol = b.nmol;
    a.nq = b.nq[c];
    a.eri = b.d_eri;
    const unsigned grid = (unsigned)((b.nq[c] + threads - 1) / threads);
    switch (c) {
        case 0: k_batch_eri<0, 0><<<grid, threads, 0, ctx->stream>>>(a); break;
        case 1: k_batch_eri<1, 0><<<grid, threads, 0, ctx->stream>>>(a); break;
        case 2: k_batch_eri<1, 1><<<grid, threads, 0, ctx->stream>>>(a); break;
        case 3: k_batch_eri<2, 0><<<grid, threads, 0, ctx->stream>>>(a); break;
        case 4: k_batch_eri<2, 1><<<grid, threads, 0, ctx->stream>>>(a); break;
        default: k_batch_eri<2, 2><<<grid, threads, 0, ctx->stream>>>(a); break;
    }
}

static int scf_grid(myqc_ctx *ctx, const BatchState &b) {
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
    return std::max(1, std::min(b.nmol, 4 * cus));
}

// A start from the guess whose DIIS error is still above kCheckErr after
// kCheckIt iterations restarts from the core guess there (not at
// learned_iter).  Measured on the seeded 4096 batch: 25 / 1e-3 keeps 99.05 %
// converged (20 / 1e-3: 98.9 %, but 16 % faster).
constexpr int kCheckIt = 25;
constexpr double kCheckErr = 1e-3;

// restart: a start from the guess that does not converge within
// learned_iter iterations restarts from the core guess (same launch)
static void launch_batch_scf(myqc_ctx *ctx, BatchState &b, bool guess, double e_tol, double d_tol, int max_iter,
                             int threads, const int *order = nullptr, int count = -1, bool restart = false,
                             int learned_iter = 0) {
    BScfArgs a;
    a.mols = b.d_mols;
    a.order = order ? order : b.d_order;
    a.nmol = count >= 0 ? count : b.nmol;
    a.counter = b.d_counter;
    a.eri = b.d_eri;
    a.S = b.d_S;
    a.H = b.d_H;
    a.guess = guess ? b.d_guess : nullptr;
    a.scratch = b.d_scratch;
    a.hist = b.d_hist;
    a.Dout = b.d_D;
    a.energy = b.d_energy;
    a.iters = b.d_iters;
    a.conv = b.d_conv;
    a.e_tol = e_tol;
    a.d_tol = d_tol;
    a.max_iter = max_iter;
    a.restart = (guess && restart) ? 1 : 0;
    a.learned_iter = a.restart ? std::min(learned_iter, max_iter) : max_iter;
    a.fallback = b.d_fail;
    a.check_it = kCheckIt;
    a.check_err = kCheckErr;
    k_zero_int<<<1, 64, 0, ctx->stream>>>(b.d_counter, 1);
    const int grid = std::max(1, std::min(scf_grid(ctx, b), a.nmol));
    k_batch_scf<<<grid, threads, scf_lds_bytes(b.nmax, threads), ctx->stream>>>(a, b.nmax);
}

// the whole pipeline: integrals, then the SCF
static void launch_batch(myqc_ctx *ctx, BatchState &b, bool guess, double e_tol, double d_tol, int max_iter) {
    k_batch_one_e<<<(b.npairs + 127) / 128, 128, 0, ctx->stream>>>(b.d_pairs, b.d_prims, b.d_ftT, b.d_atoms,
                                                                   b.d_pair_mol, b.d_mols, b.npairs, b.d_S, b.d_H);
    for (int c = 0; c < 6; ++c)
        if (b.nq[c] > 0) launch_batch_eri(ctx, b, c, b.eri_threads[c]);
    launch_batch_scf(ctx, b, guess, e_tol, d_tol, max_iter, b.scf_threads);
}

// Autotuner: times each candidate launch shape on this batch with HIP events
// and keeps the fastest (ERI: 64/128/256 threads per class; SCF: 64/128/256/512
// threads per molecule's workgroup).
static int batch_tune(myqc_ctx *ctx, BatchState &b, bool guess, double e_tol, double d_tol, int max_iter) {
    hipEvent_t e0, e1;
    MYQC_HIP(ctx, hipEventCreate(&e0));
    MYQC_HIP(ctx, hipEventCreate(&e1));
    auto timed = [&](auto &&launch) -> float {
        float best = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            (void)hipEventRecord(e0, ctx->stream);
            launch();
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            best = std::min(best, ms);
        }
        return best;
    };
    for (int c = 0; c < 6; ++c) {
        if (b.nq[c] == 0) continue;
        float best = 1e30f;
        for (int th : {64, 128, 256}) {
            const float ms = timed([&] { launch_batch_eri(ctx, b, c, th); });
            if (ms < best) {
                best = ms;
                b.eri_threads[c] = th;
            }
        }
    }
    k_batch_one_e<<<(b.npairs + 127) / 128, 128, 0, ctx->stream>>>(b.d_pairs, b.d_prims, b.d_ftT, b.d_atoms,
                                                                   b.d_pair_mol, b.d_mols, b.npairs, b.d_S, b.d_H);
    float best = 1e30f;
    for (int th : {64, 128, 256, 512}) {
        const float ms = timed([&] { launch_batch_scf(ctx, b, guess, e_tol, d_tol, max_iter, th); });
        if (ms < best) {
            best = ms;
            b.scf_threads = th;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    MYQC_HIP(ctx, hipGetLastError());
    b.tuned = true;
    return 0;
}

template <typename T>
static int upload(myqc_ctx *ctx, T **dst, const std::vector<T> &src) {
    MYQC_CHECK(dalloc(ctx, (void **)dst, sizeof(T) * std::max<size_t>(src.size(), 1)));
    if (!src.empty())
        MYQC_HIP(ctx, hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_batch_set(myqc_ctx *ctx, int nmol, const int *mol_nnuc, const int *mol_norb, const int *mol_nocc,
                              const int *Z, const double *xyz_bohr, const double *set_alpha, const int *setinfo,
                              const double *bas, const int *basinfo, const int *offsets, const double *Ftab) {
    if (!ctx || nmol <= 0 || !mol_nnuc || !mol_norb || !mol_nocc || !Z || !xyz_bohr || !set_alpha || !setinfo ||
        !bas || !basinfo || !offsets || !Ftab)
        return MYQC_E_ARG;
    free_batch(ctx);
    ctx->batch = new BatchState();
    BatchState &b = *ctx->batch;
    std::vector<ShellPair> pairs;
    std::vector<PrimPair> prims;
    std::vector<int> cls[3], pair_mol;
    std::vector<double> atoms;
    b.mols.resize(nmol);
    int atom0 = 0;
    long long eri0 = 0, mat0 = 0;
    for (int m = 0; m < nmol; ++m) {
        const int nnuc = mol_nnuc[m], norb = mol_norb[m];
        if (nnuc <= 0 || norb <= 0 || norb > kBatchMaxN || mol_nocc[m] < 1 || mol_nocc[m] > norb)
            return set_error(ctx, MYQC_E_ARG, "myqc_batch_set: molecule %d: %d atoms, %d functions, %d occupied "
                             "(1..%d functions per molecule)", m, nnuc, norb, mol_nocc[m], kBatchMaxN);
        // this molecule's slices: offsets[4 m + {0: set_alpha, 1: setinfo, 2: bas, 3: basinfo}]
        const int *o = offsets + 4 * m;
        std::vector<double> xyz(3 * nnuc);
        for (int a = 0; a < nnuc; ++a)
            for (int c = 0; c < 3; ++c) xyz[3 * a + c] = xyz_bohr[3 * atom0 + a + nnuc * c];
        std::vector<Shell> shells;
        std::vector<ShellPair> mp;
        std::vector<PrimPair> mpr;
        std::vector<int> ids[3];
        const int *si = setinfo + o[1];
        const std::string why = build_system(nnuc, xyz.data(), si[0], set_alpha + o[0], si, bas + o[2], norb,
                                             basinfo + o[3], shells, mp, mpr, ids);
        if (!why.empty()) return set_error(ctx, MYQC_E_ARG, "myqc_batch_set: molecule %d: %s", m, why.c_str());
        BMol &mol = b.mols[m];
        mol.n = norb;
        mol.nocc = mol_nocc[m];
        mol.nnuc = nnuc;
        mol.atom0 = atom0;
        mol.eri0 = eri0;
        mol.mat0 = mat0;
        double en = 0.0;
        for (int a = 0; a < nnuc; ++a) {
            atoms.push_back(Z[atom0 + a]);
            for (int c = 0; c < 3; ++c) atoms.push_back(xyz[3 * a + c]);
            for (int c = 0; c < a; ++c) {
                const double dx = xyz[3 * a] - xyz[3 * c], dy = xyz[3 * a + 1] - xyz[3 * c + 1],
                             dz = xyz[3 * a + 2] - xyz[3 * c + 2];
                en += (double)Z[atom0 + a] * Z[atom0 + c] / sqrt(dx * dx + dy * dy + dz * dz);
            }
        }
        mol.e_nuc = en;
        const int p0 = (int)pairs.size(), q0 = (int)prims.size();
        for (auto sp : mp) {
            sp.pp0 += q0;
            pairs.push_back(sp);
            pair_mol.push_back(m);
        }
        prims.insert(prims.end(), mpr.begin(), mpr.end());
        for (int c = 0; c < 3; ++c) {
            mol.cls0[c] = (int)cls[c].size();
            mol.ncls[c] = (int)ids[c].size();
            for (int id : ids[c]) cls[c].push_back(p0 + id);
        }
        atom0 += nnuc;
        eri0 += tri(tri(norb));
        mat0 += (long long)norb * norb;
        b.nmax = std::max(b.nmax, norb);
    }
    // atom pairs A <= B for the learned guess (H, He: 1 function; Li..Ne: 5,
    // atoms' functions contiguous in atom order)
    std::vector<PairRec> prs;
    std::vector<int> mol_atoms(2 * nmol);
    b.guess_ok = true;
    for (int m = 0; m < nmol; ++m) {
        const BMol &mol = b.mols[m];
        mol_atoms[2 * m] = mol.atom0;
        mol_atoms[2 * m + 1] = mol.nnuc;
        std::vector<int> ao(mol.nnuc), nf(mol.nnuc);
        int a0 = 0;
        for (int a = 0; a < mol.nnuc; ++a) {
            const int z = Z[mol.atom0 + a];
            if (z < 1 || z > 10) b.guess_ok = false;
            nf[a] = z <= 2 ? 1 : 5;
            ao[a] = a0;
            a0 += nf[a];
        }
        if (a0 != mol.n) b.guess_ok = false;
        for (int a = 0; a < mol.nnuc; ++a)
            for (int c = a; c < mol.nnuc; ++c) {
                PairRec r;
                r.mol = m;
                r.a = mol.atom0 + a;
                r.b = mol.atom0 + c;
                r.aoA = ao[a];
                r.aoB = ao[c];
                r.nA = nf[a];
                r.nB = nf[c];
                r.n = mol.n;
                r.mat0 = mol.mat0;
                prs.push_back(r);
                b.pair_mol_of.push_back(m);
            }
    }
    b.npair = (int)prs.size();
    MYQC_CHECK(upload(ctx, &b.d_pairrec, prs));
    MYQC_CHECK(upload(ctx, &b.d_mol_atoms, mol_atoms));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_fail, sizeof(int) * nmol));
    b.nmol = nmol;
    b.npairs = (int)pairs.size();
    b.total_eri = eri0;
    b.total_mat = mat0;
    b.order.resize(nmol);
    std::iota(b.order.begin(), b.order.end(), 0);
    std::stable_sort(b.order.begin(), b.order.end(), [&](int x, int y) { return b.mols[x].n > b.mols[y].n; });
    for (int c = 0; c < 6; ++c) {
        std::vector<long long> qc(nmol + 1, 0);
        for (int m = 0; m < nmol; ++m) {
            const long long nb = b.mols[m].ncls[kCB[c]], nk = b.mols[m].ncls[kCK[c]];
            qc[m + 1] = qc[m] + (kCB[c] == kCK[c] ? tri(nb) : nb * nk);
        }
        b.nq[c] = qc[nmol];
        MYQC_CHECK(upload(ctx, &b.d_qcum[c], qc));
    }
    const std::vector<double> ftT = build_ftT(Ftab);
    MYQC_CHECK(upload(ctx, &b.d_mols, b.mols));
    MYQC_CHECK(upload(ctx, &b.d_order, b.order));
    MYQC_CHECK(upload(ctx, &b.d_pair_mol, pair_mol));
    for (int c = 0; c < 3; ++c) MYQC_CHECK(upload(ctx, &b.d_cls[c], cls[c]));
    MYQC_CHECK(upload(ctx, &b.d_pairs, pairs));
    MYQC_CHECK(upload(ctx, &b.d_prims, prims));
    for (const ShellPair &sp : pairs)
        if (sp.npp > kKetPrims)
            return set_error(ctx, MYQC_E_ARG, "myqc_batch_set: a shell pair has %d primitive pairs (at most %d)", sp.npp,
                             kKetPrims);
    for (int c = 0; c < 3; ++c) {
        b.ncls[c] = (int)cls[c].size();
        if (b.ncls[c] == 0) continue;
        MYQC_CHECK(dalloc(ctx, (void **)&b.d_ksoa[c], sizeof(double) * kKetPrims * kKetFields * (size_t)b.ncls[c]));
        k_ket_soa<<<dim3((b.ncls[c] + 255) / 256, kKetPrims), 256, 0, ctx->stream>>>(b.d_pairs, b.d_prims, b.d_cls[c],
                                                                                  b.ncls[c], b.d_ksoa[c]);
        MYQC_HIP(ctx, hipGetLastError());
    }
    MYQC_CHECK(upload(ctx, &b.d_atoms, atoms));
    b.natom = (int)(atoms.size() / 4);
    MYQC_CHECK(upload(ctx, &b.d_ftT, ftT));
    const size_t M = (size_t)std::max<long long>(b.total_mat, 1);
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_eri, sizeof(double) * std::max<long long>(b.total_eri, 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_S, sizeof(double) * M));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_H, sizeof(double) * M));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_D, sizeof(double) * M));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_guess, sizeof(double) * M));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_scratch, sizeof(double) * 4 * M));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_hist, sizeof(double) * 2 * kBatchDiis * M));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_energy, sizeof(double) * nmol));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_iters, sizeof(int) * nmol));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_conv, sizeof(int) * nmol));
    MYQC_CHECK(dalloc(ctx, (void **)&b.d_counter, sizeof(int) * 64));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int lds_limit = 65536;
    (void)hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
    if (scf_lds_bytes(b.nmax, 512) > (size_t)lds_limit)
        return set_error(ctx, MYQC_E_ARG, "myqc_batch_set: n=%d needs %zu B of LDS per workgroup (limit %d)", b.nmax,
                         scf_lds_bytes(b.nmax, 512), lds_limit);
    return 0;
}

extern "C" int myqc_batch_rhf(myqc_ctx *ctx, const double *D_guess, double e_tol, double d_tol, int max_iter,
                              double *energy, int *iterations, int *converged) {
    if (!ctx || !energy || e_tol <= 0 || d_tol <= 0 || max_iter < 1) return MYQC_E_ARG;
    if (!ctx->batch || ctx->batch->nmol == 0) return set_error(ctx, MYQC_E_STATE, "myqc_batch_rhf: no batch set");
    BatchState &b = *ctx->batch;
    const bool guess = D_guess != nullptr;
    if (guess)
        MYQC_HIP(ctx, hipMemcpyAsync(b.d_guess, D_guess, sizeof(double) * b.total_mat, hipMemcpyHostToDevice,
                                     ctx->stream));
    if (!b.tuned) MYQC_CHECK(batch_tune(ctx, b, guess, e_tol, d_tol, max_iter));
    if (!b.graph || b.graph_guess != guess || b.g_etol != e_tol || b.g_dtol != d_tol || b.g_maxit != max_iter) {
        if (b.graph) (void)hipGraphExecDestroy(b.graph);
        b.graph = nullptr;
        hipGraph_t g;
        MYQC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        launch_batch(ctx, b, guess, e_tol, d_tol, max_iter);
        MYQC_HIP(ctx, hipStreamEndCapture(ctx->stream, &g));
        hipError_t e = hipGraphInstantiate(&b.graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b.graph = nullptr;
            return set_error(ctx, MYQC_E_HIP, "myqc_batch_rhf: graph instantiation failed: %s", hipGetErrorString(e));
        }
        b.graph_guess = guess;
        b.g_etol = e_tol;
        b.g_dtol = d_tol;
        b.g_maxit = max_iter;
    }
    MYQC_HIP(ctx, hipGraphLaunch(b.graph, ctx->stream));
    ++b.replays;
    MYQC_HIP(ctx, d2h(ctx, energy, b.d_energy, sizeof(double) * b.nmol));
    if (iterations) MYQC_HIP(ctx, d2h(ctx, iterations, b.d_iters, sizeof(int) * b.nmol));
    std::vector<int> conv(b.nmol);
    MYQC_HIP(ctx, d2h(ctx, conv.data(), b.d_conv, sizeof(int) * b.nmol));
    if (converged) std::copy(conv.begin(), conv.end(), converged);
    return 0;
}

extern "C" int myqc_batch_density(myqc_ctx *ctx, double *D) {
    if (!ctx || !D) return MYQC_E_ARG;
    if (!ctx->batch || ctx->batch->nmol == 0) return set_error(ctx, MYQC_E_STATE, "myqc_batch_density: no batch set");
    MYQC_HIP(ctx, d2h(ctx, D, ctx->batch->d_D, sizeof(double) * ctx->batch->total_mat));
    return 0;
}

extern "C" int myqc_batch_info(myqc_ctx *ctx, int *eri_threads, int *scf_threads, int *graph_replays,
                               long long *n_integrals) {
    if (!ctx) return MYQC_E_ARG;
    if (!ctx->batch) return set_error(ctx, MYQC_E_STATE, "myqc_batch_info: no batch set");
    const BatchState &b = *ctx->batch;
    if (eri_threads)
        for (int c = 0; c < 6; ++c) eri_threads[c] = b.eri_threads[c];
    if (scf_threads) *scf_threads = b.scf_threads;
    if (graph_replays) *graph_replays = b.replays;
    if (n_integrals) *n_integrals = b.total_eri;
    return 0;
}

extern "C" int myqc_batch_pair_data(myqc_ctx *ctx, int *npair, float *features, float *targets, float *mask) {
    if (!ctx || !npair) return MYQC_E_ARG;
    if (!ctx->batch || ctx->batch->nmol == 0) return set_error(ctx, MYQC_E_STATE, "myqc_batch_pair_data: no batch set");
    BatchState &b = *ctx->batch;
    if (!b.guess_ok)
        return set_error(ctx, MYQC_E_ARG, "myqc_batch_pair_data: the learned guess covers H..Ne with STO-3G only");
    *npair = b.npair;
    if (!features && !targets && !mask) return 0;
    MYQC_CHECK(guess_export(ctx, b.d_pairrec, b.npair, b.d_atoms, b.d_mol_atoms, b.nmol, b.natom, b.d_D, features, targets, mask));
    if (mask) {
        // pairs of molecules that did not converge carry no target
        std::vector<int> conv(b.nmol);
        MYQC_HIP(ctx, d2h(ctx, conv.data(), b.d_conv, sizeof(int) * b.nmol));
        for (int p = 0; p < b.npair; ++p)
            if (conv[b.pair_mol_of[p]] != 1)
                for (int c = 0; c < kGuessOut; ++c) mask[(size_t)p * kGuessOut + c] = 0.f;
    }
    return 0;
}

extern "C" int myqc_batch_rhf_learned(myqc_ctx *ctx, double e_tol, double d_tol, int max_iter, double *energy,
                                      int *iterations, int *converged, int *fallback) {
    if (!ctx || !energy || e_tol <= 0 || d_tol <= 0 || max_iter < 1) return MYQC_E_ARG;
    if (!ctx->batch || ctx->batch->nmol == 0)
        return set_error(ctx, MYQC_E_STATE, "myqc_batch_rhf_learned: no batch set");
    BatchState &b = *ctx->batch;
    if (!b.guess_ok)
        return set_error(ctx, MYQC_E_ARG, "myqc_batch_rhf_learned: the learned guess covers H..Ne with STO-3G only");
    MYQC_CHECK(guess_reserve(ctx, b.npair, b.natom));
    if (!b.tuned) MYQC_CHECK(batch_tune(ctx, b, false, e_tol, d_tol, max_iter));
    // the learned start gets a shorter budget (typical solves take about 25
    // iterations); a molecule it leaves unconverged restarts from the core
    // guess inside the same persistent launch
    const int learned_iter = std::min(max_iter, 48);
    if (!b.graph_learned || b.gl_etol != e_tol || b.gl_dtol != d_tol || b.gl_maxit != max_iter ||
        b.gl_gen != ctx->mlp_gen) {
        if (b.graph_learned) (void)hipGraphExecDestroy(b.graph_learned);
        b.graph_learned = nullptr;
        hipGraph_t g;
        MYQC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        k_batch_one_e<<<(b.npairs + 127) / 128, 128, 0, ctx->stream>>>(b.d_pairs, b.d_prims, b.d_ftT, b.d_atoms,
                                                                       b.d_pair_mol, b.d_mols, b.npairs, b.d_S, b.d_H);
        for (int c = 0; c < 6; ++c)
            if (b.nq[c] > 0) launch_batch_eri(ctx, b, c, b.eri_threads[c]);
        guess_launch(ctx, b.d_pairrec, b.npair, b.d_atoms, b.d_mol_atoms, b.nmol, b.natom, b.d_guess);
        launch_batch_scf(ctx, b, true, e_tol, d_tol, max_iter, b.scf_threads, nullptr, -1, true, learned_iter);
        MYQC_HIP(ctx, hipStreamEndCapture(ctx->stream, &g));
        hipError_t e = hipGraphInstantiate(&b.graph_learned, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b.graph_learned = nullptr;
            return set_error(ctx, MYQC_E_HIP, "myqc_batch_rhf_learned: graph instantiation failed: %s",
                             hipGetErrorString(e));
        }
        b.gl_etol = e_tol;
        b.gl_dtol = d_tol;
        b.gl_maxit = max_iter;
        b.gl_gen = ctx->mlp_gen;
    }
    MYQC_HIP(ctx, hipGraphLaunch(b.graph_learned, ctx->stream));
    ++b.replays;
    MYQC_HIP(ctx, d2h(ctx, energy, b.d_energy, sizeof(double) * b.nmol));
    if (iterations) MYQC_HIP(ctx, d2h(ctx, iterations, b.d_iters, sizeof(int) * b.nmol));
    if (converged) MYQC_HIP(ctx, d2h(ctx, converged, b.d_conv, sizeof(int) * b.nmol));
    if (fallback) MYQC_HIP(ctx, d2h(ctx, fallback, b.d_fail, sizeof(int) * b.nmol));
    return 0;
}

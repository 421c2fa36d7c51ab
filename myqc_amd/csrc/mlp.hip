// mlp.hip -- the learned initial guess: an atom-pair MLP that predicts
// converged RHF/STO-3G density blocks from geometry, trained on the GPU by
// the kernels below on densities the batched GPU SCF produces.  No
// counterpart in the reference (its SCF starts from the core Hamiltonian,
// src/scf/scf.f90:410-512, which stays the fallback).
//
// Pairs and invariants.  For each atom pair A <= B of a batch molecule the
// density block D[AO(A), AO(B)] is described in the pair's bond frame
// (u = (R_B - R_A)/|R_B - R_A|), which makes it rotation invariant:
//   A != B: ss (1s/2s x 1s/2s, up to 4), s_A-p_B sigma (2), p_A-s_B sigma (2),
//           p-p sigma-sigma and pi-pi (2)                               -> 10
//   A == B: 1s1s, 1s2s, 2s2s, isotropic p-p                             -> 4
// and rebuilt from them as D[s,p_B] = sigma u, D[p_A,p_B] = pi I +
// (sigma - pi) u u^T, on-site D[p,p] = iso I, on-site s-p = 0.
//
// Features (kFeat = 200, FP32): one-hot Z_A and Z_B (H..Ne), 16 Gaussians of
// r_AB (centres 0..7.5 bohr), an on-site flag, and for each of A and B the
// Gaussian-smeared counts of neighbours of each element (8 centres 1..8
// bohr), zero-padded.
//
// Network: three hidden layers of `width` with SiLU, linear output; masked
// mean-squared error; Adam.  GEMMs are LDS-tiled FP32 FMA kernels (the
// layers are 200x128 .. 128x128: far too small for the matrix cores to matter).
#include <cmath>
#include <random>
#include <vector>

#include "guess.h"

namespace myqc {

constexpr int kFeat = kGuessFeat, kOut = kGuessOut, kElem = 10, kRbf = 16, kEnv = 8;

struct MlpState {
    int n_in = 0, width = 0, n_out = 0;
    long long n_par = 0;
    long long off[8];            // W1 b1 W2 b2 W3 b3 W4 b4
    float *d_par = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
    long long adam_t = 0;
    // activations for one batch of up to cap rows
    int cap = 0;
    float *d_X = nullptr, *d_T = nullptr, *d_M = nullptr, *d_Z[3] = {nullptr, nullptr, nullptr},
          *d_H[3] = {nullptr, nullptr, nullptr}, *d_Y = nullptr, *d_dA = nullptr, *d_dB = nullptr,
          *d_loss = nullptr;
    // per-atom environment features (kElem x kEnv each) of up to env_cap atoms
    float *d_env = nullptr;
    int env_cap = 0;
    void release() {
        float **ps[] = {&d_par, &d_grad, &d_m, &d_v, &d_X, &d_T, &d_M, &d_Z[0], &d_Z[1], &d_Z[2],
                        &d_H[0], &d_H[1], &d_H[2], &d_Y, &d_dA, &d_dB, &d_loss, &d_env};
        for (auto p : ps) dfree((void **)p);
        cap = 0;
        env_cap = 0;
    }
};

void free_mlp(myqc_ctx *ctx) {
    if (ctx->mlp) {
        ctx->mlp->release();
        delete ctx->mlp;
        ctx->mlp = nullptr;
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }

// C[M x N] = op(A) op(B) (+ bias[j]); row-major.  op(A)(i,k) = TA ? A[k lda + i]
// : A[i lda + k]; op(B)(k,j) = TB ? B[j ldb + k] : B[k ldb + j].  EPI 0: plain,
// 1: Z = C + bias, H = silu(Z), 2: C + bias (linear output).  64 x 64 tiles,
// each of 256 threads owning 4 x 4 outputs; K in steps of 16 through LDS.
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float *__restrict__ A, int lda,
                                               const float *__restrict__ B, int ldb, const float *__restrict__ bias,
                                               float *__restrict__ Z, float *__restrict__ C) {
    __shared__ float As[16][68], Bs[16][68];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const int kk = e & 15, r = e >> 4;
            const int gi = i0 + r, gk = k0 + kk, gj = j0 + r;
            As[kk][r] = (gi < M && gk < K) ? (TA ? A[(size_t)gk * lda + gi] : A[(size_t)gi * lda + gk]) : 0.f;
            Bs[kk][r] = (gj < N && gk < K) ? (TB ? B[(size_t)gj * ldb + gk] : B[(size_t)gk * ldb + gj]) : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[kk][ty * 4 + q];
                b[q] = Bs[kk][tx * 4 + q];
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) acc[x][y] = fmaf(a[x], b[y], acc[x][y]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            const int gi = i0 + ty * 4 + x, gj = j0 + tx * 4 + y;
            if (gi >= M || gj >= N) continue;
            float v = acc[x][y];
            if (EPI >= 1) v += bias[gj];
            if (EPI == 1) {
                Z[(size_t)gi * N + gj] = v;
                v = v * sigm(v);
            }
            C[(size_t)gi * N + gj] = v;
        }
}

// dZ = dH * silu'(Z)
__global__ void k_silu_grad(const float *__restrict__ Z, float *__restrict__ dH, long long cnt) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (long long)gridDim.x * blockDim.x) {
        const float z = Z[e], s = sigm(z);
        dH[e] *= s * (1.0f + z * (1.0f - s));
    }
}

// db[j] = sum_i dZ[i][j]   (one thread per column; fixed order)
__global__ void k_colsum(const float *__restrict__ dZ, int rows, int cols, float *__restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.f;
    for (int i = 0; i < rows; ++i) s += dZ[(size_t)i * cols + j];
    db[j] = s;
}

// masked MSE: dY = 2 M (Y - T) / sum(M); per-block partial losses into part
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             const float *__restrict__ Mk, long long cnt, float inv_w,
                                             float *__restrict__ dY, float *__restrict__ part) {
    __shared__ float sh[256];
    float s = 0.f;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (long long)gridDim.x * blockDim.x) {
        const float d = (Y[e] - T[e]) * Mk[e];
        s += d * d;
        dY[e] = 2.0f * d * inv_w;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] * inv_w;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       long long cnt, float lr, float b1, float b2, float c1, float c2, float eps) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (long long)gridDim.x * blockDim.x) {
        const float gg = g[e];
        const float mm = b1 * m[e] + (1.0f - b1) * gg, vv = b2 * v[e] + (1.0f - b2) * gg * gg;
        m[e] = mm;
        v[e] = vv;
        p[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    }
}

// rows idx[r] of the resident set -> a contiguous batch
__global__ void k_gather_rows_f(const float *__restrict__ src, const int *__restrict__ idx, int rows, int cols,
                                float *__restrict__ dst) {
    const long long cnt = (long long)rows * cols;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(e / cols), c = (int)(e - (long long)r * cols);
        dst[e] = src[(size_t)idx[r] * cols + c];
    }
}

// ---------------------------------------------------------------------------
// host: forward / backward on the resident batch buffers
// ---------------------------------------------------------------------------
static int mlp_ensure(myqc_ctx *ctx, MlpState &s, int rows) {
    if (rows <= s.cap) return 0;
    const int cap = std::max(rows, 1024);
    float **bufs[] = {&s.d_X, &s.d_T, &s.d_M, &s.d_Z[0], &s.d_Z[1], &s.d_Z[2], &s.d_H[0], &s.d_H[1], &s.d_H[2],
                      &s.d_Y, &s.d_dA, &s.d_dB};
    const size_t w = std::max({s.n_in, s.width, s.n_out});
    ++ctx->mlp_gen;   // captured graphs that use these buffers are stale now
    for (auto b : bufs) MYQC_CHECK(dalloc(ctx, (void **)b, sizeof(float) * (size_t)cap * w));
    MYQC_CHECK(dalloc(ctx, (void **)&s.d_loss, sizeof(float) * 1024));
    s.cap = cap;
    return 0;
}

template <bool TA, bool TB, int EPI>
static void sgemm(myqc_ctx *ctx, int M, int N, int K, const float *A, int lda, const float *B, int ldb,
                  const float *bias, float *Z, float *C) {
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    k_sgemm<TA, TB, EPI><<<grid, 256, 0, ctx->stream>>>(M, N, K, A, lda, B, ldb, bias, Z, C);
}

// Y = net(X) for `rows` rows in s.d_X (keeps Z, H for the backward pass)
static void mlp_forward(myqc_ctx *ctx, MlpState &s, int rows) {
    const float *P = s.d_par;
    const float *in = s.d_X;
    int k = s.n_in;
    for (int l = 0; l < 3; ++l) {
        sgemm<false, true, 1>(ctx, rows, s.width, k, in, k, P + s.off[2 * l], k, P + s.off[2 * l + 1], s.d_Z[l],
                              s.d_H[l]);
        in = s.d_H[l];
        k = s.width;
    }
    sgemm<false, true, 2>(ctx, rows, s.n_out, s.width, in, s.width, P + s.off[6], s.width, P + s.off[7], nullptr,
                          s.d_Y);
}

// gradient of the masked MSE into s.d_grad; loss partials in s.d_loss
static int mlp_backward(myqc_ctx *ctx, MlpState &s, int rows, float inv_w, int *nblk) {
    const long long cnt = (long long)rows * s.n_out;
    const int blocks = (int)std::min<long long>(1024, (cnt + 255) / 256);
    *nblk = blocks;
    k_mse<<<blocks, 256, 0, ctx->stream>>>(s.d_Y, s.d_T, s.d_M, cnt, inv_w, s.d_dA, s.d_loss);
    float *G = s.d_grad;
    const float *P = s.d_par;
    // output layer: dW4 = dY^T H3, db4 = colsum dY, dH3 = dY W4
    float *dcur = s.d_dA, *dnext = s.d_dB;
    int width_out = s.n_out;
    for (int l = 3; l >= 0; --l) {
        const float *in = l == 0 ? s.d_X : s.d_H[l - 1];
        const int k = l == 0 ? s.n_in : s.width;
        sgemm<true, false, 0>(ctx, width_out, k, rows, dcur, width_out, in, k, nullptr, nullptr, G + s.off[2 * l]);
        k_colsum<<<(width_out + 127) / 128, 128, 0, ctx->stream>>>(dcur, rows, width_out, G + s.off[2 * l + 1]);
        if (l > 0) {
            sgemm<false, false, 0>(ctx, rows, k, width_out, dcur, width_out, P + s.off[2 * l], k, nullptr, nullptr,
                                   dnext);
            k_silu_grad<<<1024, 256, 0, ctx->stream>>>(s.d_Z[l - 1], dnext, (long long)rows * s.width);
            std::swap(dcur, dnext);
            width_out = s.width;
        }
    }
    MYQC_HIP(ctx, hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// pair features, targets and density reconstruction for the batch
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gauss(float r, float mu, float w) {
    const float d = (r - mu) / w;
    return __expf(-0.5f * d * d);
}

// atoms: [Z, x, y, z] per atom (batch table); mol_atoms: [atom0, nnuc] per molecule
// Environment features of every atom: env[a][z][k] = sum over the other
// atoms c of its molecule with Z_c = z + 1 of gauss(|R_c - R_a|, 1 + k, 0.7),
// c in input order.  A thread per (atom, z, k); mol_atoms (first atom and
// count of each of the nmol molecules, in atom order) gives the atom's
// molecule by bisection.
__global__ void k_atom_env(const double *__restrict__ atoms, const int *__restrict__ mol_atoms, int nmol, int natoms,
                           float *__restrict__ env) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)natoms * kElem * kEnv) return;
    const int a = (int)(id / (kElem * kEnv)), slot = (int)(id - (long long)a * kElem * kEnv);
    const int z = slot / kEnv, k = slot - z * kEnv;
    int lo = 0, hi = nmol - 1;
    while (lo < hi) {   // the last molecule whose first atom is <= a
        const int mid = (lo + hi + 1) >> 1;
        if (mol_atoms[2 * mid] <= a) lo = mid;
        else hi = mid - 1;
    }
    const int a0 = mol_atoms[2 * lo], na = mol_atoms[2 * lo + 1];
    const double *X = atoms + 4 * a;
    float e = 0.f;
    for (int c = a0; c < a0 + na; ++c) {
        const double *C = atoms + 4 * c;
        if (c == a || (int)C[0] != z + 1) continue;
        const double ex = C[1] - X[1], ey = C[2] - X[2], ez = C[3] - X[3];
        e += gauss((float)sqrt(ex * ex + ey * ey + ez * ez), 1.0f + k, 0.7f);
    }
    env[id] = e;
}

// The network input of each pair (A, B): one-hot Z_A and Z_B, 16 Gaussians of
// R_AB, the on-site flag, the environments of A and B (k_atom_env), zeros up
// to kFeat.  A thread per (pair, feature), so that the rows are written
// coalesced.
__global__ void k_pair_features(const PairRec *__restrict__ pr, int npair, const double *__restrict__ atoms,
                                const float *__restrict__ env, float *__restrict__ F) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)npair * kFeat) return;
    const int p = (int)(id / kFeat), c = (int)(id - (long long)p * kFeat);
    const PairRec q = pr[p];
    const double *A = atoms + 4 * q.a, *B = atoms + 4 * q.b;
    float v;
    if (c < kElem) {
        v = c == (int)A[0] - 1 ? 1.f : 0.f;
    } else if (c < 2 * kElem) {
        v = c - kElem == (int)B[0] - 1 ? 1.f : 0.f;
    } else if (c < 2 * kElem + kRbf) {
        const double dx = B[1] - A[1], dy = B[2] - A[2], dz = B[3] - A[3];
        const float r = (float)sqrt(dx * dx + dy * dy + dz * dz);
        v = gauss(r, 0.5f * (c - 2 * kElem), 0.5f);
    } else if (c == 2 * kElem + kRbf) {
        v = q.a == q.b ? 1.f : 0.f;
    } else if (c < 2 * kElem + kRbf + 1 + 2 * kElem * kEnv) {
        const int e = c - (2 * kElem + kRbf + 1), side = e / (kElem * kEnv);
        v = env[(size_t)(side ? q.b : q.a) * kElem * kEnv + (e - side * kElem * kEnv)];
    } else {
        v = 0.f;   // the input row's padding up to kFeat
    }
    F[id] = v;
}

__device__ __forceinline__ void bond_axis(const double *A, const double *B, double u[3]) {
    const double dx = B[1] - A[1], dy = B[2] - A[2], dz = B[3] - A[3];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    u[0] = dx / r;
    u[1] = dy / r;
    u[2] = dz / r;
}

// the pair's invariants of the density D (batch blocks) -> T, with mask M
__global__ void k_pair_targets(const PairRec *__restrict__ pr, int npair, const double *__restrict__ atoms,
                               const double *__restrict__ D, float *__restrict__ T, float *__restrict__ Mk) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    const PairRec q = pr[p];
    float *t = T + (size_t)p * kOut, *mk = Mk + (size_t)p * kOut;
    for (int c = 0; c < kOut; ++c) t[c] = mk[c] = 0.f;
    const int n = q.n;
    auto d = [&](int i, int j) { return D[(size_t)q.mat0 + (q.aoA + i) + (size_t)n * (q.aoB + j)]; };
    const int sA = q.nA == 5 ? 2 : 1, sB = q.nB == 5 ? 2 : 1;
    if (q.a == q.b) {
        t[0] = (float)d(0, 0);
        mk[0] = 1.f;
        if (sA == 2) {
            t[1] = (float)d(0, 1);
            t[2] = (float)d(1, 1);
            t[3] = (float)((d(2, 2) + d(3, 3) + d(4, 4)) / 3.0);
            mk[1] = mk[2] = mk[3] = 1.f;
        }
        return;
    }
    double u[3];
    bond_axis(atoms + 4 * q.a, atoms + 4 * q.b, u);
    for (int i = 0; i < sA; ++i)
        for (int j = 0; j < sB; ++j) {
            t[i * 2 + j] = (float)d(i, j);
            mk[i * 2 + j] = 1.f;
        }
    if (q.nB == 5)
        for (int i = 0; i < sA; ++i) {
            t[4 + i] = (float)(d(i, 2) * u[0] + d(i, 3) * u[1] + d(i, 4) * u[2]);
            mk[4 + i] = 1.f;
        }
    if (q.nA == 5)
        for (int j = 0; j < sB; ++j) {
            t[6 + j] = (float)(d(2, j) * u[0] + d(3, j) * u[1] + d(4, j) * u[2]);
            mk[6 + j] = 1.f;
        }
    if (q.nA == 5 && q.nB == 5) {
        double ss = 0.0, tr = 0.0;
        for (int x = 0; x < 3; ++x) {
            tr += d(2 + x, 2 + x);
            for (int y = 0; y < 3; ++y) ss += u[x] * d(2 + x, 2 + y) * u[y];
        }
        t[8] = (float)ss;
        t[9] = (float)(0.5 * (tr - ss));
        mk[8] = mk[9] = 1.f;
    }
}

// predicted invariants Y -> density blocks D[A,B] and D[B,A] (batch layout)
__global__ void k_pair_reconstruct(const PairRec *__restrict__ pr, int npair, const double *__restrict__ atoms,
                                   const float *__restrict__ Y, double *__restrict__ D) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npair) return;
    const PairRec q = pr[p];
    const float *y = Y + (size_t)p * kOut;
    const int n = q.n;
    auto put = [&](int i, int j, double v) {
        D[(size_t)q.mat0 + (q.aoA + i) + (size_t)n * (q.aoB + j)] = v;
        D[(size_t)q.mat0 + (q.aoB + j) + (size_t)n * (q.aoA + i)] = v;
    };
    if (q.a == q.b) {
        put(0, 0, y[0]);
        if (q.nA == 5) {
            put(0, 1, y[1]);
            put(1, 1, y[2]);
            for (int x = 0; x < 3; ++x) {
                for (int z = 0; z < 3; ++z) put(2 + x, 2 + z, x == z ? y[3] : 0.0);
                put(0, 2 + x, 0.0);
                put(1, 2 + x, 0.0);
            }
        }
        return;
    }
    double u[3];
    bond_axis(atoms + 4 * q.a, atoms + 4 * q.b, u);
    const int sA = q.nA == 5 ? 2 : 1, sB = q.nB == 5 ? 2 : 1;
    for (int i = 0; i < sA; ++i)
        for (int j = 0; j < sB; ++j) put(i, j, y[i * 2 + j]);
    if (q.nB == 5)
        for (int i = 0; i < sA; ++i)
            for (int x = 0; x < 3; ++x) put(i, 2 + x, y[4 + i] * u[x]);
    if (q.nA == 5)
        for (int j = 0; j < sB; ++j)
            for (int x = 0; x < 3; ++x) put(2 + x, j, y[6 + j] * u[x]);
    if (q.nA == 5 && q.nB == 5)
        for (int x = 0; x < 3; ++x)
            for (int z = 0; z < 3; ++z) put(2 + x, 2 + z, (x == z ? y[9] : 0.0) + (y[8] - y[9]) * u[x] * u[z]);
}

int guess_reserve(myqc_ctx *ctx, int npair, int natoms) {
    if (!ctx->mlp) return set_error(ctx, MYQC_E_STATE, "learned guess: no network (myqc_mlp_create)");
    MlpState &s = *ctx->mlp;
    if (s.n_in != kFeat || s.n_out != kOut)
        return set_error(ctx, MYQC_E_ARG, "learned guess: the network must map %d features to %d outputs", kFeat, kOut);
    if (natoms > s.env_cap) {
        ++ctx->mlp_gen;   // captured graphs that use the buffer are stale now
        MYQC_CHECK(dalloc(ctx, (void **)&s.d_env, sizeof(float) * kElem * kEnv * (size_t)natoms));
        s.env_cap = natoms;
    }
    return mlp_ensure(ctx, s, npair);
}

// the environments of all atoms, then the pairs' features
static void pair_features(myqc_ctx *ctx, const PairRec *d_pairs, int npair, const double *d_atoms,
                          const int *d_mol_atoms, int nmol, int natoms, float *env, float *X) {
    const long long ne = (long long)natoms * kElem * kEnv;
    k_atom_env<<<(unsigned)((ne + 255) / 256), 256, 0, ctx->stream>>>(d_atoms, d_mol_atoms, nmol, natoms, env);
    const long long cnt = (long long)npair * kFeat;
    k_pair_features<<<(unsigned)((cnt + 255) / 256), 256, 0, ctx->stream>>>(d_pairs, npair, d_atoms, env, X);
}

int guess_launch(myqc_ctx *ctx, const PairRec *d_pairs, int npair, const double *d_atoms, const int *d_mol_atoms,
                 int nmol, int natoms, double *d_D) {
    MlpState &s = *ctx->mlp;
    const int g = (npair + 127) / 128;
    pair_features(ctx, d_pairs, npair, d_atoms, d_mol_atoms, nmol, natoms, s.d_env, s.d_X);
    mlp_forward(ctx, s, npair);
    k_pair_reconstruct<<<g, 128, 0, ctx->stream>>>(d_pairs, npair, d_atoms, s.d_Y, d_D);
    return 0;
}

int guess_export(myqc_ctx *ctx, const PairRec *d_pairs, int npair, const double *d_atoms, const int *d_mol_atoms,
                 int nmol, int natoms, const double *d_D, float *X, float *T, float *M) {
    float *dX = nullptr, *dT = nullptr, *dM = nullptr, *dE = nullptr;
    int rc = dalloc(ctx, (void **)&dX, sizeof(float) * (size_t)npair * kFeat);
    if (!rc) rc = dalloc(ctx, (void **)&dE, sizeof(float) * kElem * kEnv * (size_t)std::max(natoms, 1));
    if (!rc) rc = dalloc(ctx, (void **)&dT, sizeof(float) * (size_t)npair * kOut);
    if (!rc) rc = dalloc(ctx, (void **)&dM, sizeof(float) * (size_t)npair * kOut);
    if (!rc) {
        const int g = (npair + 127) / 128;
        pair_features(ctx, d_pairs, npair, d_atoms, d_mol_atoms, nmol, natoms, dE, dX);
        k_pair_targets<<<g, 128, 0, ctx->stream>>>(d_pairs, npair, d_atoms, d_D, dT, dM);
        if (hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "pair data: launch failed");
    }
    if (!rc && X && d2h(ctx, X, dX, sizeof(float) * (size_t)npair * kFeat) != hipSuccess) rc = MYQC_E_HIP;
    if (!rc && T && d2h(ctx, T, dT, sizeof(float) * (size_t)npair * kOut) != hipSuccess) rc = MYQC_E_HIP;
    if (!rc && M && d2h(ctx, M, dM, sizeof(float) * (size_t)npair * kOut) != hipSuccess) rc = MYQC_E_HIP;
    if (rc == MYQC_E_HIP) set_error(ctx, MYQC_E_HIP, "pair data: download failed");
    dfree((void **)&dX);
    dfree((void **)&dT);
    dfree((void **)&dM);
    dfree((void **)&dE);
    return rc;
}

}  // namespace myqc

using namespace myqc;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int myqc_mlp_create(myqc_ctx *ctx, int n_in, int width, int n_out, unsigned long long seed) {
    if (!ctx || n_in <= 0 || width <= 0 || n_out <= 0 || width > 4096) return MYQC_E_ARG;
    free_mlp(ctx);
    ++ctx->mlp_gen;
    ctx->mlp = new MlpState();
    MlpState &s = *ctx->mlp;
    s.n_in = n_in;
    s.width = width;
    s.n_out = n_out;
    const int fin[4] = {n_in, width, width, width}, fout[4] = {width, width, width, n_out};
    long long o = 0;
    for (int l = 0; l < 4; ++l) {
        s.off[2 * l] = o;
        o += (long long)fin[l] * fout[l];
        s.off[2 * l + 1] = o;
        o += fout[l];
    }
    s.n_par = o;
    std::vector<float> p(o, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < 4; ++l) {
        const float lim = std::sqrt(6.0f / (float)(fin[l] + fout[l]));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (long long e = 0; e < (long long)fin[l] * fout[l]; ++e) p[s.off[2 * l] + e] = U(rng);
    }
    MYQC_CHECK(dalloc(ctx, (void **)&s.d_par, sizeof(float) * o));
    MYQC_CHECK(dalloc(ctx, (void **)&s.d_grad, sizeof(float) * o));
    MYQC_CHECK(dalloc(ctx, (void **)&s.d_m, sizeof(float) * o));
    MYQC_CHECK(dalloc(ctx, (void **)&s.d_v, sizeof(float) * o));
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_par, p.data(), sizeof(float) * o, hipMemcpyHostToDevice, ctx->stream));
    std::vector<float> z(o, 0.f);
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_m, z.data(), sizeof(float) * o, hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_v, z.data(), sizeof(float) * o, hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.adam_t = 0;
    return 0;
}

extern "C" int myqc_mlp_params(myqc_ctx *ctx, float *get, const float *set, long long *count) {
    if (!ctx) return MYQC_E_ARG;
    if (!ctx->mlp) return set_error(ctx, MYQC_E_STATE, "myqc_mlp_params: no network (myqc_mlp_create)");
    MlpState &s = *ctx->mlp;
    if (count) *count = s.n_par;
    if (set) {
        MYQC_HIP(ctx, hipMemcpyAsync(s.d_par, set, sizeof(float) * s.n_par, hipMemcpyHostToDevice, ctx->stream));
        MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (get) MYQC_HIP(ctx, d2h(ctx, get, s.d_par, sizeof(float) * s.n_par));
    return 0;
}

extern "C" int myqc_mlp_forward(myqc_ctx *ctx, const float *X, int rows, float *Y) {
    if (!ctx || !X || !Y || rows <= 0) return MYQC_E_ARG;
    if (!ctx->mlp) return set_error(ctx, MYQC_E_STATE, "myqc_mlp_forward: no network");
    MlpState &s = *ctx->mlp;
    MYQC_CHECK(mlp_ensure(ctx, s, rows));
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_X, X, sizeof(float) * (size_t)rows * s.n_in, hipMemcpyHostToDevice, ctx->stream));
    mlp_forward(ctx, s, rows);
    MYQC_HIP(ctx, hipGetLastError());
    MYQC_HIP(ctx, d2h(ctx, Y, s.d_Y, sizeof(float) * (size_t)rows * s.n_out));
    return 0;
}

extern "C" int myqc_mlp_gradient(myqc_ctx *ctx, const float *X, const float *T, const float *Mk, int rows, float *loss,
                                 float *grad) {
    if (!ctx || !X || !T || !Mk || rows <= 0) return MYQC_E_ARG;
    if (!ctx->mlp) return set_error(ctx, MYQC_E_STATE, "myqc_mlp_gradient: no network");
    MlpState &s = *ctx->mlp;
    MYQC_CHECK(mlp_ensure(ctx, s, rows));
    const size_t nt = (size_t)rows * s.n_out;
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_X, X, sizeof(float) * (size_t)rows * s.n_in, hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_T, T, sizeof(float) * nt, hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(s.d_M, Mk, sizeof(float) * nt, hipMemcpyHostToDevice, ctx->stream));
    double w = 0.0;
    for (size_t e = 0; e < nt; ++e) w += Mk[e];
    mlp_forward(ctx, s, rows);
    int nb = 0;
    MYQC_CHECK(mlp_backward(ctx, s, rows, (float)(1.0 / std::max(w, 1.0)), &nb));
    std::vector<float> part(nb);
    MYQC_HIP(ctx, d2h(ctx, part.data(), s.d_loss, sizeof(float) * nb));
    double l = 0.0;
    for (float v : part) l += v;
    if (loss) *loss = (float)l;
    if (grad) MYQC_HIP(ctx, d2h(ctx, grad, s.d_grad, sizeof(float) * s.n_par));
    return 0;
}

// Adam over `epochs` passes of shuffled minibatches of the (rows x n_in)
// inputs X with targets T and masks Mk; the data stay resident on the device.
extern "C" int myqc_mlp_train(myqc_ctx *ctx, const float *X, const float *T, const float *Mk, int rows, int epochs,
                              int batch, float lr, unsigned long long seed, float *loss_per_epoch) {
    if (!ctx || !X || !T || !Mk || rows <= 0 || epochs <= 0 || batch <= 0) return MYQC_E_ARG;
    if (!ctx->mlp) return set_error(ctx, MYQC_E_STATE, "myqc_mlp_train: no network");
    MlpState &s = *ctx->mlp;
    batch = std::min(batch, rows);
    MYQC_CHECK(mlp_ensure(ctx, s, batch));
    float *dX = nullptr, *dT = nullptr, *dM = nullptr;
    int *dIdx = nullptr;
    int rc = dalloc(ctx, (void **)&dX, sizeof(float) * (size_t)rows * s.n_in);
    if (!rc) rc = dalloc(ctx, (void **)&dT, sizeof(float) * (size_t)rows * s.n_out);
    if (!rc) rc = dalloc(ctx, (void **)&dM, sizeof(float) * (size_t)rows * s.n_out);
    if (!rc) rc = dalloc(ctx, (void **)&dIdx, sizeof(int) * rows);
    auto cleanup = [&]() {
        dfree((void **)&dX);
        dfree((void **)&dT);
        dfree((void **)&dM);
        dfree((void **)&dIdx);
    };
    if (rc) {
        cleanup();
        return rc;
    }
    if (hipMemcpyAsync(dX, X, sizeof(float) * (size_t)rows * s.n_in, hipMemcpyHostToDevice, ctx->stream) ||
        hipMemcpyAsync(dT, T, sizeof(float) * (size_t)rows * s.n_out, hipMemcpyHostToDevice, ctx->stream) ||
        hipMemcpyAsync(dM, Mk, sizeof(float) * (size_t)rows * s.n_out, hipMemcpyHostToDevice, ctx->stream)) {
        cleanup();
        return set_error(ctx, MYQC_E_HIP, "myqc_mlp_train: upload failed");
    }
    std::vector<int> idx(rows);
    for (int i = 0; i < rows; ++i) idx[i] = i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<float> part(1024);
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(idx.begin(), idx.end(), rng);
        if (hipMemcpyAsync(dIdx, idx.data(), sizeof(int) * rows, hipMemcpyHostToDevice, ctx->stream)) {
            cleanup();
            return set_error(ctx, MYQC_E_HIP, "myqc_mlp_train: upload failed");
        }
        double ep_loss = 0.0;
        int nbatches = 0;
        for (int r0 = 0; r0 + batch <= rows; r0 += batch) {
            k_gather_rows_f<<<512, 256, 0, ctx->stream>>>(dX, dIdx + r0, batch, s.n_in, s.d_X);
            k_gather_rows_f<<<512, 256, 0, ctx->stream>>>(dT, dIdx + r0, batch, s.n_out, s.d_T);
            k_gather_rows_f<<<512, 256, 0, ctx->stream>>>(dM, dIdx + r0, batch, s.n_out, s.d_M);
            mlp_forward(ctx, s, batch);
            int nb = 0;
            // loss = sum of masked squared errors / (batch n_out)
            rc = mlp_backward(ctx, s, batch, 1.0f / (float)((size_t)batch * s.n_out), &nb);
            if (rc) {
                cleanup();
                return rc;
            }
            ++s.adam_t;
            const float c1 = 1.0f / (1.0f - powf(b1, (float)s.adam_t)), c2 = 1.0f / (1.0f - powf(b2, (float)s.adam_t));
            k_adam<<<512, 256, 0, ctx->stream>>>(s.d_par, s.d_grad, s.d_m, s.d_v, s.n_par, lr, b1, b2, c1, c2, eps);
            if (loss_per_epoch && r0 == 0) {
                // first batch of the epoch: read its loss (one synchronisation per epoch)
                if (d2h(ctx, part.data(), s.d_loss, sizeof(float) * nb) != hipSuccess) {
                    cleanup();
                    return set_error(ctx, MYQC_E_HIP, "myqc_mlp_train: download failed");
                }
                for (int q = 0; q < nb; ++q) ep_loss += part[q];
            }
            ++nbatches;
        }
        if (loss_per_epoch) loss_per_epoch[ep] = (float)ep_loss;
        if (hipGetLastError() != hipSuccess) {
            cleanup();
            return set_error(ctx, MYQC_E_HIP, "myqc_mlp_train: kernel launch failed");
        }
    }
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    cleanup();
    return 0;
}

// The network's AO density for one molecule of any size (H..Ne, STO-3G):
// pair list on the host, features, forward pass and reconstruction on the
// device.  xyz_bohr is (nnuc, 3) column-major; D is n x n.
extern "C" int myqc_guess_density(myqc_ctx *ctx, int nnuc, const int *Z, const double *xyz_bohr, double *D) {
    if (!ctx || nnuc <= 0 || !Z || !xyz_bohr || !D) return MYQC_E_ARG;
    std::vector<int> ao(nnuc), nf(nnuc);
    int n = 0;
    for (int a = 0; a < nnuc; ++a) {
        if (Z[a] < 1 || Z[a] > 10) return set_error(ctx, MYQC_E_ARG, "myqc_guess_density: Z=%d outside H..Ne", Z[a]);
        nf[a] = Z[a] <= 2 ? 1 : 5;
        ao[a] = n;
        n += nf[a];
    }
    // with a system set, the molecule must be that system, so that D has the
    // context's basis size (a caller that sizes D by it cannot be overrun)
    if (ctx->have_system) {
        if (nnuc != ctx->nnuc || n != ctx->n)
            return set_error(ctx, MYQC_E_ARG, "myqc_guess_density: molecule (%d atoms, n=%d) is not the system set (%d atoms, n=%d)",
                             nnuc, n, ctx->nnuc, ctx->n);
        for (int a = 0; a < nnuc; ++a)
            if (Z[a] != ctx->Z[a])
                return set_error(ctx, MYQC_E_ARG, "myqc_guess_density: Z[%d]=%d differs from the system set", a, Z[a]);
    }
    std::vector<PairRec> prs;
    for (int a = 0; a < nnuc; ++a)
        for (int c = a; c < nnuc; ++c) prs.push_back(PairRec{0, a, c, ao[a], ao[c], nf[a], nf[c], n, 0});
    std::vector<double> atoms(4 * (size_t)nnuc);
    for (int a = 0; a < nnuc; ++a) {
        atoms[4 * a] = Z[a];
        for (int d = 0; d < 3; ++d) atoms[4 * a + 1 + d] = xyz_bohr[a + (size_t)nnuc * d];
    }
    const int npair = (int)prs.size(), mol_atoms[2] = {0, nnuc};
    MYQC_CHECK(guess_reserve(ctx, npair, nnuc));
    PairRec *dP = nullptr;
    double *dA = nullptr, *dD = nullptr;
    int *dM = nullptr;
    int rc = dtmp(ctx, TMP_GPAIRS, &dP, sizeof(PairRec) * npair);
    if (!rc) rc = dtmp(ctx, TMP_GATOMS, &dA, sizeof(double) * atoms.size());
    if (!rc) rc = dtmp(ctx, TMP_GDENS, &dD, sizeof(double) * (size_t)n * n);
    if (!rc) rc = dtmp(ctx, TMP_GMOL, &dM, sizeof(int) * 2);
    if (!rc && (hipMemcpyAsync(dP, prs.data(), sizeof(PairRec) * npair, hipMemcpyHostToDevice, ctx->stream) ||
                hipMemcpyAsync(dA, atoms.data(), sizeof(double) * atoms.size(), hipMemcpyHostToDevice, ctx->stream) ||
                hipMemcpyAsync(dM, mol_atoms, sizeof(int) * 2, hipMemcpyHostToDevice, ctx->stream)))
        rc = set_error(ctx, MYQC_E_HIP, "myqc_guess_density: upload failed");
    if (!rc) rc = guess_launch(ctx, dP, npair, dA, dM, 1, nnuc, dD);
    if (!rc && hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "myqc_guess_density: launch failed");
    if (!rc && d2h(ctx, D, dD, sizeof(double) * (size_t)n * n) != hipSuccess)
        rc = set_error(ctx, MYQC_E_HIP, "myqc_guess_density: download failed");
    return rc;
}

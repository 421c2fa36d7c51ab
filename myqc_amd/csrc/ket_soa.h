// ket_soa.h -- the primitive pairs of a list of shell pairs as structure of
// arrays, for the ERI kernels that give consecutive lanes consecutive kets of
// such a list: field f (p, P[3], PA[3], K, EIJ, 1/p) of primitive j of list
// entry r sits at [(j * kKetFields + f) * np + r], so a wavefront reads each
// field with one coalesced load.  STO-3G pairs have 3 x 3 primitive pairs;
// callers refuse more when they build the arrays.
#pragma once

#include "eri_math.h"

namespace myqc {

constexpr int kKetFields = 10, kKetPrims = 9;

// grid (ceil(np / 256), kKetPrims) x 256
static __global__ void k_ket_soa(const ShellPair *__restrict__ pairs, const PrimPair *__restrict__ prims,
                                 const int *__restrict__ ids, int np, double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (r >= np) return;
    const ShellPair sp = pairs[ids[r]];
    double f[kKetFields] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (j < sp.npp) {
        const PrimPair q = prims[sp.pp0 + j];
        const double g[kKetFields] = {q.p, q.P[0], q.P[1], q.P[2], q.PA[0], q.PA[1], q.PA[2], q.K, q.EIJ, q.ip};
#pragma unroll
        for (int k = 0; k < kKetFields; ++k) f[k] = g[k];
    }
#pragma unroll
    for (int k = 0; k < kKetFields; ++k) out[((size_t)j * kKetFields + k) * np + r] = f[k];
}

// primitive j of list entry r, for quartet_prims_f
struct KetSoa {
    const double *__restrict__ s;
    long long np, r;
    __device__ __forceinline__ PrimPair operator()(int j) const {
        const double *b = s + (size_t)j * kKetFields * np + r;
        PrimPair q;
        q.p = b[0];
        q.P[0] = b[np];
        q.P[1] = b[2 * np];
        q.P[2] = b[3 * np];
        q.PA[0] = b[4 * np];
        q.PA[1] = b[5 * np];
        q.PA[2] = b[6 * np];
        q.K = b[7 * np];
        q.EIJ = b[8 * np];
        q.beta = 0.0;
        q.ip = b[9 * np];
        return q;
    }
};

}  // namespace myqc

// many.h -- argument blocks of the kernels that work on several systems of
// one basis size at once (the SCF of many clusters in lockstep): per-system
// device pointers passed by value, blockIdx selecting the system.
#pragma once

#include <vector>

#include "context.h"

namespace myqc {

constexpr int kMany = 32;   // systems per launch at most

// C_z = alpha A_z B_z (mfma_f64.hip)
struct GemmMany {
    const double *A[kMany];
    const double *B[kMany];
    double *C[kMany];
};
int gemm_f64_many(myqc_ctx *ctx, int m, int nb, int k, const GemmMany &g, int nz, double alpha);

// purification of each system's orthonormal-basis Fock matrix (sp2.hip)
struct Sp2Many {
    const double *F[kMany];   // orthonormal-basis Fock matrices
    double *P[kMany];         // projector out (and even steps' input)
    double *W[kMany];         // odd steps' input / even steps' output
    double *lohi[kMany];      // 2 n Gershgorin intervals
    double *slots[kMany];     // bounds, traces, control words (slots + 8)
    double *tp[kMany];        // per-tile trace partials of two steps
    double nocc[kMany];
};
int sp2_purify_many(myqc_ctx *ctx, const Sp2Many &a, int K, int n, int *status, std::vector<int> &ok, int *steps);
int gershgorin_many(myqc_ctx *ctx, const Sp2Many &a, int K, int n);

}  // namespace myqc

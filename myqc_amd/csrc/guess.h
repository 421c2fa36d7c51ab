// guess.h -- atom-pair records and the launches of the learned-guess kernels
// (mlp.hip) used by the batched SCF (batch.hip).
#pragma once

#include "context.h"

namespace myqc {

constexpr int kGuessFeat = 200, kGuessOut = 16;

// one atom pair A <= B of a batch molecule
struct PairRec {
    int mol, a, b;           // molecule; atoms as global indices into the batch atom table
    int aoA, aoB, nA, nB;    // molecule-local first AO and function counts (1 or 5)
    int n;                   // molecule's number of functions
    long long mat0;          // molecule's n x n block offset
};

// network rows for npair pairs (allocates the activation buffers; call
// before graph capture)
int guess_reserve(myqc_ctx *ctx, int npair, int natoms);
// features of the pairs into the network input, forward pass, and the
// predicted density blocks into D (batch layout); capturable.  d_mol_atoms:
// first atom and atom count of each of the nmol molecules (natoms in all)
int guess_launch(myqc_ctx *ctx, const PairRec *d_pairs, int npair, const double *d_atoms, const int *d_mol_atoms,
                 int nmol, int natoms, double *d_D);
// features, targets and masks of the pairs for the densities D, to the host
int guess_export(myqc_ctx *ctx, const PairRec *d_pairs, int npair, const double *d_atoms, const int *d_mol_atoms,
                 int nmol, int natoms, const double *d_D, float *X, float *T, float *M);

}  // namespace myqc

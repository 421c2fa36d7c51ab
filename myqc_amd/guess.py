"""The learned initial guess: host side.

The network, its training and the batch guess run on the GPU through the C
ABI (myqc_mlp_*, myqc_batch_pair_data, myqc_batch_rhf_learned; kernels in
csrc/mlp.hip).  This module holds what surrounds them:

  * training data: seeded config-4 molecules (2-8 atoms from H..Ne on random
    trees) and perturbations of them (every bond +-0.1 A, every bond angle
    +-5 degrees), plus seeded clusters of 2-5 waters at the density of the
    40-water workload (every atom displaced within 0.1 A), solved by the
    batched GPU SCF, with the atom-pair
    features, bond-frame density invariants and masks the library extracts;
  * training and saving/loading of the parameters (myqc_amd/data/guess_mlp.npz);
  * a numpy statement of the pair features, invariants and density
    reconstruction (the device kernels are checked against it; the product
    path for one molecule of any size is myqc_guess_density).

Feature and invariant definitions match csrc/mlp.hip exactly (the GPU test
compares them).
"""
from __future__ import annotations

import os

import numpy as np

from . import molecules as M

N_FEAT, N_OUT, WIDTH = 200, 16, 128
N_ELEM, N_RBF, N_ENV = 10, 16, 8
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "guess_mlp.npz")


# -- data -------------------------------------------------------------------
def perturb(rng, xyz_ang, amp=0.1):
    """Displace every atom uniformly within a ball of radius `amp` angstrom."""
    d = rng.normal(size=xyz_ang.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = amp * rng.uniform(size=(len(xyz_ang), 1)) ** (1.0 / 3.0)
    return xyz_ang + d * r


def perturb_tree(rng, xyz_ang, parent, d_bond=0.1, d_angle=5.0):
    """Perturb a molecule built on a tree (parent[k] < k) in its internal
    coordinates: every bond to a parent changes length by U(-d_bond, d_bond)
    angstrom and tilts by up to d_angle / 2 degrees about a random axis normal
    to it, so that every bond angle moves by at most d_angle.  Atoms are
    rebuilt in order, each from its parent's new position."""
    xyz = np.asarray(xyz_ang, dtype=np.float64)
    new = xyz.copy()
    for k in range(1, len(xyz)):
        j = parent[k]
        b = xyz[k] - xyz[j]
        L = np.linalg.norm(b)
        u = b / L
        a = rng.normal(size=3)
        a -= a.dot(u) * u
        a /= np.linalg.norm(a)
        th = np.radians(rng.uniform(0.0, 0.5 * d_angle))
        u2 = u * np.cos(th) + np.cross(a, u) * np.sin(th)
        new[k] = new[j] + u2 * (L + rng.uniform(-d_bond, d_bond))
    return new


def molecules(n_base, n_perturb, seed):
    """n_base config-4 molecules, each with n_perturb copies perturbed by
    +-0.1 A in every bond and +-5 degrees in every bond angle (bohr)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_base):
        Z, x, parent = M.random_molecule_tree(rng)
        for _ in range(n_perturb):
            out.append((Z, M.to_bohr(Z, perturb_tree(rng, x, parent))))
    return out


def water_clusters(count, seed, n_min=2, n_max=5):
    """Seeded clusters of n_min..n_max waters at the density of the 40-water
    workload (40 waters in a 12 A cube), each perturbed like the rest of the
    training set (bohr)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nw = int(rng.integers(n_min, n_max + 1))
        box = 12.0 * (nw / 40.0) ** (1.0 / 3.0)
        Z, x = M.water_cluster(nw, seed=int(rng.integers(1 << 31)), box=box)
        out.append((Z, M.to_bohr(Z, perturb(rng, x))))
    return out


def dataset(ctx, mols, chunk=4096, e_tol=1e-10, d_tol=1e-8, max_iter=100):
    """Solve `mols` with the batched GPU SCF (core-Hamiltonian guess) and
    return the pair features X, targets T and masks M of converged molecules,
    with the mean iteration count and converged fraction."""
    Xs, Ts, Ms, iters, conv = [], [], [], [], []
    for c0 in range(0, len(mols), chunk):
        ctx.batch_set(mols[c0:c0 + chunk])
        _, it, cv = ctx.batch_rhf(e_tol=e_tol, d_tol=d_tol, max_iter=max_iter)
        X, T, Mk = ctx.batch_pair_data(N_FEAT, N_OUT)
        Xs.append(X)
        Ts.append(T)
        Ms.append(Mk)
        iters.append(it[cv == 1])
        conv.append(cv == 1)
    X, T, Mk = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)
    keep = Mk.sum(axis=1) > 0
    return X[keep], T[keep], Mk[keep], float(np.mean(np.concatenate(iters))), float(np.mean(np.concatenate(conv)))


def train(ctx, X, T, Mk, epochs=60, batch=1024, lr=1e-3, seed=0, width=WIDTH):
    ctx.mlp_create(N_FEAT, width, N_OUT, seed)
    return ctx.mlp_train(X, T, Mk, epochs, batch, lr, seed)


def save(ctx, path=DATA, **meta):
    np.savez(path, params=ctx.mlp_params(), n_in=N_FEAT, width=WIDTH, n_out=N_OUT,
             **{k: np.asarray(v) for k, v in meta.items()})


def load(ctx, path=DATA):
    """Load trained parameters into the context's network; False if absent."""
    if not os.path.exists(path):
        return False
    f = np.load(path, allow_pickle=False)
    ctx.mlp_create(int(f["n_in"]), int(f["width"]), int(f["n_out"]), 0)
    ctx.mlp_params(set_to=f["params"])
    return True


# -- one molecule (numpy around the GPU forward pass) --------------------------
def _nfun(z):
    if not 1 <= z <= 10:
        raise ValueError(f"the learned guess covers H..Ne (Z = {z})")
    return 1 if z <= 2 else 5


def pairs(Z):
    """Atom pairs a <= b with first AOs and function counts."""
    nf = [_nfun(int(z)) for z in Z]
    ao = np.concatenate([[0], np.cumsum(nf)[:-1]]).astype(int)
    return [(a, b, ao[a], ao[b], nf[a], nf[b]) for a in range(len(Z)) for b in range(a, len(Z))]


def _gauss(r, mu, w):
    return np.exp(-0.5 * ((r - mu) / w) ** 2)


def features(Z, xyz_bohr):
    """(npair, N_FEAT) features of csrc/mlp.hip's k_pair_features."""
    Z = np.asarray(Z, int)
    x = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    na = len(Z)
    R = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    env = np.zeros((na, N_ELEM, N_ENV))
    mu = 1.0 + np.arange(N_ENV)
    for a in range(na):
        for c in range(na):
            if c != a:
                env[a, Z[c] - 1] += _gauss(R[a, c], mu, 0.7)
    pl = pairs(Z)
    F = np.zeros((len(pl), N_FEAT), np.float32)
    for p, (a, b, _, _, _, _) in enumerate(pl):
        F[p, Z[a] - 1] = 1.0
        F[p, N_ELEM + Z[b] - 1] = 1.0
        F[p, 2 * N_ELEM:2 * N_ELEM + N_RBF] = _gauss(R[a, b], 0.5 * np.arange(N_RBF), 0.5)
        F[p, 2 * N_ELEM + N_RBF] = 1.0 if a == b else 0.0
        o = 2 * N_ELEM + N_RBF + 1
        F[p, o:o + N_ELEM * N_ENV] = env[a].ravel()
        F[p, o + N_ELEM * N_ENV:o + 2 * N_ELEM * N_ENV] = env[b].ravel()
    return F


def _axis(x, a, b):
    d = x[b] - x[a]
    return d / np.linalg.norm(d)


def targets(Z, xyz_bohr, D):
    """Bond-frame invariants and masks of a density (k_pair_targets)."""
    x = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    pl = pairs(Z)
    T = np.zeros((len(pl), N_OUT), np.float32)
    Mk = np.zeros((len(pl), N_OUT), np.float32)
    for p, (a, b, oa, ob, na, nb) in enumerate(pl):
        blk = D[oa:oa + na, ob:ob + nb]
        sa, sb = (2 if na == 5 else 1), (2 if nb == 5 else 1)
        if a == b:
            T[p, 0], Mk[p, 0] = blk[0, 0], 1
            if na == 5:
                T[p, 1:4] = blk[0, 1], blk[1, 1], np.trace(blk[2:, 2:]) / 3.0
                Mk[p, 1:4] = 1
            continue
        u = _axis(x, a, b)
        for i in range(sa):
            for j in range(sb):
                T[p, 2 * i + j], Mk[p, 2 * i + j] = blk[i, j], 1
        if nb == 5:
            for i in range(sa):
                T[p, 4 + i], Mk[p, 4 + i] = blk[i, 2:] @ u, 1
        if na == 5:
            for j in range(sb):
                T[p, 6 + j], Mk[p, 6 + j] = blk[2:, j] @ u, 1
        if na == 5 and nb == 5:
            ss = u @ blk[2:, 2:] @ u
            T[p, 8], T[p, 9] = ss, 0.5 * (np.trace(blk[2:, 2:]) - ss)
            Mk[p, 8:10] = 1
    return T, Mk


def reconstruct(Z, xyz_bohr, Y):
    """Density from predicted invariants (k_pair_reconstruct)."""
    x = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    pl = pairs(Z)
    n = sum(_nfun(int(z)) for z in Z)
    D = np.zeros((n, n))
    Y = np.asarray(Y, dtype=np.float64)
    for p, (a, b, oa, ob, na, nb) in enumerate(pl):
        y = Y[p]
        blk = np.zeros((na, nb))
        if a == b:
            blk[0, 0] = y[0]
            if na == 5:
                blk[0, 1] = blk[1, 0] = y[1]
                blk[1, 1] = y[2]
                blk[2:, 2:] = y[3] * np.eye(3)
        else:
            u = _axis(x, a, b)
            sa, sb = (2 if na == 5 else 1), (2 if nb == 5 else 1)
            for i in range(sa):
                for j in range(sb):
                    blk[i, j] = y[2 * i + j]
            if nb == 5:
                for i in range(sa):
                    blk[i, 2:] = y[4 + i] * u
            if na == 5:
                for j in range(sb):
                    blk[2:, j] = y[6 + j] * u
            if na == 5 and nb == 5:
                blk[2:, 2:] = y[9] * np.eye(3) + (y[8] - y[9]) * np.outer(u, u)
        D[oa:oa + na, ob:ob + nb] = blk
        D[ob:ob + nb, oa:oa + na] = blk.T
    return D


def predict_density_host(ctx, Z, xyz_bohr):
    """The network's density for one molecule with the features and the
    reconstruction evaluated here (the forward pass on the GPU): the host
    statement that myqc_guess_density, which does all of it on the device,
    is checked against."""
    return reconstruct(Z, xyz_bohr, ctx.mlp_forward(features(Z, xyz_bohr), N_OUT))

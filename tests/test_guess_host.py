"""Host-side pieces of the learned guess (myqc_amd/guess.py): the
bond-frame density invariants are rotation invariant and the density
reconstruction inverts them."""
import numpy as np

from myqc_amd import guess as G
from myqc_amd import molecules as M


def _mol(seed):
    rng = np.random.default_rng(seed)
    Z, x = M.random_molecule(rng)
    return Z, M.to_bohr(Z, x)


def _rot(seed):
    q = np.random.default_rng(seed).normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def _ao_rotation(Z, R):
    """AO-space representation of the rotation R (p functions rotate)."""
    blocks = []
    for z in Z:
        if z <= 2:
            blocks.append(np.eye(1))
        else:
            B = np.eye(5)
            B[2:, 2:] = R
            blocks.append(B)
    n = sum(b.shape[0] for b in blocks)
    U = np.zeros((n, n))
    o = 0
    for b in blocks:
        k = b.shape[0]
        U[o:o + k, o:o + k] = b
        o += k
    return U


def test_reconstruction_inverts_invariants():
    for seed in range(5):
        Z, x = _mol(seed)
        rng = np.random.default_rng(seed)
        npair = len(G.pairs(Z))
        Y = rng.standard_normal((npair, G.N_OUT))
        D = G.reconstruct(Z, x, Y)
        assert np.allclose(D, D.T)
        T, Mk = G.targets(Z, x, D)
        assert np.max(np.abs((T - Y) * Mk)) <= 1e-6        # float32 targets
        assert Mk.sum() > 0


def test_features_and_targets_rotation_invariant():
    for seed in range(4):
        Z, x = _mol(seed + 10)
        rng = np.random.default_rng(seed)
        n = sum(1 if z <= 2 else 5 for z in Z)
        A = rng.standard_normal((n, n))
        D = A + A.T
        R = _rot(seed)
        U = _ao_rotation(Z, R)
        xr = x @ R.T
        Dr = U @ D @ U.T
        F0, F1 = G.features(Z, x), G.features(Z, xr)
        assert np.max(np.abs(F0 - F1)) <= 1e-5
        T0, M0 = G.targets(Z, x, D)
        T1, M1 = G.targets(Z, xr, Dr)
        assert np.array_equal(M0, M1)
        assert np.max(np.abs(T0 - T1)) <= 1e-5 * max(1.0, np.max(np.abs(T0)))


def test_feature_layout():
    Z, x = _mol(3)
    F = G.features(Z, x)
    pl = G.pairs(Z)
    assert F.shape == (len(pl), G.N_FEAT)
    for p, (a, b, *_rest) in enumerate(pl):
        assert F[p, Z[a] - 1] == 1.0 and F[p, G.N_ELEM + Z[b] - 1] == 1.0
        assert F[p, 2 * G.N_ELEM + G.N_RBF] == (1.0 if a == b else 0.0)
    assert np.all(F[:, 197:] == 0.0)

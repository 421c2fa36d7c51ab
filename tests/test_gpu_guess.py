"""The learned initial guess: the HIP training kernels' gradients against
the CPU statement of the same network (oracle/mlp.py, itself pinned by finite
differences), the device pair features/invariants against the host statement
(myqc_amd/guess.py), training, and what the trained network gains: SCF
iterations saved on held-out molecules, with the core-Hamiltonian guess as
fallback and the energies unchanged."""
import numpy as np
import pytest

from myqc_amd import guess as G
from myqc_amd import molecules as M

pytestmark = pytest.mark.gpu


def test_mlp_gradient_matches_cpu_reference(ctx):
    from oracle import mlp
    rng = np.random.default_rng(3)
    ctx.mlp_create(G.N_FEAT, G.WIDTH, G.N_OUT, 7)
    p = ctx.mlp_params()
    X = rng.standard_normal((37, G.N_FEAT)).astype(np.float32)
    T = rng.standard_normal((37, G.N_OUT)).astype(np.float32)
    Mk = (rng.uniform(size=(37, G.N_OUT)) < 0.6).astype(np.float32)
    loss, g = ctx.mlp_gradient(X, T, Mk)
    L, gr = mlp.loss_grad(p, X, T, Mk, G.N_FEAT, G.WIDTH, G.N_OUT)
    print(f"[mlp] loss {loss:.6e} vs {L:.6e}; grad rel err {np.linalg.norm(g - gr) / np.linalg.norm(gr):.2e}")
    assert abs(loss - L) <= 1e-5 * L
    assert np.linalg.norm(g - gr) <= 1e-4 * np.linalg.norm(gr)
    # per layer, so a wrong block cannot hide behind a large one
    o = 0
    for fin, fout in ((G.N_FEAT, G.WIDTH), (G.WIDTH, G.WIDTH), (G.WIDTH, G.WIDTH), (G.WIDTH, G.N_OUT)):
        for size in (fin * fout, fout):
            sl = slice(o, o + size)
            assert np.linalg.norm(g[sl] - gr[sl]) <= 1e-4 * np.linalg.norm(gr[sl]) + 1e-9
            o += size
    Y = ctx.mlp_forward(X, G.N_OUT)
    Yr = mlp.forward(p, X, G.N_FEAT, G.WIDTH, G.N_OUT)
    assert np.max(np.abs(Y - Yr)) <= 1e-5 * max(1.0, np.max(np.abs(Yr)))


def test_device_pair_data_matches_host_statement(ctx):
    rng = np.random.default_rng(31)
    mols = []
    for _ in range(12):
        Z, x = M.random_molecule(rng)
        mols.append((Z, M.to_bohr(Z, x)))
    ctx.batch_set(mols)
    _, _, cv = ctx.batch_rhf()
    Ds = ctx.batch_density()
    X, T, Mk = ctx.batch_pair_data(G.N_FEAT, G.N_OUT)
    o = 0
    for m, (Z, xyz) in enumerate(mols):
        F = G.features(Z, xyz)
        k = len(F)
        assert np.max(np.abs(X[o:o + k] - F)) <= 2e-5
        if cv[m] == 1:
            Th, Mh = G.targets(Z, xyz, Ds[m])
            assert np.array_equal(Mk[o:o + k], Mh)
            assert np.max(np.abs(T[o:o + k] - Th)) <= 1e-6
        else:
            assert not Mk[o:o + k].any()
        o += k
    assert o == len(X)


def test_device_guess_density_matches_host_statement(ctx):
    from myqc_amd.hip import MyqcError
    ctx.mlp_create(G.N_FEAT, G.WIDTH, G.N_OUT, 11)
    for seed in (1, 2):
        Z, x = M.random_molecule(np.random.default_rng(seed))
        xyz = M.to_bohr(Z, x)
        ctx.set_system(Z, xyz)
        D = ctx.guess_density(Z, xyz)
        Dh = G.predict_density_host(ctx, Z, xyz)
        assert np.max(np.abs(D - Dh)) <= 1e-5 * max(1.0, np.max(np.abs(Dh)))
    Z, x = M.water_cluster(6, seed=6)          # beyond the batch limit: one molecule of 42 functions
    ctx.set_system(Z, M.to_bohr(Z, x))
    D = ctx.guess_density(Z, M.to_bohr(Z, x))
    assert D.shape == (42, 42) and np.allclose(D, D.T)
    # a molecule other than the system set is refused (its n would not be the context's)
    Z2, x2 = M.random_molecule(np.random.default_rng(1))
    with pytest.raises(MyqcError):
        ctx.guess_density(Z2, M.to_bohr(Z2, x2))


def test_training_reduces_loss(ctx):
    X, T, Mk, it, cf = G.dataset(ctx, G.molecules(60, 5, seed=99))
    hist = G.train(ctx, X, T, Mk, epochs=30, batch=256, lr=1e-3, seed=3)
    print(f"[mlp] {len(X)} pairs; loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.2 * hist[0]


def test_learned_guess_saves_iterations(ctx):
    """The trained network (myqc_amd/data/guess_mlp.npz, or one trained here
    if that file is absent) against the core-Hamiltonian guess on held-out
    molecules: fewer SCF iterations (also in total, counting the core-guess
    reruns of the molecules that fall back), the same energies, few
    fallbacks."""
    if not G.load(ctx):
        X, T, Mk, _, _ = G.dataset(ctx, G.molecules(400, 5, seed=20000))
        G.train(ctx, X, T, Mk, epochs=40, batch=512, lr=1e-3, seed=1)
    mols = G.molecules(300, 1, seed=777)
    ctx.batch_set(mols)
    Ec, itc, cvc = ctx.batch_rhf(e_tol=1e-10, d_tol=1e-8, max_iter=100)
    El, itl, cvl, fb = ctx.batch_rhf_learned(e_tol=1e-10, d_tol=1e-8, max_iter=100)
    both = (cvc == 1) & (cvl == 1)
    same = both & (np.abs(El - Ec) <= 1e-8)
    kept = same & (fb == 0)
    lower = int(np.sum(both & (El < Ec - 1e-8)))
    print(f"[guess] {int(kept.sum())} molecules without fallback in the same state: iterations core "
          f"{np.mean(itc[kept]):.2f} learned {np.mean(itl[kept]):.2f}; fallback {np.mean(fb):.3f}; "
          f"different state {int(both.sum() - same.sum())}/{int(both.sum())} ({lower} lower with the learned start)")
    # over every molecule both forms converge, the learned runs' counts include
    # the core-guess rerun of the molecules that fell back
    print(f"[guess] total iterations over {int(both.sum())} molecules: core {int(itc[both].sum())} "
          f"learned incl. fallbacks {int(itl[both].sum())}")
    assert both.sum() >= 0.9 * len(mols)
    assert same.sum() >= 0.85 * both.sum()
    assert np.mean(itl[kept]) <= 0.95 * np.mean(itc[kept])
    assert itl[both].sum() < itc[both].sum()
    assert np.mean(fb) <= 0.05

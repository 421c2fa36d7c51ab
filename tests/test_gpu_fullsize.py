"""The headline workloads at their full sizes, checked through properties that
do not need the reference at that size (it would take days on one core):

  * the 40-water cluster (seed 40, n = 280): G from the Schwarz-screened
    (1e-12) store against G from the unscreened store for a random symmetric
    D; the mixed-precision and learned + mixed energies against plain FP64;
    the converged density against the unscreened Fock matrix (Roothaan
    commutator, idempotency, electron count);
  * the 24-water cluster (seed 24, n = 168): (ia|jb) of the FP64 transform
    against direct contractions of the dense integrals on sampled indices,
    and E(MBPT2) of the split-bf16 transform against FP64;
  * the seeded 4096-molecule batch: every sampled molecule against the
    single-molecule path.

References for what is computed: src/I2G/RHFI2G.f90:80-90 (G), src/scf/scf.f90
(RHF loop), src/ao2mo/ao2mo.f90:1306-1439 and src/mp2/mp2.f90:79-144 (MP2).
"""
import numpy as np
import pytest

from myqc_amd import molecules as M
from myqc_amd.basis import build_basis, read_ftab
from myqc_amd.hip import BF16X3, FP32, FP64, MyqcError, enuc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def w40(ctx):
    Z, x = M.water_cluster(40, seed=40)
    xyz = M.to_bohr(Z, x)
    basis = build_basis(Z)
    assert basis.norb == 280
    return Z, xyz, basis


def test_40water_screened_fock_matches_unscreened(ctx, w40):
    """Screened digestion schedule at n = 280 (units split across workgroups,
    dig_rows = 20) against the unscreened store of every unique quartet."""
    Z, xyz, basis = w40
    n = basis.norb
    ctx.set_system(Z, xyz, basis, read_ftab())
    rng = np.random.default_rng(40)
    D = rng.uniform(-1.0, 1.0, (n, n))
    D = 0.5 * (D + D.T)
    ns_screened = ctx.eri_build(1e-12)
    G12 = ctx.fock_G(D)
    print(f"[40w] screened store: {ns_screened} unique slots", flush=True)
    ns_all = ctx.eri_build(0.0)
    G0 = ctx.fock_G(D)
    scale = np.max(np.abs(G0))
    err = np.max(np.abs(G12 - G0))
    print(f"[40w] unscreened store: {ns_all} slots; max|G(1e-12) - G(0)| = {err:.3e} of max|G| {scale:.3e}",
          flush=True)
    assert ns_all > ns_screened
    assert err <= 1e-9 * scale
    assert np.max(np.abs(G0 - G0.T)) <= 1e-12 * scale


def test_40water_mixed_and_learned_energies_match_fp64(ctx, w40):
    """Plain FP64, mixed (FP32 dD digestion) and learned + mixed solves of the
    headline cluster reach the same energy within 1e-8 Eh; the FP64 solution
    is a Roothaan solution of the unscreened Fock matrix."""
    from myqc_amd import guess as G
    Z, xyz, basis = w40
    n = basis.norb
    nocc = sum(Z) // 2
    ctx.set_system(Z, xyz, basis, read_ftab())
    S, H = ctx.int1e()
    ctx.eri_build(1e-12)
    e_nuc = enuc(Z, xyz)
    r64 = ctx.rhf(H, S, nocc, e_nuc, e_tol=1e-10, d_tol=1e-8, max_iter=200, precision=FP64)
    r32 = ctx.rhf(H, S, nocc, e_nuc, e_tol=1e-10, d_tol=1e-8, max_iter=200, precision=FP32)
    print(f"[40w] E fp64 {r64['E']:.10f} ({r64['iterations']} it), mixed {r32['E']:.10f} ({r32['iterations']} it)",
          flush=True)
    assert abs(r32["E"] - r64["E"]) <= 1e-8
    if G.load(ctx):
        D0 = ctx.guess_density(Z, xyz)
        rl = ctx.rhf(H, S, nocc, e_nuc, e_tol=1e-10, d_tol=1e-8, max_iter=200, precision=FP32, D_guess=D0)
        print(f"[40w] learned + mixed {rl['E']:.10f} ({rl['iterations']} it)", flush=True)
        assert abs(rl["E"] - r64["E"]) <= 1e-8
    C = r64["C"]
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    assert abs(np.trace(D @ S) - 2 * nocc) <= 1e-8
    assert np.max(np.abs(D @ S @ D - 2.0 * D)) <= 1e-8
    assert np.max(np.abs(C.T @ S @ C - np.eye(n))) <= 1e-9
    ctx.eri_build(0.0)
    F = H + ctx.fock_G(D)
    E = 0.5 * np.sum(D * (H + F)) + e_nuc
    comm = F @ D @ S - S @ D @ F
    print(f"[40w] unscreened: E {E:.10f}, max|FDS - SDF| = {np.max(np.abs(comm)):.2e}", flush=True)
    assert abs(E - r64["E"]) <= 1e-8
    assert np.max(np.abs(comm)) <= 1e-6


def test_24water_mp2_fp64_and_bf16x3(ctx):
    """The 24-water MP2 configuration: sampled (ia|jb) of the FP64 transform
    against direct contractions of the dense integrals, E(MBPT2) of the
    split-bf16 transform within 2e-5 Eh of FP64."""
    Z, x = M.water_cluster(24, seed=24)
    xyz = M.to_bohr(Z, x)
    basis = build_basis(Z)
    n = basis.norb
    nocc = sum(Z) // 2
    assert n == 168
    ctx.set_system(Z, xyz, basis, read_ftab())
    S, H = ctx.int1e()
    ctx.eri_build(1e-12)
    r = ctx.rhf(H, S, nocc, enuc(Z, xyz), e_tol=1e-10, d_tol=1e-8, max_iter=200)
    C, eig = r["C"], r["eig"]
    iajb = ctx.ao2mo(C, nocc, FP64)
    _, _, e64 = ctx.mp2(eig, nocc)
    XX = ctx.eri_full(1e-12).reshape(n, n, n, n, order="F")
    rng = np.random.default_rng(24)
    nv = n - nocc
    for _ in range(3):
        i, j = rng.integers(0, nocc, 2)
        a, b = rng.integers(0, nv, 2)
        v = np.einsum("uvld,u,v,l,d->", XX, C[:, i], C[:, nocc + a], C[:, j], C[:, nocc + b], optimize=True)
        print(f"[24w] (ia|jb) ({i},{a}|{j},{b}) = {iajb[i, a, j, b]:.12e} direct {v:.12e}", flush=True)
        assert abs(iajb[i, a, j, b] - v) <= 1e-10
    del XX
    ctx.ao2mo(C, nocc, BF16X3, download=False)
    _, _, e16 = ctx.mp2(eig, nocc)
    print(f"[24w] E(MBPT2) fp64 {e64:.10f} bf16x3 {e16:.10f} diff {e16 - e64:.2e}", flush=True)
    assert abs(e16 - e64) <= 2e-5


def test_4096_batch_matches_single_molecule_path(ctx):
    """The seeded 4096-molecule batch (the benchmark's): at least 97 % converge,
    and every molecule of a fixed sample that converges in both paths has the
    single-molecule path's energy within 1e-8 Eh."""
    rng = np.random.default_rng(4096)
    mols = []
    for _ in range(4096):
        Z, x = M.random_molecule(rng)
        mols.append((Z, M.to_bohr(Z, x)))
    ctx.batch_set(mols)
    E, it, cv = ctx.batch_rhf(e_tol=1e-10, d_tol=1e-8, max_iter=100)
    conv = int(np.sum(cv == 1))
    print(f"[batch4096] {conv} converged, mean iterations {np.mean(it[cv == 1]):.2f}", flush=True)
    assert conv >= 0.97 * len(mols)
    idx = np.random.default_rng(7).choice(len(mols), size=48, replace=False)
    checked = 0
    for m in idx:
        Z, xyz = mols[m]
        ctx.set_system(Z, xyz)
        S, H = ctx.int1e()
        ctx.eri_build(1e-12)
        try:
            r = ctx.rhf(H, S, sum(Z) // 2, enuc(Z, xyz), e_tol=1e-10, d_tol=1e-8, max_iter=100)
        except MyqcError:
            continue
        if cv[m] != 1:
            continue
        checked += 1
        assert abs(r["E"] - E[m]) <= 1e-8, (m, Z, r["E"], E[m])
    print(f"[batch4096] {checked} of 48 sampled molecules checked against the single path", flush=True)
    assert checked >= 40

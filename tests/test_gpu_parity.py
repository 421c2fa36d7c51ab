"""GPU parity: the HIP kernels through the C ABI against the reference's own
outputs (tests/golden, written by the flang-built reference programs) and the
CPU oracle.  Tolerances: 1e-10 per integral, 1e-8 Eh on energies."""
import numpy as np
import pytest

from conftest import CONVERGED, MOLECULES, golden
from myqc_amd.basis import build_basis
from myqc_amd.hip import FP32, FP64, enuc

pytestmark = pytest.mark.gpu


def _dense_eri(f):
    """The reference's XX from a fixture's canonical (ij|gh) list, 8-fold filled."""
    n = int(f["n"])
    XX = np.zeros((n, n, n, n))
    i, j, g, h = f["eri_idx"].T
    v = f["eri_val"]
    for a, b, c, d in ((i, j, g, h), (j, i, g, h), (i, j, h, g), (j, i, h, g),
                       (g, h, i, j), (h, g, i, j), (g, h, j, i), (h, g, j, i)):
        XX[a, b, c, d] = v
    return XX


def _setup(ctx, name):
    f = golden(name)
    Z = [int(z) for z in f["Z"]]
    print(f"[case {name}] n={int(f['n'])}", flush=True)
    ctx.set_system(Z, f["xyz"], build_basis(Z))
    return f, Z


@pytest.mark.parametrize("name", MOLECULES)
def test_eri_full_matches_reference(ctx, name):
    f, _ = _setup(ctx, name)
    XX = ctx.eri_full(0.0)
    q = f["eri_idx"]
    mine = XX[q[:, 0], q[:, 1], q[:, 2], q[:, 3]]
    assert np.max(np.abs(mine - f["eri_val"])) <= 1e-10
    # 8-fold symmetry of the filled array
    assert np.array_equal(XX, XX.transpose(1, 0, 2, 3))
    assert np.array_equal(XX, XX.transpose(2, 3, 0, 1))
    if "XX_full" in f:
        assert np.max(np.abs(XX - f["XX_full"])) <= 1e-10


@pytest.mark.parametrize("name", MOLECULES)
def test_int1e_matches_reference(ctx, name):
    f, _ = _setup(ctx, name)
    S, H = ctx.int1e()
    assert np.max(np.abs(S - f["S"])) <= 1e-10
    assert np.max(np.abs(H - f["H"])) <= 1e-10


@pytest.mark.parametrize("name", MOLECULES)
@pytest.mark.parametrize("tol", [0.0, 1e-12])
def test_fock_G_matches_RHFI2G(ctx, name, tol):
    f, _ = _setup(ctx, name)
    ctx.eri_build(tol)
    G = ctx.fock_G(f["D_guess"], FP64)
    assert np.max(np.abs(G - f["G_guess"])) <= 1e-9
    # FP32 digestion of the same density: single-precision error bound
    G32 = ctx.fock_G(f["D_guess"], FP32)
    assert np.max(np.abs(G32 - f["G_guess"])) <= 1e-4 * max(1.0, np.max(np.abs(f["G_guess"])))


@pytest.mark.parametrize("name", ["co2", "h2o_2", "mol3"])
def test_density_matches_densRHF(ctx, name):
    from oracle import oracle as O
    f, _ = _setup(ctx, name)
    nocc = int(f["nocc"])
    D = ctx.density(f["C"], nocc)
    assert np.max(np.abs(D - O.density(f["C"], nocc))) <= 1e-13
    assert np.max(np.abs(ctx.density(f["C_guess"], nocc) - f["D_guess"])) <= 1e-13


@pytest.mark.parametrize("name", CONVERGED)
def test_rhf_energy_matches_reference(ctx, name):
    f, Z = _setup(ctx, name)
    ctx.eri_build(1e-12)
    S, H = ctx.int1e()
    r = ctx.rhf(H, S, int(f["nocc"]), enuc(Z, f["xyz"]), e_tol=1e-10, d_tol=1e-8)
    assert abs(r["E"] - float(f["E_scf"])) <= 1e-8
    # densities came from MFMA purification, not the diagonalisation fallback
    assert r["purifications"] >= r["iterations"] and r["diag_fallbacks"] == 0
    # orbital energies up to rounding of the converged Fock matrix
    assert np.max(np.abs(r["eig"] - f["eig"])) <= 1e-6


@pytest.mark.parametrize("name", [m for m in MOLECULES if m not in CONVERGED])
def test_rhf_converges_where_reference_loop_oscillates(ctx, name):
    """The reference's plain Roothaan loop does not converge on these inputs
    (src/scf/scf.f90:720-904); the DIIS driver must reach the stationary
    point the oracle's DIIS loop reaches (on the reference's S, H and XX)."""
    from oracle import oracle as O
    f, Z = _setup(ctx, name)
    assert not bool(f["converged"])
    ctx.eri_build(1e-12)
    S, H = ctx.int1e()
    nocc = int(f["nocc"])
    r = ctx.rhf(H, S, nocc, enuc(Z, f["xyz"]), e_tol=1e-10, d_tol=1e-8)
    XX = _dense_eri(f)
    ro = O.rhf(f["S"], f["H"], XX, nocc, O.enuc(Z, f["xyz"]), diis=True)
    assert abs(r["E"] - ro["E"]) <= 1e-8, (r["E"], ro["E"])
    # stationary: F D S - S D F = 0 for the returned orbitals
    C = r["C"]
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    F = f["H"] + O.fock_G(XX, D)
    comm = F @ D @ f["S"] - f["S"] @ D @ F
    assert np.max(np.abs(comm)) <= 1e-6


@pytest.mark.parametrize("name", ["co2", "h2o_2", "h2o_4"])
def test_rhf_mixed_precision_energy(ctx, name):
    f, Z = _setup(ctx, name)
    ctx.eri_build(1e-12)
    S, H = ctx.int1e()
    r = ctx.rhf(H, S, int(f["nocc"]), enuc(Z, f["xyz"]), precision=FP32)
    assert abs(r["E"] - float(f["E_scf"])) <= 1e-8


@pytest.mark.parametrize("name", CONVERGED)
def test_ao2mo_mp2_matches_reference(ctx, name):
    f, _ = _setup(ctx, name)
    nocc = int(f["nocc"])
    ctx.eri_build(0.0)
    iajb = ctx.ao2mo(f["C"], nocc)
    if "ijab_AB" in f:
        ref = f["ijab_AB"]            # [i][j][a][b] = (ia|jb)
        assert np.max(np.abs(iajb.transpose(0, 2, 1, 3) - ref)) <= 1e-10
    e_aa, e_ab, e_tot = ctx.mp2(f["eig"], nocc)
    assert abs(e_aa - float(f["E_aa"])) <= 1e-8
    assert abs(e_ab - float(f["E_ab"])) <= 1e-8
    assert abs(e_tot - float(f["E_mp2"])) <= 1e-8


@pytest.mark.parametrize("name", ["co2", "h2o_2"])
def test_mp2_from_records_matches_reference(ctx, name):
    f, _ = _setup(ctx, name)
    from oracle import oracle as O
    nocc = int(f["nocc"])
    XX = O.eri_full(O.System([int(z) for z in f["Z"]], f["xyz"], build_basis([int(z) for z in f["Z"]]),
                             __import__("myqc_amd.basis", fromlist=["read_ftab"]).read_ftab()))
    iajb = O.ao2mo_ovov(XX, f["C"], nocc)
    e_aa, e_ab, e_tot = ctx.mp2_records(iajb, f["eig"], nocc)
    assert abs(e_tot - float(f["E_mp2"])) <= 1e-10
    assert abs(e_aa - float(f["E_aa"])) <= 1e-10

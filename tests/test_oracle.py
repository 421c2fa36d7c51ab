"""Pins the CPU oracle: against the reference's own outputs (tests/golden,
written by the flang-built reference programs), and against values the
mathematics fixes independently of the reference (closed forms, textbook
H2/STO-3G numbers, symmetry and positivity invariants), chosen so that a
dropped term, a wrong sign or a transposed index fails at least one."""
import numpy as np
import pytest
from scipy.special import erf

from conftest import golden
from myqc_amd import molecules as M
from myqc_amd.basis import build_basis, read_ftab
from oracle import oracle as O

FT = read_ftab()
SMALL = ["co2", "heh", "h2o", "h2o_2", "mol2", "mol3"]


def _sys(name):
    f = golden(name)
    Z = [int(z) for z in f["Z"]]
    return f, Z, O.System(Z, f["xyz"], build_basis(Z), FT)


def test_boys_matches_reference_routine_bitwise():
    g = golden("boys")
    for Q, T, F in zip(g["Q"], g["T"], g["F"]):
        mine = O.boys(int(Q), float(T), FT)
        assert np.array_equal(mine, F[: Q + 1]), (Q, T)


def test_boys_close_to_closed_form_within_table_error():
    # F_0(T) = sqrt(pi/T) erf(sqrt T) / 2; the reference's tabulated values
    # (and its single-precision pi) keep it within ~1.5e-5 relative
    for Q in (0, 3, 6, 9, 12):
        for T in np.linspace(0.01, 40.0, 97):
            if 30.0 <= T < 2 * Q + 36:
                continue   # the reference's g(T) is undefined there
            exact = 0.5 * np.sqrt(np.pi / T) * erf(np.sqrt(T))
            assert abs(O.boys(Q, T, FT)[0] / exact - 1) < 2e-5, (Q, T)


def test_boys_downward_recursion_consistent():
    # F_j = (2T F_{j+1} + e^-T)/(2j+1) holds between the returned values
    for Q in (3, 6, 12):
        for T in (0.3, 2.7, 7.9):
            F = O.boys(Q, T, FT)
            for j in range(Q):
                assert abs(F[j] - (2 * T * F[j + 1] + np.exp(-T)) / (2 * j + 1)) < 1e-15


@pytest.mark.parametrize("name", SMALL)
def test_oracle_eri_matches_reference_int2e(name):
    f, Z, s = _sys(name)
    XX = O.eri_full(s)
    q = f["eri_idx"]
    assert np.max(np.abs(XX[q[:, 0], q[:, 1], q[:, 2], q[:, 3]] - f["eri_val"])) <= 1e-12


@pytest.mark.parametrize("name", SMALL)
def test_oracle_int1e_matches_reference(name):
    f, Z, s = _sys(name)
    S, H = O.int1e(s)
    assert np.max(np.abs(S - f["S"])) <= 1e-12
    assert np.max(np.abs(H - f["H"])) <= 1e-11


@pytest.mark.parametrize("name", SMALL)
def test_oracle_G_matches_RHFI2G(name):
    f, Z, s = _sys(name)
    G = O.fock_G(O.eri_full(s), f["D_guess"])
    assert np.max(np.abs(G - f["G_guess"])) <= 1e-12


def test_oracle_density_matches_dens():
    f, Z, s = _sys("co2")
    assert np.max(np.abs(O.density(f["C_guess"], 11) - f["D_guess"])) == 0.0


def test_oracle_rhf_reproduces_reference_scf_loop():
    f, Z, s = _sys("co2")
    XX = O.eri_full(s)
    S, H = O.int1e(s)
    r = O.rhf(S, H, XX, 11, O.enuc(Z, f["xyz"]))
    assert r["iterations"] == int(f["iterations"])
    assert abs(r["E"] - float(f["E_scf"])) <= 1e-10
    rd = O.rhf(S, H, XX, 11, O.enuc(Z, f["xyz"]), diis=True)
    assert abs(rd["E"] - float(f["E_scf"])) <= 1e-10


@pytest.mark.parametrize("name", ["co2", "h2o", "h2o_2"])
def test_oracle_mp2_matches_reference(name):
    f, Z, s = _sys(name)
    XX = O.eri_full(s)
    nocc = int(f["nocc"])
    iajb = O.ao2mo_ovov(XX, f["C"], nocc)
    ea, eb, et = O.mp2(iajb, f["eig"], nocc)
    assert abs(ea - float(f["E_aa"])) <= 1e-10
    assert abs(eb - float(f["E_ab"])) <= 1e-10
    assert abs(et - float(f["E_mp2"])) <= 1e-10


def test_oracle_ao2mo_matches_dense_einsum():
    f, Z, s = _sys("h2o")
    XX = O.eri_full(s)
    C = f["C"]
    nocc = int(f["nocc"])
    ref = np.einsum("uvld,ui,va,lj,db->iajb", XX, C[:, :nocc], C[:, nocc:], C[:, :nocc], C[:, nocc:])
    assert np.max(np.abs(O.ao2mo_ovov(XX, C, nocc) - ref)) <= 1e-12


def test_mp2_closed_shell_identity():
    # E(MBPT2) = sum (ia|jb)[2(ia|jb) - (ib|ja)]/D, the closed-shell formula
    f, Z, s = _sys("h2o")
    XX = O.eri_full(s)
    nocc = int(f["nocc"])
    eig = f["eig"]
    iajb = O.ao2mo_ovov(XX, f["C"], nocc)
    eo, ev = eig[:nocc], eig[nocc:]
    Dn = eo[:, None, None, None] + eo[None, None, :, None] - ev[None, :, None, None] - ev[None, None, None, :]
    e = np.sum(iajb * (2 * iajb - iajb.transpose(0, 3, 2, 1)) / Dn)
    assert abs(O.mp2(iajb, eig, nocc)[2] - e) <= 1e-12


def test_h2_textbook_values():
    """H2, STO-3G, R = 1.4 bohr (Szabo & Ostlund, Modern Quantum Chemistry,
    sec. 3.5.2): values to the 4 decimals printed there; the reference's
    single-precision pi and tabulated Boys function stay far below that."""
    Z = [1, 1]
    xyz = np.array([[0, 0, -0.7], [0, 0, 0.7]])
    s = O.System(Z, xyz, build_basis(Z), FT)
    S, H = O.int1e(s)
    XX = O.eri_full(s)
    assert abs(S[0, 1] - 0.6593) < 1e-4
    assert abs(H[0, 0] - (-1.1204)) < 1e-4
    assert abs(H[0, 1] - (-0.9584)) < 1e-4
    assert abs(XX[0, 0, 0, 0] - 0.7746) < 1e-4
    assert abs(XX[0, 0, 1, 1] - 0.5697) < 1e-4
    assert abs(XX[0, 1, 0, 1] - 0.2970) < 1e-4
    assert abs(XX[0, 0, 0, 1] - 0.4441) < 1e-4
    r = O.rhf(S, H, XX, 1, O.enuc(Z, xyz))
    assert abs(r["E"] - (-1.1167)) < 1e-4


def test_eri_invariants_on_water_dimer():
    f, Z, s = _sys("h2o_2")
    XX = O.eri_full(s)
    n = s.n
    # 8-fold symmetry
    assert np.array_equal(XX, XX.transpose(1, 0, 2, 3))
    assert np.array_equal(XX, XX.transpose(2, 3, 0, 1))
    # the (ij),(kl) supermatrix is positive semidefinite and obeys Schwarz
    Msup = XX.reshape(n * n, n * n, order="F")
    w = np.linalg.eigvalsh(0.5 * (Msup + Msup.T))
    assert w.min() > -1e-10
    d = np.sqrt(np.abs(np.diag(Msup)))
    assert np.all(np.abs(Msup) <= np.outer(d, d) * (1 + 1e-6) + 1e-12)


def test_density_trace_is_electron_count():
    f, Z, s = _sys("co2")
    D = O.density(f["C"], int(f["nocc"]))
    assert abs(np.sum(D * f["S"]) - 22.0) < 1e-10


def test_p_shell_rotation_covariance():
    # rotating a molecule about z by 90 deg maps p_x -> p_y, p_y -> -p_x
    Z = [8, 1]
    xyz = np.array([[0.1, 0.2, 0.3], [1.2, -0.4, 0.9]])
    R = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1.0]])
    s1 = O.System(Z, xyz, build_basis(Z), FT)
    s2 = O.System(Z, xyz @ R.T, build_basis(Z), FT)
    X1, X2 = O.eri_full(s1), O.eri_full(s2)
    # new p_x = -old p_y, new p_y = old p_x (AOs 2, 3): new_i = sum_p U[p, i] old_p
    U = np.eye(6)
    U[2:4, 2:4] = np.array([[0, 1], [-1, 0.0]])
    X1r = np.einsum("pi,qj,rk,sl,pqrs->ijkl", U, U, U, U, X1)
    assert np.max(np.abs(X1r - X2)) < 1e-6   # Boys table anchors move with T, so ~1e-7


def test_mlp_reference_gradient_matches_finite_differences():
    """The numpy network used to check the GPU training kernels, against
    central differences of its own loss (float64)."""
    from oracle import mlp
    rng = np.random.default_rng(5)
    n_in, width, n_out = 7, 5, 3
    npar = (n_in * width + width) + 2 * (width * width + width) + (width * n_out + n_out)
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((11, n_in))
    T = rng.standard_normal((11, n_out))
    Mk = (rng.uniform(size=(11, n_out)) < 0.7).astype(float)
    L, g = mlp.loss_grad(p, X, T, Mk, n_in, width, n_out)
    assert abs(L - np.sum(Mk * (mlp.forward(p, X, n_in, width, n_out) - T) ** 2) / Mk.sum()) <= 1e-14
    h = 1e-6
    fd = np.zeros(npar)
    for k in range(npar):
        e = np.zeros(npar)
        e[k] = h
        fd[k] = (mlp.loss_grad(p + e, X, T, Mk, n_in, width, n_out)[0]
                 - mlp.loss_grad(p - e, X, T, Mk, n_in, width, n_out)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_fixtures_reproduce_reference_example_outputs():
    """The oracle chains against the outputs the reference itself ships
    (examples/NO: UHF MOLDEN orbital energies and Cui; examples/O/singlet and
    examples/Be: RHF).  The energy of the shipped Cui goes through the
    reference's own dens + UHFI2G/RHFI2G; the example runs stopped at
    SCF_Conv 7 (density 1e-7), the fixtures at 1e-11."""
    g = golden("uhf_no")
    assert abs(float(g["E_scf"]) - float(g["example_E"])) <= 1e-9
    assert np.max(np.abs(g["eig_a"] - g["example_eig_a"])) <= 5e-7
    assert np.max(np.abs(g["eig_b"] - g["example_eig_b"])) <= 5e-7
    g = golden("ex_o1")
    assert abs(float(g["E_scf"]) - float(g["example_E"])) <= 1e-9
    assert np.max(np.abs(g["eig"] - g["example_eig"])) <= 1e-8   # MOLDEN prints 8 decimals
    g = golden("ex_be")
    assert np.max(np.abs(g["eig"] - g["example_eig"])) <= 1e-8


def _uhf_fock(g):
    n, na, nb = int(g["n"]), int(g["nalpha"]), int(g["nbeta"])
    X = g["XX_full"]
    Ca, Cb = g["Ca"], g["Cb"]
    Da, Db = Ca[:, :na] @ Ca[:, :na].T, Cb[:, :nb] @ Cb[:, :nb].T
    J = np.einsum("ijkl,kl->ij", X, Da + Db)
    Ka, Kb = np.einsum("ikjl,kl->ij", X, Da), np.einsum("ikjl,kl->ij", X, Db)
    return n, na, nb, Da, Db, g["H"] + J - Ka, g["H"] + J - Kb


@pytest.mark.parametrize("name", ["uhf_no", "uhf_oh", "uhf_o3", "uhf_h"])
def test_uhf_fixture_is_a_stationary_point(name):
    """The UHF fixtures (the reference's dens/UHFI2G chain) satisfy the UHF
    equations built independently from their dense XX: F_s = H + J[Da+Db] -
    K[D_s] commutes with D_s S, Tr(D_s S) = nocc_s, and
    E = (Da.(Fa+H) + Db.(Fb+H))/2 + E_nuc.  The one-electron H atom has
    E = H_00 / S_00 (its contracted 1s is not renormalised, so S_00 != 1)."""
    g = golden(name)
    n, na, nb, Da, Db, Fa, Fb = _uhf_fock(g)
    S = g["S"]
    for D, F, k in ((Da, Fa, na), (Db, Fb, nb)):
        assert abs(np.trace(D @ S) - k) <= 1e-9
        assert np.max(np.abs(F @ D @ S - S @ D @ F)) <= 1e-6
    E = 0.5 * (np.sum(Da * (Fa + g["H"])) + np.sum(Db * (Fb + g["H"]))) + float(g["Enuc"])
    assert abs(E - float(g["E_scf"])) <= 1e-9
    if name == "uhf_h":
        assert abs(float(g["E_scf"]) - float(g["H"][0, 0] / S[0, 0])) <= 1e-12


@pytest.mark.parametrize("name", ["uhf_no", "uhf_oh", "uhf_o3"])
def test_uhf_mp2_fixture_against_spin_orbital_formula(name):
    """The reference's mp2_uhf values in the fixtures, against the MP2 energy
    written in spin orbitals, sum_{i<j,a<b} |<ij||ab>|^2 / D over the
    combined alpha+beta orbital set, from the dense XX (a different route:
    one antisymmetrised sum instead of three spin blocks)."""
    g = golden(name)
    n, na, nb = int(g["n"]), int(g["nalpha"]), int(g["nbeta"])
    X = g["XX_full"]
    # spin orbitals: occupied (alpha then beta), virtual likewise
    Cs = {"a": g["Ca"], "b": g["Cb"]}
    es = {"a": g["eig_a"], "b": g["eig_b"]}
    occ = [("a", i) for i in range(na)] + [("b", i) for i in range(nb)]
    vir = [("a", a) for a in range(na, n)] + [("b", b) for b in range(nb, n)]

    def mo(p, q, r, s):   # (pq|rs) in chemists' notation, zero unless spins pair up
        if p[0] != q[0] or r[0] != s[0]:
            return 0.0
        return np.einsum("uvlk,u,v,l,k->", X, Cs[p[0]][:, p[1]], Cs[q[0]][:, q[1]], Cs[r[0]][:, r[1]],
                         Cs[s[0]][:, s[1]])

    E = 0.0
    for x in range(len(occ)):
        for y in range(x + 1, len(occ)):
            for u in range(len(vir)):
                for v in range(u + 1, len(vir)):
                    i, j, a, b = occ[x], occ[y], vir[u], vir[v]
                    t = mo(i, a, j, b) - mo(i, b, j, a)
                    if t == 0.0:
                        continue
                    E += t * t / (es[i[0]][i[1]] + es[j[0]][j[1]] - es[a[0]][a[1]] - es[b[0]][b[1]])
    assert abs(E - float(g["E_mp2"])) <= 1e-10
    assert abs(float(g["E_aa"]) + float(g["E_bb"]) + float(g["E_ab"]) - float(g["E_mp2"])) <= 1e-14


def test_cis_oracle_h2_singlet_triplet_closed_form():
    """UHF-CIS on closed-shell H2 (STO-3G, R = 1.4 bohr; one occupied and one
    virtual orbital per spin): the 2 x 2 CIS matrix has the triplet e2 - e1 -
    J12 and the singlet e2 - e1 - J12 + 2 K12 (Szabo & Ostlund, sec. 4.3), with
    J12 = (11|22) and K12 = (12|12) in the MO basis.  Pins the roles and signs
    of (ai|jb) and (ab|ji) in the restated streams and matrix."""
    from oracle import ref_driver as R
    Z = [1, 1]
    xyz = np.array([[0, 0, -0.7], [0, 0, 0.7]])
    s = O.System(Z, xyz, build_basis(Z), FT)
    S, H = O.int1e(s)
    XX = O.eri_full(s)
    r = O.rhf(S, H, XX, 1, O.enuc(Z, xyz))
    C, e = r["C"], r["eig"]
    mo = np.einsum("uvlk,up,vq,lr,ks->pqrs", XX, C, C, C, C)
    J12, K12 = mo[0, 0, 1, 1], mo[0, 1, 0, 1]
    st = R.cis_uhf_streams(XX, C, C, 1, 1)
    w = np.linalg.eigvalsh(R.cis_uhf_matrix(st, e, e, 1, 1))
    d = e[1] - e[0]
    assert np.allclose(w, [d - J12, d - J12 + 2 * K12], atol=1e-12, rtol=0)
    assert K12 > 1e-3   # the two states are distinct


def test_cis_oracle_oh_matches_cfour():
    """UHF-CIS of OH (examples/OH: STO-3G, MULTI=2, EXCITE=CIS, E_NUM=4) from
    the reference chain's orbitals, against the TDA (CIS) excitation energies
    in the CFOUR output the reference ships (examples/OH/cfour/out): 0,
    0.228665544, 0.400333416 Eh.  CFOUR's integrals use the exact pi and Boys
    function (its E(SCF) is 9e-6 Eh from the reference's), so the agreement
    is to 1e-5."""
    from oracle import ref_driver as R
    g = golden("uhf_oh")
    na, nb = int(g["nalpha"]), int(g["nbeta"])
    st = R.cis_uhf_streams(g["XX_full"], g["Ca"], g["Cb"], na, nb)
    w = np.linalg.eigvalsh(R.cis_uhf_matrix(st, g["eig_a"], g["eig_b"], na, nb))
    assert np.max(np.abs(w[:3] - [0.0, 0.228665544103, 0.400333416145])) <= 1e-5


@pytest.mark.parametrize("name", ["props_co2", "props_h2o", "props_heh", "props_h2o_2", "props_mol2"])
def test_reference_dipole_integrals_closed_forms(name):
    """The reference int1p's dipole integrals (fixtures): symmetric, and on
    the diagonal P_c(u,u) = R_c S(u,u) with R the centre of AO u (an s or p
    function's density is symmetric about its centre) -- pins the origin and
    sign.  The restated cphf contraction gives CO2 (centred, symmetric) a zero
    dipole."""
    g = golden(name)
    P, S = g["P"], g["S"]
    Z = [int(z) for z in g["Z"]]
    for c in range(3):
        assert np.max(np.abs(P[c] - P[c].T)) <= 1e-14
    ao_atom = []   # AO order of buildBasis: 1s for H/He, 1s 2s 2p for Li-Ne
    for a, z in enumerate(Z):
        ao_atom += [a] * (1 if z <= 2 else 5)
    R = np.asarray(g["xyz"])[ao_atom]
    for c in range(3):
        assert np.max(np.abs(np.diag(P[c]) - R[:, c] * np.diag(S))) <= 1e-12
    if name == "props_co2":
        assert np.max(np.abs(g["mu"])) <= 1e-10

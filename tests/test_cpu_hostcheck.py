"""The library's own integral math (myqc_amd/csrc/eri_math.h, one_e_math.h,
system.h) compiled host-only into tests/hostcheck/libhostcheck.so, against
the reference's outputs: the OS/HGP recursions, the reference-compatible
Boys function, the digestion weights and the 1e integrals are exercised here
without a GPU.  The GPU tests run the same code in the kernels."""
import ctypes
import os

import numpy as np
import pytest

from conftest import MOLECULES, ROOT, golden
from myqc_amd.basis import build_basis, read_ftab
from oracle import oracle as O

HC = os.path.join(ROOT, "tests", "hostcheck", "libhostcheck.so")
FT = read_ftab()
_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(HC):
        pytest.skip("host check harness not built (run __graft_entry__.build())")
    return ctypes.CDLL(HC)


def _sys(name):
    f = golden(name)
    Z = [int(z) for z in f["Z"]]
    return f, Z, O.System(Z, f["xyz"], build_basis(Z), FT)


@pytest.mark.parametrize("name", MOLECULES)
def test_host_eri_matches_reference(hc, name):
    f, Z, s = _sys(name)
    n = s.n
    XX = np.zeros(n ** 4)
    assert hc.hc_eri_full(*s.args(), XX.ctypes.data_as(_dp)) == 0
    XX = XX.reshape(n, n, n, n, order="F")
    q = f["eri_idx"]
    assert np.max(np.abs(XX[q[:, 0], q[:, 1], q[:, 2], q[:, 3]] - f["eri_val"])) <= 1e-12


@pytest.mark.parametrize("name", MOLECULES)
def test_host_digestion_matches_RHFI2G(hc, name):
    f, Z, s = _sys(name)
    n = s.n
    D = np.asfortranarray(f["D_guess"]).ravel(order="F").copy()
    G = np.zeros(n * n)
    assert hc.hc_fock_G(*s.args(), D.ctypes.data_as(_dp), G.ctypes.data_as(_dp)) == 0
    assert np.max(np.abs(G.reshape(n, n, order="F") - f["G_guess"])) <= 1e-11


def test_host_digestion_nonsymmetric_density(hc):
    # the reference's formula sees D only through D + D^T
    f, Z, s = _sys("h2o_2")
    n = s.n
    rng = np.random.default_rng(3)
    D = rng.standard_normal((n, n))
    XX = O.eri_full(s)
    ref = O.fock_G(XX, D)
    Df = np.asfortranarray(D).ravel(order="F").copy()
    G = np.zeros(n * n)
    hc.hc_fock_G(*s.args(), Df.ctypes.data_as(_dp), G.ctypes.data_as(_dp))
    assert np.max(np.abs(G.reshape(n, n, order="F") - ref)) <= 1e-11


@pytest.mark.parametrize("name", MOLECULES)
def test_host_int1e_matches_reference(hc, name):
    f, Z, s = _sys(name)
    n = s.n
    S = np.zeros(n * n)
    H = np.zeros(n * n)
    a = s.args()
    assert hc.hc_int1e(s.nnuc, s.Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), *a[1:], S.ctypes.data_as(_dp),
                       H.ctypes.data_as(_dp)) == 0
    assert np.max(np.abs(S.reshape(n, n, order="F") - f["S"])) <= 1e-12
    assert np.max(np.abs(H.reshape(n, n, order="F") - f["H"])) <= 1e-11


def test_host_boys_matches_reference_routine(hc):
    g = golden("boys")
    F = np.zeros(5)
    worst = 0.0
    for Q, T, Fr in zip(g["Q"], g["T"], g["F"]):
        Q = int(Q)
        M = min(Q, 4)
        assert hc.hc_boys(M, Q, ctypes.c_double(float(T)), FT.ctypes.data_as(_dp), F.ctypes.data_as(_dp)) == 0
        worst = max(worst, np.max(np.abs(F[: M + 1] - Fr[: M + 1]) / np.abs(Fr[: M + 1])))
    assert worst <= 1e-14


@pytest.mark.parametrize("name", ["co2", "h2o_2", "mol3"])
def test_row_owned_digestion_formula(hc, name):
    """The J/K formula the digestion kernel uses (rows owned by the first bra
    shell, kets in their stored orientation) reproduces RHFI2G's G, for the
    reference's density and for a nonsymmetric one."""
    f, Z, s = _sys(name)
    n = s.n
    rng = np.random.default_rng(7)
    for D, ref in ((f["D_guess"], f["G_guess"]), (None, None)):
        if D is None:
            D = rng.standard_normal((n, n))
            ref = O.fock_G(O.eri_full(s), D)
        Df = np.asfortranarray(D).ravel(order="F").copy()
        G = np.zeros(n * n)
        assert hc.hc_fock_G_rows(*s.args(), Df.ctypes.data_as(_dp), G.ctypes.data_as(_dp)) == 0
        assert np.max(np.abs(G.reshape(n, n, order="F") - ref)) <= 1e-11


@pytest.mark.parametrize("name", ["co2", "h2o_2", "mol3"])
def test_pair_owned_digestion_formula(hc, name):
    """The pair-owned J/K split the digestion kernel uses (each shell pair's
    copy adds J, and K only where the pair is the unique quartet's bra;
    G = A + A^T) reproduces RHFI2G's G for the reference's density and for a
    nonsymmetric one."""
    f, Z, s = _sys(name)
    n = s.n
    rng = np.random.default_rng(11)
    for D, ref in ((f["D_guess"], f["G_guess"]), (None, None)):
        if D is None:
            D = rng.standard_normal((n, n))
            ref = O.fock_G(O.eri_full(s), D)
        Df = np.asfortranarray(D).ravel(order="F").copy()
        G = np.zeros(n * n)
        assert hc.hc_fock_G_pairs(*s.args(), Df.ctypes.data_as(_dp), G.ctypes.data_as(_dp)) == 0
        assert np.max(np.abs(G.reshape(n, n, order="F") - ref)) <= 1e-11

@pytest.mark.parametrize("name", ["co2", "h2o_2", "mol3"])
def test_three_copy_digestion_formula(hc, name):
    """The three-copy row-owned split the digestion kernel uses (the bra
    pair's copies in both orientations, the ket pair's in its stored one;
    J and K weights by role; G = A + A^T) reproduces RHFI2G's G for the
    reference's density and for a nonsymmetric one."""
    f, Z, s = _sys(name)
    n = s.n
    rng = np.random.default_rng(12)
    for D, ref in ((f["D_guess"], f["G_guess"]), (None, None)):
        if D is None:
            D = rng.standard_normal((n, n))
            ref = O.fock_G(O.eri_full(s), D)
        Df = np.asfortranarray(D).ravel(order="F").copy()
        G = np.zeros(n * n)
        assert hc.hc_fock_G_3copy(*s.args(), Df.ctypes.data_as(_dp), G.ctypes.data_as(_dp)) == 0
        assert np.max(np.abs(G.reshape(n, n, order="F") - ref)) <= 1e-11



def test_boys_grid_point_without_division():
    """The Boys expansion point Tk/10 is formed as q = Tk*0.1 corrected by one
    fma (no division); for every integer 0..1000 that equals the correctly
    rounded quotient.  fma is emulated exactly with rationals (float() of a
    Fraction rounds to nearest even, as IEEE fma does)."""
    from fractions import Fraction

    def fma(a, b, c):
        return float(Fraction(a) * Fraction(b) + Fraction(c))

    for tk in range(1001):
        x = float(tk)
        q = x * 0.1
        assert fma(fma(-q, 10.0, x), 0.1, q) == x / 10.0, tk


def _plan_G(hc, s, D, tol, max_rows, target):
    n = s.n
    Df = np.asfortranarray(D).ravel(order="F").copy()
    G = np.zeros(n * n)
    st = np.zeros(6, dtype=np.int64)
    fn = hc.hc_fock_G_plan
    fn.argtypes = None
    rc = fn(*s.args(), Df.ctypes.data_as(_dp), ctypes.c_double(tol), ctypes.c_int(max_rows),
            ctypes.c_longlong(target), G.ctypes.data_as(_dp), st.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    return rc, G.reshape(n, n, order="F"), st


@pytest.mark.parametrize("name,max_rows,target", [("co2", 64, 1 << 15), ("h2o_4", 6, 200), ("mol1", 7, 500),
                                                  ("mol3", 11, 1 << 15), ("heh", 6, 1)])
def test_host_one_copy_digestion_matches_RHFI2G(hc, name, max_rows, target):
    """The one-copy layout and schedule of digest_plan.h with the digestion
    kernel's updates, restated on the host, against the reference's RHFI2G
    (the fixture's G of its first density); small LDS row budgets and
    workgroup targets force units split over workgroups.  Every stored
    element is read exactly once (rc == 0)."""
    f, Z, s = _sys(name)
    rc, G, st = _plan_G(hc, s, f["D_guess"], 0.0, max_rows, target)
    print(f"[plan] {name}: values {st[0]} quartets {st[1]} workgroups {st[2]} chunks {st[3]} rows {st[4]}")
    assert rc == 0
    assert np.max(np.abs(G - f["G_guess"])) <= 1e-11


def test_host_one_copy_digestion_water_cluster(hc):
    """Eight waters (n = 56): several column chunks per ket class, units
    split over workgroups; unscreened against the unique-integral G build
    (itself pinned to RHFI2G above), and with a Schwarz screen every stored
    element is still read exactly once."""
    from myqc_amd import molecules as M
    Z, x = M.water_cluster(8, seed=8)
    xyz = M.to_bohr(Z, x)
    s = O.System(Z, xyz, build_basis(Z), FT)
    n = s.n
    rng = np.random.default_rng(8)
    D = rng.uniform(-1.0, 1.0, (n, n))
    Df = np.asfortranarray(D).ravel(order="F").copy()
    G0 = np.zeros(n * n)
    assert hc.hc_fock_G(*s.args(), Df.ctypes.data_as(_dp), G0.ctypes.data_as(_dp)) == 0
    G0 = G0.reshape(n, n, order="F")
    rc, G, st = _plan_G(hc, s, D, 0.0, 7, 3000)
    print(f"[plan] 8 waters: values {st[0]} workgroups {st[2]} chunks {st[3]}")
    assert rc == 0 and st[3] > st[2]
    assert np.max(np.abs(G - G0)) <= 1e-11 * max(1.0, np.max(np.abs(G0)))
    rc, G8, st8 = _plan_G(hc, s, D, 1e-8, 11, 1 << 15)
    assert rc == 0 and st8[0] < st[0]
    assert np.max(np.abs(G8 - G0)) <= 1e-5

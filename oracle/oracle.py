"""Python front end of the CPU oracle (oracle/libmyqc_oracle.so).

TEST INFRASTRUCTURE ONLY -- the checker.  Only tests/, the fixture generator
under tests/golden/ and bench.py's CPU-baseline leg may import this module.

Functions mirror the reference programs they restate:
  eri_full      int2e proc2e + fillsym (src/integrals/int2e.f90:78-351,540-613)
  int1e         int1e proc1e (src/integrals/int1e.f90:132-280)
  fock_G        RHFI2G (src/I2G/RHFI2G.f90:80-90)
  density       densRHF (src/dens/dens.f90:115-124)
  ao2mo_ovov    slow_ao2mo_MP2_RHF (src/ao2mo/ao2mo.f90:465-602)
  mp2           mp2_rhf (src/mp2/mp2.f90:79-144)
  rhf           the scf.f90 RHF driver (src/scf/scf.f90:61-216,410-512,720-904,
                1157-1267) or, with diis=True, the same fixed point reached by
                DIIS-accelerated Roothaan steps.
"""
from __future__ import annotations

import ctypes
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libmyqc_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.oracle_boys.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.oracle_eri_full.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _ip, _dp, _dp]
        L.oracle_eri_sample.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _ip, _dp,
                                        ctypes.c_int, _ip, _dp]
        L.oracle_eri_sample.restype = ctypes.c_longlong
        L.oracle_int1e.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _ip, _dp,
                                   _dp, _dp]
        L.oracle_fock_G.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.oracle_density.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.oracle_mp2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        _LIB = L
    return _LIB


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _P(a, t=_dp):
    return a.ctypes.data_as(t)


class System:
    """Molecule + basis in the reference layouts (xyz is (nnuc,3), passed
    column-major as nucpos/getenv hold it)."""

    def __init__(self, Z, xyz_bohr, basis, ftab):
        self.Z = _i(Z)
        self.xyzF = _d(np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3).ravel(order="F"))
        self.xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.basis = basis
        self.alpha = _d(basis.set_alpha)
        self.setinfo = _i(basis.setinfo)
        self.bas = _d(basis.bas)
        self.basinfo = _i(basis.basinfo)
        self.ftab = _d(ftab)
        self.n = basis.norb
        self.nset = basis.nset
        self.nnuc = len(Z)

    def args(self):
        return (self.nnuc, _P(self.xyzF), self.nset, _P(self.alpha), _P(self.setinfo, _ip), _P(self.bas),
                self.n, _P(self.basinfo, _ip), _P(self.ftab))


def boys(Q, T, ftab):
    F = np.zeros(17)
    rc = lib().oracle_boys(Q, float(T), _P(_d(ftab)), _P(F))
    if rc != 0:
        raise ValueError(f"oracle_boys rc={rc}")
    return F[: Q + 1]


def eri_full(sysm: System) -> np.ndarray:
    n = sysm.n
    XX = np.zeros(n ** 4)
    rc = lib().oracle_eri_full(*sysm.args(), _P(XX))
    if rc != 0:
        raise RuntimeError("oracle_eri_full failed")
    return XX.reshape(n, n, n, n, order="F")


def eri_sample_time(sysm: System, npairs: int, seed: int = 0):
    """Time the reference's per-bra-set-pair work for a random sample of
    ``npairs`` ordered bra set pairs.  Returns (seconds, set quartets, scale)
    where scale = nset^2 / npairs extrapolates to the whole proc2e loop."""
    rng = np.random.default_rng(seed)
    ns = sysm.nset
    sel = rng.choice(ns * ns, size=min(npairs, ns * ns), replace=False)
    ab = _i(np.stack([sel // ns, sel % ns], axis=1).ravel())
    cs = ctypes.c_double(0.0)
    t0 = time.perf_counter()
    cnt = lib().oracle_eri_sample(*sysm.args(), len(sel), _P(ab, _ip), ctypes.byref(cs))
    t1 = time.perf_counter()
    return t1 - t0, int(cnt), ns * ns / len(sel)


def int1e(sysm: System):
    n = sysm.n
    S = np.zeros(n * n)
    H = np.zeros(n * n)
    a = sysm.args()
    rc = lib().oracle_int1e(sysm.nnuc, _P(sysm.Z, _ip), *a[1:], _P(S), _P(H))
    if rc != 0:
        raise RuntimeError("oracle_int1e failed")
    return S.reshape(n, n, order="F"), H.reshape(n, n, order="F")


def fock_G(XX, D):
    n = D.shape[0]
    XXf = _d(np.asarray(XX).ravel(order="F"))
    Df = _d(np.asarray(D).ravel(order="F"))
    G = np.zeros(n * n)
    lib().oracle_fock_G(n, _P(XXf), _P(Df), _P(G))
    return G.reshape(n, n, order="F")


def density(C, nocc):
    n = C.shape[0]
    Cf = _d(np.asarray(C).ravel(order="F"))
    D = np.zeros(n * n)
    lib().oracle_density(n, _P(Cf), nocc, _P(D))
    return D.reshape(n, n, order="F")


def ao2mo_ovov(XX, C, nocc):
    """(ia|jb) by the reference's four quarter steps (ao2mo.f90:506,524,543,562)."""
    Co = C[:, :nocc]
    Cv = C[:, nocc:]
    L = np.einsum("tqrs,tp->pqrs", XX, Co, optimize=True)
    M = np.einsum("ptrs,tq->pqrs", L, Cv, optimize=True)
    N = np.einsum("pqts,tr->pqrs", M, Co, optimize=True)
    return np.einsum("pqrt,ts->pqrs", N, Cv, optimize=True)


def mp2(iajb, eig, nocc):
    """iajb[i,a,j,b] = (ia|jb).  Returns (E(AA), E(AB), E(MBPT2))."""
    nv = iajb.shape[1]
    rec = np.empty((nocc, nocc, nv, nv))
    for i in range(nocc):
        for j in range(nocc):
            rec[i, j] = iajb[i, :, j, :]
    # record (i,j) holds (a,b) with a fastest
    flat = _d(np.stack([rec[i, j].ravel(order="F") for i in range(nocc) for j in range(nocc)]).ravel())
    ea = ctypes.c_double()
    eb = ctypes.c_double()
    lib().oracle_mp2(nocc, nv, _P(_d(eig)), _P(flat), ctypes.byref(ea), ctypes.byref(eb))
    return ea.value, eb.value, 2 * ea.value + eb.value


def enuc(Z, xyz):
    """getEnr (src/scf/scf.f90:1122-1152)."""
    e = 0.0
    for a in range(len(Z)):
        for b in range(a):
            r = np.sqrt((xyz[a, 0] - xyz[b, 0]) ** 2 + (xyz[a, 1] - xyz[b, 1]) ** 2 + (xyz[a, 2] - xyz[b, 2]) ** 2)
            e += Z[a] * Z[b] / r
    return e


def dsygv(A, B):
    """DSYGV(1,'V','L') on the lower triangles (scf.f90:450-478,802-848)."""
    import scipy.linalg
    return scipy.linalg.eigh(np.tril(A), np.tril(B), lower=True, type=1, driver="gv")


def energy(D, F, H):
    """E_elec = 1/2 sum D (F + H) (scf.f90:775-781)."""
    return 0.5 * float(np.sum(D * (F + H)))


def rhf(S, H, XX, nocc, Enuc, conv=11, diis=False, max_iter=500, C0=None):
    """RHF SCF on the oracle's G build.  Without DIIS this is scf.f90's loop:
    core-H guess, then Roothaan steps until max|D_new - D_old| < 10^-conv, the
    reported energy being E[D_old] (scf.f90:775-788,1227-1258)."""
    if C0 is None:
        _, C = dsygv(H, S)
    else:
        C = C0
    D = density(C, nocc)
    G = fock_G(XX, D)
    lim = 10.0 ** (-conv)
    hist_F, hist_e = [], []
    energies = []
    it = 0
    eig = None
    while True:
        F = H + G
        E = energy(D, F, H) + Enuc
        energies.append(E)
        Fx = F
        if diis:
            err = F @ D @ S - S @ D @ F
            hist_F.append(F.copy())
            hist_e.append(err)
            if len(hist_F) > 8:
                hist_F.pop(0)
                hist_e.pop(0)
            m = len(hist_F)
            if m >= 2:
                B = -np.ones((m + 1, m + 1))
                B[m, m] = 0.0
                for i in range(m):
                    for j in range(m):
                        B[i, j] = np.sum(hist_e[i] * hist_e[j])
                rhs = np.zeros(m + 1)
                rhs[m] = -1.0
                try:
                    c = np.linalg.solve(B, rhs)
                    Fx = sum(c[i] * hist_F[i] for i in range(m))
                except np.linalg.LinAlgError:
                    Fx = F
        eig, C = dsygv(Fx, S)
        Dn = density(C, nocc)
        G = fock_G(XX, Dn)
        mdiff = float(np.max(np.abs(D - Dn)))
        D = Dn
        it += 1
        if it > 1 and mdiff < lim:
            break
        if it >= max_iter:
            break
    if diis:
        # report the energy and orbitals of the converged density
        F = H + G
        energies.append(energy(D, F, H) + Enuc)
        eig, C = dsygv(F, S)
    return dict(E=energies[-1], energies=np.array(energies), C=C, eig=eig, D=D, iterations=it)

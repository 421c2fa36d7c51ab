"""Python host mirror of the C ABI (include/myqc_hip.h) over ctypes.

Each method names the reference program it replaces.  Arrays are numpy
float64 in Fortran order, exactly as the Fortran host passes them.  There is
no CPU fallback: if libmyqc_hip.so is missing, or no gfx950 device is usable,
construction raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .basis import Basis, build_basis, read_ftab

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmyqc_hip.so")

FP64, FP32, BF16X3 = 0, 1, 2
T_ERI, T_JK, T_AO2MO, T_MP2 = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

EXPORTS = {
    "myqc_init": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "myqc_finalize": ([_vp], ctypes.c_int),
    "myqc_last_error": ([_vp], ctypes.c_char_p),
    "myqc_set_system": ([_vp, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_int, _ip, _dp,
                         ctypes.c_int], ctypes.c_int),
    "myqc_eri_full": ([_vp, ctypes.c_double, _dp], ctypes.c_int),
    "myqc_eri_build": ([_vp, ctypes.c_double, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "myqc_eri_tune": ([_vp, ctypes.c_int, _ip, _dp], ctypes.c_int),
    "myqc_eri_set_shapes": ([_vp, _ip], ctypes.c_int),
    "myqc_fock_G": ([_vp, _dp, _dp, ctypes.c_int], ctypes.c_int),
    "myqc_fock_G_uhf": ([_vp, _dp, _dp, _dp, _dp], ctypes.c_int),
    "myqc_density_rhf": ([_vp, _dp, ctypes.c_int, _dp], ctypes.c_int),
    "myqc_ao2mo_ovov": ([_vp, _dp, ctypes.c_int, _dp, ctypes.c_int], ctypes.c_int),
    "myqc_mp2_rhf": ([_vp, _dp, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
    "myqc_mp2_rhf_records": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp], ctypes.c_int),
    "myqc_mp2_rhf_records_noalloc": ([_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "myqc_int1e": ([_vp, _dp, _dp], ctypes.c_int),
    "myqc_int1p": ([_vp, _dp], ctypes.c_int),
    "myqc_dipole_rhf": ([_vp, _dp, _dp], ctypes.c_int),
    "myqc_density_uhf": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp], ctypes.c_int),
    "myqc_uhf_scf": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, ctypes.c_double,
                      ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _ip], ctypes.c_int),
    "myqc_ao2mo_ovov_uhf": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int], ctypes.c_int),
    "myqc_mp2_uhf": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp], ctypes.c_int),
    "myqc_ao2mo_cis_uhf": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp], ctypes.c_int),
    "myqc_cis_uhf": ([_vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, _dp, _dp], ctypes.c_int),
    "myqc_mp2_uhf_records": ([_vp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              _dp, _dp, _dp, _dp], ctypes.c_int),
    "myqc_rhf_scf": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double,
                      ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip], ctypes.c_int),
    "myqc_rhf_scf_density": ([_vp, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double,
                              ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip], ctypes.c_int),
    "myqc_scf_info": ([_vp, _ip, _ip], ctypes.c_int),
    "myqc_rhf_scf_many": ([ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_dp), ctypes.POINTER(_dp), _ip, _dp,
                           ctypes.POINTER(_dp), ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp,
                           ctypes.POINTER(_dp), ctypes.POINTER(_dp), _ip], ctypes.c_int),
    "myqc_batch_set": ([_vp, ctypes.c_int, _ip, _ip, _ip, _ip, _dp, _dp, _ip, _dp, _ip, _ip, _dp], ctypes.c_int),
    "myqc_batch_rhf": ([_vp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _ip, _ip], ctypes.c_int),
    "myqc_batch_density": ([_vp, _dp], ctypes.c_int),
    "myqc_batch_info": ([_vp, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "myqc_mlp_create": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong], ctypes.c_int),
    "myqc_mlp_params": ([_vp, _fp, _fp, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "myqc_mlp_forward": ([_vp, _fp, ctypes.c_int, _fp], ctypes.c_int),
    "myqc_mlp_gradient": ([_vp, _fp, _fp, _fp, ctypes.c_int, _fp, _fp], ctypes.c_int),
    "myqc_mlp_train": ([_vp, _fp, _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                        ctypes.c_ulonglong, _fp], ctypes.c_int),
    "myqc_batch_pair_data": ([_vp, _ip, _fp, _fp, _fp], ctypes.c_int),
    "myqc_guess_density": ([_vp, ctypes.c_int, _ip, _dp, _dp], ctypes.c_int),
    "myqc_batch_rhf_learned": ([_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, _ip, _ip, _ip],
                               ctypes.c_int),
    "myqc_timer_reset": ([_vp, ctypes.c_int], ctypes.c_int),
    "myqc_timer_get": ([_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _dp, _dp], ctypes.c_int),
}

_LIB = None


class MyqcError(RuntimeError):
    """A nonzero return code from the C ABI (the reference would `touch error`
    and STOP, src/myQC/env.f90:36-41)."""

    def __init__(self, code, msg):
        super().__init__(f"myqc error {code}: {msg}")
        self.code = code


def load(path: str | None = None) -> ctypes.CDLL:
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"{p} not built: run `python -m myqc_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(p)
    for name, (argt, rest) in EXPORTS.items():
        fn = getattr(L, name)
        fn.argtypes = argt
        fn.restype = rest
    if path is None:
        _LIB = L
    return L


def _f(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class Context:
    """One device, one stream (myqc_init / myqc_finalize)."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = _vp()
        rc = self.L.myqc_init(device, ctypes.byref(h))
        if rc != 0:
            raise MyqcError(rc, f"myqc_init(device={device}) failed: no usable gfx950 device")
        self.h = h
        self.n = 0
        self.basis: Basis | None = None

    def _chk(self, rc):
        if rc != 0:
            raise MyqcError(rc, self.L.myqc_last_error(self.h).decode())

    def close(self):
        if self.h is not None:
            self.L.myqc_finalize(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs ----------------------------------------------------------
    def set_system(self, Z, xyz_bohr, basis: Basis | None = None, ftab=None):
        basis = basis or build_basis(list(Z))
        ftab = read_ftab() if ftab is None else np.asarray(ftab, dtype=np.float64)
        self._Z = np.ascontiguousarray(Z, dtype=np.int32)
        self._xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3).ravel(order="F").copy()
        self._alpha = np.ascontiguousarray(basis.set_alpha, dtype=np.float64)
        self._setinfo = np.ascontiguousarray(basis.setinfo, dtype=np.int32)
        self._bas = np.ascontiguousarray(basis.bas, dtype=np.float64)
        self._basinfo = np.ascontiguousarray(basis.basinfo, dtype=np.int32)
        self._ftab = np.ascontiguousarray(ftab, dtype=np.float64)
        self._chk(self.L.myqc_set_system(self.h, len(self._Z), _p(self._Z, _ip), _p(self._xyz), basis.nset,
                                         _p(self._alpha), _p(self._setinfo, _ip), _p(self._bas), basis.norb,
                                         _p(self._basinfo, _ip), _p(self._ftab), 0))
        self.n = basis.norb
        self.basis = basis

    # -- drop-ins ----------------------------------------------------------
    def eri_full(self, schwarz_tol: float = 0.0) -> np.ndarray:
        """int2e: dense 8-fold filled XX(i,j,g,h)."""
        n = self.n
        XX = np.zeros(n ** 4)
        self._chk(self.L.myqc_eri_full(self.h, schwarz_tol, _p(XX)))
        return XX.reshape(n, n, n, n, order="F")

    def eri_build(self, schwarz_tol: float = 1e-12) -> int:
        ns = ctypes.c_longlong()
        self._chk(self.L.myqc_eri_build(self.h, schwarz_tol, ctypes.byref(ns)))
        return ns.value

    def eri_tune(self, reps: int = 2):
        """Times the ERI launch shapes on the current system and keeps the
        fastest per class for the process; returns ({class: (lanes,
        threads)}, {class: ms})."""
        sh = np.zeros(12, dtype=np.int32)
        ms = np.zeros(6)
        self._chk(self.L.myqc_eri_tune(self.h, reps, _p(sh, _ip), _p(ms)))
        return {c: (int(sh[2 * c]), int(sh[2 * c + 1])) for c in range(6)}, {c: float(ms[c]) for c in range(6)}

    def eri_set_shapes(self, shapes):
        """shapes: {class: (lanes, threads)} (classes left out keep theirs)."""
        sh = np.zeros(12, dtype=np.int32)
        for c, (sp, tpb) in shapes.items():
            sh[2 * c], sh[2 * c + 1] = sp, tpb
        self._chk(self.L.myqc_eri_set_shapes(self.h, _p(sh, _ip)))

    def fock_G(self, D, precision: int = FP64) -> np.ndarray:
        """RHFI2G: G from the stored integrals."""
        n = self.n
        Df = _f(D)
        G = np.zeros((n, n), order="F")
        self._chk(self.L.myqc_fock_G(self.h, _p(Df), _p(G), precision))
        return G

    def fock_G_uhf(self, Da, Db):
        """UHFI2G: the two spin G matrices from the stored integrals."""
        n = self.n
        Ga = np.zeros((n, n), order="F")
        Gb = np.zeros((n, n), order="F")
        self._chk(self.L.myqc_fock_G_uhf(self.h, _p(_f(Da)), _p(_f(Db)), _p(Ga), _p(Gb)))
        return Ga, Gb

    def density(self, C, nocc: int) -> np.ndarray:
        n = self.n
        Cf = _f(C)
        D = np.zeros((n, n), order="F")
        self._chk(self.L.myqc_density_rhf(self.h, _p(Cf), nocc, _p(D)))
        return D

    def int1e(self):
        n = self.n
        S = np.zeros((n, n), order="F")
        H = np.zeros((n, n), order="F")
        self._chk(self.L.myqc_int1e(self.h, _p(S), _p(H)))
        return S, H

    def int1p(self):
        """int1p: the dipole integral matrices, shape (3, n, n) (x, y, z)."""
        n = self.n
        P = np.zeros(3 * n * n)
        self._chk(self.L.myqc_int1p(self.h, _p(P)))
        return np.array([P[c * n * n:(c + 1) * n * n].reshape(n, n, order="F") for c in range(3)])

    def dipole(self, D):
        """cphf_dipole_rhf: the dipole moment (a.u.) of the total density D."""
        mu = np.zeros(3)
        self._chk(self.L.myqc_dipole_rhf(self.h, _p(_f(D)), _p(mu)))
        return mu

    def ao2mo(self, C, nocc: int, precision: int = FP64, download: bool = True):
        """ao2mo: returns iajb[i, a, j, b] = (ia|jb) (or None if not downloaded)."""
        n = self.n
        nv = n - nocc
        Cf = _f(C)
        if download:
            rec = np.zeros(nocc * nocc * nv * nv)
            self._chk(self.L.myqc_ao2mo_ovov(self.h, _p(Cf), nocc, _p(rec), precision))
            # record (i,j): (a,b) with a fastest
            r = rec.reshape(nocc, nocc, nv, nv)          # [i][j][b][a]
            return np.ascontiguousarray(r.transpose(0, 3, 1, 2))   # [i][a][j][b]
        self._chk(self.L.myqc_ao2mo_ovov(self.h, _p(Cf), nocc, None, precision))
        return None

    def mp2(self, eig, nocc: int):
        """mp2: (E(AA), E(AB), E(MBPT2))."""
        e = np.ascontiguousarray(eig, dtype=np.float64)
        a, b, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.myqc_mp2_rhf(self.h, _p(e), nocc, ctypes.byref(a), ctypes.byref(b), ctypes.byref(t)))
        return a.value, b.value, t.value

    def mp2_records(self, iajb, eig, nocc: int):
        """mp2 on host (ia|jb), iajb[i, a, j, b]: uploads ijab_AB-ordered records."""
        nv = iajb.shape[1]
        rec = np.ascontiguousarray(np.asarray(iajb, dtype=np.float64).transpose(0, 2, 3, 1))  # [i][j][b][a]
        e = np.ascontiguousarray(eig, dtype=np.float64)
        a, b, t = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.myqc_mp2_rhf_records(self.h, _p(rec), _p(e), nocc, nv, ctypes.byref(a), ctypes.byref(b),
                                              ctypes.byref(t)))
        return a.value, b.value, t.value

    def rhf(self, H, S, nocc: int, e_nuc: float, C_guess=None, e_tol=1e-10, d_tol=1e-8, max_iter=200,
            precision: int = FP64, D_guess=None):
        """In-process RHF from the core-Hamiltonian guess, orbitals C_guess or
        an AO density D_guess; returns dict(E, C, eig, iterations, ...)."""
        n = self.n
        Hf, Sf = _f(H), _f(S)
        C = np.zeros((n, n), order="F")
        eig = np.zeros(n)
        E = ctypes.c_double()
        it = ctypes.c_int()
        if D_guess is not None:
            Dg = _f(D_guess)
            self._chk(self.L.myqc_rhf_scf_density(self.h, _p(Hf), _p(Sf), nocc, e_nuc, _p(Dg), e_tol, d_tol,
                                                  max_iter, precision, ctypes.byref(E), _p(C), _p(eig),
                                                  ctypes.byref(it)))
        else:
            Cg = _f(C_guess) if C_guess is not None else None
            self._chk(self.L.myqc_rhf_scf(self.h, _p(Hf), _p(Sf), nocc, e_nuc, _p(Cg) if Cg is not None else None,
                                          e_tol, d_tol, max_iter, precision, ctypes.byref(E), _p(C), _p(eig),
                                          ctypes.byref(it)))
        npur, ndiag = ctypes.c_int(), ctypes.c_int()
        self._chk(self.L.myqc_scf_info(self.h, ctypes.byref(npur), ctypes.byref(ndiag)))
        return dict(E=E.value, C=C, eig=eig, iterations=it.value, purifications=npur.value,
                    diag_fallbacks=ndiag.value)

    # -- UHF ----------------------------------------------------------------

    def density_uhf(self, Ca, Cb, nalpha: int, nbeta: int):
        n = self.n
        Da = np.zeros((n, n), order="F")
        Db = np.zeros((n, n), order="F")
        self._chk(self.L.myqc_density_uhf(self.h, _p(_f(Ca)), _p(_f(Cb)), nalpha, nbeta, _p(Da), _p(Db)))
        return Da, Db

    def uhf(self, H, S, nalpha: int, nbeta: int, e_nuc: float, Ca_guess=None, Cb_guess=None, e_tol=1e-10,
            d_tol=1e-8, max_iter=300):
        """In-process UHF from the core-Hamiltonian orbitals or (Ca_guess,
        Cb_guess); returns dict(E, Ca, Cb, eig_a, eig_b, iterations)."""
        n = self.n
        Ca = np.zeros((n, n), order="F")
        Cb = np.zeros((n, n), order="F")
        ea = np.zeros(n)
        eb = np.zeros(n)
        E = ctypes.c_double()
        it = ctypes.c_int()
        ga = _f(Ca_guess) if Ca_guess is not None else None
        gb = _f(Cb_guess) if Cb_guess is not None else None
        self._chk(self.L.myqc_uhf_scf(self.h, _p(_f(H)), _p(_f(S)), nalpha, nbeta, e_nuc,
                                      _p(ga) if ga is not None else None, _p(gb) if gb is not None else None,
                                      e_tol, d_tol, max_iter, ctypes.byref(E), _p(Ca), _p(Cb), _p(ea), _p(eb),
                                      ctypes.byref(it)))
        return dict(E=E.value, Ca=Ca, Cb=Cb, eig_a=ea, eig_b=eb, iterations=it.value)

    @staticmethod
    def _uhf_shapes(n, nalpha, nbeta):
        va, vb = n - nalpha, n - nbeta
        return {"AA": (nalpha * (nalpha - 1) // 2 if nalpha > 0 else 0, va, va),
                "BB": (nbeta * (nbeta - 1) // 2 if nbeta > 0 else 0, vb, vb),
                "AB": (nalpha * nbeta, vb, va)}

    def ao2mo_uhf(self, Ca, Cb, nalpha: int, nbeta: int, precision: int = FP64, download: bool = True):
        """slow_ao2mo_MP2_UHF: the record streams of ijab_AA, ijab_BB and
        ijab_AB, each shaped (records, nvrt', nvrt) so that [p, b, a] is
        (ia|jb) of record p (or None if not downloaded)."""
        shp = self._uhf_shapes(self.n, nalpha, nbeta)
        if not download:
            self._chk(self.L.myqc_ao2mo_ovov_uhf(self.h, _p(_f(Ca)), _p(_f(Cb)), nalpha, nbeta, None, None, None,
                                                 precision))
            return None
        out = {k: np.zeros(max(int(np.prod(v)), 1)) for k, v in shp.items()}
        self._chk(self.L.myqc_ao2mo_ovov_uhf(self.h, _p(_f(Ca)), _p(_f(Cb)), nalpha, nbeta, _p(out["AA"]),
                                             _p(out["BB"]), _p(out["AB"]), precision))
        return {k: out[k][:int(np.prod(shp[k]))].reshape(shp[k]) for k in shp}

    def mp2_uhf(self, eig_a, eig_b, nalpha: int, nbeta: int):
        """mp2_uhf on the sets of the last ao2mo_uhf: (E(AA), E(BB), E(AB), E(MBPT2))."""
        ea = np.ascontiguousarray(eig_a, dtype=np.float64)
        eb = np.ascontiguousarray(eig_b, dtype=np.float64)
        r = [ctypes.c_double() for _ in range(4)]
        self._chk(self.L.myqc_mp2_uhf(self.h, _p(ea), _p(eb), nalpha, nbeta, *[ctypes.byref(x) for x in r]))
        return tuple(x.value for x in r)

    def mp2_uhf_records(self, recs, eig_a, eig_b, nalpha: int, nbeta: int):
        """mp2_uhf on host record streams (as ao2mo_uhf returns them)."""
        n = len(eig_a)
        ea = np.ascontiguousarray(eig_a, dtype=np.float64)
        eb = np.ascontiguousarray(eig_b, dtype=np.float64)
        rr = [np.ascontiguousarray(recs[k], dtype=np.float64).ravel() for k in ("AA", "BB", "AB")]
        rr = [x if x.size else np.zeros(1) for x in rr]
        r = [ctypes.c_double() for _ in range(4)]
        self._chk(self.L.myqc_mp2_uhf_records(self.h, _p(rr[0]), _p(rr[1]), _p(rr[2]), _p(ea), _p(eb), nalpha,
                                              nbeta, n - nalpha, n - nbeta, *[ctypes.byref(x) for x in r]))
        return tuple(x.value for x in r)

    CIS_STREAMS = ("ajib_AA", "ajbi_AA", "ajib_AB", "ajib_BB", "ajbi_BB")

    def ao2mo_cis_uhf(self, Ca, Cb, nalpha: int, nbeta: int, download: bool = True):
        """slow_ao2mo_CIS_UHF: the five record streams, each shaped
        (records, elements) = (upS', upS) (or None if not downloaded)."""
        n = self.n
        upA, upB = nalpha * (n - nalpha), nbeta * (n - nbeta)
        shp = {"ajib_AA": (upA, upA), "ajbi_AA": (upA, upA), "ajib_AB": (upB, upA), "ajib_BB": (upB, upB),
               "ajbi_BB": (upB, upB)}
        if not download:
            self._chk(self.L.myqc_ao2mo_cis_uhf(self.h, _p(_f(Ca)), _p(_f(Cb)), nalpha, nbeta, None, None, None, None,
                                                None))
            return None
        out = {k: np.zeros(max(v[0] * v[1], 1)) for k, v in shp.items()}
        self._chk(self.L.myqc_ao2mo_cis_uhf(self.h, _p(_f(Ca)), _p(_f(Cb)), nalpha, nbeta,
                                            *[_p(out[k]) for k in self.CIS_STREAMS]))
        return {k: out[k][:shp[k][0] * shp[k][1]].reshape(shp[k]) for k in shp}

    def cis_uhf(self, eig_a, eig_b, nalpha: int, nbeta: int, nroots: int, streams=None):
        """cis_scf_uhf: the lowest nroots excitation energies and eigenvectors
        ((upA + upB) x nroots), from the given record streams or the sets of
        the last ao2mo_cis_uhf."""
        n = len(eig_a)
        dim = nalpha * (n - nalpha) + nbeta * (n - nbeta)
        k = min(nroots, dim)
        E = np.zeros(max(k, 1))
        V = np.zeros((dim, max(k, 1)), order="F")
        ea = np.ascontiguousarray(eig_a, dtype=np.float64)
        eb = np.ascontiguousarray(eig_b, dtype=np.float64)
        if streams is None:
            ptrs = [None] * 5
        else:
            arrs = [np.ascontiguousarray(streams[s], dtype=np.float64).ravel() for s in self.CIS_STREAMS]
            arrs = [a if a.size else np.zeros(1) for a in arrs]
            ptrs = [_p(a) for a in arrs]
        self._chk(self.L.myqc_cis_uhf(self.h, *ptrs, _p(ea), _p(eb), nalpha, nbeta, n - nalpha, n - nbeta, nroots,
                                      _p(E), _p(V)))
        return E[:k], V[:, :k]

    # -- batched small molecules -------------------------------------------
    def batch_set(self, molecules, ftab=None):
        """molecules: list of (Z, xyz_bohr[, charge]); each one's reference
        basis arrays (buildBasis layouts) are concatenated with offsets."""
        ftab = read_ftab() if ftab is None else np.asarray(ftab, dtype=np.float64)
        nnuc, norb, nocc, Z, xyz, al, si, bs, bi, off = [], [], [], [], [], [], [], [], [], []
        o = [0, 0, 0, 0]
        for mol in molecules:
            z, x = list(mol[0]), np.asarray(mol[1], dtype=np.float64).reshape(-1, 3)
            charge = mol[2] if len(mol) > 2 else 0
            b = build_basis(z)
            nnuc.append(len(z))
            norb.append(b.norb)
            nocc.append((sum(z) - charge) // 2)
            Z.extend(z)
            xyz.append(x.ravel(order="F"))
            off.extend(o)
            for k, arr in enumerate((b.set_alpha, b.setinfo, b.bas, b.basinfo)):
                (al, si, bs, bi)[k].append(np.asarray(arr))
                o[k] += len(arr)
        self._b = dict(
            nnuc=np.array(nnuc, np.int32), norb=np.array(norb, np.int32), nocc=np.array(nocc, np.int32),
            Z=np.array(Z, np.int32), xyz=np.concatenate(xyz), al=np.concatenate(al).astype(np.float64),
            si=np.concatenate(si).astype(np.int32), bs=np.concatenate(bs).astype(np.float64),
            bi=np.concatenate(bi).astype(np.int32), off=np.array(off, np.int32),
            ftab=np.ascontiguousarray(ftab, dtype=np.float64))
        d = self._b
        self._chk(self.L.myqc_batch_set(self.h, len(molecules), _p(d["nnuc"], _ip), _p(d["norb"], _ip),
                                        _p(d["nocc"], _ip), _p(d["Z"], _ip), _p(d["xyz"]), _p(d["al"]),
                                        _p(d["si"], _ip), _p(d["bs"]), _p(d["bi"], _ip), _p(d["off"], _ip),
                                        _p(d["ftab"])))
        self.batch_norb = d["norb"].copy()
        self.batch_nocc = d["nocc"].copy()

    def batch_rhf(self, D_guess=None, e_tol=1e-10, d_tol=1e-8, max_iter=100):
        """One persistent launch over the batch; D_guess: list of n_m x n_m
        densities (the learned guess) or None for the core-Hamiltonian guess.
        Returns (energies, iterations, converged)."""
        nm = len(self.batch_norb)
        E = np.zeros(nm)
        it = np.zeros(nm, np.int32)
        cv = np.zeros(nm, np.int32)
        g = None
        if D_guess is not None:
            g = np.concatenate([np.asarray(D, dtype=np.float64).ravel(order="F") for D in D_guess])
        self._chk(self.L.myqc_batch_rhf(self.h, None if g is None else _p(g), e_tol, d_tol, max_iter, _p(E),
                                        _p(it, _ip), _p(cv, _ip)))
        return E, it, cv

    def batch_density(self):
        """The densities of the last batch_rhf, one n_m x n_m array each."""
        tot = int(np.sum(self.batch_norb.astype(np.int64) ** 2))
        D = np.zeros(tot)
        self._chk(self.L.myqc_batch_density(self.h, _p(D)))
        out, o = [], 0
        for n in self.batch_norb:
            out.append(D[o:o + n * n].reshape(n, n, order="F"))
            o += n * n
        return out

    def batch_info(self):
        eri = np.zeros(6, np.int32)
        scf, rep, nint = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        self._chk(self.L.myqc_batch_info(self.h, _p(eri, _ip), ctypes.byref(scf), ctypes.byref(rep),
                                         ctypes.byref(nint)))
        return dict(eri_threads=eri.tolist(), scf_threads=scf.value, graph_replays=rep.value,
                    n_integrals=nint.value)

    # -- learned guess --------------------------------------------------------
    def mlp_create(self, n_in: int, width: int, n_out: int, seed: int = 0):
        self._chk(self.L.myqc_mlp_create(self.h, n_in, width, n_out, seed))

    def mlp_params(self, set_to=None):
        cnt = ctypes.c_longlong()
        self._chk(self.L.myqc_mlp_params(self.h, None, None, ctypes.byref(cnt)))
        if set_to is not None:
            p = np.ascontiguousarray(set_to, dtype=np.float32)
            assert p.size == cnt.value, (p.size, cnt.value)
            self._chk(self.L.myqc_mlp_params(self.h, None, _p(p, _fp), None))
        out = np.zeros(cnt.value, np.float32)
        self._chk(self.L.myqc_mlp_params(self.h, _p(out, _fp), None, None))
        return out

    def mlp_forward(self, X, n_out: int):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], n_out), np.float32)
        self._chk(self.L.myqc_mlp_forward(self.h, _p(X, _fp), X.shape[0], _p(Y, _fp)))
        return Y

    def mlp_gradient(self, X, T, M):
        X, T, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M))
        cnt = ctypes.c_longlong()
        self._chk(self.L.myqc_mlp_params(self.h, None, None, ctypes.byref(cnt)))
        g = np.zeros(cnt.value, np.float32)
        loss = ctypes.c_float()
        self._chk(self.L.myqc_mlp_gradient(self.h, _p(X, _fp), _p(T, _fp), _p(M, _fp), X.shape[0],
                                           ctypes.byref(loss), _p(g, _fp)))
        return loss.value, g

    def mlp_train(self, X, T, M, epochs: int, batch: int, lr: float, seed: int = 0):
        X, T, M = (np.ascontiguousarray(a, dtype=np.float32) for a in (X, T, M))
        hist = np.zeros(epochs, np.float32)
        self._chk(self.L.myqc_mlp_train(self.h, _p(X, _fp), _p(T, _fp), _p(M, _fp), X.shape[0], epochs, batch, lr,
                                        seed, _p(hist, _fp)))
        return hist

    def guess_density(self, Z, xyz_bohr):
        """The network's density for one molecule (computed on the device)."""
        Zc = np.ascontiguousarray(Z, dtype=np.int32)
        x = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3).ravel(order="F").copy()
        n = int(sum(1 if z <= 2 else 5 for z in Zc))
        D = np.zeros((n, n), order="F")
        self._chk(self.L.myqc_guess_density(self.h, len(Zc), _p(Zc, _ip), _p(x), _p(D)))
        return D

    def batch_pair_data(self, n_feat: int = 200, n_out: int = 16):
        npair = ctypes.c_int()
        self._chk(self.L.myqc_batch_pair_data(self.h, ctypes.byref(npair), None, None, None))
        X = np.zeros((npair.value, n_feat), np.float32)
        T = np.zeros((npair.value, n_out), np.float32)
        M = np.zeros((npair.value, n_out), np.float32)
        self._chk(self.L.myqc_batch_pair_data(self.h, ctypes.byref(npair), _p(X, _fp), _p(T, _fp), _p(M, _fp)))
        return X, T, M

    def batch_rhf_learned(self, e_tol=1e-10, d_tol=1e-8, max_iter=100):
        """Batch SCF from the network's densities, with core-H fallback.
        Returns (energies, iterations, converged, fallback)."""
        nm = len(self.batch_norb)
        E = np.zeros(nm)
        it, cv, fb = (np.zeros(nm, np.int32) for _ in range(3))
        self._chk(self.L.myqc_batch_rhf_learned(self.h, e_tol, d_tol, max_iter, _p(E), _p(it, _ip), _p(cv, _ip),
                                                _p(fb, _ip)))
        return E, it, cv, fb

    def timer_reset(self, enable: bool = True):
        self._chk(self.L.myqc_timer_reset(self.h, 1 if enable else 0))

    def timer_get(self, which: int):
        n = ctypes.c_longlong()
        ms = ctypes.c_double()
        w = ctypes.c_double()
        self._chk(self.L.myqc_timer_get(self.h, which, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(w)))
        return n.value, ms.value, w.value


def enuc(Z, xyz) -> float:
    """Nuclear repulsion (src/scf/scf.f90:1122-1152)."""
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    z = np.asarray(Z, dtype=np.float64)
    a, b = np.tril_indices(len(z), -1)
    r = np.sqrt(((xyz[a] - xyz[b]) ** 2).sum(axis=1))
    return float(np.sum(z[a] * z[b] / r))


def rhf_many(ctxs, Hs, Ss, noccs, e_nucs, D_guesses=None, e_tol=1e-10, d_tol=1e-8, max_iter=200,
             precision: int = FP64, want_orbitals: bool = True):
    """RHF of several systems of one basis size in one lockstep solve
    (myqc_rhf_scf_many): ctxs[k] holds system k (set_system and eri_build
    done); D_guesses[k] an AO density or None for the core guess.  Returns a
    list of dict(E, C, eig, iterations); raises MyqcError when a system did
    not converge."""
    K = len(ctxs)
    L = ctxs[0].L
    n = ctxs[0].n
    hs = (_vp * K)(*[c.h for c in ctxs])
    Hf = [_f(H) for H in Hs]
    Sf = [_f(S) for S in Ss]
    Dg = [None if (D_guesses is None or D_guesses[k] is None) else _f(D_guesses[k]) for k in range(K)]
    Cs = [np.zeros((n, n), order="F") for _ in range(K)] if want_orbitals else None
    eigs = [np.zeros(n) for _ in range(K)] if want_orbitals else None
    arr = lambda xs: (_dp * K)(*[_p(x) if x is not None else None for x in xs])
    E = np.zeros(K)
    it = np.zeros(K, dtype=np.int32)
    occ = np.ascontiguousarray(noccs, dtype=np.int32)
    en = np.ascontiguousarray(e_nucs, dtype=np.float64)
    rc = L.myqc_rhf_scf_many(hs, K, arr(Hf), arr(Sf), _p(occ, _ip), _p(en), arr(Dg), e_tol, d_tol, max_iter,
                             precision, _p(E), arr(Cs) if Cs else None, arr(eigs) if eigs else None, _p(it, _ip))
    ctxs[0]._chk(rc)
    return [dict(E=float(E[k]), C=Cs[k] if Cs else None, eig=eigs[k] if eigs else None, iterations=int(it[k]))
            for k in range(K)]

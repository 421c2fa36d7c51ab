"""myqc_mp2_rhf_records_noalloc, the entry point that allocates nothing
(device-resident records, energies and results; the partial sums in the
library's fixed per-GPU scratch), against myqc_mp2_rhf_records, the entry
point it varies, on two calls in a row (the second smaller), through the C
ABI; and its size limit.  Reference sums: src/mp2/mp2.f90:79-144."""
import ctypes

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

_hip = None


def _rt():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
    return _hip


def _to_device(a):
    rt = _rt()
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
    assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0        # host to device
    return p


def _from_device(p, count):
    out = np.zeros(count)
    assert _rt().hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
    return out


def _records_of(iajb):
    """ijab_AB record stream (record (i,j) i-major, a fastest) of iajb[i,a,j,b]."""
    return np.ascontiguousarray(np.asarray(iajb).transpose(0, 2, 3, 1)).ravel()


def _both(ctx, rec, eig, nocc, nvrt):
    e_ref = ctx.mp2_records(np.ascontiguousarray(rec).reshape(nocc, nocc, nvrt, nvrt).transpose(0, 3, 1, 2),
                            eig, nocc)
    d_rec, d_eig, d_out = _to_device(rec), _to_device(eig), _to_device(np.full(3, np.nan))
    try:
        rc = ctx.L.myqc_mp2_rhf_records_noalloc(ctx.h, d_rec, d_eig, nocc, nvrt, d_out)
        assert rc == 0, ctx.L.myqc_last_error(ctx.h).decode()
        got = _from_device(d_out, 3)
    finally:
        for p in (d_rec, d_eig, d_out):
            _rt().hipFree(p)
    return np.array(e_ref), got


def test_noalloc_mp2_matches_allocating_entry(ctx):
    from myqc_amd.hip import enuc
    # call 1: (H2O)4, n = 28, from this library's own SCF and transform
    g = golden("h2o_4")
    Z = [int(z) for z in g["Z"]]
    ctx.set_system(Z, g["xyz"])
    S, H = ctx.int1e()
    ctx.eri_build(1e-12)
    nocc = int(g["nocc"])
    r = ctx.rhf(H, S, nocc, enuc(Z, g["xyz"]), e_tol=1e-10, d_tol=1e-8)
    nv = int(g["n"]) - nocc
    rec1 = _records_of(ctx.ao2mo(r["C"], nocc))
    ref1, got1 = _both(ctx, rec1, r["eig"], nocc, nv)
    print(f"[noalloc] (H2O)4 o={nocc} v={nv}: {got1} vs {ref1}", flush=True)
    assert np.max(np.abs(got1 - ref1)) <= 1e-12 * np.max(np.abs(ref1))
    # call 2, smaller: CO2's reference records (the reference's own mp2 input)
    c = golden("co2")
    o2 = int(c["nocc"])
    v2 = int(c["n"]) - o2
    # the fixture holds ijab_AB[i, j, a, b] = (ia|jb); the stream is [i][j][b][a]
    rec2 = np.ascontiguousarray(np.asarray(c["ijab_AB"], dtype=np.float64).transpose(0, 1, 3, 2)).ravel()
    assert rec2.size == o2 * o2 * v2 * v2
    ref2, got2 = _both(ctx, rec2, c["eig"], o2, v2)
    print(f"[noalloc] CO2 o={o2} v={v2}: {got2} vs {ref2} (reference {float(c['E_mp2']):.12f})", flush=True)
    assert np.max(np.abs(got2 - ref2)) <= 1e-12 * np.max(np.abs(ref2))
    assert abs(got2[2] - float(c["E_mp2"])) <= 1e-8


def test_noalloc_mp2_rejects_beyond_scratch(ctx):
    d = _to_device(np.zeros(4))
    try:
        rc = ctx.L.myqc_mp2_rhf_records_noalloc(ctx.h, d, d, 257, 1, d)
        assert rc == -1 and "256" in ctx.L.myqc_last_error(ctx.h).decode()
    finally:
        _rt().hipFree(d)

"""The UHF family on the GPU, through the C ABI: the in-process UHF SCF
(src/scf/scf.f90:221-404, :909-1117), the UHF transform
(slow_ao2mo_MP2_UHF, src/ao2mo/ao2mo.f90:614-904) and mp2_uhf
(src/mp2/mp2.f90:154-232), against the committed fixtures of the reference's
own programs (uhf_*.npz: its dens/UHFI2G loop and its mp2; uhf_no.npz also
the NO outputs the reference ships), its mp2 program run on our records, and
closed-shell identities with the RHF path.  UHF-CIS (slow_ao2mo_CIS_UHF,
src/ao2mo/ao2mo.f90:919-1227; cis_scf_uhf, src/CIS/cis.f90:87-223) against
the oracle's restated streams and matrix, and OH against the CFOUR output the
reference ships.  The Fortran programs scf, ao2mo, mp2 and cis run with
REF=UHF at the end.
"""
from __future__ import annotations

import os
import re
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "myqc_amd", "bin")
DATA = os.path.join(ROOT, "myqc_amd", "data")
REF_MP2 = os.path.join(ROOT, "oracle", "_ref", "mp2")

pytestmark = pytest.mark.gpu

UHF = ["uhf_no", "uhf_oh", "uhf_o3", "uhf_h"]


def _setup(ctx, g):
    Z = [int(z) for z in g["Z"]]
    ctx.set_system(Z, g["xyz"])
    ctx.eri_build(0.0)
    return Z


def _oracle_records(g, Ca, Cb):
    from oracle import ref_driver as R
    return R.uhf_records(g["XX_full"], Ca, Cb, int(g["nalpha"]), int(g["nbeta"]))


@pytest.mark.parametrize("name", UHF)
def test_uhf_scf_matches_reference_chain(ctx, name):
    """Core-Hamiltonian start (initUHF): the energy and both spins' orbital
    energies of the reference chain at SCF_Conv 11."""
    g = golden(name)
    _setup(ctx, g)
    S, H = ctx.int1e()
    assert np.max(np.abs(S - g["S"])) <= 1e-10 and np.max(np.abs(H - g["H"])) <= 1e-10
    na, nb = int(g["nalpha"]), int(g["nbeta"])
    print(f"[uhf] {name} n={int(g['n'])} na={na} nb={nb}", flush=True)
    r = ctx.uhf(H, S, na, nb, float(g["Enuc"]), e_tol=1e-11, d_tol=1e-9)
    print(f"[uhf] {name}: E {r['E']:.12f} ref {float(g['E_scf']):.12f} in {r['iterations']} iterations", flush=True)
    assert abs(r["E"] - float(g["E_scf"])) <= 1e-8
    assert np.max(np.abs(r["eig_a"] - g["eig_a"])) <= 1e-6
    assert np.max(np.abs(r["eig_b"] - g["eig_b"])) <= 1e-6
    # orthonormal orbitals of each spin
    for C in (r["Ca"], r["Cb"]):
        assert np.max(np.abs(C.T @ S @ C - np.eye(len(S)))) <= 1e-9
    Da, Db = ctx.density_uhf(r["Ca"], r["Cb"], na, nb)
    assert abs(np.trace(Da @ S) - na) <= 1e-9 and abs(np.trace(Db @ S) - nb) <= 1e-9
    if "example_E" in g.files:
        # the reference's own NO run (examples/NO, SCF_Conv 7)
        assert abs(r["E"] - float(g["example_E"])) <= 1e-8
        assert np.max(np.abs(r["eig_a"] - g["example_eig_a"])) <= 1e-6
        assert np.max(np.abs(r["eig_b"] - g["example_eig_b"])) <= 1e-6


def test_uhf_restart_from_reference_cui(ctx):
    """Started from the Cui the reference wrote for NO (the restart hook,
    scf.f90:314-320): the same state, found in a few iterations."""
    g = golden("uhf_no")
    _setup(ctx, g)
    S, H = ctx.int1e()
    r = ctx.uhf(H, S, int(g["nalpha"]), int(g["nbeta"]), float(g["Enuc"]), Ca_guess=g["example_Ca"],
                Cb_guess=g["example_Cb"], e_tol=1e-11, d_tol=1e-9)
    print(f"[uhf] NO from the reference's Cui: E {r['E']:.12f} in {r['iterations']} iterations", flush=True)
    assert abs(r["E"] - float(g["example_E"])) <= 1e-8
    assert r["iterations"] <= 20


def test_uhf_closed_shell_equals_rhf(ctx):
    """A closed shell run as UHF (MULTI=1) is the RHF state: same energy,
    Da = Db = D_RHF / 2."""
    g = golden("h2o")
    _setup(ctx, g)
    S, H = ctx.int1e()
    nocc = int(g["nocc"])
    from myqc_amd.hip import enuc
    en = enuc([int(z) for z in g["Z"]], g["xyz"])
    rr = ctx.rhf(H, S, nocc, en, e_tol=1e-11, d_tol=1e-9)
    ru = ctx.uhf(H, S, nocc, nocc, en, e_tol=1e-11, d_tol=1e-9)
    assert abs(ru["E"] - rr["E"]) <= 1e-9
    Da, Db = ctx.density_uhf(ru["Ca"], ru["Cb"], nocc, nocc)
    D = ctx.density(rr["C"], nocc)
    assert np.max(np.abs(Da - 0.5 * D)) <= 1e-7 and np.max(np.abs(Db - 0.5 * D)) <= 1e-7


@pytest.mark.parametrize("name", ["uhf_no", "uhf_oh", "uhf_o3"])
def test_uhf_ao2mo_records_match_oracle(ctx, name):
    """The three record streams against the four quarter steps on the
    reference's dense XX (einsum), for the fixture's orbitals."""
    g = golden(name)
    _setup(ctx, g)
    na, nb = int(g["nalpha"]), int(g["nbeta"])
    recs = ctx.ao2mo_uhf(g["Ca"], g["Cb"], na, nb)
    ref = _oracle_records(g, g["Ca"], g["Cb"])
    for key in ("AA", "BB", "AB"):
        assert recs[key].shape[0] == len(ref[key]), key
        if not ref[key] or ref[key][0].size == 0:
            continue
        want = np.array([r.T for r in ref[key]])       # [p, b, a]
        scale = max(1.0, np.max(np.abs(want)))
        assert np.max(np.abs(recs[key] - want)) <= 1e-10 * scale, key


@pytest.mark.parametrize("name", UHF)
def test_uhf_mp2_matches_reference(ctx, name):
    """E(AA), E(BB), E(AB), E(MBPT2) for the fixture's orbitals against the
    reference's mp2 on the same orbitals; device-resident and records entry
    points agree."""
    g = golden(name)
    _setup(ctx, g)
    na, nb = int(g["nalpha"]), int(g["nbeta"])
    recs = ctx.ao2mo_uhf(g["Ca"], g["Cb"], na, nb)
    e = ctx.mp2_uhf(g["eig_a"], g["eig_b"], na, nb)
    want = [float(g[k]) for k in ("E_aa", "E_bb", "E_ab", "E_mp2")]
    print(f"[uhf-mp2] {name}: {e} ref {want}", flush=True)
    for x, y in zip(e, want):
        assert abs(x - y) <= 1e-10
    e2 = ctx.mp2_uhf_records(recs, g["eig_a"], g["eig_b"], na, nb)
    for x, y in zip(e2, e):
        assert abs(x - y) <= 1e-13


def test_uhf_transform_with_equal_orbitals_is_rhf(ctx):
    """Ca = Cb = C (closed shell): ijab_AA and ijab_AB are the RHF records and
    E(AA) + E(BB) + E(AB) is the RHF E(MBPT2)."""
    g = golden("h2o")
    _setup(ctx, g)
    C, nocc, eig = g["C"], int(g["nocc"]), g["eig"]
    iajb = ctx.ao2mo(C, nocc)                       # [i][a][j][b]
    e_rhf = ctx.mp2(eig, nocc)
    recs = ctx.ao2mo_uhf(C, C, nocc, nocc)
    p = 0
    for i in range(nocc):
        for j in range(i + 1, nocc):
            assert np.max(np.abs(recs["AA"][p].T - iajb[i, :, j, :])) <= 1e-12
            p += 1
    ab = recs["AB"].reshape(nocc, nocc, *recs["AB"].shape[1:])
    assert np.max(np.abs(ab.transpose(0, 3, 1, 2) - iajb)) <= 1e-12
    e = ctx.mp2_uhf(eig, eig, nocc, nocc)
    assert abs(e[0] - e[1]) <= 1e-14 and abs(e[0] - e_rhf[0]) <= 1e-12
    assert abs(e[3] - e_rhf[2]) <= 1e-11


def test_uhf_ao2mo_bf16x3_error_bound(ctx):
    """Split-bf16 UHF transform: within the split's stated error of FP64."""
    from myqc_amd.hip import BF16X3
    g = golden("h2o_2")
    _setup(ctx, g)
    C, nocc, eig = g["C"], int(g["nocc"]), g["eig"]
    # an open-shell-like pair: beta with one electron fewer
    r64 = ctx.ao2mo_uhf(C, C, nocc, nocc - 1)
    e64 = ctx.mp2_uhf(eig, eig, nocc, nocc - 1)
    r16 = ctx.ao2mo_uhf(C, C, nocc, nocc - 1, precision=BF16X3)
    e16 = ctx.mp2_uhf(eig, eig, nocc, nocc - 1)
    for key in ("AA", "BB", "AB"):
        scale = np.max(np.abs(r64[key]))
        assert np.max(np.abs(r16[key] - r64[key])) <= 3e-4 * scale, key
    assert abs(e16[3] - e64[3]) <= 2e-5


# -- the Fortran programs with REF=UHF ------------------------------------------

def _records(path):
    buf = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(buf):
        (m,) = struct.unpack("<i", buf[pos:pos + 4])
        out.append(np.frombuffer(buf[pos + 4:pos + 4 + m], dtype="<f8"))
        pos += 8 + m
    return out


def _write_inputs(d, Z, xyz_bohr, mult, scf_conv=11, calc=1, excite=0, e_num=0):
    """envdat as the reference parser writes it for REF=UHF
    (src/parser/parser.f90:503-531): nelcA = nelcB + MULTI - 1.  options are
    0-based: calc (1), ref (3), SCF_Conv (8), multi (10), excite (13), E_NUM (15)."""
    nelc = int(sum(Z))
    nb = (nelc - (mult - 1)) // 2
    na = nb + mult - 1
    opts = [1, calc, 0, 1, 0, 1, 1000000, 1, scf_conv, 0, mult, 1, 1, excite, 0, e_num, 0]
    with open(os.path.join(d, "envdat"), "w") as f:
        f.write(f" {len(Z)}\n {na} {nb}\n 17\n " + " ".join(map(str, opts)) + "\n")
    with open(os.path.join(d, "nucpos"), "w") as f:
        for z, r in zip(Z, xyz_bohr):
            f.write(f" {int(z)} {float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")
    with open(os.path.join(d, "fmem"), "w") as f:
        f.write(" 1000000.0\n")
    shutil.copy(os.path.join(DATA, "Ftab"), os.path.join(d, "Ftab"))
    shutil.copy(os.path.join(DATA, "mybasis"), os.path.join(d, "mybasis"))
    return na, nb


def _run(d, exe):
    print(f"[program] {os.path.basename(exe)} in {d}", flush=True)
    r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not os.path.exists(os.path.join(d, "error")), \
        f"{exe} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    return r.stdout


def _number(text, label):
    m = re.search(re.escape(label) + r"\s*=?\s*([-+0-9.EeDd]+)", text)
    assert m, f"no '{label}' in output:\n{text}"
    return float(m.group(1).replace("D", "E").replace("d", "e"))


def test_uhf_programs_end_to_end(tmp_path):
    """int1e, scf, ao2mo and mp2 with REF=UHF on NO (MULTI=2): the energy and
    orbital energies of the reference chain, the file layouts the reference
    reads (Cui and eig with two lines, Da with two records, the three ijab
    streams), and the reference's own mp2 program on our files."""
    if not all(os.path.exists(os.path.join(BIN, p)) for p in ("int1e", "scf", "ao2mo", "mp2")):
        pytest.fail("Fortran host programs not built (myqc_amd/build.py)")
    g = golden("uhf_no")
    n = int(g["n"])
    d = str(tmp_path)
    na, nb = _write_inputs(d, g["Z"], g["xyz"], 2)
    assert (na, nb) == (int(g["nalpha"]), int(g["nbeta"]))
    _run(d, os.path.join(BIN, "int1e"))
    out = _run(d, os.path.join(BIN, "scf"))
    E = _number(out, "Total SCF energy (a.u)")
    assert abs(E - float(g["E_scf"])) <= 1e-8, (E, float(g["E_scf"]))
    lines = [ln for ln in open(os.path.join(d, "eig")).read().splitlines() if ln.strip()]
    vals = np.array(" ".join(lines).split(), dtype=np.float64)
    assert vals.size == 2 * n
    assert np.max(np.abs(vals[:n] - g["eig_a"])) <= 1e-6 and np.max(np.abs(vals[n:] - g["eig_b"])) <= 1e-6
    cv = np.array(open(os.path.join(d, "Cui")).read().split(), dtype=np.float64)
    assert cv.size == 2 * n * n
    Ca, Cb = cv[:n * n].reshape(n, n, order="F"), cv[n * n:].reshape(n, n, order="F")
    Dr = _records(os.path.join(d, "Da"))
    assert len(Dr) == 2
    assert np.max(np.abs(Dr[0].reshape(n, n, order="F") - Ca[:, :na] @ Ca[:, :na].T)) <= 1e-12
    assert np.max(np.abs(Dr[1].reshape(n, n, order="F") - Cb[:, :nb] @ Cb[:, :nb].T)) <= 1e-12

    _run(d, os.path.join(BIN, "ao2mo"))
    ref = _oracle_records(g, Ca, Cb)
    for key in ("AA", "BB", "AB"):
        got = _records(os.path.join(d, "ijab_" + key))
        assert len(got) == len(ref[key]), key
        for x, y in zip(got, ref[key]):
            assert np.max(np.abs(x - y.ravel(order="F"))) <= 1e-10, key

    out = _run(d, os.path.join(BIN, "mp2"))
    ours = [_number(out, lab) for lab in ("E(AA) =", "E(BB) =", "E(AB) =", "E(MBPT2) =")]
    assert abs(ours[3] - float(g["E_mp2"])) <= 1e-8, (ours, float(g["E_mp2"]))
    if os.path.exists(REF_MP2):
        theirs = _run(d, REF_MP2)
        for lab, v in zip(("E(AA) =", "E(BB) =", "E(AB) =", "E(MBPT2) ="), ours):
            assert abs(_number(theirs, lab) - v) <= 1e-10, lab


# -- UHF-CIS ----------------------------------------------------------------------

CFOUR_OH = [0.0, 0.228665544103, 0.400333416145]   # examples/OH/cfour/out, TDA roots 1-3


@pytest.mark.parametrize("name", ["uhf_no", "uhf_oh", "uhf_o3"])
def test_cis_streams_match_oracle(ctx, name):
    """The five streams of slow_ao2mo_CIS_UHF against its quarter steps on the
    reference's dense XX, for the fixture's orbitals."""
    from oracle import ref_driver as R
    g = golden(name)
    _setup(ctx, g)
    na, nb = int(g["nalpha"]), int(g["nbeta"])
    got = ctx.ao2mo_cis_uhf(g["Ca"], g["Cb"], na, nb)
    want = R.cis_uhf_streams(g["XX_full"], g["Ca"], g["Cb"], na, nb)
    for key in ctx.CIS_STREAMS:
        assert got[key].shape == want[key].shape, key
        if want[key].size:
            scale = max(1.0, np.max(np.abs(want[key])))
            assert np.max(np.abs(got[key] - want[key])) <= 1e-10 * scale, key


@pytest.mark.parametrize("name", ["uhf_no", "uhf_oh", "uhf_o3"])
def test_cis_matches_oracle(ctx, name):
    """Every excitation energy of the UHF-CIS matrix against numpy's
    eigensolver on the restated matrix; eigenvectors up to sign for isolated
    roots; device-resident and record-stream entry points agree.  OH also
    against the CFOUR output the reference ships."""
    from oracle import ref_driver as R
    g = golden(name)
    _setup(ctx, g)
    n, na, nb = int(g["n"]), int(g["nalpha"]), int(g["nbeta"])
    dim = na * (n - na) + nb * (n - nb)
    streams = ctx.ao2mo_cis_uhf(g["Ca"], g["Cb"], na, nb)
    E, V = ctx.cis_uhf(g["eig_a"], g["eig_b"], na, nb, dim)
    A = R.cis_uhf_matrix(R.cis_uhf_streams(g["XX_full"], g["Ca"], g["Cb"], na, nb), g["eig_a"], g["eig_b"], na, nb)
    w, U = np.linalg.eigh(A)
    print(f"[cis] {name}: dim {dim}, lowest {E[:4]}", flush=True)
    assert np.max(np.abs(E - w)) <= 1e-10
    for k in range(dim):
        gap = min([abs(w[k] - w[j]) for j in range(dim) if j != k] or [1.0])
        if gap > 1e-6:
            assert abs(abs(V[:, k] @ U[:, k]) - 1.0) <= 1e-8, k
    E2, V2 = ctx.cis_uhf(g["eig_a"], g["eig_b"], na, nb, 3, streams=streams)
    assert np.max(np.abs(E2 - E[:3])) <= 1e-12
    if name == "uhf_oh":
        assert np.max(np.abs(E[:3] - CFOUR_OH)) <= 1e-5


def test_cis_programs_end_to_end(tmp_path):
    """int1e, scf, ao2mo and cis on the reference's OH example (REF=UHF,
    MULTI=2, CALC=SCF, EXCITE=CIS, E_NUM=4): the five stream files, the
    ints_AA/ints_BB files, and the printed eigenvalues against the oracle on
    this run's own orbitals and against CFOUR."""
    from oracle import ref_driver as R
    if not all(os.path.exists(os.path.join(BIN, p)) for p in ("int1e", "scf", "ao2mo", "cis")):
        pytest.fail("Fortran host programs not built (myqc_amd/build.py)")
    g = golden("uhf_oh")
    n = int(g["n"])
    d = str(tmp_path)
    na, nb = _write_inputs(d, g["Z"], g["xyz"], 2, calc=0, excite=1, e_num=4)
    _run(d, os.path.join(BIN, "int1e"))
    _run(d, os.path.join(BIN, "scf"))
    _run(d, os.path.join(BIN, "ao2mo"))
    out = _run(d, os.path.join(BIN, "cis"))
    vals = np.array([float(x) for x in re.findall(r"Eigenvalue \(a\.u\.\)\s+([-+0-9.Ee]+)", out)])
    assert vals.size == 4, out
    cv = np.array(open(os.path.join(d, "Cui")).read().split(), dtype=np.float64)
    Ca, Cb = cv[:n * n].reshape(n, n, order="F"), cv[n * n:].reshape(n, n, order="F")
    ev = np.array(" ".join(open(os.path.join(d, "eig")).read().split()).split(), dtype=np.float64)
    st = R.cis_uhf_streams(g["XX_full"], Ca, Cb, na, nb)
    for key in ("ajib_AA", "ajbi_AA", "ajib_AB", "ajib_BB", "ajbi_BB"):
        recs = _records(os.path.join(d, key))
        assert len(recs) == st[key].shape[0], key
        for x, y in zip(recs, st[key]):
            assert np.max(np.abs(x - y)) <= 1e-10, key
    ints = _records(os.path.join(d, "ints_AA"))
    assert np.max(np.abs(np.array(ints) - (st["ajib_AA"] - st["ajbi_AA"]))) <= 1e-10
    w = np.linalg.eigvalsh(R.cis_uhf_matrix(st, ev[:n], ev[n:], na, nb))
    assert np.max(np.abs(vals - w[:4])) <= 1e-9
    assert np.max(np.abs(vals[:3] - CFOUR_OH)) <= 1e-5

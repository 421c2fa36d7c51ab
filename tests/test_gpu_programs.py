"""The Fortran host programs (int1e, int2e, scf, ao2mo, mp2) run end to end
over the iso_c_binding C ABI, reading and writing the reference's files
(envdat/nucpos/fmem, mybasis, Ftab -> Suv/Huv, XX, Cui/Da/eig, ijab_AA/AB),
checked against the reference's own outputs for the same inputs (committed
fixtures) and the oracle.

Reference program contracts: src/integrals/int1e.f90, src/integrals/int2e.f90
(XX: one unformatted record, :166,306), src/scf/scf.f90 (RHF, :61-216),
src/ao2mo/ao2mo.f90 (:465-602, records :566-582), src/mp2/mp2.f90 (:79-144).
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
from molden_cmp import compare_molden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "myqc_amd", "bin")
DATA = os.path.join(ROOT, "myqc_amd", "data")

pytestmark = pytest.mark.gpu


def _records(path):
    """Fortran sequential unformatted records (4-byte markers, subrecords)."""
    buf = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(buf):
        (m,) = struct.unpack("<i", buf[pos:pos + 4])
        chunks = []
        while True:
            ln = abs(m)
            chunks.append(buf[pos + 4:pos + 4 + ln])
            pos += 8 + ln
            if m >= 0:
                break
            (m,) = struct.unpack("<i", buf[pos:pos + 4])
        out.append(np.frombuffer(b"".join(chunks), dtype="<f8"))
    return out


def _text_matrix(path, n):
    return np.array(open(path).read().split(), dtype=np.float64)[: n * n].reshape(n, n, order="F")


def _write_inputs(d, Z, xyz_bohr, scf_conv=11, ref=0, nab=None):
    """The files the reference's parser writes (src/myQC/env.f90:16-73):
    envdat = nnuc / nelcA nelcB / 17 / options, nucpos = Z x y z (bohr)."""
    nelc = int(sum(Z))
    na, nb = nab if nab is not None else (nelc // 2, nelc // 2)
    opts = [1, 1, 0, ref, 0, 1, 1000000, 1, scf_conv, 0, 1, 1, 1, 0, 0, 0, 0]
    with open(os.path.join(d, "envdat"), "w") as f:
        f.write(f" {len(Z)}\n {na} {nb}\n 17\n " + " ".join(map(str, opts)) + "\n")
    with open(os.path.join(d, "nucpos"), "w") as f:
        for z, r in zip(Z, xyz_bohr):
            f.write(f" {int(z)} {float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")
    with open(os.path.join(d, "fmem"), "w") as f:
        f.write(" 1000000.0\n")
    shutil.copy(os.path.join(DATA, "Ftab"), os.path.join(d, "Ftab"))
    shutil.copy(os.path.join(DATA, "mybasis"), os.path.join(d, "mybasis"))


def _run(d, prog):
    exe = os.path.join(BIN, prog)
    print(f"[program] {prog} in {d}", flush=True)
    r = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not os.path.exists(os.path.join(d, "error")), \
        f"{prog} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}"
    return r.stdout


def _number(text, label):
    m = re.search(re.escape(label) + r"\s*=?\s*([-+0-9.EeDd]+)", text)
    assert m, f"no '{label}' in output:\n{text}"
    return float(m.group(1).replace("D", "E").replace("d", "e"))


@pytest.mark.parametrize("mol", ["co2", "h2o"])
def test_programs_end_to_end(mol, tmp_path):
    if not all(os.path.exists(os.path.join(BIN, p)) for p in ("int1e", "int2e", "scf", "ao2mo", "mp2")):
        pytest.fail("Fortran host programs not built (myqc_amd/build.py)")
    from oracle import oracle as O

    g = golden(mol)
    n, nocc = int(g["n"]), int(g["nocc"])
    d = str(tmp_path)
    _write_inputs(d, g["Z"], g["xyz"])

    _run(d, "int1e")
    S = _text_matrix(os.path.join(d, "Suv"), n)
    H = _text_matrix(os.path.join(d, "Huv"), n)
    assert np.max(np.abs(S - g["S"])) <= 1e-10
    assert np.max(np.abs(H - g["H"])) <= 1e-10

    _run(d, "int2e")
    recs = _records(os.path.join(d, "XX"))
    assert len(recs) == 1 and recs[0].size == n ** 4
    XX = recs[0].reshape(n, n, n, n, order="F")
    assert np.max(np.abs(XX - g["XX_full"])) <= 1e-10
    # int2e skips when XX exists (int2e.f90:58-67)
    assert "Reading two electron integrals" in _run(d, "int2e")

    out = _run(d, "scf")
    E = _number(out, "Total SCF energy (a.u)")
    assert abs(E - float(g["E_scf"])) <= 1e-8, (E, float(g["E_scf"]))
    eig = np.array(open(os.path.join(d, "eig")).read().split(), dtype=np.float64)
    assert np.max(np.abs(eig - g["eig"])) <= 1e-6
    C = _text_matrix(os.path.join(d, "Cui"), n)
    # orthonormal orbitals, and the density record is 2 C_occ C_occ^T
    assert np.max(np.abs(C.T @ g["S"] @ C - np.eye(n))) <= 1e-9
    Da = _records(os.path.join(d, "Da"))[0].reshape(n, n, order="F")
    assert np.max(np.abs(Da - 2.0 * C[:, :nocc] @ C[:, :nocc].T)) <= 1e-12
    assert abs(np.trace(Da @ g["S"]) - 2 * nocc) <= 1e-9

    _run(d, "ao2mo")
    nv = n - nocc
    ab = _records(os.path.join(d, "ijab_AB"))
    aa = _records(os.path.join(d, "ijab_AA"))
    assert len(ab) == nocc * nocc and len(aa) == nocc * (nocc - 1) // 2
    # (ia|jb) of this run's own orbitals, by the oracle's quarter transforms
    iajb = O.ao2mo_ovov(g["XX_full"], C, nocc)
    k = 0
    for i in range(nocc):
        for j in range(nocc):
            rec = ab[k].reshape(nv, nv, order="F")
            assert np.max(np.abs(rec - iajb[i, :, j, :])) <= 1e-10
            k += 1
    rec01 = aa[0].reshape(nv, nv, order="F")
    assert np.max(np.abs(rec01 - iajb[0, :, 1, :])) <= 1e-10

    # the fmem ledger is back where the reference's programs leave it
    # (src/myQC/env.f90:77-112)
    files = golden("files")
    fm = float(open(os.path.join(d, "fmem")).read().split()[0])
    if f"fmem_{mol}_int2e" in files.files:
        assert abs(fm - float(files[f"fmem_{mol}_int2e"])) <= 1e-6
    assert abs(fm - 1000000.0) <= 1e-6

    out = _run(d, "mp2")
    # the reference's mp2 prints the correlation energies (mp2.f90:136-141)
    for label, key in (("E(AA) =", "E_aa"), ("E(AB) =", "E_ab"), ("E(MBPT2) =", "E_mp2")):
        v = _number(out, label)
        assert abs(v - float(g[key])) <= 1e-8, (label, v, float(g[key]))


def test_program_error_sentinel(tmp_path):
    """A failing program writes the `error` sentinel and stops nonzero, and
    later programs stop at once when it exists (src/myQC/env.f90:36-41)."""
    d = str(tmp_path)
    g = golden("h2o")
    _write_inputs(d, g["Z"], g["xyz"])
    os.remove(os.path.join(d, "Ftab"))
    r = subprocess.run([os.path.join(BIN, "int2e")], cwd=d, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and os.path.exists(os.path.join(d, "error"))
    r = subprocess.run([os.path.join(BIN, "int1e")], cwd=d, capture_output=True, text=True, timeout=60)
    assert not os.path.exists(os.path.join(d, "Suv"))


def _write_record(path, a):
    """One Fortran sequential unformatted record (under 2 GiB)."""
    b = np.asfortranarray(a, dtype="<f8").ravel(order="F").tobytes()
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(b)) + b + struct.pack("<i", len(b)))


def test_int2e_record_above_2gib(ctx, tmp_path):
    """XX of a system with n >= 128 (19 waters, n = 133: 2.5 GB), which
    int2e.f90:166,306 writes as one record: by default in gfortran's subrecord
    layout (the reference's own build uses gfortran, src/Makefile:11), and
    with MYQC_XX_RECORD=single as one record the way this image's flang
    runtime writes it.  Both read back to the C ABI's dense integrals, and
    when the reference's own RHFI2G has been built from its sources here with
    flang (oracle/_ref), that program reads the single-record XX
    (src/I2G/RHFI2G.f90:72-74) and returns the G that myqc_fock_G gives."""
    from myqc_amd import molecules as M
    from myqc_amd.basis import build_basis
    Z, x = M.water_cluster(19, seed=19)
    xyz = M.to_bohr(Z, x)
    n = build_basis(Z).norb
    assert n >= 128
    d = str(tmp_path)
    _write_inputs(d, Z, xyz)
    nbytes = 8 * n ** 4
    ctx.set_system(Z, xyz)
    ref = ctx.eri_full(0.0).ravel(order="F")
    rng = np.random.default_rng(3)
    idx = rng.integers(0, n ** 4, size=200000)
    path = os.path.join(d, "XX")
    # default: gfortran subrecords of at most 2^31 - 9 bytes (the layout of
    # the reference's gfortran build), head markers negative but on the last,
    # tail markers negative but on the first
    _run(d, "int2e")
    maxsub = 2147483639
    k = -(-nbytes // maxsub)
    assert os.path.getsize(path) == nbytes + 8 * k
    with open(path, "rb") as f:
        pos = 0
        for i in range(k):
            ln = min(maxsub, nbytes - i * maxsub)
            f.seek(pos)
            (head,) = struct.unpack("<i", f.read(4))
            f.seek(pos + 4 + ln)
            (tail,) = struct.unpack("<i", f.read(4))
            print(f"[XX] n={n} subrecord {i}: markers {head} {tail}", flush=True)
            assert head == (ln if i == k - 1 else -ln) and tail == (ln if i == 0 else -ln)
            pos += 8 + ln
    # values: the data bytes of the subrecords in order
    raw = np.memmap(path, dtype=np.uint8, mode="r")
    data = np.concatenate([raw[4 + i * (maxsub + 8): 4 + i * (maxsub + 8) + min(maxsub, nbytes - i * maxsub)]
                           for i in range(k)]).view("<f8")
    assert np.array_equal(data[idx], ref[idx])
    del raw, data
    os.remove(path)
    # MYQC_XX_RECORD=single: one record, as this image's Fortran runtime
    # (flang, the compiler the reference is built with here) writes it, its
    # 4-byte markers holding the length modulo 2^32
    env = dict(os.environ, MYQC_XX_RECORD="single")
    print(f"[program] int2e (single record) in {d}", flush=True)
    r = subprocess.run([os.path.join(BIN, "int2e")], cwd=d, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    size = os.path.getsize(path)
    with open(path, "rb") as f:
        (head,) = struct.unpack("<I", f.read(4))
        f.seek(-4, 2)
        (tail,) = struct.unpack("<I", f.read(4))
    print(f"[XX] n={n} file {size} bytes, markers {head} {tail}", flush=True)
    assert size == nbytes + 8 and head == tail == nbytes % 2 ** 32
    XX = np.memmap(path, dtype="<f8", mode="r", offset=4, shape=(n ** 4,))
    assert np.array_equal(XX[idx], ref[idx])
    del XX, ref
    rhfi2g = os.path.join(ROOT, "oracle", "_ref", "RHFI2G")
    if not os.path.exists(rhfi2g):
        pytest.skip("reference RHFI2G not built here (oracle/build_ref.sh)")
    D = rng.standard_normal((n, n))
    D = 0.5 * (D + D.T)
    _write_record(os.path.join(d, "Da"), D)
    print(f"[program] reference RHFI2G in {d}", flush=True)
    r = subprocess.run([rhfi2g], cwd=d, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and not os.path.exists(os.path.join(d, "error")), r.stdout + r.stderr
    G_ref = _records(os.path.join(d, "Guv"))[0].reshape(n, n, order="F")
    ctx.eri_build(0.0)
    G = ctx.fock_G(D)
    err = np.max(np.abs(G - G_ref))
    print(f"[XX] reference RHFI2G on our XX: max |G - G_ref| = {err:.2e} (max |G| {np.max(np.abs(G_ref)):.2e})")
    assert err <= 1e-9 * max(1.0, np.max(np.abs(G_ref)))


@pytest.mark.parametrize("mol", ["co2", "h2o_2"])
def test_uhf_fock_matches_reference_UHFI2G(ctx, mol, tmp_path):
    """myqc_fock_G_uhf against the reference's own UHFI2G program (built from
    src/I2G/UHFI2G.f90 into oracle/_ref) reading the XX our int2e writes, for
    random symmetric spin densities; with Da = Db = D/2 both equal RHF's G(D)."""
    g = golden(mol)
    n = int(g["n"])
    Z = [int(z) for z in g["Z"]]
    ctx.set_system(Z, g["xyz"])
    ctx.eri_build(0.0)
    rng = np.random.default_rng(5)
    Da, Db = (0.5 * (M + M.T) for M in (rng.standard_normal((n, n)), rng.standard_normal((n, n))))
    Ga, Gb = ctx.fock_G_uhf(Da, Db)
    Gh, Gh2 = ctx.fock_G_uhf(0.5 * Da, 0.5 * Da)
    G = ctx.fock_G(Da)
    assert np.max(np.abs(Gh - G)) <= 1e-12 * max(1.0, np.max(np.abs(G))) and np.array_equal(Gh, Gh2)
    uhfi2g = os.path.join(ROOT, "oracle", "_ref", "UHFI2G")
    if not os.path.exists(uhfi2g):
        pytest.skip("reference UHFI2G not built here (oracle/build_ref.sh)")
    d = str(tmp_path)
    _write_inputs(d, g["Z"], g["xyz"])
    _run(d, "int2e")
    with open(os.path.join(d, "Da"), "wb") as f:
        for M in (Da, Db):
            b = np.asfortranarray(M, dtype="<f8").ravel(order="F").tobytes()
            f.write(struct.pack("<i", len(b)) + b + struct.pack("<i", len(b)))
    print(f"[program] reference UHFI2G in {d}", flush=True)
    r = subprocess.run([uhfi2g], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not os.path.exists(os.path.join(d, "error")), r.stdout + r.stderr
    GaR, GbR = (x.reshape(n, n, order="F") for x in _records(os.path.join(d, "Guv")))
    scale = max(1.0, np.max(np.abs(GaR)))
    assert np.max(np.abs(Ga - GaR)) <= 1e-10 * scale
    assert np.max(np.abs(Gb - GbR)) <= 1e-10 * scale


def test_props_programs_end_to_end(tmp_path):
    """int1e, scf, int1p and cphf on H2O: Pxuv against the reference's own
    int1p, and the printed dipole (6 decimals, as cphf prints it) against the
    restated contraction over the reference's integrals."""
    if not all(os.path.exists(os.path.join(BIN, p)) for p in ("int1e", "scf", "int1p", "cphf")):
        pytest.fail("Fortran host programs not built (myqc_amd/build.py)")
    g = golden("h2o")
    pr = golden("props_h2o")
    n = int(g["n"])
    d = str(tmp_path)
    _write_inputs(d, g["Z"], g["xyz"])
    _run(d, "int1e")
    _run(d, "scf")
    _run(d, "int1p")
    vals = np.array(open(os.path.join(d, "Pxuv")).read().split(), dtype=np.float64)
    assert vals.size == 3 * n * n
    P = vals.reshape(3, n * n)
    for c in range(3):
        assert np.max(np.abs(P[c].reshape(n, n, order="F") - pr["P"][c])) <= 1e-10
    assert "Reading Dipole integrals" in _run(d, "int1p")
    out = _run(d, "cphf")
    lines = out.splitlines()
    k = next(i for i, ln in enumerate(lines) if "X" in ln and "Y" in ln and "Z" in ln)
    mu = np.array([float(x) for x in lines[k + 1].split()])
    assert np.max(np.abs(mu - pr["mu"])) <= 1e-6 + 1e-12


@pytest.mark.parametrize("key,fixture,ref,e_tol,c_tol", [("be", "ex_be", 0, 2e-8, 1e-6),
                                                         ("o1", "ex_o1", 0, 2e-8, 1e-6),
                                                         ("no", "uhf_no", 1, 1e-6, 1e-4)])
def test_molden_matches_reference_examples(key, fixture, ref, e_tol, c_tol, tmp_path):
    """The scf program's MOLDEN file against the ones the reference ships
    (examples/Be, examples/O/singlet, examples/NO; makeMOLDEN,
    src/scf/scf.f90:1343-1466): atoms, basis blocks, and per spin the orbital
    energies (printed to 8 decimals), occupations and coefficients (as the
    projector of each level, so signs and rotations among degenerate orbitals
    do not matter; for NO, whose UHF solution breaks the axial symmetry, as
    invariants under rotations about the axis), at the examples' default
    SCF_Conv of 7.  NO's orbital energies are held to 1e-6, the bound the UHF
    tests hold them to against the same example (its density criterion leaves
    them uncertain well above the 8 printed decimals)."""
    g = golden(fixture)
    d = str(tmp_path)
    Z = [int(z) for z in g["Z"]]
    nab = (int(g["nalpha"]), int(g["nbeta"])) if ref == 1 else None
    _write_inputs(d, Z, g["xyz"], scf_conv=7, ref=ref, nab=nab)
    _run(d, "int1e")
    _run(d, "scf")
    ours = open(os.path.join(d, "MOLDEN")).read()
    linear = None
    if ref == 1 and np.allclose(np.asarray(g["xyz"])[:, :2], 0.0):
        # NO: the unpaired electron's pi orbital may lie along x or y
        linear, k = [], 0
        for z in Z:
            m = 1 if z <= 2 else 5
            linear.append(list(range(k, k + m)))
            k += m
    compare_molden(ours, str(golden("files")[f"molden_{key}"]), e_tol, c_tol, linear)

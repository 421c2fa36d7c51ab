"""Host-side file contracts of the programs, checked on the CPU (no device
call): the basinfo/setinfo arrays of buildBasis, the fmem ledger's failure
path, and the gfortran subrecord layout of large unformatted records.

References: src/myQC/basis.f90:23-226 (buildBasis, files :213-218);
src/myQC/env.f90:77-112 (setenv/nmem) with the checks of
src/integrals/int1e.f90:81-85 and src/scf/scf.f90:114-133; the XX record of
src/integrals/int2e.f90:166,306 in the layout of the reference's gfortran
build (src/Makefile:11).  The expected files are the reference's own outputs
(tests/golden/files.npz, written by tests/golden/make_file_fixtures.py).
"""
from __future__ import annotations

import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

BIN = os.path.join(ROOT, "myqc_amd", "bin")
DATA = os.path.join(ROOT, "myqc_amd", "data")
SUBREC = os.path.join(ROOT, "tests", "hostcheck", "subrec_check")
MOLS = ("co2", "h2o_2", "mol1", "mol3")


def _ints_before_blank(path):
    vals = []
    for line in open(path):
        if not line.strip():
            break
        vals += [int(t) for t in line.split()]
    return np.array(vals, dtype=np.int64)


def _inputs(d, Z, xyz, fmem="1000000.0", ref=0, nelc=None):
    nelc = int(sum(Z)) if nelc is None else nelc
    opts = [1, 1, 0, ref, 0, 1, 1000000, 1, 11, 0, 1, 1, 1, 0, 0, 0, 0]
    with open(os.path.join(d, "envdat"), "w") as f:
        f.write(f" {len(Z)}\n {nelc // 2} {nelc // 2}\n 17\n " + " ".join(map(str, opts)) + "\n")
    with open(os.path.join(d, "nucpos"), "w") as f:
        for z, r in zip(Z, xyz):
            f.write(f" {int(z)} {float(r[0])!r} {float(r[1])!r} {float(r[2])!r}\n")
    with open(os.path.join(d, "fmem"), "w") as f:
        f.write(f" {fmem}\n")
    shutil.copy(os.path.join(DATA, "Ftab"), os.path.join(d, "Ftab"))
    shutil.copy(os.path.join(DATA, "mybasis"), os.path.join(d, "mybasis"))


@pytest.mark.parametrize("mol", MOLS)
def test_basis_arrays_match_reference_files(mol):
    """myqc_amd.basis (the arrays the C ABI takes) equals what the reference's
    buildBasis writes to basinfo and setinfo for the same molecule."""
    from myqc_amd.basis import build_basis
    f = golden("files")
    b = build_basis([int(z) for z in golden(mol)["Z"]])
    assert np.array_equal(np.asarray(b.basinfo, dtype=np.int64), f[f"basinfo_{mol}"].astype(np.int64))
    assert np.array_equal(np.asarray(b.setinfo, dtype=np.int64), f[f"setinfo_{mol}"].astype(np.int64))


@pytest.mark.parametrize("mol", MOLS)
def test_program_basis_files_and_ledger_failure(mol, tmp_path):
    """The int1e program writes basinfo/setinfo as the reference does, and with
    fmem too small for S and H it fails the way the reference's int1e does
    (the `error` sentinel, int1e.f90:81-85) before any device call.  The scf
    program fails the same way (scf.f90:114-133)."""
    if not os.path.exists(os.path.join(BIN, "int1e")):
        pytest.fail("Fortran host programs not built (myqc_amd/build.py)")
    f = golden("files")
    g = golden(mol)
    d = str(tmp_path)
    _inputs(d, [int(z) for z in g["Z"]], g["xyz"], fmem="1.0000000000000000E-006")
    r = subprocess.run([os.path.join(BIN, "int1e")], cwd=d, capture_output=True, text=True, timeout=60)
    assert bool(f["lowmem_int1e_error"])
    assert r.returncode != 0 and os.path.exists(os.path.join(d, "error")), r.stdout + r.stderr
    assert not os.path.exists(os.path.join(d, "Suv"))
    assert np.array_equal(_ints_before_blank(os.path.join(d, "basinfo")), f[f"basinfo_{mol}"])
    assert np.array_equal(_ints_before_blank(os.path.join(d, "setinfo")), f[f"setinfo_{mol}"])
    os.remove(os.path.join(d, "error"))
    r = subprocess.run([os.path.join(BIN, "scf")], cwd=d, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and os.path.exists(os.path.join(d, "error")), r.stdout + r.stderr


def _subrecords(buf):
    """(head, tail, payload) of each subrecord, in file order."""
    out, pos = [], 0
    while pos < len(buf):
        (h,) = struct.unpack("<i", buf[pos:pos + 4])
        ln = abs(h)
        (t,) = struct.unpack("<i", buf[pos + 4 + ln:pos + 8 + ln])
        out.append((h, t, buf[pos + 4:pos + 4 + ln]))
        pos += 8 + ln
    assert pos == len(buf)
    return out


@pytest.mark.parametrize("nbytes,maxsub", [(100, 24), (96, 24), (24, 24), (23, 24), (1, 1), (8, 3), (4096, 4096)])
def test_record_subrecord_layout(nbytes, maxsub, tmp_path):
    """write_record_subrecords (the programs' writer of XX) in gfortran's
    layout, with the subrecord limit lowered so the split shows at small
    sizes: full subrecords of maxsub bytes then the rest; head markers
    negative except on the last, tail markers negative except on the first;
    one subrecord is a plain record (+len, +len)."""
    if not os.path.exists(SUBREC):
        pytest.fail("tests/hostcheck/subrec_check not built (__graft_entry__.build)")
    path = str(tmp_path / "rec")
    subprocess.run([SUBREC, path, str(nbytes), str(maxsub)], check=True, timeout=60)
    subs = _subrecords(open(path, "rb").read())
    k = -(-nbytes // maxsub)
    assert len(subs) == k
    for i, (h, t, p) in enumerate(subs):
        ln = maxsub if i < k - 1 else nbytes - maxsub * (k - 1)
        assert len(p) == ln
        assert h == (ln if i == k - 1 else -ln)
        assert t == (ln if i == 0 else -ln)
    payload = b"".join(p for _, _, p in subs)
    assert payload == bytes(i % 256 for i in range(nbytes))



@pytest.mark.parametrize("key,nnuc_uhf", [("be", 0), ("o1", 0), ("no", 1)])
def test_molden_writer_reproduces_reference_files(key, nnuc_uhf, tmp_path):
    """write_molden (the scf program's MOLDEN writer) fed the atoms, orbital
    energies and coefficients of a MOLDEN file the reference ships
    (examples/Be, O/singlet, NO) writes that file again: the same sections,
    fixed-format lines, basis blocks from mybasis and orbitals."""
    from molden_cmp import compare_molden, parse_molden
    exe = os.path.join(ROOT, "tests", "hostcheck", "molden_check")
    if not os.path.exists(exe):
        pytest.fail("tests/hostcheck/molden_check not built (__graft_entry__.build)")
    ref = str(golden("files")[f"molden_{key}"])
    atoms, _, mos = parse_molden(ref)
    n = len(mos[0][3])
    d = str(tmp_path)
    shutil.copy(os.path.join(DATA, "mybasis"), os.path.join(d, "mybasis"))
    with open(os.path.join(d, "molden_in"), "w") as f:
        f.write(f"{len(atoms)} {n} {nnuc_uhf}\n")
        for a in atoms:
            f.write(f"{a[2]} {a[3][0]!r} {a[3][1]!r} {a[3][2]!r}\n")
        for spin in ("Alpha", "Beta") if nnuc_uhf else ("Alpha",):
            o = [m for m in mos if m[1] == spin]
            f.write(" ".join(repr(m[0]) for m in o) + "\n")
            f.write(" ".join(repr(float(x)) for m in o for x in m[3]) + "\n")
    subprocess.run([exe], cwd=d, check=True, timeout=60)
    ours = open(os.path.join(d, "MOLDEN")).read()
    compare_molden(ours, ref, 1e-12, 1e-12)
    # the fixed-format lines are the reference's, character for character
    import re
    pat = re.compile(r"^( Ene= {1,}-?\d+\.\d{8}| [ \d]{4} {4} *-?\d+\.\d{10}|[A-Z][a-z ] {4}[ \d]{3} {4}[ \d]{3} .*)$")
    fixed = lambda t: [ln for ln in t.splitlines() if pat.match(ln)]
    assert len(fixed(ref)) == len(parse_molden(ref)[0]) + sum(1 + len(m[3]) for m in mos)
    assert fixed(ours) == fixed(ref)

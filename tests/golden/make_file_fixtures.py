"""Writes tests/golden/files.npz: the reference's own file outputs for the
host-side contracts of the programs, from the reference's programs built from
its sources by oracle/build_ref.sh into oracle/_ref/.  Run in a container
that has /root/reference:

    make -C oracle && python tests/golden/make_file_fixtures.py

Contents:
  basinfo_<mol>, setinfo_<mol>: the integer arrays buildBasis writes to
      basinfo/setinfo (src/myQC/basis.f90:213-218) when the reference's int1e
      runs on fixture <mol>;
  fmem_<mol>_int1e, fmem_<mol>_int2e: the fmem figure after the reference's
      int1e and then int2e (the ledger of src/myQC/env.f90:77-112), from
      1000000.0;
  lowmem_int1e_error: whether the reference's int1e writes `error` when fmem
      cannot hold S and H (int1e.f90:81-85);
  molden_<example>: the MOLDEN files the reference ships in examples/
      (Be, O/singlet, NO), written by its makeMOLDEN (scf.f90:1343-1466).
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ref_driver as R  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = "/root/reference/examples"
MOLS = ("co2", "h2o_2", "mol1", "mol3")


def ints_before_blank(path):
    vals = []
    for line in open(path):
        if not line.strip():
            break
        vals += [int(t) for t in line.split()]
    return np.array(vals, dtype=np.int32)


def main():
    if not R.have_ref():
        sys.exit("reference programs missing: run `make -C oracle` in a container with /root/reference")
    out = {}
    for mol in MOLS:
        g = np.load(os.path.join(OUT, mol + ".npz"))
        Z = [int(z) for z in g["Z"]]
        d = tempfile.mkdtemp(prefix="myqc_files_", dir=R.REF)
        try:
            R.write_env(d, Z, g["xyz"], charge=int(g["charge"]))
            R._run(d, "int1e")
            out[f"basinfo_{mol}"] = ints_before_blank(os.path.join(d, "basinfo"))
            out[f"setinfo_{mol}"] = ints_before_blank(os.path.join(d, "setinfo"))
            out[f"fmem_{mol}_int1e"] = np.float64(open(os.path.join(d, "fmem")).read().split()[0])
            R._run(d, "int2e")
            out[f"fmem_{mol}_int2e"] = np.float64(open(os.path.join(d, "fmem")).read().split()[0])
        finally:
            shutil.rmtree(d)
        print(mol, out[f"basinfo_{mol}"][:2], out[f"fmem_{mol}_int1e"], out[f"fmem_{mol}_int2e"])
    g = np.load(os.path.join(OUT, "h2o_2.npz"))
    d = tempfile.mkdtemp(prefix="myqc_files_", dir=R.REF)
    try:
        R.write_env(d, [int(z) for z in g["Z"]], g["xyz"])
        with open(os.path.join(d, "fmem"), "w") as f:
            f.write(" 1.0000000000000000E-006\n")
        import subprocess
        subprocess.run([os.path.join(R.REF, "int1e")], cwd=d, capture_output=True, text=True)
        out["lowmem_int1e_error"] = np.bool_(os.path.exists(os.path.join(d, "error")))
    finally:
        shutil.rmtree(d)
    print("lowmem int1e error:", out["lowmem_int1e_error"])
    for key, ex in (("be", "Be"), ("o1", "O/singlet"), ("no", "NO")):
        out[f"molden_{key}"] = np.array(open(os.path.join(EXAMPLES, ex, "MOLDEN")).read())
    np.savez_compressed(os.path.join(OUT, "files.npz"), **out)
    print("wrote files.npz")


if __name__ == "__main__":
    main()

"""Writes the golden fixtures under tests/golden/ from the reference's own
programs, built from its sources by oracle/build_ref.sh (flang) into
oracle/_ref/.  Run in a container that has /root/reference:

    make -C oracle && python tests/golden/make_fixtures.py

Each <name>.npz holds inputs (Z, xyz in bohr, charge) and the reference's
outputs: S and H (int1e), a set of canonical ERIs (int2e XX), the first
density/G pair (dens + RHFI2G on the core-H guess), the converged SCF energy,
orbital energies and MO coefficients (the scf.f90 loop, restated in
oracle/ref_driver.py because scf needs LAPACK), and E(AA)/E(AB)/E(MBPT2)
(mp2, on ijab_AB from the restated ao2mo transform).

uhf_<mol>.npz hold the UHF chain (the restated UHF loop over the reference's
dens/UHFI2G, its mp2 on the restated UHF records) for the reference's example
inputs NO, OH, O (triplet) and H; uhf_no.npz also the reference's own NO
outputs (examples/NO: MOLDEN orbital energies, Cui) and the energy of that Cui.
ex_o1.npz / ex_be.npz add the reference's RHF example outputs (MOLDEN orbital
energies; for O also its Cui and their energy).

props_<mol>.npz hold the reference's int1p dipole integrals for fixture <mol>,
its converged density and the dipole moment of the restated cphf.

boys.npz holds the reference Boys routine's F_0..F_Q(T) for Q in
{0,3,6,9,12} (oracle/boys_dump.f90 over the compiled auxilary module).
"""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from myqc_amd import molecules as M  # noqa: E402
from oracle.ref_driver import REF, have_ref, run_reference  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def canonical_quads(n):
    out = []
    for i in range(n):
        for j in range(i, n):
            for g in range(i, n):
                for h in range(j if g == i else g, n):
                    out.append((i, j, g, h))
    return np.array(out, dtype=np.int32).reshape(-1, 4)


def write(name, Z, xyz_ang, charge=0, max_eri=20000, seed=0, mp2=True):
    xyz = M.to_bohr(Z, xyz_ang)
    r = run_reference(Z, xyz, charge=charge, mp2=mp2)
    n = r["n"]
    q = canonical_quads(n)
    if len(q) > max_eri:
        rng = np.random.default_rng(seed)
        q = q[np.sort(rng.choice(len(q), size=max_eri, replace=False))]
    XX = r["XX"]
    vals = XX[q[:, 0], q[:, 1], q[:, 2], q[:, 3]]
    d = dict(Z=np.array(Z, dtype=np.int32), xyz=xyz, charge=np.int32(charge), n=np.int32(n),
             S=r["S"], H=r["H"], eri_idx=q, eri_val=vals, D_guess=r["D_guess"], G_guess=r["G_guess"],
             C_guess=r["C_guess"], E_scf=np.float64(r["E_scf"]), energies=r["energies"], eig=r["eig"], C=r["C"],
             iterations=np.int32(r["iterations"]), converged=np.bool_(r["converged"]), Enuc=np.float64(r["Enuc"]),
             nocc=np.int32(r["nocc"]))
    if n <= 15:
        d["XX_full"] = XX
    if mp2:
        d.update(E_aa=np.float64(r["E_aa"]), E_ab=np.float64(r["E_ab"]), E_mp2=np.float64(r["E_mp2"]))
        if n <= 15:
            d["ijab_AB"] = r["ijab_AB"]
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
    print(f"{name}: n={n} E_scf={r['E_scf']:.12f} iters={r['iterations']} conv={r['converged']}"
          + (f" E_mp2={r['E_mp2']:.12f}" if mp2 else ""), flush=True)


EXAMPLES = "/root/reference/examples"


def molden_energies(path):
    """Orbital energies of a MOLDEN file the reference wrote (its makeMOLDEN,
    src/scf/scf.f90:1343-1466): (alpha, beta) arrays."""
    import re
    txt = open(path).read()
    out = {"Alpha": [], "Beta": []}
    for m in re.finditer(r"Ene=\s*(\S+)\s*\n\s*Spin=\s*(\w+)", txt):
        out[m.group(2)].append(float(m.group(1)))
    return np.array(out["Alpha"]), np.array(out["Beta"])


def write_uhf(name, Z, xyz_ang, charge=0, mult=2, example=None):
    """UHF chain of the reference (restated loop over its dens/UHFI2G, its mp2)
    at SCF_Conv 11; with example = a directory of the reference's examples,
    also its own MOLDEN orbital energies, its Cui and their energy."""
    from oracle import ref_driver as R
    xyz = M.to_bohr(Z, xyz_ang)
    r = R.run_reference_uhf(Z, xyz, charge=charge, mult=mult, scf_conv=11, mp2=True)
    n = r["n"]
    d = dict(Z=np.array(Z, dtype=np.int32), xyz=xyz, charge=np.int32(charge), mult=np.int32(mult),
             n=np.int32(n), nalpha=np.int32(r["nalpha"]), nbeta=np.int32(r["nbeta"]), S=r["S"], H=r["H"],
             Enuc=np.float64(r["Enuc"]), E_scf=np.float64(r["E_scf"]), eig_a=r["eig_a"], eig_b=r["eig_b"],
             Ca=r["Ca"], Cb=r["Cb"], iterations=np.int32(r["iterations"]), converged=np.bool_(r["converged"]),
             E_aa=np.float64(r["E_aa"]), E_bb=np.float64(r["E_bb"]), E_ab=np.float64(r["E_ab"]),
             E_mp2=np.float64(r["E_mp2"]))
    if n <= 15:
        d["XX_full"] = r["XX"]
    if example:
        ea, eb = molden_energies(os.path.join(EXAMPLES, example, "MOLDEN"))
        Ca, Cb = R.read_text_matrices(os.path.join(EXAMPLES, example, "Cui"), n, 2)
        wd, n2, S, H, XX = R.prepare(Z, xyz, charge, mult, 1, 11, False)
        try:
            E_ex = R.uhf_energy_of(wd, H, R.enuc(Z, xyz), Ca, Cb, n)
        finally:
            subprocess.run(["rm", "-rf", wd], check=True)
        d.update(example_eig_a=ea, example_eig_b=eb, example_Ca=Ca, example_Cb=Cb, example_E=np.float64(E_ex))
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
    print(f"{name}: n={n} na={r['nalpha']} nb={r['nbeta']} E_scf={r['E_scf']:.12f} iters={r['iterations']} "
          f"conv={r['converged']} E_mp2={r['E_mp2']:.12f}" + (f" E(example Cui)={d['example_E']:.12f}" if example else ""),
          flush=True)


def write_rhf_example(name, Z, xyz_ang, example, with_cui):
    """RHF example outputs of the reference: MOLDEN orbital energies (and, with
    with_cui, its Cui and the energy of those orbitals through its dens and
    RHFI2G), beside the reference chain's own converged values."""
    from oracle import ref_driver as R
    write(name, Z, xyz_ang, mp2=False)
    g = dict(np.load(os.path.join(OUT, name + ".npz")))
    ea, _ = molden_energies(os.path.join(EXAMPLES, example, "MOLDEN"))
    g["example_eig"] = ea
    if with_cui:
        n = int(g["n"])
        xyz = M.to_bohr(Z, xyz_ang)
        C = R.read_text_matrices(os.path.join(EXAMPLES, example, "Cui"), n, 1)[0]
        wd, _, S, H, XX = R.prepare(Z, xyz, 0, 1, 0, 11, False)
        try:
            g["example_E"] = np.float64(R.rhf_energy_of(wd, H, R.enuc(Z, xyz), C, n))
        finally:
            subprocess.run(["rm", "-rf", wd], check=True)
        g["example_C"] = C
    np.savez_compressed(os.path.join(OUT, name + ".npz"), **g)
    print(f"{name}: example eig {ea[:3]}" + (f" E(example Cui)={g['example_E']:.12f}" if with_cui else ""))


def write_props(name, mol):
    """The reference's int1p dipole integrals (Pxuv, three list-directed
    records) for fixture `mol`, its converged RHF density, and the dipole
    moment of cphf_dipole_rhf (cphf.f90:106-121) restated on them.  The
    reference's cphf itself cannot run under flang's runtime: it opens
    basinfo on unit 100, closes unit 1 instead (cphf.f90:46-49) and then reads
    basinfo again where it reopens unit 100 on Pxuv (:97-100)."""
    from oracle import ref_driver as R
    g = dict(np.load(os.path.join(OUT, mol + ".npz")))
    Z = [int(z) for z in g["Z"]]
    xyz = g["xyz"]
    n, nocc = int(g["n"]), int(g["nocc"])
    wd, _, S, H, XX = R.prepare(Z, xyz, int(g["charge"]), 1, 0, 11, False)
    try:
        R._run(wd, "int1p")
        P = np.array(R.read_text_matrices(os.path.join(wd, "Pxuv"), n, 3))
    finally:
        subprocess.run(["rm", "-rf", wd], check=True)
    C = g["C"]
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    mu = R.dipole_rhf(P, D, Z, xyz)
    np.savez_compressed(os.path.join(OUT, name + ".npz"), Z=g["Z"], xyz=xyz, n=np.int32(n), S=S, D=D, P=P, mu=mu)
    print(f"{name}: n={n} dipole {mu}")


def boys():
    ftab = os.path.join(ROOT, "myqc_amd", "data", "Ftab")
    subprocess.run(["cp", ftab, os.path.join(REF, "Ftab")], check=True)
    txt = subprocess.run([os.path.join(REF, "boys_dump")], cwd=REF, capture_output=True, text=True, check=True).stdout
    Q, T, F = [], [], []
    for line in txt.splitlines():
        t = line.split()
        Q.append(int(t[0]))
        T.append(float(t[1]))
        f = np.zeros(13)
        vals = [float(x) for x in t[2:]]
        f[: len(vals)] = vals
        F.append(f)
    np.savez_compressed(os.path.join(OUT, "boys.npz"), Q=np.array(Q, dtype=np.int32), T=np.array(T), F=np.array(F))
    print(f"boys: {len(Q)} points")


def main():
    if not have_ref():
        sys.exit("reference programs missing: run `make -C oracle` in a container with /root/reference")
    boys()
    write("co2", *M.co2())
    write("heh", [2, 1], np.array([[0, 0, 0], [0, 0, 1.0]]), charge=1)
    write("h2o", *M.water())
    write("h2o_2", *M.water_cluster(2, seed=2, box=4.0))
    write("h2o_4", *M.water_cluster(4, seed=4, box=5.5))
    # the reference's own examples (src inputs under examples/)
    write_uhf("uhf_no", [7, 8], np.array([[0, 0, 0], [0, 0, 1.160871722356318]]), mult=2, example="NO")
    write_uhf("uhf_oh", [8, 1], np.array([[0, 0, 0], [0, 0, 1.0]]), mult=2)
    write_uhf("uhf_o3", [8], np.array([[0.0, 0.0, 0.0]]), mult=3)
    write_uhf("uhf_h", [1], np.array([[0.0, 0.0, 0.0]]), mult=2)
    write_rhf_example("ex_o1", [8], np.array([[0.0, 0.0, 0.0]]), "O/singlet", True)
    write_rhf_example("ex_be", [4], np.array([[0.0, 0.0, 0.0]]), "Be", False)
    for mol in ("co2", "h2o", "heh", "h2o_2", "mol2"):
        write_props("props_" + mol, mol)
    rng = np.random.default_rng(4096)
    k = 0
    while k < 6:
        Z, x = M.random_molecule(rng)
        try:
            write(f"mol{k}", Z, x, mp2=True)
            k += 1
        except Exception as e:  # reference failure (e.g. no convergence) -> skip that draw
            print("skip:", e)


if __name__ == "__main__":
    main()

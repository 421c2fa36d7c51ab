"""Molecule geometries for the RHF/MP2 two-electron path.

The reference reads a ZMAT (Cartesian, Angstrom by default), recentres it on a
mass-weighted centre of mass and converts to bohr with its own factor
(src/parser/parser.f90:14,458-487); the host programs then read ``nucpos`` in
bohr (src/myQC/env.f90:64-67).  ``to_bohr`` restates that conversion with the
accumulator initialised (src/parser/parser.f90:472 leaves it uninitialised,
which is benign for translation-invariant energies).

Generators for the synthetic workloads:
  * ``co2()`` -- linear CO2, C at the origin, O at +-1.16 A on z (15 functions).
  * ``water_cluster(nw, seed)`` -- nw rigid waters (O-H 0.96 A, HOH 104.5 deg),
    O uniform in a 12 A cube with O-O >= 2.6 A rejection, uniform random
    orientation from a normalised 4-normal quaternion; atom order O,H,H.
  * ``random_molecule(rng)`` -- 2-8 atoms from H, He, Li-Ne on a random tree,
    bonds U[0.9,1.6] A, nonbonded pairs >= 0.9 A, made closed-shell by adding
    (or substituting) one H.
"""
from __future__ import annotations

import numpy as np

# src/parser/parser.f90:14
A2B = 1.8897161646320724
# src/parser/parser.f90:459 (H..Ne)
MASS = [1.0, 4.0, 7.0, 9.0, 11.0, 12.0, 14.0, 16.0, 19.0, 20.0]
SYMBOLS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]


def to_bohr(Z, xyz_ang, units_bohr: bool = False) -> np.ndarray:
    """Recentre on the mass-weighted centre and convert to bohr, returning an
    (nnuc, 3) float64 array (src/parser/parser.f90:467-487)."""
    xyz = np.array(xyz_ang, dtype=np.float64).reshape(-1, 3).copy()
    m = np.array([MASS[z - 1] for z in Z])
    com = np.zeros(3)
    tot = 0.0
    for i in range(len(Z)):
        com += m[i] * xyz[i]
        tot += m[i]
    com = com / tot
    xyz = xyz - com
    if not units_bohr:
        xyz = xyz * A2B
    return xyz


def co2():
    return [6, 8, 8], np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.16], [0.0, 0.0, -1.16]])


def water():
    Z, x = _water_frame()
    return Z, x


def _water_frame():
    r = 0.96
    th = np.deg2rad(104.5) / 2.0
    x = np.array([[0.0, 0.0, 0.0],
                  [r * np.sin(th), 0.0, r * np.cos(th)],
                  [-r * np.sin(th), 0.0, r * np.cos(th)]])
    return [8, 1, 1], x


def _quat_rot(q):
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def water_cluster(nw: int, seed: int, box: float = 12.0, rmin: float = 2.6):
    """nw waters at seeded uniform O positions in a ``box`` A cube (O-O >= rmin)."""
    rng = np.random.default_rng(seed)
    Zw, frame = _water_frame()
    Os = []
    tries = 0
    while len(Os) < nw:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("water_cluster: could not place waters")
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - o) >= rmin for o in Os):
            Os.append(p)
    Z = []
    xyz = []
    for o in Os:
        R = _quat_rot(rng.normal(size=4))
        for zi, f in zip(Zw, frame):
            Z.append(zi)
            xyz.append(o + R @ f)
    return Z, np.array(xyz)


def random_molecule(rng: np.random.Generator, max_atoms: int = 8):
    """2..max_atoms atoms from H..Ne on a random tree, made closed-shell."""
    Z, xyz, _ = random_molecule_tree(rng, max_atoms)
    return Z, xyz


def random_molecule_tree(rng: np.random.Generator, max_atoms: int = 8):
    """random_molecule with its tree: parent[k] is the atom atom k was bonded
    to when placed (-1 for the first); the same draws from rng."""
    while True:
        na = int(rng.integers(2, max_atoms + 1))
        Z = [int(rng.integers(1, 11)) for _ in range(na)]
        xyz = [np.zeros(3)]
        parent = [-1]
        ok = True
        for k in range(1, na):
            placed = False
            for _ in range(100):
                j = int(rng.integers(0, k))
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                p = xyz[j] + d * rng.uniform(0.9, 1.6)
                if all(np.linalg.norm(p - q) >= 0.9 for q in xyz):
                    xyz.append(p)
                    parent.append(j)
                    placed = True
                    break
            if not placed:
                ok = False
                break
        if not ok:
            continue
        if sum(Z) % 2 == 1:
            if na < max_atoms:
                placed = False
                for _ in range(100):
                    j = int(rng.integers(0, na))
                    d = rng.normal(size=3)
                    d /= np.linalg.norm(d)
                    p = xyz[j] + d * rng.uniform(0.9, 1.6)
                    if all(np.linalg.norm(p - q) >= 0.9 for q in xyz):
                        xyz.append(p)
                        parent.append(j)
                        Z.append(1)
                        placed = True
                        break
                if not placed:
                    continue
            else:
                Z[-1] = 1 if Z[-1] != 1 else 2
        if sum(Z) % 2 == 1:
            continue
        return Z, np.array(xyz), parent


def nbasis(Z) -> int:
    """STO-3G function count: 1 for H/He, 5 for Li..Ne."""
    return sum(1 if z <= 2 else 5 for z in Z)

"""MOLDEN files as makeMOLDEN writes them (src/scf/scf.f90:1343-1466):
parsing and comparison for the tests."""
import numpy as np


def parse_molden(text):
    """Sections of a MOLDEN file as makeMOLDEN writes it
    (src/scf/scf.f90:1343-1466): atom lines, GTO tokens, and orbitals
    (energy, spin, occupation, coefficients)."""
    lines = text.splitlines()
    i = 0
    assert lines[i].strip() == "[Molden Format]" and lines[i + 1].strip() == "[ATOMS] AU"
    i += 2
    atoms = []
    while lines[i].strip() != "[Molden Format]":
        t = lines[i].split()
        atoms.append((t[0], int(t[1]), int(t[2]), [float(x) for x in t[3:6]]))
        i += 1
    assert lines[i + 1].strip() == "[GTO]"
    i += 2
    gto = []
    while lines[i].strip() != "[MO]":
        gto.append(lines[i].split())
        i += 1
    i += 1
    mos = []
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        assert lines[i].strip() == "Sym= A", lines[i]
        e = float(lines[i + 1].split()[1])
        spin = lines[i + 2].split()[1]
        occ = lines[i + 3].split()[1]
        i += 4
        c = []
        while i < len(lines) and lines[i].strip() and lines[i].split()[0].isdigit():
            c.append(float(lines[i].split()[1]))
            i += 1
        mos.append((e, spin, occ, np.array(c)))
    return atoms, gto, mos


def _axial_invariants(C, blocks):
    """Per orbital (column of C) and per atom block of AO indices
    (s..., px, py, pz or s only): |c_s|, |c_pz| and |(c_px, c_py)|."""
    rows = []
    for blk in blocks:
        if len(blk) == 1:
            rows.append(np.abs(C[blk[0]]))
        else:
            s = blk[:-3]
            px, py, pz = blk[-3:]
            rows += [np.abs(C[k]) for k in s]
            rows.append(np.abs(C[pz]))
            rows.append(np.hypot(C[px], C[py]))
    return np.array(rows)


def compare_molden(ours, ref, e_tol, c_tol, linear_z=None):
    """linear_z: for a molecule on the z axis, the AO indices of each atom
    (1 for H/He; 1s, 2s, px, py, pz for Li..Ne), to compare orbitals through
    their axial invariants."""
    a1, g1, m1 = parse_molden(ours)
    a2, g2, m2 = parse_molden(ref)
    assert [a[:3] for a in a1] == [a[:3] for a in a2]
    assert np.max(np.abs(np.array([a[3] for a in a1]) - np.array([a[3] for a in a2]))) <= 1e-8
    # the GTO blocks: same lines, text tokens equal, numbers equal (list-directed
    # output differs between Fortran runtimes only in spacing and digits shown)
    g1 = [t for t in g1 if t]
    g2 = [t for t in g2 if t]
    assert len(g1) == len(g2)
    for t1, t2 in zip(g1, g2):
        assert len(t1) == len(t2)
        for x, y in zip(t1, t2):
            try:
                fx, fy = float(x.replace("D", "E")), float(y.replace("D", "E"))
            except ValueError:
                assert x == y
                continue
            assert abs(fx - fy) <= 1e-12 * max(1.0, abs(fy))
    assert len(m1) == len(m2)
    for spin in ("Alpha", "Beta"):
        o1 = [m for m in m1 if m[1] == spin]
        o2 = [m for m in m2 if m[1] == spin]
        assert len(o1) == len(o2)
        if not o1:
            continue
        e1 = np.array([m[0] for m in o1])
        e2 = np.array([m[0] for m in o2])
        assert np.max(np.abs(e1 - e2)) <= e_tol
        assert [m[2] for m in o1] == [m[2] for m in o2]
        C1 = np.array([m[3] for m in o1]).T
        C2 = np.array([m[3] for m in o2]).T
        if linear_z is not None:
            # a molecule on the z axis: a UHF solution that breaks the axial
            # symmetry is determined only up to a rotation about z, so each
            # orbital is compared through its invariants per atom (|s|,
            # |p_z|, and the length of (p_x, p_y))
            inv1, inv2 = _axial_invariants(C1, linear_z), _axial_invariants(C2, linear_z)
            assert np.max(np.abs(inv1 - inv2)) <= c_tol, (spin, np.max(np.abs(inv1 - inv2)))
            continue
        # coefficients up to sign and to rotations within a degenerate level:
        # the projector onto each level's orbitals
        k = 0
        while k < len(e2):
            j = k + 1
            while j < len(e2) and abs(e2[j] - e2[k]) <= 1e-6:
                j += 1
            P1 = C1[:, k:j] @ C1[:, k:j].T
            P2 = C2[:, k:j] @ C2[:, k:j].T
            assert np.max(np.abs(P1 - P2)) <= c_tol, (spin, k, j, np.max(np.abs(P1 - P2)))
            k = j


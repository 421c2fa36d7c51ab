"""The many-system RHF (myqc_rhf_scf_many): several systems of one basis size
solved in lockstep on one stream must give what the single-system solve
(myqc_rhf_scf / myqc_rhf_scf_density) gives for each of them: the same
energy within 1e-8 Eh, orbitals spanning the same occupied space, and the
orbital energies.  The loop it batches is src/scf/scf.f90's RHF (DIIS,
convergence on dE and the density change); the batching has no counterpart
in the reference."""
import numpy as np
import pytest

from myqc_amd import molecules as M
from myqc_amd.basis import build_basis, read_ftab
from myqc_amd.hip import FP32, FP64, Context, MyqcError, enuc, rhf_many

pytestmark = pytest.mark.gpu


def _systems(waters, seeds):
    out = []
    for s in seeds:
        Z, x = M.water_cluster(waters, seed=s)
        xyz = M.to_bohr(Z, x)
        out.append((Z, xyz, build_basis(Z), sum(Z) // 2, enuc(Z, xyz)))
    return out


def _prepare(ctxs, systems):
    ftab = read_ftab()
    HS = []
    for c, (Z, xyz, basis, _, _) in zip(ctxs, systems):
        c.set_system(Z, xyz, basis, ftab)
        S, H = c.int1e()
        c.eri_build(1e-12)
        HS.append((H, S))
    return HS


def _occupied_projector(C, S, nocc):
    Co = C[:, :nocc]
    return Co @ Co.T @ S


@pytest.mark.parametrize("precision,learned", [(FP64, False), (FP32, False), (FP64, True), (FP32, True)])
def test_many_matches_single_system_solves(ctx, precision, learned):
    from myqc_amd import guess as G
    systems = _systems(6, [6, 7, 8, 9, 10])
    ctxs = [Context(0) for _ in systems]
    try:
        if learned and not G.load(ctxs[0]):
            pytest.skip("no trained guess network")
        if learned:
            for c in ctxs[1:]:
                assert G.load(c)
        HS = _prepare(ctxs, systems)
        D0 = [c.guess_density(s[0], s[1]) for c, s in zip(ctxs, systems)] if learned else None
        res = rhf_many(ctxs, [h for h, _ in HS], [s for _, s in HS], [s[3] for s in systems],
                       [s[4] for s in systems], D_guesses=D0, precision=precision)
        for k, (c, s) in enumerate(zip(ctxs, systems)):
            H, S = HS[k]
            r1 = c.rhf(H, S, s[3], s[4], e_tol=1e-10, d_tol=1e-8, max_iter=200, precision=precision,
                       D_guess=D0[k] if learned else None)
            print(f"[many] system {k}: E {res[k]['E']:.10f} ({res[k]['iterations']} it) single "
                  f"{r1['E']:.10f} ({r1['iterations']} it)", flush=True)
            assert abs(res[k]["E"] - r1["E"]) <= 1e-8
            assert abs(res[k]["iterations"] - r1["iterations"]) <= 2
            P1 = _occupied_projector(r1["C"], S, s[3])
            Pm = _occupied_projector(res[k]["C"], S, s[3])
            assert np.max(np.abs(Pm - P1)) <= 1e-6
            assert np.max(np.abs(res[k]["eig"] - r1["eig"])) <= 1e-6
            C = res[k]["C"]
            assert np.max(np.abs(C.T @ S @ C - np.eye(len(C)))) <= 1e-9
    finally:
        for c in ctxs:
            c.close()


def test_many_40water_clusters_match_single_fp64(ctx):
    """Four clusters of the headline size (n = 280), learned + mixed when the
    network is there, against each one's plain FP64 single-system solve."""
    from myqc_amd import guess as G
    systems = _systems(40, [40, 41, 42, 43])
    ctxs = [Context(0) for _ in systems]
    try:
        learned = all(G.load(c) for c in ctxs)
        HS = _prepare(ctxs, systems)
        D0 = [c.guess_density(s[0], s[1]) for c, s in zip(ctxs, systems)] if learned else None
        res = rhf_many(ctxs, [h for h, _ in HS], [s for _, s in HS], [s[3] for s in systems],
                       [s[4] for s in systems], D_guesses=D0, precision=FP32, want_orbitals=False)
        for k, (c, s) in enumerate(zip(ctxs, systems)):
            H, S = HS[k]
            r1 = c.rhf(H, S, s[3], s[4], e_tol=1e-10, d_tol=1e-8, max_iter=200, precision=FP64)
            print(f"[many40] cluster {40 + k}: E {res[k]['E']:.10f} ({res[k]['iterations']} it) fp64 "
                  f"{r1['E']:.10f} ({r1['iterations']} it)", flush=True)
            assert abs(res[k]["E"] - r1["E"]) <= 1e-8
    finally:
        for c in ctxs:
            c.close()


def test_many_refuses_mixed_sizes_and_repeated_contexts(ctx):
    a, b = Context(0), Context(0)
    try:
        s6 = _systems(6, [6])[0]
        s5 = _systems(5, [5])[0]
        HS = _prepare([a, b], [s6, s5])
        with pytest.raises(MyqcError):
            rhf_many([a, b], [HS[0][0], HS[1][0]], [HS[0][1], HS[1][1]], [s6[3], s5[3]], [s6[4], s5[4]])
        with pytest.raises(MyqcError):
            rhf_many([a, a], [HS[0][0]] * 2, [HS[0][1]] * 2, [s6[3]] * 2, [s6[4]] * 2)
        # the contexts still work on their own streams afterwards
        r = a.rhf(HS[0][0], HS[0][1], s6[3], s6[4])
        assert np.isfinite(r["E"])
    finally:
        a.close()
        b.close()


def test_pipelined_solver_matches_single_solves(ctx):
    """myqc_amd.many.ManySolver: three batches (of 3, 3 and 2 molecules) on
    two alternating context sets, each batch set up while the previous one
    solves; every energy equals the single-system solve's."""
    from myqc_amd.many import ManySolver
    seeds = [[11, 12, 13], [14, 15, 16], [17, 18]]
    batches = []
    for bs in seeds:
        b = []
        for s in bs:
            Z, x = M.water_cluster(6, seed=s)
            b.append((Z, M.to_bohr(Z, x)))
        batches.append(b)
    solver = ManySolver(3, precision=FP32, workers=3, pipeline=True)
    try:
        results = list(solver.solve(batches))
    finally:
        solver.close()
    assert [len(r) for r in results] == [3, 3, 2]
    for b, res in zip(batches, results):
        for (Z, xyz), r in zip(b, res):
            ctx.set_system(Z, xyz)
            S, H = ctx.int1e()
            ctx.eri_build(1e-12)
            r1 = ctx.rhf(H, S, sum(Z) // 2, enuc(Z, xyz), precision=FP64)
            assert abs(r["E"] - r1["E"]) <= 1e-8

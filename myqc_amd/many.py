"""Many molecules of one basis size, solved in batches of up to 32 by the
lockstep RHF (myqc_rhf_scf_many).  Each batch is set up (one-electron
integrals, the screened two-electron integrals and, with the learned guess,
the network's density) on host threads, one context and HIP stream per
molecule, then solved in one lockstep SCF.  With pipeline=True the next
batch's set-up runs beside the current batch's SCF, on a second set of
contexts (the two sets take turns); with the GPU's few hardware queues
shared by both, that pays only when the set-up is mostly host work.

    solver = ManySolver(k=16, learned=True, precision=FP32)
    for results in solver.solve(batches):   # batches: lists of (Z, xyz_bohr)
        ...
    solver.close()

The reference solves one molecule per run (src/scf/scf.f90); this driver is
the host-side scheduling of the batched path and has no counterpart there.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

from .basis import build_basis, read_ftab
from .hip import FP64, Context, enuc, rhf_many


class ManySolver:
    """Two sets of k contexts on one device, used in turn by consecutive
    batches; `workers` host threads set a batch up.  learned: start from the
    trained network's density (each context loads it once); precision: FP64
    or FP32 (mixed); pipeline: set up the next batch while the current one
    solves."""

    def __init__(self, k: int, device: int = 0, learned: bool = False, precision: int = FP64,
                 workers: int = 8, schwarz_tol: float = 1e-12, e_tol: float = 1e-10, d_tol: float = 1e-8,
                 max_iter: int = 200, pipeline: bool = False):
        if not 1 <= k <= 32:
            raise ValueError(f"ManySolver: k={k} (1..32)")
        self.k, self.learned, self.precision, self.pipeline = k, learned, precision, pipeline
        self.tol, self.e_tol, self.d_tol, self.max_iter = schwarz_tol, e_tol, d_tol, max_iter
        self.ftab = read_ftab()
        self.sets = [[Context(device) for _ in range(k)] for _ in range(2)]
        if learned:
            from . import guess as G
            for s in self.sets:
                for c in s:
                    if not G.load(c):
                        self.close()
                        raise RuntimeError("ManySolver: no trained guess network (myqc_amd/data/guess_mlp.npz)")
        self.pool = ThreadPoolExecutor(max_workers=max(1, min(workers, k)))
        self.lead = ThreadPoolExecutor(max_workers=1)

    def _setup_one(self, ctx, mol):
        Z, xyz = mol
        ctx.set_system(Z, xyz, build_basis(list(Z)), self.ftab)
        S, H = ctx.int1e()
        ctx.eri_build(self.tol)
        D0 = ctx.guess_density(Z, xyz) if self.learned else None
        return H, S, D0, sum(Z) // 2, enuc(Z, xyz)

    def _setup(self, s, mols):
        ctxs = self.sets[s]
        t0 = time.perf_counter()
        out = list(self.pool.map(lambda i: self._setup_one(ctxs[i], mols[i]), range(len(mols))))
        return out, time.perf_counter() - t0

    def solve(self, batches, want_orbitals: bool = False):
        """Yields, per batch, the list of dict(E, C, eig, iterations) of
        hip.rhf_many.  Every batch holds 1..k molecules of one basis size."""
        batches = list(batches)
        for b in batches:
            if not 1 <= len(b) <= self.k:
                raise ValueError(f"ManySolver.solve: a batch of {len(b)} molecules (1..{self.k})")
        if not batches:
            return
        self.timings = []   # per batch: (set-up seconds, lockstep SCF seconds)
        fut = self.lead.submit(self._setup, 0, batches[0])
        for b in range(len(batches)):
            hs, t_setup = fut.result()
            if b + 1 < len(batches) and self.pipeline:
                fut = self.lead.submit(self._setup, (b + 1) % 2, batches[b + 1])
            ctxs = self.sets[b % 2][:len(hs)]
            t0 = time.perf_counter()
            res = rhf_many(ctxs, [h[0] for h in hs], [h[1] for h in hs], [h[3] for h in hs], [h[4] for h in hs],
                           D_guesses=[h[2] for h in hs] if self.learned else None, e_tol=self.e_tol,
                           d_tol=self.d_tol, max_iter=self.max_iter, precision=self.precision,
                           want_orbitals=want_orbitals)
            self.timings.append((t_setup, time.perf_counter() - t0))
            if b + 1 < len(batches) and not self.pipeline:
                fut = self.lead.submit(self._setup, (b + 1) % 2, batches[b + 1])
            yield res

    def close(self):
        for ex in ("pool", "lead"):
            if hasattr(self, ex):
                getattr(self, ex).shutdown()
        for s in getattr(self, "sets", []):
            for c in s:
                c.close()

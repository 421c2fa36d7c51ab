"""Train the learned-guess network on the GPU and save it.

    python -m myqc_amd.train_guess [--base 1000] [--perturb 20] [--epochs 60] [--out PATH]

Data: `base` seeded config-4 molecules (seed 20000) with `perturb` perturbed
copies each (every bond +-0.1 A, every bond angle +-5 degrees): 20k
geometries by default,
plus `waters` seeded clusters of 2-5 waters (seed 20002), solved by the
batched GPU SCF from the core-Hamiltonian guess.  Training:
Adam on the library's own forward/backward kernels, half the epochs at
lr 1e-3 and half at 2e-4.  Held-out check: 512 fresh molecules (seed 20001)
solved from the core-Hamiltonian guess and from the network's densities.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import guess as G
from .hip import Context


def evaluate(ctx, mols):
    ctx.batch_set(mols)
    Ec, itc, cvc = ctx.batch_rhf(e_tol=1e-10, d_tol=1e-8, max_iter=100)
    El, itl, cvl, fb = ctx.batch_rhf_learned(e_tol=1e-10, d_tol=1e-8, max_iter=100)
    both = (cvc == 1) & (cvl == 1)
    same = both & (np.abs(El - Ec) <= 1e-8)
    return dict(molecules=len(mols), core_iterations=float(np.mean(itc[cvc == 1])),
                learned_iterations=float(np.mean(itl[cvl == 1])), converged_core=int(np.sum(cvc == 1)),
                converged_learned=int(np.sum(cvl == 1)), fallback_rate=float(np.mean(fb)),
                same_state=int(np.sum(same)), both_converged=int(np.sum(both)),
                lower_with_learned=int(np.sum(both & (El < Ec - 1e-8))),
                higher_with_learned=int(np.sum(both & (El > Ec + 1e-8))),
                max_dE_same_state=float(np.max(np.abs(El - Ec)[same])) if same.any() else None)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--base", type=int, default=1000)
    ap.add_argument("--perturb", type=int, default=20)
    ap.add_argument("--waters", type=int, default=4000, help="small water clusters added to the data")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--out", default=G.DATA)
    args = ap.parse_args(argv)
    with Context(0) as ctx:
        t0 = time.perf_counter()
        mols = G.molecules(args.base, args.perturb, seed=20000) + G.water_clusters(args.waters, seed=20002)
        X, T, Mk, it, cf = G.dataset(ctx, mols)
        t_data = time.perf_counter() - t0
        print(f"[train] {len(mols)} geometries solved ({cf:.3f} converged, {it:.1f} iterations) -> {len(X)} pairs "
              f"in {t_data:.1f} s", flush=True)
        t0 = time.perf_counter()
        h1 = G.train(ctx, X, T, Mk, epochs=args.epochs // 2, batch=args.batch, lr=1e-3, seed=1)
        h2 = ctx.mlp_train(X, T, Mk, args.epochs - args.epochs // 2, args.batch, 2e-4, 2)
        t_train = time.perf_counter() - t0
        hist = np.concatenate([h1, h2])
        print(f"[train] loss {hist[0]:.3e} -> {hist[-1]:.3e} in {t_train:.1f} s", flush=True)
        ev = evaluate(ctx, G.molecules(512, 1, seed=20001))
        print("[train] held-out " + json.dumps(ev), flush=True)
        evw = evaluate(ctx, G.water_clusters(256, seed=20003))
        print("[train] held-out water clusters " + json.dumps(evw), flush=True)
        G.save(ctx, args.out, loss=hist, geometries=len(mols), pairs=len(X), train_seconds=t_train,
               data_seconds=t_data, heldout=json.dumps(ev), heldout_waters=json.dumps(evw))
        print(f"[train] saved {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

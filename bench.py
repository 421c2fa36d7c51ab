"""Benchmark: converged RHF/STO-3G solves per second on one MI355X.

Workload: the seeded 40-water cluster (280 functions; O-H 0.96 A, HOH
104.5 deg, O uniform in a 12 A cube with O-O >= 2.6 A, seed 40), Schwarz
cutoff 1e-12.  One step = one full solve from the molecule: one-electron
integrals, Schwarz factors, the screened unique two-electron integrals, and
the SCF to |dE| < 1e-10 Eh.  Forms: plain FP64 digestion; mixed
(incremental FP32 dD Fock builds with FP64 rebuilds and an FP64 finish); and
both again from the learned guess (the trained network's density, computed on
the device inside the timed solve, myqc_amd/data/guess_mlp.npz); and the
learned + mixed form batched, over 16 seeded clusters of the same size solved
side by side (one context and HIP stream each; one step = one batch of 16).
`value` is the rate of the fastest form whose energies all agree with plain
FP64 to 1e-8 Eh.
Extra legs: RHF-MP2 of 24 waters (ao2mo FP64 vs split-bf16) and the
4096-molecule batch (batched persistent launch, with and without the learned
guess, against the plain single-molecule path).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu]

`roofline` is the J/K digestion kernel's: algorithmic bytes per pass
(8 N_stored + 16 n^2) over its HIP-event time.  `roofline.traffic` is the HBM
bytes per digestion launch that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
passes measured, scaled by the same counters' reading on a plain copy of
known size (--copy-calibration runs one), as recorded in
profiles/k_digest_40water_traffic.json; null when that file is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_digest_40water_traffic.json")
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0   # dense MFMA


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def solve(ctx, Z, xyz, basis, ftab, nocc, e_nuc, precision, learned=False):
    """One full solve from the molecule; with `learned` the network's
    density (computed on the device) is the starting point."""
    ctx.set_system(Z, xyz, basis, ftab)
    S, H = ctx.int1e()
    nst = ctx.eri_build(1e-12)
    D0 = ctx.guess_density(Z, xyz) if learned else None
    r = ctx.rhf(H, S, nocc, e_nuc, e_tol=1e-10, d_tol=1e-8, max_iter=200, precision=precision, D_guess=D0)
    r["n_stored"] = nst
    return r


def concurrent_leg(waters, k, steps, warmup):
    """The learned + mixed form over a batch of k seeded clusters of the same
    size (seeds waters .. waters+k-1, the headline cluster first), solved side
    by side: one context (one HIP stream) and one host thread per cluster, so
    one cluster's latency-bound phases (purification steps, the final
    eigen-decomposition, host set-up) overlap the others' integral and
    digestion kernels.  Every energy is checked against the same cluster's
    plain FP64 solve, run beforehand outside the timed region.  One step is
    one batch of k solves; the k streams start the timed steps together and
    the time runs until the last one has finished its last solve."""
    import threading

    from myqc_amd import guess as G
    from myqc_amd import molecules as M
    from myqc_amd.basis import build_basis, read_ftab
    from myqc_amd.hip import FP32, FP64, Context, enuc

    ftab = read_ftab()
    mols = []
    for s in range(k):
        Z, x = M.water_cluster(waters, seed=waters + s)
        xyz = M.to_bohr(Z, x)
        mols.append((Z, xyz, build_basis(Z), sum(Z) // 2, enuc(Z, xyz)))
    ctxs = [Context(0) for _ in range(k)]
    for c in ctxs:
        if not G.load(c):
            for cc in ctxs:
                cc.close()
            return None
    e64 = [solve(c, *m[:3], ftab, m[3], m[4], FP64)["E"] for c, m in zip(ctxs, mols)]
    energies = [[] for _ in range(k)]
    errors = []
    start = threading.Barrier(k + 1)
    done = [0.0] * k

    def run(i):
        # warmup solves, then the timed ones back to back: each cluster's
        # stream keeps its own pace, so the host-only phases of the clusters
        # (set-up, integral ordering, schedule) do not line up
        try:
            m = mols[i]
            for _ in range(warmup):
                solve(ctxs[i], m[0], m[1], m[2], ftab, m[3], m[4], FP32, learned=True)
        except Exception as e:
            errors.append(e)
        start.wait()
        try:
            for _ in range(steps):
                r = solve(ctxs[i], m[0], m[1], m[2], ftab, m[3], m[4], FP32, learned=True)
                energies[i].append((r["E"], r["iterations"]))
        except Exception as e:  # reported after the join
            errors.append(e)
        done[i] = time.perf_counter()

    ths = [threading.Thread(target=run, args=(i,)) for i in range(k)]
    for t in ths:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    for t in ths:
        t.join()
    dt = max(done) - t0
    if errors:
        raise errors[0]
    dmax = max(abs(e[0] - e64[i]) for i in range(k) for e in energies[i])
    log(f"[bench] concurrent: {k} clusters x {steps} solves in {dt * 1e3:.1f} ms, max |E - E_fp64| = {dmax:.2e}")
    for c in ctxs:
        c.close()
    return dict(workload=f"{k} RHF/STO-3G {waters}-water clusters (seeds {waters}-{waters + k - 1}) side by side, "
                         f"one context and stream each", clusters=k, sec=dt / (steps * k), ms_per_batch=dt / steps * 1e3,
                max_dE_vs_fp64=dmax, mean_iterations=float(np.mean([e[1] for es in energies for e in es])),
                E_first=energies[0][-1][0], E_first_fp64=e64[0])


def lockstep_leg(waters, k, steps, warmup, forms):
    """Batches of k seeded clusters of the same size (seeds waters ..
    waters+k-1, the headline cluster first) solved together by
    myqc_amd.many.ManySolver: host threads set the batch up (one-electron
    integrals, the screened two-electron integrals and, in the learned form,
    the network's density; one context and stream per cluster), then one
    lockstep SCF of the k clusters (myqc_rhf_scf_many), where every step but
    the digestion passes is one launch for all clusters and the digestion
    passes fan out over four streams.  The next batch is set up on a second
    set of contexts while the current one solves.  One step is one batch of
    k full solves from the molecules.
    Every energy is checked against the same cluster's plain FP64
    single-system solve, run beforehand outside the timed region."""
    from myqc_amd import guess as G
    from myqc_amd import molecules as M
    from myqc_amd.basis import build_basis, read_ftab
    from myqc_amd.hip import FP32, FP64, Context, enuc
    from myqc_amd.many import ManySolver

    ftab = read_ftab()
    mols = []
    for s in range(k):
        Z, x = M.water_cluster(waters, seed=waters + s)
        mols.append((Z, M.to_bohr(Z, x)))
    ref = Context(0)
    have_net = G.load(ref)
    e64 = [solve(ref, Z, xyz, build_basis(Z), ftab, sum(Z) // 2, enuc(Z, xyz), FP64)["E"] for Z, xyz in mols]
    ref.close()
    out = {}
    for name, prec, learned in (("fp64", FP64, False), ("learned_mixed", FP32, True)):
        if name not in forms or (learned and not have_net):
            continue
        solver = ManySolver(k, learned=learned, precision=prec, workers=8, pipeline=True)
        # at least two warm-up batches, so that both context sets hold their
        # buffers before the timed batches
        for _ in solver.solve([mols] * max(warmup, 2)):
            pass
        t0 = time.perf_counter()
        rs = list(solver.solve([mols] * steps))
        dt = time.perf_counter() - t0
        tim = np.array(solver.timings)
        solver.close()
        log(f"[bench] lockstep {name}: per batch set-up {1e3 * tim[:, 0].mean():.1f} ms, "
            f"lockstep SCF {1e3 * tim[:, 1].mean():.1f} ms")
        dmax = max(abs(r[i]["E"] - e64[i]) for r in rs for i in range(k))
        its = float(np.mean([r[i]["iterations"] for r in rs for i in range(k)]))
        log(f"[bench] lockstep {name}: {k} clusters x {steps} batches in {dt * 1e3:.1f} ms "
            f"({k * steps / dt:.1f} solves/s), max |E - E_fp64| = {dmax:.2e}, mean iterations {its:.2f}")
        out[name] = dict(sec=dt / (steps * k), ms_per_batch=dt / steps * 1e3, max_dE_vs_fp64=dmax,
                         mean_iterations=its, E_first=rs[-1][0]["E"], E_first_fp64=e64[0],
                         setup_ms=1e3 * float(tim[:, 0].mean()), scf_ms=1e3 * float(tim[:, 1].mean()))
    return dict(workload=f"{k} RHF/STO-3G {waters}-water clusters (seeds {waters}-{waters + k - 1}) per batch, "
                         f"set up side by side, one lockstep SCF per batch", clusters=k, forms=out)


def cpu_baseline(Z, xyz, basis, ftab, n_iter, budget_s=12.0, batch_mols=None, cores=16):
    """CPU numbers beside the GPU rate (oracle/baseline.py), on this host:

    * the headline solve by the oracle's port, single core: int2e (proc2e)
      from a stratified sample of bra set pairs (ratio estimator on a
      screen-based cost model, bootstrap standard error), plus the RHFI2G n^4
      pass timed at n = 56 and 84 and scaled to n, times the solve's
      iteration count;
    * the reference's own programs (built from their sources into
      oracle/_ref) against the port: int2e on CO2, H2O and (H2O)2, one
      RHFI2G pass at n = 56; the headline solve scaled by those ratios;
    * the batch: the reference's solve (int1e, int2e, the SCF loop spawning
      dens and RHFI2G) over molecules of the seeded batch, one process per
      core, for about 20 s."""
    from myqc_amd import molecules as M
    from oracle import baseline as BL
    from oracle import oracle as O
    sysm = O.System(Z, xyz, basis, ftab)
    est = BL.stratified_int2e(sysm, basis, xyz, budget_s=budget_s, seed=2)
    g = BL.rhfi2g_time(56, 84, basis.norb)
    t_port = est["seconds"] + n_iter * g["scaled"]
    out = dict(value=1.0 / t_port, unit="RHF solves/s", cores=1, kind="port",
               sample=(f"oracle port: int2e on {est['pairs_sampled']} of {est['pairs_total']} bra set pairs in "
                       f"{est['strata']} cost strata ({est['sample_seconds']:.1f} s timed), "
                       f"{est['seconds']:.0f} +- {est['stderr']:.0f} s (bootstrap); RHFI2G n^4 pass timed at "
                       f"n=56, 84 (exponent {g['exponent']:.2f}) and scaled to n={basis.norb}: "
                       f"{g['scaled']:.1f} s x {n_iter} iterations; {t_port:.0f} s per solve"),
               int2e_s=est["seconds"], int2e_stderr_s=est["stderr"], rhfi2g_pass_s=g["scaled"])
    mols = []
    for k in (1, 2):
        Zw, xw = M.water_cluster(k, seed=k)
        mols.append((f"(H2O){k}", Zw, M.to_bohr(Zw, xw)))
    mols.append(("CO2", [6, 8, 8], M.to_bohr([6, 8, 8], np.array([[0, 0, 0], [0, 0, 1.16], [0, 0, -1.16]]))))
    cal = BL.calibrate(mols)
    if cal:
        t_ref = est["seconds"] * cal["int2e_ref_over_port"] + n_iter * g["scaled"] * cal["rhfi2g_ref_over_port"]
        out["reference_scaled"] = {"value": 1.0 / t_ref, "unit": "RHF solves/s", "cores": 1,
                                   "kind": "reference (scaled)",
                                   "int2e_ref_over_port": cal["int2e_ref_over_port"],
                                   "rhfi2g_ref_over_port": cal["rhfi2g_ref_over_port"],
                                   "calibration": cal["rows"], "seconds_per_solve": t_ref}
    if batch_mols:
        rate = BL.batch_reference_rate(batch_mols, cores, budget_s=20.0)
        out["batch_reference"] = {"value": rate["value"], "unit": "converged solves/s", "cores": cores,
                                  "kind": "reference",
                                  "sample": f"{rate['molecules']} molecules of the seeded batch finished in "
                                            f"{rate['seconds']:.1f} s on {cores} processes, "
                                            f"{rate['converged']} converged, {rate['mean_seconds']:.1f} s each "
                                            f"(the quicker ones: the rate leans in the CPU's favour)"}
    return out


def mp2_leg(ctx, waters, reps=2):
    """RHF-MP2 of a seeded water cluster: the ao2mo quarter transforms in FP64
    (hand-written FP64 MFMA kernels) and split-bf16 (bf16 MFMA, FP32
    accumulation), each timed over its quarter-step launches with HIP events;
    E(MBPT2) of both forms."""
    from myqc_amd import molecules as M
    from myqc_amd.basis import build_basis, read_ftab
    from myqc_amd.hip import BF16X3, FP64, T_AO2MO, enuc
    Z, x = M.water_cluster(waters, seed=waters)
    xyz = M.to_bohr(Z, x)
    basis = build_basis(Z)
    nocc = sum(Z) // 2
    ctx.set_system(Z, xyz, basis, read_ftab())
    S, H = ctx.int1e()
    ctx.eri_build(1e-12)
    r = ctx.rhf(H, S, nocc, enuc(Z, xyz), e_tol=1e-10, d_tol=1e-8, max_iter=200)
    n = basis.norb
    out = {"workload": f"RHF-MP2/STO-3G {waters}-water cluster (seed {waters}), n={n}, o={nocc}, v={n - nocc}",
           "E_scf": r["E"]}
    for name, prec in (("fp64", FP64), ("bf16x3", BF16X3)):
        ctx.ao2mo(r["C"], nocc, prec, download=False)   # warmup
        ctx.timer_reset(True)
        for _ in range(reps):
            ctx.ao2mo(r["C"], nocc, prec, download=False)
        cnt, ms, flops = ctx.timer_get(T_AO2MO)
        ctx.timer_reset(False)
        _, _, e2 = ctx.mp2(r["eig"], nocc)
        tf = (flops / cnt) / (ms / cnt * 1e-3) / 1e12
        peak = FP64_PEAK_TFLOPS if prec == FP64 else BF16_PEAK_TFLOPS
        out[name] = {"E_mp2": e2, "ao2mo_ms": ms / cnt, "TFLOPs": tf, "peak": peak, "frac": tf / peak}
        log(f"[bench] mp2 {name}: E(MBPT2)={e2:.10f} ao2mo {ms / cnt:.2f} ms {tf:.1f} TFLOP/s")
    out["dE_mp2_bf16x3_vs_fp64"] = out["bf16x3"]["E_mp2"] - out["fp64"]["E_mp2"]
    return out


def batch_leg(ctx, nmol, steps=3, plain_sample=64):
    """The seeded 4096-molecule batch (2-8 atoms from H..Ne on random trees):
    converged RHF solves/s of the batched persistent launch (graph replays,
    integrals + SCF from the core-Hamiltonian guess) against the plain
    single-molecule FP64 path timed on a sample of the same molecules."""
    from myqc_amd import molecules as M
    from myqc_amd.hip import MyqcError, enuc
    rng = np.random.default_rng(nmol)
    mols = []
    for _ in range(nmol):
        Z, x = M.random_molecule(rng)
        mols.append((Z, M.to_bohr(Z, x)))
    ctx.batch_set(mols)
    t0 = time.perf_counter()
    E, it, cv = ctx.batch_rhf(e_tol=1e-10, d_tol=1e-8, max_iter=100)   # tunes launch shapes, captures the graph
    t_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    for _ in range(steps):
        E, it, cv = ctx.batch_rhf(e_tol=1e-10, d_tol=1e-8, max_iter=100)
    dt = (time.perf_counter() - t0) / steps
    conv = int(np.sum(cv == 1))
    learned = None
    from myqc_amd import guess as G
    if G.load(ctx):
        El, itl, cvl, fb = ctx.batch_rhf_learned(e_tol=1e-10, d_tol=1e-8, max_iter=100)   # captures its graph
        t0 = time.perf_counter()
        for _ in range(steps):
            El, itl, cvl, fb = ctx.batch_rhf_learned(e_tol=1e-10, d_tol=1e-8, max_iter=100)
        dtl = (time.perf_counter() - t0) / steps
        both = (cv == 1) & (cvl == 1)
        learned = {"solves_per_s": int(np.sum(cvl == 1)) / dtl, "ms_per_batch": dtl * 1e3,
                   "converged": int(np.sum(cvl == 1)), "mean_iterations": float(np.mean(itl[cvl == 1])),
                   "fallback_rate": float(np.mean(fb)),
                   "same_state_as_core_guess": int(np.sum(both & (np.abs(El - E) <= 1e-8))),
                   "both_converged": int(np.sum(both))}
    info = ctx.batch_info()
    # plain path on a sample: set_system + int1e + screened ERIs + SCF per molecule
    idx = np.random.default_rng(1).choice(nmol, size=min(plain_sample, nmol), replace=False)
    t0 = time.perf_counter()
    done, worst = 0, 0.0
    for m in idx:
        Z, xyz = mols[m]
        ctx.set_system(Z, xyz)
        S, H = ctx.int1e()
        ctx.eri_build(1e-12)
        try:
            r = ctx.rhf(H, S, sum(Z) // 2, enuc(Z, xyz), e_tol=1e-10, d_tol=1e-8, max_iter=100)
        except MyqcError:
            continue
        done += 1
        if cv[m] == 1:
            worst = max(worst, abs(r["E"] - E[m]))
    t_plain = (time.perf_counter() - t0) / len(idx)
    out = {"workload": f"{nmol} seeded molecules (seed {nmol}), 2-8 atoms H..Ne, n = {min(ctx.batch_norb)}.."
                       f"{max(ctx.batch_norb)} (mean {np.mean(ctx.batch_norb):.1f})",
           "batched_solves_per_s": conv / dt, "ms_per_batch": dt * 1e3, "converged": conv,
           "mean_iterations": float(np.mean(it[cv == 1])), "first_call_ms": t_first * 1e3,
           "plain_fp64_solves_per_s": 1.0 / t_plain, "plain_sample": int(len(idx)),
           "speedup_vs_plain": (conv / dt) * t_plain, "max_dE_vs_plain": worst,
           "learned": learned,
           "eri_threads": info["eri_threads"], "scf_threads": info["scf_threads"],
           "stored_integrals": info["n_integrals"]}
    log(f"[bench] batch {nmol}: {out['batched_solves_per_s']:.0f} solves/s ({dt * 1e3:.1f} ms per batch, "
        f"{conv} converged, {out['mean_iterations']:.1f} iterations); plain {out['plain_fp64_solves_per_s']:.1f}/s")
    return out


def copy_calibration(mb):
    """One device-to-device copy of `mb` MiB, for scaling HBM counters."""
    import torch
    src = torch.ones(mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: {mb} MiB read + {mb} MiB written")


def measured_traffic(waters):
    if waters != 40 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    return t.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--waters", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed,learned,learned_mixed",
                    help="comma list of fp64, mixed, learned, learned_mixed")
    ap.add_argument("--mp2-waters", type=int, default=24, help="water cluster of the MP2 leg (0: skip)")
    ap.add_argument("--batch", type=int, default=4096, help="molecules in the batched leg (0: skip)")
    ap.add_argument("--concurrent", type=int, default=16,
                    help="clusters solved side by side in the concurrent learned+mixed form (0: skip)")
    ap.add_argument("--lockstep", type=int, default=32,
                    help="clusters per batch of the lockstep many-system SCF (0: skip; at most 32)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="first run one device copy of MB MiB (scale for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested); the path is single-GPU by design",
              file=sys.stderr)
        sys.exit(2)

    from myqc_amd import molecules as M
    from myqc_amd.basis import build_basis, read_ftab
    from myqc_amd.hip import FP32, FP64, T_ERI, T_JK, Context, enuc

    Z, x = M.water_cluster(args.waters, seed=args.waters)
    xyz = M.to_bohr(Z, x)
    basis = build_basis(Z)
    ftab = read_ftab()
    nocc = sum(Z) // 2
    e_nuc = enuc(Z, xyz)
    log(f"[bench] {args.waters} waters, n={basis.norb}, nocc={nocc}")
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    ctx = Context(0)
    from myqc_amd import guess as G
    have_net = G.load(ctx)
    # the ERI launch shapes, searched once on the headline cluster (untimed;
    # kept for the process, so every later build uses them)
    ctx.set_system(Z, xyz, basis, ftab)
    ctx.eri_build(1e-12)
    eri_shapes, eri_shape_ms = ctx.eri_tune(reps=2)
    log(f"[bench] ERI launch shapes (lanes, threads) per class: {eri_shapes}")
    res = {}
    all_forms = (("fp64", FP64, False), ("mixed", FP32, False), ("learned", FP64, True),
                 ("learned_mixed", FP32, True))
    forms = [f for f in all_forms if f[0] in args.forms.split(",") and (have_net or not f[2])]
    for name, prec, learned in forms:
        for w in range(args.warmup):
            r = solve(ctx, Z, xyz, basis, ftab, nocc, e_nuc, prec, learned)
            log(f"[bench] {name} warmup {w}: E={r['E']:.10f} iters={r['iterations']}")
        ctx.timer_reset(True)
        t0 = time.perf_counter()
        for s in range(args.steps):
            r = solve(ctx, Z, xyz, basis, ftab, nocc, e_nuc, prec, learned)
            log(f"[bench] {name} step {s}: E={r['E']:.10f} iters={r['iterations']}")
        dt = time.perf_counter() - t0
        eri = ctx.timer_get(T_ERI)
        jk = ctx.timer_get(T_JK)
        ctx.timer_reset(False)
        res[name] = dict(E=r["E"], iters=r["iterations"], sec=dt / args.steps, eri=eri, jk=jk,
                         n_stored=r["n_stored"])
    mp2 = mp2_leg(ctx, args.mp2_waters) if args.mp2_waters else None
    batch = batch_leg(ctx, args.batch) if args.batch else None
    ctx.close()
    conc = None
    if args.concurrent > 1 and have_net:
        conc = concurrent_leg(args.waters, args.concurrent, args.steps, args.warmup)
    lock = None
    if args.lockstep > 1:
        lock = lockstep_leg(args.waters, min(args.lockstep, 32), args.steps, args.warmup, args.forms.split(","))
    p = res.get("fp64") or next(iter(res.values()))
    # the fastest form whose energies agree with plain FP64 to 1e-8 Eh
    best_name, best = "fp64" if "fp64" in res else next(iter(res)), p
    for name, r in res.items():
        if abs(r["E"] - p["E"]) <= 1e-8 and r["sec"] < best["sec"]:
            best_name, best = name, r
    if conc and conc["max_dE_vs_fp64"] <= 1e-8 and conc["sec"] < best["sec"]:
        best_name, best = "learned_mixed_concurrent", conc
    for name, r in (lock["forms"].items() if lock else ()):
        if r["max_dE_vs_fp64"] <= 1e-8 and r["sec"] < best["sec"]:
            best_name, best = name + "_lockstep", r
    eri_n, eri_ms, eri_flops = p["eri"]
    jk_n, jk_ms, jk_bytes = p["jk"]
    eri_avg_ms = eri_ms / max(eri_n, 1)
    eri_tflops = (eri_flops / max(eri_n, 1)) / (eri_avg_ms * 1e-3) / 1e12
    jk_gbs = (jk_bytes / max(jk_n, 1)) / ((jk_ms / max(jk_n, 1)) * 1e-3) / 1e9
    gint = (p["n_stored"] / 1e9) / (eri_avg_ms * 1e-3)
    out = {
        "metric": "RHF solves/s at |dE|<=1e-8 Eh (40-water STO-3G); ERI Gint/s",
        "value": 1.0 / best["sec"],
        "unit": "RHF solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best.get("ms_per_batch", best["sec"] * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if "mixed" in best_name else "f64",
        "data": "synthetic",
        "config": {"workload": (conc["workload"] + f", n={basis.norb}, Schwarz 1e-12, DIIS to |dE|<1e-10"
                                if best_name == "learned_mixed_concurrent" else
                                lock["workload"] + f", n={basis.norb}, Schwarz 1e-12, DIIS to |dE|<1e-10"
                                if best_name.endswith("_lockstep") else
                                f"RHF/STO-3G {args.waters}-water cluster (seed {args.waters}), n={basis.norb}, "
                                f"Schwarz 1e-12, DIIS to |dE|<1e-10"),
                   "form": {"fp64": "plain FP64", "mixed": "mixed (FP32 dD digestion)",
                            "learned": "learned guess, FP64",
                            "learned_mixed": "learned guess + mixed (FP32 dD digestion)",
                            "learned_mixed_concurrent": "learned guess + mixed, clusters solved side by side",
                            "fp64_lockstep": "plain FP64, clusters batched in one lockstep SCF",
                            "learned_mixed_lockstep": "learned guess + mixed, clusters batched in one lockstep SCF",
                            }[best_name]},
        "plain_fp64": {"solves_per_s": 1.0 / p["sec"], "E": p["E"], "iterations": p["iters"]},
        "forms": {name: {"solves_per_s": 1.0 / r["sec"], "E": r["E"], "iterations": r["iters"],
                         "dE_vs_fp64": r["E"] - p["E"]} for name, r in res.items()},
        "eri": {"Gint_per_s": gint, "stored_integrals": p["n_stored"], "ms_per_build": eri_avg_ms,
                "TFLOPs": eri_tflops,
                "shapes": {["ss|ss", "ps|ss", "ps|ps", "pp|ss", "pp|ps", "pp|pp"][c]: list(v)
                           for c, v in eri_shapes.items()},
                "class_ms_tuned": {["ss|ss", "ps|ss", "ps|ps", "pp|ss", "pp|ps", "pp|pp"][c]: v
                                   for c, v in eri_shape_ms.items()}},
        "jk": {"GB_per_s": jk_gbs, "ms_per_pass": jk_ms / max(jk_n, 1), "passes": jk_n},
        "roofline": {"bound": "hbm", "achieved": jk_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": jk_gbs / HBM_PEAK_GBS, "traffic": measured_traffic(args.waters),
                     "kernel": "J/K digestion (k_digest)"},
        "roofline_eri": {"bound": "fp64-valu", "achieved": eri_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": eri_tflops / FP64_PEAK_TFLOPS},
    }
    if conc:
        out["concurrent"] = {k_: v for k_, v in conc.items() if k_ != "sec"} | {"solves_per_s": 1.0 / conc["sec"]}
    if lock:
        out["lockstep"] = {"workload": lock["workload"], "clusters": lock["clusters"],
                           "forms": {n_: {k_: v for k_, v in r.items() if k_ != "sec"} | {"solves_per_s": 1.0 / r["sec"]}
                                     for n_, r in lock["forms"].items()}}
        if "fp64" in lock["forms"]:
            out["plain_batched"] = {"solves_per_s": 1.0 / lock["forms"]["fp64"]["sec"],
                                    "vs_plain_fp64": p["sec"] / lock["forms"]["fp64"]["sec"]}
    out["vs_plain_fp64"] = p["sec"] / best["sec"]
    if mp2:
        out["mp2"] = mp2
    if batch:
        out["batch"] = batch
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle port, sampled; reference programs for calibration and the batch)")
        bm = None
        if args.batch:
            brng = np.random.default_rng(args.batch)
            bm = []
            for _ in range(args.batch):
                Zb, xb = M.random_molecule(brng)
                bm.append((Zb, M.to_bohr(Zb, xb)))
            bm = [bm[i] for i in np.random.default_rng(5).permutation(len(bm))]
        out["cpu_baseline"] = cpu_baseline(Z, xyz, basis, ftab, p["iters"], batch_mols=bm,
                                           cores=min(16, os.cpu_count() or 1))
    print(json.dumps(out))


if __name__ == "__main__":
    main()

/*
 * myqc_hip.h -- C ABI of libmyqc_hip, the MI355X (gfx950) two-electron path
 * of myQC.
 *
 * Conventions
 *  - Arrays are Fortran column-major and are passed unchanged from
 *    iso_c_binding callers (real(c_double) :: A(*)).
 *  - Basis arrays are the reference's flat layouts exactly as buildBasis
 *    produces them (src/myQC/basis.f90:23-226): bas(set*OpS+k), basinfo =
 *    [OpS, norb, {pri,ang,ori,nprim,center}...], setinfo = [nset, setl,
 *    {norb_in_set, maxL, center, orb0..orb3}...].
 *  - The caller owns host buffers; the context owns device buffers.
 *  - Every entry point returns 0 on success or a negative MYQC_E_* code; no
 *    exception or abort crosses the ABI.  myqc_last_error() gives the text.
 *    A Fortran caller then writes the `error` sentinel and STOPs, as the
 *    reference does (src/myQC/env.f90:36-41).
 *  - Calls are synchronous at return and must come from one host thread per
 *    context.
 */
#ifndef MYQC_HIP_H
#define MYQC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct myqc_ctx myqc_ctx;

/* error codes */
#define MYQC_OK 0
#define MYQC_E_ARG (-1)      /* invalid argument / size                     */
#define MYQC_E_STATE (-2)    /* call order (e.g. no system set)             */
#define MYQC_E_HIP (-3)      /* HIP runtime or library failure              */
#define MYQC_E_NOMEM (-4)    /* device allocation failed                    */
#define MYQC_E_NOCONV (-5)   /* SCF did not converge                        */
#define MYQC_E_DEVICE (-6)   /* no usable gfx950 device                     */

/* set_system flags */
#define MYQC_REF_COMPAT 0    /* reference numerics: FP32-rounded pi, Ftab Boys at Q=3*sum(set L) */

/* precision selectors */
#define MYQC_FP64 0
#define MYQC_FP32 1          /* J/K digestion of FP32-stored integrals        */
#define MYQC_BF16X3 2        /* ao2mo: split bf16 hi/lo MFMA products, FP32 accumulation */

/* Context: one device, one HIP stream.  No counterpart in the reference
 * (its programs own their arrays, src/myQC/env.f90). */
int myqc_init(int device, myqc_ctx **out);
int myqc_finalize(myqc_ctx *ctx);
const char *myqc_last_error(const myqc_ctx *ctx);

/* Molecule + basis + Boys table.  Replaces the inputs int2e reads
 * (getenv: src/myQC/env.f90:16-73; buildBasis: src/myQC/basis.f90:23;
 * Ftab: src/integrals/int2e.f90:161-163).  xyz_bohr is (nnuc,3) col-major as
 * getenv holds it; Ftab is Ft(0:120,0:22) col-major (Tk fastest). */
int myqc_set_system(myqc_ctx *ctx, int nnuc, const int *Z, const double *xyz_bohr, int nset,
                    const double *set_alpha, const int *setinfo, const double *bas, int norb,
                    const int *basinfo, const double *Ftab, int flags);

/* int2e drop-in: the dense, 8-fold filled (ij|gh) = XX(i,j,g,h), n^4
 * col-major (i fastest).  Replaces proc2e + fillsym
 * (src/integrals/int2e.f90:78-351,540-613).  schwarz_tol = 0 computes every
 * shell quartet (the reference has no Schwarz screen). */
int myqc_eri_full(myqc_ctx *ctx, double schwarz_tol, double *XX);

/* Device-resident, Schwarz-screened unique integrals for J/K digestion
 * (the work int2e hands to RHFI2G through the XX file).  Returns the number
 * of stored integral slots. */
int myqc_eri_build(myqc_ctx *ctx, double schwarz_tol, long long *n_stored);

/* Launch-shape search of myqc_eri_build's kernels: for each of the six
 * quartet classes (ss|ss, ps|ss, ps|ps, pp|ss, pp|ps, pp|pp) present in the
 * system of the last myqc_eri_build, every candidate of lanes per shell
 * quartet (1, 2, 4, 8) x threads per workgroup (64, 128, 256) is timed (best
 * of `reps`), and the fastest is kept for all later builds in the process.
 * shapes (may be NULL): 12 ints, {lanes, threads} per class; ms (may be
 * NULL): 6 doubles, the chosen shape's time.  No counterpart in the
 * reference (whose int2e loop, src/integrals/int2e.f90:192-284, has one
 * fixed order); for callers that build many systems of one kind. */
int myqc_eri_tune(myqc_ctx *ctx, int reps, int *shapes, double *ms);
/* Sets the shapes myqc_eri_tune would choose (12 ints as above; a class
 * with 0 lanes keeps its shape). */
int myqc_eri_set_shapes(myqc_ctx *ctx, const int *shapes);

/* RHFI2G drop-in: G(i,j) = sum_kl D(k,l)[(ij|kl) - 1/4 (ik|jl) - 1/4 (il|jk)]
 * (src/I2G/RHFI2G.f90:80-90) from the integrals of myqc_eri_build.
 * precision MYQC_FP64, or MYQC_FP32: the incremental form for a density
 * change D, over FP32 copies of the integrals, skipping integrals whose
 * Schwarz bound times max|D| is below 1e-12. */
int myqc_fock_G(myqc_ctx *ctx, const double *D, double *G, int precision);

/* UHFI2G (src/I2G/UHFI2G.f90:80-93): the two spin Fock contributions
 *   Ga(i,j) = sum_kl (Da + Db)(k,l) XX(i,j,k,l) - Da(k,l) XX(i,k,j,l), Gb likewise
 * from the integrals of myqc_eri_build (FP64; n x n column-major, D as the
 * spin densities C_occ C_occ^T of each spin, symmetrised). */
int myqc_fock_G_uhf(myqc_ctx *ctx, const double *Da, const double *Db, double *Ga, double *Gb);

/* int1p drop-in (src/integrals/int1p.f90:140-355): the dipole integrals
 * P_c(u,v) = <u| r_c |v> about the coordinate origin, c = x, y, z, as three
 * n x n column-major matrices one after another (the three records of Pxuv,
 * :230-233). */
int myqc_int1p(myqc_ctx *ctx, double *P);

/* cphf_dipole_rhf drop-in (src/props/cphf.f90:106-122; :80-131): the SCF dipole
 * moment mu_c = -sum_uv D(u,v) P_c(v,u) + sum_A Z_A R_A,c (atomic units) for
 * the total density D (n x n; the Da record of an RHF run). */
int myqc_dipole_rhf(myqc_ctx *ctx, const double *D, double *mu);

/* densRHF drop-in: D(u,v) = 2 sum_{i<nocc} C(u,i) C(v,i)
 * (src/dens/dens.f90:115-124). */
int myqc_density_rhf(myqc_ctx *ctx, const double *C, int nocc, double *D);

/* slow_ao2mo_MP2_RHF drop-in: (ia|jb) for all i,j < nocc, a,b < nvrt
 * (src/ao2mo/ao2mo.f90:465-602).  iajb is written in ijab_AB record order:
 * record (i,j) i-major, each record nvrt x nvrt with a fastest
 * (src/ao2mo/ao2mo.f90:575-582); iajb may be NULL to keep it on the device
 * for myqc_mp2_rhf.  precision MYQC_FP64 or MYQC_BF16X3. */
int myqc_ao2mo_ovov(myqc_ctx *ctx, const double *C, int nocc, double *iajb, int precision);

/* mp2_rhf drop-in: E(AA), E(AB) and E(MBPT2) = 2 E(AA) + E(AB)
 * (src/mp2/mp2.f90:79-144) over the (ia|jb) of the last myqc_ao2mo_ovov. */
int myqc_mp2_rhf(myqc_ctx *ctx, const double *eig, int nocc, double *e_aa, double *e_ab, double *e_tot);

/* slow_ao2mo_MP2_UHF drop-in (src/ao2mo/ao2mo.f90:614-904): the three spin
 * sets (ia|jb) with i,a from (Ca, nalpha) and j,b from (Cb, nbeta) as the set
 * requires, written as the record streams of the reference's files, each
 * record nvrt x nvrt' with a fastest:
 *   ijab_aa: records (i,j), i < j < nalpha, i-major (:718-724);
 *   ijab_bb: records (i,j), i < j < nbeta (:804-810);
 *   ijab_ab: records (i,j), i < nalpha, j < nbeta, i-major, nvrt_a x nvrt_b (:890-896).
 * Any output may be NULL; all three sets stay on the device for myqc_mp2_uhf.
 * A set with no occupied or no virtual orbital is empty.  precision
 * MYQC_FP64 or MYQC_BF16X3. */
int myqc_ao2mo_ovov_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int nalpha, int nbeta, double *ijab_aa,
                        double *ijab_bb, double *ijab_ab, int precision);

/* mp2_uhf drop-in (src/mp2/mp2.f90:154-232) over the sets of the last
 * myqc_ao2mo_ovov_uhf:
 *   E(AA) = sum_{i<j, a<b} ((ia|jb) - (ib|ja))^2 / (ea_i + ea_j - ea_a - ea_b), E(BB) likewise,
 *   E(AB) = sum_{ijab} (ia|jb)^2 / (ea_i + eb_j - ea_a - eb_b),  E(MBPT2) = E(AA) + E(BB) + E(AB).
 * eig_a, eig_b: the n orbital energies of each spin (the two lines of eig). */
int myqc_mp2_uhf(myqc_ctx *ctx, const double *eig_a, const double *eig_b, int nalpha, int nbeta, double *e_aa,
                 double *e_bb, double *e_ab, double *e_tot);

/* mp2_uhf drop-in for a separate mp2 program: the same sums over the record
 * streams of ijab_AA, ijab_BB and ijab_AB (layouts as myqc_ao2mo_ovov_uhf
 * writes them; src/mp2/mp2.f90:176-222), uploaded to the device.
 * nalpha + nvrt_a must equal nbeta + nvrt_b (= n). */
int myqc_mp2_uhf_records(myqc_ctx *ctx, const double *ijab_aa, const double *ijab_bb, const double *ijab_ab,
                         const double *eig_a, const double *eig_b, int nalpha, int nbeta, int nvrt_a, int nvrt_b,
                         double *e_aa, double *e_bb, double *e_ab, double *e_tot);

/* slow_ao2mo_CIS_UHF drop-in (src/ao2mo/ao2mo.f90:919-1227): the five record
 * streams of the reference's CIS files, each record (j,b) j-major with b
 * fastest, each record over (i,a) with a fastest (upS = nocc_S nvrt_S):
 *   ajib_aa (upA x upA): (ai|jb), alpha;     ajbi_aa (upA x upA): (ab|ji), alpha;
 *   ajib_ab (upB records of upA): (a_A i_A|j_B b_B);
 *   ajib_bb, ajbi_bb (upB x upB): beta.
 * Any output may be NULL; all sets stay on the device for myqc_cis_uhf.  FP64. */
int myqc_ao2mo_cis_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int nalpha, int nbeta, double *ajib_aa,
                       double *ajbi_aa, double *ajib_ab, double *ajib_bb, double *ajbi_bb);

/* cis_scf_uhf drop-in (src/CIS/cis.f90:87-223): the UHF-CIS matrix
 *   A(ia, jb) = (ai|jb) - (ab|ji) + delta_ij delta_ab (e_a - e_i) within a spin,
 *   A(ia_A, jb_B) = (ai|jb)_AB,
 * of dimension upA + upB, diagonalised whole; the lowest nroots eigenvalues
 * (excitation energies, ascending) into energies and their eigenvectors
 * (column-major, (upA + upB) x nroots; alpha (i,a) with a fastest, then beta)
 * into vectors (either may be NULL; nroots is clamped to upA + upB, as
 * options(15) is at cis.f90:212-213).  The five record streams are those of
 * myqc_ao2mo_cis_uhf; all NULL takes the sets it left on the device.
 * eig_a, eig_b: the n orbital energies of each spin. */
int myqc_cis_uhf(myqc_ctx *ctx, const double *ajib_aa, const double *ajbi_aa, const double *ajib_ab,
                 const double *ajib_bb, const double *ajbi_bb, const double *eig_a, const double *eig_b, int nalpha,
                 int nbeta, int nvrt_a, int nvrt_b, int nroots, double *energies, double *vectors);

/* mp2_rhf drop-in for a separate mp2 program: the same sums over (ia|jb)
 * records read from an ijab_AB file (record (i,j) i-major, nvrt x nvrt with a
 * fastest; src/mp2/mp2.f90:99-135), uploaded to the device. */
int myqc_mp2_rhf_records(myqc_ctx *ctx, const double *iajb_records, const double *eig, int nocc, int nvrt,
                         double *e_aa, double *e_ab, double *e_tot);

/* myqc_mp2_rhf_records for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same sums over the
 * ijab_AB record stream, with the records (nocc^2 records of nvrt x nvrt, a
 * fastest), the n = nocc + nvrt orbital energies and the three results
 * {E(AA), E(AB), E(MBPT2)} all in caller-owned DEVICE memory (d_out: 3
 * doubles).  The partial sums live in a fixed device scratch of this GPU that
 * holds at most 65536 records: nocc > 256 returns MYQC_E_ARG.  Nothing is
 * allocated; complete at return.  Calls that overlap on one GPU share the
 * scratch and race. */
int myqc_mp2_rhf_records_noalloc(myqc_ctx *ctx, const double *d_records, const double *d_eig, int nocc, int nvrt,
                                 double *d_out);

/* Core Hamiltonian and overlap, int1e drop-in (src/integrals/int1e.f90:132-280). */
int myqc_int1e(myqc_ctx *ctx, double *S, double *H);

/* In-process RHF: the scf.f90 RHF loop (src/scf/scf.f90:61-216,720-904)
 * with the dens/RHFI2G spawns replaced by device-resident D, G and the
 * integrals of myqc_eri_build, plus DIIS.  C_guess may be NULL (core-H guess,
 * scf.f90:410-512).  Converged when |dE| < e_tol and max|dD| < d_tol.
 * precision MYQC_FP64, or MYQC_FP32 for incremental dD Fock builds with FP32
 * digestion and a final FP64 rebuild.  Outputs: total energy (with E_nuc),
 * MO coefficients C (n x n), orbital energies eig (n), iteration count. */
int myqc_rhf_scf(myqc_ctx *ctx, const double *H, const double *S, int nocc, double e_nuc, const double *C_guess,
                 double e_tol, double d_tol, int max_iter, int precision, double *energy, double *C, double *eig,
                 int *iterations);

/* myqc_rhf_scf started from an AO density D_guess (n x n, the 2 C_occ C_occ^T
 * convention of densRHF) instead of orbitals: the entry for the learned
 * guess.  No counterpart in the reference (restarts read Cui,
 * src/scf/scf.f90:148-156).  Runs from a supplied guess (this entry, or
 * C_guess above) level-shift the virtual levels by 0.5 Eh while the DIIS
 * error is above 1; the converged state is unaffected. */
int myqc_rhf_scf_density(myqc_ctx *ctx, const double *H, const double *S, int nocc, double e_nuc,
                         const double *D_guess, double e_tol, double d_tol, int max_iter, int precision,
                         double *energy, double *C, double *eig, int *iterations);

/* myqc_rhf_scf_density (D_guess[k] given) or myqc_rhf_scf from the core
 * guess (D_guess NULL or D_guess[k] NULL) for nsys <= 32 systems of one basis
 * size at once, each on its own context (myqc_set_system and myqc_eri_build
 * done), solved in lockstep on the first context's stream: the Fock builds one
 * system after another, every other step of the loop one launch for all
 * systems, one host synchronisation per iteration for all.  No counterpart in
 * the reference, which solves one molecule per run; for callers that solve
 * many molecules of one size (conformers, clusters).  Each system converges
 * on its own criteria, as in myqc_rhf_scf; outputs per system k: energy[k],
 * C[k] (n x n), eig[k] (n), iterations[k] (C, eig may be NULL, or hold NULL
 * entries).  MYQC_E_NOCONV when any system did not converge (the others'
 * outputs are valid).  The contexts must not be used by other threads during
 * the call. */
int myqc_rhf_scf_many(myqc_ctx *const *ctxs, int nsys, const double *const *H, const double *const *S,
                      const int *nocc, const double *e_nuc, const double *const *D_guess, double e_tol, double d_tol,
                      int max_iter, int precision, double *energy, double *const *C, double *const *eig,
                      int *iterations);

/* densUHF drop-in: Da(u,v) = sum_{i<nalpha} Ca(u,i) Ca(v,i), Db likewise over
 * nbeta (src/dens/dens.f90:213-228). */
int myqc_density_uhf(myqc_ctx *ctx, const double *Ca, const double *Cb, int nalpha, int nbeta, double *Da,
                     double *Db);

/* In-process UHF: the scf.f90 UHF loop (src/scf/scf.f90:221-404 UHF,
 * :909-1117 UHFiter) with the dens/UHFI2G spawns replaced by device-resident
 * Da, Db, Ga, Gb and the integrals of myqc_eri_build, plus DIIS over both
 * spins.  Ca_guess and Cb_guess are both NULL (both spins from the
 * core-Hamiltonian orbitals, initUHF :518-622) or both given (a Cui restart
 * file's two matrices, :314-320).  E = (Da.(Fa+H) + Db.(Fb+H))/2 + e_nuc
 * (:969-975).  Converged when |dE| < e_tol and max|dD| (both spins) < d_tol.
 * FP64.  Outputs: energy, Ca, Cb (n x n), eig_a, eig_b (n), iterations; each
 * may be NULL. */
int myqc_uhf_scf(myqc_ctx *ctx, const double *H, const double *S, int nalpha, int nbeta, double e_nuc,
                 const double *Ca_guess, const double *Cb_guess, double e_tol, double d_tol, int max_iter,
                 double *energy, double *Ca, double *Cb, double *eig_a, double *eig_b, int *iterations);

/* Diagnostics of the last myqc_rhf_scf: densities obtained by purification
 * and the iterations that fell back to diagonalisation.  No counterpart in
 * the reference (it diagonalises every iteration, src/scf/scf.f90:848). */
int myqc_scf_info(myqc_ctx *ctx, int *purifications, int *diag_fallbacks);

/* Batched RHF of many small molecules (1..40 functions each) in one
 * persistent launch.  No counterpart in the reference, which solves one
 * molecule per run of its programs; for callers that screen or train on
 * thousands of molecules.  myqc_batch_set takes, per molecule, exactly the
 * arrays myqc_set_system takes, concatenated in molecule order:
 *   mol_nnuc[m], mol_norb[m], mol_nocc[m];
 *   Z (sum nnuc) and xyz_bohr (each molecule an nnuc x 3 col-major block);
 *   set_alpha, setinfo, bas, basinfo (buildBasis layouts) with
 *   offsets[4 m + {0,1,2,3}] = the molecule's first element in each.
 * myqc_batch_rhf runs integrals + SCF for the whole batch (DIIS, |dE| <
 * e_tol and max|dD| < d_tol as myqc_rhf_scf) from the core-Hamiltonian guess
 * or, with D_guess != NULL, from the given densities (n_m x n_m blocks,
 * concatenated).  The first call tunes the launch shapes and captures the
 * pipeline in a hipGraph; later calls replay it.  converged[m] is 1, 0 (not
 * within max_iter) or -1 (singular overlap). */
int myqc_batch_set(myqc_ctx *ctx, int nmol, const int *mol_nnuc, const int *mol_norb, const int *mol_nocc,
                   const int *Z, const double *xyz_bohr, const double *set_alpha, const int *setinfo,
                   const double *bas, const int *basinfo, const int *offsets, const double *Ftab);
int myqc_batch_rhf(myqc_ctx *ctx, const double *D_guess, double e_tol, double d_tol, int max_iter, double *energy,
                   int *iterations, int *converged);
/* The converged densities of the last myqc_batch_rhf (n_m x n_m blocks). */
int myqc_batch_density(myqc_ctx *ctx, double *D);
/* Chosen launch shapes (threads per ERI class kernel, per SCF workgroup),
 * graph replays so far and stored integrals of the batch. */
int myqc_batch_info(myqc_ctx *ctx, int *eri_threads, int *scf_threads, int *graph_replays, long long *n_integrals);

/* Learned initial guess.  No counterpart in the reference (its SCF starts
 * from the core Hamiltonian, src/scf/scf.f90:410-512, which stays the
 * fallback); for callers that solve many similar molecules.  An MLP with
 * three SiLU hidden layers of `width` maps n_in atom-pair features to n_out
 * bond-frame density invariants (the batch guess uses n_in = 200,
 * n_out = 16).  Arrays are row-major FP32 (rows x n_in, rows x n_out).
 *   myqc_mlp_create    Xavier-uniform weights from `seed`, zero biases
 *   myqc_mlp_params    copy the flat parameter vector out (get) and/or in
 *                      (set); count = its length
 *   myqc_mlp_forward   inference
 *   myqc_mlp_gradient  masked mean-squared error sum(M (Y - T)^2) / sum(M)
 *                      and its gradient over the parameters
 *   myqc_mlp_train     Adam over `epochs` shuffled passes of minibatches,
 *                      the data resident on the device; loss_per_epoch
 *                      (may be NULL) gets each epoch's first-batch loss
 *                      sum(M (Y - T)^2) / (batch n_out) */
int myqc_mlp_create(myqc_ctx *ctx, int n_in, int width, int n_out, unsigned long long seed);
int myqc_mlp_params(myqc_ctx *ctx, float *get, const float *set, long long *count);
int myqc_mlp_forward(myqc_ctx *ctx, const float *X, int rows, float *Y);
int myqc_mlp_gradient(myqc_ctx *ctx, const float *X, const float *T, const float *M, int rows, float *loss,
                      float *grad);
int myqc_mlp_train(myqc_ctx *ctx, const float *X, const float *T, const float *M, int rows, int epochs, int batch,
                   float lr, unsigned long long seed, float *loss_per_epoch);
/* Atom pairs (A <= B) of the batch molecules: their features, and the
 * targets and masks taken from the densities of the last myqc_batch_rhf
 * (masks zero for molecules that did not converge).  With all three
 * pointers NULL only *npair is returned. */
int myqc_batch_pair_data(myqc_ctx *ctx, int *npair, float *features, float *targets, float *mask);
/* The network's AO density (n x n) for one molecule of H..Ne atoms, as the
 * D_guess of myqc_rhf_scf_density; xyz_bohr is (nnuc, 3) column-major.  n is
 * the molecule's STO-3G size, 1 function per H/He and 5 per Li..Ne atom (the
 * AO order of buildBasis).  When a system is set on the context the molecule
 * must be that system (same nnuc and Z, so n is the context's n); otherwise
 * MYQC_E_ARG. */
int myqc_guess_density(myqc_ctx *ctx, int nnuc, const int *Z, const double *xyz_bohr, double *D);
/* myqc_batch_rhf started from the network's densities, computed on the
 * device inside the same captured graph; molecules that do not converge from
 * it within min(max_iter, 48) iterations are solved again from the
 * core-Hamiltonian guess (fallback[m] = 1, iterations[m] counts both runs). */
int myqc_batch_rhf_learned(myqc_ctx *ctx, double e_tol, double d_tol, int max_iter, double *energy, int *iterations,
                           int *converged, int *fallback);

/* Kernel timing accumulated on the context's stream with HIP events since
 * the last reset: which = 0 ERI (the shell-quartet class kernels of a build),
 * 1 J/K digestion (the digestion kernel alone; work = bytes s*N_stored + 16 n^2,
 * s = 8 FP64 or 4 FP32), 2 ao2mo (work = flops of the four quarter steps),
 * 3 MP2.  Returns launches, total milliseconds and algorithmic work.
 * No counterpart in the reference (it prints CPU_TIME per program). */
int myqc_timer_reset(myqc_ctx *ctx, int enable);
int myqc_timer_get(myqc_ctx *ctx, int which, long long *launches, double *ms, double *work);

#ifdef __cplusplus
}
#endif
#endif

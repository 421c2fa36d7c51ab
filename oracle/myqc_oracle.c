/*
 * myqc_oracle.c -- CPU restatement of the reference's two-electron path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle: only tests/, the golden
 * fixture generator and bench.py's CPU-baseline leg may load it, and only as
 * the checker.  The product (libmyqc_hip) never links or calls it.
 *
 * It follows the reference algorithm, not the GPU one: McMurchie-Davidson
 * Hermite expansions per primitive *set* quartet, with every numerical quirk
 * that decides agreement at 1e-10:
 *   - Pi rounded to single precision (every `Pi = 3.1415926535897931` without
 *     a D0 suffix: src/integrals/int2e.f90:621, auxilary.f90:169,196,711,
 *     int1e.f90:324,394,500);
 *   - Boys F_m from the tabulated Ftab with a 7-term Taylor expansion at
 *     Q = 3*(sum of set max-L), downward recursion (auxilary.f90:85-215);
 *   - the primitive-pair screen EIJ*EGH < 1e-14 (int2e.f90:257) and the
 *     |E| < 1e-16 Hermite-term drop (auxilary.f90:573,611-615);
 *   - contraction coefficients used as given (auxilary.f90:620-622).
 *
 * Index layouts are the reference's: bas(set*OpS+k), basinfo, setinfo,
 * xyz(nnuc,3) column-major in bohr, XX(i,j,g,h) column-major with i fastest.
 */
#define _POSIX_C_SOURCE 199309L   /* clock_gettime under -std=c99 */
#include <math.h>
#include <time.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define FT(tk, m) Ft[(tk) + 121 * (m)]

/* `REAL(KIND=8),PARAMETER :: Pi = 3.1415926535897931` is a default-real literal */
static double pi_sp(void) { return (double)3.1415926535897931f; }

static double factR8(int n) {          /* auxilary.f90:290-312 */
    double v = 1.0;
    for (int i = n; i >= 2; --i) v *= i;
    return v;
}

static double ipow(double x, int k) {  /* integer power by repeated products */
    double r = 1.0;
    for (int i = 0; i < k; ++i) r *= x;
    return r;
}

/* auxilary.f90:265-285.  For 30 <= T the reference returns an undefined value;
 * the contribution it feeds is e^{-T} g / T <= 3.1e-15 |g|.  0.49 is used. */
static double boysG(double T) {
    if (T >= 12.0 && T < 15.0)
        return 0.4999489092 - 0.2473631686 * pow(T, -1.0) + 0.321180909 * pow(T, -2.0)
               - 0.3811559346 * pow(T, -3.0);
    if (T >= 15.0 && T < 18.0) return 0.4998436875 - 0.24249438 * pow(T, -1.0) + 0.24642845 * pow(T, -2.0);
    if (T >= 18.0 && T < 24.0) return 0.499093162 - 0.2152832 * pow(T, -1.0);
    return 0.490;
}

/* auxilary.f90:85-215: F_0..F_Q(T).  Returns -1 if Q > 16 (the reference STOPs). */
int oracle_boys(int Q, double T, const double *Ft, double *F) {
    if (Q > 16) return -1;
    for (int j = 0; j <= Q; ++j) F[j] = 0.0;
    if (T >= 0.0 && T < 12.0) {
        int Tk = (int)lround(T * 10.0);            /* NINT: half away from zero */
        double x = Tk / 10.0 - T;
        for (int k = 0; k <= 6; ++k) F[Q] = F[Q] + FT(Tk, Q + k) * ipow(x, k) / factR8(k);
        double e = exp(-T);
        for (int j = Q - 1; j >= 0; --j) F[j] = (2 * T * F[j + 1] + e) / (2.0 * j + 1.0);
    } else if (T >= 12.0 && T < 2 * Q + 36) {
        double g = boysG(T);
        double pi = pi_sp();
        F[0] = 0.5 * pow(pi, 0.5) * pow(T, -0.5) - exp(-T) * g / T;
        for (int j = 1; j <= Q; ++j) F[j] = pow(2.0 * T, -1.0) * ((2.0 * (j - 1) + 1) * F[j - 1] - exp(-T));
    } else if (T >= 2 * Q + 36) {
        double pi = pi_sp();
        F[0] = 0.5 * pow(pi, 0.5) * pow(T, -0.5);
        for (int j = 1; j <= Q; ++j) F[j] = pow(2.0 * T, -1.0) * (2.0 * (j - 1) + 1.0) * F[j - 1];
    } else {
        return -2;
    }
    return 0;
}

/* auxilary.f90:704-726 */
static double gtoD(int l, double a) {
    double pi = pi_sp();
    if (l == 0) return pow(2.0 * a / pi, 3.0 / 4.0);
    return pow(128.0 * pow(a, 5.0) / pow(pi, 3.0), 1.0 / 4.0);
}

/* 1-D Hermite coefficients E^{ja,kb}_t (auxilary.f90:406-536): the reference
 * reduces the b index when ja <= kb (with PB) and the a index otherwise (PA).
 * E[t][ja][kb], ja,kb <= LMAXE, t <= ja+kb. */
#define LMAXE 3
#define TMAX (2 * LMAXE + 1)
typedef struct { double v[TMAX + 2][LMAXE + 1][LMAXE + 1]; } Ecoef;

static double Eget(const Ecoef *E, int t, int j, int k) {
    if (t < 0 || j < 0 || k < 0 || t > j + k) return 0.0;
    return E->v[t][j][k];
}

static void hermite_E(Ecoef *E, double PA, double PB, double pp, int amax, int bmax) {
    memset(E, 0, sizeof(*E));
    for (int s = 0; s <= amax + bmax; ++s)
        for (int j = 0; j <= amax; ++j) {
            int k = s - j;
            if (k < 0 || k > bmax) continue;
            for (int t = 0; t <= j + k; ++t) {
                double v;
                if (j == 0 && k == 0) v = (t == 0) ? 1.0 : 0.0;
                else if (j <= k)
                    v = Eget(E, t - 1, j, k - 1) / (2.0 * pp) + PB * Eget(E, t, j, k - 1) + (t + 1) * Eget(E, t + 1, j, k - 1);
                else
                    v = Eget(E, t - 1, j - 1, k) / (2.0 * pp) + PA * Eget(E, t, j - 1, k) + (t + 1) * Eget(E, t + 1, j - 1, k);
                E->v[t][j][k] = v;
            }
        }
}

typedef struct {
    int nset, setl, OpS, norb;
    const int *setinfo;   /* reference layout, with the 2-entry header */
    const int *basinfo;
    const double *bas, *alpha, *xyz;
    int nnuc;
} Sys;

static int set_norb(const Sys *s, int a) { return s->setinfo[2 + a * s->setl + 0]; }
static int set_maxL(const Sys *s, int a) { return s->setinfo[2 + a * s->setl + 1]; }
static int set_center(const Sys *s, int a) { return s->setinfo[2 + a * s->setl + 2]; }
static int set_orb(const Sys *s, int a, int i) { return s->setinfo[2 + a * s->setl + 3 + i]; }
static int orb_ang(const Sys *s, int o) { return s->basinfo[2 + 5 * o + 1]; }
static int orb_ori(const Sys *s, int o) { return s->basinfo[2 + 5 * o + 2]; }
static double X(const Sys *s, int atom, int c) { return s->xyz[atom + s->nnuc * c]; }

static void orb_lvec(const Sys *s, int o, int l[3]) {
    int ang = orb_ang(s, o), ori = orb_ori(s, o);
    if (ori == -1) { l[0] = l[1] = l[2] = ang; }
    else { l[0] = l[1] = l[2] = 0; l[ori] = ang; }
}

/* A set pair's nonzero Hermite terms (getDk, auxilary.f90:541-633) */
typedef struct {
    int n;
    double p, EIJ, P[3];
    double D[64 * 4];
    int C[64 * 4][3];
    int O[64 * 4][2];
} Pair;

static void make_pair(const Sys *s, int a, int b, Pair *pr, int extra_b) {
    double aa = s->alpha[a], bb = s->alpha[b];
    int u = set_center(s, a), v = set_center(s, b);
    double p = aa + bb, mm = aa * bb, AB[3], PA[3], PB[3];
    for (int i = 0; i < 3; ++i) {
        AB[i] = X(s, u, i) - X(s, v, i);
        pr->P[i] = (aa * X(s, u, i) + bb * X(s, v, i)) / p;
        PA[i] = pr->P[i] - X(s, u, i);
        PB[i] = pr->P[i] - X(s, v, i);
    }
    pr->p = p;
    pr->EIJ = exp(-mm * (pow(AB[0], 2.0) + pow(AB[1], 2.0) + pow(AB[2], 2.0)) / p);
    Ecoef E[3];
    int la = set_maxL(s, a), lb = set_maxL(s, b);
    for (int c = 0; c < 3; ++c) hermite_E(&E[c], PA[c], PB[c], p, la, lb + extra_b);
    const double tol = 0.1e-15;
    pr->n = 0;
    for (int i = 0; i < set_norb(s, a); ++i) {
        int oa = set_orb(s, a, i), ll[3];
        orb_lvec(s, oa, ll);
        for (int j = 0; j < set_norb(s, b); ++j) {
            int ob = set_orb(s, b, j), rr[3];
            orb_lvec(s, ob, rr);
            int Nm = ll[0] + rr[0], Lm = ll[1] + rr[1], Mm = ll[2] + rr[2];
            for (int M = 0; M <= Mm; ++M) {
                if (fabs(Eget(&E[2], M, ll[2], rr[2])) < tol) continue;
                for (int L = 0; L <= Lm; ++L) {
                    if (fabs(Eget(&E[1], L, ll[1], rr[1])) < tol) continue;
                    for (int N = 0; N <= Nm; ++N) {
                        if (fabs(Eget(&E[0], N, ll[0], rr[0])) < tol) continue;
                        double d = Eget(&E[0], N, ll[0], rr[0]) * Eget(&E[1], L, ll[1], rr[1]) * Eget(&E[2], M, ll[2], rr[2]);
                        d = d * pr->EIJ * gtoD(orb_ang(s, oa), aa) * gtoD(orb_ang(s, ob), bb);
                        d = d * s->bas[a * s->OpS + i] * s->bas[b * s->OpS + j];
                        int k = pr->n++;
                        pr->D[k] = d;
                        pr->C[k][0] = N; pr->C[k][1] = L; pr->C[k][2] = M;
                        pr->O[k][0] = oa; pr->O[k][1] = ob;
                    }
                }
            }
        }
    }
}

/* Hermite Coulomb integrals R^{(j)}_{NLM} (RNLMj, auxilary.f90:22-80):
 * N is reduced first, then L, then M. */
#define RMAX 7
typedef struct {
    double v[RMAX][RMAX][RMAX][2 * RMAX];
    unsigned char seen[RMAX][RMAX][RMAX][2 * RMAX];
} Rtab;

static double Rget(Rtab *R, double a, double b, double c, int N, int L, int M, int j, double al, const double *Fj) {
    if (N < 0 || L < 0 || M < 0) return 0.0;
    if (R->seen[N][L][M][j]) return R->v[N][L][M][j];
    double v;
    if (N == 0 && L == 0 && M == 0) v = ipow(-2.0 * al, j) * Fj[j];
    else if (N != 0)
        v = a * Rget(R, a, b, c, N - 1, L, M, j + 1, al, Fj) + (N - 1) * Rget(R, a, b, c, N - 2, L, M, j + 1, al, Fj);
    else if (L != 0)
        v = b * Rget(R, a, b, c, 0, L - 1, M, j + 1, al, Fj) + (L - 1) * Rget(R, a, b, c, 0, L - 2, M, j + 1, al, Fj);
    else
        v = c * Rget(R, a, b, c, 0, 0, M - 1, j + 1, al, Fj) + (M - 1) * Rget(R, a, b, c, 0, 0, M - 2, j + 1, al, Fj);
    R->seen[N][L][M][j] = 1;
    R->v[N][L][M][j] = v;
    return v;
}

static void init_sys(Sys *s, int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                     const double *bas, int norb, const int *basinfo) {
    s->nnuc = nnuc; s->xyz = xyz; s->nset = nset; s->alpha = alpha; s->setinfo = setinfo;
    s->setl = setinfo[1]; s->bas = bas; s->norb = norb; s->basinfo = basinfo; s->OpS = basinfo[0];
}

/* One primitive set quartet into canonical XX elements (clmnew, int2e.f90:618-726) */
static void clmnew(const Pair *ab, const Pair *cd, int Q, const double *Ft, double *XX, int n) {
    double pi = pi_sp();
    double p = ab->p, q = cd->p;
    double PQ[3] = {ab->P[0] - cd->P[0], ab->P[1] - cd->P[1], ab->P[2] - cd->P[2]};
    double ll = 2 * pow(pi, 2.5) / (p * q * sqrt(p + q));
    double TT = p * q * (pow(PQ[0], 2.0) + pow(PQ[1], 2.0) + pow(PQ[2], 2.0)) / (p + q);
    double Fj[17];
    oracle_boys(Q, TT, Ft, Fj);
    static Rtab R;   /* oracle is single-threaded */
    memset(R.seen, 0, sizeof(R.seen));
    double al = p * q / (p + q);
    size_t n2 = (size_t)n * n, n3 = n2 * n;
    for (int k = 0; k < ab->n; ++k) {
        int i = ab->O[k][0], j = ab->O[k][1];
        if (j < i) continue;
        for (int kp = 0; kp < cd->n; ++kp) {
            int g = cd->O[kp][0], h = cd->O[kp][1];
            if (h < g) continue;
            if (g * n + h < i * n + j) continue;
            int Np = cd->C[kp][0], Lp = cd->C[kp][1], Mp = cd->C[kp][2];
            double r = Rget(&R, PQ[0], PQ[1], PQ[2], ab->C[k][0] + Np, ab->C[k][1] + Lp, ab->C[k][2] + Mp, 0, al, Fj);
            double sgn = ((Np + Lp + Mp) & 1) ? -1.0 : 1.0;
            XX[i + n * j + n2 * g + n3 * h] += sgn * ll * cd->D[kp] * r * ab->D[k];
        }
    }
}

/* proc2e (int2e.f90:78-351) with fillsym (int2e.f90:540-613):
 * the dense 8-fold-filled XX(n,n,n,n). */
int oracle_eri_full(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                    const double *bas, int norb, const int *basinfo, const double *Ft, double *XX) {
    Sys s;
    init_sys(&s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo);
    int n = norb;
    size_t n2 = (size_t)n * n, n3 = n2 * n, n4 = n3 * n;
    memset(XX, 0, n4 * sizeof(double));
    Pair *pairs = (Pair *)malloc(sizeof(Pair) * (size_t)nset * nset);
    if (!pairs) return -1;
    for (int a = 0; a < nset; ++a)
        for (int b = 0; b < nset; ++b) make_pair(&s, a, b, &pairs[a * nset + b], 0);
    for (int a = 0; a < nset; ++a)
        for (int b = 0; b < nset; ++b) {
            const Pair *ab = &pairs[a * nset + b];
            for (int c = 0; c < nset; ++c)
                for (int d = 0; d < nset; ++d) {
                    const Pair *cd = &pairs[c * nset + d];
                    if (cd->EIJ * ab->EIJ < 1.0e-14) continue;
                    int Q = 3 * (set_maxL(&s, a) + set_maxL(&s, b) + set_maxL(&s, c) + set_maxL(&s, d));
                    clmnew(ab, cd, Q, Ft, XX, n);
                }
        }
    free(pairs);
#define XI(i, j, g, h) XX[(size_t)(i) + n * (size_t)(j) + n2 * (size_t)(g) + n3 * (size_t)(h)]
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j)
            for (int g = i; g < n; ++g)
                for (int h = (g == i ? j : g); h < n; ++h) {
                    double v = XI(i, j, g, h);
                    XI(i, j, h, g) = v; XI(j, i, g, h) = v; XI(j, i, h, g) = v;
                    XI(g, h, i, j) = v; XI(g, h, j, i) = v; XI(h, g, i, j) = v; XI(h, g, j, i) = v;
                }
#undef XI
    return 0;
}

/* Timing sample of proc2e's work: the bra set pairs listed in `abs` (pairs of
 * set indices) against every ket set pair, accumulated into a scratch that is
 * indexed like XX's canonical elements for those bra orbitals.  Returns the
 * number of primitive set quartets evaluated. */
/* the same with times[t] = seconds spent on bra set pair t (times may be NULL) */
long long oracle_eri_sample_timed(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                                  const double *bas, int norb, const int *basinfo, const double *Ft,
                                  int nab, const int *abs_, double *checksum, double *times);

long long oracle_eri_sample(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                            const double *bas, int norb, const int *basinfo, const double *Ft,
                            int nab, const int *abs_, double *checksum) {
    return oracle_eri_sample_timed(nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ft, nab, abs_, checksum,
                                   NULL);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

long long oracle_eri_sample_timed(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                                  const double *bas, int norb, const int *basinfo, const double *Ft,
                                  int nab, const int *abs_, double *checksum, double *times) {
    Sys s;
    init_sys(&s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo);
    int n = norb;
    Pair *pairs = (Pair *)malloc(sizeof(Pair) * (size_t)nset * nset);
    if (!pairs) return -1;
    for (int a = 0; a < nset; ++a)
        for (int b = 0; b < nset; ++b) make_pair(&s, a, b, &pairs[a * nset + b], 0);
    /* scratch: every bra orbital pair (i,j) of the sampled sets maps to slot
     * (i mod 4, j mod 4) x (g,h) -- same arithmetic, bounded memory */
    size_t n2 = (size_t)n * n;
    double *acc = (double *)calloc(16 * n2, sizeof(double));
    long long cnt = 0;
    double pi = pi_sp();
    static Rtab R;
    for (int t = 0; t < nab; ++t) {
        const double t0 = times ? now_s() : 0.0;
        int a = abs_[2 * t], b = abs_[2 * t + 1];
        const Pair *ab = &pairs[a * nset + b];
        for (int c = 0; c < nset; ++c)
            for (int d = 0; d < nset; ++d) {
                const Pair *cd = &pairs[c * nset + d];
                if (cd->EIJ * ab->EIJ < 1.0e-14) continue;
                int Q = 3 * (set_maxL(&s, a) + set_maxL(&s, b) + set_maxL(&s, c) + set_maxL(&s, d));
                double p = ab->p, q = cd->p;
                double PQ[3] = {ab->P[0] - cd->P[0], ab->P[1] - cd->P[1], ab->P[2] - cd->P[2]};
                double ll = 2 * pow(pi, 2.5) / (p * q * sqrt(p + q));
                double TT = p * q * (pow(PQ[0], 2.0) + pow(PQ[1], 2.0) + pow(PQ[2], 2.0)) / (p + q);
                double Fj[17];
                oracle_boys(Q, TT, Ft, Fj);
                memset(R.seen, 0, sizeof(R.seen));
                double al = p * q / (p + q);
                for (int k = 0; k < ab->n; ++k) {
                    int i = ab->O[k][0], j = ab->O[k][1];
                    if (j < i) continue;
                    for (int kp = 0; kp < cd->n; ++kp) {
                        int g = cd->O[kp][0], h = cd->O[kp][1];
                        if (h < g) continue;
                        if (g * n + h < i * n + j) continue;
                        int Np = cd->C[kp][0], Lp = cd->C[kp][1], Mp = cd->C[kp][2];
                        double r = Rget(&R, PQ[0], PQ[1], PQ[2], ab->C[k][0] + Np, ab->C[k][1] + Lp,
                                        ab->C[k][2] + Mp, 0, al, Fj);
                        double sgn = ((Np + Lp + Mp) & 1) ? -1.0 : 1.0;
                        acc[((i & 3) * 4 + (j & 3)) * n2 + (size_t)g * n + h] += sgn * ll * cd->D[kp] * r * ab->D[k];
                    }
                }
                ++cnt;
            }
        if (times) times[t] = now_s() - t0;
    }
    double cs = 0.0;
    for (size_t i = 0; i < 16 * n2; ++i) cs += acc[i];
    *checksum = cs;
    free(acc);
    free(pairs);
    return cnt;
}

/* int1e proc1e (int1e.f90:132-280): overlap S and core Hamiltonian H = T + V */
int oracle_int1e(int nnuc, const int *Z, const double *xyz, int nset, const double *alpha, const int *setinfo,
                 const double *bas, int norb, const int *basinfo, const double *Ft, double *S, double *H) {
    Sys s;
    init_sys(&s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo);
    int n = norb;
    double pi = pi_sp();
    memset(S, 0, sizeof(double) * n * n);
    memset(H, 0, sizeof(double) * n * n);
    static Rtab R;
    for (int a = 0; a < nset; ++a)
        for (int b = 0; b < nset; ++b) {
            Pair pr;
            make_pair(&s, a, b, &pr, 2);
            if (pr.EIJ < 1.0e-14) continue;
            double aa = s.alpha[a], bb = s.alpha[b], p = pr.p, EIJ = pr.EIJ;
            int u = set_center(&s, a), v = set_center(&s, b);
            double PA[3], PB[3];
            for (int i = 0; i < 3; ++i) { PA[i] = pr.P[i] - X(&s, u, i); PB[i] = pr.P[i] - X(&s, v, i); }
            Ecoef E[3];
            int la = set_maxL(&s, a), lb = set_maxL(&s, b);
            for (int c = 0; c < 3; ++c) hermite_E(&E[c], PA[c], PB[c], p, la, lb + 2);
            for (int i = 0; i < set_norb(&s, a); ++i) {
                int oa = set_orb(&s, a, i), na[3];
                orb_lvec(&s, oa, na);
                for (int j = 0; j < set_norb(&s, b); ++j) {
                    int ob = set_orb(&s, b, j), nb[3];
                    orb_lvec(&s, ob, nb);
                    /* overlap (int1e.f90:374-380) */
                    double t = EIJ * pow(pi / p, 3.0 / 2.0) * s.bas[a * s.OpS + i] * s.bas[b * s.OpS + j];
                    t = t * gtoD(orb_ang(&s, oa), aa);
                    t = t * gtoD(orb_ang(&s, ob), bb);
                    t = t * Eget(&E[0], 0, na[0], nb[0]) * Eget(&E[1], 0, na[1], nb[1]) * Eget(&E[2], 0, na[2], nb[2]);
                    S[oa + n * ob] += t;
                    /* kinetic (int1e.f90:442-469) */
                    double val = 0.0;
                    for (int c = 0; c < 3; ++c) {
                        int o1 = c == 0 ? 1 : 0, o2 = c == 2 ? 1 : 2;
                        double tt = nb[c] * (nb[c] - 1) * Eget(&E[c], 0, na[c], nb[c] - 2);
                        tt = tt - 2.0 * bb * nb[c] * Eget(&E[c], 0, na[c], nb[c]);
                        tt = tt - 2.0 * bb * (nb[c] + 1) * Eget(&E[c], 0, na[c], nb[c]);
                        tt = tt + 4.0 * pow(bb, 2.0) * Eget(&E[c], 0, na[c], nb[c] + 2);
                        tt = tt * Eget(&E[o1], 0, na[o1], nb[o1]) * Eget(&E[o2], 0, na[o2], nb[o2]);
                        val = val + tt;
                    }
                    val = val * (-0.5) * EIJ * pow(pi / p, 3.0 / 2.0);
                    val = val * s.bas[a * s.OpS + i] * s.bas[b * s.OpS + j];
                    val = val * gtoD(orb_ang(&s, oa), aa);
                    val = val * gtoD(orb_ang(&s, ob), bb);
                    H[oa + n * ob] += val;
                }
            }
            /* nuclear attraction (coulomb, int1e.f90:479-582) */
            int Q = 3 * (la + lb);
            for (int c = 0; c < nnuc; ++c) {
                double CP[3];
                for (int i = 0; i < 3; ++i) CP[i] = X(&s, c, i) - pr.P[i];
                double TT = p * (pow(CP[0], 2.0) + pow(CP[1], 2.0) + pow(CP[2], 2.0));
                double Fj[17];
                oracle_boys(Q, TT, Ft, Fj);
                memset(R.seen, 0, sizeof(R.seen));
                for (int k = 0; k < pr.n; ++k) {
                    double r = Rget(&R, -CP[0], -CP[1], -CP[2], pr.C[k][0], pr.C[k][1], pr.C[k][2], 0, p, Fj);
                    double temp = Z[c] * (2.0 * pi / p) * r * pr.D[k];
                    H[pr.O[k][0] + n * pr.O[k][1]] -= temp;
                }
            }
        }
    return 0;
}

/* RHFI2G (src/I2G/RHFI2G.f90:80-90) */
void oracle_fock_G(int n, const double *XX, const double *D, double *G) {
    size_t n2 = (size_t)n * n, n3 = n2 * n;
#define XI(i, j, g, h) XX[(size_t)(i) + n * (size_t)(j) + n2 * (size_t)(g) + n3 * (size_t)(h)]
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double t = 0.0;
            for (int l = 0; l < n; ++l)
                for (int k = 0; k < n; ++k)
                    t = t + D[k + (size_t)n * l] * (XI(i, j, k, l) - 0.25 * XI(i, k, j, l) - 0.25 * XI(i, l, j, k));
            G[i + (size_t)n * j] = t;
        }
#undef XI
}

/* densRHF (src/dens/dens.f90:115-124) */
void oracle_density(int n, const double *C, int nocc, double *D) {
    for (int v = 0; v < n; ++v)
        for (int u = 0; u < n; ++u) {
            double t = 0.0;
            for (int i = 0; i < nocc; ++i) t = t + C[u + (size_t)n * i] * C[v + (size_t)n * i];
            D[u + (size_t)n * v] = t * 2;
        }
}

/* mp2_rhf (src/mp2/mp2.f90:79-144) over iajb[i][j] records of (a fastest)
 * v x v values: rec(a,b) = (ia|jb). */
void oracle_mp2(int nocc, int nvrt, const double *eig, const double *iajb, double *e_aa, double *e_ab) {
    double sum1 = 0.0, sum2 = 0.0;
    size_t rs = (size_t)nvrt * nvrt;
#define R(i, j, a, b) iajb[((size_t)(i) * nocc + (j)) * rs + (a) + (size_t)nvrt * (b)]
#define DEN(i, j, a, b) (eig[i] + eig[j] - eig[(a) + nocc] - eig[(b) + nocc])
    for (int i = 0; i <= nocc - 2; ++i) {
        for (int j = 0; j <= i; ++j)
            for (int a = 0; a < nvrt; ++a)
                for (int b = 0; b < nvrt; ++b) sum2 = sum2 + pow(R(i, j, a, b), 2.0) / DEN(i, j, a, b);
        for (int j = i + 1; j <= nocc - 1; ++j) {
            for (int a = 0; a <= nvrt - 2; ++a) {
                for (int b = 0; b <= a; ++b) sum2 = sum2 + pow(R(i, j, a, b), 2.0) / DEN(i, j, a, b);
                for (int b = a + 1; b < nvrt; ++b) {
                    sum2 = sum2 + pow(R(i, j, a, b), 2.0) / DEN(i, j, a, b);
                    sum1 = sum1 + pow(R(i, j, a, b) - R(i, j, b, a), 2.0) / DEN(i, j, a, b);
                }
            }
            for (int b = 0; b < nvrt; ++b) sum2 = sum2 + pow(R(i, j, nvrt - 1, b), 2.0) / DEN(i, j, nvrt - 1, b);
        }
    }
    for (int j = 0; j < nocc; ++j)
        for (int a = 0; a < nvrt; ++a)
            for (int b = 0; b < nvrt; ++b) sum2 = sum2 + pow(R(nocc - 1, j, a, b), 2.0) / DEN(nocc - 1, j, a, b);
#undef R
#undef DEN
    *e_aa = sum1;
    *e_ab = sum2;
}

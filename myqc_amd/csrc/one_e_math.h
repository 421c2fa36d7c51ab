// one_e_math.h -- overlap, kinetic and nuclear-attraction integrals of one
// contracted s/p shell pair (host and device).  Values follow int1e's proc1e
// (src/integrals/int1e.f90:132-280): overlap (:321-386) and kinetic
// (:391-474) from 1-D Hermite overlap coefficients E^{ij}_0, nuclear
// attraction (:479-582) by Obara-Saika recursion over F_m(p |PC|^2) at the
// reference's Boys order Q = 3*(sum of the two sets' max L), single-precision
// pi, and the EIJ < 1e-14 pair screen (:237-241).
#pragma once

#include "eri_math.h"

namespace myqc {

// E^{ij}_t for i <= 1, j <= 3, t <= 4 (Hermite expansion coefficients; t = 0
// is the 1-D overlap factor without (pi/p)^{1/2})
MYQC_HD double e1d_t(int i, int j, int t, double PA, double PB, double p) {
    if (i < 0 || j < 0 || t < 0 || t > 4) return 0.0;
    // E[t][i][j] by the Hermite recursion on j, then i
    double E[5][2][4];
    for (int t = 0; t < 5; ++t)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 4; ++b) E[t][a][b] = 0.0;
    E[0][0][0] = 1.0;
    const double h = 0.5 / p;
    for (int b = 1; b <= 3; ++b)
        for (int t = 0; t <= b; ++t)
            E[t][0][b] = (t > 0 ? h * E[t - 1][0][b - 1] : 0.0) + PB * E[t][0][b - 1] + (t + 1 <= 4 ? (t + 1) * E[t + 1][0][b - 1] : 0.0);
    for (int b = 0; b <= 3; ++b)
        for (int t = 0; t <= b + 1; ++t)
            E[t][1][b] = (t > 0 ? h * E[t - 1][0][b] : 0.0) + PA * E[t][0][b] + (t + 1 <= 4 ? (t + 1) * E[t + 1][0][b] : 0.0);
    return E[t][i][j];
}

MYQC_HD double e1d(int i, int j, double PA, double PB, double p) { return e1d_t(i, j, 0, PA, PB, p); }

// Dipole integrals <a| r_c |b> (c = x, y, z, about the coordinate origin) of
// one shell pair, as int1p's dipole (src/integrals/int1p.f90:265-355):
//   (pi/p)^{3/2} K (E^c_1 + P_c E^c_0) prod_{d != c} E^d_0,
// with single-precision pi and the EIJ < 1e-14 pair screen (:214-218).
MYQC_HD void dipole_pair(const ShellPair &sp, const PrimPair *__restrict__ prims, double dip[3][3][3]) {
    const int na = sp.la ? 3 : 1, nb = sp.lb ? 3 : 1;
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) dip[c][a][b] = 0.0;
    for (int ip = 0; ip < sp.npp; ++ip) {
        const PrimPair pr = prims[sp.pp0 + ip];
        if (pr.EIJ < 1.0e-14) continue;
        const double p = pr.p;
        double PB[3];
        for (int d = 0; d < 3; ++d) PB[d] = pr.PA[d] + sp.AB[d];
        const double pref = pow(PI_SP / p, 1.5) * pr.K;
        for (int a = 0; a < na; ++a)
            for (int b = 0; b < nb; ++b) {
                int la[3] = {0, 0, 0}, lb[3] = {0, 0, 0};
                if (sp.la) la[a] = 1;
                if (sp.lb) lb[b] = 1;
                double E0[3], E1[3];
                for (int d = 0; d < 3; ++d) {
                    E0[d] = e1d_t(la[d], lb[d], 0, pr.PA[d], PB[d], p);
                    E1[d] = e1d_t(la[d], lb[d], 1, pr.PA[d], PB[d], p);
                }
                for (int c = 0; c < 3; ++c)
                    dip[c][a][b] += pref * (E1[c] + pr.P[c] * E0[c]) * E0[(c + 1) % 3] * E0[(c + 2) % 3];
            }
    }
}

// S, H = T + V contributions of one shell pair; atoms[4c..4c+3] = Z, x, y, z.
// Only the nuclei c0, c0 + cstep, ... enter V, and only the primitive pairs
// c0, c0 + cstep, ... enter S and T, so that cstep callers (the lanes of a
// wavefront) summing their s and h cover the pair once; the defaults take
// everything.
// LA, LB are the pair's angular momenta (the first shell carries the larger).
template <int LA, int LB>
MYQC_HD void one_e_pair_t(const ShellPair &sp, const PrimPair *__restrict__ prims, const double *__restrict__ ftT,
                          const double *__restrict__ atoms, int nnuc, double s[3][3], double h[3][3], int c0,
                          int cstep) {
    constexpr int na = LA ? 3 : 1, nb = LB ? 3 : 1;
    double kin[3][3], v[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) s[a][b] = kin[a][b] = v[a][b] = 0.0;
    const int Q = 3 * sp.bl;
    for (int ip = 0; ip < sp.npp; ++ip) {
        const PrimPair pr = prims[sp.pp0 + ip];
        if (pr.EIJ < 1.0e-14) continue;                  // int1e.f90:238-241
        const double p = pr.p;
        double PB[3];
        for (int d = 0; d < 3; ++d) PB[d] = pr.PA[d] + sp.AB[d];   // P - B = (P - A) + (A - B)
        const double beta = pr.beta;
        const bool st = ip % cstep == c0;
        const double pref = st ? pow(PI_SP / p, 1.5) * pr.K : 0.0;
        if (st) {
#pragma unroll
            for (int a = 0; a < na; ++a)
#pragma unroll
                for (int b = 0; b < nb; ++b) {
                    int la[3] = {0, 0, 0}, lb[3] = {0, 0, 0};
                    if (LA) la[a] = 1;
                    if (LB) lb[b] = 1;
                    double Ex[3], kx[3];
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const int i = la[d], j = lb[d];
                        Ex[d] = e1d(i, j, pr.PA[d], PB[d], p);
                        // int1e.f90:442-462
                        kx[d] = j * (j - 1) * e1d(i, j - 2, pr.PA[d], PB[d], p) - 2.0 * beta * j * Ex[d] -
                                2.0 * beta * (j + 1) * Ex[d] + 4.0 * beta * beta * e1d(i, j + 2, pr.PA[d], PB[d], p);
                    }
                    s[a][b] += pref * Ex[0] * Ex[1] * Ex[2];
                    kin[a][b] += -0.5 * pref * (kx[0] * Ex[1] * Ex[2] + Ex[0] * kx[1] * Ex[2] + Ex[0] * Ex[1] * kx[2]);
                }
        }
        // nuclear attraction: [a|V|s]^(m) for a up to level la+lb, then HRR
        for (int c = c0; c < nnuc; c += cstep) {
            const double Zc = atoms[4 * c];
            double PC[3];
            for (int d = 0; d < 3; ++d) PC[d] = pr.P[d] - atoms[4 * c + 1 + d];
            const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
            double F[3];
            boys_ref<2>(Q, T, ftT, F);
            const double base = -Zc * (2.0 * PI_SP / p) * pr.K;
            double v0[3];
            for (int m = 0; m < 3; ++m) v0[m] = base * F[m];
            double v1[3][2];   // [p_i]^(m), m <= 1
            for (int i = 0; i < 3; ++i)
                for (int m = 0; m < 2; ++m) v1[i][m] = pr.PA[i] * v0[m] - PC[i] * v0[m + 1];
            const double hh = 0.5 / p;
            if constexpr (!LA && !LB) {
                v[0][0] += v0[0];
            } else if constexpr (LA && !LB) {
#pragma unroll
                for (int a = 0; a < 3; ++a) v[a][0] += v1[a][0];
            } else {
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int b = 0; b < 3; ++b) {
                        // [p_a + 1_b|s] then (p_a|p_b) = [p_a+1_b| + AB_b [p_a|
                        const double d2 = pr.PA[b] * v1[a][0] - PC[b] * v1[a][1] + (a == b ? hh * (v0[0] - v0[1]) : 0.0);
                        v[a][b] += d2 + sp.AB[b] * v1[a][0];
                    }
            }
        }
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) h[a][b] = kin[a][b] + v[a][b];
}

// S, H of one shell pair with runtime angular momenta (first shell larger)
MYQC_HD void one_e_pair(const ShellPair &sp, const PrimPair *__restrict__ prims, const double *__restrict__ ftT,
                        const double *__restrict__ atoms, int nnuc, double s[3][3], double h[3][3], int c0 = 0,
                        int cstep = 1) {
    if (sp.la && sp.lb) one_e_pair_t<1, 1>(sp, prims, ftT, atoms, nnuc, s, h, c0, cstep);
    else if (sp.la) one_e_pair_t<1, 0>(sp, prims, ftT, atoms, nnuc, s, h, c0, cstep);
    else one_e_pair_t<0, 0>(sp, prims, ftT, atoms, nnuc, s, h, c0, cstep);
}

}  // namespace myqc

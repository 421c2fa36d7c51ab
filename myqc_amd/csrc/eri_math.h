// eri_math.h -- contracted s/p shell-quartet ERIs by Obara-Saika vertical
// recursion and Head-Gordon-Pople horizontal recursion, FP64.
//
// Everything here is __host__ __device__ so the same code runs in the gfx950
// kernels and in the host-side check against the oracle.
//
// Numerics follow the reference's integral path where they change values at
// the 1e-10 level (src/integrals/int2e.f90, src/integrals/auxilary.f90):
//   * pi rounded to single precision (the `Pi = 3.1415926535897931` literals
//     without a D0 suffix, int2e.f90:621, auxilary.f90:169,196,711);
//   * Boys F_m(T): for T < 12 a 7-term Taylor expansion of F_Q from the
//     tabulated Ftab at Tk = NINT(10 T), Q = 3 * (sum of the four sets' max L),
//     then downward recursion (auxilary.f90:130-162); for 12 <= T < 2Q+36 the
//     asymptotic F_0 with the piecewise g(T), upward recursion
//     (auxilary.f90:167-189,265-285); beyond, the simplified upward form
//     (auxilary.f90:194-215);
//   * primitive quartets with EIJ*EGH < 1e-14 skipped (int2e.f90:257).
// The recursion itself (OS/HGP rather than McMurchie-Davidson) differs from
// the reference only at rounding level.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include <cmath>
#include <vector>

#define MYQC_HD __host__ __device__ __forceinline__

namespace myqc {

// `REAL(KIND=8),PARAMETER :: Pi = 3.1415926535897931` holds the single-precision value
constexpr double PI_SP = (double)3.1415926535897931f;

constexpr int FT_NT = 121;   // Tk = 0..120  (T = 0 .. 12 step 0.1)
constexpr int FT_NM = 19;    // m  = 0..18 of the record's 23: the largest order read is Q + 6 = 18 (Q <= 12)
// The staged table also holds e^{-k/10}, k = 0..600, after the Boys values
// (FT_E0 on): e^{-T} for T < 60 (the largest T at which it is used, 2Q + 36
// with Q <= 12) from one entry and a short series
constexpr int FT_NE = 601;
constexpr int FT_E0 = FT_NT * FT_NM;
// and, from FT_G0 on, the coefficients of BoysG's g(T) = c0 + c1/T + c2/T^2 +
// c3/T^3 on its four ranges of T (12-15, 15-18, 18-24, 24 on)
constexpr int FT_G0 = FT_E0 + FT_NE;
constexpr int FT_TOTAL = FT_G0 + 16;

// Cartesian components up to l = 2: 0 s | 1 x 2 y 3 z | 4 xx 5 xy 6 xz 7 yy 8 yz 9 zz
MYQC_HD constexpr int cexp(int idx, int d) {
    // exponents (x,y,z) per component
    return (d == 0) ? (idx == 1 ? 1 : idx == 4 ? 2 : (idx == 5 || idx == 6) ? 1 : 0)
         : (d == 1) ? (idx == 2 ? 1 : idx == 7 ? 2 : (idx == 5 || idx == 8) ? 1 : 0)
                    : (idx == 3 ? 1 : idx == 9 ? 2 : (idx == 6 || idx == 8) ? 1 : 0);
}
MYQC_HD constexpr int lvl(int idx) { return idx == 0 ? 0 : idx < 4 ? 1 : 2; }
MYQC_HD constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
MYQC_HD constexpr int nupto(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }  // components of level <= l
MYQC_HD constexpr int first(int l) { return nupto(l - 1 < 0 ? -1 : l - 1) * (l > 0); }
// index of the component with exponents (x,y,z), total <= 2
MYQC_HD constexpr int cidx(int x, int y, int z) {
    return (x + y + z == 0) ? 0
         : (x + y + z == 1) ? (x ? 1 : y ? 2 : 3)
         : (x == 2 ? 4 : (x == 1 && y == 1) ? 5 : (x == 1 && z == 1) ? 6 : y == 2 ? 7 : (y == 1 && z == 1) ? 8 : 9);
}
// component minus one unit along d (-1 if none)
MYQC_HD constexpr int lower(int idx, int d) {
    return cexp(idx, d) == 0 ? -1
         : cidx(cexp(idx, 0) - (d == 0), cexp(idx, 1) - (d == 1), cexp(idx, 2) - (d == 2));
}
// component plus one unit along d (level <= 2 only)
MYQC_HD constexpr int raise(int idx, int d) {
    return cidx(cexp(idx, 0) + (d == 0), cexp(idx, 1) + (d == 1), cexp(idx, 2) + (d == 2));
}
// the direction used to build a component by recursion: its first nonzero axis
MYQC_HD constexpr int bdir(int idx) { return cexp(idx, 0) ? 0 : cexp(idx, 1) ? 1 : 2; }

// 1/sqrt(x) for x > 0 in the normal range: the device's reciprocal square
// root estimate y refined by one second-order step, y + y e (1/2 + 3/8 e)
// with e = 1 - x y^2 (the series of (1 - e)^(-1/2)), or the host's
// 1/sqrt.  No handling of 0, infinity or denormals: every caller passes a
// sum of Gaussian exponents or a Boys argument above 12.
MYQC_HD double rsqrt_hd(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
#else
    return 1.0 / sqrt(x);
#endif
}

// The predicate on some active lane of the wavefront (device; the value is
// wave-uniform, so a branch on it jumps over code no lane needs instead of
// running it under a mask), or the predicate itself (host).
MYQC_HD bool any_lane(bool p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(p) != 0;
#else
    return p;
#endif
}

// 1/x for x > 0 in the normal range: the device's reciprocal estimate refined
// by two Newton-Raphson steps (within about an ulp of the quotient, without
// the full division sequence), or the host's division
MYQC_HD double recip_hd(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}

// The staged table ftT of (Boys table transposed, ftT[Tk*FT_NM + m]; then
// e^{-k/10} at FT_E0 + k; then g's coefficients at FT_G0 + 4 range + power),
// built from the Ftab record (Ft(Tk, m) at Tk + 121 m).  Host only.
inline std::vector<double> build_ftT(const double *Ftab) {
    std::vector<double> ftT(FT_TOTAL);
    for (int tk = 0; tk < FT_NT; ++tk)
        for (int m = 0; m < FT_NM; ++m) ftT[tk * FT_NM + m] = Ftab[tk + FT_NT * m];
    for (int k = 0; k < FT_NE; ++k) ftT[FT_E0 + k] = std::exp(-0.1 * k);
    // BoysG (auxilary.f90:271-283); beyond T = 30 the reference's g is
    // undefined and 0.49 is used (changes F_0 by at most 3e-15 |g|)
    const double gc[16] = {0.4999489092, -0.2473631686, 0.321180909, -0.3811559346,
                           0.4998436875, -0.24249438,   0.24642845,  0.0,
                           0.499093162,  -0.2152832,    0.0,         0.0,
                           0.490,        0.0,           0.0,         0.0};
    for (int k = 0; k < 16; ++k) ftT[FT_G0 + k] = gc[k];
    return ftT;
}

// Boys F_0..F_M(T) as the reference computes them at order Q (M <= Q <= 16).
// ftT: the staged table (build_ftT).
template <int M>
MYQC_HD void boys_ref(int Q, double T, const double *__restrict__ ftT, double (&F)[M + 1]) {
    // Tk = NINT(10 T) (T >= 0; T clamped to 60, beyond every table use) and
    // x = Tk / 10 - T, |x| <= 0.05.  Tk / 10 without a division: the product
    // and one fma correction equal the correctly rounded quotient for every
    // integer 0..1000 (checked exhaustively)
    const double Tc = T < 60.0 ? T : 60.0;
    const int Tk = (int)floor(Tc * 10.0 + 0.5);
    const double tkd = (double)Tk, q10 = tkd * 0.1;
    const double x = fma(fma(-q10, 10.0, tkd), 0.1, q10) - Tc;
    // e^{-T} feeds the downward recursion below T = 12 (when Q > 0) and the
    // asymptotic form up to 2Q + 36; one evaluation serves both branches (the
    // lanes of a wavefront usually fall on both sides of T = 12):
    // e^{-T} = e^{-Tk/10} e^x with the Taylor series of e^x to x^7 (the next
    // term is below 1e-15 relative)
    double e = 0.0;
    if ((Q > 0 || T >= 12.0) && T < 2.0 * Q + 36.0) {
        double s = 1.0 / 5040.0;
        s = fma(s, x, 1.0 / 720.0);
        s = fma(s, x, 1.0 / 120.0);
        s = fma(s, x, 1.0 / 24.0);
        s = fma(s, x, 1.0 / 6.0);
        s = fma(s, x, 0.5);
        s = fma(s, x, 1.0);
        s = fma(s, x, 1.0);
        e = ftT[FT_E0 + Tk] * s;
    }
    if (T < 12.0) {
        const double *r = ftT + Tk * FT_NM + Q;
        // sum_{k=0..6} Ft(Tk,Q+k) x^k / k!   (auxilary.f90:152-155), in
        // Horner form r0 + x (r1 + x/2 (r2 + x/3 (r3 + ...)))
        double fq = r[6];
#pragma unroll
        for (int k = 5; k >= 0; --k) fq = fma(fq, x * (1.0 / (k + 1)), r[k]);
        // with Q = 0 (sets of s functions only) there is no recursion: F_0 is
        // the Taylor sum itself
        const int top = Q < M ? Q : M;
        double t2 = 2.0 * T;
        // downward recursion F_j = (2T F_{j+1} + e^{-T}) / (2j+1)  (auxilary.f90:158-160),
        // with 1/(2j+1) as constants; the reference divides (rounding-level difference)
        // Q <= 12: sets have max L <= 1 (checked when the system is built), so
        // Q = 3 (sum of four set L's) <= 12.  Q is a multiple of 3, so the
        // steps j = Q-1 .. M run in blocks of three (j = 3b+2, 3b+1, 3b for
        // Q > 3b), each behind one branch that a wavefront skips when none of
        // its lanes needs it (any_lane)
#pragma unroll
        for (int b = 3; b >= 0; --b) {
            if (3 * b + 2 < M) continue;
            if (any_lane(Q > 3 * b)) {
                const bool on = Q > 3 * b;
#pragma unroll
                for (int j = 3 * b + 2; j >= 3 * b; --j)
                    if (j >= M) {
                        const double v = (t2 * fq + e) * (1.0 / (2.0 * j + 1.0));
                        fq = on ? v : fq;
                    }
            }
        }
        // fq now holds F_top, top = min(Q, M); the reference never forms F_j for
        // j > Q, and callers only read F_j for j <= Q
#pragma unroll
        for (int j = M; j >= 0; --j) {
            if (j > top) F[j] = 0.0;
            else if (j == top) F[j] = fq;
            else F[j] = (t2 * F[j + 1] + e) * (1.0 / (2.0 * j + 1.0));
        }
    } else {
        const double half_sqrt_pi = 0.5 * sqrt(PI_SP);
        const double rt = rsqrt_hd(T);
        const double iT = rt * rt;
        const double it2 = 0.5 * iT;
        if (T < 2.0 * Q + 36.0) {
            // auxilary.f90:181-187 with BoysG (:271-283): g's coefficients
            // for T's range from the staged table, one polynomial for every
            // lane (no branch per range)
            const int rg = (T >= 15.0) + (T >= 18.0) + (T >= 24.0);
            const double *gc = ftT + FT_G0 + 4 * rg;
            const double g = fma(fma(fma(gc[3], iT, gc[2]), iT, gc[1]), iT, gc[0]);
            F[0] = half_sqrt_pi * rt - e * g * iT;
#pragma unroll
            for (int j = 1; j <= M; ++j) F[j] = it2 * ((2.0 * (j - 1) + 1.0) * F[j - 1] - e);
        } else {
            F[0] = half_sqrt_pi * rt;
#pragma unroll
            for (int j = 1; j <= M; ++j) F[j] = it2 * (2.0 * (j - 1) + 1.0) * F[j - 1];
        }
    }
}

// One primitive pair of a shell pair (first shell carries the larger L).
struct PrimPair {
    double p;        // alpha + beta
    double P[3];     // Gaussian product centre
    double PA[3];    // P - A (A = centre of the first shell)
    double K;        // EIJ * c_a * c_b  (normalised contraction weights folded in)
    double EIJ;      // exp(-alpha beta |AB|^2 / p): the reference's pair screen factor
    double beta;     // exponent of the second shell's primitive (kinetic energy)
    double ip;       // 1 / p
    // default construction leaves the fields unset: arrays of primitive pairs
    // are sized before they are filled
    MYQC_HD PrimPair() {}
};

// Contracted shell pair.
struct ShellPair {
    double AB[3];    // A - B
    int pp0;         // first PrimPair
    int npp;         // number of primitive pairs
    int ao_a, ao_b;  // first AO of each shell
    int la, lb;      // angular momentum of each shell (la >= lb)
    int bl;          // sum of the two shells' set max-L (Boys order Q = 3*(bl_bra+bl_ket))
    int same;        // both shells identical
    int sa, sb;      // shell indices of the first and second shell
};

// Contracted [e0|f0] for e in levels LA..LA+LB and f in LC..LC+LD, accumulated
// over all primitive quartets of the shell quartet (bra, ket).
template <int LA, int LB, int LC, int LD>
struct QuartetAcc {
    static constexpr int EB = LA + LB, EK = LC + LD, M = EB + EK;
    static constexpr int E0 = first(LA), NE = nupto(EB) - E0;
    static constexpr int F0 = first(LC), NF = nupto(EK) - F0;
    double v[NE][NF];
};

// One primitive quartet (b k) of a shell quartet added into acc: the Boys
// values, the bra vertical recursion [e0|00]^(m), the ket vertical recursion
// [e0|f0]^(m), and the targets.  Q: the Boys order rule of the shell
// quartet; c0 = 2 pi^2.5.
template <int LA, int LB, int LC, int LD>
MYQC_HD __forceinline__ void prim_quartet(const PrimPair &b, const PrimPair &k, int Q, double c0,
                                          const double *__restrict__ ftT, QuartetAcc<LA, LB, LC, LD> &acc) {
    using A = QuartetAcc<LA, LB, LC, LD>;
    constexpr int EB = A::EB, EK = A::EK, M = A::M;
    constexpr int NB = nupto(EB);
    const double p = b.p, q = k.p, pq = p + q;
    const double ipq = recip_hd(pq);
    const double rho = p * q * ipq;
    const double rsq = rsqrt_hd(pq);
    double PQ[3], WP[3], WQ[3], QC[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        PQ[d] = b.P[d] - k.P[d];
        // W = (pP + qQ)/(p+q):  W - P = -q/(p+q) PQ,  W - Q = p/(p+q) PQ
        WP[d] = -q * ipq * PQ[d];
        WQ[d] = p * ipq * PQ[d];
        QC[d] = k.PA[d];
    }
    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    const double pref = c0 * b.ip * k.ip * rsq * b.K * k.K;
    double F[M + 1];
    boys_ref<M>(Q, T, ftT, F);
    // bra vertical recursion [e0|00]^(m)
    double Vb[NB][M + 1];
#pragma unroll
    for (int m = 0; m <= M; ++m) Vb[0][m] = pref * F[m];
    if constexpr (EB > 0) {
        const double i2p = 0.5 * b.ip, rp = q * ipq;
#pragma unroll
        for (int idx = 1; idx < NB; ++idx) {
            const int d = bdir(idx), lo = lower(idx, d);
            const int am1 = cexp(lo, d);
#pragma unroll
            for (int m = 0; m <= M - lvl(idx); ++m) {
                double v = b.PA[d] * Vb[lo][m] + WP[d] * Vb[lo][m + 1];
                if (am1 > 0) {
                    const int lo2 = lower(lo, d);
                    v += am1 * i2p * (Vb[lo2][m] - rp * Vb[lo2][m + 1]);
                }
                Vb[idx][m] = v;
            }
        }
    }
    // ket vertical recursion [e0|f0]^(m), then accumulate the targets
    if constexpr (EK == 0) {
#pragma unroll
        for (int e = A::E0; e < NB; ++e) acc.v[e - A::E0][0] += Vb[e][0];
    } else {
        constexpr int NK = nupto(EK);
        const double i2q = 0.5 * k.ip, rq = p * ipq, i2pq = 0.5 * ipq;
        double W[NB][NK][M + 1];
#pragma unroll
        for (int e = 0; e < NB; ++e)
#pragma unroll
            for (int m = 0; m <= M - lvl(e); ++m) W[e][0][m] = Vb[e][m];
#pragma unroll
        for (int f = 1; f < NK; ++f) {
            const int d = bdir(f), lo = lower(f, d);
            const int cm1 = cexp(lo, d);
#pragma unroll
            for (int e = 0; e < NB; ++e) {
#pragma unroll
                for (int m = 0; m <= M - lvl(e) - lvl(f); ++m) {
                    double v = QC[d] * W[e][lo][m] + WQ[d] * W[e][lo][m + 1];
                    if (cm1 > 0) {
                        const int lo2 = lower(lo, d);
                        v += cm1 * i2q * (W[e][lo2][m] - rq * W[e][lo2][m + 1]);
                    }
                    if (cexp(e, d) > 0) {
                        const int el = lower(e, d);
                        v += cexp(e, d) * i2pq * W[el][lo][m + 1];
                    }
                    W[e][f][m] = v;
                }
            }
        }
#pragma unroll
        for (int e = A::E0; e < NB; ++e)
#pragma unroll
            for (int f = A::F0; f < NK; ++f) acc.v[e - A::E0][f - A::F0] += W[e][f][0];
    }
}

// SCREEN = false skips the reference's primitive-pair screen: used for the
// Schwarz factors, since a primitive pair dropped from (ab|ab) by
// EIJ^2 < 1e-14 can survive EIJ*EGH >= 1e-14 against a compact ket pair.
// The ket's primitive pairs come from kprim(j), j < ket.npp.
template <int LA, int LB, int LC, int LD, bool SCREEN = true, typename KF>
MYQC_HD void quartet_prims_f(const ShellPair &bra, const ShellPair &ket, const PrimPair *__restrict__ ppb,
                             KF &&kprim, const double *__restrict__ ftT, QuartetAcc<LA, LB, LC, LD> &acc) {
    using A = QuartetAcc<LA, LB, LC, LD>;
    const int Q = 3 * (bra.bl + ket.bl);
#pragma unroll
    for (int e = 0; e < A::NE; ++e)
#pragma unroll
        for (int f = 0; f < A::NF; ++f) acc.v[e][f] = 0.0;
    const double c0 = 2.0 * pow(PI_SP, 2.5);
    // ket primitives outer: a lane's own ket data is loaded once per quartet,
    // and the inner loop reads the bra's primitives, whose addresses the lanes
    // of a wavefront (one bra pair, consecutive kets) mostly share
    for (int j = 0; j < ket.npp; ++j) {
        const PrimPair k = kprim(j);
        for (int i = 0; i < bra.npp; ++i) {
            const PrimPair b = ppb[bra.pp0 + i];
            if (SCREEN && b.EIJ * k.EIJ < 1.0e-14) continue;          // int2e.f90:257
            prim_quartet<LA, LB, LC, LD>(b, k, Q, c0, ftT, acc);
        }
    }
}

// The same sum split over SPLIT lanes: lane `sub` of the quartet's group
// takes the primitive quartets t = j npp_bra + i with t = sub (mod SPLIT);
// the group's partial sums are added afterwards (split_reduce on the device).
template <int LA, int LB, int LC, int LD, int SPLIT, typename KF>
MYQC_HD void quartet_prims_split(const ShellPair &bra, const ShellPair &ket, const PrimPair *__restrict__ ppb,
                                 KF &&kprim, const double *__restrict__ ftT, QuartetAcc<LA, LB, LC, LD> &acc,
                                 int sub) {
    using A = QuartetAcc<LA, LB, LC, LD>;
    const int Q = 3 * (bra.bl + ket.bl);
#pragma unroll
    for (int e = 0; e < A::NE; ++e)
#pragma unroll
        for (int f = 0; f < A::NF; ++f) acc.v[e][f] = 0.0;
    const double c0 = 2.0 * pow(PI_SP, 2.5);
    const int nbp = bra.npp, total = nbp * ket.npp;
    int j = 0, i = sub;
    while (i >= nbp && nbp > 0) {
        i -= nbp;
        ++j;
    }
    for (int t = sub; t < total; t += SPLIT) {
        const PrimPair k = kprim(j);
        const PrimPair b = ppb[bra.pp0 + i];
        if (!(b.EIJ * k.EIJ < 1.0e-14)) prim_quartet<LA, LB, LC, LD>(b, k, Q, c0, ftT, acc);   // int2e.f90:257
        i += SPLIT;
        while (i >= nbp) {
            i -= nbp;
            ++j;
        }
    }
}

template <int LA, int LB, int LC, int LD, bool SCREEN = true>
MYQC_HD void quartet_prims(const ShellPair &bra, const ShellPair &ket, const PrimPair *__restrict__ ppb,
                           const PrimPair *__restrict__ ppk, const double *__restrict__ ftT,
                           QuartetAcc<LA, LB, LC, LD> &acc) {
    quartet_prims_f<LA, LB, LC, LD, SCREEN>(bra, ket, ppb, [&](int j) { return ppk[ket.pp0 + j]; }, ftT, acc);
}

// Horizontal recursion: contracted (ab|cd) with a,b,c,d component indices
// within each shell (s: 0; p: 0,1,2 = x,y,z).  out[((a*NBb + b)*NCc + c)*NDd + d].
template <int LA, int LB, int LC, int LD>
MYQC_HD void quartet_hrr(const QuartetAcc<LA, LB, LC, LD> &acc, const double AB[3], const double CD[3],
                         double *out) {
    using A = QuartetAcc<LA, LB, LC, LD>;
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);
    // ket HRR for every bra level-(LA..LA+LB) component e: I[e][c][d]
    double I[A::NE][nc][nd];
#pragma unroll
    for (int e = 0; e < A::NE; ++e)
#pragma unroll
        for (int c = 0; c < nc; ++c)
#pragma unroll
            for (int d = 0; d < nd; ++d) {
                const int ci = first(LC) + c;       // component index of c
                if constexpr (LD == 0) {
                    I[e][c][d] = acc.v[e][ci - A::F0];
                } else {
                    // (e0|c p_d) = [e0|(c+1_d)0] + CD_d [e0|c0]
                    I[e][c][d] = acc.v[e][raise(ci, d) - A::F0] + CD[d] * acc.v[e][ci - A::F0];
                }
            }
#pragma unroll
    for (int a = 0; a < na; ++a)
#pragma unroll
        for (int b = 0; b < nb; ++b)
#pragma unroll
            for (int c = 0; c < nc; ++c)
#pragma unroll
                for (int d = 0; d < nd; ++d) {
                    const int ai = first(LA) + a;
                    double v;
                    if constexpr (LB == 0) {
                        v = I[ai - A::E0][c][d];
                    } else {
                        v = I[raise(ai, b) - A::E0][c][d] + AB[b] * I[ai - A::E0][c][d];
                    }
                    out[((a * nb + b) * nc + c) * nd + d] = v;
                }
}

// Degeneracy weight of element (i,j|k,l) of a shell-quartet block so that
// every unique integral is visited once in the J/K digestion: 0 for a repeat
// within a block whose shells or pairs coincide, else 1/2 per index equality
// (i==j, k==l, (ij)==(kl)).
MYQC_HD double block_weight(int a, int b, int c, int d, int same_ab, int same_cd, int same_pair, int nb, int nd,
                            int i, int j, int k, int l) {
    if (same_ab && a > b) return 0.0;
    if (same_cd && c > d) return 0.0;
    if (same_pair && (a * nb + b) > (c * nd + d)) return 0.0;
    double w = 1.0;
    if (i == j) w *= 0.5;
    if (k == l) w *= 0.5;
    if ((i == k && j == l) || (i == l && j == k)) w *= 0.5;
    return w;
}

}  // namespace myqc

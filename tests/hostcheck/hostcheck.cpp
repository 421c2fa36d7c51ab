// hostcheck.cpp -- TEST INFRASTRUCTURE: runs the library's own integral math
// (eri_math.h, one_e_math.h, system.h) on the CPU so it can be compared with
// the oracle without a GPU.  Built host-only; never part of the product path.
#include <cstring>
#include <string>
#include <vector>

#include "../../myqc_amd/csrc/eri_math.h"
#include "../../myqc_amd/csrc/one_e_math.h"
#include "../../myqc_amd/csrc/system.h"

using namespace myqc;

namespace {

struct Sys {
    std::vector<Shell> shells;
    std::vector<ShellPair> pairs;
    std::vector<PrimPair> prims;
    std::vector<int> ids[3];
    std::vector<double> ftT, xyz;
    int n = 0;
};

std::string make(Sys &s, int nnuc, const double *xyzF, int nset, const double *alpha, const int *setinfo,
                 const double *bas, int norb, const int *basinfo, const double *Ftab) {
    s.xyz.resize(3 * nnuc);
    for (int a = 0; a < nnuc; ++a)
        for (int c = 0; c < 3; ++c) s.xyz[3 * a + c] = xyzF[a + nnuc * c];
    s.n = norb;
    s.ftT = build_ftT(Ftab);
    return build_system(nnuc, s.xyz.data(), nset, alpha, setinfo, bas, norb, basinfo, s.shells, s.pairs, s.prims, s.ids);
}

int pclass(const ShellPair &p) { return p.la + p.lb; }

template <int LA, int LB, int LC, int LD, typename Fn>
void one(const Sys &s, int pb, int pk, Fn &&fn) {
    QuartetAcc<LA, LB, LC, LD> acc;
    const ShellPair &bra = s.pairs[pb], &ket = s.pairs[pk];
    quartet_prims<LA, LB, LC, LD>(bra, ket, s.prims.data(), s.prims.data(), s.ftT.data(), acc);
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);
    double out[na * nb * nc * nd];
    quartet_hrr<LA, LB, LC, LD>(acc, bra.AB, ket.AB, out);
    for (int ia = 0; ia < na; ++ia)
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < nc; ++ic)
                for (int id = 0; id < nd; ++id) {
                    const int i = bra.ao_a + ia, j = bra.ao_b + ib, k = ket.ao_a + ic, l = ket.ao_b + id;
                    const double w = block_weight(ia, ib, ic, id, bra.same, ket.same, pb == pk, nb, nd, i, j, k, l);
                    fn(i, j, k, l, out[((ia * nb + ib) * nc + ic) * nd + id], w);
                }
}

// visit every unique shell quartet once, bra = pair of higher class
template <typename Fn>
void all_quartets(const Sys &s, Fn &&fn) {
    const int np = (int)s.pairs.size();
    for (int P = 0; P < np; ++P)
        for (int Q = P; Q < np; ++Q) {
            int pb = P, pk = Q;
            if (pclass(s.pairs[P]) < pclass(s.pairs[Q])) std::swap(pb, pk);
            const int cb = pclass(s.pairs[pb]), ck = pclass(s.pairs[pk]);
            if (cb == 0 && ck == 0) one<0, 0, 0, 0>(s, pb, pk, fn);
            else if (cb == 1 && ck == 0) one<1, 0, 0, 0>(s, pb, pk, fn);
            else if (cb == 1 && ck == 1) one<1, 0, 1, 0>(s, pb, pk, fn);
            else if (cb == 2 && ck == 0) one<1, 1, 0, 0>(s, pb, pk, fn);
            else if (cb == 2 && ck == 1) one<1, 1, 1, 0>(s, pb, pk, fn);
            else one<1, 1, 1, 1>(s, pb, pk, fn);
        }
}

}  // namespace

extern "C" int hc_eri_full(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                           const double *bas, int norb, const int *basinfo, const double *Ftab, double *XX) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    const size_t n = norb;
    memset(XX, 0, sizeof(double) * n * n * n * n);
    all_quartets(s, [&](size_t i, size_t j, size_t k, size_t l, double v, double w) {
        if (w == 0.0) return;
        XX[i + n * (j + n * (k + n * l))] = v;
        XX[j + n * (i + n * (k + n * l))] = v;
        XX[i + n * (j + n * (l + n * k))] = v;
        XX[j + n * (i + n * (l + n * k))] = v;
        XX[k + n * (l + n * (i + n * j))] = v;
        XX[l + n * (k + n * (i + n * j))] = v;
        XX[k + n * (l + n * (j + n * i))] = v;
        XX[l + n * (k + n * (j + n * i))] = v;
    });
    return 0;
}

// the unique-integral form of the G build: each unique integral's six updates
// with its degeneracy weight (block_weight, shared with the dense XX fill)
extern "C" int hc_fock_G(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                         const double *bas, int norb, const int *basinfo, const double *Ftab, const double *D,
                         double *G) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    const size_t n = norb;
    std::vector<double> Ds(n * n), A(n * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) Ds[i + n * j] = 0.5 * (D[i + n * j] + D[j + n * i]);
    all_quartets(s, [&](size_t i, size_t j, size_t k, size_t l, double v, double w) {
        const double x = v * w;
        A[i + n * j] += 2.0 * Ds[k + n * l] * x;
        A[k + n * l] += 2.0 * Ds[i + n * j] * x;
        A[i + n * k] -= 0.5 * Ds[j + n * l] * x;
        A[j + n * k] -= 0.5 * Ds[i + n * l] * x;
        A[i + n * l] -= 0.5 * Ds[j + n * k] * x;
        A[j + n * l] -= 0.5 * Ds[i + n * k] * x;
    });
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) G[i + n * j] = A[i + n * j] + A[j + n * i];
    return 0;
}

extern "C" int hc_int1e(int nnuc, const int *Z, const double *xyz, int nset, const double *alpha, const int *setinfo,
                        const double *bas, int norb, const int *basinfo, const double *Ftab, double *S, double *H) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    std::vector<double> atoms(4 * nnuc);
    for (int c = 0; c < nnuc; ++c) {
        atoms[4 * c] = Z[c];
        for (int d = 0; d < 3; ++d) atoms[4 * c + 1 + d] = s.xyz[3 * c + d];
    }
    const size_t n = norb;
    for (const ShellPair &sp : s.pairs) {
        double ss[3][3], hh[3][3];
        one_e_pair(sp, s.prims.data(), s.ftT.data(), atoms.data(), nnuc, ss, hh);
        for (int a = 0; a < (sp.la ? 3 : 1); ++a)
            for (int b = 0; b < (sp.lb ? 3 : 1); ++b) {
                const size_t i = sp.ao_a + a, j = sp.ao_b + b;
                S[i + n * j] = S[j + n * i] = ss[a][b];
                H[i + n * j] = H[j + n * i] = hh[a][b];
            }
    }
    return 0;
}

extern "C" int hc_boys(int M, int Q, double T, const double *Ftab, double *F) {
    const std::vector<double> ftT = build_ftT(Ftab);
    switch (M) {
        case 0: { double f[1]; boys_ref<0>(Q, T, ftT.data(), f); memcpy(F, f, sizeof f); break; }
        case 1: { double f[2]; boys_ref<1>(Q, T, ftT.data(), f); memcpy(F, f, sizeof f); break; }
        case 2: { double f[3]; boys_ref<2>(Q, T, ftT.data(), f); memcpy(F, f, sizeof f); break; }
        case 3: { double f[4]; boys_ref<3>(Q, T, ftT.data(), f); memcpy(F, f, sizeof f); break; }
        case 4: { double f[5]; boys_ref<4>(Q, T, ftT.data(), f); memcpy(F, f, sizeof f); break; }
        default: return -1;
    }
    return 0;
}

// the row-owned digestion of jk.hip: for every ordered bra shell pair (A,B)
// and every shell pair (C,D) in its stored orientation, row i in A gets
//   J: G(i,j) += f D(k,l) (ij|kl),  f = 2 unless C == D
//   K: G(i,k) -= D(j,l) (ij|kl)/2,  and if C != D  G(i,l) -= D(j,k) (ij|kl)/2
// then G is symmetrised.
extern "C" int hc_fock_G_rows(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                              const double *bas, int norb, const int *basinfo, const double *Ftab, const double *D,
                              double *G) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    const size_t n = norb;
    std::vector<double> XX(n * n * n * n);
    if (hc_eri_full(nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab, XX.data()) != 0) return -1;
    std::vector<double> Ds(n * n), Gr(n * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) Ds[i + n * j] = 0.5 * (D[i + n * j] + D[j + n * i]);
    auto X = [&](size_t i, size_t j, size_t k, size_t l) { return XX[i + n * (j + n * (k + n * l))]; };
    const int ns = (int)s.shells.size();
    for (int A = 0; A < ns; ++A)
        for (int B = 0; B < ns; ++B)
            for (const ShellPair &q : s.pairs) {
                const int na = s.shells[A].L ? 3 : 1, nb = s.shells[B].L ? 3 : 1;
                const int nc = q.la ? 3 : 1, nd = q.lb ? 3 : 1;
                const double fj = q.same ? 1.0 : 2.0;
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < nb; ++b)
                        for (int c = 0; c < nc; ++c)
                            for (int d = 0; d < nd; ++d) {
                                const size_t i = s.shells[A].ao + a, j = s.shells[B].ao + b, k = q.ao_a + c,
                                             l = q.ao_b + d;
                                const double v = X(i, j, k, l);
                                Gr[i + n * j] += fj * Ds[k + n * l] * v;
                                Gr[i + n * k] -= 0.5 * Ds[j + n * l] * v;
                                if (!q.same) Gr[i + n * l] -= 0.5 * Ds[j + n * k] * v;
                            }
            }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) G[i + n * j] = 0.5 * (Gr[i + n * j] + Gr[j + n * i]);
    return 0;
}

// The pair-owned digestion of the stored integrals, restated on the host:
// every shell pair P owns a copy of (P|Q) for each ket Q (degeneracy-weighted
// values); the copy adds J to P's block, and the four K terms only when P is
// the bra of the unique quartet (higher class, or same class and rank(P) <=
// rank(Q)); the diagonal quartet (P|P) also adds its ket-side J; G = A + A^T.
extern "C" int hc_fock_G_pairs(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                               const double *bas, int norb, const int *basinfo, const double *Ftab, const double *D,
                               double *G) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    const size_t n = norb;
    std::vector<double> XX(n * n * n * n);
    if (hc_eri_full(nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab, XX.data()) != 0) return -1;
    std::vector<double> Ds(n * n), A(n * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) Ds[i + n * j] = 0.5 * (D[i + n * j] + D[j + n * i]);
    auto X = [&](size_t i, size_t j, size_t k, size_t l) { return XX[i + n * (j + n * (k + n * l))]; };
    std::vector<int> rank(s.pairs.size());
    for (int c = 0; c < 3; ++c)
        for (size_t r = 0; r < s.ids[c].size(); ++r) rank[s.ids[c][r]] = (int)r;
    for (int P = 0; P < (int)s.pairs.size(); ++P) {
        const ShellPair &p = s.pairs[P];
        const int cp = pclass(p);
        for (int ck = 0; ck < 3; ++ck)
            for (int Q : s.ids[ck]) {
                const ShellPair &q = s.pairs[Q];
                const bool bra = cp > ck || (cp == ck && rank[P] <= rank[Q]);
                const bool diag = P == Q;
                const int na = p.la ? 3 : 1, nb = p.lb ? 3 : 1, nc = q.la ? 3 : 1, nd = q.lb ? 3 : 1;
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < nb; ++b)
                        for (int c = 0; c < nc; ++c)
                            for (int d = 0; d < nd; ++d) {
                                const size_t i = p.ao_a + a, j = p.ao_b + b, k = q.ao_a + c, l = q.ao_b + d;
                                const double w = block_weight(a, b, c, d, p.same, q.same, diag, nb, nd, (int)i,
                                                              (int)j, (int)k, (int)l);
                                if (w == 0.0) continue;
                                const double x = X(i, j, k, l) * w;
                                A[i + n * j] += 2.0 * Ds[k + n * l] * x;
                                if (bra) {
                                    A[i + n * k] -= 0.5 * Ds[j + n * l] * x;
                                    A[j + n * k] -= 0.5 * Ds[i + n * l] * x;
                                    A[i + n * l] -= 0.5 * Ds[j + n * k] * x;
                                    A[j + n * l] -= 0.5 * Ds[i + n * k] * x;
                                }
                                if (diag) A[k + n * l] += 2.0 * Ds[i + n * j] * x;
                            }
            }
    }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) G[i + n * j] = A[i + n * j] + A[j + n * i];
    return 0;
}

// The three-copy row-owned digestion, restated on the host.  Each unique
// quartet {R, S} (R the bra: higher class, or same class and rank(R) <=
// rank(S)) is held by the bra pair in both orientations and by the ket pair
// in its stored orientation only.  A copy whose row pair is R and partner S
// adds, for the rows of its first shell, with G = A + A^T at the end:
//   bra role (R bra, or R == S): J x (forward ? (R.same ? 1/2 : 1) : 0),
//                                K x (R == S ? 1/2 : 1)
//   ket role (R ket):            J x (R.same ? 1/2 : 1), no K
// with J, K the row formula of hc_fock_G_rows (raw integrals).
extern "C" int hc_fock_G_3copy(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                               const double *bas, int norb, const int *basinfo, const double *Ftab, const double *D,
                               double *G) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    const size_t n = norb;
    std::vector<double> XX(n * n * n * n);
    if (hc_eri_full(nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab, XX.data()) != 0) return -1;
    std::vector<double> Ds(n * n), A(n * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) Ds[i + n * j] = 0.5 * (D[i + n * j] + D[j + n * i]);
    auto X = [&](size_t i, size_t j, size_t k, size_t l) { return XX[i + n * (j + n * (k + n * l))]; };
    std::vector<int> rank(s.pairs.size());
    for (int c = 0; c < 3; ++c)
        for (size_t r = 0; r < s.ids[c].size(); ++r) rank[s.ids[c][r]] = (int)r;
    for (int R = 0; R < (int)s.pairs.size(); ++R) {
        const ShellPair &p = s.pairs[R];
        const int cr = pclass(p);
        for (int cs = 0; cs < 3; ++cs)
            for (int S : s.ids[cs]) {
                const ShellPair &q = s.pairs[S];
                const bool bra = cr > cs || (cr == cs && rank[R] <= rank[S]);
                for (int o = 0; o < 2; ++o) {
                    if (o == 1 && (p.same || !bra)) continue;   // reverse copies: bra role, distinct shells
                    const double wj = !bra ? (p.same ? 0.5 : 1.0) : (o == 0 ? (p.same ? 0.5 : 1.0) : 0.0);
                    const double wk = !bra ? 0.0 : (R == S ? 0.5 : 1.0);
                    const int sx = o == 0 ? p.ao_a : p.ao_b, sy = o == 0 ? p.ao_b : p.ao_a;
                    const int nx = (o == 0 ? p.la : p.lb) ? 3 : 1, ny = (o == 0 ? p.lb : p.la) ? 3 : 1;
                    const int nc = q.la ? 3 : 1, nd = q.lb ? 3 : 1;
                    const double fj = q.same ? 1.0 : 2.0;
                    for (int a = 0; a < nx; ++a)
                        for (int b = 0; b < ny; ++b)
                            for (int c = 0; c < nc; ++c)
                                for (int d = 0; d < nd; ++d) {
                                    const size_t i = sx + a, j = sy + b, k = q.ao_a + c, l = q.ao_b + d;
                                    const double v = X(i, j, k, l);
                                    A[i + n * j] += wj * fj * Ds[k + n * l] * v;
                                    A[i + n * k] -= wk * 0.5 * Ds[j + n * l] * v;
                                    if (!q.same) A[i + n * l] -= wk * 0.5 * Ds[j + n * k] * v;
                                }
                }
            }
    }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) G[i + n * j] = A[i + n * j] + A[j + n * i];
    return 0;
}

// The one-copy digestion of jk.hip over the layout and schedule of
// digest_plan.h, restated on the host: the stored values (one pre-weighted
// copy per unique screened quartet, at the plan's offsets), then every
// workgroup's column chunks and units with the kernel's six updates, the
// ket-side partial slots, and the two gathers.  Every stored element must be
// read exactly once (returns -2 otherwise).  G = A + A^T.
#include "../../myqc_amd/csrc/digest_plan.h"

namespace {

template <int LA, int LB>
double schwarz_pair(const Sys &s, int p) {
    QuartetAcc<LA, LB, LA, LB> acc;
    const ShellPair &sp = s.pairs[p];
    quartet_prims<LA, LB, LA, LB>(sp, sp, s.prims.data(), s.prims.data(), s.ftT.data(), acc);
    constexpr int na = ncart(LA), nb = ncart(LB);
    double out[na * nb * na * nb];
    quartet_hrr<LA, LB, LA, LB>(acc, sp.AB, sp.AB, out);
    double m = 0.0;
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) m = std::max(m, fabs(out[((a * nb + b) * na + a) * nb + b]));
    return sqrt(m);
}

template <typename Fn>
void block_of(const Sys &s, int pb, int pk, Fn &&fn) {
    const int cb = pclass(s.pairs[pb]), ck = pclass(s.pairs[pk]);
    if (cb == 0 && ck == 0) one<0, 0, 0, 0>(s, pb, pk, fn);
    else if (cb == 1 && ck == 0) one<1, 0, 0, 0>(s, pb, pk, fn);
    else if (cb == 1 && ck == 1) one<1, 0, 1, 0>(s, pb, pk, fn);
    else if (cb == 2 && ck == 0) one<1, 1, 0, 0>(s, pb, pk, fn);
    else if (cb == 2 && ck == 1) one<1, 1, 1, 0>(s, pb, pk, fn);
    else if (cb == 2 && ck == 2) one<1, 1, 1, 1>(s, pb, pk, fn);
}

}  // namespace

extern "C" int hc_fock_G_plan(int nnuc, const double *xyz, int nset, const double *alpha, const int *setinfo,
                              const double *bas, int norb, const int *basinfo, const double *Ftab, const double *D,
                              double tol, int max_rows, long long target, double *G, long long *stats) {
    Sys s;
    if (!make(s, nnuc, xyz, nset, alpha, setinfo, bas, norb, basinfo, Ftab).empty()) return -1;
    const int n = norb, np = (int)s.pairs.size(), ns = (int)s.shells.size();
    std::vector<double> q(np);
    for (int p = 0; p < np; ++p) {
        const int c = pclass(s.pairs[p]);
        q[p] = c == 0 ? schwarz_pair<0, 0>(s, p) : c == 1 ? schwarz_pair<1, 0>(s, p) : schwarz_pair<1, 1>(s, p);
    }
    std::vector<int> ids[3];
    for (int c = 0; c < 3; ++c) {
        ids[c] = s.ids[c];
        std::vector<std::pair<double, int>> key;
        for (int p : ids[c]) key.push_back({-q[p], p});
        std::sort(key.begin(), key.end());
        for (size_t r = 0; r < key.size(); ++r) ids[c][r] = key[r].second;
    }
    std::vector<int> sao(ns), sL(ns);
    for (int X = 0; X < ns; ++X) {
        sao[X] = s.shells[X].ao;
        sL[X] = s.shells[X].L;
    }
    DigestPlan pl;
    if (!build_digest_plan(pl, n, sao, sL, s.pairs, ids, q, tol, max_rows, target).empty()) return -1;
    // stored values
    std::vector<double> x(std::max<long long>(pl.n_values, 1), 0.0);
    std::vector<int> seen(x.size(), 0);
    for (int p = 0; p < np; ++p) {
        const ShellPair &bp = s.pairs[p];
        const int nb = bp.lb ? 3 : 1;
        for (int ck = 0; ck < 3; ++ck)
            for (int r = pl.lo[(size_t)p * 3 + ck]; r < pl.hi[(size_t)p * 3 + ck]; ++r) {
                const int k = ids[ck][r];
                const ShellPair &kp = s.pairs[k];
                const int nc = kp.la ? 3 : 1, nd = kp.lb ? 3 : 1;
                const long long base = pl.off[(size_t)p * 3 + ck] + r;
                const long long len = pl.hi[(size_t)p * 3 + ck] - pl.lo[(size_t)p * 3 + ck];
                block_of(s, p, k, [&](int i, int j, int kk, int l, double v, double w) {
                    const int a = i - bp.ao_a, b = j - bp.ao_b, c = kk - kp.ao_a, d = l - kp.ao_b;
                    x[base + (((a * nb + b) * nc + c) * nd + d) * len] = v * w;
                });
            }
    }
    // the symmetrised density and the workgroups
    std::vector<double> Ds((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Ds[i + (size_t)n * j] = 0.5 * (D[i + (size_t)n * j] + D[j + (size_t)n * i]);
    std::vector<double> partial((size_t)pl.n_rows * n, 0.0), jkp((size_t)std::max(pl.n_jk_slots, 1) * kJkSlot, 0.0);
    const int nwg = (int)(pl.wg.size() / kWgInts);
    for (int w = 0; w < nwg; ++w) {
        const int *wr = &pl.wg[(size_t)w * kWgInts];
        const int aoA = wr[0] & 0x7fff, LA = wr[0] >> 15, u0 = wr[1], u1 = wr[2], c0 = wr[3], c1 = wr[4];
        const int nrows = wr[5], row0 = wr[6];
        const int na = LA ? 3 : 1;
        if (nrows > pl.max_rows) return -3;
        std::vector<double> acc((size_t)nrows * n, 0.0), drow((size_t)nrows * n);
        for (int rr = 0; rr < nrows; ++rr)
            for (int j = 0; j < n; ++j) drow[(size_t)rr * n + j] = Ds[j + (size_t)n * pl.row_ao[row0 + rr]];
        for (int c = c0; c < c1; ++c) {
            const int ck = pl.chunk[(size_t)c * 4], r0 = pl.chunk[(size_t)c * 4 + 1], nk = pl.chunk[(size_t)c * 4 + 2];
            const int slot = pl.chunk[(size_t)c * 4 + 3], ncd = dig_ncd(ck);
            if (nk > dig_width(ck)) return -3;
            for (int u = u0; u < u1; ++u) {
                const int *ur = &pl.unit[(size_t)u * kUnitInts];
                const int lo = ur[3 + 2 * ck], hi = ur[4 + 2 * ck];
                const int aoB = ur[1] & 0x7fff, LB = ur[1] >> 15, sb = ur[2];
                const int nb = LB ? 3 : 1;
                const long long off = pl.unit_off[(size_t)u * 3 + ck];
                for (int kk = 0; kk < nk; ++kk) {
                    const int r = r0 + kk;
                    if (r < lo || r >= hi) continue;
                    const ShellPair &kp = s.pairs[ids[ck][r]];
                    const int nc = kp.la ? 3 : 1, nd = kp.lb ? 3 : 1, aoC = kp.ao_a, aoD = kp.ao_b;
                    for (int a = 0; a < na; ++a)
                        for (int b = 0; b < nb; ++b)
                            for (int cc = 0; cc < nc; ++cc)
                                for (int d = 0; d < nd; ++d) {
                                    const long long e = off + r + (long long)(((a * nb + b) * nc + cc) * nd + d) * (hi - lo);
                                    if (e < 0 || e >= (long long)x.size()) return -3;
                                    ++seen[e];
                                    const double v = x[e];
                                    auto DA = [&](int row, int col) { return drow[(size_t)row * n + col]; };
                                    acc[(size_t)a * n + aoB + b] += 2.0 * Ds[(aoC + cc) + (size_t)n * (aoD + d)] * v;
                                    jkp[(size_t)slot * kJkSlot + (size_t)kk * ncd + cc * nd + d] +=
                                        2.0 * DA(a, aoB + b) * v;
                                    acc[(size_t)a * n + aoC + cc] -= 0.5 * DA(sb + b, aoD + d) * v;
                                    acc[(size_t)a * n + aoD + d] -= 0.5 * DA(sb + b, aoC + cc) * v;
                                    acc[(size_t)(sb + b) * n + aoC + cc] -= 0.5 * DA(a, aoD + d) * v;
                                    acc[(size_t)(sb + b) * n + aoD + d] -= 0.5 * DA(a, aoC + cc) * v;
                                }
                }
            }
        }
        for (int rr = 0; rr < nrows; ++rr)
            for (int j = 0; j < n; ++j) partial[(size_t)(row0 + rr) * n + j] = acc[(size_t)rr * n + j];
        (void)aoA;
    }
    // gathers, by the kernels' groups (each group's sum added in)
    const int np_c[3] = {(int)ids[0].size(), (int)ids[1].size(), (int)ids[2].size()};
    std::vector<double> jk(std::max(pl.jk_base[3], 1), 0.0);
    std::vector<int> jk_cover(pl.jk_list.size(), 0);
    for (size_t g = 0; g < pl.jk_group.size() / 3; ++g) {
        const int b = pl.jk_group[3 * g], k0 = pl.jk_group[3 * g + 1], k1 = pl.jk_group[3 * g + 2];
        if (k1 - k0 > kJkGroup || k0 < pl.jk_list_off[b] || k1 > pl.jk_list_off[b + 1]) return -4;
        const int c = b >= pl.jk_block0[2] ? 2 : b >= pl.jk_block0[1] ? 1 : 0;
        const int ncd = dig_ncd(c), W = dig_width(c), r0 = (b - pl.jk_block0[c]) * W;
        for (int pos = 0; pos < W * ncd && r0 + pos / ncd < np_c[c]; ++pos) {
            double sum = 0.0;
            for (int k = k0; k < k1; ++k) sum += jkp[(size_t)pl.jk_list[k] * kJkSlot + pos];
            jk[pl.jk_base[c] + (size_t)r0 * ncd + pos] += sum;
        }
        for (int k = k0; k < k1; ++k) ++jk_cover[k];
    }
    for (int v : jk_cover)
        if (v != 1) return -4;
    std::vector<int> ao_shell(n);
    for (int X = 0; X < ns; ++X)
        for (int k = 0; k < (sL[X] ? 3 : 1); ++k) ao_shell[sao[X] + k] = X;
    std::vector<double> A((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double sum = 0.0;
            for (int k = pl.row_list_off[i]; k < pl.row_list_off[i + 1]; ++k) sum += partial[(size_t)pl.row_list[k] * n + j];
            const int X = ao_shell[i], Y = ao_shell[j];
            const int o = pl.jkoff[(size_t)X * ns + Y];
            if (o >= 0) sum += jk[o + (i - sao[X]) * (sL[Y] ? 3 : 1) + (j - sao[Y])];
            A[i + (size_t)n * j] = sum;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) G[i + (size_t)n * j] = A[i + (size_t)n * j] + A[j + (size_t)n * i];
    if (stats) {
        stats[0] = pl.n_values;
        stats[1] = pl.n_quartets;
        stats[2] = nwg;
        stats[3] = (long long)(pl.chunk.size() / 4);
        stats[4] = pl.n_rows;
        stats[5] = pl.n_jk_slots;
    }
    for (size_t e = 0; e < seen.size() && pl.n_values > 0; ++e)
        if (seen[e] != 1) return -2;
    return 0;
}

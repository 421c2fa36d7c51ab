// system.h -- host-side construction of contracted shells, shell pairs and
// primitive pairs from the reference's basis arrays (no HIP API calls, so the
// same code serves the library and the host-side check).
//
// Input layouts are buildBasis's (src/myQC/basis.f90:23-226):
//   setinfo = [nset, setl, {norb_in_set, maxL, centre, orb0..}...]
//   basinfo = [OpS, norb, {pri, ang, ori, nprim, centre}...]
//   bas(set*OpS + k): coefficient of the k-th orbital of a set.
// A shell is an s function, or the three p functions of an sp set group; the
// reference's 2s and 2p share exponents but not coefficients, so they are
// separate shells here, each remembering its sets' max L (bL), which fixes
// the Boys order Q = 3*(sum of bL) of every quartet (int2e.f90:196-245,654-656).
#pragma once

#include <algorithm>
#include <cmath>
#include <map>
#include <thread>
#include <string>
#include <vector>

#include "eri_math.h"

namespace myqc {

struct Shell {
    int center, L, bL, ao;
    std::vector<double> alpha, coef;   // coef includes the primitive normalisation
};

// primitive normalisation (auxilary.f90:704-726) with the single-precision pi
inline double gto_norm(int l, double a) {
    if (l == 0) return pow(2.0 * a / PI_SP, 3.0 / 4.0);
    return pow(128.0 * pow(a, 5.0) / pow(PI_SP, 3.0), 1.0 / 4.0);
}

// Returns an empty string on success, else the reason.
inline std::string build_system(int nnuc, const double *xyz /*[atom][3]*/, int nset, const double *set_alpha,
                                const int *setinfo, const double *bas, int norb, const int *basinfo,
                                std::vector<Shell> &shells, std::vector<ShellPair> &pairs,
                                std::vector<PrimPair> &prims, std::vector<int> pair_ids[3]) {
    char buf[256];
    if (setinfo[0] != nset || basinfo[1] != norb) return "setinfo/basinfo headers disagree with nset/norb";
    const int setl = setinfo[1], OpS = basinfo[0];
    if (setl != 3 + OpS || OpS < 1 || OpS > 8) return "bad set length / orbitals per set";
    std::map<int, int> shell_of;   // AO of the s function or of the p_x function -> shell
    shells.clear();
    for (int s = 0; s < nset; ++s) {
        const int *si = setinfo + 2 + s * setl;
        const int nos = si[0], maxL = si[1], cen = si[2];
        if (nos < 1 || nos > OpS || cen < 0 || cen >= nnuc || maxL < 0 || maxL > 1) {
            snprintf(buf, sizeof buf, "set %d: unsupported layout (norb %d, maxL %d, centre %d)", s, nos, maxL, cen);
            return buf;
        }
        for (int k = 0; k < nos; ++k) {
            const int o = si[3 + k];
            if (o < 0 || o >= norb) {
                snprintf(buf, sizeof buf, "set %d: orbital %d out of range", s, o);
                return buf;
            }
            const int ang = basinfo[2 + 5 * o + 1], ori = basinfo[2 + 5 * o + 2];
            if (ang > 1) {
                snprintf(buf, sizeof buf, "orbital %d: angular momentum %d unsupported", o, ang);
                return buf;
            }
            if (ang == 1 && ori > 0) continue;   // p_y, p_z belong to the p_x shell
            if (ang == 1 && (o + 2 >= norb || basinfo[2 + 5 * (o + 1) + 2] != 1 || basinfo[2 + 5 * (o + 2) + 2] != 2)) {
                snprintf(buf, sizeof buf, "orbital %d: p functions not ordered x,y,z", o);
                return buf;
            }
            auto it = shell_of.find(o);
            if (it == shell_of.end()) {
                Shell sh;
                sh.center = cen;
                sh.L = ang;
                sh.bL = maxL;
                sh.ao = o;
                it = shell_of.emplace(o, (int)shells.size()).first;
                shells.push_back(sh);
            }
            Shell &sh = shells[it->second];
            if (sh.center != cen || sh.bL != maxL) {
                snprintf(buf, sizeof buf, "orbital %d: sets disagree on centre or max L", o);
                return buf;
            }
            sh.alpha.push_back(set_alpha[s]);
            sh.coef.push_back(bas[s * OpS + k] * gto_norm(ang, set_alpha[s]));
        }
    }
    std::sort(shells.begin(), shells.end(), [](const Shell &x, const Shell &y) { return x.ao < y.ao; });
    for (const Shell &sh : shells)
        if (sh.alpha.empty() || sh.alpha.size() > 6) {
            snprintf(buf, sizeof buf, "shell at AO %d: %zu primitives (1..6 supported)", sh.ao, sh.alpha.size());
            return buf;
        }
    pairs.clear();
    prims.clear();
    for (int c = 0; c < 3; ++c) pair_ids[c].clear();
    const int ns = (int)shells.size();
    // pairs first (cheap), then their primitive pairs in parallel: each pair
    // owns the slice [pp0, pp0 + npp) of prims, which is sized once
    {
        size_t npairs = 0, nprims = 0;
        for (int A = 0; A < ns; ++A)
            for (int B = A; B < ns; ++B) {
                ++npairs;
                nprims += shells[A].alpha.size() * shells[B].alpha.size();
            }
        pairs.reserve(npairs);
        prims.resize(nprims);
    }
    size_t pnext = 0;
    for (int A = 0; A < ns; ++A)
        for (int B = A; B < ns; ++B) {
            int f = A, g = B;
            if (shells[A].L < shells[B].L) std::swap(f, g);   // first shell carries the larger L
            const Shell &sa = shells[f], &sb = shells[g];
            ShellPair sp;
            const double *RA = xyz + 3 * sa.center, *RB = xyz + 3 * sb.center;
            for (int d = 0; d < 3; ++d) sp.AB[d] = RA[d] - RB[d];
            sp.pp0 = (int)pnext;
            sp.ao_a = sa.ao;
            sp.ao_b = sb.ao;
            sp.la = sa.L;
            sp.lb = sb.L;
            sp.bl = sa.bL + sb.bL;
            sp.same = (A == B);
            sp.sa = f;
            sp.sb = g;
            sp.npp = (int)(sa.alpha.size() * sb.alpha.size());
            pnext += sp.npp;
            pair_ids[sa.L + sb.L].push_back((int)pairs.size());
            pairs.push_back(sp);
        }
    auto fill = [&](size_t p0, size_t p1) {
        for (size_t q = p0; q < p1; ++q) {
            const ShellPair &sp = pairs[q];
            const Shell &sa = shells[sp.sa], &sb = shells[sp.sb];
            const double *RA = xyz + 3 * sa.center, *RB = xyz + 3 * sb.center;
            const double ab2 = sp.AB[0] * sp.AB[0] + sp.AB[1] * sp.AB[1] + sp.AB[2] * sp.AB[2];
            PrimPair *out = prims.data() + sp.pp0;
            for (size_t i = 0; i < sa.alpha.size(); ++i)
                for (size_t j = 0; j < sb.alpha.size(); ++j) {
                    const double aa = sa.alpha[i], bb = sb.alpha[j], p = aa + bb;
                    PrimPair &pp = *out++;
                    pp.p = p;
                    for (int d = 0; d < 3; ++d) {
                        pp.P[d] = (aa * RA[d] + bb * RB[d]) / p;
                        pp.PA[d] = pp.P[d] - RA[d];
                    }
                    pp.EIJ = exp(-aa * bb * ab2 / p);       // int2e.f90:223
                    pp.K = pp.EIJ * sa.coef[i] * sb.coef[j];
                    pp.beta = bb;
                    pp.ip = 1.0 / p;
                }
        }
    };
    const size_t np = pairs.size();
    const unsigned nt = (unsigned)std::max<size_t>(1, std::min<size_t>({16, std::thread::hardware_concurrency(),
                                                                         np / 2048 + 1}));
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(fill, np * t / nt, np * (t + 1) / nt);
    fill(0, np / nt);
    for (auto &th : pool) th.join();
    return "";
}

}  // namespace myqc

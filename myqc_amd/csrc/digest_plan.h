// digest_plan.h -- host-side layout of the stored integrals and the J/K
// digestion schedule (no HIP API calls, so the host-side check runs the same
// code without a device).
//
// Storage: each unique shell quartet (P|Q) that passes the Schwarz screen is
// stored ONCE, under its bra pair P (the pair of the higher class, or of the
// lower Schwarz rank within one class), in P's stored orientation (A, B) =
// (sa, sb), at the Schwarz rank of the ket Q within Q's class.  The kets of P
// in class ck form the rank range [lo, hi): all screened kets of a lower
// class (a prefix of the descending Schwarz order), in P's own class those
// from P's rank on.  P's values for class ck are element-major: element e of
// the block of ket rank r at off + r + e (hi - lo), so the lanes of a
// wavefront, which take consecutive ranks, read and write consecutive
// addresses.  The values are pre-weighted by the degeneracy weight of
// each element (block_weight), so that the digestion applies, for every
// stored element x = w (ab|cd) with the symmetrised density D, the six
// updates of
//     G = A + A^T,  A(a,b) += 2 D(c,d) x,  A(c,d) += 2 D(a,b) x,
//     A(a,c) -= D(b,d) x / 2,  A(a,d) -= D(b,c) x / 2,
//     A(b,c) -= D(a,d) x / 2,  A(b,d) -= D(a,c) x / 2
// (the G of RHFI2G, src/I2G/RHFI2G.f90:80-90, over the unique integrals).
//
// Schedule: a workgroup owns the rows of G of one row shell A and of the B
// shells of a run of its pairs (A, B) ("units"), at most max_rows AOs in all,
// so that every update but A(c,d) lands in rows it holds in LDS.  Its kets are
// walked in column chunks: one ket class and W(ck) consecutive ranks (a
// wavefront's lanes take KPL(ck) kets each); for each chunk the wavefront
// visits every unit whose rank range meets it, and the ket-side update
// A(c,d) is summed in registers over those units and written once per chunk
// to a partial slot.  A gather then sums, per AO row, the partial rows of the
// workgroups that hold it, and per ket, the partial slots of the chunks that
// cover it.  No global atomics; the summation order is fixed up to the LDS
// atomics inside a workgroup.
//
// Balance: a wavefront's time goes mostly to its rows (one load round trip
// and one set of LDS updates each, whatever the number of active lanes), so
// a chunk costs about the loads per lane of its rows.  A workgroup's chunks
// are listed by descending cost and handed to its wavefronts as they become
// free (longest first); the workgroups are listed by descending cost, so
// that the long ones start first and the short ones fill the tail.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "eri_math.h"

namespace myqc {

// kets per lane and per column chunk of each ket class
MYQC_HD constexpr int dig_kpl(int ck) { return ck == 0 ? 4 : ck == 1 ? 2 : 1; }
MYQC_HD constexpr int dig_width(int ck) { return 64 * dig_kpl(ck); }
MYQC_HD constexpr int dig_ncd(int ck) { return ck == 0 ? 1 : ck == 1 ? 3 : 9; }
constexpr int kJkSlot = 576;    // doubles per ket-side partial slot (the largest W(ck) ncd(ck))

// unit record, ints: pair, aoB | LB << 15, LDS row of B's first AO, lo0, hi0,
// lo1, hi1, lo2, hi2, Schwarz factor of the pair (float bits, rounded up)
constexpr int kUnitInts = 10;
// workgroup record, ints: aoA | LA << 15, first unit, end unit, first chunk,
// end chunk, rows, first partial row
constexpr int kWgInts = 7;
// chunk record, ints: class, first rank, kets, ket-side partial slot
constexpr int kChunkInts = 4;
// the ket-side gather's groups: at most this many partial slots of one rank
// block are summed by one workgroup
constexpr int kJkGroup = 32;

struct DigestPlan {
    int n = 0, ns = 0, np = 0;
    int max_rows = 0;                    // LDS rows of one workgroup at most
    std::vector<int> rank, pcls;         // per pair: rank within its class, class
    std::vector<int> lo, hi;             // [p*3+ck] ket ranks of which p is the bra
    std::vector<long long> off;          // [p*3+ck] element e of ket rank r at off + r + e*(hi - lo)
    long long n_values = 0;              // stored (pre-weighted) values
    long long n_quartets = 0;
    std::vector<int> unit, wg, chunk;    // records as described above
    std::vector<long long> unit_off;     // [u*3+ck] = off of the unit's pair
    std::vector<int> row_ao;             // [partial row] = AO held by that row
    std::vector<int> row_list_off, row_list;   // per AO: the partial rows holding it (CSR)
    int n_rows = 0;                      // partial rows in all
    int n_jk_slots = 0;
    std::vector<int> jk_list_off, jk_list;     // per (class, block of W ranks): chunk slots (CSR)
    std::vector<int> jk_block0;          // first block index of each class (3 + 1 entries)
    std::vector<int> jk_base;            // first J-ket value of each class (3 + 1 entries)
    std::vector<int> jkoff;              // [X*ns + Y]: J-ket block of pair (X,Y) if stored so, else -1
    std::vector<int> jk_group;           // {rank block, first, end} into jk_list, at most kJkGroup slots
};

inline int pair_nf(const ShellPair &p) { return (p.la ? 3 : 1) * (p.lb ? 3 : 1); }

// the pairs of each row shell A = sa, by ascending AO of B = sb
inline std::vector<std::vector<int>> pairs_by_row_shell(int ns, const std::vector<int> &shells_ao,
                                                         const std::vector<ShellPair> &pairs) {
    std::vector<std::vector<int>> of_a(ns);
    for (int p = 0; p < (int)pairs.size(); ++p) of_a[pairs[p].sa].push_back(p);
    for (auto &v : of_a)
        std::sort(v.begin(), v.end(), [&](int x, int y) { return shells_ao[pairs[x].sb] < shells_ao[pairs[y].sb]; });
    return of_a;
}

// The layout (ranks, ket ranges, value offsets).  pair_ids: per class, pair
// ids in descending Schwarz order; schwarz: per pair; shells_ao: per shell;
// tol: the Schwarz threshold.
inline std::string build_digest_layout(DigestPlan &pl, int n, const std::vector<int> &shells_ao,
                                       const std::vector<ShellPair> &pairs, const std::vector<int> pair_ids[3],
                                       const std::vector<double> &schwarz, double tol) {
    const int ns = (int)shells_ao.size(), np = (int)pairs.size();
    pl = DigestPlan();
    pl.n = n;
    pl.ns = ns;
    pl.np = np;
    pl.rank.assign(np, 0);
    pl.pcls.assign(np, 0);
    std::vector<double> qs[3];
    for (int c = 0; c < 3; ++c) {
        qs[c].resize(pair_ids[c].size());
        for (size_t r = 0; r < pair_ids[c].size(); ++r) {
            pl.rank[pair_ids[c][r]] = (int)r;
            pl.pcls[pair_ids[c][r]] = c;
            qs[c][r] = schwarz[pair_ids[c][r]];
        }
    }
    // screened prefix of each class for every pair: kets with q_P q_Q >= tol
    pl.lo.assign((size_t)np * 3, 0);
    pl.hi.assign((size_t)np * 3, 0);
    for (int cb = 0; cb < 3; ++cb)
        for (int ck = 0; ck < 3; ++ck) {
            const auto &qk = qs[ck];
            size_t ptr = qk.size();
            for (size_t r = 0; r < pair_ids[cb].size(); ++r) {
                const int p = pair_ids[cb][r];
                const double qp = schwarz[p];
                long long cnt;
                if (tol <= 0.0) cnt = (long long)qk.size();
                else if (qp <= 0.0) cnt = 0;
                else {
                    const double thr = tol / qp;
                    while (ptr > 0 && !(qk[ptr - 1] >= thr)) --ptr;
                    cnt = (long long)ptr;
                }
                int l = 0, h = (int)cnt;
                if (ck > cb) h = 0;                    // p is never the bra of a higher class
                else if (ck == cb) l = (int)r;         // in its own class, from its own rank on
                if (h < l) h = l;
                pl.lo[(size_t)p * 3 + ck] = l;
                pl.hi[(size_t)p * 3 + ck] = h;
            }
        }
    // value offsets, pairs grouped by row shell
    const auto of_a = pairs_by_row_shell(ns, shells_ao, pairs);
    pl.off.assign((size_t)np * 3, 0);
    long long off = 0;
    for (int A = 0; A < ns; ++A)
        for (int p : of_a[A]) {
            const int nf = pair_nf(pairs[p]);
            for (int ck = 0; ck < 3; ++ck) {
                const int l = pl.lo[(size_t)p * 3 + ck], h = pl.hi[(size_t)p * 3 + ck];
                const int blk = nf * dig_ncd(ck);
                pl.off[(size_t)p * 3 + ck] = off - l;
                off += (long long)(h - l) * blk;
                pl.n_quartets += h - l;
            }
        }
    pl.n_values = off;
    return "";
}

// The schedule over a layout: max_rows: LDS rows a workgroup may hold;
// target: stored values per workgroup (a run of units is split over several
// workgroups when longer); shells_L: per shell.
inline std::string build_digest_schedule(DigestPlan &pl, const std::vector<int> &shells_ao,
                                         const std::vector<int> &shells_L, const std::vector<ShellPair> &pairs,
                                         const std::vector<int> pair_ids[3], const std::vector<double> &schwarz,
                                         int max_rows, long long target) {
    const int ns = pl.ns, np = pl.np, n = pl.n;
    pl.max_rows = max_rows;
    if (max_rows < 6) return "digestion: too few LDS rows for one p-shell row pair";
    const auto of_a = pairs_by_row_shell(ns, shells_ao, pairs);
    // J-ket layout and the pair -> J-ket block map
    pl.jk_base.assign(4, 0);
    for (int c = 0; c < 3; ++c) pl.jk_base[c + 1] = pl.jk_base[c] + (int)pair_ids[c].size() * dig_ncd(c);
    pl.jk_block0.assign(4, 0);
    for (int c = 0; c < 3; ++c)
        pl.jk_block0[c + 1] = pl.jk_block0[c] + ((int)pair_ids[c].size() + dig_width(c) - 1) / dig_width(c);
    pl.jkoff.assign((size_t)ns * ns, -1);
    for (int p = 0; p < np; ++p)
        pl.jkoff[(size_t)pairs[p].sa * ns + pairs[p].sb] =
            pl.jk_base[pl.pcls[p]] + pl.rank[p] * dig_ncd(pl.pcls[p]);
    std::vector<std::vector<int>> jk_slots(pl.jk_block0[3]);
    // work vectors, reused from run to run
    std::vector<int> rowaos, chunks;   // chunks: (ck, r0, nk) triples
    std::vector<long long> cost, bc, diff, walk, wg_walk;
    std::vector<size_t> order;
    // workgroups
    for (int A = 0; A < ns; ++A) {
        const auto &units = of_a[A];
        const int nA = shells_L[A] ? 3 : 1;
        size_t u = 0;
        while (u < units.size()) {
            // a run of units whose distinct B rows fit beside A's
            size_t u1 = u;
            int rows = nA;
            while (u1 < units.size()) {
                const int B = pairs[units[u1]].sb;
                const int add = (B == A) ? 0 : (shells_L[B] ? 3 : 1);
                if (rows + add > max_rows) break;
                rows += add;
                ++u1;
            }
            if (u1 == u) return "digestion: a unit does not fit the LDS rows";
            // unit records and their LDS rows: A's rows first, then each B's
            const int ubase = (int)(pl.unit.size() / kUnitInts);
            rowaos.clear();
            for (int k = 0; k < nA; ++k) rowaos.push_back(shells_ao[A] + k);
            for (size_t k = u; k < u1; ++k) {
                const int p = units[k], B = pairs[p].sb;
                int slot = 0;
                if (B != A) {
                    slot = (int)rowaos.size();
                    for (int m = 0; m < (shells_L[B] ? 3 : 1); ++m) rowaos.push_back(shells_ao[B] + m);
                }
                float qf = (float)schwarz[p];
                if ((double)qf < schwarz[p]) qf = std::nextafter(qf, 3.0e38f);
                int qbits;
                memcpy(&qbits, &qf, 4);
                const int rec[kUnitInts] = {p,
                                            shells_ao[B] | (shells_L[B] << 15),
                                            slot,
                                            pl.lo[(size_t)p * 3 + 0],
                                            pl.hi[(size_t)p * 3 + 0],
                                            pl.lo[(size_t)p * 3 + 1],
                                            pl.hi[(size_t)p * 3 + 1],
                                            pl.lo[(size_t)p * 3 + 2],
                                            pl.hi[(size_t)p * 3 + 2],
                                            qbits};
                pl.unit.insert(pl.unit.end(), rec, rec + kUnitInts);
                for (int ck = 0; ck < 3; ++ck) pl.unit_off.push_back(pl.off[(size_t)p * 3 + ck]);
            }
            const int uend = (int)(pl.unit.size() / kUnitInts);
            // column chunks: aligned blocks of W(ck) ranks that meet some
            // unit's range, with the values each holds (the ends of a unit's
            // range add to their blocks, the blocks between through a
            // difference array)
            chunks.clear();
            cost.clear();
            for (int ck = 0; ck < 3; ++ck) {
                const int W = dig_width(ck);
                int lmin = 1 << 30, hmax = 0;
                for (int k = ubase; k < uend; ++k) {
                    const int l = pl.unit[(size_t)k * kUnitInts + 3 + 2 * ck], h = pl.unit[(size_t)k * kUnitInts + 4 + 2 * ck];
                    if (h > l) {
                        lmin = std::min(lmin, l);
                        hmax = std::max(hmax, h);
                    }
                }
                if (hmax == 0) continue;
                const int B0 = lmin / W, nblk = (hmax - 1) / W - B0 + 1;
                bc.assign(nblk, 0);
                diff.assign(nblk + 1, 0);
                for (int k = ubase; k < uend; ++k) {
                    const int l = pl.unit[(size_t)k * kUnitInts + 3 + 2 * ck], h = pl.unit[(size_t)k * kUnitInts + 4 + 2 * ck];
                    if (h <= l) continue;
                    const long long w = (long long)pair_nf(pairs[pl.unit[(size_t)k * kUnitInts]]) * dig_ncd(ck);
                    const int bl = l / W, bh = (h - 1) / W;
                    if (bl == bh) {
                        bc[bl - B0] += (h - l) * w;
                        continue;
                    }
                    bc[bl - B0] += ((long long)(bl + 1) * W - l) * w;
                    bc[bh - B0] += ((long long)h - (long long)bh * W) * w;
                    diff[bl + 1 - B0] += W * w;
                    diff[bh - B0] -= W * w;
                }
                long long run = 0;
                for (int b = 0; b < nblk; ++b) {
                    run += diff[b];
                    const long long c = bc[b] + run;
                    if (c == 0) continue;
                    const int b0 = (B0 + b) * W, b1 = std::min(b0 + W, (int)pair_ids[ck].size());
                    chunks.push_back(ck);
                    chunks.push_back(b0);
                    chunks.push_back(b1 - b0);
                    cost.push_back(c);
                }
            }
            // each chunk's walk cost: loads per lane over the rows of every
            // unit that meets it
            walk.assign(cost.size(), 0);
            for (size_t k = 0; k < cost.size(); ++k) {
                const int ck = chunks[3 * k], r0 = chunks[3 * k + 1], nk = chunks[3 * k + 2];
                for (int q = ubase; q < uend; ++q) {
                    const int *ur = &pl.unit[(size_t)q * kUnitInts];
                    if (ur[4 + 2 * ck] <= r0 || ur[3 + 2 * ck] >= r0 + nk) continue;
                    walk[k] += (long long)nA * ((ur[1] >> 15) ? 3 : 1) * dig_ncd(ck) * dig_kpl(ck);
                }
            }
            // split the chunks into workgroups of about `target` values
            size_t c0 = 0;
            const size_t nch = cost.size();
            while (c0 < nch) {
                size_t c1 = c0;
                long long acc = 0, wsum = 0;
                while (c1 < nch && (c1 == c0 || acc + cost[c1] <= target)) acc += cost[c1++];
                order.clear();
                for (size_t k = c0; k < c1; ++k) {
                    order.push_back(k);
                    wsum += walk[k];
                }
                std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return walk[x] > walk[y]; });
                const int chunk0 = (int)(pl.chunk.size() / kChunkInts);
                for (const size_t k : order) {
                    const int ck = chunks[3 * k], r0 = chunks[3 * k + 1], nk = chunks[3 * k + 2];
                    const int slot = pl.n_jk_slots++;
                    pl.chunk.push_back(ck);
                    pl.chunk.push_back(r0);
                    pl.chunk.push_back(nk);
                    pl.chunk.push_back(slot);
                    jk_slots[pl.jk_block0[ck] + r0 / dig_width(ck)].push_back(slot);
                }
                const int rec[kWgInts] = {shells_ao[A] | (shells_L[A] << 15), ubase, uend, chunk0,
                                          (int)(pl.chunk.size() / kChunkInts), (int)rowaos.size(), pl.n_rows};
                pl.wg.insert(pl.wg.end(), rec, rec + kWgInts);
                wg_walk.push_back(wsum);
                pl.row_ao.insert(pl.row_ao.end(), rowaos.begin(), rowaos.end());
                pl.n_rows += (int)rowaos.size();
                c0 = c1;
            }
            u = u1;
        }
    }
    // workgroups by descending walk cost (each record carries its own
    // units, chunks and partial rows)
    {
        const size_t nwg = wg_walk.size();
        std::vector<size_t> perm(nwg);
        for (size_t w = 0; w < nwg; ++w) perm[w] = w;
        std::stable_sort(perm.begin(), perm.end(), [&](size_t x, size_t y) { return wg_walk[x] > wg_walk[y]; });
        std::vector<int> wg(pl.wg.size());
        for (size_t w = 0; w < nwg; ++w)
            std::copy(pl.wg.begin() + perm[w] * kWgInts, pl.wg.begin() + (perm[w] + 1) * kWgInts,
                      wg.begin() + w * kWgInts);
        pl.wg.swap(wg);
    }
    // per AO, the partial rows that hold it
    std::vector<std::vector<int>> rows_of(n);
    for (int r = 0; r < pl.n_rows; ++r) rows_of[pl.row_ao[r]].push_back(r);
    pl.row_list_off.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
        pl.row_list_off[i + 1] = pl.row_list_off[i] + (int)rows_of[i].size();
        pl.row_list.insert(pl.row_list.end(), rows_of[i].begin(), rows_of[i].end());
    }
    pl.jk_list_off.assign(jk_slots.size() + 1, 0);
    for (size_t b = 0; b < jk_slots.size(); ++b) {
        pl.jk_list_off[b + 1] = pl.jk_list_off[b] + (int)jk_slots[b].size();
        pl.jk_list.insert(pl.jk_list.end(), jk_slots[b].begin(), jk_slots[b].end());
    }
    // the ket-side gather's work split: runs of at most kJkGroup slots of
    // one rank block
    pl.jk_group.clear();
    for (size_t b = 0; b < jk_slots.size(); ++b)
        for (int k = pl.jk_list_off[b]; k < pl.jk_list_off[b + 1]; k += kJkGroup)
            pl.jk_group.insert(pl.jk_group.end(), {(int)b, k, std::min(k + kJkGroup, pl.jk_list_off[b + 1])});
    return "";
}

inline std::string build_digest_plan(DigestPlan &pl, int n, const std::vector<int> &shells_ao,
                                     const std::vector<int> &shells_L, const std::vector<ShellPair> &pairs,
                                     const std::vector<int> pair_ids[3], const std::vector<double> &schwarz,
                                     double tol, int max_rows, long long target) {
    std::string e = build_digest_layout(pl, n, shells_ao, pairs, pair_ids, schwarz, tol);
    if (!e.empty()) return e;
    return build_digest_schedule(pl, shells_ao, shells_L, pairs, pair_ids, schwarz, max_rows, target);
}

}  // namespace myqc

// eri.hip -- system setup (shells, shell pairs, primitive pairs), Schwarz
// factors, and the contracted shell-quartet ERI kernels.
//
// Work decomposition: one thread per unique contracted shell quartet of one
// angular class; the thread loops over the quartet's primitive quartets (up
// to 3x3x3x3 for STO-3G), builds [e0|f0] by OS vertical recursion in
// registers, contracts, then applies the HGP horizontal recursion.  Unique
// quartets are enumerated implicitly from Schwarz-sorted pair lists, so no
// quartet list is materialised: thread t finds its bra rank by binary search
// over per-bra prefix counts, and its ket rank follows.
//
// Reference: src/integrals/int2e.f90:78-351 (proc2e), :618-726 (clmnew),
// :540-613 (fillsym); src/integrals/auxilary.f90 (Boys, gtoD, getDk).
#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "context.h"
#include "ket_soa.h"
#include "system.h"

namespace myqc {

int set_error(myqc_ctx *ctx, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

// Device allocations keep their capacity: a buffer is only replaced when a
// request outgrows it, so repeated solves of same-sized systems reuse their
// memory instead of paying hipFree/hipMalloc (and the 1-2 GB integral store)
// every time.  The capacities are kept per address in one process-wide
// registry (contexts may run on several host threads), and dfree drops the
// entry of the buffer it frees, so an address the allocator hands out again
// never inherits a stale size.
namespace {
std::mutex g_cap_mutex;
std::unordered_map<void *, size_t> g_cap;
}  // namespace

static size_t cap_lookup(void *p) {
    std::lock_guard<std::mutex> lk(g_cap_mutex);
    auto it = g_cap.find(p);
    return it == g_cap.end() ? 0 : it->second;
}
static void cap_set(void *p, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_cap_mutex);
    if (bytes) g_cap[p] = bytes;
    else g_cap.erase(p);
}

int dalloc(myqc_ctx *ctx, void **p, size_t bytes) {
    if (*p) {
        if (bytes > 0 && cap_lookup(*p) >= bytes) return 0;
        dfree(p);
    }
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(ctx, MYQC_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    cap_set(*p, bytes);
    return 0;
}

size_t dcapacity(const void *p) { return p ? cap_lookup(const_cast<void *>(p)) : 0; }

// Small device-to-host reads (per-iteration scalars, status words) land in
// the context's pinned staging buffer first; larger ones go straight to the
// caller's memory.  Complete at return either way.
hipError_t d2h(myqc_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (ctx->h_pin && bytes <= kPinBytes) {
        hipError_t e = hipMemcpyAsync(ctx->h_pin, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess) memcpy(dst, ctx->h_pin, bytes);
        return e;
    }
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->stream);
}

void dfree(void **p) {
    if (*p) {
        cap_set(*p, 0);
        (void)hipFree(*p);
    }
    *p = nullptr;
}

int timer_begin(myqc_ctx *ctx, int cat, hipEvent_t *ev) {
    *ev = nullptr;
    if (!ctx->timing) return 0;
    MYQC_HIP(ctx, hipEventCreate(ev));
    MYQC_HIP(ctx, hipEventRecord(*ev, ctx->stream));
    return 0;
}

int timer_end(myqc_ctx *ctx, int cat, hipEvent_t ev, double work) {
    if (!ctx->timing || !ev) return 0;
    hipEvent_t b;
    MYQC_HIP(ctx, hipEventCreate(&b));
    MYQC_HIP(ctx, hipEventRecord(b, ctx->stream));
    ctx->events.push_back(EventRec{cat, ev, b, work});
    return 0;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <int C> struct PairL;
template <> struct PairL<0> { static constexpr int a = 0, b = 0; };
template <> struct PairL<1> { static constexpr int a = 1, b = 0; };
template <> struct PairL<2> { static constexpr int a = 1, b = 1; };

struct EriArgs {
    const ShellPair *pairs;
    const PrimPair *prims;
    const double *ksoa;        // ket primitives of the ket class (k_ket_soa)
    long long nket;            // pairs in the ket class
    const double *ftT;
    const int *sorted_b;
    const int *sorted_k;
    const long long *cum;
    int nbra;
    long long nq;
    int same_class;
    // MODE 0: expanded copies
    const long long *segoff;   // [(X*ns + Y)*3 + ket class]
    const int *rank;           // rank of a pair within its class's sorted list
    int ns;
    double *x64;
    float *x32;
    // MODE 1: dense XX
    double *XX;
    int n;
};

// one copy of a quartet block, permuted to the ordered bra (X,Y) with the ket
// in its stored orientation: dst[((x*ny + y)*nu + u)*nw + w]
template <int NX, int NY, int NU, int NW, typename F>
__device__ __forceinline__ void put_copy(double *__restrict__ x64, float *__restrict__ x32, long long base, F &&src) {
#pragma unroll
    for (int x = 0; x < NX; ++x)
#pragma unroll
        for (int y = 0; y < NY; ++y)
#pragma unroll
            for (int u = 0; u < NU; ++u)
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const long long e = base + ((x * NY + y) * NU + u) * NW + w;
                    const double v = src(x, y, u, w);
                    x64[e] = v;
                    if (x32) x32[e] = (float)v;
                }
}

__device__ __forceinline__ void find_quartet(const long long *__restrict__ cum, int nbra, long long t, int same,
                                             int &rb, long long &rk) {
    // last r with cum[r] <= t
    int lo = 0, hi = nbra;   // cum[0] = 0 <= t < cum[nbra]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (cum[mid] <= t) lo = mid;
        else hi = mid;
    }
    rb = lo;
    rk = t - cum[lo] + (same ? lo : 0);
}

__device__ __forceinline__ void load_ft(const double *__restrict__ ftT, double *s_ft) {
    for (int i = threadIdx.x; i < FT_NT * FT_NM; i += blockDim.x) s_ft[i] = ftT[i];
    __syncthreads();
}

// MODE 0: store the block's copies for digestion;  MODE 1: scatter into dense XX (8-fold)
template <int CB, int CK, int MODE>
__global__ __launch_bounds__(256) void k_eri(EriArgs a) {
    __shared__ double s_ft[FT_NT * FT_NM];
    load_ft(a.ftT, s_ft);
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nq) return;
    constexpr int LA = PairL<CB>::a, LB = PairL<CB>::b, LC = PairL<CK>::a, LD = PairL<CK>::b;
    constexpr int na = ncart(LA), nb = ncart(LB), nc = ncart(LC), nd = ncart(LD);
    constexpr int BLK = na * nb * nc * nd;
    int rb;
    long long rk;
    find_quartet(a.cum, a.nbra, t, a.same_class, rb, rk);
    const int pb = a.sorted_b[rb], pk = a.sorted_k[rk];
    const ShellPair ket = a.pairs[pk];
    QuartetAcc<LA, LB, LC, LD> acc;
    // A wavefront whose lanes share the bra pair (most do: its lanes take
    // consecutive kets of one bra) reads the bra's pair and primitive data
    // through scalar loads; the classes with few kets per bra, and waves that
    // straddle two bras, load them per lane.
    constexpr bool kUniformPath = CB + CK <= 2;
    const int pb0 = __builtin_amdgcn_readfirstlane(pb);
    ShellPair bra;
    if (kUniformPath && __all(pb == pb0)) {
        bra = a.pairs[pb0];
        quartet_prims_f<LA, LB, LC, LD>(bra, ket, a.prims, KetSoa{a.ksoa, a.nket, rk}, s_ft, acc);
    } else {
        bra = a.pairs[pb];
        quartet_prims_f<LA, LB, LC, LD>(bra, ket, a.prims, KetSoa{a.ksoa, a.nket, rk}, s_ft, acc);
    }
    double CD[3] = {ket.AB[0], ket.AB[1], ket.AB[2]};
    double out[BLK];
    quartet_hrr<LA, LB, LC, LD>(acc, bra.AB, CD, out);
    if (MODE == 0) {
        // copies for the row owners A (bra (A,B)), B (bra (B,A)) and C (ket (C,D))
        const int A = bra.sa, B = bra.sb, C = ket.sa, D = ket.sb, ns = a.ns;
        const long long rq = a.rank[pk], rp = a.rank[pb];
        auto at = [&](int ia, int ib, int ic, int id) { return out[((ia * nb + ib) * nc + ic) * nd + id]; };
        put_copy<na, nb, nc, nd>(a.x64, a.x32, a.segoff[((long long)A * ns + B) * 3 + CK] + rq * BLK,
                                 [&](int x, int y, int u, int w) { return at(x, y, u, w); });
        // the bra's reverse segment holds only the kets of which the bra pair
        // is the bra: in the same class those from its own rank on
        if (!bra.same)
            put_copy<nb, na, nc, nd>(a.x64, a.x32,
                                     a.segoff[((long long)B * ns + A) * 3 + CK] + (rq - (CB == CK ? rp : 0)) * BLK,
                                     [&](int x, int y, int u, int w) { return at(y, x, u, w); });
        // the ket pair's copy, in its stored orientation only
        if (pb != pk)
            put_copy<nc, nd, na, nb>(a.x64, a.x32, a.segoff[((long long)C * ns + D) * 3 + CB] + rp * BLK,
                                     [&](int x, int y, int u, int w) { return at(u, w, x, y); });
        return;
    }
    const int same_pair = (pb == pk);
#pragma unroll
    for (int ia = 0; ia < na; ++ia)
#pragma unroll
        for (int ib = 0; ib < nb; ++ib)
#pragma unroll
            for (int ic = 0; ic < nc; ++ic)
#pragma unroll
                for (int id = 0; id < nd; ++id) {
                    const int e = ((ia * nb + ib) * nc + ic) * nd + id;
                    const int i = bra.ao_a + ia, j = bra.ao_b + ib, k = ket.ao_a + ic, l = ket.ao_b + id;
                    const double w = block_weight(ia, ib, ic, id, bra.same, ket.same, same_pair, nb, nd, i, j, k, l);
                    if (w != 0.0) {
                        const size_t n = a.n;
                        const double v = out[e];
                        const size_t I = i, J = j, K = k, L = l;
                        a.XX[I + n * (J + n * (K + n * L))] = v;
                        a.XX[J + n * (I + n * (K + n * L))] = v;
                        a.XX[I + n * (J + n * (L + n * K))] = v;
                        a.XX[J + n * (I + n * (L + n * K))] = v;
                        a.XX[K + n * (L + n * (I + n * J))] = v;
                        a.XX[L + n * (K + n * (I + n * J))] = v;
                        a.XX[K + n * (L + n * (J + n * I))] = v;
                        a.XX[L + n * (K + n * (J + n * I))] = v;
                    }
                }
}

// Schwarz factor of each pair of one class: sqrt(max_ab |(ab|ab)|)
template <int C>
__global__ __launch_bounds__(256) void k_schwarz(const ShellPair *__restrict__ pairs, const PrimPair *__restrict__ prims,
                                                 const double *__restrict__ ftT, const int *__restrict__ ids, int np,
                                                 double *__restrict__ q) {
    __shared__ double s_ft[FT_NT * FT_NM];
    load_ft(ftT, s_ft);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    constexpr int LA = PairL<C>::a, LB = PairL<C>::b;
    constexpr int na = ncart(LA), nb = ncart(LB);
    const int p = ids[t];
    const ShellPair sp = pairs[p];
    QuartetAcc<LA, LB, LA, LB> acc;
    quartet_prims<LA, LB, LA, LB, false>(sp, sp, prims, prims, s_ft, acc);
    double out[na * nb * na * nb];
    quartet_hrr<LA, LB, LA, LB>(acc, sp.AB, sp.AB, out);
    double m = 0.0;
#pragma unroll
    for (int ia = 0; ia < na; ++ia)
#pragma unroll
        for (int ib = 0; ib < nb; ++ib) m = fmax(m, fabs(out[((ia * nb + ib) * na + ia) * nb + ib]));
    q[p] = sqrt(m);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static const int kCB[6] = {0, 1, 1, 2, 2, 2};
static const int kCK[6] = {0, 0, 1, 0, 1, 2};
static const int kBLK[6] = {1, 3, 9, 9, 27, 81};

// Drops the stored integrals; with keep_memory the device buffers stay
// allocated (their capacity is reused by the next build).
void free_eri(myqc_ctx *ctx, bool keep_memory) {
    for (auto &c : ctx->cls) {
        if (!keep_memory) dfree((void **)&c.d_cum);
        c.nq = 0;
    }
    void **ptrs[] = {(void **)&ctx->d_segoff, (void **)&ctx->d_rank,     (void **)&ctx->d_x64,
                     (void **)&ctx->d_x32,    (void **)&ctx->d_wg,       (void **)&ctx->d_unit,
                     (void **)&ctx->d_unit_off, (void **)&ctx->d_shell, (void **)&ctx->d_ao_shell,
                     (void **)&ctx->d_partial, (void **)&ctx->d_Dpair,
                     (void **)&ctx->d_Dpair32, (void **)&ctx->d_chunk};
    if (!keep_memory) {
        for (auto p : ptrs) dfree(p);
        for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_meta[c]);
        for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_qrank[c]);
        dfree((void **)&ctx->d_chunkq);
    }
    ctx->n_wg = ctx->n_units = ctx->n_chunks = 0;
    ctx->n_expanded = 0;
    ctx->have_eri = false;
    ctx->have_eri32 = false;
}

int eri_schwarz(myqc_ctx *ctx) {
    if (ctx->have_schwarz) return 0;
    const int np = (int)ctx->pairs.size();
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_schwarz, sizeof(double) * np));
    for (int c = 0; c < 3; ++c) {
        const auto &ids = ctx->pair_ids[c];
        if (ids.empty()) continue;
        MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_sorted[c], sizeof(int) * ids.size()));
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_sorted[c], ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
        const int nblk = ((int)ids.size() + 255) / 256;
        if (c == 0)
            k_schwarz<0><<<nblk, 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims, ctx->d_ftT, ctx->d_sorted[c],
                                                         (int)ids.size(), ctx->d_schwarz);
        else if (c == 1)
            k_schwarz<1><<<nblk, 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims, ctx->d_ftT, ctx->d_sorted[c],
                                                         (int)ids.size(), ctx->d_schwarz);
        else
            k_schwarz<2><<<nblk, 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims, ctx->d_ftT, ctx->d_sorted[c],
                                                         (int)ids.size(), ctx->d_schwarz);
        MYQC_HIP(ctx, hipGetLastError());
    }
    ctx->schwarz.resize(np);
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->schwarz.data(), ctx->d_schwarz, sizeof(double) * np, hipMemcpyDeviceToHost,
                                 ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (const ShellPair &sp : ctx->pairs)
        if (sp.npp > kKetPrims)
            return set_error(ctx, MYQC_E_ARG, "a shell pair has %d primitive pairs (at most %d: STO-3G shells)",
                             sp.npp, kKetPrims);
    // sort each class by Schwarz factor, largest first (ties by pair id); one
    // host thread per class
    {
        auto sort_class = [ctx](int c) {
            auto &ids = ctx->pair_ids[c];
            std::vector<std::pair<double, int>> key(ids.size());
            for (size_t r = 0; r < ids.size(); ++r) key[r] = {-ctx->schwarz[ids[r]], ids[r]};
            std::sort(key.begin(), key.end());   // descending factor, ties by pair id
            for (size_t r = 0; r < ids.size(); ++r) ids[r] = key[r].second;
        };
        std::thread t1(sort_class, 1), t2(sort_class, 2);
        sort_class(0);
        t1.join();
        t2.join();
    }
    for (int c = 0; c < 3; ++c) {
        auto &ids = ctx->pair_ids[c];
        if (!ids.empty()) {
            MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_sorted[c], ids.data(), sizeof(int) * ids.size(),
                                         hipMemcpyHostToDevice, ctx->stream));
            const int np = (int)ids.size();
            MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_ksoa[c], sizeof(double) * kKetPrims * kKetFields * (size_t)np));
            k_ket_soa<<<dim3((np + 255) / 256, kKetPrims), 256, 0, ctx->stream>>>(ctx->d_pairs, ctx->d_prims,
                                                                                  ctx->d_sorted[c], np,
                                                                                  ctx->d_ksoa[c]);
            MYQC_HIP(ctx, hipGetLastError());
        }
    }
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_schwarz = true;
    return 0;
}

// Per-bra counts of surviving kets and their prefix sums for class (cb, ck).
static void class_counts(myqc_ctx *ctx, int cb, int ck, double tol, std::vector<long long> &cum) {
    const auto &B = ctx->pair_ids[cb];
    const auto &K = ctx->pair_ids[ck];
    cum.assign(B.size() + 1, 0);
    std::vector<double> qk(K.size());
    for (size_t i = 0; i < K.size(); ++i) qk[i] = ctx->schwarz[K[i]];
    // bras in descending Schwarz order need ascending thresholds: the number
    // of kets with qk >= tol / qb only shrinks, so one pointer walks down
    size_t ptr = qk.size();
    for (size_t r = 0; r < B.size(); ++r) {
        const double qb = ctx->schwarz[B[r]];
        long long ub;
        if (tol <= 0.0) ub = (long long)K.size();
        else if (qb <= 0.0) ub = 0;
        else {
            const double need = tol / qb;
            while (ptr > 0 && !(qk[ptr - 1] >= need)) --ptr;
            ub = (long long)ptr;
        }
        long long cnt = (cb == ck) ? std::max(0LL, ub - (long long)r) : ub;
        cum[r + 1] = cum[r] + cnt;
    }
}

template <int CB, int CK, int MODE>
static void launch_eri(myqc_ctx *ctx, const QuartetClass &qc, double *XX, hipStream_t s) {
    EriArgs a;
    a.pairs = ctx->d_pairs;
    a.prims = ctx->d_prims;
    a.ftT = ctx->d_ftT;
    a.sorted_b = ctx->d_sorted[CB];
    a.sorted_k = ctx->d_sorted[CK];
    a.ksoa = ctx->d_ksoa[CK];
    a.nket = (long long)ctx->pair_ids[CK].size();
    a.cum = qc.d_cum;
    a.nbra = qc.nbra;
    a.nq = qc.nq;
    a.same_class = (CB == CK);
    a.segoff = ctx->d_segoff;
    a.rank = ctx->d_rank;
    a.ns = ctx->ns;
    a.x64 = ctx->d_x64;
    a.x32 = ctx->d_x32;
    a.XX = XX;
    a.n = ctx->n;
    const long long nblk = (qc.nq + 255) / 256;
    k_eri<CB, CK, MODE><<<(unsigned)nblk, 256, 0, s>>>(a);
}

template <int MODE>
static void launch_class(myqc_ctx *ctx, int c, double *XX, hipStream_t s) {
    const QuartetClass &qc = ctx->cls[c];
    if (qc.nq == 0) return;
    switch (c) {
        case 0: launch_eri<0, 0, MODE>(ctx, qc, XX, s); break;
        case 1: launch_eri<1, 0, MODE>(ctx, qc, XX, s); break;
        case 2: launch_eri<1, 1, MODE>(ctx, qc, XX, s); break;
        case 3: launch_eri<2, 0, MODE>(ctx, qc, XX, s); break;
        case 4: launch_eri<2, 1, MODE>(ctx, qc, XX, s); break;
        case 5: launch_eri<2, 2, MODE>(ctx, qc, XX, s); break;
    }
}

// FP64 operations per contracted shell quartet of each class (FMA = 2,
// transcendental = 1), for the roofline's flops: the FP64 VALU instruction
// counters (ADD, MUL, FMA, TRANS _F64, each x 64 lanes) of one build of the
// 40-water STO-3G cluster at Schwarz 1e-12, divided by the class's quartet
// count.  They include lanes that sit idle while a neighbour computes a
// primitive quartet that its own pair screen drops, and both sides of a
// divergent Boys branch, so they count issued rather than strictly useful work.
static const double kQuartetFlops[6] = {4836.7, 4930.0, 6708.2, 6929.1, 11986.4, 33463.8};

static int build_classes(myqc_ctx *ctx, double tol) {
    std::vector<long long> cum;
    ctx->n_quartets = 0;
    ctx->n_stored = 0;
    ctx->eri_flops = 0.0;
    for (int c = 0; c < 6; ++c) {
        QuartetClass &qc = ctx->cls[c];
        qc.cb = kCB[c];
        qc.ck = kCK[c];
        qc.blk = kBLK[c];
        qc.nbra = (int)ctx->pair_ids[qc.cb].size();
        if (qc.nbra == 0 || ctx->pair_ids[qc.ck].empty()) {
            qc.nq = 0;
            continue;
        }
        class_counts(ctx, qc.cb, qc.ck, tol, cum);
        qc.nq = cum.back();
        MYQC_CHECK(dalloc(ctx, (void **)&qc.d_cum, sizeof(long long) * cum.size()));
        MYQC_HIP(ctx, hipMemcpyAsync(qc.d_cum, cum.data(), sizeof(long long) * cum.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
        ctx->n_quartets += qc.nq;
        ctx->n_stored += qc.nq * qc.blk;
        ctx->eri_flops += (double)qc.nq * kQuartetFlops[c];
        if (getenv("MYQC_HOST_TIMING"))
            fprintf(stderr, "[host] eri class (%d,%d): %lld quartets, %.6e flops by the table\n", qc.cb, qc.ck, qc.nq,
                    (double)qc.nq * kQuartetFlops[c]);
    }
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// Segments of the expanded storage and the digestion schedule.  For every
// ordered shell pair (X,Y) and ket class ck, the kets of class ck whose
// Schwarz product with (X,Y) passes `tol` form a prefix of that class's
// sorted pair list; their blocks (X Y|ket) are stored contiguously.
// kets per lane of a digestion chunk (jk.hip)
__host__ __device__ constexpr int kets_per_lane(int row) { return row == 1 ? 8 : row == 3 ? 4 : row == 9 ? 2 : 1; }

static int build_segments(myqc_ctx *ctx, double tol) {
    const int ns = (int)ctx->shells.size();
    ctx->ns = ns;
    const int np = (int)ctx->pairs.size();
    // pair of each unordered shell pair, and its Schwarz factor
    std::vector<int> pair_of((size_t)ns * ns, -1);
    for (int p = 0; p < np; ++p) {
        pair_of[(size_t)ctx->pairs[p].sa * ns + ctx->pairs[p].sb] = p;
        pair_of[(size_t)ctx->pairs[p].sb * ns + ctx->pairs[p].sa] = p;
    }
    std::vector<int> rank(np, 0);
    std::vector<double> qsorted[3];
    for (int c = 0; c < 3; ++c) {
        const auto &ids = ctx->pair_ids[c];
        qsorted[c].resize(ids.size());
        for (size_t r = 0; r < ids.size(); ++r) {
            rank[ids[r]] = (int)r;
            qsorted[c][r] = ctx->schwarz[ids[r]];
        }
    }
    // screened ket count of every pair in each ket class: kets with
    // q_P q_Q >= tol form a prefix of the Schwarz-sorted class list
    // (each class's pairs in descending Schwarz order need ascending
    // thresholds tol / q_p, so one pointer per ket class walks down once)
    std::vector<int> pcnt((size_t)np * 3);
    for (int cb = 0; cb < 3; ++cb) {
        const auto &ids = ctx->pair_ids[cb];
        for (int ck = 0; ck < 3; ++ck) {
            const auto &qk = qsorted[ck];
            size_t ptr = qk.size();
            for (size_t r = 0; r < ids.size(); ++r) {
                const int p = ids[r];
                const double qp = ctx->schwarz[p];
                long long cnt;
                if (tol <= 0.0) cnt = (long long)qk.size();
                else if (qp <= 0.0) cnt = 0;
                else {
                    const double thr = tol / qp;
                    while (ptr > 0 && !(qk[ptr - 1] >= thr)) --ptr;
                    cnt = (long long)ptr;
                }
                pcnt[(size_t)p * 3 + ck] = (int)cnt;
            }
        }
    }
    // a pair's forward segment (its stored orientation) holds every screened
    // partner; its reverse segment only the partners of which it is the bra
    // (lower classes, and in its own class those from its own rank on)
    std::vector<int> pcls(np, 0);
    for (int c = 0; c < 3; ++c)
        for (int id : ctx->pair_ids[c]) pcls[id] = c;
    ctx->segoff.assign((size_t)ns * ns * 3, 0);
    ctx->segcnt.assign((size_t)ns * ns * 3, 0);
    ctx->segstart.assign((size_t)ns * ns * 3, 0);
    long long off = 0;
    for (int X = 0; X < ns; ++X)
        for (int Y = 0; Y < ns; ++Y) {
            const int p = pair_of[(size_t)X * ns + Y];
            const int nxy = (ctx->shells[X].L ? 3 : 1) * (ctx->shells[Y].L ? 3 : 1);
            const bool fwd = p >= 0 && ctx->pairs[p].sa == X;
            for (int ck = 0; ck < 3; ++ck) {
                long long cnt = p >= 0 ? pcnt[(size_t)p * 3 + ck] : 0, start = 0;
                if (p >= 0 && !fwd) {
                    if (ck > pcls[p]) cnt = 0;
                    else if (ck == pcls[p]) {
                        start = rank[p];
                        cnt = std::max(0LL, cnt - start);
                    }
                }
                const size_t seg = ((size_t)X * ns + Y) * 3 + ck;
                ctx->segoff[seg] = off;
                ctx->segcnt[seg] = (int)cnt;
                ctx->segstart[seg] = (int)start;
                off += cnt * nxy * (ck == 0 ? 1 : ck == 1 ? 3 : 9);
            }
        }
    ctx->n_expanded = off;
    ctx->pair_of.swap(pair_of);
    ctx->pair_rank = rank;
    ctx->pair_cls.swap(pcls);
    const size_t nseg = (size_t)ns * ns * 3;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_segoff, sizeof(long long) * nseg));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_rank, sizeof(int) * np));
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_segoff, ctx->segoff.data(), sizeof(long long) * nseg, hipMemcpyHostToDevice,
                                 ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_rank, ctx->pair_rank.data(), sizeof(int) * np, hipMemcpyHostToDevice,
                                 ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// The digestion schedule over the segments of build_segments.  It is built
// on the host while the ERI kernels that fill the segments run.
static int build_schedule(myqc_ctx *ctx) {
    const int ns = ctx->ns;
    // digestion schedule: per row shell X, its units Y (all three ket
    // classes) packed into workgroups of about kTarget values whose Y shells
    // span at most dig_rows AOs (the D rows a workgroup stages in LDS: with
    // the 3 owned rows of G, (3 + dig_rows) n doubles stay within
    // kLdsBudget so that three workgroups fit a CU); each unit's kets are
    // cut into chunks of at most 64 of one class, one wavefront's work
    const long long kTarget = 1 << 15;
    const size_t kLdsBudget = 52 * 1024;
    const int n = ctx->n;
    ctx->dig_rows = std::max(3, std::min(64, (int)(kLdsBudget / (sizeof(double) * (size_t)n)) - 3));
    std::vector<int> wg, unit, chunk, shell(4 * ns);
    std::vector<long long> unit_off, wg_cost;
    // per chunk {q_row q_first, q_row q_last, q_row, 0} (the incremental screen)
    std::vector<float> chunkq;
    {
        long long nchunk = 0, nunit = 0;
        for (size_t seg = 0; seg < ctx->segcnt.size(); ++seg) {
            nchunk += (ctx->segcnt[seg] + 63) / 64 + 2;
            nunit += (seg % 3 == 0) && (ctx->segcnt[seg] + ctx->segcnt[seg + 1] + ctx->segcnt[seg + 2] > 0);
        }
        chunk.reserve(4 * (size_t)nchunk);
        chunkq.reserve(4 * (size_t)nchunk);
        unit.reserve(4 * (size_t)nunit);
        unit_off.reserve(3 * (size_t)nunit);
    }
    for (int X = 0; X < ns; ++X) {
        shell[4 * X + 0] = ctx->shells[X].ao;
        shell[4 * X + 1] = ctx->shells[X].L;
        shell[4 * X + 2] = (int)(wg.size() / 6);
        const int nX = ctx->shells[X].L ? 3 : 1;
        long long acc = 0;
        int ch0 = (int)(chunk.size() / 4), row0 = -1, row1 = -1;
        auto close = [&]() {
            const int ch1 = (int)(chunk.size() / 4);
            if (ch1 > ch0) {
                wg.push_back(X);
                wg.push_back(ch0);
                wg.push_back(ch1);
                wg.push_back(row0);
                wg.push_back(row1 - row0);
                wg.push_back((int)(wg.size() / 6));   // partial slot: the workgroups of X stay contiguous
                wg_cost.push_back(acc);
            }
            ch0 = ch1;
            row0 = row1 = -1;
            acc = 0;
        };
        for (int Y = 0; Y < ns; ++Y) {
            const size_t seg = ((size_t)X * ns + Y) * 3;
            const int cnt[3] = {ctx->segcnt[seg], ctx->segcnt[seg + 1], ctx->segcnt[seg + 2]};
            if (cnt[0] + cnt[1] + cnt[2] == 0) continue;
            const int aoY = ctx->shells[Y].ao, nY = ctx->shells[Y].L ? 3 : 1;
            if (row0 >= 0 && aoY + nY - row0 > ctx->dig_rows) close();
            unit.push_back(Y);
            for (int ck = 0; ck < 3; ++ck) unit.push_back(cnt[ck]);
            const int p = ctx->pair_of[(size_t)X * ns + Y];
            const bool fwd = ctx->pairs[p].sa == X;
            const int cp = ctx->pair_cls[p], rp = ctx->pair_rank[p];
            const double qrow = ctx->schwarz[p];
            for (int ck = 0; ck < 3; ++ck) {
                const int start = ctx->segstart[seg + ck], end = start + cnt[ck];
                const int row = nY * (ck == 0 ? 1 : ck == 1 ? 3 : 9);
                // offset such that the block of the ket with rank r sits at off + r * block
                const long long off = ctx->segoff[seg + ck] - (long long)start * nX * row;
                unit_off.push_back(off);
                const int step = 64 * kets_per_lane(row);
                // role ranges of the segment's kets: [r0, r1) with flags
                // 1 = bra role, 2 = forward orientation, 4 = the diagonal quartet
                int cut[4] = {start, end, end, end}, flag[3] = {0, 0, 0};
                const int fo = fwd ? 2 : 0;
                if (ck < cp) flag[0] = 1 | fo;
                else if (ck > cp) flag[0] = 0 | fo;
                else {
                    // same class: kets below the pair's rank are bras of it,
                    // the pair itself is the diagonal quartet, the rest its kets
                    const int d0 = std::min(std::max(rp, start), end);
                    cut[1] = d0;
                    cut[2] = std::min(d0 + (rp >= start && rp < end ? 1 : 0), end);
                    flag[0] = 0 | fo;
                    flag[1] = 1 | 4 | fo;
                    flag[2] = 1 | fo;
                }
                for (int g = 0; g < 3; ++g)
                    for (int r0 = cut[g]; r0 < cut[g + 1]; r0 += step) {
                        // a workgroup closes after about kTarget values, also
                        // inside a unit: chunks are independent
                        if (row0 < 0) row0 = aoY;
                        row1 = aoY + nY;
                        const int nk = std::min(step, cut[g + 1] - r0);
                        chunk.push_back((int)(unsigned)(off & 0xffffffffLL));
                        chunk.push_back((int)(off >> 32));
                        chunk.push_back(aoY | (ctx->shells[Y].L << 15) | (ck << 16) | (nk << 18) | (flag[g] << 28));
                        chunk.push_back(r0);
                        {
                            const auto &ids = ctx->pair_ids[ck];
                            const float up = std::numeric_limits<float>::infinity();
                            chunkq.push_back(std::nextafter((float)(qrow * ctx->schwarz[ids[r0]]), up));
                            chunkq.push_back(std::nextafter((float)(qrow * ctx->schwarz[ids[r0 + nk - 1]]), up));
                            chunkq.push_back(std::nextafter((float)qrow, up));
                            chunkq.push_back(0.0f);
                        }
                        acc += (long long)nX * row * nk;
                        if (acc >= kTarget) close();
                    }
            }
        }
        close();
        shell[4 * X + 3] = (int)(wg.size() / 6);
    }
    // launch order: most values first, so that the long workgroups do not
    // form the tail of the pass
    {
        const size_t nw = wg_cost.size();
        std::vector<int> order(nw);
        for (size_t i = 0; i < nw; ++i) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return wg_cost[p] > wg_cost[q]; });
        std::vector<int> sorted(wg.size());
        for (size_t i = 0; i < nw; ++i)
            for (int k = 0; k < 6; ++k) sorted[6 * i + k] = wg[6 * (size_t)order[i] + k];
        wg.swap(sorted);
    }
    ctx->n_wg = (int)(wg.size() / 6);
    ctx->n_chunks = (int)(chunk.size() / 4);
    ctx->n_units = (int)(unit.size() / 4);
    // ket metadata in Schwarz-rank order per class
    std::vector<int> meta[3];
    for (int c = 0; c < 3; ++c) {
        const auto &ids = ctx->pair_ids[c];
        if (ids.empty()) continue;
        meta[c].resize(ids.size());
        for (size_t r = 0; r < ids.size(); ++r) {
            const ShellPair &sp = ctx->pairs[ids[r]];
            meta[c][r] = sp.ao_a | (sp.ao_b << 15) | (sp.same << 30);
        }
        MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_meta[c], sizeof(int) * meta[c].size()));
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_meta[c], meta[c].data(), sizeof(int) * meta[c].size(),
                                     hipMemcpyHostToDevice, ctx->stream));
    }
    // Schwarz factors by rank (rounded up, so the screen never drops a
    // partner that the exact factor keeps)
    std::vector<float> qrank[3];
    for (int c = 0; c < 3; ++c) {
        const auto &ids = ctx->pair_ids[c];
        if (ids.empty()) continue;
        qrank[c].resize(ids.size());
        for (size_t r = 0; r < ids.size(); ++r)
            qrank[c][r] = std::nextafter((float)ctx->schwarz[ids[r]], std::numeric_limits<float>::infinity());
        MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_qrank[c], sizeof(float) * qrank[c].size()));
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_qrank[c], qrank[c].data(), sizeof(float) * qrank[c].size(),
                                     hipMemcpyHostToDevice, ctx->stream));
    }
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_chunkq, sizeof(float) * std::max<size_t>(chunkq.size(), 4)));
    if (!chunkq.empty())
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_chunkq, chunkq.data(), sizeof(float) * chunkq.size(),
                                     hipMemcpyHostToDevice, ctx->stream));
    std::vector<int> ao_shell(ctx->n);
    for (int X = 0; X < ns; ++X)
        for (int k = 0; k < (ctx->shells[X].L ? 3 : 1); ++k) ao_shell[ctx->shells[X].ao + k] = X;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_wg, sizeof(int) * std::max<size_t>(wg.size(), 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_unit, sizeof(int) * std::max<size_t>(unit.size(), 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_unit_off, sizeof(long long) * std::max<size_t>(unit_off.size(), 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_chunk, sizeof(int) * std::max<size_t>(chunk.size(), 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_shell, sizeof(int) * shell.size()));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_ao_shell, sizeof(int) * ao_shell.size()));
    if (!wg.empty())
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_wg, wg.data(), sizeof(int) * wg.size(), hipMemcpyHostToDevice, ctx->stream));
    if (!unit.empty()) {
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_unit, unit.data(), sizeof(int) * unit.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_unit_off, unit_off.data(), sizeof(long long) * unit_off.size(),
                                     hipMemcpyHostToDevice, ctx->stream));
        MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_chunk, chunk.data(), sizeof(int) * chunk.size(), hipMemcpyHostToDevice,
                                     ctx->stream));
    }
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_shell, shell.data(), sizeof(int) * shell.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_ao_shell, ao_shell.data(), sizeof(int) * ao_shell.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// the side streams and fork/join events, created on first use
static int ensure_side_streams(myqc_ctx *ctx) {
    if (ctx->fork_ev) return 0;
    for (int k = 0; k < 2; ++k) {
        MYQC_HIP(ctx, hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
        MYQC_HIP(ctx, hipEventCreateWithFlags(&ctx->join_ev[k], hipEventDisableTiming));
    }
    MYQC_HIP(ctx, hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
    return 0;
}

int eri_store(myqc_ctx *ctx, double tol, bool want32) {
    HostClock hc(ctx);
    MYQC_CHECK(eri_schwarz(ctx));
    hc.mark("eri_store: Schwarz factors");
    free_eri(ctx, true);
    MYQC_CHECK(build_classes(ctx, tol));
    hc.mark("eri_store: quartet classes");
    MYQC_CHECK(build_segments(ctx, tol));
    hc.mark("eri_store: segments");
    const size_t nx = (size_t)std::max<long long>(ctx->n_expanded, 1);
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_x64, sizeof(double) * nx));
    if (want32) MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_x32, sizeof(float) * nx));
    else dfree((void **)&ctx->d_x32);   // a kept buffer may be smaller than this build
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_Dpair, sizeof(double) * 9 * ctx->pairs.size()));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_Dpair32, sizeof(float) * 9 * ctx->pairs.size()));
    MYQC_CHECK(ensure_side_streams(ctx));
    hipEvent_t ev;
    MYQC_CHECK(timer_begin(ctx, T_ERI, &ev));
    // the classes write disjoint blocks of the store, so they may overlap:
    // the L-bra classes (few waves, long ones) start first on side stream 0,
    // (Ls|Ls) on side stream 1, the two large s-ket classes on the context
    // stream; the context stream then waits for both side streams
    MYQC_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
    for (int k = 0; k < 2; ++k) MYQC_HIP(ctx, hipStreamWaitEvent(ctx->side[k], ctx->fork_ev, 0));
    static const int kOrder[6] = {5, 4, 3, 2, 0, 1};
    for (int c : kOrder) {
        if (ctx->cls[c].nq == 0) continue;
        launch_class<0>(ctx, c, nullptr, c >= 3 ? ctx->side[0] : c == 2 ? ctx->side[1] : ctx->stream);
        MYQC_HIP(ctx, hipGetLastError());
    }
    for (int k = 0; k < 2; ++k) {
        MYQC_HIP(ctx, hipEventRecord(ctx->join_ev[k], ctx->side[k]));
        MYQC_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->join_ev[k], 0));
    }
    MYQC_CHECK(timer_end(ctx, T_ERI, ev, ctx->eri_flops));
    hc.mark("eri_store: allocations and launches");
    // the digestion schedule is built while the kernels run
    MYQC_CHECK(build_schedule(ctx));
    const size_t part = (size_t)std::max(ctx->n_wg, 1) * 3 * ctx->n;
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_partial, sizeof(double) * part));
    hc.mark("eri_store: schedule (host, beside the kernels)");
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hc.mark("eri_store: kernels");
    ctx->have_eri = true;
    ctx->have_eri32 = want32;
    ctx->eri_tol = tol;
    return 0;
}

int eri_fill_dense(myqc_ctx *ctx, double tol, double *d_XX) {
    MYQC_CHECK(eri_schwarz(ctx));
    const size_t n = ctx->n;
    MYQC_HIP(ctx, hipMemsetAsync(d_XX, 0, sizeof(double) * n * n * n * n, ctx->stream));
    // the dense fill shares the class enumeration with the stored build
    QuartetClass saved[6];
    bool had = ctx->have_eri;
    const long long saved_stored = ctx->n_stored, saved_quartets = ctx->n_quartets;
    const double saved_flops = ctx->eri_flops;
    for (int c = 0; c < 6; ++c) saved[c] = ctx->cls[c];
    for (int c = 0; c < 6; ++c) ctx->cls[c] = QuartetClass{};
    int rc = build_classes(ctx, tol);
    if (rc == 0) {
        hipEvent_t ev;
        rc = timer_begin(ctx, T_ERI, &ev);
        for (int c = 0; c < 6 && rc == 0; ++c) {
            launch_class<1>(ctx, c, d_XX, ctx->stream);
            if (hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "ERI dense-fill launch failed");
        }
        if (rc == 0) rc = timer_end(ctx, T_ERI, ev, ctx->eri_flops);
        if (rc == 0 && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "ERI dense fill failed");
    }
    for (int c = 0; c < 6; ++c) dfree((void **)&ctx->cls[c].d_cum);
    for (int c = 0; c < 6; ++c) ctx->cls[c] = saved[c];
    ctx->have_eri = had;
    ctx->n_stored = saved_stored;
    ctx->n_quartets = saved_quartets;
    ctx->eri_flops = saved_flops;
    return rc;
}

}  // namespace myqc

using namespace myqc;

extern "C" int myqc_set_system(myqc_ctx *ctx, int nnuc, const int *Z, const double *xyz_bohr, int nset,
                               const double *set_alpha, const int *setinfo, const double *bas, int norb,
                               const int *basinfo, const double *Ftab, int flags) {
    if (!ctx) return MYQC_E_ARG;
    if (nnuc <= 0 || nset <= 0 || norb <= 0 || !Z || !xyz_bohr || !set_alpha || !setinfo || !bas || !basinfo || !Ftab)
        return set_error(ctx, MYQC_E_ARG, "myqc_set_system: invalid arguments");
    if (flags != MYQC_REF_COMPAT) return set_error(ctx, MYQC_E_ARG, "myqc_set_system: unknown flags %d", flags);
    HostClock hc(ctx);
    free_eri(ctx, true);
    ctx->have_system = false;
    ctx->have_schwarz = false;
    ctx->n = norb;
    ctx->nnuc = nnuc;
    ctx->Z.assign(Z, Z + nnuc);
    ctx->xyz.resize(3 * nnuc);
    for (int a = 0; a < nnuc; ++a)
        for (int c = 0; c < 3; ++c) ctx->xyz[3 * a + c] = xyz_bohr[a + nnuc * c];
    const std::string why = build_system(nnuc, ctx->xyz.data(), nset, set_alpha, setinfo, bas, norb, basinfo,
                                         ctx->shells, ctx->pairs, ctx->prims, ctx->pair_ids);
    if (!why.empty()) return set_error(ctx, MYQC_E_ARG, "myqc_set_system: %s", why.c_str());
    hc.mark("set_system: shells and pairs");
    // Boys table transposed to [Tk][m] so the 7 Taylor rows are contiguous
    ctx->ftT.assign(FT_NT * FT_NM, 0.0);
    for (int tk = 0; tk < FT_NT; ++tk)
        for (int m = 0; m < FT_NM; ++m) ctx->ftT[tk * FT_NM + m] = Ftab[tk + FT_NT * m];
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_pairs, sizeof(ShellPair) * ctx->pairs.size()));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_prims, sizeof(PrimPair) * ctx->prims.size()));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_ftT, sizeof(double) * ctx->ftT.size()));
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_pairs, ctx->pairs.data(), sizeof(ShellPair) * ctx->pairs.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_prims, ctx->prims.data(), sizeof(PrimPair) * ctx->prims.size(),
                                 hipMemcpyHostToDevice, ctx->stream));
    MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_ftT, ctx->ftT.data(), sizeof(double) * ctx->ftT.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hc.mark("set_system: uploads");
    ctx->have_system = true;
    ctx->work_n = -1;
    return 0;
}

extern "C" int myqc_eri_full(myqc_ctx *ctx, double schwarz_tol, double *XX) {
    if (!ctx || !XX) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_eri_full: no system set");
    const size_t n = ctx->n, n4 = n * n * n * n;
    double *d_XX = nullptr;
    MYQC_CHECK(dalloc(ctx, (void **)&d_XX, sizeof(double) * n4));
    int rc = eri_fill_dense(ctx, schwarz_tol, d_XX);
    if (rc == 0) {
        hipError_t e = d2h(ctx, XX, d_XX, sizeof(double) * n4);
        if (e != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "myqc_eri_full: copy failed: %s", hipGetErrorString(e));
    }
    dfree((void **)&d_XX);
    return rc;
}

extern "C" int myqc_eri_build(myqc_ctx *ctx, double schwarz_tol, long long *n_stored) {
    if (!ctx) return MYQC_E_ARG;
    if (!ctx->have_system) return set_error(ctx, MYQC_E_STATE, "myqc_eri_build: no system set");
    if (schwarz_tol < 0) return set_error(ctx, MYQC_E_ARG, "myqc_eri_build: negative tolerance");
    MYQC_CHECK(eri_store(ctx, schwarz_tol, true));
    if (n_stored) *n_stored = ctx->n_stored;
    return 0;
}

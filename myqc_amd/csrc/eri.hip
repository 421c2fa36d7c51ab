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
#include "wave.h"
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
    // MODE 0: the stored (pre-weighted) block of each quartet
    const long long *segoff;   // [bra pair*3 + ket class]: element e of ket rank r at segoff + r + e*len
    double *x64;
    float *x32;
    // MODE 1: dense XX
    double *XX;
    int n;
};

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
    for (int i = threadIdx.x; i < FT_TOTAL; i += blockDim.x) s_ft[i] = ftT[i];
    __syncthreads();
}

// MODE 0: store the block's copies for digestion;  MODE 1: scatter into dense XX (8-fold).
// SPLIT lanes per shell quartet: they share its primitive quartets, add their
// sums (group_sum), and each stores every SPLIT-th element of the block.
// Waves per SIMD asked of the register allocator per class (the staged
// table's 23 KB of LDS allows seven 256-thread workgroups per CU): ss|ss 6,
// ps|ss 5, ps|ps and pp|ss 4, the rest unconstrained.
constexpr int eri_waves(int cb, int ck) { return cb + ck == 0 ? 6 : cb + ck == 1 ? 5 : cb + ck == 2 ? 4 : 1; }
template <int CB, int CK, int MODE, int SPLIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(eri_waves(CB, CK)))) void k_eri(EriArgs a) {
    __shared__ double s_ft[FT_TOTAL];
    load_ft(a.ftT, s_ft);
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long t = tid / SPLIT;
    const int sub = (int)(tid % SPLIT);
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
    if constexpr (SPLIT > 1) {
        bra = a.pairs[pb];
        quartet_prims_split<LA, LB, LC, LD, SPLIT>(bra, ket, a.prims, KetSoa{a.ksoa, a.nket, rk}, s_ft, acc, sub);
        using AQ = QuartetAcc<LA, LB, LC, LD>;
#pragma unroll
        for (int e = 0; e < AQ::NE; ++e)
#pragma unroll
            for (int f = 0; f < AQ::NF; ++f) acc.v[e][f] = group_sum<SPLIT>(acc.v[e][f]);
    } else if (kUniformPath && __all(pb == pb0)) {
        bra = a.pairs[pb0];
        quartet_prims_f<LA, LB, LC, LD>(bra, ket, a.prims, KetSoa{a.ksoa, a.nket, rk}, s_ft, acc);
    } else {
        bra = a.pairs[pb];
        quartet_prims_f<LA, LB, LC, LD>(bra, ket, a.prims, KetSoa{a.ksoa, a.nket, rk}, s_ft, acc);
    }
    double CD[3] = {ket.AB[0], ket.AB[1], ket.AB[2]};
    double out[BLK];
    quartet_hrr<LA, LB, LC, LD>(acc, bra.AB, CD, out);
    const int same_pair = (pb == pk);
    if (MODE == 0) {
        // one copy under the bra pair, each element times its degeneracy
        // weight (the digestion then applies the same six updates to all);
        // element-major over the bra's kets of this class (len of them)
        const long long base = a.segoff[(long long)pb * 3 + CK] + rk;
        const long long len = a.cum[rb + 1] - a.cum[rb];
#pragma unroll
        for (int ia = 0; ia < na; ++ia)
#pragma unroll
            for (int ib = 0; ib < nb; ++ib)
#pragma unroll
                for (int ic = 0; ic < nc; ++ic)
#pragma unroll
                    for (int id = 0; id < nd; ++id) {
                        const int e = ((ia * nb + ib) * nc + ic) * nd + id;
                        if (SPLIT > 1 && e % SPLIT != sub) continue;
                        const double w = block_weight(ia, ib, ic, id, bra.same, ket.same, same_pair, nb, nd,
                                                      bra.ao_a + ia, bra.ao_b + ib, ket.ao_a + ic, ket.ao_b + id);
                        const double v = out[e] * w;
                        a.x64[base + e * len] = v;
                        if (a.x32) a.x32[base + e * len] = (float)v;
                    }
        return;
    }
#pragma unroll
    for (int ia = 0; ia < na; ++ia)
#pragma unroll
        for (int ib = 0; ib < nb; ++ib)
#pragma unroll
            for (int ic = 0; ic < nc; ++ic)
#pragma unroll
                for (int id = 0; id < nd; ++id) {
                    const int e = ((ia * nb + ib) * nc + ic) * nd + id;
                    if (SPLIT > 1 && e % SPLIT != sub) continue;
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
    __shared__ double s_ft[FT_TOTAL];
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
    void **ptrs[] = {(void **)&ctx->d_segoff,   (void **)&ctx->d_rank,        (void **)&ctx->d_x64,
                     (void **)&ctx->d_x32,      (void **)&ctx->d_wg,          (void **)&ctx->d_unit,
                     (void **)&ctx->d_unit_off, (void **)&ctx->d_shell,       (void **)&ctx->d_ao_shell,
                     (void **)&ctx->d_partial,  (void **)&ctx->d_Dpair,       (void **)&ctx->d_Dpair32,
                     (void **)&ctx->d_chunk,    (void **)&ctx->d_row_ao,      (void **)&ctx->d_row_list,
                     (void **)&ctx->d_row_list_off, (void **)&ctx->d_jk_list, (void **)&ctx->d_jk_list_off,
                     (void **)&ctx->d_jkoff,    (void **)&ctx->d_jk_part,     (void **)&ctx->d_jk,
                     (void **)&ctx->d_jk_group};
    if (!keep_memory) {
        for (auto p : ptrs) dfree(p);
        for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_meta[c]);
        for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_qrank[c]);
    }
    ctx->n_wg = 0;
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

// Launch shape of each quartet class in the stored build: lanes per shell
// quartet (1, 2, 4, 8) and threads per workgroup (64, 128, 256).  The
// defaults hold until myqc_eri_tune measures the candidates on a system and
// keeps the fastest of each class for the process.
struct EriShape {
    int split = 1, tpb = 256;
};
static std::mutex g_shape_mu;
static EriShape g_shape[6];

static EriShape class_shape(int c) {
    std::lock_guard<std::mutex> g(g_shape_mu);
    return g_shape[c];
}

template <int CB, int CK, int MODE>
static void launch_eri(myqc_ctx *ctx, const QuartetClass &qc, double *XX, hipStream_t s, EriShape sh = EriShape{}) {
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
    a.x64 = ctx->d_x64;
    a.x32 = ctx->d_x32;
    a.XX = XX;
    a.n = ctx->n;
    if (MODE == 1) sh = EriShape{};   // the dense fill (small systems) keeps one lane per quartet
    const long long nblk = (qc.nq * sh.split + sh.tpb - 1) / sh.tpb;
    switch (sh.split) {
        case 2: k_eri<CB, CK, MODE, 2><<<(unsigned)nblk, sh.tpb, 0, s>>>(a); break;
        case 4: k_eri<CB, CK, MODE, 4><<<(unsigned)nblk, sh.tpb, 0, s>>>(a); break;
        case 8: k_eri<CB, CK, MODE, 8><<<(unsigned)nblk, sh.tpb, 0, s>>>(a); break;
        default: k_eri<CB, CK, MODE, 1><<<(unsigned)nblk, sh.tpb, 0, s>>>(a); break;
    }
}

template <int MODE>
static void launch_class(myqc_ctx *ctx, int c, double *XX, hipStream_t s, EriShape sh) {
    const QuartetClass &qc = ctx->cls[c];
    if (qc.nq == 0) return;
    switch (c) {
        case 0: launch_eri<0, 0, MODE>(ctx, qc, XX, s, sh); break;
        case 1: launch_eri<1, 0, MODE>(ctx, qc, XX, s, sh); break;
        case 2: launch_eri<1, 1, MODE>(ctx, qc, XX, s, sh); break;
        case 3: launch_eri<2, 0, MODE>(ctx, qc, XX, s, sh); break;
        case 4: launch_eri<2, 1, MODE>(ctx, qc, XX, s, sh); break;
        case 5: launch_eri<2, 2, MODE>(ctx, qc, XX, s, sh); break;
    }
}

// FP64 operations per contracted shell quartet of each class (FMA = 2,
// transcendental = 1), for the roofline's flops: the FP64 VALU instruction
// counters (ADD, MUL, FMA, TRANS _F64, each x 64 lanes) of one build of the
// 40-water STO-3G cluster at Schwarz 1e-12, divided by the class's quartet
// count.  They include lanes that sit idle while a neighbour computes a
// primitive quartet that its own pair screen drops, and both sides of a
// divergent Boys branch, so they count issued rather than strictly useful work.
static const double kQuartetFlops[6] = {3589.3, 4008.7, 5969.1, 6161.0, 11273.7, 34674.3};

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

// copies a host vector to a device buffer kept by the context
template <typename T>
static int upload(myqc_ctx *ctx, T **dst, const std::vector<T> &src) {
    MYQC_CHECK(dalloc(ctx, (void **)dst, sizeof(T) * std::max<size_t>(src.size(), 1)));
    if (!src.empty())
        MYQC_HIP(ctx, hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, ctx->stream));
    return 0;
}

// LDS the digestion kernel may use per workgroup: three workgroups per CU
constexpr size_t kDigestLds = 52 * 1024;

// The layout of the stored integrals (digest_plan.h): each screened unique
// quartet's block offset under its bra pair.  The ERI kernels need only this.
static int build_layout(myqc_ctx *ctx, double tol) {
    const int ns = (int)ctx->shells.size();
    ctx->ns = ns;
    std::vector<int> sao(ns);
    for (int X = 0; X < ns; ++X) sao[X] = ctx->shells[X].ao;
    const std::string e = build_digest_layout(ctx->plan, ctx->n, sao, ctx->pairs, ctx->pair_ids, ctx->schwarz, tol);
    if (!e.empty()) return set_error(ctx, MYQC_E_ARG, "%s", e.c_str());
    if (ctx->plan.n_values != ctx->n_stored || ctx->plan.n_quartets != ctx->n_quartets)
        return set_error(ctx, MYQC_E_STATE, "integral layout (%lld values) disagrees with the quartet classes (%lld)",
                         ctx->plan.n_values, ctx->n_stored);
    MYQC_CHECK(upload(ctx, &ctx->d_segoff, ctx->plan.off));
    MYQC_CHECK(upload(ctx, &ctx->d_rank, ctx->plan.rank));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// The digestion schedule over the layout, built on the host while the ERI
// kernels that fill the store run, and its device copies.
static int build_schedule(myqc_ctx *ctx) {
    const int ns = ctx->ns, n = ctx->n;
    std::vector<int> sao(ns), sL(ns);
    for (int X = 0; X < ns; ++X) {
        sao[X] = ctx->shells[X].ao;
        sL[X] = ctx->shells[X].L;
    }
    // LDS rows: the G rows (FP64) and the D rows (FP64 for the plain pass)
    const int max_rows = std::min(64, (int)(kDigestLds / (2 * sizeof(double) * (size_t)n)));
    if (max_rows < 6)
        return set_error(ctx, MYQC_E_ARG, "digestion: n=%d leaves fewer than 6 rows of G and D in %zu bytes of LDS", n,
                         kDigestLds);
    // workgroups of about 64k stored values: measured on the 40-water
    // cluster, 2^15 and 2^17 were 8 % and 10 % slower per pass (fewer chunks
    // per workgroup to balance its wavefronts, or too few workgroups)
    DigestPlan &pl = ctx->plan;
    const std::string e =
        build_digest_schedule(pl, sao, sL, ctx->pairs, ctx->pair_ids, ctx->schwarz, max_rows, 1LL << 16);
    if (!e.empty()) return set_error(ctx, MYQC_E_ARG, "%s", e.c_str());
    ctx->n_wg = (int)(pl.wg.size() / kWgInts);
    MYQC_CHECK(upload(ctx, &ctx->d_wg, pl.wg));
    MYQC_CHECK(upload(ctx, &ctx->d_unit, pl.unit));
    MYQC_CHECK(upload(ctx, &ctx->d_unit_off, pl.unit_off));
    MYQC_CHECK(upload(ctx, &ctx->d_chunk, pl.chunk));
    MYQC_CHECK(upload(ctx, &ctx->d_row_ao, pl.row_ao));
    MYQC_CHECK(upload(ctx, &ctx->d_row_list, pl.row_list));
    MYQC_CHECK(upload(ctx, &ctx->d_row_list_off, pl.row_list_off));
    MYQC_CHECK(upload(ctx, &ctx->d_jk_list, pl.jk_list));
    MYQC_CHECK(upload(ctx, &ctx->d_jk_list_off, pl.jk_list_off));
    MYQC_CHECK(upload(ctx, &ctx->d_jk_group, pl.jk_group));
    MYQC_CHECK(upload(ctx, &ctx->d_jkoff, pl.jkoff));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_jk_part, sizeof(double) * kJkSlot * std::max<size_t>(pl.n_jk_slots, 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_jk, sizeof(double) * std::max(pl.jk_base[3], 1)));
    MYQC_CHECK(dalloc(ctx, (void **)&ctx->d_partial, sizeof(double) * (size_t)std::max(pl.n_rows, 1) * n));
    // ket metadata and Schwarz factors in rank order per class
    for (int c = 0; c < 3; ++c) {
        const auto &ids = ctx->pair_ids[c];
        if (ids.empty()) continue;
        std::vector<int> meta(ids.size());
        std::vector<float> qrank(ids.size());
        for (size_t r = 0; r < ids.size(); ++r) {
            const ShellPair &sp = ctx->pairs[ids[r]];
            meta[r] = sp.ao_a | (sp.ao_b << 15);
            // rounded up, so that the screen never drops a ket the exact factor keeps
            qrank[r] = std::nextafter((float)ctx->schwarz[ids[r]], std::numeric_limits<float>::infinity());
        }
        MYQC_CHECK(upload(ctx, &ctx->d_meta[c], meta));
        MYQC_CHECK(upload(ctx, &ctx->d_qrank[c], qrank));
    }
    std::vector<int> shell(4 * ns, 0), ao_shell(n);
    for (int X = 0; X < ns; ++X) {
        shell[4 * X] = ctx->shells[X].ao;
        shell[4 * X + 1] = ctx->shells[X].L;
        for (int k = 0; k < (ctx->shells[X].L ? 3 : 1); ++k) ao_shell[ctx->shells[X].ao + k] = X;
    }
    MYQC_CHECK(upload(ctx, &ctx->d_shell, shell));
    MYQC_CHECK(upload(ctx, &ctx->d_ao_shell, ao_shell));
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
    MYQC_CHECK(build_layout(ctx, tol));
    hc.mark("eri_store: layout");
    const size_t nx = (size_t)std::max<long long>(ctx->n_stored, 1);
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
        launch_class<0>(ctx, c, nullptr, c >= 3 ? ctx->side[0] : c == 2 ? ctx->side[1] : ctx->stream,
                        class_shape(c));
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
    hc.mark("eri_store: schedule (host, beside the kernels)");
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    hc.mark("eri_store: kernels");
    ctx->have_eri = true;
    ctx->have_eri32 = want32;
    ctx->eri_tol = tol;
    return 0;
}

// The launch-shape search of the stored build: every (lanes per quartet,
// threads per workgroup) candidate of every class with quartets in the
// current system is timed on the context stream (the fastest of `reps`
// launches; each launch rewrites the same values into the store), and the
// fastest shape of each class is kept for the process.
static const int kSplits[4] = {1, 2, 4, 8}, kTpbs[3] = {64, 128, 256};

int eri_tune(myqc_ctx *ctx, int reps, int *shape_out, double *ms_out) {
    if (!ctx->have_eri) return set_error(ctx, MYQC_E_STATE, "myqc_eri_tune: call myqc_eri_build first");
    if (reps < 1 || reps > 100) return set_error(ctx, MYQC_E_ARG, "myqc_eri_tune: reps=%d (1..100)", reps);
    hipEvent_t e0, e1;
    MYQC_HIP(ctx, hipEventCreate(&e0));
    MYQC_HIP(ctx, hipEventCreate(&e1));
    int rc = 0;
    EriShape best[6];
    double best_ms[6];
    for (int c = 0; c < 6 && rc == 0; ++c) {
        best[c] = class_shape(c);
        best_ms[c] = 0.0;
        if (ctx->cls[c].nq == 0) continue;
        best_ms[c] = 1e300;
        for (int sp : kSplits)
            for (int tpb : kTpbs) {
                const EriShape sh{sp, tpb};
                double t = 1e300;
                for (int r = 0; r < reps && rc == 0; ++r) {
                    float ms = 0.0f;
                    if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "event");
                    launch_class<0>(ctx, c, nullptr, ctx->stream, sh);
                    if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, ctx->stream) != hipSuccess ||
                                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                        rc = set_error(ctx, MYQC_E_HIP, "myqc_eri_tune: class %d launch failed", c);
                    t = std::min(t, (double)ms);
                }
                if (t < best_ms[c]) {
                    best_ms[c] = t;
                    best[c] = sh;
                }
            }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(g_shape_mu);
        for (int c = 0; c < 6; ++c)
            if (ctx->cls[c].nq > 0) g_shape[c] = best[c];
    }
    for (int c = 0; c < 6; ++c) {
        if (shape_out) {
            shape_out[2 * c] = best[c].split;
            shape_out[2 * c + 1] = best[c].tpb;
        }
        if (ms_out) ms_out[c] = best_ms[c];
    }
    return 0;
}

// sets the launch shapes directly (a shape of 0 lanes keeps that class's)
int eri_set_shapes(myqc_ctx *ctx, const int *shape) {
    for (int c = 0; c < 6; ++c) {
        const int sp = shape[2 * c], tpb = shape[2 * c + 1];
        if (sp == 0) continue;
        if ((sp != 1 && sp != 2 && sp != 4 && sp != 8) || (tpb != 64 && tpb != 128 && tpb != 256))
            return set_error(ctx, MYQC_E_ARG, "myqc_eri_set_shapes: class %d: %d lanes, %d threads", c, sp, tpb);
    }
    std::lock_guard<std::mutex> g(g_shape_mu);
    for (int c = 0; c < 6; ++c)
        if (shape[2 * c] != 0) g_shape[c] = EriShape{shape[2 * c], shape[2 * c + 1]};
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
            launch_class<1>(ctx, c, d_XX, ctx->stream, EriShape{});
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
    ctx->ftT = build_ftT(Ftab);
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

extern "C" int myqc_eri_tune(myqc_ctx *ctx, int reps, int *shapes, double *ms) {
    if (!ctx) return MYQC_E_ARG;
    return eri_tune(ctx, reps, shapes, ms);
}

extern "C" int myqc_eri_set_shapes(myqc_ctx *ctx, const int *shapes) {
    if (!ctx || !shapes) return MYQC_E_ARG;
    return eri_set_shapes(ctx, shapes);
}

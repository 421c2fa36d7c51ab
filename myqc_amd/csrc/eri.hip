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
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>

#include "context.h"
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

int dalloc(myqc_ctx *ctx, void **p, size_t bytes) {
    if (*p) {
        hipFree(*p);
        *p = nullptr;
    }
    if (bytes == 0) return 0;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(ctx, MYQC_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    return 0;
}

hipError_t d2h(myqc_ctx *ctx, void *dst, const void *src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ctx->stream);
}

void dfree(void **p) {
    if (*p) hipFree(*p);
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
    const double *ftT;
    const int *sorted_b;
    const int *sorted_k;
    const long long *cum;
    int nbra;
    long long nq;
    int same_class;
    double *v64;
    float *v32;
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
    for (int i = threadIdx.x; i < FT_NT * FT_NM; i += blockDim.x) s_ft[i] = ftT[i];
    __syncthreads();
}

// MODE 0: store weighted block values;  MODE 1: scatter into dense XX (8-fold)
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
    const ShellPair bra = a.pairs[pb], ket = a.pairs[pk];
    QuartetAcc<LA, LB, LC, LD> acc;
    quartet_prims<LA, LB, LC, LD>(bra, ket, a.prims, a.prims, s_ft, acc);
    double CD[3] = {ket.AB[0], ket.AB[1], ket.AB[2]};
    double out[BLK];
    quartet_hrr<LA, LB, LC, LD>(acc, bra.AB, CD, out);
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
                    if (MODE == 0) {
                        const double v = out[e] * w;
                        a.v64[t * BLK + e] = v;
                        if (a.v32) a.v32[t * BLK + e] = (float)v;
                    } else {
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

void free_eri(myqc_ctx *ctx) {
    for (auto &c : ctx->cls) {
        dfree((void **)&c.d_cum);
        dfree((void **)&c.d_v64);
        dfree((void **)&c.d_v32);
        c.nq = 0;
    }
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
    // sort each class by Schwarz factor, largest first (ties by pair id)
    for (int c = 0; c < 3; ++c) {
        auto &ids = ctx->pair_ids[c];
        std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) { return ctx->schwarz[x] > ctx->schwarz[y]; });
        if (!ids.empty())
            MYQC_HIP(ctx, hipMemcpyAsync(ctx->d_sorted[c], ids.data(), sizeof(int) * ids.size(),
                                         hipMemcpyHostToDevice, ctx->stream));
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
    for (size_t r = 0; r < B.size(); ++r) {
        const double qb = ctx->schwarz[B[r]];
        long long ub;
        if (tol <= 0.0) ub = (long long)K.size();
        else if (qb <= 0.0) ub = 0;
        else {
            const double need = tol / qb;
            // number of kets with qk >= need (qk sorted descending)
            ub = std::lower_bound(qk.begin(), qk.end(), need, [](double a, double b) { return a >= b; }) - qk.begin();
        }
        long long cnt = (cb == ck) ? std::max(0LL, ub - (long long)r) : ub;
        cum[r + 1] = cum[r] + cnt;
    }
}

template <int CB, int CK, int MODE>
static void launch_eri(myqc_ctx *ctx, const QuartetClass &qc, double *v64, float *v32, double *XX) {
    EriArgs a;
    a.pairs = ctx->d_pairs;
    a.prims = ctx->d_prims;
    a.ftT = ctx->d_ftT;
    a.sorted_b = ctx->d_sorted[CB];
    a.sorted_k = ctx->d_sorted[CK];
    a.cum = qc.d_cum;
    a.nbra = qc.nbra;
    a.nq = qc.nq;
    a.same_class = (CB == CK);
    a.v64 = v64;
    a.v32 = v32;
    a.XX = XX;
    a.n = ctx->n;
    const long long nblk = (qc.nq + 255) / 256;
    k_eri<CB, CK, MODE><<<(unsigned)nblk, 256, 0, ctx->stream>>>(a);
}

template <int MODE>
static void launch_class(myqc_ctx *ctx, int c, double *v64, float *v32, double *XX) {
    const QuartetClass &qc = ctx->cls[c];
    if (qc.nq == 0) return;
    switch (c) {
        case 0: launch_eri<0, 0, MODE>(ctx, qc, v64, v32, XX); break;
        case 1: launch_eri<1, 0, MODE>(ctx, qc, v64, v32, XX); break;
        case 2: launch_eri<1, 1, MODE>(ctx, qc, v64, v32, XX); break;
        case 3: launch_eri<2, 0, MODE>(ctx, qc, v64, v32, XX); break;
        case 4: launch_eri<2, 1, MODE>(ctx, qc, v64, v32, XX); break;
        case 5: launch_eri<2, 2, MODE>(ctx, qc, v64, v32, XX); break;
    }
}

// Static FP64 operation count per primitive quartet and per contracted
// quartet of each class (FMA = 2), for the roofline's algorithmic flops.
// Counted from the recursions in eri_math.h; see eri_opcount() in the tests.
static const double kPrimFlops[6] = {70, 95, 160, 165, 290, 540};
static const double kContrFlops[6] = {2, 6, 30, 30, 110, 400};

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
        // every STO-3G shell pair has 9 primitive pairs
        ctx->eri_flops += (double)qc.nq * (81.0 * kPrimFlops[c] + kContrFlops[c]);
    }
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int eri_store(myqc_ctx *ctx, double tol, bool want32) {
    MYQC_CHECK(eri_schwarz(ctx));
    free_eri(ctx);
    MYQC_CHECK(build_classes(ctx, tol));
    hipEvent_t ev;
    MYQC_CHECK(timer_begin(ctx, T_ERI, &ev));
    for (int c = 0; c < 6; ++c) {
        QuartetClass &qc = ctx->cls[c];
        if (qc.nq == 0) continue;
        MYQC_CHECK(dalloc(ctx, (void **)&qc.d_v64, sizeof(double) * qc.nq * qc.blk));
        if (want32) MYQC_CHECK(dalloc(ctx, (void **)&qc.d_v32, sizeof(float) * qc.nq * qc.blk));
        launch_class<0>(ctx, c, qc.d_v64, want32 ? qc.d_v32 : nullptr, nullptr);
        MYQC_HIP(ctx, hipGetLastError());
    }
    MYQC_CHECK(timer_end(ctx, T_ERI, ev, ctx->eri_flops));
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
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
    for (int c = 0; c < 6; ++c) saved[c] = ctx->cls[c];
    for (int c = 0; c < 6; ++c) ctx->cls[c] = QuartetClass{};
    int rc = build_classes(ctx, tol);
    if (rc == 0) {
        hipEvent_t ev;
        rc = timer_begin(ctx, T_ERI, &ev);
        for (int c = 0; c < 6 && rc == 0; ++c) {
            launch_class<1>(ctx, c, nullptr, nullptr, d_XX);
            if (hipGetLastError() != hipSuccess) rc = set_error(ctx, MYQC_E_HIP, "ERI dense-fill launch failed");
        }
        if (rc == 0) rc = timer_end(ctx, T_ERI, ev, ctx->eri_flops);
        if (rc == 0 && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = set_error(ctx, MYQC_E_HIP, "ERI dense fill failed");
    }
    for (int c = 0; c < 6; ++c) dfree((void **)&ctx->cls[c].d_cum);
    for (int c = 0; c < 6; ++c) ctx->cls[c] = saved[c];
    ctx->have_eri = had;
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
    free_eri(ctx);
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

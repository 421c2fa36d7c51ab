// context.h -- the myqc_ctx object behind the C ABI: one device, one stream,
// device-resident basis data, stored integrals and SCF work arrays.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <array>
#include <set>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <string>
#include <unordered_map>
#include <cstdint>
#include <vector>

#include "../../include/myqc_hip.h"
#include "digest_plan.h"
#include "eri_math.h"
#include "system.h"

namespace myqc {

// Unique shell quartets of one (bra pair class, ket pair class), enumerated
// from Schwarz-sorted pair lists: bra rank r owns ket ranks [k0(r), k0(r)+cnt(r)).
struct QuartetClass {
    int cb = 0, ck = 0;
    long long nq = 0;
    int nbra = 0;
    int blk = 0;
    long long *d_cum = nullptr;   // nbra + 1 prefix sums of counts
};

struct ScfBuffers;
struct ManyBuffers;
struct BatchState;
struct MlpState;

enum TimerCat { T_ERI = 0, T_JK = 1, T_AO2MO = 2, T_MP2 = 3, T_NCAT = 4 };

struct EventRec {
    int cat;
    hipEvent_t a, b;
    double work;
};

}  // namespace myqc

struct myqc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // two side streams (default priority) and their fork/join events: the
    // ERI classes with few waves run beside the large ones
    hipStream_t side[2] = {nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[2] = {nullptr, nullptr};
    rocblas_handle blas = nullptr;
    std::string err;

    // system
    bool have_system = false;
    int n = 0, nnuc = 0;
    std::vector<int> Z;
    std::vector<double> xyz;              // [atom][3]
    std::vector<myqc::Shell> shells;
    std::vector<myqc::ShellPair> pairs;
    std::vector<myqc::PrimPair> prims;
    std::vector<int> pair_ids[3];         // per pair class, sorted by Schwarz value (descending)
    std::vector<double> schwarz;          // per pair
    std::vector<double> ftT;              // Ftab transposed: [Tk][m]
    myqc::ShellPair *d_pairs = nullptr;
    myqc::PrimPair *d_prims = nullptr;
    double *d_ftT = nullptr;
    int *d_sorted[3] = {nullptr, nullptr, nullptr};
    // per class, the primitive pairs of the sorted pairs as structure of
    // arrays: field f of primitive j of rank r at [(j * kKetFields + f) * np + r]
    double *d_ksoa[3] = {nullptr, nullptr, nullptr};
    double *d_schwarz = nullptr;
    bool have_schwarz = false;

    // stored integrals: each unique screened shell quartet once, pre-weighted,
    // under its bra pair (the layout and schedule of digest_plan.h)
    bool have_eri = false;
    bool have_eri32 = false;
    double eri_tol = 0.0;
    myqc::QuartetClass cls[6];
    long long n_stored = 0;               // stored values (one per unique integral slot)
    long long n_quartets = 0;
    double eri_flops = 0.0;               // algorithmic FP64 work of the last build
    int ns = 0;                           // number of shells
    myqc::DigestPlan plan;                // layout and digestion schedule (host)
    long long *d_segoff = nullptr;        // [p*3 + ket class]: block of ket rank r at segoff + r*BLK
    int *d_rank = nullptr;                // rank of each pair within its class's sorted list
    double *d_x64 = nullptr;              // stored values
    float *d_x32 = nullptr;
    // digestion schedule on the device
    int n_wg = 0;
    int *d_wg = nullptr;                  // kWgInts per workgroup
    int *d_unit = nullptr;                // kUnitInts per unit
    long long *d_unit_off = nullptr;      // [u][3] value offsets
    int *d_chunk = nullptr;               // {class, first rank, kets, ket-side slot} per column chunk
    int *d_row_ao = nullptr;              // AO of each partial row
    int *d_row_list = nullptr, *d_row_list_off = nullptr;   // per AO, its partial rows (CSR)
    int *d_jk_list = nullptr, *d_jk_list_off = nullptr;     // per rank block, its chunk slots (CSR)
    int *d_jk_group = nullptr;                              // the ket-side gather's groups
    int *d_jkoff = nullptr;               // [X*ns + Y]: J-ket block of pair (X,Y) if stored so, else -1
    double *d_jk_part = nullptr;          // ket-side J partial slots (kJkSlot values each, FP64 or FP32)
    double *d_jk = nullptr;               // ket-side J by class and rank
    int *d_meta[3] = {nullptr, nullptr, nullptr};   // per class, by Schwarz rank: aoC | aoD << 15
    float *d_qrank[3] = {nullptr, nullptr, nullptr}; // per class, the Schwarz factor of each rank (descending)
    double *d_dmax = nullptr;             // max |dD| of the current incremental pass (float bits in the first word)
    int *d_shell = nullptr;               // [s] = {ao, L, 0, 0}
    int *d_ao_shell = nullptr;            // shell of each AO
    double *d_partial = nullptr;          // partial rows x n (FP64, or FP32 in the incremental passes)
    double *d_Dpair = nullptr;            // density blocks of the kets, by class and rank (see jk.hip)
    float *d_Dpair32 = nullptr;

    // work arrays (n x n unless noted)
    double *d_D = nullptr, *d_A = nullptr, *d_G = nullptr;
    float *d_D32 = nullptr, *d_A32 = nullptr;
    int work_n = -1;

    // ao2mo output kept on the device for mp2
    double *d_iajb = nullptr;             // [i][a][j][b] with i fastest (o*v*o*v)
    double *d_q12 = nullptr, *d_q3 = nullptr;   // the RHF transform's quarter-step intermediates (kept)
    uint16_t *d_planes = nullptr;         // split-bf16 GEMM: hi/lo planes of the small operand
    size_t iajb_elems = 0;
    int iajb_nocc = 0, iajb_nv = 0;
    // UHF (ia|jb) sets AA, BB, AB, each [i][a][j][b] with i fastest, and
    // their (o1, nv1, o2, nv2)
    double *d_uhf[3] = {nullptr, nullptr, nullptr};
    int uhf_dims[3][4] = {};
    bool have_uhf = false;
    // same-spin (ab|ij) for UHF-CIS (alpha, beta), [a][b][i][j] with a
    // fastest, and their (nvrt, nocc)
    double *d_vvoo[2] = {nullptr, nullptr};
    int vvoo_dims[2][2] = {};
    bool have_vvoo = false;

    // last RHF run: purification steps and diagonalisation fallbacks
    myqc::ScfBuffers *scf = nullptr;       // RHF (and UHF alpha) work arrays, kept across calls
    myqc::ScfBuffers *scf_b = nullptr;     // UHF beta work arrays
    myqc::ManyBuffers *many = nullptr;     // the many-system SCF's shared buffers (this context leading)
    myqc::BatchState *batch = nullptr;     // batched small-molecule RHF (batch.hip)
    myqc::MlpState *mlp = nullptr;         // learned-guess network (mlp.hip)
    int mlp_gen = 0;                      // bumped whenever the network's buffers move
    int scf_purify = 0, scf_diag = 0;
    int scf_sp2_steps = 0;                // steps the last purification took

    // per-call device temporaries of the solve path (int1e, the learned
    // guess), kept across calls with their capacity: a repeated call makes no
    // hipMalloc/hipFree (hipFree waits for the whole device, which would
    // serialise contexts that run side by side on other streams)
    void *d_tmp[8] = {};

    // pinned host staging for small device-to-host reads (d2h)
    void *h_pin = nullptr;
    // this device's address of the file-scope scratch of the no-allocation
    // entry points (capi.hip), from hipGetSymbolAddress at myqc_init
    double *d_scratch = nullptr;
    // the rocBLAS/rocSOLVER workspace, owned by the context and handed to the
    // handle with rocblas_set_workspace: the library allocates nothing inside
    // a call (blas_reserve sizes it before each new problem shape)
    void *d_blasws = nullptr;
    size_t blasws_bytes = 0;
    std::set<std::array<int, 8>> blas_shapes;   // call shapes the workspace has been sized for

    // timing
    bool timing = false;
    std::vector<myqc::EventRec> events;
    long long t_n[myqc::T_NCAT] = {0, 0, 0, 0};
    double t_ms[myqc::T_NCAT] = {0, 0, 0, 0};
    double t_work[myqc::T_NCAT] = {0, 0, 0, 0};
};

namespace myqc {

int set_error(myqc_ctx *ctx, int code, const char *fmt, ...);

#define MYQC_HIP(ctx, call)                                                                          \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                        \
            return myqc::set_error(ctx, MYQC_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                   __FILE__, __LINE__);                                              \
    } while (0)

#define MYQC_BLAS(ctx, call)                                                                         \
    do {                                                                                             \
        rocblas_status s_ = (call);                                                                  \
        if (s_ != rocblas_status_success)                                                            \
            return myqc::set_error(ctx, MYQC_E_HIP, "%s failed: rocblas status %d (%s:%d)", #call,     \
                                   (int)s_, __FILE__, __LINE__);                                     \
    } while (0)

#define MYQC_CHECK(x)          \
    do {                       \
        int rc_ = (x);         \
        if (rc_ != 0) return rc_; \
    } while (0)

// bytes of the pinned staging buffer (ctx->h_pin)
constexpr size_t kPinBytes = 4096;
// records (i,j) the no-allocation MP2 entry point takes at most (nocc <= 256)
constexpr size_t kNoAllocRecords = 65536;

// device-to-host copy ordered on ctx->stream, complete at return
hipError_t d2h(myqc_ctx *ctx, void *dst, const void *src, size_t bytes);

// device allocation helper: keeps a buffer whose capacity covers `bytes`,
// otherwise frees it and allocates `bytes`
int dalloc(myqc_ctx *ctx, void **p, size_t bytes);
void dfree(void **p);
// capacity of a dalloc allocation (0 for an unknown address)
size_t dcapacity(const void *p);

// Grows the context's rocBLAS workspace to what `calls` needs: the calls run
// in rocBLAS's device-memory size-query mode (nothing executes), and the
// workspace is reallocated and handed to the handle only when it must grow.
template <typename F>
int blas_reserve(myqc_ctx *ctx, F calls);
template <typename F>
int blas_reserve(myqc_ctx *ctx, F calls) {
    if (rocblas_start_device_memory_size_query(ctx->blas) != rocblas_status_success)
        return set_error(ctx, MYQC_E_HIP, "rocblas_start_device_memory_size_query failed");
    calls();
    size_t need = 0;
    if (rocblas_stop_device_memory_size_query(ctx->blas, &need) != rocblas_status_success)
        return set_error(ctx, MYQC_E_HIP, "rocblas_stop_device_memory_size_query failed");
    if (ctx->d_blasws && need <= ctx->blasws_bytes) return 0;
    // 1 MiB granules, at least 4 MiB; the old workspace may still be read by
    // queued work, so the stream drains before it is replaced
    const size_t gran = size_t(1) << 20;
    size_t bytes = ((std::max(need, size_t(4) << 20) + gran - 1) / gran) * gran;
    MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    MYQC_CHECK(dalloc(ctx, &ctx->d_blasws, bytes));
    ctx->blasws_bytes = bytes;
    if (rocblas_set_workspace(ctx->blas, ctx->d_blasws, bytes) != rocblas_status_success)
        return set_error(ctx, MYQC_E_HIP, "rocblas_set_workspace failed");
    return 0;
}

// slots of ctx->d_tmp
enum TmpSlot { TMP_S = 0, TMP_H, TMP_ATOMS, TMP_GPAIRS, TMP_GATOMS, TMP_GDENS, TMP_GMOL, TMP_NSLOT };
static_assert(TMP_NSLOT <= 8, "d_tmp slots");
// a temporary of at least `bytes` in slot `slot`, kept by the context
template <typename T>
int dtmp(myqc_ctx *ctx, int slot, T **p, size_t bytes) {
    MYQC_CHECK(dalloc(ctx, &ctx->d_tmp[slot], bytes));
    *p = static_cast<T *>(ctx->d_tmp[slot]);
    return 0;
}

// timing helpers: bracket a launch sequence on ctx->stream
int timer_begin(myqc_ctx *ctx, int cat, hipEvent_t *ev);
int timer_end(myqc_ctx *ctx, int cat, hipEvent_t ev, double work);

// Host-side phase timing to stderr when MYQC_HOST_TIMING is set (profiling
// of the setup work around the kernels).
struct HostClock {
    bool on;
    double t0;
    static double now() {
        timespec ts;
        clock_gettime(CLOCK_MONOTONIC, &ts);
        return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
    }
    explicit HostClock(myqc_ctx *) : on(getenv("MYQC_HOST_TIMING") != nullptr), t0(now()) {}
    void mark(const char *what) {
        if (!on) return;
        const double t = now();
        fprintf(stderr, "[host] %-40s %8.3f ms\n", what, t - t0);
        t0 = t;
    }
};

// internals shared between translation units
int eri_schwarz(myqc_ctx *ctx);
int eri_store(myqc_ctx *ctx, double tol, bool want32);
int eri_fill_dense(myqc_ctx *ctx, double tol, double *d_XX);
// G(D) into d_G (precision FP64), or d_G += G32(D - Dprev) (precision FP32;
// Dprev may be NULL for D itself)
int digest_G(myqc_ctx *ctx, const double *d_D, double *d_G, int precision, const double *d_Dprev = nullptr);
int digest_G_uhf(myqc_ctx *ctx, const double *d_Da, const double *d_Db, double *d_Ga, double *d_Gb);
int ensure_work(myqc_ctx *ctx);
void free_eri(myqc_ctx *ctx, bool keep_memory = false);
void free_scf(myqc_ctx *ctx);
void free_many(myqc_ctx *ctx);
void free_batch(myqc_ctx *ctx);
void free_mlp(myqc_ctx *ctx);

}  // namespace myqc

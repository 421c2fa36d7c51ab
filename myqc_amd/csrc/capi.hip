// capi.hip -- context lifetime, errors and kernel timers of the C ABI.
#include <cstring>

#include <rocblas/rocblas.h>

#include "context.h"

using namespace myqc;

// Device scratch of the entry points that allocate nothing
// (myqc_mp2_rhf_records_noalloc): 2 doubles per (i,j) record, for at most
// kNoAllocRecords records.  Each device the library runs on holds one copy
// until the process ends; every call on that device shares it.  Host code
// uses only the address hipGetSymbolAddress gives for it (ctx->d_scratch),
// never the array's name.
static __device__ __align__(256) double scratch[2 * kNoAllocRecords];

extern "C" int myqc_init(int device, myqc_ctx **out) {
    if (!out) return MYQC_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return MYQC_E_DEVICE;
    if (device < 0 || device >= ndev) return MYQC_E_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MYQC_E_DEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return MYQC_E_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return MYQC_E_DEVICE;
    myqc_ctx *ctx = new myqc_ctx();
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return MYQC_E_HIP;
    }
    if (rocblas_create_handle(&ctx->blas) != rocblas_status_success ||
        rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success) {
        hipStreamDestroy(ctx->stream);
        delete ctx;
        return MYQC_E_HIP;
    }
    // this device's copy of the no-allocation scratch (the device is current)
    if (hipGetSymbolAddress(reinterpret_cast<void **>(&ctx->d_scratch), HIP_SYMBOL(scratch)) != hipSuccess ||
        !ctx->d_scratch) {
        myqc_finalize(ctx);
        return MYQC_E_HIP;
    }
    // pinned staging for small reads, and a context-owned rocBLAS workspace
    // from the start (rocBLAS then never allocates inside a call)
    if (hipHostMalloc(&ctx->h_pin, kPinBytes, hipHostMallocDefault) != hipSuccess ||
        blas_reserve(ctx, [] {}) != 0) {
        myqc_finalize(ctx);
        return MYQC_E_HIP;
    }
    *out = ctx;
    return 0;
}

extern "C" int myqc_finalize(myqc_ctx *ctx) {
    if (!ctx) return MYQC_E_ARG;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    free_eri(ctx);
    free_scf(ctx);
    free_many(ctx);
    free_batch(ctx);
    free_mlp(ctx);
    void **ptrs[] = {(void **)&ctx->d_pairs, (void **)&ctx->d_prims, (void **)&ctx->d_ftT, (void **)&ctx->d_schwarz,
                     (void **)&ctx->d_D,     (void **)&ctx->d_A,     (void **)&ctx->d_G,   (void **)&ctx->d_D32,
                     (void **)&ctx->d_A32,   (void **)&ctx->d_iajb, (void **)&ctx->d_planes, (void **)&ctx->d_dmax,
                     (void **)&ctx->d_q12,   (void **)&ctx->d_q3};
    for (auto p : ptrs) dfree(p);
    for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_sorted[c]);
    for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_ksoa[c]);
    for (int c = 0; c < 3; ++c) dfree((void **)&ctx->d_uhf[c]);
    for (int c = 0; c < 2; ++c) dfree((void **)&ctx->d_vvoo[c]);
    for (auto &t : ctx->d_tmp) dfree(&t);
    for (auto &e : ctx->events) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    for (int k = 0; k < 2; ++k) {
        if (ctx->side[k]) hipStreamDestroy(ctx->side[k]);
        if (ctx->join_ev[k]) hipEventDestroy(ctx->join_ev[k]);
    }
    if (ctx->fork_ev) hipEventDestroy(ctx->fork_ev);
    if (ctx->blas) rocblas_destroy_handle(ctx->blas);
    dfree(&ctx->d_blasws);
    if (ctx->h_pin) hipHostFree(ctx->h_pin);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

extern "C" const char *myqc_last_error(const myqc_ctx *ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

extern "C" int myqc_timer_reset(myqc_ctx *ctx, int enable) {
    if (!ctx) return MYQC_E_ARG;
    hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->events) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    ctx->events.clear();
    for (int c = 0; c < T_NCAT; ++c) {
        ctx->t_n[c] = 0;
        ctx->t_ms[c] = 0.0;
        ctx->t_work[c] = 0.0;
    }
    ctx->timing = enable != 0;
    return 0;
}

extern "C" int myqc_timer_get(myqc_ctx *ctx, int which, long long *launches, double *ms, double *work) {
    if (!ctx || which < 0 || which >= T_NCAT) return MYQC_E_ARG;
    // resolve pending event pairs
    if (!ctx->events.empty()) {
        MYQC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (auto &e : ctx->events) {
            float t = 0.0f;
            MYQC_HIP(ctx, hipEventElapsedTime(&t, e.a, e.b));
            ctx->t_n[e.cat] += 1;
            ctx->t_ms[e.cat] += t;
            ctx->t_work[e.cat] += e.work;
            hipEventDestroy(e.a);
            hipEventDestroy(e.b);
        }
        ctx->events.clear();
    }
    if (launches) *launches = ctx->t_n[which];
    if (ms) *ms = ctx->t_ms[which];
    if (work) *work = ctx->t_work[which];
    return 0;
}

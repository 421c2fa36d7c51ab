// mfma_probe.hip -- TEST INFRASTRUCTURE: one v_mfma_f64_16x16x4f64 (and one
// v_mfma_f32_16x16x32_bf16) on caller-given A and B, every lane's outputs
// recorded, so a test can check the lane layouts the SP2 and split-bf16 GEMM
// kernels assume.  All accesses are within the 64-lane arrays below.
#include <hip/hip_runtime.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void k_probe(const double *A, const double *B, double *raw) {
    const int l = threadIdx.x;
    // assumed: lane l feeds A[l%16][l/16] and B[l/16][l%16]  (A row-major 16x4, B row-major 4x16)
    const double a = A[(l % 16) * 4 + l / 16];
    const double b = B[(l / 16) * 16 + l % 16];
    d4_t c = {0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    for (int v = 0; v < 4; ++v) raw[l * 4 + v] = c[v];
}

extern "C" int mfma_probe(const double *A, const double *B, double *raw) {
    double *dA = nullptr, *dB = nullptr, *dR = nullptr;
    if (hipMalloc(&dA, 64 * 8) != hipSuccess || hipMalloc(&dB, 64 * 8) != hipSuccess ||
        hipMalloc(&dR, 256 * 8) != hipSuccess)
        return -1;
    int rc = 0;
    if (hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = -2;
    if (rc == 0) {
        k_probe<<<1, 64>>>(dA, dB, dR);
        if (hipMemcpy(raw, dR, 256 * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = -3;
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dR);
    return rc;
}

// assumed: lane l feeds A[l%16][8(l/16)+e] and B[8(l/16)+e][l%16], e = 0..7
// (A row-major 16x32, B row-major 32x16, values exact in bf16)
__global__ __launch_bounds__(64) void k_probe_bf16(const float *A, const float *B, float *raw) {
    const int l = threadIdx.x;
    bf16x8 a, b;
    for (int e = 0; e < 8; ++e) {
        a[e] = (__bf16)A[(l % 16) * 32 + 8 * (l / 16) + e];
        b[e] = (__bf16)B[(8 * (l / 16) + e) * 16 + l % 16];
    }
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    for (int v = 0; v < 4; ++v) raw[l * 4 + v] = c[v];
}

extern "C" int mfma_probe_bf16(const float *A, const float *B, float *raw) {
    float *dA = nullptr, *dB = nullptr, *dR = nullptr;
    if (hipMalloc(&dA, 512 * 4) != hipSuccess || hipMalloc(&dB, 512 * 4) != hipSuccess ||
        hipMalloc(&dR, 256 * 4) != hipSuccess)
        return -1;
    int rc = 0;
    if (hipMemcpy(dA, A, 512 * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dB, B, 512 * 4, hipMemcpyHostToDevice) != hipSuccess)
        rc = -2;
    if (rc == 0) {
        k_probe_bf16<<<1, 64>>>(dA, dB, dR);
        if (hipMemcpy(raw, dR, 256 * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = -3;
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dR);
    return rc;
}

// Measured peaks of the pipes the E-step runs on (dg_device_peaks): the FP16
// MFMA pipe (v_mfma_f32_32x32x16_f16), the transcendental pipe (v_exp_f32)
// and vector FP64 FMA.  Each kernel keeps many independent dependency chains
// per wave so the pipe, not latency, bounds it; the grid is several waves per
// SIMD on every CU.  Results are written out so nothing is dead code.
#include <hip/hip_runtime.h>

#include "dg_internal.h"

namespace {

typedef _Float16 ph8 __attribute__((ext_vector_type(8)));
typedef float pf16 __attribute__((ext_vector_type(16)));

constexpr int kMfmaChains = 4;      // accumulators per wave (4 x 16 VGPRs)
constexpr int kExpChains = 8;
constexpr int kF64Chains = 8;

__global__ __launch_bounds__(256) void k_peak_mfma(int iters, float seed, float *out) {
    const int l = threadIdx.x & 63;
    ph8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = (_Float16)(seed * (l + i));
        b[i] = (_Float16)(seed * (l - i));
    }
    pf16 c[kMfmaChains];
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k)
#pragma unroll
        for (int r = 0; r < 16; ++r) c[k][r] = 0.f;
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kMfmaChains; ++k)
            c[k] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c[k], 0, 0, 0);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kMfmaChains; ++k)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += c[k][r];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_exp(int iters, float seed, float *out) {
    float x[kExpChains];
#pragma unroll
    for (int k = 0; k < kExpChains; ++k) x[k] = seed * (float)(threadIdx.x + k) * 1e-3f;
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kExpChains; ++k) x[k] = __builtin_amdgcn_exp2f(-x[k]);
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < kExpChains; ++k) s += x[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_peak_f64(int iters, double seed, double *out) {
    double x[kF64Chains];
    const double m = 1.0 - 1e-9 * seed, a = 1e-12 * seed;
#pragma unroll
    for (int k = 0; k < kF64Chains; ++k) x[k] = seed * (threadIdx.x + k);
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kF64Chains; ++k) x[k] = fma(x[k], m, a);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kF64Chains; ++k) s += x[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

// Times one launch of each kernel after a warm-up launch; grids of `wgs`
// 256-thread workgroups.  Returns DG_OK and the three rates, or a DG_E* code.
int dg_peaks_run(dg_ctx *c, int wgs, double *mfma_tflops, double *exp_per_s, double *f64_tflops) {
    const int nthreads = wgs * 256;
    float *out = nullptr;
    double *out64 = nullptr;
    if (hipMalloc((void **)&out, sizeof(float) * nthreads) != hipSuccess) return DG_ENOMEM;
    if (hipMalloc((void **)&out64, sizeof(double) * nthreads) != hipSuccess) {
        (void)hipFree(out);
        return DG_ENOMEM;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    auto timed = [&](auto launch) {
        launch();                                    // warm-up
        (void)hipEventRecord(e0, c->stream);
        launch();
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return (double)ms * 1e-3;
    };
    const int itM = 4096, itE = 16384, itF = 16384;
    const double tM = timed([&] {
        hipLaunchKernelGGL(k_peak_mfma, dim3(wgs), dim3(256), 0, c->stream, itM, 1e-3f, out);
    });
    const double tE = timed([&] {
        hipLaunchKernelGGL(k_peak_exp, dim3(wgs), dim3(256), 0, c->stream, itE, 1.0f, out);
    });
    const double tF = timed([&] {
        hipLaunchKernelGGL(k_peak_f64, dim3(wgs), dim3(256), 0, c->stream, itF, 1.0, out64);
    });
    const hipError_t err = hipGetLastError();
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(out);
    (void)hipFree(out64);
    if (err != hipSuccess) return dg_hip_check(c, err, "dg_device_peaks");
    const double waves = (double)wgs * 4.0;
    *mfma_tflops = waves * itM * kMfmaChains * (32.0 * 32.0 * 16.0 * 2.0) / tM * 1e-12;
    *exp_per_s = (double)nthreads * itE * kExpChains / tE;
    *f64_tflops = (double)nthreads * itF * kF64Chains * 2.0 / tF * 1e-12;
    return DG_OK;
}

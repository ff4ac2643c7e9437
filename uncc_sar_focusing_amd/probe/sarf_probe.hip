// sarf_probe.hip -- microbenchmarks of the gfx950 rates the backprojection
// roofline is priced against: FP32 FMA issue rate (non-packed, as every
// kernel of libsarf is built), the v_sin/v_cos transcendental rate, and a
// streaming copy of known size (the reference point for HBM byte counters).
// Built as its own library (libsarf_probe.so); not part of the focusing API.
#include <hip/hip_runtime.h>

#include <cstddef>

namespace {

constexpr int kThreads = 256;
constexpr int kFmaChains = 16;
constexpr int kTransChains = 8;

__global__ __launch_bounds__(kThreads) void k_fma(float *out, int iters, float a, float b) {
    float x[kFmaChains];
#pragma unroll
    for (int i = 0; i < kFmaChains; i++) x[i] = threadIdx.x * 1e-3f + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < kFmaChains; i++) x[i] = fmaf(x[i], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < kFmaChains; i++) s += x[i];
    if (s == -1.f) out[blockIdx.x * kThreads + threadIdx.x] = s;  // keeps the chains live
}

__global__ __launch_bounds__(kThreads) void k_trans(float *out, int iters) {
    float x[kTransChains];
#pragma unroll
    for (int i = 0; i < kTransChains; i++) x[i] = (threadIdx.x + 7 * i) * 1e-4f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < kTransChains; i += 2) {
            x[i] = __builtin_amdgcn_sinf(x[i]);          // v_sin_f32 (revolutions)
            x[i + 1] = __builtin_amdgcn_cosf(x[i + 1]);  // v_cos_f32
        }
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < kTransChains; i++) s += x[i];
    if (s == -1.f) out[blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void k_copy(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, size_t n) {
    const size_t stride = (size_t)gridDim.x * kThreads;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

int blocks_for(int device, int per_cu) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus * per_cu;
}

// Times `launch` once after one warm-up; returns ms or a negative value.
template <class F>
float time_ms(F launch) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.f;
    }
    launch();
    (void)hipEventRecord(a, nullptr);
    launch();
    (void)hipEventRecord(b, nullptr);
    float ms = -1.f;
    if (hipEventSynchronize(b) != hipSuccess || hipGetLastError() != hipSuccess ||
        hipEventElapsedTime(&ms, a, b) != hipSuccess)
        ms = -1.f;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

}  // namespace

extern "C" {

// Sustained FP32 FMA rate in TFLOP/s (2 flops per FMA), non-packed.
int sarf_probe_fma(int device, double *tflops) {
    if (!tflops || hipSetDevice(device) != hipSuccess) return -1;
    const int blocks = blocks_for(device, 32), iters = 4096;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * blocks * kThreads) != hipSuccess) return -4;
    const float ms = time_ms([&] {
        hipLaunchKernelGGL(k_fma, dim3(blocks), dim3(kThreads), 0, nullptr, out, iters, 0.9999999f,
                           1e-7f);
    });
    (void)hipFree(out);
    if (ms <= 0.f) return -3;
    *tflops = 2.0 * kFmaChains * (double)iters * blocks * kThreads / (ms * 1e-3) / 1e12;
    return 0;
}

// Sustained v_sin_f32 / v_cos_f32 rate in T ops/s.
int sarf_probe_trans(int device, double *tops) {
    if (!tops || hipSetDevice(device) != hipSuccess) return -1;
    const int blocks = blocks_for(device, 32), iters = 1024;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * blocks * kThreads) != hipSuccess) return -4;
    const float ms = time_ms([&] {
        hipLaunchKernelGGL(k_trans, dim3(blocks), dim3(kThreads), 0, nullptr, out, iters);
    });
    (void)hipFree(out);
    if (ms <= 0.f) return -3;
    *tops = (double)kTransChains * iters * blocks * kThreads / (ms * 1e-3) / 1e12;
    return 0;
}

// Copy of `bytes` (multiple of 16) from one buffer to another, `reps` timed
// launches after one warm-up; GB/s counts bytes read + bytes written.
// Each launch moves exactly `bytes` in and `bytes` out.
int sarf_probe_copy(int device, size_t bytes, int reps, double *gbs) {
    if (!gbs || reps < 1 || bytes < 16 || bytes % 16 || hipSetDevice(device) != hipSuccess)
        return -1;
    float4 *src = nullptr, *dst = nullptr;
    if (hipMalloc(&src, bytes) != hipSuccess) return -4;
    if (hipMalloc(&dst, bytes) != hipSuccess) {
        (void)hipFree(src);
        return -4;
    }
    (void)hipMemset(src, 0, bytes);
    const size_t n = bytes / 16;
    const int blocks = blocks_for(device, 8);
    const float ms = time_ms([&] {
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kThreads), 0, nullptr, src, dst, n);
    });
    (void)hipFree(src);
    (void)hipFree(dst);
    if (ms <= 0.f) return -3;
    *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return 0;
}

}  // extern "C"

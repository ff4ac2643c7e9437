// sarf_synth.hip -- point-scatterer phase history on the GPU.
//
// Data model of matlab/formImagePointTargetBP.m:63-81:
//   ph[n,k] = sum_s A_s exp(-j 4 pi f_k (|a_n - p_s| - R0_n) / c),
//   f_k = f0_n + k df.
// The phase argument is formed and reduced (to revolutions in [-0.5, 0.5])
// in FP64, the rotation uses v_sin/v_cos on the reduced FP32 argument, and
// the sum over scatterers is FP32.  A workgroup handles 256 samples of one
// pulse; scatterer ranges for the pulse are computed once per chunk of 256
// scatterers into LDS.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

__global__ __launch_bounds__(256) void synth_ph(int K, int Np, int S,
                                                const double *__restrict__ scat,
                                                const float2 *__restrict__ amp,
                                                const double *__restrict__ ant,
                                                const double *__restrict__ r0,
                                                const double *__restrict__ f0, double df,
                                                float2 *__restrict__ ph) {
    __shared__ double sdr[256];
    __shared__ float2 samp[256];
    const int n = blockIdx.y;
    const int k = blockIdx.x * 256 + threadIdx.x;
    const double ax = ant[3 * n], ay = ant[3 * n + 1], az = ant[3 * n + 2];
    const double fk = f0[n] + k * df;
    const double scale = -2.0 * fk / kClight;  // revolutions per metre of dR
    float accr = 0.f, acci = 0.f;
    for (int s0 = 0; s0 < S; s0 += 256) {
        __syncthreads();
        const int s = s0 + threadIdx.x;
        if (s < S) {
            const double dx = ax - scat[3 * s], dy = ay - scat[3 * s + 1], dz = az - scat[3 * s + 2];
            sdr[threadIdx.x] = sqrt(dx * dx + dy * dy + dz * dz) - r0[n];
            samp[threadIdx.x] = amp[s];
        }
        __syncthreads();
        const int ns = min(256, S - s0);
        for (int i = 0; i < ns; i++) {
            double rev = scale * sdr[i];
            rev -= rint(rev);
            const float r = (float)rev;
            const float sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);
            const float2 A = samp[i];
            accr = fmaf(A.x, cs, fmaf(-A.y, sn, accr));
            acci = fmaf(A.x, sn, fmaf(A.y, cs, acci));
        }
    }
    if (k < K) {
        float2 *o = ph + (size_t)n * K + k;
        const float2 old = *o;
        *o = make_float2(old.x + accr, old.y + acci);
    }
}

// ---------------------------------------------------------------------------
// Many scatterers: type-1 non-uniform FFT per pulse.
//   ph[k] = sum_s c_s e^{-j 2 pi k u_s},  u_s = dR_s / maxWr (cycles/sample),
//   c_s = A_s e^{-j 2 pi f0 2 dR_s / c}.
// With k = k' + kc (kc = K/2) the e^{-j 2 pi kc u_s} factor joins c_s and
// k' is centred.  Each scatterer is spread onto an M-point grid (M >= 2K) at
// x_s = M u_s with the "exponential of semicircle" kernel
//   phi(x) = exp(beta (sqrt(1 - (2x/w)^2) - 1)),  |x| < w/2,  w = 8,
// b = sum_s c_s phi(x_s - m); then B_l = sum_m b_m e^{-j 2 pi l m / M} and
// ph[k' + kc] = B_{k' mod M} / phi_hat(k'/M), phi_hat by Gauss-Legendre
// quadrature on the host.  Spreading runs in LDS (one workgroup per pulse),
// the grid transform reuses the range-compression FFT (fed conj(b) (-1)^m,
// so it returns conj(B)/M), and a last kernel deconvolves and adds into ph.
// ---------------------------------------------------------------------------
constexpr int kNufftW = 8;
constexpr float kNufftBeta = 2.30f * kNufftW;

__device__ __forceinline__ float es_kernel(float x) {
    const float z = 2.0f * x / kNufftW;
    const float r = 1.0f - z * z;
    return r > 0.f ? __expf(kNufftBeta * (sqrtf(r) - 1.0f)) : 0.f;
}

__global__ __launch_bounds__(1024) void nufft_spread(int K, int M, int S,
                                                     const double *__restrict__ scat,
                                                     const float2 *__restrict__ amp,
                                                     const double *__restrict__ ant,
                                                     const double *__restrict__ r0,
                                                     const double *__restrict__ f0, double df,
                                                     float2 *__restrict__ grid) {
    extern __shared__ float sg[];  // 2*M floats (re, im)
    const int n = blockIdx.x;
    for (int i = threadIdx.x; i < 2 * M; i += blockDim.x) sg[i] = 0.f;
    __syncthreads();
    const double ax = ant[3 * n], ay = ant[3 * n + 1], az = ant[3 * n + 2];
    const double maxWr = kClight / (2.0 * df);
    const double kc = (double)(K / 2);
    const double rev_per_m = 2.0 * f0[n] / kClight;
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
        const double dx = ax - scat[3 * s], dy = ay - scat[3 * s + 1], dz = az - scat[3 * s + 2];
        const double dR = sqrt(dx * dx + dy * dy + dz * dz) - r0[n];
        double u = dR / maxWr;
        u -= rint(u);  // periodic in u (period 1), like the sampled spectrum
        double rev = -(rev_per_m * dR + kc * u);
        rev -= rint(rev);
        const float sn = __builtin_amdgcn_sinf((float)rev), cs = __builtin_amdgcn_cosf((float)rev);
        const float2 A = amp[s];
        const float cr = A.x * cs - A.y * sn, ci = A.x * sn + A.y * cs;
        const double x = u * M;  // grid units, in [-M/2, M/2)
        const int m0 = (int)ceil(x - 0.5 * kNufftW);
#pragma unroll
        for (int t = 0; t < kNufftW; t++) {
            const int m = m0 + t;
            const float wgt = es_kernel((float)(x - m));
            const int mm = ((m % M) + M) % M;
            atomicAdd(&sg[2 * mm], wgt * cr);
            atomicAdd(&sg[2 * mm + 1], wgt * ci);
        }
    }
    __syncthreads();
    // conj(b) (-1)^m: the range-compression FFT then returns conj(B)/M
    for (int m = threadIdx.x; m < M; m += blockDim.x) {
        const float sgn = (m & 1) ? -1.f : 1.f;
        grid[(size_t)n * M + m] = make_float2(sgn * sg[2 * m], -sgn * sg[2 * m + 1]);
    }
}

__global__ __launch_bounds__(256) void nufft_deconv_add(int K, int M, const float2 *__restrict__ G,
                                                        const float *__restrict__ inv_phihat,
                                                        float2 *__restrict__ ph) {
    const int n = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int kp = k - K / 2;
    const int l = ((kp % M) + M) % M;
    const float2 g = G[(size_t)n * M + l];  // conj(B_l) / M
    const float sc = (float)M * inv_phihat[k];
    float2 *o = ph + (size_t)n * K + k;
    const float2 old = *o;
    *o = make_float2(old.x + g.x * sc, old.y - g.y * sc);
}

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on P_n).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; i++) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

}  // namespace

int launch_range_compress(sarf_ctx *ctx, const sarf_rc_desc *d, const void *ph,
                          const float *taper, const double *f0, const double *r0, void *prof,
                          hipStream_t stream);

}  // namespace sarf

using namespace sarf;

extern "C" int sarf_synth_phase_history_nufft(sarf_ctx *ctx, int K, int Np, int S,
                                              const double *scat, const float *amp,
                                              const double *ant, const double *r0,
                                              const double *f0, double df, void *ph,
                                              void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (K < 2 || Np < 1 || S < 0)
        return fail(SARF_ERR_INVALID, "synth_nufft: bad sizes K=%d Np=%d S=%d", K, Np, S);
    if (S == 0) return SARF_OK;
    if (!scat || !amp || !ant || !r0 || !f0 || !ph || !(df > 0))
        return fail(SARF_ERR_INVALID, "synth_nufft: bad argument");
    int M = 32;
    while (M < 2 * K) M <<= 1;
    if (M > 16384) return fail(SARF_ERR_UNSUPPORTED, "synth_nufft: K=%d too large", K);
    // 1/phi_hat(k'/M) for k = 0..K-1, k' = k - K/2
    std::vector<double> gx, gw;
    gauss_legendre(64, gx, gw);
    std::vector<float> inv(K);
    const double h = 0.5 * kNufftW;
    for (int k = 0; k < K; k++) {
        const double xi = (double)(k - K / 2) / M;
        double acc = 0.0;
        for (size_t q = 0; q < gx.size(); q++) {
            const double x = h * gx[q];  // [-w/2, w/2]
            const double z = 2.0 * x / kNufftW;
            const double phi = std::exp(kNufftBeta * (std::sqrt(std::max(0.0, 1.0 - z * z)) - 1.0));
            acc += gw[q] * h * phi * std::cos(2.0 * kPi * xi * x);
        }
        inv[k] = (float)(1.0 / acc);
    }
    hipStream_t s = (hipStream_t)stream;
    void *bufg = nullptr, *bufG = nullptr, *bufi = nullptr;
    SARF_TRY(scratch_get(ctx, "nufft.grid", sizeof(float2) * (size_t)M * Np, &bufg));
    SARF_TRY(scratch_get(ctx, "nufft.G", sizeof(float2) * (size_t)M * Np, &bufG));
    SARF_TRY(scratch_get(ctx, "nufft.inv", sizeof(float) * K, &bufi));
    SARF_HIP(hipMemcpyAsync(bufi, inv.data(), sizeof(float) * K, hipMemcpyHostToDevice, s));
    SARF_HIP(hipStreamSynchronize(s));  // inv is pageable host memory
    hipLaunchKernelGGL(nufft_spread, dim3(Np), dim3(1024), sizeof(float) * 2 * M, s, K, M, S, scat,
                       (const float2 *)amp, ant, r0, f0, df, (float2 *)bufg);
    SARF_HIP(hipGetLastError());
    sarf_rc_desc rc{M, Np, M, 0};
    SARF_TRY(launch_range_compress(ctx, &rc, bufg, nullptr, nullptr, nullptr, bufG, s));
    hipLaunchKernelGGL(nufft_deconv_add, dim3((K + 255) / 256, Np), dim3(256), 0, s, K, M,
                       (const float2 *)bufG, (const float *)bufi, (float2 *)ph);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

extern "C" int sarf_synth_phase_history(sarf_ctx *ctx, int K, int Np, int S, const double *scat,
                                        const float *amp, const double *ant, const double *r0,
                                        const double *f0, double df, void *ph, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (K < 1 || Np < 1 || S < 0 || Np > 65535)
        return fail(SARF_ERR_INVALID, "synth: bad sizes K=%d Np=%d S=%d", K, Np, S);
    if (S == 0) return SARF_OK;
    if (!scat || !amp || !ant || !r0 || !f0 || !ph)
        return fail(SARF_ERR_INVALID, "synth: null argument");
    dim3 grid((K + 255) / 256, Np);
    hipLaunchKernelGGL(synth_ph, grid, dim3(256), 0, (hipStream_t)stream, K, Np, S, scat,
                       (const float2 *)amp, ant, r0, f0, df, (float2 *)ph);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

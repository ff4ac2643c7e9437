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

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_synth_phase_history(sarf_ctx *ctx, int K, int Np, int S, const double *scat,
                                        const float *amp, const double *ant, const double *r0,
                                        const double *f0, double df, void *ph, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
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

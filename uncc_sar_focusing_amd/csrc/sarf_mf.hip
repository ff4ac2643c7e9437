// sarf_mf.hip -- matched-filter image formation.
//
// Semantics of run_mf / computeDifferentialRangeAndPhaseCorrectionsMF
// (src/cpuBackProjection/cpuBackProjection.hpp:165-228) and mfBasic.m:66-89:
//   img[x*Ny + y] += sum_n sum_k ph[n][k] exp(j 4 pi f_{n,k} dR_n / c),
//   f_{n,k} = f0_n + k df_n, dR_n = |a_n - p| - R0_n, target z = 0,
// for every pixel and pulse (no swath gate, no interpolation).
//
// gfx950 formulation.  A workgroup (4 wave64) owns a 16 x 64 pixel tile:
// lanes run along y, each wave owns 4 columns.  The cost is the O(K) sum,
// so each pixel and pulse affords an exact FP64 set-up: dR = |a_n - p| - R0
// (antenna terms staged in LDS per 32-pulse chunk), the revolutions
// th0 = frac(2 f0 dR / c) and w = exp(j 2 pi frac(2 df dR / c)) (FP64
// sincos, rounded to float2).  Since 2 f_k dR / c = th0 + k (2 df dR / c),
// the pulse's sum is e0 * sum_k ph_k w^k, e0 = exp(j 2 pi th0): a Horner
// recurrence over the pulse's samples (4 FP32 FMA per sample and pixel),
// samples staged in LDS and read as broadcast pairs.  Rounding w to float2
// costs about 4e-8 rad per step, so the phase error after K steps stays
// below K * 1e-7 rad at any pixel spacing.
#include <algorithm>
#include <cmath>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

constexpr int kMfThreads = 256;
constexpr int kMfTY = 64;      // lanes along y
constexpr int kMfPX = 4;       // columns per wave
constexpr int kMfTX = 4 * kMfPX;
constexpr int kMfChunk = 32;   // pulses set up per chunk
constexpr int kMfStage = 1024; // samples staged in LDS at a time (even)
constexpr int kMfMaxK = 8192;

struct MfArgs {
    int Nx, Ny, K, Np;
    int tiles_x, tiles_y, splits;
    int accumulate;
    long split_stride;
    double left, bottom, dx, dy;
};

__global__ __launch_bounds__(kMfThreads) void mf_mixed(MfArgs a, const float2 *__restrict__ ph,
                                                       const double *__restrict__ ant,
                                                       const double *__restrict__ r0,
                                                       const double *__restrict__ f0,
                                                       const double *__restrict__ df,
                                                       float2 *__restrict__ img) {
    __shared__ float4 samp[kMfStage / 2];  // (ph_{2i}, ph_{2i+1})
    __shared__ double4 pA[kMfChunk];       // a_x, a_y, a_z^2, R0
    __shared__ double2 pB[kMfChunk];       // 2 f0 / c, 2 df / c
    const int ntiles = a.tiles_x * a.tiles_y;
    const int split = blockIdx.x / ntiles;
    const int t = blockIdx.x - split * ntiles;
    const int tx = t % a.tiles_x, ty = t / a.tiles_x;
    const int per = (a.Np + a.splits - 1) / a.splits;
    const int nbeg = min(a.Np, split * per), nend = min(a.Np, nbeg + per);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int xb = tx * kMfTX, yb = ty * kMfTY;
    const int y = yb + lane;
    const double py = a.bottom + y * a.dy;
    double px[kMfPX];
#pragma unroll
    for (int j = 0; j < kMfPX; j++) px[j] = a.left + (xb + wave * kMfPX + j) * a.dx;
    float2 acc[kMfPX];
#pragma unroll
    for (int j = 0; j < kMfPX; j++) acc[j] = make_float2(0.f, 0.f);
    const int K = a.K;
    const int nst = (K + kMfStage - 1) / kMfStage;

    for (int n0 = nbeg; n0 < nend; n0 += kMfChunk) {
        const int nc = min(kMfChunk, nend - n0);
        __syncthreads();
        if (tid < nc) {
            const int n = n0 + tid;
            const double az = ant[3 * n + 2];
            pA[tid] = make_double4(ant[3 * n], ant[3 * n + 1], az * az, r0[n]);
            pB[tid] = make_double2(2.0 * f0[n] / kClight, 2.0 * df[n] / kClight);
        }
        for (int j = 0; j < nc; j++) {
            const int n = n0 + j;
            float2 e0[kMfPX], w[kMfPX], hs[kMfPX];
            for (int st = nst - 1; st >= 0; st--) {
                const int kb = st * kMfStage, ke = min(K, kb + kMfStage);
                const int L = ke - kb, Lp = (L + 1) & ~1;
                __syncthreads();
                float2 *s2 = (float2 *)samp;
                const float2 *src = ph + (size_t)n * K + kb;
                for (int i = tid; i < Lp; i += kMfThreads)
                    s2[i] = i < L ? src[i] : make_float2(0.f, 0.f);
                __syncthreads();
                if (st == nst - 1) {
                    const double4 A = pA[j];
                    const double2 B = pB[j];
                    const double ddy = A.y - py;
#pragma unroll
                    for (int q = 0; q < kMfPX; q++) {
                        const double ddx = A.x - px[q];
                        const double dR = sqrt(fma(ddx, ddx, fma(ddy, ddy, A.z))) - A.w;
                        const double t0 = B.x * dR, t1 = B.y * dR;
                        const float th0 = (float)(t0 - floor(t0));
                        e0[q] = make_float2(__builtin_amdgcn_cosf(th0), __builtin_amdgcn_sinf(th0));
                        double sn, cs;
                        sincos(2.0 * kPi * (t1 - floor(t1)), &sn, &cs);
                        w[q] = make_float2((float)cs, (float)sn);
                        hs[q] = make_float2(0.f, 0.f);
                    }
                }
                // Horner, highest sample first: h <- h w + ph_k
                for (int i = Lp / 2 - 1; i >= 0; i--) {
                    const float4 pp = samp[i];
#pragma unroll
                    for (int q = 0; q < kMfPX; q++) {
                        float hr = fmaf(hs[q].x, w[q].x, fmaf(-hs[q].y, w[q].y, pp.z));
                        float hi = fmaf(hs[q].x, w[q].y, fmaf(hs[q].y, w[q].x, pp.w));
                        hs[q].x = fmaf(hr, w[q].x, fmaf(-hi, w[q].y, pp.x));
                        hs[q].y = fmaf(hr, w[q].y, fmaf(hi, w[q].x, pp.y));
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < kMfPX; q++) acc[q] = cadd(acc[q], cmul(e0[q], hs[q]));
        }
    }
    if (y >= a.Ny) return;
    float2 *o = img + (size_t)split * a.split_stride;
#pragma unroll
    for (int q = 0; q < kMfPX; q++) {
        const int x = xb + wave * kMfPX + q;
        if (x >= a.Nx) continue;
        const size_t idx = (size_t)x * a.Ny + y;
        o[idx] = a.accumulate ? cadd(o[idx], acc[q]) : acc[q];
    }
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_matched_filter(sarf_ctx *ctx, const sarf_mf_desc *d, const void *ph,
                                   const double *ant, const double *r0, const double *f0,
                                   const double *df, void *img, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!d || !ph || !ant || !r0 || !f0 || !df || !img)
        return fail(SARF_ERR_INVALID, "sarf_matched_filter: null argument");
    if (d->Nx < 2 || d->Ny < 2 || d->K < 1 || d->Np < 1)
        return fail(SARF_ERR_INVALID, "sarf_matched_filter: bad sizes Nx=%d Ny=%d K=%d Np=%d",
                    d->Nx, d->Ny, d->K, d->Np);
    if (d->K > kMfMaxK)
        return fail(SARF_ERR_UNSUPPORTED, "sarf_matched_filter: K=%d > %d", d->K, kMfMaxK);
    if (!(d->Wx > 0) || !(d->Wy > 0) || !std::isfinite(d->x0) || !std::isfinite(d->y0))
        return fail(SARF_ERR_INVALID, "sarf_matched_filter: bad geometry Wx=%g Wy=%g", d->Wx,
                    d->Wy);
    hipStream_t s = (hipStream_t)stream;
    MfArgs a{};
    a.Nx = d->Nx;
    a.Ny = d->Ny;
    a.K = d->K;
    a.Np = d->Np;
    a.tiles_x = (d->Nx + kMfTX - 1) / kMfTX;
    a.tiles_y = (d->Ny + kMfTY - 1) / kMfTY;
    a.dx = d->Wx / (d->Nx - 1);
    a.dy = d->Wy / (d->Ny - 1);
    a.left = d->x0 - d->Wx / 2;
    a.bottom = d->y0 - d->Wy / 2;
    const int ntiles = a.tiles_x * a.tiles_y;
    int splits = d->pulse_split;
    if (splits <= 0) {
        splits = 1;
        const int want = 4 * (ctx->cu_count > 0 ? ctx->cu_count : 256);
        while (splits < 64 && ntiles * splits < want && d->Np / (2 * splits) >= 4) splits *= 2;
    }
    splits = std::max(1, std::min(splits, d->Np));
    const long npix = (long)d->Nx * d->Ny;
    float2 *out = (float2 *)img;
    a.splits = splits;
    a.accumulate = (d->flags & SARF_BP_ACCUMULATE) ? 1 : 0;
    if (splits > 1) {
        void *buf = nullptr;
        SARF_TRY(scratch_get(ctx, "mf.partials", sizeof(float2) * npix * splits, &buf));
        out = (float2 *)buf;
        a.split_stride = npix;
        a.accumulate = 0;
    }
    hipLaunchKernelGGL(mf_mixed, dim3(ntiles * splits), dim3(kMfThreads), 0, s, a,
                       (const float2 *)ph, ant, r0, f0, df, out);
    SARF_HIP(hipGetLastError());
    if (splits > 1)
        SARF_TRY(launch_reduce_partials(out, npix, splits, npix,
                                        (d->flags & SARF_BP_ACCUMULATE) ? 1 : 0, (float2 *)img, s));
    return SARF_OK;
}

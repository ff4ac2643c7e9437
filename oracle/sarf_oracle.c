/*
 * sarf_oracle.c -- CPU restatement of the reference SAR focusing algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the reference results that GPU outputs are
 * compared with (or as the timed CPU baseline),
 * never as the product path.  The product is the HIP library under
 * uncc_sar_focusing_amd/csrc, which has no CPU fallback.
 *
 * Two restatements of the same algorithm are provided:
 *
 *   orc_*_f64        : all arithmetic in IEEE double, exact linspace pixel
 *                      grid (x_i = x0 - Wx/2 + i*Wx/(Nx-1)).  This is the
 *                      FP64 oracle every GPU result is compared with.
 *   orc_*_f32_faithful: the reference's own operation order and float-isms
 *                      (float dR, float-accumulated pixel grid, float range
 *                      bins, float interpolation, float accumulation).  It
 *                      is the CPU timing baseline and measures the
 *                      reference's own FP32 error.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/cpuBackProjection/cpuBackProjection.hpp:24-42   interp1 (search rule)
 *   src/cpuBackProjection/cpuBackProjection.hpp:45,47   range index <-> value
 *   src/cpuBackProjection/cpuBackProjection.hpp:49-93   BP inner loop
 *   src/cpuBackProjection/cpuBackProjection.hpp:95-163  run_bp (bins, RC, BP)
 *   src/cpuBackProjection/cpuBackProjection.hpp:165-228 matched filter
 *   src/cpuBackProjection/cpuBackProjection_fft.hpp:46-48,81-84  1/N, fftshift
 *   matlab/bpBasic.m:61-97   zero-padded ifft(ph, Nfft), r_vec, gate, interp1
 *   include/uncc_complex.hpp:81-84,186-191  polar(), float-truncated cmul
 *
 * Parity status of each exported function is stated beside it; the checks
 * that pin it live in tests/test_oracle.py.
 *
 * Layouts (all complex arrays interleaved re,im):
 *   phase history ph[n*K + k]            K samples per pulse, pulses contiguous
 *   range profiles prof[n*Nfft + m]      Nfft bins per pulse
 *   antenna ant[3*n + {0,1,2}]           x,y,z per pulse
 *   image img[x*Ny + y]                  Ny x Nx column-major, y fastest
 *                                         (cpuBackProjection.hpp:87)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_CLIGHT 299792458.0        /* uncc_sar_focusing.hpp:34 */
#define ORC_PI 3.141592653589793      /* uncc_sar_focusing.hpp:36 */

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2 decimation-in-time, double and float.             */
/* Pinned against numpy.fft and a direct O(N^2) DFT in tests.               */
/* sign = -1 forward (e^{-j}), +1 inverse (e^{+j}), unscaled.                */
/* ------------------------------------------------------------------------ */
static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void bitrev_permute_d(double *x, int n) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr; x[2 * j + 1] = ti;
        }
    }
}

int orc_fft_f64(double *x, int n, int sign) {
    if (!is_pow2(n)) return -1;
    bitrev_permute_d(x, n);
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            double ang = sign * 2.0 * ORC_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (int s = 0; s < n; s += len) {
                double *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                double tr = b[0] * wr - b[1] * wi, ti = b[0] * wi + b[1] * wr;
                b[0] = a[0] - tr; b[1] = a[1] - ti;
                a[0] += tr; a[1] += ti;
            }
        }
    }
    return 0;
}

static void bitrev_permute_f(float *x, int n) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            float tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j]; x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr; x[2 * j + 1] = ti;
        }
    }
}

/* float FFT with double-precision twiddles rounded to float (as FFTW's
 * single-precision plans use precomputed accurate twiddles). */
int orc_fft_f32(float *x, int n, int sign) {
    if (!is_pow2(n)) return -1;
    bitrev_permute_f(x, n);
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            double ang = sign * 2.0 * ORC_PI * (double)k / (double)len;
            float wr = (float)cos(ang), wi = (float)sin(ang);
            for (int s = 0; s < n; s += len) {
                float *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                float tr = b[0] * wr - b[1] * wi, ti = b[0] * wi + b[1] * wr;
                b[0] = a[0] - tr; b[1] = a[1] - ti;
                a[0] += tr; a[1] += ti;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Range bins: r[i] = (i/N - 0.5) * maxWr                                    */
/* cpuBackProjection.hpp:45,124-135; bpBasic.m:61.                           */
/* ------------------------------------------------------------------------ */
void orc_range_bins_f64(int nfft, double maxWr, double *r) {
    for (int i = 0; i < nfft; i++) r[i] = ((double)i / nfft - 0.5) * maxWr;
}

/* float form of the reference macro: ((float) idx / N - 0.5f) * maxWr */
void orc_range_bins_f32(int nfft, float maxWr, float *r) {
    for (int i = 0; i < nfft; i++) r[i] = ((float)i / nfft - 0.5f) * maxWr;
}

/* ------------------------------------------------------------------------ */
/* Range compression: prof = fftshift(ifft(ph, Nfft))                        */
/*   zero-pad at the end (bpBasic.m:82), inverse DFT scaled by 1/Nfft        */
/*   (cpuBackProjection_fft.hpp:46-48), cshift((N+1)/2) (:81-84):            */
/*   prof[m] = X[(m + (N+1)/2) mod N].                                       */
/* Optional taper (per-sample real weight, matlab/rangeCompress.m:6-12) and  */
/* phase ramp exp(-j 4 pi f0 R0 / c) per pulse (rangeCompress.m:17).         */
/* ------------------------------------------------------------------------ */
/* Unscaled inverse DFT of length n by definition (any n), O(n^2). */
static void dft_inverse_direct(double *x, int n) {
    double *y = (double *)calloc(2 * (size_t)n, sizeof(double));
    for (int m = 0; m < n; m++) {
        double ar = 0, ai = 0;
        for (int k = 0; k < n; k++) {
            double a = 2.0 * ORC_PI * (double)(((long)k * m) % n) / n;
            double c = cos(a), s = sin(a);
            ar += x[2 * k] * c - x[2 * k + 1] * s;
            ai += x[2 * k] * s + x[2 * k + 1] * c;
        }
        y[2 * m] = ar;
        y[2 * m + 1] = ai;
    }
    memcpy(x, y, sizeof(double) * 2 * n);
    free(y);
}

int orc_range_compress_f64(int K, int Np, int nfft, const double *ph,
                           const double *taper /* K or NULL */,
                           const double *ramp_f0 /* Np or NULL */,
                           const double *ramp_r0 /* Np or NULL */,
                           double *prof) {
    if (K > nfft || K <= 0 || Np <= 0) return -1;
    int shift = (nfft + 1) / 2;
    int err = 0;
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 2 * nfft);
#pragma omp for schedule(static)
        for (int n = 0; n < Np; n++) {
            memset(buf, 0, sizeof(double) * 2 * nfft);
            for (int k = 0; k < K; k++) {
                double w = taper ? taper[k] : 1.0;
                buf[2 * k] = ph[2 * ((size_t)n * K + k)] * w;
                buf[2 * k + 1] = ph[2 * ((size_t)n * K + k) + 1] * w;
            }
            if (is_pow2(nfft))
                orc_fft_f64(buf, nfft, +1);
            else
                dft_inverse_direct(buf, nfft);
            double rr = 1.0 / nfft, ri = 0.0;
            if (ramp_f0 && ramp_r0) {
                double a = -4.0 * ORC_PI * ramp_f0[n] * ramp_r0[n] / ORC_CLIGHT;
                rr = cos(a) / nfft; ri = sin(a) / nfft;
            }
            double *out = prof + 2 * (size_t)n * nfft;
            for (int m = 0; m < nfft; m++) {
                int src = (m + shift) % nfft;
                double xr = buf[2 * src], xi = buf[2 * src + 1];
                out[2 * m] = xr * rr - xi * ri;
                out[2 * m + 1] = xr * ri + xi * rr;
            }
        }
        free(buf);
    }
    return err;
}

/* float restatement (FFTW single precision stand-in: own radix-2 FFT). */
int orc_range_compress_f32(int K, int Np, int nfft, const float *ph, float *prof) {
    if (!is_pow2(nfft) || K > nfft || K <= 0 || Np <= 0) return -1;
    int shift = (nfft + 1) / 2;
#pragma omp parallel
    {
        float *buf = (float *)malloc(sizeof(float) * 2 * nfft);
#pragma omp for schedule(static)
        for (int n = 0; n < Np; n++) {
            memset(buf, 0, sizeof(float) * 2 * nfft);
            memcpy(buf, ph + 2 * (size_t)n * K, sizeof(float) * 2 * K);
            orc_fft_f32(buf, nfft, +1);
            float *out = prof + 2 * (size_t)n * nfft;
            for (int m = 0; m < nfft; m++) {
                int src = (m + shift) % nfft;
                out[2 * m] = buf[2 * src] / nfft;      /* x /= x.size() */
                out[2 * m + 1] = buf[2 * src + 1] / nfft;
            }
        }
        free(buf);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Backprojection, FP64.                                                     */
/* For every pulse n, pixel (x,y) at z = 0 (cpuBackProjection.hpp:58):       */
/*   dR = |a_n - p| - R0_n                                  (:70-74)         */
/*   if r[0] < dR < r[N-1]:                                 (:76)            */
/*     phi = 4 pi f0_n dR / c                               (:78-79)         */
/*     i   = first bin with r[i] > dR, searched from floor((dR/maxWr+.5)N)   */
/*                                                          (:28-29,:83)     */
/*     a   = (dR - r[i-1]) / (r[i] - r[i-1])                (:38-39)         */
/*     img[x*Ny+y] += (a s[i] + (1-a) s[i-1]) e^{j phi}     (:40,:87)        */
/* Pixel grid x_i = x0 - Wx/2 + i Wx/(Nx-1) by multiplication (the           */
/* reference accumulates it in float, :62,:89-91; bpBasic.m:100 linspace).   */
/* pulse_begin/pulse_end select a contiguous pulse range (pulse-split        */
/* additivity); accumulate != 0 adds into img instead of overwriting.        */
/* ------------------------------------------------------------------------ */
static inline int interp_search_d(const double *r, int nfft, double dR, double idx) {
    int ri = (int)floor(idx) - 1;
    if (ri < -1) ri = -1;
    while (++ri < nfft && r[ri] <= dR) {}
    return ri;
}

int orc_bp_f64(int Nx, int Ny, int nfft, int Np, int pulse_begin, int pulse_end,
               double x0, double y0, double Wx, double Wy, double maxWr,
               const double *prof, const double *ant, const double *r0,
               const double *f0, int accumulate, double *img) {
    if (Nx < 2 || Ny < 2 || nfft < 2 || Np <= 0) return -1;
    if (pulse_begin < 0) pulse_begin = 0;
    if (pulse_end > Np || pulse_end < 0) pulse_end = Np;
    double *r = (double *)malloc(sizeof(double) * nfft);
    orc_range_bins_f64(nfft, maxWr, r);
    double rmin = r[0], rmax = r[nfft - 1];
    double dx = Wx / (Nx - 1), dy = Wy / (Ny - 1);
    double left = x0 - Wx / 2, bottom = y0 - Wy / 2;
    if (!accumulate) memset(img, 0, sizeof(double) * 2 * (size_t)Nx * Ny);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Nx; ix++) {
        double px = left + ix * dx;
        for (int n = pulse_begin; n < pulse_end; n++) {
            const double ax = ant[3 * n], ay = ant[3 * n + 1], az = ant[3 * n + 2];
            const double *s = prof + 2 * (size_t)n * nfft;
            const double k4 = 4.0 * ORC_PI * f0[n] / ORC_CLIGHT;
            for (int iy = 0; iy < Ny; iy++) {
                double py = bottom + iy * dy;
                double ddx = ax - px, ddy = ay - py, ddz = az - 0.0;
                double dR = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - r0[n];
                if (dR > rmin && dR < rmax) {
                    double ph = k4 * dR;
                    double idx = (dR / maxWr + 0.5) * nfft;
                    int ri = interp_search_d(r, nfft, dR, idx);
                    if (ri == nfft || ri == 0) continue;  /* interp1 error path */
                    double al = (dR - r[ri - 1]) / (r[ri] - r[ri - 1]);
                    double vr = al * s[2 * ri] + (1.0 - al) * s[2 * (ri - 1)];
                    double vi = al * s[2 * ri + 1] + (1.0 - al) * s[2 * (ri - 1) + 1];
                    double c = cos(ph), sn = sin(ph);
                    double *o = img + 2 * ((size_t)ix * Ny + iy);
                    o[0] += vr * c - vi * sn;
                    o[1] += vr * sn + vi * c;
                }
            }
        }
    }
    free(r);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Backprojection, FP32-faithful restatement of cpuBackProjection.hpp:49-93  */
/* Same loop order (pulse outer, x, y), float target_x/target_y advanced by  */
/* += delta (:62,:68,:89-91), float dR (:70-74), phase product in double     */
/* cast to float then float cos/sin (:78-79, uncc_complex.hpp:81-84), float  */
/* index (:83 with the :47 macro), linear search (:28-29), float alpha/lerp  */
/* (:38-40), float complex multiply-accumulate (uncc_complex.hpp:186-191).   */
/* The x loop is split over OpenMP threads inside each pulse; every pixel    */
/* still sums pulses in the same order, so the result does not depend on the */
/* thread count.                                                            */
/* ------------------------------------------------------------------------ */
int orc_bp_f32_faithful(int Nx, int Ny, int nfft, int Np, int pulse_begin, int pulse_end,
                        float x0, float y0, float Wx, float Wy, float maxWr,
                        const float *prof, const float *antx, const float *anty,
                        const float *antz, const float *r0, const float *f0,
                        float *img) {
    if (Nx < 2 || Ny < 2 || nfft < 2 || Np <= 0) return -1;
    if (pulse_begin < 0) pulse_begin = 0;
    if (pulse_end > Np || pulse_end < 0) pulse_end = Np;
    float *r = (float *)malloc(sizeof(float) * nfft);
    orc_range_bins_f32(nfft, maxWr, r);
    float rmin = INFINITY, rmax = -INFINITY;
    for (int i = 0; i < nfft; i++) {
        if (rmin > r[i]) rmin = r[i];
        if (rmax < r[i]) rmax = r[i];
    }
    const float delta_x = Wx / (Nx - 1);
    const float delta_y = Wy / (Ny - 1);
    memset(img, 0, sizeof(float) * 2 * (size_t)Nx * Ny);
    /* The float-accumulated grid (target += delta) is precomputed once so
     * that threads can start at any column with the same float values the
     * sequential loop produces. */
    float *gx = (float *)malloc(sizeof(float) * Nx);
    float *gy = (float *)malloc(sizeof(float) * Ny);
    {
        float t = x0 - (Wx / 2);
        for (int i = 0; i < Nx; i++) { gx[i] = t; t += delta_x; }
        t = y0 - (Wy / 2);
        for (int i = 0; i < Ny; i++) { gy[i] = t; t += delta_y; }
    }
#pragma omp parallel
    for (int n = pulse_begin; n < pulse_end; n++) {
        const float *s = prof + 2 * (size_t)n * nfft;
        const float ax = antx[n], ay = anty[n], az = antz[n];
#pragma omp for schedule(static)
        for (int ix = 0; ix < Nx; ix++) {
            const float target_x = gx[ix];
            const float target_z = 0;
            for (int iy = 0; iy < Ny; iy++) {
                const float target_y = gy[iy];
                float dR_val = sqrtf((ax - target_x) * (ax - target_x) +
                                     (ay - target_y) * (ay - target_y) +
                                     (az - target_z) * (az - target_z)) - r0[n];
                if (dR_val > rmin && dR_val < rmax) {
                    float phase = (float)((4.0 * ORC_PI * f0[n] * dR_val) / ORC_CLIGHT);
                    float pc_r = 1.0f * cosf(phase), pc_i = 1.0f * sinf(phase);
                    float dR_idx = (dR_val / maxWr + 0.5f) * nfft;
                    int ri = (int)floorf(dR_idx) - 1;
                    if (ri < -1) ri = -1;
                    while (++ri < nfft && r[ri] <= dR_val) {}
                    if (ri == nfft || ri == 0) continue;
                    float alpha = (dR_val - r[ri - 1]) / (r[ri] - r[ri - 1]);
                    float vr = alpha * s[2 * ri] + (1.0f - alpha) * s[2 * (ri - 1)];
                    float vi = alpha * s[2 * ri + 1] + (1.0f - alpha) * s[2 * (ri - 1) + 1];
                    float *o = img + 2 * ((size_t)ix * Ny + iy);
                    const float cr = vr * pc_r - vi * pc_i;
                    const float ci = vr * pc_i + vi * pc_r;
                    o[0] += cr;
                    o[1] += ci;
                }
            }
        }
    }
    free(gx); free(gy); free(r);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Matched filter (exact image formation), FP64:                             */
/*   img += ph[k] e^{j 4 pi f_k dR / c}  for every k     (cpuBackProjection  */
/*   .hpp:166-201; matlab/mfBasic.m).  f_k = f0_n + k df.                    */
/* All pulses are used (the reference's run_mf skips pulses 0-1 because its  */
/* progress-print condition wraps the call, :218; that is not kept).         */
/* ------------------------------------------------------------------------ */
int orc_mf_f64(int Nx, int Ny, int K, int Np, double x0, double y0, double Wx,
               double Wy, double df, const double *ph, const double *ant,
               const double *r0, const double *f0, double *img) {
    if (Nx < 2 || Ny < 2 || K <= 0 || Np <= 0) return -1;
    double dx = Wx / (Nx - 1), dy = Wy / (Ny - 1);
    double left = x0 - Wx / 2, bottom = y0 - Wy / 2;
    memset(img, 0, sizeof(double) * 2 * (size_t)Nx * Ny);
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < Nx; ix++) {
        double px = left + ix * dx;
        for (int iy = 0; iy < Ny; iy++) {
            double py = bottom + iy * dy;
            double accr = 0, acci = 0;
            for (int n = 0; n < Np; n++) {
                double ddx = ant[3 * n] - px, ddy = ant[3 * n + 1] - py, ddz = ant[3 * n + 2];
                double dR = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - r0[n];
                const double *p = ph + 2 * (size_t)n * K;
                for (int k = 0; k < K; k++) {
                    double a = 4.0 * ORC_PI * (f0[n] + k * df) * dR / ORC_CLIGHT;
                    double c = cos(a), s = sin(a);
                    accr += p[2 * k] * c - p[2 * k + 1] * s;
                    acci += p[2 * k] * s + p[2 * k + 1] * c;
                }
            }
            img[2 * ((size_t)ix * Ny + iy)] = accr;
            img[2 * ((size_t)ix * Ny + iy) + 1] = acci;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Point-scatterer phase history, FP64 (matlab/formImagePointTargetBP.m:63-81)*/
/*   ph[n,k] = sum_t A_t e^{-j 4 pi f_k (|a_n - p_t| - R0_n) / c}            */
/* with f_k = f0_n + k df and A_t complex.                                   */
/* ------------------------------------------------------------------------ */
int orc_synth_ph_f64(int K, int Np, int S, const double *scat /* S x 3 */,
                     const double *amp /* S complex */, const double *ant,
                     const double *r0, const double *f0, double df, double *ph) {
#pragma omp parallel for schedule(static)
    for (int n = 0; n < Np; n++) {
        double *out = ph + 2 * (size_t)n * K;
        memset(out, 0, sizeof(double) * 2 * K);
        for (int t = 0; t < S; t++) {
            double ddx = ant[3 * n] - scat[3 * t], ddy = ant[3 * n + 1] - scat[3 * t + 1];
            double ddz = ant[3 * n + 2] - scat[3 * t + 2];
            double dR = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - r0[n];
            for (int k = 0; k < K; k++) {
                /* phase in revolutions, reduced exactly before the sincos */
                double rev = -2.0 * (f0[n] + k * df) * dR / ORC_CLIGHT;
                rev -= nearbyint(rev);
                double c = cos(2.0 * ORC_PI * rev), s = sin(2.0 * ORC_PI * rev);
                out[2 * k] += amp[2 * t] * c - amp[2 * t + 1] * s;
                out[2 * k + 1] += amp[2 * t] * s + amp[2 * t + 1] * c;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Image entropy E = -sum p ln p, p = |I|^2 / sum |I|^2  (FP64).             */
/* Equivalent to ln S - T/S with S = sum|I|^2, T = sum |I|^2 ln |I|^2.       */
/* ------------------------------------------------------------------------ */
double orc_entropy_f64(const double *img, long npix) {
    double S = 0, T = 0;
    for (long i = 0; i < npix; i++) {
        double m = img[2 * i] * img[2 * i] + img[2 * i + 1] * img[2 * i + 1];
        S += m;
        if (m > 0) T += m * log(m);
    }
    if (S <= 0) return 0.0;
    return log(S) - T / S;
}

/* ------------------------------------------------------------------------ */
/* Adjoint backprojection, FP64: gradient of E(I(a)) with respect to the     */
/* antenna positions a_n (R0_n held fixed).                                  */
/*   dE/da_n = sum_p 2 w_p Re( conj(I_p) (v'_n + j k_n v_n) e^{j k_n dR} )   */
/*             * (a_n - p) / |a_n - p|                                       */
/*   w_p = dE/d|I_p|^2 = -(ln(|I_p|^2/S) + E) / S                            */
/* v_n is the linearly interpolated profile (interp1 rule above), v'_n its   */
/* slope (s[i]-s[i-1])/(r[i]-r[i-1]).  img must be the forward image formed  */
/* from the same inputs (orc_bp_f64).                                        */
/* ------------------------------------------------------------------------ */
int orc_bp_adjoint_f64(int Nx, int Ny, int nfft, int Np, double x0, double y0,
                       double Wx, double Wy, double maxWr, const double *prof,
                       const double *ant, const double *r0, const double *f0,
                       const double *img, double *grad /* Np x 3 */) {
    long npix = (long)Nx * Ny;
    double S = 0, T = 0;
    for (long i = 0; i < npix; i++) {
        double m = img[2 * i] * img[2 * i] + img[2 * i + 1] * img[2 * i + 1];
        S += m;
        if (m > 0) T += m * log(m);
    }
    if (S <= 0) return -1;
    double E = log(S) - T / S;
    double *r = (double *)malloc(sizeof(double) * nfft);
    orc_range_bins_f64(nfft, maxWr, r);
    double rmin = r[0], rmax = r[nfft - 1];
    double dx = Wx / (Nx - 1), dy = Wy / (Ny - 1);
    double left = x0 - Wx / 2, bottom = y0 - Wy / 2;
#pragma omp parallel for schedule(static)
    for (int n = 0; n < Np; n++) {
        const double ax = ant[3 * n], ay = ant[3 * n + 1], az = ant[3 * n + 2];
        const double *s = prof + 2 * (size_t)n * nfft;
        const double k4 = 4.0 * ORC_PI * f0[n] / ORC_CLIGHT;
        double gx = 0, gy = 0, gz = 0;
        for (int ix = 0; ix < Nx; ix++) {
            double px = left + ix * dx;
            for (int iy = 0; iy < Ny; iy++) {
                double py = bottom + iy * dy;
                double ddx = ax - px, ddy = ay - py, ddz = az;
                double R = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                double dR = R - r0[n];
                if (!(dR > rmin && dR < rmax)) continue;
                double idx = (dR / maxWr + 0.5) * nfft;
                int ri = interp_search_d(r, nfft, dR, idx);
                if (ri == nfft || ri == 0) continue;
                double h = r[ri] - r[ri - 1];
                double al = (dR - r[ri - 1]) / h;
                double vr = al * s[2 * ri] + (1.0 - al) * s[2 * (ri - 1)];
                double vi = al * s[2 * ri + 1] + (1.0 - al) * s[2 * (ri - 1) + 1];
                double dvr = (s[2 * ri] - s[2 * (ri - 1)]) / h;
                double dvi = (s[2 * ri + 1] - s[2 * (ri - 1) + 1]) / h;
                /* d/d(dR) [v e^{j k dR}] = (v' + j k v) e^{j k dR} */
                double qr = dvr - k4 * vi, qi = dvi + k4 * vr;
                double ph = k4 * dR, c = cos(ph), sn = sin(ph);
                double er = qr * c - qi * sn, ei = qr * sn + qi * c;
                long p = (long)ix * Ny + iy;
                double Ir = img[2 * p], Ii = img[2 * p + 1];
                double m = Ir * Ir + Ii * Ii;
                double w = (m > 0) ? -(log(m / S) + E) / S : 0.0;
                double cpn = 2.0 * w * (Ir * er + Ii * ei); /* Re(conj(I) e) */
                gx += cpn * ddx / R;
                gy += cpn * ddy / R;
                gz += cpn * ddz / R;
            }
        }
        grad[3 * n] = gx; grad[3 * n + 1] = gy; grad[3 * n + 2] = gz;
    }
    free(r);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 8-bit dB display image, as the reference's writeBMPFile computes it      */
/* (include/uncc_sar_globals.hpp:66-118), in float with no contraction:     */
/*   m = max |I| (scaled abs of include/uncc_complex.hpp:93-102),           */
/*   pixelf = (float)(255.0 / dyn) * (20 * log10f(|I| / m) + dyn),          */
/*   byte = pixelf < 0 ? 0 : (unsigned char)pixelf,                         */
/* rows top-down with row r showing y = Ny-1-r (the flipX branch, :85-98):  */
/* out[r*Nx + x].  Parity: pinned by closed-form levels in test_oracle.py.  */
/* ------------------------------------------------------------------------ */
static float scaled_abs_f(float x, float y) {
    float s = fmaxf(fabsf(x), fabsf(y));
    if (s == 0.0f) return s;
    x /= s;
    y /= s;
    float xx = x * x, yy = y * y;
    return s * sqrtf(xx + yy);
}

int orc_db8(int Nx, int Ny, const float *img, float dyn, unsigned char *out) {
    if (Nx < 1 || Ny < 1 || !(dyn > 0)) return -1;
    long n = (long)Nx * Ny;
    float m = 0.0f;
    for (long i = 0; i < n; i++) {
        float a = scaled_abs_f(img[2 * i], img[2 * i + 1]);
        if (a > m) m = a;
    }
    const float scale = (float)(255.0 / (double)dyn);
    for (int r = 0; r < Ny; r++) {
        int y = Ny - 1 - r;
        for (int x = 0; x < Nx; x++) {
            long i = (long)x * Ny + y;
            float a = scaled_abs_f(img[2 * i], img[2 * i + 1]);
            float l = log10f(a / m);
            float t = 20.0f * l;
            float u = t + dyn;
            float v = scale * u;
            out[(long)r * Nx + x] = (v < 0.0f || v != v) ? 0 : (unsigned char)v;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exhaustive check of a level table for the dB quantisation above: for     */
/* every float r in [0, 1] (bit patterns 0 .. 0x3f800000), the byte of the   */
/* reference expression (uncc_sar_globals.hpp:99-104, as orc_db8 evaluates  */
/* it for r = |I|/max) must equal the number of levels[1..255] <= r.        */
/* Returns the number of mismatching r; *first_bad gets the smallest        */
/* mismatching bit pattern (0xffffffff if none).                            */
/* ------------------------------------------------------------------------ */
static unsigned char db8_byte_of_ratio(float r, float dyn, float scale) {
    float l = log10f(r);
    float t = 20.0f * l;
    float u = t + dyn;
    float v = scale * u;
    return (v < 0.0f || v != v) ? 0 : (unsigned char)v;
}

long orc_db8_level_mismatches(float dyn, const float *levels, unsigned int *first_bad) {
    const float scale = (float)(255.0 / (double)dyn);
    const unsigned int top = 0x3f800000u;
    long bad = 0;
    unsigned int first = 0xffffffffu;
#pragma omp parallel for schedule(static, 1 << 20) reduction(+ : bad) reduction(min : first)
    for (long u = 0; u <= (long)top; u++) {
        unsigned int bits = (unsigned int)u;
        float r;
        memcpy(&r, &bits, 4);
        int lo = 0;
        for (int step = 128; step > 0; step >>= 1)
            if (lo + step <= 255 && levels[lo + step] <= r) lo += step;
        if (lo != (int)db8_byte_of_ratio(r, dyn, scale)) {
            bad++;
            if (bits < first) first = bits;
        }
    }
    *first_bad = first;
    return bad;
}

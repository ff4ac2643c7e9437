// sarf_pga.hip -- phase gradient autofocus (PGA) on a focused image.
//
// Semantics of matlab/PGA.m:151-232 (the reference's C++ restatement is
// src/autofocus/cpu/autofocus_cpu.cpp:249-474), image rows = range y,
// columns = azimuth x; our image is img[x*Ny + y].  Per iteration:
//   1. per row y: idx_y = argmax_x |A(y,x)| (first maximum); the row is
//      circularly shifted by s_y = ceil(Nx/2) - idx_y so its maximum lands
//      on column ceil(Nx/2) (0-based), PGA.m:159-164;
//   2. window from the noncoherent column energy of the unshifted image:
//      w_x = sum_y |A(y,x)|^2, dB = 10 log10(w/max w); the mask keeps the
//      columns between the last below -db_down before the centre and the
//      first below it after (PGA.m:167-188);
//   3. G = fft(ifftshift(mask .* shifted rows)) along x (PGA.m:195-196);
//   4. phi_dot_k = sum_y Im(conj G_k G_{k+1}) / sum_y |G_k|^2 (PGA.m:197-201);
//   5. phi = unwrap([0 cumsum(phi_dot)]), linear trend removed by least
//      squares over x = 0..Nx-1, rms recorded, af_ph += phi (PGA.m:204-217);
//   6. A <- ifft(fft(A) .* e^{-j phi}) along x (PGA.m:229-232).
// GPU formulation: every pass reads rows in y-contiguous order; the shift,
// mask and ifftshift of step 3 fold into an index remap plus one phase factor
// per (k, y); step 6 is one circular convolution A <- A (*) m with
// m = ifft(e^{-j phi}).  Transforms are direct sums with exact integer
// twiddle indices (any Nx), FP32 products, FP64 reductions.
#include <algorithm>
#include <cmath>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

constexpr int kPgaMaxN = 4096;

__global__ __launch_bounds__(256) void pga_row_argmax(const float2 *__restrict__ A, int Nx, int Ny,
                                                      int *__restrict__ shift) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= Ny) return;
    float best = -1.f;
    int bi = 0;
    for (int x = 0; x < Nx; x++) {
        const float2 v = A[(size_t)x * Ny + y];
        const float m = v.x * v.x + v.y * v.y;
        if (m > best) {
            best = m;
            bi = x;
        }
    }
    const int center = (Nx + 1) / 2;  // 0-based target column = 1-based ceil(Nx/2)
    shift[y] = ((center - bi) % Nx + Nx) % Nx;
}

__global__ __launch_bounds__(256) void pga_col_energy(const float2 *__restrict__ A, int Nx, int Ny,
                                                      double *__restrict__ w) {
    __shared__ double red[256];
    const int x = blockIdx.x;
    double s = 0.0;
    for (int y = threadIdx.x; y < Ny; y += blockDim.x) {
        const float2 v = A[(size_t)x * Ny + y];
        s += (double)v.x * v.x + (double)v.y * v.y;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) w[x] = red[0];
}

// One thread: window limits from the column energy (mask[j] = 1 on
// [left, right], 0-based).
__global__ void pga_window(const double *__restrict__ w, int Nx, double db_down,
                           float *__restrict__ mask) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double wmax = 0.0;
    for (int x = 0; x < Nx; x++) wmax = fmax(wmax, w[x]);
    // dB relative to the maximum; max(dB) = 0, so the cut-off is -|db_down|
    const double cut = -fabs(db_down);
    const int center = (Nx + 1) / 2;  // 1-based ceil(Nx/2)
    int left = 0, right = Nx - 1;
    for (int j = center - 1; j >= 0; j--) {
        if (10.0 * log10(w[j] / wmax) < cut) {
            left = j + 1;
            break;
        }
    }
    for (int j = center; j < Nx; j++) {
        if (10.0 * log10(w[j] / wmax) < cut) {
            right = j - 1;
            break;
        }
    }
    for (int j = 0; j < Nx; j++) mask[j] = (j >= left && j <= right) ? 1.f : 0.f;
}

// G[k][y] = e^{-j 2 pi k (s_y - fl)/N} sum_x mask[(x + s_y) mod N] A[x][y] e^{-j 2 pi k x / N}
// (fl = floor(N/2)); a lane owns one y and KB consecutive k.
template <int KB>
__global__ __launch_bounds__(256) void pga_dft(const float2 *__restrict__ A, int Nx, int Ny,
                                               const int *__restrict__ shift,
                                               const float *__restrict__ mask,
                                               const float2 *__restrict__ tw /* e^{+j 2 pi i/N} */,
                                               float2 *__restrict__ G) {
    __shared__ float2 stw[kPgaMaxN];
    __shared__ float smask[kPgaMaxN];
    for (int i = threadIdx.x; i < Nx; i += blockDim.x) {
        stw[i] = tw[i];
        smask[i] = mask[i];
    }
    __syncthreads();
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int k0 = blockIdx.y * KB;
    if (y >= Ny) return;
    const int s = shift[y];
    float ar[KB], ai[KB];
    int idx[KB], step[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
        ar[q] = ai[q] = 0.f;
        idx[q] = 0;
        step[q] = (k0 + q) % Nx;  // k mod N keeps every index below N
    }
    int mi = s;  // (x + s) mod N
    for (int x = 0; x < Nx; x++) {
        const float m = smask[mi];
        if (++mi == Nx) mi = 0;
        if (m != 0.f) {
            const float2 v = A[(size_t)x * Ny + y];
#pragma unroll
            for (int q = 0; q < KB; q++) {
                const float2 w = stw[idx[q]];  // e^{+j 2 pi k x/N}; use its conjugate
                ar[q] = fmaf(v.x, w.x, fmaf(v.y, w.y, ar[q]));
                ai[q] = fmaf(v.y, w.x, fmaf(-v.x, w.y, ai[q]));
            }
        }
#pragma unroll
        for (int q = 0; q < KB; q++) {
            idx[q] += step[q];
            if (idx[q] >= Nx) idx[q] -= Nx;
        }
    }
    const int fl = Nx / 2;
#pragma unroll
    for (int q = 0; q < KB; q++) {
        const int k = k0 + q;
        if (k >= Nx) break;
        // e^{-j 2 pi k (s - fl) / N}
        const long e = ((long)k * (((s - fl) % Nx + Nx) % Nx)) % Nx;
        const float2 w = stw[e];
        const float pr = ar[q] * w.x + ai[q] * w.y;
        const float pi = ai[q] * w.x - ar[q] * w.y;
        G[(size_t)k * Ny + y] = make_float2(pr, pi);
    }
}

// num[k] = sum_y Im(conj(G_k) G_{k+1}), den[k] = sum_y |G_k|^2, k < N-1.
__global__ __launch_bounds__(256) void pga_phidot(const float2 *__restrict__ G, int Nx, int Ny,
                                                  double *__restrict__ phidot) {
    __shared__ double rn[256], rd[256];
    const int k = blockIdx.x;
    double num = 0.0, den = 0.0;
    for (int y = threadIdx.x; y < Ny; y += blockDim.x) {
        const float2 a = G[(size_t)k * Ny + y], b = G[(size_t)(k + 1) * Ny + y];
        num += (double)a.x * b.y - (double)a.y * b.x;
        den += (double)a.x * a.x + (double)a.y * a.y;
    }
    rn[threadIdx.x] = num;
    rd[threadIdx.x] = den;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            rn[threadIdx.x] += rn[threadIdx.x + o];
            rd[threadIdx.x] += rd[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) phidot[k] = rd[0] > 0.0 ? rn[0] / rd[0] : 0.0;
}

// One thread: phi = unwrap([0 cumsum(phidot)]), detrend, rms, af_ph += phi.
// Writes the detrended phase to phi_d.  unwrap follows numpy/MATLAB: a step
// d is replaced by d - 2 pi round-to-(-pi, pi].
__global__ void pga_integrate(const double *__restrict__ phidot, int Nx, double *__restrict__ phi_d,
                              double *__restrict__ af_ph, double *__restrict__ rms_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double twopi = 2.0 * kPi;
    double raw_prev = 0.0, un = 0.0, corr = 0.0;
    phi_d[0] = 0.0;
    double sx = 0, sy = 0, sxx = 0, sxy = 0;
    for (int i = 0; i < Nx; i++) {
        const double raw = (i == 0) ? 0.0 : raw_prev + phidot[i - 1];
        if (i > 0) {
            const double d = raw - raw_prev;
            double dd = fmod(d + kPi, twopi);
            if (dd < 0) dd += twopi;
            dd -= kPi;
            if (dd == -kPi && d > 0) dd = kPi;
            if (fabs(d) >= kPi) corr += dd - d;
        }
        un = raw + corr;
        phi_d[i] = un;
        raw_prev = raw;
        sx += i;
        sy += un;
        sxx += (double)i * i;
        sxy += (double)i * un;
    }
    const double n = Nx;
    const double den = n * sxx - sx * sx;
    const double slope = den != 0 ? (n * sxy - sx * sy) / den : 0.0;
    const double icpt = (sy - slope * sx) / n;
    double r = 0.0;
    for (int i = 0; i < Nx; i++) {
        const double v = phi_d[i] - (slope * i + icpt);
        phi_d[i] = v;
        af_ph[i] += v;
        r += v * v;
    }
    *rms_out = sqrt(r / n);
}

// m[x] = (1/N) sum_k e^{-j phi_k} e^{+j 2 pi k x / N}  (ifft of e^{-j phi}).
__global__ __launch_bounds__(256) void pga_kernel(const double *__restrict__ phi, int Nx,
                                                  const double2 *__restrict__ tw,
                                                  float2 *__restrict__ m) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= Nx) return;
    double ar = 0.0, ai = 0.0;
    long idx = 0;
    for (int k = 0; k < Nx; k++) {
        const double c = cos(phi[k]), s = -sin(phi[k]);
        const double2 w = tw[idx];
        ar += c * w.x - s * w.y;
        ai += c * w.y + s * w.x;
        idx += x;
        if (idx >= Nx) idx -= Nx;
    }
    m[x] = make_float2((float)(ar / Nx), (float)(ai / Nx));
}

// out[x][y] = sum_x' m[(x - x') mod N] A[x'][y]; a lane owns one y and XB
// consecutive output columns.
template <int XB>
__global__ __launch_bounds__(256) void pga_apply(const float2 *__restrict__ A, int Nx, int Ny,
                                                 const float2 *__restrict__ m,
                                                 float2 *__restrict__ out) {
    __shared__ float2 sm[kPgaMaxN];
    for (int i = threadIdx.x; i < Nx; i += blockDim.x) sm[i] = m[i];
    __syncthreads();
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int x0 = blockIdx.y * XB;
    if (y >= Ny) return;
    float ar[XB], ai[XB];
    int idx[XB];
#pragma unroll
    for (int q = 0; q < XB; q++) {
        ar[q] = ai[q] = 0.f;
        idx[q] = (x0 + q) % Nx;  // (x - x') mod N at x' = 0
    }
    for (int xp = 0; xp < Nx; xp++) {
        const float2 v = A[(size_t)xp * Ny + y];
#pragma unroll
        for (int q = 0; q < XB; q++) {
            const float2 w = sm[idx[q]];
            ar[q] = fmaf(v.x, w.x, fmaf(-v.y, w.y, ar[q]));
            ai[q] = fmaf(v.x, w.y, fmaf(v.y, w.x, ai[q]));
            if (--idx[q] < 0) idx[q] += Nx;
        }
    }
#pragma unroll
    for (int q = 0; q < XB; q++) {
        if (x0 + q < Nx) out[(size_t)(x0 + q) * Ny + y] = make_float2(ar[q], ai[q]);
    }
}

__global__ void zero_f64(double *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_pga(sarf_ctx *ctx, const sarf_pga_desc *d, void *img, double *af_ph,
                        double *rms, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    if (!d || !img || !af_ph || !rms) return fail(SARF_ERR_INVALID, "sarf_pga: null argument");
    const int Nx = d->Nx, Ny = d->Ny;
    if (Nx < 2 || Ny < 1 || d->iters < 0)
        return fail(SARF_ERR_INVALID, "sarf_pga: Nx=%d Ny=%d iters=%d", Nx, Ny, d->iters);
    if (Nx > kPgaMaxN)
        return fail(SARF_ERR_UNSUPPORTED, "sarf_pga: Nx=%d > %d", Nx, kPgaMaxN);
    hipStream_t s = (hipStream_t)stream;
    const size_t npix = (size_t)Nx * Ny;
    void *bufG, *bufT, *bufI;
    SARF_TRY(scratch_get(ctx, "pga.G", sizeof(float2) * npix, &bufG));
    SARF_TRY(scratch_get(ctx, "pga.tmp", sizeof(float2) * npix, &bufT));
    const size_t small = sizeof(int) * Ny + sizeof(double) * (3 * (size_t)Nx) +
                         sizeof(float) * Nx + sizeof(float2) * Nx + 256;
    SARF_TRY(scratch_get(ctx, "pga.small", small, &bufI));
    int *shift = (int *)bufI;
    double *w = (double *)(((uintptr_t)(shift + Ny) + 15) & ~(uintptr_t)15);
    double *phidot = w + Nx;
    double *phi = phidot + Nx;
    float *mask = (float *)(phi + Nx);
    float2 *m = (float2 *)(((uintptr_t)(mask + Nx) + 15) & ~(uintptr_t)15);
    const float2 *tw32 = nullptr;
    const double2 *tw64 = nullptr;
    SARF_TRY(twiddles_any_f32(ctx, Nx, &tw32));
    SARF_TRY(twiddles_any_f64(ctx, Nx, &tw64));
    float2 *A = (float2 *)img, *G = (float2 *)bufG, *T = (float2 *)bufT;
    hipLaunchKernelGGL(zero_f64, dim3((Nx + 255) / 256), dim3(256), 0, s, af_ph, Nx);
    constexpr int KB = 8;
    const dim3 gy((Ny + 255) / 256);
    for (int it = 0; it < d->iters; it++) {
        hipLaunchKernelGGL(pga_row_argmax, gy, dim3(256), 0, s, A, Nx, Ny, shift);
        hipLaunchKernelGGL(pga_col_energy, dim3(Nx), dim3(256), 0, s, A, Nx, Ny, w);
        hipLaunchKernelGGL(pga_window, dim3(1), dim3(64), 0, s, w, Nx, d->db_down, mask);
        hipLaunchKernelGGL((pga_dft<KB>), dim3((Ny + 255) / 256, (Nx + KB - 1) / KB), dim3(256), 0,
                           s, A, Nx, Ny, shift, mask, tw32, G);
        hipLaunchKernelGGL(pga_phidot, dim3(Nx - 1), dim3(256), 0, s, G, Nx, Ny, phidot);
        hipLaunchKernelGGL(pga_integrate, dim3(1), dim3(64), 0, s, phidot, Nx, phi, af_ph, rms + it);
        hipLaunchKernelGGL(pga_kernel, dim3((Nx + 255) / 256), dim3(256), 0, s, phi, Nx, tw64, m);
        hipLaunchKernelGGL((pga_apply<KB>), dim3((Ny + 255) / 256, (Nx + KB - 1) / KB), dim3(256), 0,
                           s, A, Nx, Ny, m, T);
        SARF_HIP(hipGetLastError());
        SARF_HIP(hipMemcpyAsync(A, T, sizeof(float2) * npix, hipMemcpyDeviceToDevice, s));
    }
    return SARF_OK;
}

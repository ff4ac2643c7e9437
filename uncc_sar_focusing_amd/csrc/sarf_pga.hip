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

constexpr int kArgSeg = 16;  // x segments per row in the first argmax pass

// Pass 1: per (segment, y) the first maximum of |A(y, x)|^2 over the
// segment's x range; grid (ceil(Ny/256), kArgSeg).
__global__ __launch_bounds__(256) void pga_row_argmax_seg(const float2 *__restrict__ A, int Nx,
                                                          int Ny, float *__restrict__ segmax,
                                                          int *__restrict__ segidx) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int seg = blockIdx.y;
    if (y >= Ny) return;
    const int per = (Nx + kArgSeg - 1) / kArgSeg;
    const int xb = seg * per, xe = min(Nx, xb + per);
    float best = -1.f;
    int bi = xb;
    for (int x = xb; x < xe; x++) {
        const float2 v = A[(size_t)x * Ny + y];
        const float m = v.x * v.x + v.y * v.y;
        if (m > best) {
            best = m;
            bi = x;
        }
    }
    segmax[(size_t)seg * Ny + y] = best;
    segidx[(size_t)seg * Ny + y] = bi;
}

// Pass 2: first maximum over the segments (in x order) -> circular shift
// s_y = ceil(Nx/2) - idx_y (mod Nx).
__global__ __launch_bounds__(256) void pga_row_shift(const float *__restrict__ segmax,
                                                     const int *__restrict__ segidx, int Nx,
                                                     int Ny, int *__restrict__ shift) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= Ny) return;
    float best = -1.f;
    int bi = 0;
    for (int s = 0; s < kArgSeg; s++) {
        const float m = segmax[(size_t)s * Ny + y];
        if (m > best) {
            best = m;
            bi = segidx[(size_t)s * Ny + y];
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

// One workgroup: window limits from the column energy (mask[j] = 1 on
// [left, right], 0-based): left = 1 + the last column before the centre
// below the cut-off, right = the first column from the centre on below it,
// minus 1 (parallel max/min reductions).
__global__ __launch_bounds__(1024) void pga_window(const double *__restrict__ w, int Nx,
                                                   double db_down, float *__restrict__ mask) {
    __shared__ double rmax[1024];
    __shared__ int rl[1024], rr[1024];
    const int tid = threadIdx.x;
    double m = 0.0;
    for (int x = tid; x < Nx; x += blockDim.x) m = fmax(m, w[x]);
    rmax[tid] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) rmax[tid] = fmax(rmax[tid], rmax[tid + o]);
        __syncthreads();
    }
    const double wmax = rmax[0];
    // dB relative to the maximum; max(dB) = 0, so the cut-off is -|db_down|
    const double cut = -fabs(db_down);
    const int center = (Nx + 1) / 2;  // 1-based ceil(Nx/2)
    int lastb = -1, firstb = Nx;      // last below in [0, center), first below in [center, Nx)
    for (int j = tid; j < Nx; j += blockDim.x) {
        if (10.0 * log10(w[j] / wmax) < cut) {
            if (j < center) lastb = max(lastb, j);
            else firstb = min(firstb, j);
        }
    }
    rl[tid] = lastb;
    rr[tid] = firstb;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) {
            rl[tid] = max(rl[tid], rl[tid + o]);
            rr[tid] = min(rr[tid], rr[tid + o]);
        }
        __syncthreads();
    }
    const int left = rl[0] + 1;                       // 0 when none is below
    const int right = rr[0] < Nx ? rr[0] - 1 : Nx - 1;
    for (int j = tid; j < Nx; j += blockDim.x) mask[j] = (j >= left && j <= right) ? 1.f : 0.f;
}

// G[k][y] = e^{-j 2 pi k (s_y - fl)/N} sum_x mask[(x + s_y) mod N] A[x][y] e^{-j 2 pi k x / N}
// (fl = floor(N/2)).  A workgroup covers 64 y (one per lane) and KB
// consecutive k; its four waves each sum a quarter of x and the quarters
// are added through LDS, so a wave streams at most ceil(N/4) columns.
template <int KB>
__global__ __launch_bounds__(256) void pga_dft(const float2 *__restrict__ A, int Nx, int Ny,
                                               const int *__restrict__ shift,
                                               const float *__restrict__ mask,
                                               const float2 *__restrict__ tw /* e^{+j 2 pi i/N} */,
                                               float2 *__restrict__ G) {
    __shared__ float2 stw[kPgaMaxN];
    __shared__ float smask[kPgaMaxN];
    __shared__ float2 red[3][KB][64];
    for (int i = threadIdx.x; i < Nx; i += blockDim.x) {
        stw[i] = tw[i];
        smask[i] = mask[i];
    }
    __syncthreads();
    const int ly = threadIdx.x & 63;
    const int part = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int y = blockIdx.x * 64 + ly;
    const int k0 = blockIdx.y * KB;
    const bool live = y < Ny;
    const int s = live ? shift[y] : 0;
    const int per = (Nx + 3) / 4;
    const int xb = min(Nx, part * per), xe = min(Nx, xb + per);
    float ar[KB], ai[KB];
    int idx[KB], step[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
        ar[q] = ai[q] = 0.f;
        step[q] = (k0 + q) % Nx;  // k mod N keeps every index below N
        idx[q] = (int)(((long)step[q] * xb) % Nx);
    }
    if (live) {
        int mi = (xb + s) % Nx;  // (x + s) mod N
        for (int x = xb; x < xe; x++) {
            const float m = smask[mi];
            if (++mi == Nx) mi = 0;
            float2 v = A[(size_t)x * Ny + y];
            v.x *= m;
            v.y *= m;
#pragma unroll
            for (int q = 0; q < KB; q++) {
                const float2 w = stw[idx[q]];  // e^{+j 2 pi k x/N}; use its conjugate
                ar[q] = fmaf(v.x, w.x, fmaf(v.y, w.y, ar[q]));
                ai[q] = fmaf(v.y, w.x, fmaf(-v.x, w.y, ai[q]));
                idx[q] += step[q];
                if (idx[q] >= Nx) idx[q] -= Nx;
            }
        }
    }
    if (part > 0) {
#pragma unroll
        for (int q = 0; q < KB; q++) red[part - 1][q][ly] = make_float2(ar[q], ai[q]);
    }
    __syncthreads();
    if (part != 0 || !live) return;
    const int fl = Nx / 2;
#pragma unroll
    for (int q = 0; q < KB; q++) {
        const int k = k0 + q;
        if (k >= Nx) break;
        const float2 p1 = red[0][q][ly], p2 = red[1][q][ly], p3 = red[2][q][ly];
        const float sr = (ar[q] + p1.x) + (p2.x + p3.x);
        const float si = (ai[q] + p1.y) + (p2.y + p3.y);
        // e^{-j 2 pi k (s - fl) / N}
        const long e = ((long)k * (((s - fl) % Nx + Nx) % Nx)) % Nx;
        const float2 w = stw[e];
        G[(size_t)k * Ny + y] = make_float2(sr * w.x + si * w.y, si * w.x - sr * w.y);
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

constexpr int kIntThreads = 1024;  // pga_integrate block; kPgaMaxN / 1024 = 4 per thread

// numpy/MATLAB unwrap of one step d: d - 2 pi round-to-(-pi, pi] when |d| >= pi.
__device__ inline double unwrap_step(double d) {
    if (fabs(d) < kPi) return d;
    const double twopi = 2.0 * kPi;
    double dd = fmod(d + kPi, twopi);
    if (dd < 0) dd += twopi;
    dd -= kPi;
    if (dd == -kPi && d > 0) dd = kPi;
    return dd;
}

__device__ inline double block_sum_1024(double v, double *red) {
    const int tid = threadIdx.x;
    __syncthreads();
    red[tid] = v;
    __syncthreads();
    for (int o = kIntThreads / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    return red[0];
}

// One workgroup of 1024: phi = unwrap([0 cumsum(phidot)]), detrend, rms,
// af_ph += phi; writes e^{-j phi} (detrended) for the correction kernel.
// unwrap of a cumulative sum is the cumulative sum of the unwrapped steps,
// so phi is one block-wide prefix scan of unwrap_step(phidot).
__global__ __launch_bounds__(kIntThreads) void pga_integrate(const double *__restrict__ phidot,
                                                             int Nx,
                                                             double2 *__restrict__ ephi,
                                                             double *__restrict__ af_ph,
                                                             double *__restrict__ rms_out) {
    __shared__ double sc[kIntThreads];
    __shared__ double red[kIntThreads];
    const int tid = threadIdx.x;
    const int per = (Nx + kIntThreads - 1) / kIntThreads;  // <= 4
    const int b = tid * per;
    double loc[4];
    double run = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = b + j;
        if (j < per && i < Nx && i > 0) run += unwrap_step(phidot[i - 1]);
        loc[j] = run;
    }
    sc[tid] = run;
    __syncthreads();
    for (int o = 1; o < kIntThreads; o <<= 1) {  // inclusive Hillis-Steele scan
        const double v = tid >= o ? sc[tid - o] : 0.0;
        __syncthreads();
        sc[tid] += v;
        __syncthreads();
    }
    const double off = tid > 0 ? sc[tid - 1] : 0.0;
    double sy = 0.0, sxy = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = b + j;
        loc[j] += off;
        if (j < per && i < Nx) {
            sy += loc[j];
            sxy += (double)i * loc[j];
        }
    }
    sy = block_sum_1024(sy, red);
    sxy = block_sum_1024(sxy, red);
    const double n = Nx;
    const double sx = 0.5 * n * (n - 1.0), sxx = (n - 1.0) * n * (2.0 * n - 1.0) / 6.0;
    const double den = n * sxx - sx * sx;
    const double slope = den != 0 ? (n * sxy - sx * sy) / den : 0.0;
    const double icpt = (sy - slope * sx) / n;
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = b + j;
        if (j < per && i < Nx) {
            const double v = loc[j] - (slope * i + icpt);
            af_ph[i] += v;
            double sn, cs;
            sincos(v, &sn, &cs);
            ephi[i] = make_double2(cs, -sn);
            r += v * v;
        }
    }
    r = block_sum_1024(r, red);
    if (tid == 0) *rms_out = sqrt(r / n);
}

// m[x] = (1/N) sum_k e^{-j phi_k} e^{+j 2 pi k x / N}  (ifft of e^{-j phi});
// one workgroup per x, k strided over its lanes.
__global__ __launch_bounds__(256) void pga_kernel(const double2 *__restrict__ ephi, int Nx,
                                                  const double2 *__restrict__ tw,
                                                  float2 *__restrict__ m) {
    __shared__ double rr[256], ri[256];
    const int x = blockIdx.x;
    const int tid = threadIdx.x;
    double ar = 0.0, ai = 0.0;
    long idx = ((long)tid * x) % Nx;
    const long inc = ((long)blockDim.x * x) % Nx;
    for (int k = tid; k < Nx; k += blockDim.x) {
        const double2 e = ephi[k];
        const double2 w = tw[idx];
        ar += e.x * w.x - e.y * w.y;
        ai += e.x * w.y + e.y * w.x;
        idx += inc;
        if (idx >= Nx) idx -= Nx;
    }
    rr[tid] = ar;
    ri[tid] = ai;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) {
            rr[tid] += rr[tid + o];
            ri[tid] += ri[tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) m[x] = make_float2((float)(rr[0] / Nx), (float)(ri[0] / Nx));
}

// out[x][y] = sum_x' m[(x - x') mod N] A[x'][y].  A workgroup covers 64 y
// and XB consecutive output columns; its four waves each sum a quarter of
// x' and the quarters are added through LDS.
template <int XB>
__global__ __launch_bounds__(256) void pga_apply(const float2 *__restrict__ A, int Nx, int Ny,
                                                 const float2 *__restrict__ m,
                                                 float2 *__restrict__ out) {
    __shared__ float2 sm[kPgaMaxN];
    __shared__ float2 red[3][XB][64];
    for (int i = threadIdx.x; i < Nx; i += blockDim.x) sm[i] = m[i];
    __syncthreads();
    const int ly = threadIdx.x & 63;
    const int part = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int y = blockIdx.x * 64 + ly;
    const int x0 = blockIdx.y * XB;
    const bool live = y < Ny;
    const int per = (Nx + 3) / 4;
    const int xb = min(Nx, part * per), xe = min(Nx, xb + per);
    float ar[XB], ai[XB];
    int idx[XB];
#pragma unroll
    for (int q = 0; q < XB; q++) {
        ar[q] = ai[q] = 0.f;
        idx[q] = ((x0 + q - xb) % Nx + Nx) % Nx;  // (x - x') mod N at x' = xb
    }
    if (live) {
        for (int xp = xb; xp < xe; xp++) {
            const float2 v = A[(size_t)xp * Ny + y];
#pragma unroll
            for (int q = 0; q < XB; q++) {
                const float2 w = sm[idx[q]];
                ar[q] = fmaf(v.x, w.x, fmaf(-v.y, w.y, ar[q]));
                ai[q] = fmaf(v.x, w.y, fmaf(v.y, w.x, ai[q]));
                if (--idx[q] < 0) idx[q] += Nx;
            }
        }
    }
    if (part > 0) {
#pragma unroll
        for (int q = 0; q < XB; q++) red[part - 1][q][ly] = make_float2(ar[q], ai[q]);
    }
    __syncthreads();
    if (part != 0 || !live) return;
#pragma unroll
    for (int q = 0; q < XB; q++) {
        if (x0 + q >= Nx) break;
        const float2 p1 = red[0][q][ly], p2 = red[1][q][ly], p3 = red[2][q][ly];
        out[(size_t)(x0 + q) * Ny + y] =
            make_float2((ar[q] + p1.x) + (p2.x + p3.x), (ai[q] + p1.y) + (p2.y + p3.y));
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
    auto align16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
    const size_t o_segmax = align16(sizeof(int) * Ny);
    const size_t o_segidx = o_segmax + align16(sizeof(float) * kArgSeg * Ny);
    const size_t o_w = o_segidx + align16(sizeof(int) * kArgSeg * Ny);
    const size_t o_phidot = o_w + sizeof(double) * Nx;
    const size_t o_ephi = o_phidot + sizeof(double) * Nx;
    const size_t o_mask = o_ephi + sizeof(double2) * Nx;
    const size_t o_m = o_mask + align16(sizeof(float) * Nx);
    const size_t small = o_m + sizeof(float2) * Nx;
    SARF_TRY(scratch_get(ctx, "pga.small", small, &bufI));
    char *base = (char *)bufI;
    int *shift = (int *)base;
    float *segmax = (float *)(base + o_segmax);
    int *segidx = (int *)(base + o_segidx);
    double *w = (double *)(base + o_w);
    double *phidot = (double *)(base + o_phidot);
    double2 *ephi = (double2 *)(base + o_ephi);
    float *mask = (float *)(base + o_mask);
    float2 *m = (float2 *)(base + o_m);
    const float2 *tw32 = nullptr;
    const double2 *tw64 = nullptr;
    SARF_TRY(twiddles_any_f32(ctx, Nx, &tw32));
    SARF_TRY(twiddles_any_f64(ctx, Nx, &tw64));
    float2 *G = (float2 *)bufG;
    // ping-pong between the image and the scratch copy; the result is
    // copied back to the image once at the end when it lands in the copy
    float2 *src = (float2 *)img, *dst = (float2 *)bufT;
    hipLaunchKernelGGL(zero_f64, dim3((Nx + 255) / 256), dim3(256), 0, s, af_ph, Nx);
    const bool wide = npix >= ((size_t)1 << 18);
    const int KB = wide ? 16 : 8;
    const dim3 gy((Ny + 255) / 256);
    const dim3 gxy((Ny + 63) / 64, (Nx + KB - 1) / KB);
    for (int it = 0; it < d->iters; it++) {
        hipLaunchKernelGGL(pga_row_argmax_seg, dim3(gy.x, kArgSeg), dim3(256), 0, s, src, Nx, Ny,
                           segmax, segidx);
        hipLaunchKernelGGL(pga_row_shift, gy, dim3(256), 0, s, segmax, segidx, Nx, Ny, shift);
        hipLaunchKernelGGL(pga_col_energy, dim3(Nx), dim3(256), 0, s, src, Nx, Ny, w);
        hipLaunchKernelGGL(pga_window, dim3(1), dim3(1024), 0, s, w, Nx, d->db_down, mask);
        if (wide)
            hipLaunchKernelGGL((pga_dft<16>), gxy, dim3(256), 0, s, src, Nx, Ny, shift, mask, tw32, G);
        else
            hipLaunchKernelGGL((pga_dft<8>), gxy, dim3(256), 0, s, src, Nx, Ny, shift, mask, tw32, G);
        hipLaunchKernelGGL(pga_phidot, dim3(Nx - 1), dim3(256), 0, s, G, Nx, Ny, phidot);
        hipLaunchKernelGGL(pga_integrate, dim3(1), dim3(kIntThreads), 0, s, phidot, Nx, ephi, af_ph,
                           rms + it);
        hipLaunchKernelGGL(pga_kernel, dim3(Nx), dim3(256), 0, s, ephi, Nx, tw64, m);
        if (wide)
            hipLaunchKernelGGL((pga_apply<16>), gxy, dim3(256), 0, s, src, Nx, Ny, m, dst);
        else
            hipLaunchKernelGGL((pga_apply<8>), gxy, dim3(256), 0, s, src, Nx, Ny, m, dst);
        SARF_HIP(hipGetLastError());
        std::swap(src, dst);
    }
    if (src != (float2 *)img)
        SARF_HIP(hipMemcpyAsync(img, src, sizeof(float2) * npix, hipMemcpyDeviceToDevice, s));
    return SARF_OK;
}

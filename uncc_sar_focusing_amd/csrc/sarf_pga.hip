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
//
// GPU formulation.  The image is transposed once into rows along x
// (At[y*Nx + x]) for all iterations and transposed back at the end.  Every
// transform along x is a length-N DFT (any N <= 4096) done by Bluestein's
// chirp-z identity inside one workgroup per row: nk = (n^2 + k^2 - (k-n)^2)/2
// turns the DFT into a circular convolution with the chirp
// w_m = exp(-j pi m^2 / N), evaluated with two power-of-two FFTs of length
// M >= 2N - 1 (M >= 32).  From M = 1024, M/16 lanes per row hold 16 values
// each in registers and run self-sorting radix-16 passes through a padded LDS
// row (two rows per workgroup at M = 1024, 2048); one transform's output
// feeds the next in registers.  Shorter rows use one 256-thread workgroup per
// row with the row in LDS (radix-2, in place), which has the lower latency
// there.  FFT(chirp) is computed once per call (radix-2 in LDS).  The
// shift, mask and ifftshift of step 3 fold into the load (index remap) and
// one phase factor per (k, y); step 6 is a forward transform, the e^{-j phi}
// multiply and an inverse transform in the same workgroup, in place.  Reductions over rows (column energy, phase
// differences) are two-pass FP64 sums.
#include <algorithm>
#include <cmath>

#include "sarf_fft.hpp"

namespace sarf {
namespace {

constexpr int kPgaMaxN = 4096;
constexpr int kThreads = 256;
constexpr int kRowChunks = 16;  // y chunks of the two-pass column reductions

// In-place radix-2 decimation-in-time FFT of buf[0..M), M = 2^L, by one
// workgroup of kThreads: forward uses e^{-j 2 pi t/M}, inverse e^{+j 2 pi t/M}
// (unnormalised).  tw[t] = e^{+j 2 pi t / M}, t < M.
__device__ void lds_fft(float2 *buf, int L, const float2 *__restrict__ tw, bool inverse) {
    const int M = 1 << L;
    const int tid = threadIdx.x;
    for (int i = tid; i < M; i += kThreads) {
        const int r = (int)(__brev((unsigned)i) >> (32 - L));
        if (r > i) {
            const float2 t = buf[i];
            buf[i] = buf[r];
            buf[r] = t;
        }
    }
    __syncthreads();
    for (int s = 1; s <= L; s++) {
        const int h = 1 << (s - 1);
        const int tstride = M >> s;
        for (int b = tid; b < M / 2; b += kThreads) {
            const int j = b & (h - 1);
            const int i0 = ((b >> (s - 1)) << s) + j;
            const int i1 = i0 + h;
            float2 w = tw[j * tstride];
            if (!inverse) w.y = -w.y;
            const float2 u = buf[i0];
            const float2 v = cmul(buf[i1], w);
            buf[i0] = make_float2(u.x + v.x, u.y + v.y);
            buf[i1] = make_float2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
}

// Bluestein chirp w_m = e^{-j pi m^2 / N} from ch2n[t] = e^{+j 2 pi t/(2N)}.
__device__ __forceinline__ float2 chirp(int m, int N, const float2 *__restrict__ ch2n) {
    const long t = ((long)m * m) % (2L * N);
    const float2 c = ch2n[t];
    return make_float2(c.x, -c.y);
}

// buf[0..N) holds a_n = x_n w_n, buf[N..M) zeros; on return buf[k] (k < N)
// holds X_k / w_k * M, i.e. X_k = w_k buf[k] / M is the DFT of x.
__device__ __forceinline__ void bluestein_core(float2 *buf, int L, const float2 *__restrict__ twM,
                                               const float2 *__restrict__ H) {
    const int M = 1 << L;
    lds_fft(buf, L, twM, false);
    for (int i = threadIdx.x; i < M; i += kThreads) buf[i] = cmul(buf[i], H[i]);
    __syncthreads();
    lds_fft(buf, L, twM, true);
}

// H = FFT_M(h), h_m = conj(w_m) for m < N, h_{M-m} = conj(w_m) for 0 < m < N.
__global__ __launch_bounds__(kThreads) void pga_chirp_spectrum(int N, int L,
                                                               const float2 *__restrict__ ch2n,
                                                               const float2 *__restrict__ twM,
                                                               float2 *__restrict__ H) {
    extern __shared__ float2 buf[];
    const int M = 1 << L;
    for (int i = threadIdx.x; i < M; i += kThreads) {
        float2 v = make_float2(0.f, 0.f);
        if (i < N) v = chirp(i, N, ch2n);
        else if (i > M - N) v = chirp(M - i, N, ch2n);
        buf[i] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    lds_fft(buf, L, twM, false);
    for (int i = threadIdx.x; i < M; i += kThreads) H[i] = buf[i];
}

// dst[c*R + r] = src[r*C + c] for an R x C complex matrix (64 x 64 tiles).
__global__ __launch_bounds__(kThreads) void transpose_c64(const float2 *__restrict__ src, int R,
                                                          int C, float2 *__restrict__ dst) {
    __shared__ float2 tile[64][65];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    for (int i = ly; i < 64; i += kThreads / 64) {
        const int r = r0 + i, c = c0 + lx;
        if (r < R && c < C) tile[i][lx] = src[(size_t)r * C + c];
    }
    __syncthreads();
    for (int i = ly; i < 64; i += kThreads / 64) {
        const int c = c0 + i, r = r0 + lx;
        if (r < R && c < C) dst[(size_t)c * R + r] = tile[lx][i];
    }
}

// Per row y (one workgroup): first maximum of |At[y][x]|^2 over x ->
// circular shift s_y = ceil(N/2) - idx (mod N).
__global__ __launch_bounds__(kThreads) void pga_row_argmax(const float2 *__restrict__ At, int N,
                                                           int *__restrict__ shift) {
    __shared__ float bm[kThreads];
    __shared__ int bi[kThreads];
    const int y = blockIdx.x, tid = threadIdx.x;
    const float2 *row = At + (size_t)y * N;
    float best = -1.f;
    int idx = 0;
    for (int x = tid; x < N; x += kThreads) {
        const float2 v = row[x];
        const float m = v.x * v.x + v.y * v.y;
        if (m > best) {
            best = m;
            idx = x;
        }
    }
    bm[tid] = best;
    bi[tid] = idx;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
        if (tid < o) {
            const float m2 = bm[tid + o];
            const int i2 = bi[tid + o];
            if (m2 > bm[tid] || (m2 == bm[tid] && i2 < bi[tid])) {
                bm[tid] = m2;
                bi[tid] = i2;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int center = (N + 1) / 2;  // 0-based target column = 1-based ceil(N/2)
        shift[y] = ((center - bi[0]) % N + N) % N;
    }
}

// Two-pass column sums over rows.  Pass 1, grid (ceil(N/64), kRowChunks):
// part[chunk*N + x] = sum over the chunk's rows of |At[y][x]|^2 (mode 0), or
// for G (mode 1) num/den parts of phi_dot: Im(conj G_k G_{k+1}), |G_k|^2.
__global__ __launch_bounds__(kThreads) void pga_col_partial(const float2 *__restrict__ At, int N,
                                                            int Ny, int mode,
                                                            double *__restrict__ pa,
                                                            double *__restrict__ pb) {
    __shared__ double ra[4][64], rb[4][64];
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lx;
    const int per = (Ny + kRowChunks - 1) / kRowChunks;
    const int yb = blockIdx.y * per, ye = min(Ny, yb + per);
    double a = 0.0, b = 0.0;
    if (x < N) {
        for (int y = yb + ly; y < ye; y += 4) {
            const float2 g = At[(size_t)y * N + x];
            if (mode == 0) {
                a += (double)g.x * g.x + (double)g.y * g.y;
            } else if (x + 1 < N) {
                const float2 g1 = At[(size_t)y * N + x + 1];
                a += (double)g.x * g1.y - (double)g.y * g1.x;
                b += (double)g.x * g.x + (double)g.y * g.y;
            }
        }
    }
    ra[ly][lx] = a;
    rb[ly][lx] = b;
    __syncthreads();
    if (ly == 0 && x < N) {
        pa[(size_t)blockIdx.y * N + x] = (ra[0][lx] + ra[1][lx]) + (ra[2][lx] + ra[3][lx]);
        if (mode == 1) pb[(size_t)blockIdx.y * N + x] = (rb[0][lx] + rb[1][lx]) + (rb[2][lx] + rb[3][lx]);
    }
}

// Pass 2: out[x] = sum of the chunk parts (mode 0: energy), or
// phi_dot[k] = num / den for k < N-1 (mode 1).
__global__ __launch_bounds__(kThreads) void pga_col_final(const double *__restrict__ pa,
                                                          const double *__restrict__ pb, int N,
                                                          int mode, double *__restrict__ out) {
    const int x = blockIdx.x * kThreads + threadIdx.x;
    if (x >= N) return;
    double a = 0.0, b = 0.0;
    for (int c = 0; c < kRowChunks; c++) {
        a += pa[(size_t)c * N + x];
        if (mode == 1) b += pb[(size_t)c * N + x];
    }
    if (mode == 0) out[x] = a;
    else if (x < N - 1) out[x] = b > 0.0 ? a / b : 0.0;
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

// Short rows (M <= 512): one workgroup of kThreads per row, the row in LDS
// (dynamic, M float2), radix-2 transforms -- more lanes per row than the
// register-resident kernels give, so lower latency at these lengths.
// Step 3 for row y:
// G[y][k] = e^{-j 2 pi k (s - fl)/N} DFT_x( mask[(x + s) mod N] At[y][x] )_k,
// fl = floor(N/2): the shift by s, the mask and ifftshift as one remap.
__global__ __launch_bounds__(kThreads) void pga_row_dft_lds(const float2 *__restrict__ At, int N, int L,
                                                        const int *__restrict__ shift,
                                                        const float *__restrict__ mask,
                                                        const float2 *__restrict__ ch2n,
                                                        const float2 *__restrict__ twM,
                                                        const float2 *__restrict__ twN,
                                                        const float2 *__restrict__ H,
                                                        float2 *__restrict__ G) {
    extern __shared__ float2 buf[];
    const int y = blockIdx.x, M = 1 << L;
    const int s = shift[y];
    const float2 *row = At + (size_t)y * N;
    for (int x = threadIdx.x; x < M; x += kThreads) {
        float2 v = make_float2(0.f, 0.f);
        if (x < N) {
            int mi = x + s;
            if (mi >= N) mi -= N;
            const float m = mask[mi];
            const float2 a = row[x];
            v = cmul(make_float2(m * a.x, m * a.y), chirp(x, N, ch2n));
        }
        buf[x] = v;
    }
    __syncthreads();
    bluestein_core(buf, L, twM, H);
    const float invM = 1.0f / (float)M;
    const int fl = N / 2;
    const long sh = ((long)(s - fl) % N + N) % N;
    for (int k = threadIdx.x; k < N; k += kThreads) {
        float2 X = cmul(buf[k], chirp(k, N, ch2n));
        X.x *= invM;
        X.y *= invM;
        const float2 e = twN[((long)k * sh) % N];  // e^{+j 2 pi k sh / N}; use its conjugate
        G[(size_t)y * N + k] = cmul(X, make_float2(e.x, -e.y));
    }
}

// Step 6 for row y, in place: At[y] <- IDFT( DFT(At[y]) .* e^{-j phi} ),
// the inverse as conj(DFT(conj(.)))/N through the same Bluestein core.
__global__ __launch_bounds__(kThreads) void pga_row_apply_lds(float2 *__restrict__ At, int N, int L,
                                                          const double2 *__restrict__ ephi,
                                                          const float2 *__restrict__ ch2n,
                                                          const float2 *__restrict__ twM,
                                                          const float2 *__restrict__ H) {
    extern __shared__ float2 buf[];
    const int y = blockIdx.x, M = 1 << L;
    float2 *row = At + (size_t)y * N;
    for (int x = threadIdx.x; x < M; x += kThreads)
        buf[x] = x < N ? cmul(row[x], chirp(x, N, ch2n)) : make_float2(0.f, 0.f);
    __syncthreads();
    bluestein_core(buf, L, twM, H);
    const float invM = 1.0f / (float)M;
    // Z_k = X_k e^{-j phi_k}; next input conj(Z_k) w_k, zero-padded (each
    // lane reads and rewrites only its own elements)
    for (int k = threadIdx.x; k < M; k += kThreads) {
        float2 v = make_float2(0.f, 0.f);
        if (k < N) {
            const float2 w = chirp(k, N, ch2n);
            const float2 X = cmul(buf[k], w);
            const double2 e = ephi[k];
            const float2 Z = cmul(make_float2(X.x * invM, X.y * invM),
                                  make_float2((float)e.x, (float)e.y));
            v = cmul(make_float2(Z.x, -Z.y), w);
        }
        buf[k] = v;
    }
    __syncthreads();
    bluestein_core(buf, L, twM, H);
    const float scale = invM / (float)N;
    for (int n = threadIdx.x; n < N; n += kThreads) {
        const float2 u = cmul(buf[n], chirp(n, N, ch2n));
        row[n] = make_float2(u.x * scale, -u.y * scale);
    }
}

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Bluestein's middle step on the forward transform of the chirped row held
// in pass-0 order (v = conj of the input, as ifft_regs computes the forward
// transform as conj(ifft(conj))): multiply by H and run the inverse
// transform; on return v[r] = b_k (k = j + r*T) and X_k = w_k b_k / M.
template <int LOG2M>
__device__ __forceinline__ void bluestein_regs(float2 (&v)[16], float2 *Lr, int j,
                                               const float2 *__restrict__ twM,
                                               const float2 *__restrict__ H) {
    constexpr int T = fft::Shape<LOG2M>::T;
    fft::ifft_regs<LOG2M>(v, Lr, j, twM);  // v[r] = conj(A_k)
    float2 u[16];
#pragma unroll
    for (int r = 0; r < 16; r++) u[r] = cmul(conjf2(v[r]), H[j + r * T]);
    fft::to_input_order<LOG2M>(u, v);
    fft::ifft_regs<LOG2M>(v, Lr, j, twM);
}

// Step 3 for rows y (P rows per workgroup, M/16 lanes each, M = 2^LOG2M):
// G[y][k] = e^{-j 2 pi k (s - fl)/N} DFT_x( mask[(x + s) mod N] At[y][x] )_k,
// fl = floor(N/2): the shift by s, the mask and ifftshift as one remap.
template <int LOG2M>
__global__ __launch_bounds__(fft::group_threads(LOG2M)) void pga_row_dft(
    const float2 *__restrict__ At, int N, int Ny, const int *__restrict__ shift,
    const float *__restrict__ mask, const float2 *__restrict__ ch2n,
    const float2 *__restrict__ twM, const float2 *__restrict__ twN,
    const float2 *__restrict__ H, float2 *__restrict__ G) {
    using S = fft::Shape<LOG2M>;
    constexpr int M = S::N, T = S::T, P = fft::rows_per_group(LOG2M), PAD = fft::padded(LOG2M);
    __shared__ float2 lds[P * PAD];
    const int lp = threadIdx.x / T, j = threadIdx.x % T;
    const int y = blockIdx.x * P + lp;
    const bool active = y < Ny;
    float2 *Lr = lds + lp * PAD;
    const int s = active ? shift[y] : 0;
    const float2 *row = At + (size_t)(active ? y : 0) * N;
    float2 v[16];
#pragma unroll
    for (int b = 0; b < S::NB0; b++)
#pragma unroll
        for (int r = 0; r < S::R0; r++) {
            const int n = j + b * T + r * (M / S::R0);
            float2 a = make_float2(0.f, 0.f);
            if (active && n < N) {
                int mi = n + s;
                if (mi >= N) mi -= N;
                const float m = mask[mi];
                const float2 x = row[n];
                a = cmul(make_float2(m * x.x, m * x.y), chirp(n, N, ch2n));
            }
            v[b * S::R0 + r] = conjf2(a);
        }
    bluestein_regs<LOG2M>(v, Lr, j, twM, H);
    const float invM = 1.0f / (float)M;
    const int fl = N / 2;
    const long sh = ((long)(s - fl) % N + N) % N;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int k = j + r * T;
        if (active && k < N) {
            float2 X = cmul(v[r], chirp(k, N, ch2n));
            X.x *= invM;
            X.y *= invM;
            const float2 e = twN[((long)k * sh) % N];  // e^{+j 2 pi k sh / N}; use its conjugate
            G[(size_t)y * N + k] = cmul(X, conjf2(e));
        }
    }
}

// Step 6 for rows y, in place: At[y] <- IDFT( DFT(At[y]) .* e^{-j phi} ),
// the inverse as conj(DFT(conj(.)))/N through the same Bluestein step.
template <int LOG2M>
__global__ __launch_bounds__(fft::group_threads(LOG2M)) void pga_row_apply(
    float2 *__restrict__ At, int N, int Ny, const double2 *__restrict__ ephi,
    const float2 *__restrict__ ch2n, const float2 *__restrict__ twM,
    const float2 *__restrict__ H) {
    using S = fft::Shape<LOG2M>;
    constexpr int M = S::N, T = S::T, P = fft::rows_per_group(LOG2M), PAD = fft::padded(LOG2M);
    __shared__ float2 lds[P * PAD];
    const int lp = threadIdx.x / T, j = threadIdx.x % T;
    const int y = blockIdx.x * P + lp;
    const bool active = y < Ny;
    float2 *Lr = lds + lp * PAD;
    float2 *row = At + (size_t)(active ? y : 0) * N;
    float2 v[16];
#pragma unroll
    for (int b = 0; b < S::NB0; b++)
#pragma unroll
        for (int r = 0; r < S::R0; r++) {
            const int n = j + b * T + r * (M / S::R0);
            const float2 a = (active && n < N) ? cmul(row[n], chirp(n, N, ch2n))
                                                : make_float2(0.f, 0.f);
            v[b * S::R0 + r] = conjf2(a);
        }
    bluestein_regs<LOG2M>(v, Lr, j, twM, H);
    const float invM = 1.0f / (float)M;
    // Z_k = X_k e^{-j phi_k}; the next forward transform's input is
    // conj(Z_k) w_k, zero-padded, fed conjugated in pass-0 order
    float2 u[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int k = j + r * T;
        float2 nx = make_float2(0.f, 0.f);
        if (k < N) {
            const float2 w = chirp(k, N, ch2n);
            const float2 X = cmul(v[r], w);
            const double2 e = ephi[k];
            const float2 Z = cmul(make_float2(X.x * invM, X.y * invM),
                                  make_float2((float)e.x, (float)e.y));
            nx = cmul(conjf2(Z), w);
        }
        u[r] = conjf2(nx);
    }
    fft::to_input_order<LOG2M>(u, v);
    bluestein_regs<LOG2M>(v, Lr, j, twM, H);
    const float scale = invM / (float)N;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int n = j + r * T;
        if (active && n < N) {
            const float2 uo = cmul(v[r], chirp(n, N, ch2n));
            row[n] = make_float2(uo.x * scale, -uo.y * scale);
        }
    }
}

template <int LOG2M>
void launch_rows(int Ny, bool apply, float2 *At, int N, const int *shift, const float *mask,
                 const double2 *ephi, const float2 *ch2n, const float2 *twM, const float2 *twN,
                 const float2 *H, float2 *G, hipStream_t s) {
    constexpr int P = fft::rows_per_group(LOG2M);
    const dim3 grid((Ny + P - 1) / P), block(fft::group_threads(LOG2M));
    if (apply)
        hipLaunchKernelGGL(pga_row_apply<LOG2M>, grid, block, 0, s, At, N, Ny, ephi, ch2n, twM, H);
    else
        hipLaunchKernelGGL(pga_row_dft<LOG2M>, grid, block, 0, s, At, N, Ny, shift, mask, ch2n,
                           twM, twN, H, G);
}

// Row transforms (steps 3 and 6) for a transform length 2^L, 5 <= L <= 13:
// one workgroup per row in LDS up to L = 9 (measured faster there), the
// register-resident kernels from L = 10.
int launch_row_kernel(int L, int Ny, bool apply, float2 *At, int N, const int *shift,
                      const float *mask, const double2 *ephi, const float2 *ch2n,
                      const float2 *twM, const float2 *twN, const float2 *H, float2 *G,
                      hipStream_t s) {
    if (L <= 9) {
        const size_t lds = sizeof(float2) * ((size_t)1 << L);
        if (apply)
            hipLaunchKernelGGL(pga_row_apply_lds, dim3(Ny), dim3(kThreads), lds, s, At, N, L, ephi,
                               ch2n, twM, H);
        else
            hipLaunchKernelGGL(pga_row_dft_lds, dim3(Ny), dim3(kThreads), lds, s, At, N, L, shift,
                               mask, ch2n, twM, twN, H, G);
        return SARF_OK;
    }
    switch (L) {
#define SARF_PGA_ROWS(l) \
        case l: launch_rows<l>(Ny, apply, At, N, shift, mask, ephi, ch2n, twM, twN, H, G, s); break;
        SARF_PGA_ROWS(10) SARF_PGA_ROWS(11) SARF_PGA_ROWS(12) SARF_PGA_ROWS(13)
#undef SARF_PGA_ROWS
        default: return fail(SARF_ERR_UNSUPPORTED, "sarf_pga: transform length 2^%d", L);
    }
    return SARF_OK;
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
// af_ph += phi; writes e^{-j phi} (detrended) for the correction.
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
    const int N = Nx;
    int L = 5;  // the row kernels take 32 <= M <= 8192
    while ((1 << L) < 2 * N - 1) L++;
    const int M = 1 << L;  // <= 8192
    const size_t npix = (size_t)Nx * Ny;
    const size_t lds = sizeof(float2) * (size_t)M;
    void *bufA, *bufG, *bufS;
    SARF_TRY(scratch_get(ctx, "pga.At", sizeof(float2) * npix, &bufA));
    SARF_TRY(scratch_get(ctx, "pga.G", sizeof(float2) * npix, &bufG));
    auto align16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
    const size_t o_parta = align16(sizeof(int) * Ny);
    const size_t o_partb = o_parta + sizeof(double) * kRowChunks * N;
    const size_t o_w = o_partb + sizeof(double) * kRowChunks * N;
    const size_t o_phidot = o_w + sizeof(double) * N;
    const size_t o_ephi = o_phidot + sizeof(double) * N;
    const size_t o_mask = o_ephi + sizeof(double2) * N;
    const size_t o_H = o_mask + align16(sizeof(float) * N);
    const size_t small = o_H + sizeof(float2) * M;
    SARF_TRY(scratch_get(ctx, "pga.small", small, &bufS));
    char *base = (char *)bufS;
    int *shift = (int *)base;
    double *pa = (double *)(base + o_parta);
    double *pb = (double *)(base + o_partb);
    double *w = (double *)(base + o_w);
    double *phidot = (double *)(base + o_phidot);
    double2 *ephi = (double2 *)(base + o_ephi);
    float *mask = (float *)(base + o_mask);
    float2 *H = (float2 *)(base + o_H);
    const float2 *twM = nullptr, *twN = nullptr, *ch2n = nullptr;
    SARF_TRY(twiddles_any_f32(ctx, M, &twM));
    SARF_TRY(twiddles_any_f32(ctx, N, &twN));
    SARF_TRY(twiddles_any_f32(ctx, 2L * N, &ch2n));
    float2 *A = (float2 *)img, *At = (float2 *)bufA, *G = (float2 *)bufG;

    hipLaunchKernelGGL(zero_f64, dim3((Nx + 255) / 256), dim3(256), 0, s, af_ph, Nx);
    if (d->iters == 0) {
        SARF_HIP(hipGetLastError());
        return SARF_OK;
    }
    hipLaunchKernelGGL(pga_chirp_spectrum, dim3(1), dim3(kThreads), lds, s, N, L, ch2n, twM, H);
    // img is Nx rows (x) of Ny values; At is Ny rows of Nx values
    hipLaunchKernelGGL(transpose_c64, dim3((Ny + 63) / 64, (Nx + 63) / 64), dim3(kThreads), 0, s,
                       A, Nx, Ny, At);
    const dim3 gcol((N + 63) / 64, kRowChunks);
    for (int it = 0; it < d->iters; it++) {
        hipLaunchKernelGGL(pga_row_argmax, dim3(Ny), dim3(kThreads), 0, s, At, N, shift);
        hipLaunchKernelGGL(pga_col_partial, gcol, dim3(kThreads), 0, s, At, N, Ny, 0, pa, pb);
        hipLaunchKernelGGL(pga_col_final, dim3((N + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                           pa, pb, N, 0, w);
        hipLaunchKernelGGL(pga_window, dim3(1), dim3(1024), 0, s, w, N, d->db_down, mask);
        SARF_TRY(launch_row_kernel(L, Ny, false, At, N, shift, mask, ephi, ch2n, twM, twN, H, G,
                                   s));
        hipLaunchKernelGGL(pga_col_partial, gcol, dim3(kThreads), 0, s, G, N, Ny, 1, pa, pb);
        hipLaunchKernelGGL(pga_col_final, dim3((N + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                           pa, pb, N, 1, phidot);
        hipLaunchKernelGGL(pga_integrate, dim3(1), dim3(kIntThreads), 0, s, phidot, N, ephi, af_ph,
                           rms + it);
        SARF_TRY(launch_row_kernel(L, Ny, true, At, N, shift, mask, ephi, ch2n, twM, twN, H, G,
                                   s));
        SARF_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(transpose_c64, dim3((Nx + 63) / 64, (Ny + 63) / 64), dim3(kThreads), 0, s,
                       At, Ny, Nx, A);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

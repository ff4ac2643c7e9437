// sarf_pga.hip -- phase gradient autofocus (PGA) on focused images.
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
// chirp-z identity: nk = (n^2 + k^2 - (k-n)^2)/2 turns the DFT into a
// circular convolution with the chirp w_m = exp(-j pi m^2 / N), evaluated
// with two power-of-two FFTs of length M >= 2N - 1 (M >= 32).  From
// M = 1024, M/16 lanes per row hold 16 values each in registers and run
// self-sorting radix-16 passes through a padded LDS row (4 and 2 rows per
// workgroup at M = 1024, 2048); one transform's output feeds the next in
// registers.  Shorter rows use one 256-thread workgroup per row with the row
// in LDS (radix-2, in place), which has the lower latency there.  The chirp
// and FFT(chirp) are tables per row length, evaluated once in FP64 on the
// host and kept in the context.  The shift, mask and
// ifftshift of step 3 fold into the load (index remap) and one phase factor
// per (k, y); step 6 is a forward transform, the e^{-j phi} multiply and an
// inverse transform in the same workgroup, in place.  Reductions over rows
// (column energy, phase differences) are two-pass FP64 sums.
//
// Every step is a device body on one image's state (PgaImage).  sarf_pga
// launches the bodies for one image with the state as a kernel argument;
// sarf_pga_scenes launches each step once for a whole batch of images of
// different sizes: the state of every image sits in device memory, each
// launch covers the concatenated blocks of all images (a block finds its
// image by binary search over the per-image block offsets), and the row
// transforms run one launch per transform length.  Both run the same
// arithmetic in the same order, so their results are identical.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sarf_fft.hpp"

namespace sarf {
namespace {

constexpr int kPgaMaxN = 4096;
constexpr int kThreads = 256;
constexpr int kRowChunks = 16;     // y chunks of the two-pass column reductions
constexpr int kIntThreads = 1024;  // window block
constexpr int kSpecThreads = 512;  // integrate block: kPgaMaxN / 512 = 8 values per thread,
                                   // and the M / 16 <= 512 lanes of the spectrum transforms
constexpr int kMinL = 5, kMaxL = 13;

// Device state of one image's PGA call.
struct PgaImage {
    float2 *img, *At, *G;  // image (Nx rows of Ny), its transpose, row spectra
    double *af_ph, *rms;   // outputs: Nx, iters
    int *shift;            // Ny
    double *pa, *pb;       // kRowChunks x Nx partial column sums
    float *mask;           // Nx
    float2 *B;             // M: spectrum of this iteration's correction kernel
    const float2 *wtab;    // N: chirp w_n = e^{-j pi n^2 / N} (context table)
    const float2 *H;       // M: FFT of the conjugate chirp (context table)
    const float2 *twM, *twN;  // e^{+j 2 pi t/M}, /N
    int *stop;             // 1 once an iteration's rms fell below rms_stop
    double rms_stop;       // > 0: early stop threshold; 0: fixed iteration count
    int Nx, Ny, L, iters;
};

// Early stop: every step after the iteration whose rms fell below the
// threshold returns at once (uniform per workgroup; kernel boundaries order
// the flag's write and reads).
__device__ __forceinline__ bool stopped(const PgaImage &p) {
    return p.rms_stop > 0.0 && *p.stop != 0;
}

// In-place radix-2 decimation-in-time FFT of buf[0..M), M = 2^L, by one
// workgroup (any size): forward uses e^{-j 2 pi t/M}, inverse e^{+j 2 pi t/M}
// (unnormalised).  tw[t] = e^{+j 2 pi t / M}, t < M.
__device__ void lds_fft(float2 *buf, int L, const float2 *__restrict__ tw, bool inverse) {
    const int M = 1 << L;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < M; i += nt) {
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
        for (int b = tid; b < M / 2; b += nt) {
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

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// buf[0..N) holds a_n = x_n w_n, buf[N..M) zeros; on return buf[k] (k < N)
// holds X_k / w_k * M, i.e. X_k = w_k buf[k] / M is the DFT of x.
__device__ __forceinline__ void bluestein_core(float2 *buf, int L, const float2 *__restrict__ twM,
                                               const float2 *__restrict__ H) {
    const int M = 1 << L;
    lds_fft(buf, L, twM, false);
    for (int i = threadIdx.x; i < M; i += blockDim.x) buf[i] = cmul(buf[i], H[i]);
    __syncthreads();
    lds_fft(buf, L, twM, true);
}

// ---------------------------------------------------------------- bodies

// af_ph = 0, the stop flag cleared and, with an rms threshold, rms = NaN
// (iterations that do not run report NaN); any workgroup size.
__device__ void init_body(const PgaImage &p) {
    for (int i = threadIdx.x; i < p.Nx; i += blockDim.x) p.af_ph[i] = 0.0;
    if (threadIdx.x == 0) *p.stop = 0;
    if (p.rms_stop > 0.0)
        for (int i = threadIdx.x; i < p.iters; i += blockDim.x) p.rms[i] = __builtin_nan("");
}

// dst[c*R + r] = src[r*C + c] for an R x C complex matrix; 64 x 64 tile
// (bx, by) = (column tile, row tile).
__device__ void transpose_body(const float2 *__restrict__ src, int R, int C,
                               float2 *__restrict__ dst, int bx, int by) {
    __shared__ float2 tile[64][65];
    constexpr int kStep = kThreads / 64, kRows = 64 / kStep;  // rows i = ly + k kStep
    const int r0 = by * 64, c0 = bx * 64;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    {
        // every load of the thread issued before the first is consumed
        // (a loop with the LDS store in it keeps one load in flight)
        const int c = c0 + lx;
        float2 v[kRows];
#pragma unroll
        for (int k = 0; k < kRows; k++) {
            const int r = r0 + ly + k * kStep;
            v[k] = (r < R && c < C) ? src[(size_t)r * C + c] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int k = 0; k < kRows; k++) tile[ly + k * kStep][lx] = v[k];
    }
    __syncthreads();
    const int r = r0 + lx;
#pragma unroll
    for (int k = 0; k < kRows; k++) {
        const int i = ly + k * kStep, c = c0 + i;
        if (r < R && c < C) dst[(size_t)c * R + r] = tile[lx][i];
    }
}

// Rows of the argmax blocks: one wave per row.
constexpr int kArgRows = kThreads / 64;
__host__ __device__ constexpr int argmax_blocks(int Ny) { return (Ny + kArgRows - 1) / kArgRows; }

// Argmax block blk, wave w: row y = blk * kArgRows + w, first maximum of
// |At[y][x]|^2 over x -> circular shift s_y = ceil(N/2) - idx (mod N).  Each
// lane scans x = lane, lane + 64, ...; the lanes then combine by butterflies
// keeping the larger value, or the smaller index of equal values, so the
// result is the row's first maximum whatever the lane order.  No barriers.
__device__ void argmax_body(const PgaImage &p, int blk) {
    if (stopped(p)) return;
    const int N = p.Nx, lane = threadIdx.x & 63;
    const int y = blk * kArgRows + (threadIdx.x >> 6);
    if (y >= p.Ny) return;
    const float2 *row = p.At + (size_t)y * N;
    float best = -1.f;
    int idx = 0;
    int x = lane;
    for (; x + 192 < N; x += 256) {  // four loads in flight, compared in x order
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = row[x + 64 * u];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const float m = v[u].x * v[u].x + v[u].y * v[u].y;
            if (m > best) {
                best = m;
                idx = x + 64 * u;
            }
        }
    }
    for (; x < N; x += 64) {
        const float2 v = row[x];
        const float m = v.x * v.x + v.y * v.y;
        if (m > best) {
            best = m;
            idx = x;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const float m2 = __shfl_xor(best, o);
        const int i2 = __shfl_xor(idx, o);
        if (m2 > best || (m2 == best && i2 < idx)) {
            best = m2;
            idx = i2;
        }
    }
    if (lane == 0) {
        const int center = (N + 1) / 2;  // 0-based target column = 1-based ceil(N/2)
        p.shift[y] = ((center - idx) % N + N) % N;
    }
}

// Two-pass column sums over rows.  Pass 1, block (bx, chunk):
// pa[chunk*N + x] = sum over the chunk's rows of |At[y][x]|^2 (mode 0), or
// for G (mode 1) num/den parts of phi_dot: Im(conj G_k G_{k+1}), |G_k|^2.
// Pass 2 (the sum over chunks, in chunk order) opens the window and
// integrate steps (window_fused_body, integrate_fused_body).
__device__ void col_partial_body(const PgaImage &p, int mode, int bx, int chunk) {
    if (stopped(p)) return;
    __shared__ double ra[4][64], rb[4][64];
    const float2 *src = mode == 0 ? p.At : p.G;
    const int N = p.Nx, Ny = p.Ny;
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int x = bx * 64 + lx;
    const int per = (Ny + kRowChunks - 1) / kRowChunks;
    const int yb = chunk * per, ye = min(Ny, yb + per);
    double a = 0.0, b = 0.0;
    if (x < N) {
        // rows y = yb + ly, yb + ly + 4, ... added in that order, four rows'
        // loads issued before any is consumed
        const int x1 = x + 1 < N ? x + 1 : x;
        int y = yb + ly;
        for (; y + 12 < ye; y += 16) {
            float2 g[4], g1[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                g[u] = src[(size_t)(y + 4 * u) * N + x];
                if (mode == 1) g1[u] = src[(size_t)(y + 4 * u) * N + x1];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                if (mode == 0) {
                    a += (double)g[u].x * g[u].x + (double)g[u].y * g[u].y;
                } else if (x + 1 < N) {
                    a += (double)g[u].x * g1[u].y - (double)g[u].y * g1[u].x;
                    b += (double)g[u].x * g[u].x + (double)g[u].y * g[u].y;
                }
            }
        }
        for (; y < ye; y += 4) {
            const float2 g = src[(size_t)y * N + x];
            if (mode == 0) {
                a += (double)g.x * g.x + (double)g.y * g.y;
            } else if (x + 1 < N) {
                const float2 g1 = src[(size_t)y * N + x + 1];
                a += (double)g.x * g1.y - (double)g.y * g1.x;
                b += (double)g.x * g.x + (double)g.y * g.y;
            }
        }
    }
    ra[ly][lx] = a;
    rb[ly][lx] = b;
    __syncthreads();
    if (ly == 0 && x < N) {
        p.pa[(size_t)chunk * N + x] = (ra[0][lx] + ra[1][lx]) + (ra[2][lx] + ra[3][lx]);
        if (mode == 1) p.pb[(size_t)chunk * N + x] = (rb[0][lx] + rb[1][lx]) + (rb[2][lx] + rb[3][lx]);
    }
}

// Workgroup reductions of NT threads (NT / 64 waves): within each wave by
// lane butterflies (every lane ends with the same value, since a + b = b + a
// and max/min are exact), then over the waves' values through `slots`
// (>= NT / 64 entries, shared).
template <int NT, class T, class Op>
__device__ __forceinline__ T block_reduce(T v, T *slots, Op op) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
    __syncthreads();  // earlier readers of slots are done
    if (lane == 0) slots[w] = v;
    __syncthreads();
    T t = slots[0];
#pragma unroll
    for (int i = 1; i < NT / 64; i++) t = op(t, slots[i]);
    return t;
}
struct OpSum {
    __device__ double operator()(double a, double b) const { return a + b; }
};
struct OpMax {
    __device__ double operator()(double a, double b) const { return fmax(a, b); }
    __device__ int operator()(int a, int b) const { return max(a, b); }
};
struct OpMin {
    __device__ int operator()(int a, int b) const { return min(a, b); }
};

// One workgroup of kIntThreads: window limits from the column energy w
// (Nx, in LDS; mask[j] = 1 on [left, right], 0-based): left = 1 + the last
// column before the centre below the cut-off, right = the first column from
// the centre on below it, minus 1 (parallel max/min reductions).
__device__ void window_body(const PgaImage &p, const double *w, double db_down) {
    __shared__ double smax[kIntThreads / 64];
    __shared__ int sl[kIntThreads / 64], sr[kIntThreads / 64];
    const int Nx = p.Nx, tid = threadIdx.x;
    double m = 0.0;
    for (int x = tid; x < Nx; x += kIntThreads) m = fmax(m, w[x]);
    const double wmax = block_reduce<kIntThreads>(m, smax, OpMax{});
    // dB relative to the maximum; max(dB) = 0, so the cut-off is -|db_down|
    const double cut = -fabs(db_down);
    const int center = (Nx + 1) / 2;  // 1-based ceil(Nx/2)
    int lastb = -1, firstb = Nx;      // last below in [0, center), first below in [center, Nx)
    for (int j = tid; j < Nx; j += kIntThreads) {
        if (10.0 * log10(w[j] / wmax) < cut) {
            if (j < center) lastb = max(lastb, j);
            else firstb = min(firstb, j);
        }
    }
    lastb = block_reduce<kIntThreads>(lastb, sl, OpMax{});
    firstb = block_reduce<kIntThreads>(firstb, sr, OpMin{});
    const int left = lastb + 1;  // 0 when none is below
    const int right = firstb < Nx ? firstb - 1 : Nx - 1;
    for (int j = tid; j < Nx; j += kIntThreads) p.mask[j] = (j >= left && j <= right) ? 1.f : 0.f;
}

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

constexpr int kPer = kPgaMaxN / kSpecThreads;  // x values per thread of the integrate block

// One workgroup of kSpecThreads: phi = unwrap([0 cumsum(phidot)]) from phid
// (Nx - 1 values, LDS), detrend, rms -> rms[it], af_ph += phi (af0: this
// thread's af_ph entries x = tid * per + j, loaded ahead); writes
// e^{-j phi} of the detrended phi for the correction to eph (Nx, LDS) as
// (cos phi, sin phi) in FP32.  unwrap of a cumulative sum is the cumulative
// sum of the unwrapped steps, so phi is one block-wide prefix scan of
// unwrap_step.
__device__ __forceinline__ bool integrate_body(const PgaImage &p, int it, const double *phid,
                                               const double (&af0)[kPer], float2 *eph) {
    __shared__ double sc[kSpecThreads / 64];
    __shared__ double red[kSpecThreads / 64];
    const int Nx = p.Nx, tid = threadIdx.x;
    const int per = (Nx + kSpecThreads - 1) / kSpecThreads;  // <= kPer
    const int b = tid * per;
    double loc[kPer];
    double run = 0.0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = b + j;
        if (j < per && i < Nx && i > 0) run += unwrap_step(phid[i - 1]);
        loc[j] = run;
    }
    // exclusive prefix sum of `run` over the threads: inclusive within the
    // wave by shuffles, shifted by one lane, plus the earlier waves' totals
    const int lane = tid & 63, wv = tid >> 6;
    double x = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sc[wv] = x;
    __syncthreads();
    double off = __shfl_up(x, 1);
    if (lane == 0) off = 0.0;
    for (int i = 0; i < wv; i++) off += sc[i];
    double sy = 0.0, sxy = 0.0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = b + j;
        loc[j] += off;
        if (j < per && i < Nx) {
            sy += loc[j];
            sxy += (double)i * loc[j];
        }
    }
    sy = block_reduce<kSpecThreads>(sy, red, OpSum{});
    sxy = block_reduce<kSpecThreads>(sxy, red, OpSum{});
    const double n = Nx;
    const double sx = 0.5 * n * (n - 1.0), sxx = (n - 1.0) * n * (2.0 * n - 1.0) / 6.0;
    const double den = n * sxx - sx * sx;
    const double slope = den != 0 ? (n * sxy - sx * sy) / den : 0.0;
    const double icpt = (sy - slope * sx) / n;
    double r = 0.0;
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = b + j;
        loc[j] -= slope * i + icpt;
        if (j < per && i < Nx) r += loc[j] * loc[j];
    }
    r = block_reduce<kSpecThreads>(r, red, OpSum{});
    const double rms = sqrt(r / n);
    if (tid == 0) p.rms[it] = rms;
    // autofocus_cpu.cpp:432: an estimate below the threshold ends the loop
    // before it is applied
    if (p.rms_stop > 0.0 && rms < p.rms_stop) {
        if (tid == 0) *p.stop = 1;
        return false;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int i = b + j;
        if (j < per && i < Nx) {
            const double v = loc[j];
            p.af_ph[i] = af0[j] + v;
            double sn, cs;
            sincos(v, &sn, &cs);
            eph[i] = make_float2((float)cs, (float)sn);
        }
    }
    return true;
}

// Short rows (M <= 512): one workgroup of kThreads per row, the row in LDS
// (dynamic, M float2), radix-2 transforms -- more lanes per row than the
// register-resident kernels give, so lower latency at these lengths.
// Step 3 for row y:
// G[y][k] = e^{-j 2 pi k (s - fl)/N} DFT_x( mask[(x + s) mod N] At[y][x] )_k,
// fl = floor(N/2): the shift by s, the mask and ifftshift as one remap.
__device__ void row_dft_lds_body(const PgaImage &p, int y) {
    if (stopped(p)) return;
    extern __shared__ float2 buf[];
    const int N = p.Nx, L = p.L, M = 1 << L;
    const int s = p.shift[y];
    const float2 *row = p.At + (size_t)y * N;
    for (int x = threadIdx.x; x < M; x += kThreads) {
        float2 v = make_float2(0.f, 0.f);
        if (x < N) {
            int mi = x + s;
            if (mi >= N) mi -= N;
            const float m = p.mask[mi];
            const float2 a = row[x];
            v = cmul(make_float2(m * a.x, m * a.y), p.wtab[x]);
        }
        buf[x] = v;
    }
    __syncthreads();
    bluestein_core(buf, L, p.twM, p.H);
    const float invM = 1.0f / (float)M;
    const int fl = N / 2;
    const int sh = ((s - fl) % N + N) % N;
    for (int k = threadIdx.x; k < N; k += kThreads) {
        float2 X = cmul(buf[k], p.wtab[k]);
        X.x *= invM;
        X.y *= invM;
        const float2 e = p.twN[((unsigned)k * (unsigned)sh) % (unsigned)N];  // e^{+j 2 pi k sh / N}; use its conjugate
        p.G[(size_t)y * N + k] = cmul(X, conjf2(e));
    }
}

// Step 6 for row y, in place: At[y] <- IDFT_N( DFT_N(At[y]) .* e^{-j phi} ),
// which is the circular convolution At[y] (*)_N h, h = IDFT_N(e^{-j phi}):
// the row zero-padded to M, times B = FFT_M(h extended over -(N-1)..N-1)
// (apply_spectrum_body), back, and the bins N-1 .. 2N-2 are the result.
__device__ void row_apply_lds_body(const PgaImage &p, int y) {
    if (stopped(p)) return;
    extern __shared__ float2 buf[];
    const int N = p.Nx, L = p.L, M = 1 << L;
    float2 *row = p.At + (size_t)y * N;
    for (int x = threadIdx.x; x < M; x += kThreads)
        buf[x] = x < N ? row[x] : make_float2(0.f, 0.f);
    __syncthreads();
    lds_fft(buf, L, p.twM, false);
    for (int k = threadIdx.x; k < M; k += kThreads) buf[k] = cmul(buf[k], p.B[k]);
    __syncthreads();
    lds_fft(buf, L, p.twM, true);
    const float invM = 1.0f / (float)M;
    for (int n = threadIdx.x; n < N; n += kThreads) {
        const float2 c = buf[n + N - 1];
        row[n] = make_float2(c.x * invM, c.y * invM);
    }
}

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

// LDS of the register-resident row kernels: P padded rows.  The kernels
// declare it once for both row steps, so a block holds one copy.
template <int LOG2M>
constexpr int row_lds_elems() {
    return fft::rows_per_group(LOG2M) * fft::padded(LOG2M);
}

// Step 3 for the P rows of block blk (M/16 lanes per row, M = 2^LOG2M), as
// row_dft_lds_body; lds: row_lds_elems<LOG2M>() elements.
template <int LOG2M>
__device__ __forceinline__ void row_dft_body(const PgaImage &p, int blk, float2 *lds) {
    if (stopped(p)) return;
    using S = fft::Shape<LOG2M>;
    constexpr int M = S::N, T = S::T, P = fft::rows_per_group(LOG2M), PAD = fft::padded(LOG2M);
    const int N = p.Nx;
    const int lp = threadIdx.x / T, j = threadIdx.x % T;
    const int y = blk * P + lp;
    const bool active = y < p.Ny;
    float2 *Lr = lds + lp * PAD;
    const int s = active ? p.shift[y] : 0;
    const float2 *row = p.At + (size_t)(active ? y : 0) * N;
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
                const float m = p.mask[mi];
                const float2 x = row[n];
                a = cmul(make_float2(m * x.x, m * x.y), p.wtab[n]);
            }
            v[b * S::R0 + r] = conjf2(a);
        }
    bluestein_regs<LOG2M>(v, Lr, j, p.twM, p.H);
    const float invM = 1.0f / (float)M;
    const int fl = N / 2;
    const int sh = ((s - fl) % N + N) % N;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int k = j + r * T;
        if (active && k < N) {
            float2 X = cmul(v[r], p.wtab[k]);
            X.x *= invM;
            X.y *= invM;
            const float2 e = p.twN[((unsigned)k * (unsigned)sh) % (unsigned)N];  // e^{+j 2 pi k sh / N}; use its conjugate
            p.G[(size_t)y * N + k] = cmul(X, conjf2(e));
        }
    }
}

// Step 6 for the P rows of block blk, in place, as row_apply_lds_body: the
// circular convolution with the correction kernel by two transforms of
// length M (forward of the zero-padded row, product with B, inverse).
template <int LOG2M>
__device__ __forceinline__ void row_apply_body(const PgaImage &p, int blk, float2 *lds) {
    if (stopped(p)) return;
    using S = fft::Shape<LOG2M>;
    constexpr int M = S::N, T = S::T, P = fft::rows_per_group(LOG2M), PAD = fft::padded(LOG2M);
    const int N = p.Nx;
    const int lp = threadIdx.x / T, j = threadIdx.x % T;
    const int y = blk * P + lp;
    const bool active = y < p.Ny;
    float2 *Lr = lds + lp * PAD;
    float2 *row = p.At + (size_t)(active ? y : 0) * N;
    float2 v[16];
#pragma unroll
    for (int b = 0; b < S::NB0; b++)
#pragma unroll
        for (int r = 0; r < S::R0; r++) {
            const int n = j + b * T + r * (M / S::R0);
            const float2 a = (active && n < N) ? row[n] : make_float2(0.f, 0.f);
            v[b * S::R0 + r] = conjf2(a);
        }
    fft::ifft_regs<LOG2M>(v, Lr, j, p.twM);  // v[r] = conj(A_k), k = j + r*T
    float2 u[16];
#pragma unroll
    for (int r = 0; r < 16; r++) u[r] = cmul(conjf2(v[r]), p.B[j + r * T]);
    fft::to_input_order<LOG2M>(u, v);
    fft::ifft_regs<LOG2M>(v, Lr, j, p.twM);  // v[r] = M (a * b)_{j + r*T}
    const float invM = 1.0f / (float)M;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int k = j + r * T - (N - 1);
        if (active && k >= 0 && k < N) row[k] = make_float2(v[r].x * invM, v[r].y * invM);
    }
}

// After integrate: the correction kernel's spectrum for step 6,
//   h = IDFT_N(e^{-j phi}) = conj(DFT_N(conj(e^{-j phi}))) / N   (Bluestein),
//   b_m = h[(m - (N-1)) mod N] for m <= 2N-2, 0 up to M,  B = FFT_M(b),
// so that IDFT_N(DFT_N(x) .* e^{-j phi}) = (x * b)[N-1 .. 2N-2] for a row x
// zero-padded to M >= 2N-1.  It is one vector per image, formed by the one
// integrate workgroup, so its three length-M transforms spread over all
// kSpecThreads threads: self-sorting Stockham passes of radix 4 (one of
// radix 2 first when log2 M is odd) through the LDS row, one to four
// butterflies per thread and pass, twiddles from the sine/cosine unit.

// d * e^{j 2 pi rev}, rev in revolutions.
__device__ __forceinline__ float2 rotate_rev(float2 d, float rev) {
    const float c = __builtin_amdgcn_cosf(rev), s = __builtin_amdgcn_sinf(rev);
    return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
}

// One Stockham pass of radix R (2 or 4) and sign SIGN over buf (M values at
// padded positions, fft::lpad) by NT threads, span Ns: butterfly b (k = b
// mod Ns) reads x_r = buf[b + r M/R], rotates x_r by e^{SIGN j 2 pi r k /
// (R Ns)}, takes the R-point DFT of sign SIGN and writes output q to
// buf[(b - k) R + k + q Ns].  Every read precedes every write (barrier).
template <int SIGN, int R, int M, int NT>
__device__ __forceinline__ void stockham_pass(float2 *buf, int Ns) {
    constexpr int NBF = M / R;
    constexpr int PER = NBF > NT ? NBF / NT : 1;
    static_assert(NBF <= NT || NBF % NT == 0, "butterflies per thread");
    const int tid = threadIdx.x;
    float2 x[PER][R];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = tid + q * NT;
        if (b < NBF) {
#pragma unroll
            for (int r = 0; r < R; r++) x[q][r] = buf[fft::lpad(b + r * NBF)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int b = tid + q * NT;
        if (b < NBF) {
            const int k = b & (Ns - 1);
            const float step = (float)(SIGN * k) / (float)(R * Ns);  // exact: powers of two
#pragma unroll
            for (int r = 1; r < R; r++) x[q][r] = rotate_rev(x[q][r], step * (float)r);
            float2 y[R];
            if constexpr (R == 2) {
                y[0] = cadd(x[q][0], x[q][1]);
                y[1] = csub(x[q][0], x[q][1]);
            } else {
                const float2 s02 = cadd(x[q][0], x[q][2]), d02 = csub(x[q][0], x[q][2]);
                const float2 s13 = cadd(x[q][1], x[q][3]), d13 = csub(x[q][1], x[q][3]);
                // (SIGN j) * d13
                const float2 jd = SIGN > 0 ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
                y[0] = cadd(s02, s13);
                y[1] = cadd(d02, jd);
                y[2] = csub(s02, s13);
                y[3] = csub(d02, jd);
            }
            const int o = (b - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) buf[fft::lpad(o + r * Ns)] = y[r];
        }
    }
    __syncthreads();
}

// Unnormalised length-2^L transform of sign SIGN (e^{SIGN j 2 pi k m / M})
// of buf, in place, by all NT threads of the workgroup; starts with a
// barrier (the caller may just have written buf) and ends with one.
template <int SIGN, int L, int NT>
__device__ __forceinline__ void lds_transform(float2 *buf) {
    constexpr int M = 1 << L;
    __syncthreads();
    int Ns = 1;
    if constexpr ((L & 1) != 0) {
        stockham_pass<SIGN, 2, M, NT>(buf, 1);
        Ns = 2;
    }
#pragma unroll
    for (int p = 0; p < L / 2; p++) {
        stockham_pass<SIGN, 4, M, NT>(buf, Ns);
        Ns *= 4;
    }
}

// The spectrum's tables for the thread's positions k = tid + q * kSpecThreads
// (q < M / kSpecThreads, at least one): w_k (k < N, else 0) and H_k, loaded
// ahead of the integration.
template <int L>
struct SpecTabs {
    static constexpr int Q = (1 << L) > kSpecThreads ? (1 << L) / kSpecThreads : 1;
    float2 w[Q], h[Q];
};
template <int L>
__device__ __forceinline__ void load_spec_tabs(const PgaImage &p, SpecTabs<L> &t) {
    constexpr int M = 1 << L;
#pragma unroll
    for (int q = 0; q < SpecTabs<L>::Q; q++) {
        const int k = threadIdx.x + q * kSpecThreads;
        t.w[q] = k < p.Nx ? p.wtab[k] : make_float2(0.f, 0.f);
        t.h[q] = k < M ? p.H[k] : make_float2(0.f, 0.f);
    }
}

// eph: conj(e^{-j phi}) = (cos phi, sin phi), N values in LDS (overwritten
// with h); buf: padded(L) float2 of LDS.  All kSpecThreads threads call it.
template <int L>
__device__ __forceinline__ void spectrum_lds(int N, float2 *eph, const SpecTabs<L> &t,
                                             float2 *__restrict__ B, float2 *buf) {
    constexpr int M = 1 << L, NT = kSpecThreads, Q = SpecTabs<L>::Q;
    const int tid = threadIdx.x;
    // a_k = conj(e_k) w_k for k < N, zero up to M
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int k = tid + q * NT;
        if (k < M) buf[fft::lpad(k)] = k < N ? cmul(eph[k], t.w[q]) : make_float2(0.f, 0.f);
    }
    lds_transform<-1, L, NT>(buf);  // A = FFT_M(a)
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int k = tid + q * NT;
        if (k < M) buf[fft::lpad(k)] = cmul(buf[fft::lpad(k)], t.h[q]);
    }
    lds_transform<+1, L, NT>(buf);  // c = M IFFT_M(A .* H)
    // h_i = conj(w_i c_i) / (M N), i < N, staged in eph (read above)
    const float scale = 1.0f / ((float)M * (float)N);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int i = tid + q * NT;
        if (i < N) {
            const float2 X = cmul(buf[fft::lpad(i)], t.w[q]);
            eph[i] = make_float2(X.x * scale, -X.y * scale);
        }
    }
    __syncthreads();
    // b_m = h[(m - (N-1)) mod N] for m <= 2N-2, zero up to M
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int m = tid + q * NT;
        if (m < M) {
            float2 bv = make_float2(0.f, 0.f);
            if (m <= 2 * N - 2) {
                int i = m - (N - 1);
                if (i < 0) i += N;
                bv = eph[i];
            }
            buf[fft::lpad(m)] = bv;
        }
    }
    lds_transform<-1, L, NT>(buf);  // B = FFT_M(b)
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int k = tid + q * NT;
        if (k < M) B[k] = buf[fft::lpad(k)];
    }
}

// Fused steps (fewer launches per iteration, same arithmetic):
//   stats:     blocks b < argmax_blocks(Ny): argmax of kArgRows rows; then the
//              blocks of the column-energy partial sums (col_partial_body, mode 0);
//   window:    the column energy from its partials, then the window
//              (one workgroup of kIntThreads);
//   integrate: phi_dot from the phase-difference partials, the integration
//              and detrend, and -- if the estimate is applied -- the correction
//              kernel's spectrum (one workgroup of kSpecThreads, the transforms
//              in registers through a padded LDS row).
__device__ void stats_body(const PgaImage &p, int b) {
    const int nr = argmax_blocks(p.Ny);
    if (b < nr) {
        argmax_body(p, b);
    } else {
        const int l = b - nr, gx = (p.Nx + 63) / 64;
        col_partial_body(p, 0, l % gx, l / gx);
    }
}

__device__ void window_fused_body(const PgaImage &p, double db_down) {
    if (stopped(p)) return;
    __shared__ double w[kPgaMaxN];
    const int N = p.Nx;
    for (int x = threadIdx.x; x < N; x += blockDim.x) {
        double a = 0.0;
        for (int c = 0; c < kRowChunks; c++) a += p.pa[(size_t)c * N + x];
        w[x] = a;
    }
    __syncthreads();
    window_body(p, w, db_down);
}

// The integrate step for transform length M = 2^L; every global read it
// needs but the partial sums (the spectrum's tables, af_ph) is issued first.
// Lr: padded(L) float2; phid: kPgaMaxN doubles; eph: kPgaMaxN float2 (LDS).
template <int L>
__device__ __forceinline__ void integrate_l(const PgaImage &p, int it, float2 *Lr, double *phid,
                                            float2 *eph) {
    const int N = p.Nx, tid = threadIdx.x;
    SpecTabs<L> t;
    load_spec_tabs<L>(p, t);
    const int per = (N + kSpecThreads - 1) / kSpecThreads, x0 = tid * per;
    double af0[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) af0[j] = (j < per && x0 + j < N) ? p.af_ph[x0 + j] : 0.0;
    for (int x = tid; x < N - 1; x += blockDim.x) {
        double a = 0.0, b = 0.0;
        for (int c = 0; c < kRowChunks; c++) {
            a += p.pa[(size_t)c * N + x];
            b += p.pb[(size_t)c * N + x];
        }
        phid[x] = b > 0.0 ? a / b : 0.0;
    }
    __syncthreads();
    if (!integrate_body(p, it, phid, af0, eph)) return;
    __syncthreads();
    spectrum_lds<L>(N, eph, t, p.B, Lr);
}

__device__ __forceinline__ void integrate_fused_body(const PgaImage &p, int it, float2 *Lr,
                                                     double *phid, float2 *eph) {
    if (stopped(p)) return;
    switch (p.L) {
        case 5: integrate_l<5>(p, it, Lr, phid, eph); break;
        case 6: integrate_l<6>(p, it, Lr, phid, eph); break;
        case 7: integrate_l<7>(p, it, Lr, phid, eph); break;
        case 8: integrate_l<8>(p, it, Lr, phid, eph); break;
        case 9: integrate_l<9>(p, it, Lr, phid, eph); break;
        case 10: integrate_l<10>(p, it, Lr, phid, eph); break;
        case 11: integrate_l<11>(p, it, Lr, phid, eph); break;
        case 12: integrate_l<12>(p, it, Lr, phid, eph); break;
        case 13: integrate_l<13>(p, it, Lr, phid, eph); break;
    }
}

// Blocks of a row step for an image of Ny rows: one workgroup per row up to
// L = 9, rows_per_group rows per workgroup from L = 10.
int row_blocks(int L, int Ny) {
    if (L <= 9) return Ny;
    const int P = fft::rows_per_group(L);
    return (Ny + P - 1) / P;
}

// ------------------------------------------------- kernels: one image

__global__ __launch_bounds__(kThreads) void k1_setup(PgaImage p) { init_body(p); }
// The forward transpose's first block also runs init_body.
__global__ __launch_bounds__(kThreads) void k1_transpose(PgaImage p, int forward) {
    if (forward && blockIdx.x == 0 && blockIdx.y == 0) init_body(p);
    if (forward) transpose_body(p.img, p.Nx, p.Ny, p.At, blockIdx.x, blockIdx.y);
    else transpose_body(p.At, p.Ny, p.Nx, p.img, blockIdx.x, blockIdx.y);
}
__global__ __launch_bounds__(kThreads) void k1_stats(PgaImage p) { stats_body(p, blockIdx.x); }
__global__ __launch_bounds__(kThreads) void k1_col_partial(PgaImage p, int mode) {
    col_partial_body(p, mode, blockIdx.x, blockIdx.y);
}
__global__ __launch_bounds__(kIntThreads) void k1_window(PgaImage p, double db_down) {
    window_fused_body(p, db_down);
}
__global__ __launch_bounds__(kSpecThreads) void k1_integrate(PgaImage p, int it) {
    __shared__ float2 Lr[fft::padded(kMaxL)];
    __shared__ double phid[kPgaMaxN];
    __shared__ float2 eph[kPgaMaxN];
    integrate_fused_body(p, it, Lr, phid, eph);
}
__global__ __launch_bounds__(kThreads) void k1_rows_lds(PgaImage p, int apply) {
    if (apply) row_apply_lds_body(p, blockIdx.x);
    else row_dft_lds_body(p, blockIdx.x);
}
template <int LOG2M>
__global__ __launch_bounds__(fft::group_threads(LOG2M)) void k1_rows(PgaImage p, int apply) {
    __shared__ float2 lds[row_lds_elems<LOG2M>()];
    if (apply) row_apply_body<LOG2M>(p, blockIdx.x, lds);
    else row_dft_body<LOG2M>(p, blockIdx.x, lds);
}

// ------------------------------------------------ kernels: many images

// The kernels below copy their image's PgaImage into registers before any
// other work: a field read after a barrier could otherwise become a vector
// load (the barrier's fence may clobber memory for all the compiler knows).

// Image i owns blocks [off[i], off[i+1]) of a launch over n images.
__device__ __forceinline__ int owner(const int *__restrict__ off, int n, int blk) {
    int lo = 0, hi = n;  // off[lo] <= blk < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Per-image block offsets, one array of n + 1 per step kind.
enum OffKind { kOffTranspose = 0, kOffRows, kOffColPartial, kOffKinds };

struct PgaBatchDev {
    const PgaImage *img;  // n images
    const int *off;       // kOffKinds x (n + 1)
    const int *grp;       // row-step groups: image ids, then block offsets
    int n;
};

__global__ __launch_bounds__(kThreads) void kb_setup(PgaBatchDev b) {
    const PgaImage p = b.img[blockIdx.x];
    init_body(p);
}
// The forward transpose's first block of every image also runs init_body.
__global__ __launch_bounds__(kThreads) void kb_transpose(PgaBatchDev b, int forward) {
    const int *off = b.off + kOffTranspose * (b.n + 1);
    const int i = owner(off, b.n, blockIdx.x);
    const PgaImage p = b.img[i];
    const int l = blockIdx.x - off[i];
    if (forward && l == 0) init_body(p);
    const int R = forward ? p.Nx : p.Ny, C = forward ? p.Ny : p.Nx;
    const int gx = (C + 63) / 64;
    if (forward) transpose_body(p.img, R, C, p.At, l % gx, l / gx);
    else transpose_body(p.At, R, C, p.img, l % gx, l / gx);
}
// Blocks [0, rows): the argmax blocks of every image; then the column-energy
// partial blocks of every image.
__global__ __launch_bounds__(kThreads) void kb_stats(PgaBatchDev b, int rows) {
    if ((int)blockIdx.x < rows) {
        const int *off = b.off + kOffRows * (b.n + 1);
        const int i = owner(off, b.n, blockIdx.x);
        const PgaImage p = b.img[i];
        argmax_body(p, blockIdx.x - off[i]);
    } else {
        const int *off = b.off + kOffColPartial * (b.n + 1);
        const int bb = blockIdx.x - rows;
        const int i = owner(off, b.n, bb);
        const PgaImage p = b.img[i];
        const int l = bb - off[i], gx = (p.Nx + 63) / 64;
        col_partial_body(p, 0, l % gx, l / gx);
    }
}
__global__ __launch_bounds__(kThreads) void kb_col_partial(PgaBatchDev b, int mode) {
    const int *off = b.off + kOffColPartial * (b.n + 1);
    const int i = owner(off, b.n, blockIdx.x);
    const PgaImage p = b.img[i];
    const int l = blockIdx.x - off[i], gx = (p.Nx + 63) / 64;
    col_partial_body(p, mode, l % gx, l / gx);
}
__global__ __launch_bounds__(kIntThreads) void kb_window(PgaBatchDev b, double db_down) {
    const PgaImage p = b.img[blockIdx.x];
    window_fused_body(p, db_down);
}
__global__ __launch_bounds__(kSpecThreads) void kb_integrate(PgaBatchDev b, int it) {
    __shared__ float2 Lr[fft::padded(kMaxL)];
    __shared__ double phid[kPgaMaxN];
    __shared__ float2 eph[kPgaMaxN];
    const PgaImage p = b.img[blockIdx.x];
    integrate_fused_body(p, it, Lr, phid, eph);
}
// Row step of one transform length: grp[g0 .. g0 + m) are the images' ids,
// grp[g0 + m .. g0 + 2m] their block offsets.
__global__ __launch_bounds__(kThreads) void kb_rows_lds(PgaBatchDev b, int g0, int m, int apply) {
    const int *ids = b.grp + g0, *off = ids + m;
    const int k = owner(off, m, blockIdx.x);
    const PgaImage p = b.img[ids[k]];
    if (apply) row_apply_lds_body(p, blockIdx.x - off[k]);
    else row_dft_lds_body(p, blockIdx.x - off[k]);
}
// M <= 4096: capped at 128 VGPRs for four waves per SIMD (the LDS allows four
// blocks per CU); the few spills cost less than the lost occupancy (64 uneven
// images, 5 iterations: 6.10-6.16 -> 5.98-6.01 ms).  The single-image kernel
// spills more under the same cap and is left uncapped.
template <int LOG2M>
__global__ __launch_bounds__(fft::group_threads(LOG2M))
__attribute__((amdgpu_waves_per_eu(LOG2M <= 12 ? 4 : 1))) void kb_rows(PgaBatchDev b, int g0,
                                                                        int m, int apply) {
    const int *ids = b.grp + g0, *off = ids + m;
    const int k = owner(off, m, blockIdx.x);
    const PgaImage p = b.img[ids[k]];
    __shared__ float2 lds[row_lds_elems<LOG2M>()];
    if (apply) row_apply_body<LOG2M>(p, blockIdx.x - off[k], lds);
    else row_dft_body<LOG2M>(p, blockIdx.x - off[k], lds);
}

// ---------------------------------------------------------------- host

int transform_log2(int N) {
    int L = kMinL;  // the row kernels take 32 <= M <= 8192
    while ((1 << L) < 2 * N - 1) L++;
    return L;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

// Bytes of one image's small state, laid out by carve().
size_t small_bytes(int Nx, int Ny, int L) {
    return align256(sizeof(int) * Ny) + 2 * align256(sizeof(double) * kRowChunks * Nx) +
           align256(sizeof(float) * Nx) + align256(sizeof(int)) +
           align256(sizeof(float2) * ((size_t)1 << L));
}

void carve(PgaImage &p, char *base) {
    auto take = [&](size_t bytes) {
        char *r = base;
        base += align256(bytes);
        return r;
    };
    p.shift = (int *)take(sizeof(int) * p.Ny);
    p.pa = (double *)take(sizeof(double) * kRowChunks * p.Nx);
    p.pb = (double *)take(sizeof(double) * kRowChunks * p.Nx);
    p.mask = (float *)take(sizeof(float) * p.Nx);
    p.stop = (int *)take(sizeof(int));
    p.B = (float2 *)take(sizeof(float2) * ((size_t)1 << p.L));
}

// Row tables of length N (M = 2^L), cached in the context: [w (N) | H (M)],
// w_n = e^{-j pi n^2 / N} and H = FFT_M(g), g_m = conj(w_m) for m < N,
// g_{M-m} = conj(w_m) for 0 < m < N.  With c and C of bluestein_f64,
// w = conj(c) and g = conj(h), so H_k = conj(C_{(M-k) mod M}); FP64, rounded
// once.
int row_tables(sarf_ctx *ctx, int N, int L, const float2 **w, const float2 **H) {
    const int M = 1 << L;
    const long key = kPgaTableKey + N;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->rc_blue.find(key);
    if (it == ctx->rc_blue.end()) {
        std::vector<double> cr, ci, hr, hi;
        bluestein_f64(N, L, cr, ci, hr, hi);
        std::vector<float2> t((size_t)N + M);
        for (int n = 0; n < N; n++) t[n] = make_float2((float)cr[n], (float)-ci[n]);
        for (int k = 0; k < M; k++) {
            const int q = (M - k) & (M - 1);
            t[(size_t)N + k] = make_float2((float)hr[q], (float)-hi[q]);
        }
        void *dv = nullptr;
        hipError_t e = hipMalloc(&dv, sizeof(float2) * t.size());
        if (e != hipSuccess)
            return fail(SARF_ERR_NOMEM, "pga: row tables: %s", hipGetErrorString(e));
        e = hipMemcpy(dv, t.data(), sizeof(float2) * t.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(dv);
            return fail(SARF_ERR_HIP, "pga: row tables: %s", hipGetErrorString(e));
        }
        it = ctx->rc_blue.emplace(key, dv).first;
    }
    *w = (const float2 *)it->second;
    *H = (const float2 *)it->second + N;
    return SARF_OK;
}

int tables(sarf_ctx *ctx, PgaImage &p) {
    SARF_TRY(twiddles_any_f32(ctx, 1L << p.L, &p.twM));
    SARF_TRY(twiddles_any_f32(ctx, p.Nx, &p.twN));
    SARF_TRY(row_tables(ctx, p.Nx, p.L, &p.wtab, &p.H));
    return SARF_OK;
}

template <int LOG2M>
void launch_k1_rows(const PgaImage &p, int apply, hipStream_t s) {
    hipLaunchKernelGGL(k1_rows<LOG2M>, dim3(row_blocks(LOG2M, p.Ny)),
                       dim3(fft::group_threads(LOG2M)), 0, s, p, apply);
}

int k1_row_step(const PgaImage &p, int apply, hipStream_t s) {
    if (p.L <= 9) {
        hipLaunchKernelGGL(k1_rows_lds, dim3(p.Ny), dim3(kThreads),
                           sizeof(float2) * ((size_t)1 << p.L), s, p, apply);
        return SARF_OK;
    }
    switch (p.L) {
        case 10: launch_k1_rows<10>(p, apply, s); break;
        case 11: launch_k1_rows<11>(p, apply, s); break;
        case 12: launch_k1_rows<12>(p, apply, s); break;
        case 13: launch_k1_rows<13>(p, apply, s); break;
        default: return fail(SARF_ERR_UNSUPPORTED, "sarf_pga: transform length 2^%d", p.L);
    }
    return SARF_OK;
}

template <int LOG2M>
void launch_kb_rows(const PgaBatchDev &b, int g0, int m, int blocks, int apply, hipStream_t s) {
    hipLaunchKernelGGL(kb_rows<LOG2M>, dim3(blocks), dim3(fft::group_threads(LOG2M)), 0, s, b,
                       g0, m, apply);
}

int kb_row_step(const PgaBatchDev &b, int L, int g0, int m, int blocks, int apply, hipStream_t s) {
    if (L <= 9) {
        hipLaunchKernelGGL(kb_rows_lds, dim3(blocks), dim3(kThreads),
                           sizeof(float2) * ((size_t)1 << L), s, b, g0, m, apply);
        return SARF_OK;
    }
    switch (L) {
        case 10: launch_kb_rows<10>(b, g0, m, blocks, apply, s); break;
        case 11: launch_kb_rows<11>(b, g0, m, blocks, apply, s); break;
        case 12: launch_kb_rows<12>(b, g0, m, blocks, apply, s); break;
        case 13: launch_kb_rows<13>(b, g0, m, blocks, apply, s); break;
        default: return fail(SARF_ERR_UNSUPPORTED, "sarf_pga_scenes: transform length 2^%d", L);
    }
    return SARF_OK;
}

int check_image(const char *who, int Nx, int Ny) {
    if (Nx < 2 || Ny < 1) return fail(SARF_ERR_INVALID, "%s: Nx=%d Ny=%d", who, Nx, Ny);
    if (Nx > kPgaMaxN) return fail(SARF_ERR_UNSUPPORTED, "%s: Nx=%d > %d", who, Nx, kPgaMaxN);
    return SARF_OK;
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_pga(sarf_ctx *ctx, const sarf_pga_desc *d, void *img, double *af_ph,
                        double *rms, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!d || !img || !af_ph || !rms) return fail(SARF_ERR_INVALID, "sarf_pga: null argument");
    SARF_TRY(check_image("sarf_pga", d->Nx, d->Ny));
    if (d->iters < 0) return fail(SARF_ERR_INVALID, "sarf_pga: iters=%d", d->iters);
    if (!(d->rms_stop >= 0.0))
        return fail(SARF_ERR_INVALID, "sarf_pga: rms_stop=%g", d->rms_stop);
    hipStream_t s = (hipStream_t)stream;
    PgaImage p{};
    p.Nx = d->Nx;
    p.Ny = d->Ny;
    p.L = transform_log2(p.Nx);
    p.img = (float2 *)img;
    p.af_ph = af_ph;
    p.rms = rms;
    p.iters = d->iters;
    p.rms_stop = d->rms_stop;
    const size_t npix = (size_t)p.Nx * p.Ny;
    void *bufA, *bufG, *bufS;
    SARF_TRY(scratch_get(ctx, "pga.At", sizeof(float2) * npix, &bufA));
    SARF_TRY(scratch_get(ctx, "pga.G", sizeof(float2) * npix, &bufG));
    SARF_TRY(scratch_get(ctx, "pga.small", small_bytes(p.Nx, p.Ny, p.L), &bufS));
    p.At = (float2 *)bufA;
    p.G = (float2 *)bufG;
    carve(p, (char *)bufS);
    SARF_TRY(tables(ctx, p));
    if (d->iters == 0) hipLaunchKernelGGL(k1_setup, dim3(1), dim3(kThreads), 0, s, p);
    if (d->iters > 0) {
        // img is Nx rows (x) of Ny values; At is Ny rows of Nx values
        hipLaunchKernelGGL(k1_transpose, dim3((p.Ny + 63) / 64, (p.Nx + 63) / 64), dim3(kThreads),
                           0, s, p, 1);
        const dim3 gcp((p.Nx + 63) / 64, kRowChunks);
        for (int it = 0; it < d->iters; it++) {
            hipLaunchKernelGGL(k1_stats, dim3(argmax_blocks(p.Ny) + gcp.x * gcp.y), dim3(kThreads),
                               0, s, p);
            hipLaunchKernelGGL(k1_window, dim3(1), dim3(kIntThreads), 0, s, p, d->db_down);
            SARF_TRY(k1_row_step(p, 0, s));
            hipLaunchKernelGGL(k1_col_partial, gcp, dim3(kThreads), 0, s, p, 1);
            hipLaunchKernelGGL(k1_integrate, dim3(1), dim3(kSpecThreads), 0, s, p, it);
            SARF_TRY(k1_row_step(p, 1, s));
            SARF_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k1_transpose, dim3((p.Nx + 63) / 64, (p.Ny + 63) / 64), dim3(kThreads),
                           0, s, p, 0);
    }
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

extern "C" int sarf_pga_scenes(sarf_ctx *ctx, int nscenes, const sarf_pga_scene *scenes, int iters,
                               double db_down, double rms_stop, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (nscenes < 1 || !scenes || iters < 0 || !(rms_stop >= 0.0))
        return fail(SARF_ERR_INVALID, "sarf_pga_scenes: nscenes=%d iters=%d rms_stop=%g", nscenes,
                    iters, rms_stop);
    const int n = nscenes;
    std::vector<PgaImage> host(n);  // value-initialised: padding is zero (the cache key)
    size_t tot_pix = 0, tot_small = 0;
    for (int i = 0; i < n; i++) {
        const sarf_pga_scene &c = scenes[i];
        if (!c.img || !c.af_ph || !c.rms)
            return fail(SARF_ERR_INVALID, "sarf_pga_scenes: scene %d has a null pointer", i);
        SARF_TRY(check_image("sarf_pga_scenes", c.Nx, c.Ny));
        PgaImage &p = host[i];
        p.Nx = c.Nx;
        p.Ny = c.Ny;
        p.L = transform_log2(c.Nx);
        p.img = (float2 *)c.img;
        p.af_ph = c.af_ph;
        p.rms = c.rms;
        p.iters = iters;
        p.rms_stop = rms_stop;
        tot_pix += align256(sizeof(float2) * (size_t)c.Nx * c.Ny);
        tot_small += small_bytes(p.Nx, p.Ny, p.L);
        SARF_TRY(tables(ctx, p));
    }
    // block offsets per step kind, and the row-step groups by transform length
    std::vector<int> off((size_t)kOffKinds * (n + 1), 0);
    for (int i = 0; i < n; i++) {
        const PgaImage &p = host[i];
        const int cnt[kOffKinds] = {((p.Nx + 63) / 64) * ((p.Ny + 63) / 64),
                                    argmax_blocks(p.Ny), ((p.Nx + 63) / 64) * kRowChunks};
        for (int k = 0; k < kOffKinds; k++)
            off[(size_t)k * (n + 1) + i + 1] = off[(size_t)k * (n + 1) + i] + cnt[k];
    }
    struct Group {
        int L, g0, m, blocks;
    };
    std::vector<Group> groups;
    std::vector<int> grp;
    for (int L = kMinL; L <= kMaxL; L++) {
        std::vector<int> ids;
        for (int i = 0; i < n; i++)
            if (host[i].L == L) ids.push_back(i);
        if (ids.empty()) continue;
        Group g{L, (int)grp.size(), (int)ids.size(), 0};
        grp.insert(grp.end(), ids.begin(), ids.end());
        grp.push_back(0);
        for (int id : ids) {
            g.blocks += row_blocks(L, host[id].Ny);
            grp.push_back(g.blocks);
        }
        groups.push_back(g);
    }
    const size_t b_img = sizeof(PgaImage) * (size_t)n, b_off = sizeof(int) * off.size(),
                 b_grp = sizeof(int) * grp.size();
    const size_t b_desc = align256(b_img) + align256(b_off) + align256(b_grp);
    hipStream_t s = (hipStream_t)stream;
    // Device state of this image list -- descriptors, transposes, spectra and
    // per-image small state -- in an allocation of its own, keyed by the
    // caller's list and the parameters stored in the state: a list seen
    // before launches without copying or waiting, and a graph captured with
    // it keeps pointers no other list's call writes.
    std::vector<unsigned char> key(sizeof(sarf_pga_scene) * (size_t)n + sizeof(int) +
                                   sizeof(double));
    std::memcpy(key.data(), scenes, sizeof(sarf_pga_scene) * (size_t)n);
    std::memcpy(key.data() + sizeof(sarf_pga_scene) * (size_t)n, &iters, sizeof(int));
    std::memcpy(key.data() + sizeof(sarf_pga_scene) * (size_t)n + sizeof(int), &rms_stop,
                sizeof(double));
    void *buf = nullptr;
    bool fresh = false;
    SARF_TRY(list_buffer(ctx, 'p', key, b_desc + 2 * tot_pix + tot_small, s, "sarf_pga_scenes",
                         &buf, &fresh));
    char *bufA = (char *)buf + b_desc, *bufG = bufA + tot_pix, *bufS = bufG + tot_pix;
    {
        size_t op = 0, os = 0;
        for (int i = 0; i < n; i++) {
            PgaImage &p = host[i];
            p.At = (float2 *)(bufA + op);
            p.G = (float2 *)(bufG + op);
            op += align256(sizeof(float2) * (size_t)p.Nx * p.Ny);
            carve(p, bufS + os);
            os += small_bytes(p.Nx, p.Ny, p.L);
        }
    }
    PgaBatchDev b{};
    b.img = (const PgaImage *)buf;
    b.off = (const int *)((char *)buf + align256(b_img));
    b.grp = (const int *)((char *)buf + align256(b_img) + align256(b_off));
    b.n = n;
    if (fresh) {
        // pageable sources: the copies complete before returning
        SARF_HIP(hipMemcpyAsync((void *)b.img, host.data(), b_img, hipMemcpyHostToDevice, s));
        SARF_HIP(hipMemcpyAsync((void *)b.off, off.data(), b_off, hipMemcpyHostToDevice, s));
        SARF_HIP(hipMemcpyAsync((void *)b.grp, grp.data(), b_grp, hipMemcpyHostToDevice, s));
        SARF_HIP(hipStreamSynchronize(s));
    }
    const auto total = [&](int kind) { return off[(size_t)kind * (n + 1) + n]; };
    if (iters == 0) hipLaunchKernelGGL(kb_setup, dim3(n), dim3(kThreads), 0, s, b);
    if (iters > 0) {
        hipLaunchKernelGGL(kb_transpose, dim3(total(kOffTranspose)), dim3(kThreads), 0, s, b, 1);
        for (int it = 0; it < iters; it++) {
            hipLaunchKernelGGL(kb_stats, dim3(total(kOffRows) + total(kOffColPartial)),
                               dim3(kThreads), 0, s, b, total(kOffRows));
            hipLaunchKernelGGL(kb_window, dim3(n), dim3(kIntThreads), 0, s, b, db_down);
            for (const Group &g : groups) SARF_TRY(kb_row_step(b, g.L, g.g0, g.m, g.blocks, 0, s));
            hipLaunchKernelGGL(kb_col_partial, dim3(total(kOffColPartial)), dim3(kThreads), 0, s,
                               b, 1);
            hipLaunchKernelGGL(kb_integrate, dim3(n), dim3(kSpecThreads), 0, s, b, it);
            for (const Group &g : groups) SARF_TRY(kb_row_step(b, g.L, g.g0, g.m, g.blocks, 1, s));
            SARF_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(kb_transpose, dim3(total(kOffTranspose)), dim3(kThreads), 0, s, b, 0);
    }
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

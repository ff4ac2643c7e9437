// sarf_rc.hip -- range compression: prof[n] = fftshift(ifft(ph[n], Nfft)) / ...
//
// Semantics (reference): per pulse, zero-pad the K frequency samples to Nfft
// (matlab/bpBasic.m:82), inverse DFT scaled by 1/Nfft
// (src/cpuBackProjection/cpuBackProjection_fft.hpp:46-48), then
// fftshift = cshift((N+1)/2) (:81-84); the GPU reference does the same with
// cuFFT + cufftNormalize_1DBatch + cufftShift_1DBatch
// (cuda_BackProjectionKernels.cuh:95-135).  Optional taper and phase ramp
// follow matlab/rangeCompress.m:6-17.
//
// Design (gfx950): one workgroup holds whole pulses in LDS and runs a
// self-sorting Stockham IFFT, radix 16 per pass (16 complex values per
// lane), so each pulse is read from HBM once (K samples, the zero padding is
// never loaded) and written once (Nfft bins).  Fused into the passes:
//   * fftshift, as a (-1)^k sign on the input samples (N even):
//       X[(m + N/2) mod N] = sum_k x_k (-1)^k e^{+j 2 pi k m / N};
//   * taper, 1/N and the ramp, on the first load / last store.
// LDS rows are padded by one element every 16 so the stride-16 scatter of
// the first pass is bank-conflict free.  N up to 16384 fits (139 KiB of the
// 160 KiB LDS).  The FP64 variant (validation builds) and FP32 powers of two
// outside 32..16384 use a global-memory radix-4 Stockham; lengths that are
// not a power of two use Bluestein's identity (register-resident up to 8192,
// through the global-memory passes beyond).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "sarf_fft.hpp"

namespace sarf {
namespace {

using fft::dft;
using fft::lpad;
using fft::pass_twiddles;

// Longest non-power-of-two transform done by the register-resident
// Bluestein kernel (M <= 16384); longer ones (and FP64) go through the
// global-memory passes, up to 2^20 (M = 2^21).
constexpr int kBluesteinMaxN = 8192;
constexpr int kBluesteinGlobalMaxN = 1 << 20;

constexpr int rc_threads(int log2n) {
    return ((1 << log2n) / 16 >= 256) ? (1 << log2n) / 16 : 256;
}

// One range-compression job: Np pulses of K samples into Nfft-bin profiles,
// pulse groups numbered from g0 within its launch.
struct RcJob {
    const float2 *ph;
    float2 *prof;
    const float *taper;
    const double *f0;
    const double *r0;
    int K, Np, flags;
    int g0;
};

// Job source of a one-job launch (the job is a kernel argument).
struct OneJob {
    RcJob j;
    __device__ __forceinline__ int advance(int ji, int) const { return ji; }
    __device__ __forceinline__ const RcJob &get(int) const { return j; }
};

// Job source of a batched launch: jobs in device memory in increasing g0.
// A workgroup's group index only grows, so its job index only advances.
struct JobList {
    const RcJob *__restrict__ jobs;
    int n;
    __device__ __forceinline__ int advance(int ji, int g) const {
        while (ji + 1 < n && jobs[ji + 1].g0 <= g) ji++;
        return ji;
    }
    __device__ __forceinline__ RcJob get(int ji) const { return jobs[ji]; }
};

// ROWS: how many of a lane's 16 sample rows k = j + i*T (i < 16) can hold
// samples, i.e. ceil(K/T) rounded up to 4 or 16.  With ROWS = 4 only four
// values per lane are loaded and the next pulse group's are prefetched into
// registers; with ROWS = 16 (K > N/4) the group's samples are loaded at the
// top of its iteration.
template <int LOG2N, int ROWS, class Jobs>
__device__ __forceinline__ void rc_lds_body(const Jobs &J, int ngroups,
                                            const float2 *__restrict__ tw, float2 *lds) {
    constexpr int N = 1 << LOG2N;
    constexpr int T = N / 16;                  // lanes per pulse
    constexpr int P = T >= 256 ? 1 : 256 / T;  // pulses per workgroup
    constexpr int B = LOG2N % 4;
    constexpr int R0 = B == 0 ? 16 : (1 << B);  // radix of the first pass
    constexpr int NB0 = 16 / R0;                // first-pass butterflies per lane
    constexpr int PADN = N + N / 16;

    const int tid = threadIdx.x;
    const int lp = tid / T, j = tid % T;
    float2 *L = lds + lp * PADN;

    // Sample row i of pulse n (sample k = j + i*T) as pass 0 consumes it:
    // zero padding (k >= K is never loaded), taper, fftshift fold (-1)^k.
    auto sample = [&](const RcJob &job, int n, int i, int jl) {
        const int k = jl + i * T;
        float2 v = make_float2(0.f, 0.f);
        if (n < job.Np && k < job.K) {
            v = job.ph[(size_t)n * job.K + k];
            if (job.flags & SARF_RC_TAPER) {
                const float w = job.taper[k];
                v.x *= w;
                v.y *= w;
            }
            if (k & 1) v = make_float2(-v.x, -v.y);
        }
        return v;
    };
    // Raw prefetch of group g's rows (no use of the values here, so the
    // loads stay in flight): the address is clamped to sample 0 for rows past
    // K or pulses past Np, and pass 0 masks those rows to zero.
    auto fetch = [&](const RcJob &job, int g, float2 *x, float *w, int jl) {
        const int n = (g - job.g0) * P + lp;
#pragma unroll
        for (int i = 0; i < ROWS; i++) {
            const int k = jl + i * T;
            const bool ok = n < job.Np && k < job.K;
            x[i] = job.ph[ok ? (size_t)n * job.K + k : 0];
            w[i] = (job.flags & SARF_RC_TAPER) ? job.taper[ok ? k : 0] : 1.f;
        }
    };

    // Persistent over pulse groups (the grid is at most one resident wave of
    // workgroups).  With ROWS = 4 the next group's samples are loaded into
    // registers while this group's LDS passes run, and this group's stores
    // drain while the next group computes, so HBM traffic overlaps the FFT.
    constexpr bool kPrefetch = ROWS <= 4;
    float2 xin[kPrefetch ? ROWS : 1];
    float win[kPrefetch ? ROWS : 1];
    int ji = J.advance(0, blockIdx.x);
    if (kPrefetch) fetch(J.get(ji), blockIdx.x, xin, win, j);
    for (int g = blockIdx.x; g < ngroups; g += gridDim.x) {
        const RcJob job = J.get(ji);
        const int n = (g - job.g0) * P + lp;
        const bool active = n < job.Np;
        // Opaque copy of the lane index: keeps LLVM from hoisting every
        // pass's LDS addresses and twiddle indices out of the pulse loop,
        // which would hold them in registers across iterations.
        int jj = j;
        __asm__ volatile("" : "+v"(jj));
        float2 v[16];
        // ---- pass 0: registers -> LDS, radix R0, Ns = 1 (no twiddles) ------
        // lane input r of butterfly b is sample row i = b + r * NB0
#pragma unroll
        for (int b = 0; b < NB0; b++) {
            const int jb = jj + b * T;
#pragma unroll
            for (int r = 0; r < R0; r++) {
                const int i = b + r * NB0;
                if (kPrefetch) {
                    float2 x = make_float2(0.f, 0.f);
                    const int k = jj + i * T;
                    if (i < ROWS && active && k < job.K) {
                        const int ic = i < ROWS ? i : 0;
                        x = xin[ic];
                        x.x *= win[ic];
                        x.y *= win[ic];
                        if (k & 1) x = make_float2(-x.x, -x.y);
                    }
                    v[b * R0 + r] = x;
                } else {
                    v[b * R0 + r] = sample(job, n, i, jj);
                }
            }
            dft<+1, R0>(v + b * R0);
#pragma unroll
            for (int r = 0; r < R0; r++) L[lpad(jb * R0 + r)] = v[b * R0 + r];
        }
        const int gn = g + (int)gridDim.x;
        if (gn < ngroups) {
            ji = J.advance(ji, gn);
            if (kPrefetch) fetch(J.get(ji), gn, xin, win, jj);
        }
        __syncthreads();

        // ---- radix-16 passes ----------------------------------------------
        constexpr int NPASS16 = (LOG2N - (B == 0 ? 4 : B)) / 4;
        int Ns = R0;
#pragma unroll
        for (int pass = 0; pass < NPASS16; pass++) {
            const bool last = (pass == NPASS16 - 1);
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = L[lpad(jj + r * (N / 16))];
            const int k = jj & (Ns - 1);
            if (k != 0) {
                float2 p[16];
                pass_twiddles<LOG2N>(tw, k, Ns, p);
#pragma unroll
                for (int r = 1; r < 16; r++) v[r] = cmul(v[r], p[r]);
            }
            dft<+1, 16>(v);
            if (!last) {
                __syncthreads();  // every lane has read before anyone overwrites
                const int base = (jj - k) * 16 + k;
#pragma unroll
                for (int r = 0; r < 16; r++) L[lpad(base + r * Ns)] = v[r];
                __syncthreads();
            } else if (active) {
                // Ns = N/16: output index j + r*N/16, coalesced per r
                float sr = 1.0f / N, si = 0.0f;
                if (job.flags & SARF_RC_RAMP) {
                    double a = -4.0 * kPi * job.f0[n] * job.r0[n] / kClight;
                    sr = (float)(cos(a) / N);
                    si = (float)(sin(a) / N);
                }
                float2 *dst = job.prof + (size_t)n * N;
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    const float2 x = v[r];
                    dst[jj + r * (N / 16)] = make_float2(x.x * sr - x.y * si, x.x * si + x.y * sr);
                }
            }
            Ns *= 16;
        }
        __syncthreads();  // the last pass has read LDS before the next pass 0 writes
    }
}

template <int LOG2N>
constexpr int rc_lds_elems() {
    constexpr int N = 1 << LOG2N;
    constexpr int T = N / 16;
    return (T >= 256 ? 1 : 256 / T) * (N + N / 16);
}

// One set of pulses (the job is a kernel argument).
template <int LOG2N, int ROWS>
__global__ __launch_bounds__(rc_threads(LOG2N)) void rc_lds_f32(RcJob job, int ngroups,
                                                               const float2 *__restrict__ tw) {
    __shared__ float2 lds[rc_lds_elems<LOG2N>()];
    rc_lds_body<LOG2N, ROWS>(OneJob{job}, ngroups, tw, lds);
}

// Many sets of pulses of one transform length (jobs in device memory).
template <int LOG2N, int ROWS>
__global__ __launch_bounds__(rc_threads(LOG2N)) void rc_lds_f32_jobs(
    const RcJob *__restrict__ jobs, int njobs, int ngroups, const float2 *__restrict__ tw) {
    __shared__ float2 lds[rc_lds_elems<LOG2N>()];
    rc_lds_body<LOG2N, ROWS>(JobList{jobs, njobs}, ngroups, tw, lds);
}

// ---------------------------------------------------------------------------
// Global-memory Stockham, radix 4 (one radix-2 pass if log2 N is odd): the
// FP64 path, and FP32 transforms longer than the 16384 bins that fit in
// LDS.  Pass 0 reads the phase history (zero padding, taper, fftshift fold);
// the last pass applies 1/N and the ramp.
// ---------------------------------------------------------------------------
template <typename T2, int R>
__global__ __launch_bounds__(256) void rc_pass(
    const T2 *__restrict__ in, T2 *__restrict__ out, int N, int Ns, int Np,
    const T2 *__restrict__ tw, int first, int K, int flags,
    const float *__restrict__ taper, int last, const double *__restrict__ f0,
    const double *__restrict__ r0) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = N / R;
    const long n = gid / per;
    const int j = (int)(gid % per);
    if (n >= Np) return;
    using Re = decltype(T2{}.x);
    T2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int idx = j + r * (N / R);
        if (first) {
            T2 x{};
            if (idx < K) {
                x = in[n * K + idx];
                if (flags & SARF_RC_TAPER) {
                    x.x *= taper[idx];
                    x.y *= taper[idx];
                }
                if (idx & 1) {
                    x.x = -x.x;
                    x.y = -x.y;
                }
            }
            v[r] = x;
        } else {
            v[r] = in[n * N + idx];
        }
    }
    const int k = j % Ns;
#pragma unroll
    for (int r = 1; r < R; r++) {
        const long ti = (long)r * k * (N / (Ns * R));
        v[r] = cmul(v[r], tw[ti]);
    }
    // inverse DFT of size R (R = 2 or 4)
    if constexpr (R == 2) {
        T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else {
        T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        T2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
        T2 jb1;  // +j * b1
        jb1.x = -b1.y;
        jb1.y = b1.x;
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, jb1);
        v[3] = csub(a1, jb1);
    }
    double sr = 1.0, si = 0.0;
    if (last) {
        sr = 1.0 / N;
        if (flags & SARF_RC_RAMP) {
            const double a = -4.0 * kPi * f0[n] * r0[n] / kClight;
            sr = cos(a) / N;
            si = sin(a) / N;
        }
    }
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) {
        T2 x = v[r];
        if (last) {
            const Re xr = x.x, xi = x.y;
            x.x = xr * (Re)sr - xi * (Re)si;
            x.y = xr * (Re)si + xi * (Re)sr;
        }
        out[n * N + base + r * Ns] = x;
    }
}

// ---------------------------------------------------------------------------
// Nfft not a power of two beyond the register-resident range (FP32 above
// 8192, every FP64 length): the same Bluestein identity as rc_bluestein below,
// with the two length-M transforms done by the global-memory passes.  Per
// pulse, in rows of M:
//   rcb_pre : z_k = conj(x_k taper_k c_k) for k < K, 0 up to M
//   passes  : Z = IFFT_M(z) unscaled = conj(FFT_M(x c))
//   rcb_mul : U = conj(Z) .* C                         (C = FFT_M(h), h the
//             circularly extended conj(c); FFT_M(x c) .* C = FFT_M((x c) * h))
//   passes  : W = IFFT_M(U) unscaled = M ((x c) * h)
//   rcb_post: X_m = c_m W_m / (M N) (ramp), stored at the fftshift position.
// ---------------------------------------------------------------------------
template <typename T2>
__global__ __launch_bounds__(256) void rcb_pre(const T2 *__restrict__ ph, int K, int M, long total,
                                               int flags, const float *__restrict__ taper,
                                               const T2 *__restrict__ chirp, T2 *__restrict__ z) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const long n = gid / M;
    const int k = (int)(gid - n * M);
    T2 v{};
    if (k < K) {
        T2 x = ph[n * K + k];
        if (flags & SARF_RC_TAPER) {
            x.x *= taper[k];
            x.y *= taper[k];
        }
        v = cmul(x, chirp[k]);
        v.y = -v.y;
    }
    z[gid] = v;
}

template <typename T2>
__global__ __launch_bounds__(256) void rcb_mul(T2 *__restrict__ z, int M, long total,
                                               const T2 *__restrict__ C) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    T2 v = z[gid];
    v.y = -v.y;
    z[gid] = cmul(v, C[gid % M]);
}

template <typename T2>
__global__ __launch_bounds__(256) void rcb_post(const T2 *__restrict__ w, int M, int N, long total,
                                                int flags, const double *__restrict__ f0,
                                                const double *__restrict__ r0,
                                                const T2 *__restrict__ chirp, T2 *__restrict__ prof) {
    using Re = decltype(T2{}.x);
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const long n = gid / N;
    const int m = (int)(gid - n * N);
    double sr = 1.0 / ((double)M * N), si = 0.0;
    if (flags & SARF_RC_RAMP) {
        const double a = -4.0 * kPi * f0[n] * r0[n] / kClight;
        sr = cos(a) / ((double)M * N);
        si = sin(a) / ((double)M * N);
    }
    const T2 X = cmul(w[n * M + m], chirp[m]);
    const int half = (N + 1) / 2;
    int o = m - half;
    if (o < 0) o += N;
    T2 y;
    y.x = X.x * (Re)sr - X.y * (Re)si;
    y.y = X.x * (Re)si + X.y * (Re)sr;
    prof[n * N + o] = y;
}

// ---------------------------------------------------------------------------
// Nfft not a power of two, up to 8192 (e.g. the reference's default
// N_fft = K, charlotte_sar_api.hpp:173): Bluestein's identity
// k m = (k^2 + m^2 - (m - k)^2) / 2 turns the length-N inverse DFT into a
// circular convolution of length M = 2^ceil(log2(2N - 1)):
//   X_m = c_m sum_k (x_k c_k) conj(c_{m-k}),   c_k = e^{+j pi k^2 / N},
// formed per pulse by two register-resident radix-16 transforms of length M
// (sarf_fft.hpp): the forward transform of a_k = x_k c_k as conj(ifft(conj)),
// the product with C = FFT_M(conj(c)) (a per-N table, circularly extended),
// and the inverse transform.  c_k comes from the exact k^2 mod 2N in FP64.
// Zero padding (k >= K) is never loaded; fftshift = cshift((N+1)/2) is an
// index remap on the store (odd N included), with 1/N and the ramp.  One row
// of M/16 lanes per pulse, rows packed to >= 256 lanes per workgroup.
// ---------------------------------------------------------------------------
template <int LOG2M>
__global__ __launch_bounds__(fft::group_threads(LOG2M)) void rc_bluestein(
    const float2 *__restrict__ ph, int K, int N, int Np, int flags,
    const float *__restrict__ taper, const double *__restrict__ f0,
    const double *__restrict__ r0, const float2 *__restrict__ chirp,
    const float2 *__restrict__ C, const float2 *__restrict__ twM, float2 *__restrict__ prof) {
    using S = fft::Shape<LOG2M>;
    constexpr int M = S::N, T = S::T, P = fft::rows_per_group(LOG2M);
    constexpr int PAD = fft::padded(LOG2M);
    __shared__ float2 lds[P * PAD];
    const int lp = threadIdx.x / T, j = threadIdx.x % T;
    const int n = blockIdx.x * P + lp;
    const bool active = n < Np;
    float2 *Lr = lds + lp * PAD;
    const float2 *x = ph + (size_t)(active ? n : 0) * K;
    float2 v[16];
#pragma unroll
    for (int b = 0; b < S::NB0; b++)
#pragma unroll
        for (int r = 0; r < S::R0; r++) {
            const int k = j + b * T + r * (M / S::R0);
            float2 a = make_float2(0.f, 0.f);
            if (active && k < K) {
                float2 xv = x[k];
                if (flags & SARF_RC_TAPER) {
                    xv.x *= taper[k];
                    xv.y *= taper[k];
                }
                a = cmul(xv, chirp[k]);
            }
            v[b * S::R0 + r] = make_float2(a.x, -a.y);
        }
    fft::ifft_regs<LOG2M>(v, Lr, j, twM);  // v[r] = conj(A_k), k = j + r*T
    float2 u[16];
#pragma unroll
    for (int r = 0; r < 16; r++) u[r] = cmul(make_float2(v[r].x, -v[r].y), C[j + r * T]);
    fft::to_input_order<LOG2M>(u, v);
    fft::ifft_regs<LOG2M>(v, Lr, j, twM);  // v[r] = M (a * conj c)_{j + r*T}
    if (!active) return;
    double sr = 1.0 / ((double)M * N), si = 0.0;
    if (flags & SARF_RC_RAMP) {
        const double ang = -4.0 * kPi * f0[n] * r0[n] / kClight;
        sr = cos(ang) / ((double)M * N);
        si = sin(ang) / ((double)M * N);
    }
    const float fr = (float)sr, fi = (float)si;
    const int half = (N + 1) / 2;
    float2 *out = prof + (size_t)n * N;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = j + r * T;
        if (m < N) {
            const float2 X = cmul(v[r], chirp[m]);
            int o = m - half;
            if (o < 0) o += N;
            out[o] = make_float2(X.x * fr - X.y * fi, X.x * fi + X.y * fr);
        }
    }
}

// Workgroups per CU of a kernel (queried once per kernel).
template <class K>
int rc_occupancy(K kernel, int threads, std::atomic<int> &cache) {
    int occ = cache.load(std::memory_order_relaxed);
    if (occ == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess ||
            occ < 1)
            occ = 1;
        cache.store(occ, std::memory_order_relaxed);
    }
    return occ;
}

template <int LOG2N>
constexpr int rc_pulses_per_group() {
    return (1 << LOG2N) / 16 >= 256 ? 1 : 256 / ((1 << LOG2N) / 16);
}

template <int LOG2N, int ROWS>
int launch_lds_rows(const sarf_ctx *ctx, const sarf_rc_desc *d, const float2 *ph, const float *taper,
               const double *f0, const double *r0, const float2 *tw, float2 *prof,
               hipStream_t s) {
    constexpr int P = rc_pulses_per_group<LOG2N>();
    const int ngroups = (d->Np + P - 1) / P;
    // one resident wave of workgroups, each looping over pulse groups
    static std::atomic<int> per_cu{0};
    const int occ = rc_occupancy(rc_lds_f32<LOG2N, ROWS>, rc_threads(LOG2N), per_cu);
    const int cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    const int grid = std::min(ngroups, occ * cus);
    const RcJob job{ph, prof, taper, f0, r0, d->K, d->Np, d->flags, 0};
    hipLaunchKernelGGL((rc_lds_f32<LOG2N, ROWS>), dim3(grid), dim3(rc_threads(LOG2N)), 0, s, job,
                       ngroups, tw);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

template <int LOG2N>
int launch_lds(const sarf_ctx *ctx, const sarf_rc_desc *d, const float2 *ph, const float *taper,
               const double *f0, const double *r0, const float2 *tw, float2 *prof,
               hipStream_t s) {
    constexpr int T = (1 << LOG2N) / 16;
    if (d->K <= 4 * T) return launch_lds_rows<LOG2N, 4>(ctx, d, ph, taper, f0, r0, tw, prof, s);
    return launch_lds_rows<LOG2N, 16>(ctx, d, ph, taper, f0, r0, tw, prof, s);
}

// Sample rows a lane holds for K samples at transform length 2^log2n.
int rc_rows(int log2n, int K) { return K <= 4 * ((1 << log2n) / 16) ? 4 : 16; }

template <int LOG2N, int ROWS>
int launch_jobs_rows(const sarf_ctx *ctx, const RcJob *jobs, int njobs, int ngroups,
                     const float2 *tw, hipStream_t s) {
    static std::atomic<int> per_cu{0};
    const int occ = rc_occupancy(rc_lds_f32_jobs<LOG2N, ROWS>, rc_threads(LOG2N), per_cu);
    const int cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    const int grid = std::min(ngroups, occ * cus);
    hipLaunchKernelGGL((rc_lds_f32_jobs<LOG2N, ROWS>), dim3(grid), dim3(rc_threads(LOG2N)), 0, s,
                       jobs, njobs, ngroups, tw);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

template <int LOG2N>
int launch_jobs(const sarf_ctx *ctx, int rows, const RcJob *jobs, int njobs, int ngroups,
                const float2 *tw, hipStream_t s) {
    if (rows == 4) return launch_jobs_rows<LOG2N, 4>(ctx, jobs, njobs, ngroups, tw, s);
    return launch_jobs_rows<LOG2N, 16>(ctx, jobs, njobs, ngroups, tw, s);
}

}  // namespace

// Bluestein chirp and kernel spectrum of a length-N DFT with transform
// length M = 2^L, in FP64: c_k = e^{+j pi k^2 / N} (k < N; the phase from the
// exact k^2 mod 2N) and C = FFT_M(h) with h_d = conj(c_d) for d < N,
// h_{M-d} = conj(c_d) for 0 < d < N, zero between (iterative radix 2 over
// directly evaluated twiddles).
void bluestein_f64(int N, int L, std::vector<double> &cr, std::vector<double> &ci,
                   std::vector<double> &hr, std::vector<double> &hi) {
    const int M = 1 << L;
    cr.assign(N, 0.0);
    ci.assign(N, 0.0);
    for (int k = 0; k < N; k++) {
        const long q = ((long)k * k) % (2L * N);
        const double a = kPi * (double)q / (double)N;
        cr[k] = std::cos(a);
        ci[k] = std::sin(a);
    }
    hr.assign(M, 0.0);
    hi.assign(M, 0.0);
    for (int d = 0; d < N; d++) {
        hr[d] = cr[d];
        hi[d] = -ci[d];
        if (d > 0) {
            hr[M - d] = cr[d];
            hi[M - d] = -ci[d];
        }
    }
    // forward FFT (e^{-j 2 pi t / M}) over a table of the M/2 twiddles
    std::vector<double> wr(M / 2 > 0 ? M / 2 : 1), wi(M / 2 > 0 ? M / 2 : 1);
    for (int t = 0; t < M / 2; t++) {
        const double a = -2.0 * kPi * (double)t / (double)M;
        wr[t] = std::cos(a);
        wi[t] = std::sin(a);
    }
    for (int i = 1, r = 0; i < M; i++) {
        int bit = M >> 1;
        for (; r & bit; bit >>= 1) r ^= bit;
        r |= bit;
        if (i < r) {
            std::swap(hr[i], hr[r]);
            std::swap(hi[i], hi[r]);
        }
    }
    for (int len = 2; len <= M; len <<= 1) {
        const int half = len / 2, step = M / len;
        for (int st = 0; st < M; st += len)
            for (int q = 0; q < half; q++) {
                const double xr = wr[q * step], xi = wi[q * step];
                const int i0 = st + q, i1 = i0 + half;
                const double tr = hr[i1] * xr - hi[i1] * xi, ti = hr[i1] * xi + hi[i1] * xr;
                hr[i1] = hr[i0] - tr;
                hi[i1] = hi[i0] - ti;
                hr[i0] += tr;
                hi[i0] += ti;
            }
    }
}

namespace {

// Bluestein tables of a length-N inverse DFT with transform length M = 2^L
// (cached per N and precision in the context): chirp (N) and C (M) of
// bluestein_f64, rounded once to T2.  One allocation [chirp (N) | C (M)].
template <typename T2>
int bluestein_tables(sarf_ctx *ctx, int N, int L, const T2 **chirp, const T2 **C) {
    const int M = 1 << L;
    const long key = sizeof(T2) == sizeof(double2) ? -(long)N : (long)N;
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->rc_blue.find(key);
    if (it == ctx->rc_blue.end()) {
        std::vector<double> cr, ci, hr, hi;
        bluestein_f64(N, L, cr, ci, hr, hi);
        using Re = decltype(T2{}.x);
        std::vector<T2> h((size_t)N + M);
        for (int k = 0; k < N; k++) {
            h[k].x = (Re)cr[k];
            h[k].y = (Re)ci[k];
        }
        for (int t = 0; t < M; t++) {
            h[(size_t)N + t].x = (Re)hr[t];
            h[(size_t)N + t].y = (Re)hi[t];
        }
        void *dv = nullptr;
        hipError_t e = hipMalloc(&dv, sizeof(T2) * h.size());
        if (e != hipSuccess)
            return fail(SARF_ERR_NOMEM, "range_compress: Bluestein tables: %s", hipGetErrorString(e));
        e = hipMemcpy(dv, h.data(), sizeof(T2) * h.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(dv);
            return fail(SARF_ERR_HIP, "range_compress: Bluestein tables: %s", hipGetErrorString(e));
        }
        it = ctx->rc_blue.emplace(key, dv).first;
    }
    *chirp = (const T2 *)it->second;
    *C = (const T2 *)it->second + N;
    return SARF_OK;
}

template <int LOG2M>
void launch_bluestein_t(const sarf_rc_desc *d, const float2 *ph, const float *taper,
                        const double *f0, const double *r0, const float2 *chirp, const float2 *C,
                        const float2 *tw, float2 *prof, hipStream_t s) {
    constexpr int P = fft::rows_per_group(LOG2M);
    hipLaunchKernelGGL(rc_bluestein<LOG2M>, dim3((d->Np + P - 1) / P),
                       dim3(fft::group_threads(LOG2M)), 0, s, ph, d->K, d->Nfft, d->Np, d->flags,
                       taper, f0, r0, chirp, C, tw, prof);
}

// Stockham passes of length N = 2^log2n through global memory (ping-pong
// with `tmp` so the last pass lands in prof).
template <typename T2>
int launch_passes(const sarf_rc_desc *d, int log2n, const T2 *ph, const float *taper,
                  const double *f0, const double *r0, const T2 *tw, T2 *prof, T2 *tmp,
                  hipStream_t s) {
    const int npass = (log2n & 1) + log2n / 2;
    T2 *bufs[2] = {prof, tmp};
    int cur = (npass % 2 == 1) ? 0 : 1;  // output buffer index of pass 0
    const T2 *in = ph;
    int Ns = 1, pass = 0;
    if (log2n & 1) {
        const long threads = (long)d->Np * (d->Nfft / 2);
        hipLaunchKernelGGL((rc_pass<T2, 2>), dim3((threads + 255) / 256), dim3(256), 0, s, in,
                           bufs[cur], d->Nfft, Ns, d->Np, tw, 1, d->K, d->flags, taper,
                           npass == 1, f0, r0);
        SARF_HIP(hipGetLastError());
        in = bufs[cur];
        cur ^= 1;
        Ns *= 2;
        pass++;
    }
    for (; pass < npass; pass++) {
        const long threads = (long)d->Np * (d->Nfft / 4);
        hipLaunchKernelGGL((rc_pass<T2, 4>), dim3((threads + 255) / 256), dim3(256), 0, s, in,
                           bufs[cur], d->Nfft, Ns, d->Np, tw, pass == 0, d->K, d->flags, taper,
                           pass == npass - 1, f0, r0);
        SARF_HIP(hipGetLastError());
        in = bufs[cur];
        cur ^= 1;
        Ns *= 4;
    }
    return SARF_OK;
}

// Unscaled inverse transforms of length M = 2^L of Np rows held in x,
// ping-ponging with y; returns the buffer holding the result.
template <typename T2>
T2 *inverse_passes(T2 *x, T2 *y, int L, int Np, const T2 *tw, hipStream_t s) {
    const int M = 1 << L;
    T2 *in = x, *out = y;
    int Ns = 1;
    if (L & 1) {
        const long threads = (long)Np * (M / 2);
        hipLaunchKernelGGL((rc_pass<T2, 2>), dim3((threads + 255) / 256), dim3(256), 0, s, in, out,
                           M, Ns, Np, tw, 0, M, 0, nullptr, 0, nullptr, nullptr);
        std::swap(in, out);
        Ns *= 2;
    }
    for (; Ns < M; Ns *= 4) {
        const long threads = (long)Np * (M / 4);
        hipLaunchKernelGGL((rc_pass<T2, 4>), dim3((threads + 255) / 256), dim3(256), 0, s, in, out,
                           M, Ns, Np, tw, 0, M, 0, nullptr, 0, nullptr, nullptr);
        std::swap(in, out);
    }
    return in;
}

// Bluestein through global memory (rcb_pre / passes / rcb_mul / passes /
// rcb_post), two scratch rows of M per pulse.
template <typename T2>
int launch_bluestein_global(sarf_ctx *ctx, const sarf_rc_desc *d, int L, const T2 *ph,
                            const float *taper, const double *f0, const double *r0,
                            const T2 *tw, T2 *prof, hipStream_t s) {
    const int N = d->Nfft, M = 1 << L;
    const T2 *chirp = nullptr, *C = nullptr;
    SARF_TRY(bluestein_tables<T2>(ctx, N, L, &chirp, &C));
    const long rows = (long)d->Np * M;
    void *a = nullptr, *b = nullptr;
    SARF_TRY(scratch_get(ctx, "rcb.a", sizeof(T2) * (size_t)rows, &a));
    SARF_TRY(scratch_get(ctx, "rcb.b", sizeof(T2) * (size_t)rows, &b));
    T2 *za = (T2 *)a, *zb = (T2 *)b;
    const dim3 gm((unsigned)((rows + 255) / 256));
    hipLaunchKernelGGL(rcb_pre<T2>, gm, dim3(256), 0, s, ph, d->K, M, rows, d->flags, taper,
                       chirp, za);
    T2 *z = inverse_passes<T2>(za, zb, L, d->Np, tw, s);
    hipLaunchKernelGGL(rcb_mul<T2>, gm, dim3(256), 0, s, z, M, rows, C);
    T2 *w = inverse_passes<T2>(z, z == za ? zb : za, L, d->Np, tw, s);
    const long outs = (long)d->Np * N;
    hipLaunchKernelGGL(rcb_post<T2>, dim3((unsigned)((outs + 255) / 256)), dim3(256), 0, s, w, M, N,
                       outs, d->flags, f0, r0, chirp, prof);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

int rc_pulses_per_group(int log2n) {
    const int T = (1 << log2n) / 16;
    return T >= 256 ? 1 : 256 / T;
}

}  // namespace

int launch_range_compress(sarf_ctx *ctx, const sarf_rc_desc *d, const void *ph,
                          const float *taper, const double *f0, const double *r0,
                          void *prof, hipStream_t s) {
    const int log2n = ilog2_exact(d->Nfft);
    if (d->K < 1 || d->Np < 1 || d->K > d->Nfft)
        return fail(SARF_ERR_INVALID, "range_compress: need 1 <= K <= Nfft, Np >= 1 "
                    "(K=%d Np=%d Nfft=%d)", d->K, d->Np, d->Nfft);
    if ((d->flags & SARF_RC_TAPER) && !taper)
        return fail(SARF_ERR_INVALID, "range_compress: SARF_RC_TAPER without taper");
    if ((d->flags & SARF_RC_RAMP) && (!f0 || !r0))
        return fail(SARF_ERR_INVALID, "range_compress: SARF_RC_RAMP without f0/r0");
    const bool f64 = (d->flags & SARF_RC_F64) != 0;
    if (log2n < 0 && !f64 && d->Nfft <= kBluesteinMaxN) {
        // not a power of two: Bluestein through two register-resident
        // transforms of length M = 2^ceil(log2(2N - 1))
        int L = 5;
        while ((1 << L) < 2 * d->Nfft - 1) L++;
        const float2 *chirp = nullptr, *C = nullptr, *tw = nullptr;
        SARF_TRY(bluestein_tables<float2>(ctx, d->Nfft, L, &chirp, &C));
        SARF_TRY(twiddles_f32(ctx, L, &tw));
        const float2 *x = (const float2 *)ph;
        float2 *y = (float2 *)prof;
        switch (L) {
            case 5: launch_bluestein_t<5>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 6: launch_bluestein_t<6>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 7: launch_bluestein_t<7>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 8: launch_bluestein_t<8>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 9: launch_bluestein_t<9>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 10: launch_bluestein_t<10>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 11: launch_bluestein_t<11>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 12: launch_bluestein_t<12>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 13: launch_bluestein_t<13>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            case 14: launch_bluestein_t<14>(d, x, taper, f0, r0, chirp, C, tw, y, s); break;
            default: return fail(SARF_ERR_UNSUPPORTED, "range_compress: Nfft=%d", d->Nfft);
        }
        SARF_HIP(hipGetLastError());
        return SARF_OK;
    }
    if (log2n < 0) {
        // not a power of two, beyond the register-resident range (or FP64):
        // Bluestein through the global-memory passes
        if (d->Nfft > kBluesteinGlobalMaxN)
            return fail(SARF_ERR_UNSUPPORTED, "range_compress: Nfft=%d (lengths that are not a "
                        "power of two go up to %d)", d->Nfft, kBluesteinGlobalMaxN);
        int L = 1;
        while ((1 << L) < 2 * d->Nfft - 1) L++;
        if (f64) {
            const double2 *tw = nullptr;
            SARF_TRY(twiddles_f64(ctx, L, &tw));
            return launch_bluestein_global<double2>(ctx, d, L, (const double2 *)ph, taper, f0, r0,
                                                    tw, (double2 *)prof, s);
        }
        const float2 *tw = nullptr;
        SARF_TRY(twiddles_f32(ctx, L, &tw));
        return launch_bluestein_global<float2>(ctx, d, L, (const float2 *)ph, taper, f0, r0, tw,
                                               (float2 *)prof, s);
    }
    if (f64 || log2n > 14 || log2n < 5) {
        if (log2n < 1 || log2n > 24)
            return fail(SARF_ERR_UNSUPPORTED, "range_compress: Nfft=%d", d->Nfft);
        void *tmp = nullptr;
        if (f64) {
            const double2 *tw = nullptr;
            SARF_TRY(twiddles_f64(ctx, log2n, &tw));
            SARF_TRY(scratch_get(ctx, "rc64.tmp", sizeof(double2) * (size_t)d->Nfft * d->Np, &tmp));
            return launch_passes<double2>(d, log2n, (const double2 *)ph, taper, f0, r0, tw,
                                          (double2 *)prof, (double2 *)tmp, s);
        }
        const float2 *tw = nullptr;
        SARF_TRY(twiddles_f32(ctx, log2n, &tw));
        SARF_TRY(scratch_get(ctx, "rc32.tmp", sizeof(float2) * (size_t)d->Nfft * d->Np, &tmp));
        return launch_passes<float2>(d, log2n, (const float2 *)ph, taper, f0, r0, tw,
                                     (float2 *)prof, (float2 *)tmp, s);
    }
    const float2 *tw = nullptr;
    SARF_TRY(twiddles_f32(ctx, log2n, &tw));
    const float2 *x = (const float2 *)ph;
    float2 *y = (float2 *)prof;
    switch (log2n) {
        case 5: return launch_lds<5>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 6: return launch_lds<6>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 7: return launch_lds<7>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 8: return launch_lds<8>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 9: return launch_lds<9>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 10: return launch_lds<10>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 11: return launch_lds<11>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 12: return launch_lds<12>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 13: return launch_lds<13>(ctx, d, x, taper, f0, r0, tw, y, s);
        case 14: return launch_lds<14>(ctx, d, x, taper, f0, r0, tw, y, s);
    }
    return fail(SARF_ERR_UNSUPPORTED, "range_compress: Nfft=%d", d->Nfft);
}

}  // namespace sarf

using namespace sarf;

// Many scenes: the scenes of one transform length (and one sample-row count)
// form one persistent launch whose workgroups stride over the pulse groups
// of all of them, so small scenes share the GPU instead of leaving it idle
// between launches.
extern "C" int sarf_range_compress_scenes(sarf_ctx *ctx, int nscenes,
                                          const sarf_rc_scene *scenes, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (nscenes < 1 || !scenes)
        return fail(SARF_ERR_INVALID, "sarf_range_compress_scenes: no scenes");
    // class key = log2n * 2 + (16 sample rows ? 1 : 0) -> scene indices
    std::map<int, std::vector<int>> classes;
    for (int i = 0; i < nscenes; i++) {
        const sarf_rc_scene &c = scenes[i];
        const int log2n = ilog2_exact(c.Nfft);
        if (!c.ph || !c.prof)
            return fail(SARF_ERR_INVALID, "sarf_range_compress_scenes: scene %d has a null "
                        "pointer", i);
        if (c.K < 1 || c.Np < 1 || c.K > c.Nfft)
            return fail(SARF_ERR_INVALID, "sarf_range_compress_scenes: scene %d: need 1 <= K <= "
                        "Nfft, Np >= 1 (K=%d Np=%d Nfft=%d)", i, c.K, c.Np, c.Nfft);
        if (log2n < 5 || log2n > 14)
            return fail(SARF_ERR_UNSUPPORTED, "sarf_range_compress_scenes: scene %d: Nfft=%d is "
                        "not one of the powers of two 32..16384", i, c.Nfft);
        if (c.flags & ~(SARF_RC_TAPER | SARF_RC_RAMP))
            return fail(SARF_ERR_INVALID, "sarf_range_compress_scenes: scene %d: flags %d (FP32 "
                        "only: SARF_RC_TAPER | SARF_RC_RAMP)", i, c.flags);
        if ((c.flags & SARF_RC_TAPER) && !c.taper)
            return fail(SARF_ERR_INVALID, "sarf_range_compress_scenes: scene %d: SARF_RC_TAPER "
                        "without taper", i);
        if ((c.flags & SARF_RC_RAMP) && (!c.f0 || !c.r0))
            return fail(SARF_ERR_INVALID, "sarf_range_compress_scenes: scene %d: SARF_RC_RAMP "
                        "without f0/r0", i);
        classes[log2n * 2 + (rc_rows(log2n, c.K) == 16)].push_back(i);
    }
    // jobs grouped by class, each class numbering its pulse groups from 0
    struct Launch {
        int log2n, rows, first, count, ngroups;
    };
    std::vector<Launch> launches;
    std::vector<RcJob> jobs;
    jobs.reserve(nscenes);
    for (const auto &kv : classes) {
        const int log2n = kv.first / 2;
        const int P = rc_pulses_per_group(log2n);
        Launch l{log2n, (kv.first & 1) ? 16 : 4, (int)jobs.size(), (int)kv.second.size(), 0};
        for (int i : kv.second) {
            const sarf_rc_scene &c = scenes[i];
            jobs.push_back(RcJob{(const float2 *)c.ph, (float2 *)c.prof, c.taper, c.f0, c.r0, c.K,
                                 c.Np, c.flags, l.ngroups});
            l.ngroups += (c.Np + P - 1) / P;
        }
        launches.push_back(l);
    }
    const float2 *tws[15] = {};
    for (const Launch &l : launches) SARF_TRY(twiddles_f32(ctx, l.log2n, &tws[l.log2n]));

    const size_t bytes = sizeof(RcJob) * jobs.size();
    hipStream_t s = (hipStream_t)stream;
    // The job list depends only on the caller's descriptors: a list seen
    // before has its own device copy, so the call launches without copying
    // or waiting (and can be captured into a graph once it has run outside).
    std::vector<unsigned char> key(sizeof(sarf_rc_scene) * (size_t)nscenes);
    std::memcpy(key.data(), scenes, key.size());
    void *buf = nullptr;
    bool fresh = false;
    SARF_TRY(list_buffer(ctx, 'r', key, bytes, s, "sarf_range_compress_scenes", &buf, &fresh));
    RcJob *djobs = (RcJob *)buf;
    if (fresh) {
        // pageable source: wait for the copy before the host vector goes away
        SARF_HIP(hipMemcpyAsync(djobs, jobs.data(), bytes, hipMemcpyHostToDevice, s));
        SARF_HIP(hipStreamSynchronize(s));
    }
    for (const Launch &l : launches) {
        const RcJob *jl = djobs + l.first;
        const float2 *tw = tws[l.log2n];
        int rc = SARF_ERR_UNSUPPORTED;
        switch (l.log2n) {
            case 5: rc = launch_jobs<5>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 6: rc = launch_jobs<6>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 7: rc = launch_jobs<7>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 8: rc = launch_jobs<8>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 9: rc = launch_jobs<9>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 10: rc = launch_jobs<10>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 11: rc = launch_jobs<11>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 12: rc = launch_jobs<12>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 13: rc = launch_jobs<13>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
            case 14: rc = launch_jobs<14>(ctx, l.rows, jl, l.count, l.ngroups, tw, s); break;
        }
        if (rc != SARF_OK) return rc;
    }
    return SARF_OK;
}

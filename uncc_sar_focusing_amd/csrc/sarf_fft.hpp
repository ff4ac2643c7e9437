// sarf_fft.hpp -- register/LDS building blocks of the power-of-two FFTs
// (range compression, PGA row transforms): radix-2 DIF networks on up to 16
// registers with exact trivial rotations, and a self-sorting Stockham
// transform of length N = 2^LOG2N by N/16 lanes (16 values per lane, radix-16
// passes through a padded LDS row).
#pragma once

#include "sarf_internal.hpp"

namespace sarf {
namespace fft {

// cos(2 pi m / 16) for m = 0..15; sin(x) is cos(x - pi/2).
template <int M>
struct Tw16 {
    static constexpr float c() {
        constexpr float t[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f,
                                 0.38268343236508978f, 0.0f, -0.38268343236508978f,
                                 -0.70710678118654752f, -0.92387953251128674f, -1.0f,
                                 -0.92387953251128674f, -0.70710678118654752f,
                                 -0.38268343236508978f, 0.0f, 0.38268343236508978f,
                                 0.70710678118654752f, 0.92387953251128674f};
        return t[M & 15];
    }
    static constexpr float s() { return Tw16<(M + 12) & 15>::c(); }  // sin x = cos(x - pi/2)
};

// d * e^{SIGN j 2 pi M / 16}, with the trivial rotations done exactly.
template <int SIGN, int M>
__device__ __forceinline__ float2 rot16(float2 d) {
    constexpr int m = M & 15;
    if constexpr (m == 0) {
        return d;
    } else if constexpr (m == 8) {
        return make_float2(-d.x, -d.y);
    } else if constexpr (m == 4) {  // e^{SIGN j pi/2} = SIGN j
        return SIGN > 0 ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    } else if constexpr (m == 12) {
        return SIGN > 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
    } else {
        constexpr float c = Tw16<m>::c();
        constexpr float s = SIGN * Tw16<m>::s();
        return make_float2(d.x * c - d.y * s, d.x * s + d.y * c);
    }
}

// Radix-2 decimation-in-frequency network on R registers (R = 2,4,8,16),
// natural-order input, bit-reversed output.
template <int SIGN, int L, int S, int K>
struct DifBfly {
    static __device__ __forceinline__ void run(float2 *v) {
        if constexpr (K < L / 2) {
            float2 a = v[S + K], b = v[S + K + L / 2];
            v[S + K] = cadd(a, b);
            v[S + K + L / 2] = rot16<SIGN, (K * 16 / L)>(csub(a, b));
            DifBfly<SIGN, L, S, K + 1>::run(v);
        }
    }
};
template <int SIGN, int R, int L, int S>
struct DifStage {
    static __device__ __forceinline__ void run(float2 *v) {
        if constexpr (S < R) {
            DifBfly<SIGN, L, S, 0>::run(v);
            DifStage<SIGN, R, L, S + L>::run(v);
        }
    }
};
template <int SIGN, int R, int L>
struct DifAll {
    static __device__ __forceinline__ void run(float2 *v) {
        if constexpr (L >= 2) {
            DifStage<SIGN, R, L, 0>::run(v);
            DifAll<SIGN, R, L / 2>::run(v);
        }
    }
};
template <int R>
__host__ __device__ constexpr int bitrev(int x) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (x & 1);
        x >>= 1;
    }
    return r;
}

// In-place R-point DFT (sign SIGN) of v[0..R-1], result in natural order.
template <int SIGN, int R>
__device__ __forceinline__ void dft(float2 *v) {
    DifAll<SIGN, R, R>::run(v);
    float2 t[R];
#pragma unroll
    for (int i = 0; i < R; i++) t[i] = v[bitrev<R>(i)];
#pragma unroll
    for (int i = 0; i < R; i++) v[i] = t[i];
}

__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }

// e^{j 2 pi r rev} for r = 1..15 from the sine/cosine unit, which takes
// revolutions (rev = k / (Ns 16) is exact, so each twiddle is one rounding
// of the unit, about 1e-7).
__device__ __forceinline__ void powers16_rev(float rev, float2 *p) {
#pragma unroll
    for (int r = 1; r < 16; r++) {
        const float a = rev * (float)r;
        p[r] = make_float2(__builtin_amdgcn_cosf(a), __builtin_amdgcn_sinf(a));
    }
}

// w^r for r = 0..15 from w (r = 0 unused); error <= ~4 roundings.
__device__ __forceinline__ void powers16(float2 w, float2 *p) {
    p[1] = w;
    p[2] = cmul(w, w);
    p[3] = cmul(p[2], w);
    p[4] = cmul(p[2], p[2]);
    p[5] = cmul(p[4], w);
    p[6] = cmul(p[4], p[2]);
    p[7] = cmul(p[4], p[3]);
    p[8] = cmul(p[4], p[4]);
#pragma unroll
    for (int r = 9; r < 16; r++) p[r] = cmul(p[8], p[r - 8]);
}

// Pass twiddles of a length-2^LOG2N transform for k = j mod Ns.  Up to 2^13
// from the sine/cosine unit: measured on MI355X, a batch of range
// compressions at 2^11-2^13 points ran 7 % faster and single ones as fast,
// with rel-L2 against FP64 of 1.5-2.2e-7 instead of 2.9-3.6e-7.  At 2^14 and
// above from the table entry w = tw[k N / (Ns 16)] and its products
// (powers16), which measured faster there (2^14: 0.191 against 0.196 ms for
// 4096 x 2048 -> 16384).
template <int LOG2N>
__device__ __forceinline__ void pass_twiddles(const float2 *__restrict__ tw, int k, int Ns,
                                              float2 *p) {
    constexpr int N = 1 << LOG2N;
    if constexpr (LOG2N <= 13) {
        powers16_rev((float)k / (float)(Ns * 16), p);
    } else {
        powers16(tw[k * (N / (Ns * 16))], p);
    }
}


// Lanes and workgroup size of a transform of length 2^log2n: N/16 lanes per
// row, rows packed so a workgroup has at least 256 lanes.
constexpr int lanes(int log2n) { return (1 << log2n) / 16; }
constexpr int rows_per_group(int log2n) { return lanes(log2n) >= 256 ? 1 : 256 / lanes(log2n); }
constexpr int group_threads(int log2n) { return lanes(log2n) * rows_per_group(log2n); }
constexpr int padded(int log2n) { return (1 << log2n) + (1 << log2n) / 16; }

// First-pass radix and butterflies per lane of a length-2^LOG2N transform.
template <int LOG2N>
struct Shape {
    static constexpr int N = 1 << LOG2N;
    static constexpr int T = N / 16;
    static constexpr int B = LOG2N % 4;
    static constexpr int R0 = B == 0 ? 16 : (1 << B);
    static constexpr int NB0 = 16 / R0;
    static constexpr int NPASS16 = (LOG2N - (B == 0 ? 4 : B)) / 4;
    static_assert(LOG2N >= 5, "at least one radix-16 pass after the first");
};

// Unnormalised inverse transform (e^{+j 2 pi k m / N}) of one row by its T
// lanes (lane j < T), through the row's padded LDS buffer L (padded(LOG2N)
// float2).  tw[t] = e^{+j 2 pi t / N}.
//   in:  v[b*R0 + r] = x[j + b*T + r*(N/R0)]  (b < NB0, r < R0)
//   out: v[r]        = X[j + r*T]             (r < 16)
// Starts with a barrier, so the caller may have just read L; every thread
// of the workgroup must call it.  The forward transform is conj(ifft(conj)).
// With active = false a thread only takes part in the barriers (for a
// workgroup wider than the row's T lanes); its v is left unchanged.
template <int LOG2N>
__device__ __forceinline__ void ifft_regs(float2 (&v)[16], float2 *L, int j,
                                          const float2 *__restrict__ tw, bool active = true) {
    using S = Shape<LOG2N>;
    constexpr int N = S::N, T = S::T, R0 = S::R0, NB0 = S::NB0;
    __syncthreads();
    if (active) {
#pragma unroll
        for (int b = 0; b < NB0; b++) {
            const int jb = j + b * T;
            dft<+1, R0>(v + b * R0);
#pragma unroll
            for (int r = 0; r < R0; r++) L[lpad(jb * R0 + r)] = v[b * R0 + r];
        }
    }
    __syncthreads();
    int Ns = R0;
#pragma unroll
    for (int pass = 0; pass < S::NPASS16; pass++) {
        const bool last = (pass == S::NPASS16 - 1);
        const int k = j & (Ns - 1);
        if (active) {
#pragma unroll
            for (int r = 0; r < 16; r++) v[r] = L[lpad(j + r * (N / 16))];
            if (k != 0) {
                float2 p[16];
                pass_twiddles<LOG2N>(tw, k, Ns, p);
#pragma unroll
                for (int r = 1; r < 16; r++) v[r] = cmul(v[r], p[r]);
            }
            dft<+1, 16>(v);
        }
        if (!last) {
            __syncthreads();  // every lane has read before anyone overwrites
            if (active) {
                const int base = (j - k) * 16 + k;
#pragma unroll
                for (int r = 0; r < 16; r++) L[lpad(base + r * Ns)] = v[r];
            }
            __syncthreads();
        }
        Ns *= 16;
    }
}

// Natural-order output of ifft_regs (u[r] = X[j + r*T]) rearranged into the
// pass-0 input order of the next transform of the same length.
template <int LOG2N>
__device__ __forceinline__ void to_input_order(const float2 (&u)[16], float2 (&v)[16]) {
    using S = Shape<LOG2N>;
#pragma unroll
    for (int b = 0; b < S::NB0; b++)
#pragma unroll
        for (int r = 0; r < S::R0; r++) v[b * S::R0 + r] = u[b + r * S::NB0];
}

}  // namespace fft
}  // namespace sarf

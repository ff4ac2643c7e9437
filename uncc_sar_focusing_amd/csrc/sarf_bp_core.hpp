// sarf_bp_core.hpp -- the backprojection tile engine shared by every BP
// kernel (image formation, candidate batches, adjoint).
//
// Reference semantics (src/cpuBackProjection/cpuBackProjection.hpp:24-93):
// for every pulse n and pixel p,
//   dR  = |a_n - p| - R0_n                                   (:70-74)
//   if r_0 < dR < r_{N-1}, r_i = (i/N - 0.5) maxWr           (:76, :45)
//     pos = (dR/maxWr + 0.5) N, left bin i = floor(pos)      (:83, :28-29)
//     v   = s_i + (pos - i)(s_{i+1} - s_i)                   (:38-40)
//     img[x*Ny + y] += v * exp(j 4 pi f0_n dR / c)           (:78-79, :87)
// The interpolation rule is the CPU reference's (first bin with r > dR),
// not the GPU reference's roundf (cuda_BackProjectionKernels.cuh:68-69).
//
// Design for gfx950 (one workgroup = 256 lanes = 4 wave64 per pixel tile):
//   * Lanes run along y (contiguous stores, image is y fastest); each lane
//     owns PX pixels of one row; the tile is TX x TY.
//   * Pulses go in chunks of up to kBpChunk.  For each (tile, pulse) one lane
//     computes in FP64: u = a_n - p_c (p_c = tile centre), R_c = |u|,
//     dR_c = R_c - R0, the profile window the tile needs (from its corners,
//     exact for any geometry) and psi_c = frac(2 f0 dR_c / c) in revolutions.
//     Only these need more than FP32: at R ~ 12.5 km an FP32 range has a
//     1 mm ulp (0.4 rad at 9.6 GHz).
//   * Per pixel, FP32, q = p - p_c (a few metres), uh = u / R_c:
//         s = uh.q,  t = |q|^2 - s^2,
//         dR - dR_c = -s + t (g0 + g1 s),  g0 = 1/(2 R_c), g1 = 1/(2 R_c^2)
//     (expansion of (R_c - s) sqrt(1 + t/(R_c - s)^2) - R_c; remainder
//     t^2/(8 R^3)), so no square root per pixel.
//   * Lane orientation: lanes run along y (XL = false) or along x (XL =
//     true).  Lanes along the cross-range axis read nearly the same profile
//     bin, so the per-pixel window gathers are LDS broadcasts; lanes along
//     the line of sight spread over ~1.4 bins per lane and bank-conflict.
//   * XL tiles use the same expansion in separable form: with P the pixel-
//     axis and L the lane-axis coordinate (uh components u_p, u_l),
//         dR - dR_c = F_p(P) + F_l(L) + P C1(L) + P^2 C2(L),
//     F_p staged per (pulse, column) in LDS and F_l folded into the lane's
//     window-position and phase offsets, so a pixel needs two FMAs for its
//     local range instead of four.
//   * Profile windows are staged in LDS as intercept and slope in
//     window-local bin units, {s_i - i d_i, d_i} with d_i = s_{i+1} - s_i,
//     so the lerp at position p (i = floor p) is one 16-byte LDS read and
//     two FMAs on p itself: no fract.  The intercept's rounding is of the
//     order of ulp(p) |d_i|, the error p itself already carries.  A chunk's
//     windows are packed.
//   * Phase psi = psi_c + (2 f0/c)(dR - dR_c) in revolutions goes straight
//     into v_sin_f32 / v_cos_f32 (they take revolutions and reduce exactly
//     for |psi| <= 256; outside that range they are not valid).
//   * Fast pulses (uniform per (tile, pulse)): the window lies inside the
//     profile, so no per-pixel swath gate, and (2 f0/c)|dR - dR_c| <= 250
//     over the tile, so no phase reduction.  Other pulses take the gated
//     path, which also reduces the phase to one revolution.
#pragma once

#include "sarf_internal.hpp"

namespace sarf {

constexpr int kBpThreads = 256;
// Waves per SIMD the tile kernels are compiled for: 4 (<= 128 VGPRs) up to
// 8 pixels per lane, except lanes-along-x tiles whose column terms take the
// LDS past 40 KiB (3 workgroups per CU); the 16-pixel tiles need more
// registers than that.
template <int TY, int PX, bool XL>
constexpr int bp_waves_per_eu() {
    return PX > 8 ? 1 : (XL && (kBpThreads / TY) * PX > 32) ? 3 : 4;
}
// The same for a tile described at run time (lanes tl, pixels per lane px).
constexpr int bp_waves_per_eu_rt(int tl, int px, bool xl) {
    return px > 8 ? 1 : (xl && (kBpThreads / tl) * px > 32) ? 3 : 4;
}
constexpr int kBpChunk = 32;  // max pulses staged per chunk (the 256-lane tile stages 64)
constexpr int kBpCap = 2048;  // LDS window capacity, in bins (16 B each)
static_assert(kBpChunk <= 64 && kBpCap % kBpThreads == 0, "staging layout");

struct BpArgs {
    int Nx, Ny, Nfft, Np;
    int tiles_x, tiles_y;
    int accumulate;
    int splits;         // pulse groups per tile (> 1: partial images, reduced after)
    int tail;           // > 0 (with splits > 1): split 0 takes the first Np - tail
                        // pulses straight into the image, the other splits share
                        // the last tail pulses in partial images
    long split_stride;  // elements between partial images
    float2 *part;       // partial images (splits > 1)
    double left, bottom, dx, dy;  // pixel (i,j) = (left + i dx, bottom + j dy)
    double maxWr;
    float invbin;  // Nfft / maxWr (bins per metre)
};

// Pulse range [nbeg, nend) of split s of a tile: even groups, or, with a
// tail, one head group of the first Np - tail pulses and the tail pulses in
// even groups over the remaining splits.  Split-major launch order runs every
// tile's head first, so the short tail groups fill in as CUs free up.
__device__ __forceinline__ void split_range(const BpArgs &a, int s, int &nbeg, int &nend) {
    if (a.tail > 0 && a.splits > 1) {
        const int head = a.Np - a.tail;
        if (s == 0) {
            nbeg = 0;
            nend = head;
            return;
        }
        const int per = (a.tail + a.splits - 2) / (a.splits - 1);
        nbeg = min(a.Np, head + (s - 1) * per);
        nend = min(a.Np, nbeg + per);
        return;
    }
    const int per = (a.Np + a.splits - 1) / a.splits;
    nbeg = min(a.Np, s * per);
    nend = min(a.Np, nbeg + per);
}

// Staged state of one chunk of up to CH pulses.
template <int CH>
struct BpLdsT {
    static_assert(CH <= 64, "one wave stages a chunk");
    float4 win[kBpCap];
    float4 pA[CH];  // uh_x, uh_y, g0, g1
    float4 pB[CH];  // pos_c (window-local), psi_c, kappa, W - 2
    int4 pI[CH];    // start bin, LDS offset (bins), W, interior
    float4 pC[CH];  // separable terms (lane axis): c2l, c3l, e1, e2
    float4 pD[CH];  // separable terms: e3, c2p, c3p (pixel axis), -
    int nc;   // pulses staged
    int tot;  // window bins staged
};

// Antenna positions as given.
struct PlainPos {
    const double *ant;
    __device__ __forceinline__ double3 operator()(int n) const {
        return make_double3(ant[3 * n], ant[3 * n + 1], ant[3 * n + 2]);
    }
};

// Candidate trajectory: a_n + (sum_k coef_k basis_k[n]) dir_n -- e.g. a
// polynomial line-of-sight error a t^2 + b t^3 + c t^4 along the unit LOS --
// or, with one direction field per basis (per_basis),
// a_n + sum_k coef_k basis_k[n] dir_k,n -- e.g. absolute positions that are
// polynomials in the pulse index along x, y and z.
struct LosPolyPos {
    const double *ant, *dir, *basis, *coef;
    int nbasis, Np, per_basis;
    __device__ __forceinline__ double3 operator()(int n) const {
        if (per_basis) {
            double x = ant[3 * n], y = ant[3 * n + 1], z = ant[3 * n + 2];
            for (int k = 0; k < nbasis; k++) {
                const double w = coef[k] * basis[(size_t)k * Np + n];
                const double *d = dir + 3 * ((size_t)k * Np + n);
                x = fma(w, d[0], x);
                y = fma(w, d[1], y);
                z = fma(w, d[2], z);
            }
            return make_double3(x, y, z);
        }
        double off = 0.0;
        for (int k = 0; k < nbasis; k++) off = fma(coef[k], basis[(size_t)k * Np + n], off);
        return make_double3(fma(off, dir[3 * n], ant[3 * n]), fma(off, dir[3 * n + 1], ant[3 * n + 1]),
                            fma(off, dir[3 * n + 2], ant[3 * n + 2]));
    }
};

// XCD-aware tile order: workgroup b runs on XCD b % 8 (round-robin
// dispatch); give each XCD a contiguous run of tiles.  Tiles are numbered
// x fastest, so a run is a band of image rows at similar range whose
// profile windows overlap in that XCD's L2.  Bijective on [0, ntiles).
__device__ __forceinline__ int swizzle_tile(int b, int ntiles) {
    const int full = ntiles & ~7;
    if (b >= full) return b;
    const int per = full >> 3;
    return (b & 7) * per + (b >> 3);
}

// XL = false: the TY lanes of a group run along y and each lane owns PX
// pixels along x (tile (256/TY * PX) x TY).  XL = true: the TY lanes run
// along x and each lane owns PX pixels along y (tile TY x (256/TY * PX)).
template <int TY, int PX, bool XL = false>
struct BpTile {
    static constexpr int TL = TY;                  // lanes along the lane axis
    static constexpr int GX = kBpThreads / TL;     // lane groups along the pixel axis
    static constexpr int NP = GX * PX;             // tile extent along the pixel axis
    static constexpr int kPx = PX;                 // pixels per lane
    static constexpr int TX = XL ? TL : NP;        // tile extent along x
    static constexpr int TYE = XL ? NP : TL;       // tile extent along y
    static constexpr bool kSep = XL;               // separable local range
    // Pulses per chunk: 64 for the 256-lane tile, whose narrow windows and
    // 8-column F_p table leave room for them within four workgroups' LDS per
    // CU (fewer chunk rounds: 2048^2 x 4096 pulses 6.88 -> 6.71 ms), 32 else.
    static constexpr int CH = (XL && TL == 256 && PX <= 8) ? 64 : kBpChunk;
    using Lds = BpLdsT<CH>;
    static constexpr int kFx = kSep ? CH * NP : 4;  // LDS floats for F_p
    int ix0, iy0, ly, gx;
    double pcx, pcy, hx, hy;
    float qy;  // lane coordinate, tile-centred (y if !XL, x if XL)
    float qx[PX], q2[PX], x2[PX], accr[PX], acci[PX];  // pixel-axis coordinates

    __device__ __forceinline__ void init(const BpArgs &a, int t) {
        const int ty = t / a.tiles_x, tx = t - ty * a.tiles_x;
        ix0 = tx * TX;
        iy0 = ty * TYE;
        pcx = a.left + (ix0 + 0.5 * (TX - 1)) * a.dx;
        pcy = a.bottom + (iy0 + 0.5 * (TYE - 1)) * a.dy;
        hx = 0.5 * (TX - 1) * a.dx;
        hy = 0.5 * (TYE - 1) * a.dy;
        const int tid = threadIdx.x;
        ly = tid % TL;
        gx = tid / TL;
        const double dl = XL ? a.dx : a.dy, dp = XL ? a.dy : a.dx;
        qy = (float)((ly - 0.5 * (TL - 1)) * dl);
#pragma unroll
        for (int j = 0; j < PX; j++) {
            qx[j] = (float)((gx * PX + j - 0.5 * (NP - 1)) * dp);
            q2[j] = qx[j] * qx[j] + qy * qy;
            x2[j] = qx[j] * qx[j];
            accr[j] = 0.f;
            acci[j] = 0.f;
        }
    }

    __device__ __forceinline__ bool pixel(int j, int &x, int &y, const BpArgs &a) const {
        if (XL) {
            x = ix0 + ly;
            y = iy0 + gx * PX + j;
        } else {
            x = ix0 + gx * PX + j;
            y = iy0 + ly;
        }
        return x < a.Nx && y < a.Ny;
    }

    // Per-(tile,pulse) FP64 set-up for pulses [n0, n0 + cmax): fills L.pA,
    // L.pB, L.pI, L.pC, L.pD, packs windows, sets L.nc; then stages the
    // windows and, for separable tiles given fx, F_p per (pulse, column).
    template <class Pos>
    __device__ __forceinline__ void stage(const BpArgs &a, Lds &L, const float2 *__restrict__ prof,
                                          const Pos &pos, const double *__restrict__ r0,
                                          const double *__restrict__ f0, int n0, int cmax,
                                          float *fx = nullptr) {
        const int tid = threadIdx.x;
        __syncthreads();  // previous users of L are done
        // Wave 0, one lane per pulse of the chunk (cmax <= CH <= 64).
        int W = 0, start = 0, interior = 0;
        if (tid < cmax) {
            const int n = n0 + tid;
            const double3 an = pos(n);
            const double ux = an.x - pcx, uy = an.y - pcy, uz = an.z;
            const double Rc = sqrt(ux * ux + uy * uy + uz * uz);
            const double iR = 1.0 / Rc;
            const double dRc = Rc - r0[n];
            const double posg = (dRc / a.maxWr + 0.5) * a.Nfft;
            // |dR - dR_c| over the tile <= |uh_x| hx + |uh_y| hy + |q|^2/(2 R_c)
            const double ext = fabs(ux * iR) * hx + fabs(uy * iR) * hy + (hx * hx + hy * hy) * 0.5 * iR;
            const double span = ext * a.Nfft / a.maxWr;
            const long st = (long)floor(posg - span) - 1;
            // Clamped so a pulse always fits; the host checks the worst case so
            // the clamp does not bind for valid geometry.
            W = min((int)((long)ceil(posg + span) - st) + 2, kBpCap);
            start = (int)st;
            const double kap = 2.0 * f0[n] / kClight;  // revolutions per metre
            double psi = kap * dRc;
            psi -= rint(psi);
            L.pA[tid] = make_float4((float)(ux * iR), (float)(uy * iR), (float)(0.5 * iR),
                                    (float)(0.5 * iR * iR));
            L.pB[tid] = make_float4((float)(posg - (double)st), (float)psi, (float)kap,
                                    (float)(W - 2));
            {
                // separable form of the same third-order expansion (u_p, u_l:
                // uh along the pixel and lane axes):
                //   F_p = -u_p P + g0 (1-u_p^2) P^2 + g1 u_p (1-u_p^2) P^3  (c2p, c3p)
                //   F_l = -u_l L + g0 (1-u_l^2) L^2 + g1 u_l (1-u_l^2) L^3  (c2l, c3l)
                //   C1  = -2 u_p u_l g0 L + g1 u_p (1-3u_l^2) L^2            (e1, e2)
                //   C2  = g1 u_l (1-3u_p^2) L                                (e3)
                const double ah = ux * iR, bh = uy * iR, g0 = 0.5 * iR, g1 = 0.5 * iR * iR;
                const double up = XL ? bh : ah, ul = XL ? ah : bh;
                const double mp = 1.0 - up * up, ml = 1.0 - ul * ul;
                L.pC[tid] = make_float4((float)(g0 * ml), (float)(g1 * ul * ml),
                                        (float)(-2.0 * up * ul * g0),
                                        (float)(g1 * up * (1.0 - 3.0 * ul * ul)));
                L.pD[tid] = make_float4((float)(g1 * ul * (1.0 - 3.0 * up * up)), (float)(g0 * mp),
                                        (float)(g1 * up * mp), 0.f);
            }
            // Fast path: the window lies inside the profile (no swath gate) and
            // |psi| <= 0.5 + kap * ext stays inside v_sin/v_cos's input range
            // (|x| <= 256 revolutions), so the phase needs no reduction.
            interior = (st >= 0 && st + W <= a.Nfft && kap * ext <= 250.0) ? 1 : 0;
        }
        if (tid < 64) {
            // Windows are packed back to back: offsets are the exclusive
            // prefix sum of W over the chunk (a wave scan), and the pulses
            // whose windows fit in LDS are a prefix of the chunk.
            int incl = W;
#pragma unroll
            for (int d = 1; d < CH; d <<= 1) {
                const int t = __shfl_up(incl, d, 64);
                if (tid >= d) incl += t;
            }
            const int nc = __popcll(__ballot(tid < cmax && incl <= kBpCap));
            if (tid < cmax) L.pI[tid] = make_int4(start, incl - W, W, interior);
            if (tid == nc - 1) {
                L.nc = nc;
                L.tot = incl;
            }
        }
        __syncthreads();
        const int nc = L.nc, tot = L.tot;
        // Window fill over the chunk's bins as one flat range, in rounds of
        // kSlots bins per thread whose loads are all issued before any is
        // consumed (one memory latency per round instead of one per pulse):
        // bin f belongs to the last pulse c with offset <= f.
        constexpr int kSlots = 4;
        const float2 *__restrict__ base = prof + (size_t)n0 * a.Nfft;  // chunk's first profile
        for (int f0r = 0; f0r < tot; f0r += kSlots * kBpThreads) {
            float2 v0[kSlots], v1[kSlots];
            int wb[kSlots];
#pragma unroll
            for (int k = 0; k < kSlots; k++) {
                const int f = f0r + tid + k * kBpThreads;
                v0[k] = make_float2(0.f, 0.f);
                v1[k] = make_float2(0.f, 0.f);
                wb[k] = 0;
                if (f < tot) {
                    int c = 0;
#pragma unroll
                    for (int step = CH / 2; step >= 1; step >>= 1)
                        if (c + step < nc && L.pI[c + step].y <= f) c += step;
                    const int4 I = L.pI[c];
                    const int b = f - I.y;
                    const int g = I.x + b;  // profile bin; |g| < 2 Nfft
                    const unsigned o = (unsigned)(c * a.Nfft + g);  // < CH * Nfft
                    if (g >= 0 && g < a.Nfft) v0[k] = base[o];
                    if (g + 1 >= 0 && g + 1 < a.Nfft) v1[k] = base[o + 1];
                    wb[k] = b;
                }
            }
#pragma unroll
            for (int k = 0; k < kSlots; k++) {
                const int f = f0r + tid + k * kBpThreads;
                if (f < tot) {
                    const float dr = v1[k].x - v0[k].x, di = v1[k].y - v0[k].y;
                    const float fb = (float)wb[k];
                    L.win[f] = make_float4(fmaf(-fb, dr, v0[k].x), fmaf(-fb, di, v0[k].y), dr, di);
                }
            }
        }
        if (kSep && fx != nullptr) {
            static_assert(kBpThreads % NP == 0, "a thread keeps one column");
            const int col = tid % NP;
            const double dp = XL ? a.dy : a.dx;
            const float P = (float)((col - 0.5 * (NP - 1)) * dp);  // == qx of that column
            for (int c = tid / NP; c < nc; c += kBpThreads / NP) {
                const float4 A = L.pA[c];
                const float up = XL ? A.y : A.x;
                const float4 D = L.pD[c];
                fx[c * NP + col] = P * fmaf(P, fmaf(P, D.z, D.y), -up);
            }
        }
        __syncthreads();
    }

    // Separable form (XL tiles): per pulse the lane forms F_l, C1, C2 and
    // folds F_l into its window-position and phase offsets; per pixel
    // dl = F_p + P C1 + P^2 C2 = dR - dR_c - F_l.
    __device__ __forceinline__ void accumulate_chunk_sep(const BpArgs &a, const Lds &L,
                                                         const float *__restrict__ fx) {
        const float invbin = a.invbin;
        const int nc = L.nc;
        for (int c = 0; c < nc; c++) {
            const float4 A = L.pA[c];
            const float4 B = L.pB[c];
            const int4 I = L.pI[c];
            const float4 C = L.pC[c];
            const float e3 = L.pD[c].x;
            const float ul = XL ? A.x : A.y;
            const float Fl = qy * fmaf(qy, fmaf(qy, C.y, C.x), -ul);
            const float C1 = qy * fmaf(qy, C.w, C.z);
            const float C2 = qy * e3;
            const float bx = fmaf(Fl, invbin, B.x);  // window position at dl = 0
            const float by = fmaf(Fl, B.z, B.y);     // phase (revolutions) at dl = 0
            float F[PX];
#pragma unroll
            for (int j = 0; j < PX; j += 4) {
                const float4 f4 = *reinterpret_cast<const float4 *>(fx + c * NP + gx * PX + j);
                F[j] = f4.x;
                F[j + 1] = f4.y;
                F[j + 2] = f4.z;
                F[j + 3] = f4.w;
            }
            const float wmax = B.w;
            if (I.w) {
                // interior: no gate, no clamp (as in accumulate_chunk)
#pragma unroll
                for (int j = 0; j < PX; j++) {
                    const float dl = fmaf(x2[j], C2, fmaf(qx[j], C1, F[j]));
                    const float p = fmaf(dl, invbin, bx);
                    const int i = (int)p;
                    const float4 w = L.win[I.y + i];
                    const float vr = fmaf(p, w.z, w.x);
                    const float vi = fmaf(p, w.w, w.y);
                    const float ph = fmaf(dl, B.z, by);  // revolutions
                    const float sn = __builtin_amdgcn_sinf(ph);
                    const float cs = __builtin_amdgcn_cosf(ph);
                    accr[j] = fmaf(vr, cs, fmaf(-vi, sn, accr[j]));
                    acci[j] = fmaf(vr, sn, fmaf(vi, cs, acci[j]));
                }
            } else {
                // swath gate and one-revolution phase (as in accumulate_chunk)
                const float glo = (float)(-I.x);
                const float ghi = (float)(a.Nfft - 1 - I.x);
#pragma unroll
                for (int j = 0; j < PX; j++) {
                    const float dl = fmaf(x2[j], C2, fmaf(qx[j], C1, F[j]));
                    const float p = fmaf(dl, invbin, bx);
                    const int i = (int)__builtin_amdgcn_fmed3f(p, 0.f, wmax);
                    const float4 w = L.win[I.y + i];
                    const bool ok = p > glo && p < ghi;
                    const float vr = ok ? fmaf(p, w.z, w.x) : 0.f;
                    const float vi = ok ? fmaf(p, w.w, w.y) : 0.f;
                    const float ph = __builtin_amdgcn_fractf(fmaf(dl, B.z, by));
                    const float sn = __builtin_amdgcn_sinf(ph);
                    const float cs = __builtin_amdgcn_cosf(ph);
                    accr[j] = fmaf(vr, cs, fmaf(-vi, sn, accr[j]));
                    acci[j] = fmaf(vr, sn, fmaf(vi, cs, acci[j]));
                }
            }
        }
    }

    // Accumulate the staged chunk into the lane's pixels.
    __device__ __forceinline__ void accumulate_chunk(const BpArgs &a, const Lds &L,
                                                     const float *__restrict__ fx) {
        if constexpr (kSep) {
            accumulate_chunk_sep(a, L, fx);
            return;
        }
        const float invbin = a.invbin;
        const int nc = L.nc;
        for (int c = 0; c < nc; c++) {
            const float4 A = L.pA[c];
            const float4 B = L.pB[c];
            const int4 I = L.pI[c];
            const float ul = XL ? A.x : A.y;  // uh along the lane axis
            const float up = XL ? A.y : A.x;  // uh along the pixel axis
            const float sy = ul * qy;
            const float wmax = B.w;
            if (I.w) {
                // Interior: the window has one bin of margin on each side of
                // [pos_c - span, pos_c + span] and span bounds |dl| * invbin,
                // so 1 <= p <= W - 2 up to FP32 rounding (< 0.01 bin) and the
                // truncating conversion needs no clamp.
#pragma unroll
                for (int j = 0; j < PX; j++) {
                    const float s = fmaf(up, qx[j], sy);
                    const float tq = fmaf(-s, s, q2[j]);
                    const float g = fmaf(s, A.w, A.z);
                    const float dl = fmaf(tq, g, -s);  // dR - dR_c
                    const float p = fmaf(dl, invbin, B.x);
                    const int i = (int)p;
                    const float4 w = L.win[I.y + i];
                    const float vr = fmaf(p, w.z, w.x);
                    const float vi = fmaf(p, w.w, w.y);
                    const float ph = fmaf(dl, B.z, B.y);  // revolutions
                    const float sn = __builtin_amdgcn_sinf(ph);
                    const float cs = __builtin_amdgcn_cosf(ph);
                    accr[j] = fmaf(vr, cs, fmaf(-vi, sn, accr[j]));
                    acci[j] = fmaf(vr, sn, fmaf(vi, cs, acci[j]));
                }
            } else {
                // swath gate r_0 < dR < r_{N-1}  <=>  0 < pos_global < N - 1;
                // the phase is reduced to one revolution (high carriers or
                // large tiles take it past the sine unit's range)
                const float glo = (float)(-I.x);
                const float ghi = (float)(a.Nfft - 1 - I.x);
#pragma unroll
                for (int j = 0; j < PX; j++) {
                    const float s = fmaf(up, qx[j], sy);
                    const float tq = fmaf(-s, s, q2[j]);
                    const float g = fmaf(s, A.w, A.z);
                    const float dl = fmaf(tq, g, -s);
                    const float p = fmaf(dl, invbin, B.x);
                    const int i = (int)__builtin_amdgcn_fmed3f(p, 0.f, wmax);
                    const float4 w = L.win[I.y + i];
                    const bool ok = p > glo && p < ghi;
                    const float vr = ok ? fmaf(p, w.z, w.x) : 0.f;
                    const float vi = ok ? fmaf(p, w.w, w.y) : 0.f;
                    const float ph = __builtin_amdgcn_fractf(fmaf(dl, B.z, B.y));
                    const float sn = __builtin_amdgcn_sinf(ph);
                    const float cs = __builtin_amdgcn_cosf(ph);
                    accr[j] = fmaf(vr, cs, fmaf(-vi, sn, accr[j]));
                    acci[j] = fmaf(vr, sn, fmaf(vi, cs, acci[j]));
                }
            }
        }
    }

    // All pulses in [nbeg, nend); fx: LDS of kFx floats.
    template <class Pos>
    __device__ __forceinline__ void run(const BpArgs &a, Lds &L, float *fx,
                                        const float2 *__restrict__ prof, const Pos &pos,
                                        const double *__restrict__ r0,
                                        const double *__restrict__ f0, int nbeg, int nend) {
        for (int n0 = nbeg; n0 < nend;) {
            stage(a, L, prof, pos, r0, f0, n0, min(CH, nend - n0), fx);
            accumulate_chunk(a, L, fx);
            n0 += L.nc;
        }
    }
};

// Wave64 + workgroup sum of two doubles; result valid on thread 0.
__device__ __forceinline__ double2 block_sum2(double a, double b, double2 *red /* [4] */) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wave] = make_double2(a, b);
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
            r.x += red[w].x;
            r.y += red[w].y;
        }
    }
    return r;
}

}  // namespace sarf

// sarf_adjoint.hip -- adjoint backprojection: the gradient of the image
// entropy with respect to every antenna phase centre (no reference
// counterpart; it turns grid-search autofocus into gradient autofocus).
//
//   E = -sum_p q_p ln q_p,  q_p = |I_p|^2 / S,  S = sum |I_p|^2
//   I_p = sum_n v_n(dR_np) e^{j 2 pi kappa_n dR_np},  dR_np = |a_n - p| - R0_n
//   dE/da_n = sum_p c_np (a_n - p)/|a_n - p|,
//   c_np = 2 w_p Re( conj(I_p) (v'_n + j 2 pi kappa_n v_n) e^{j 2 pi kappa_n dR} )
//   w_p  = dE/d|I_p|^2 = -(ln q_p + E) / S
// with v the linear interpolant of the range profile (the forward rule) and
// v' its slope; R0 is the radar's fixed dechirp reference.
//
// The interpolated profile values are recomputed per pixel and pulse with
// the same tile engine as the forward kernel (nothing per pixel-pulse is
// stored).  Per (tile, pulse) the kernel reduces
//   A = sum c,  B = sum c (dR - dR_c),  Cx = sum c q_x,  Cy = sum c q_y
// and uses (a - p)/|a - p| = uh_c (1 - d/R_c) - q/R_c + O(|q|^2/R_c^2)
// (d = dR - dR_c, q = p - p_c) to get the tile's share
//   g_n = uh_c (A - B/R_c) - (Cx, Cy, 0)/R_c.
// Lanes accumulate 4 pulses x 4 sums in registers; a reduce-scatter
// butterfly over the wave leaves one sum per lane, and each wave writes its
// own share (no workgroup barrier per pulse group).  Shares land in
// partial[n][tile][wave] (FP64) and are summed in order by a second kernel
// (deterministic).  Pulse ranges are split over
// workgroups when the image has too few tiles to fill the GPU.
#include <algorithm>
#include <cmath>

#include "sarf_bp_core.hpp"

namespace sarf {

int bp_pick_tile(const sarf_bp_desc *d, double dx, double dy);
int bp_tile_shape(int id, int *tx, int *ty);
int bp_fill_args(sarf_ctx *ctx, const sarf_bp_desc *d, int tile, BpArgs *a);

namespace {

constexpr int kSub = 4;          // pulses reduced together
constexpr int kNv = 4 * kSub;    // sums per lane per group (A, B, Cx, Cy per pulse)
constexpr int kLogNv = 4;

// Reduce-scatter of kNv values over a wave64: after log2(kNv) halving steps
// lane l holds the sum over a group of kNv lanes of value index idx(l); two
// more xor steps finish the wave sum (replicated in 64/kNv lanes).
__device__ __forceinline__ float wave_reduce_scatter(float (&v)[kNv], int lane) {
#pragma unroll
    for (int step = 0; step < kLogNv; step++) {
        const int o = 1 << step;
        const int half = (kNv / 2) >> step;
        // all-ones in lanes that keep the upper half.  Kept opaque so the
        // choice stays a bit blend of two registers: written as a select of
        // v[i] / v[i + half] the compiler folds it into a lane-variable
        // index into v, which costs a compare/select chain over all kNv.
        unsigned m = 0u - (unsigned)((lane >> step) & 1);
        __asm__ volatile("" : "+v"(m));
#pragma unroll
        for (int i = 0; i < half; i++) {
            const unsigned lo = __float_as_uint(v[i]), hi = __float_as_uint(v[i + half]);
            const float keep = __uint_as_float((hi & m) | (lo & ~m));
            const float send = __uint_as_float((lo & m) | (hi & ~m));
            v[i] = keep + __shfl_xor(send, o, 64);
        }
    }
    float r = v[0];
#pragma unroll
    for (int o = kNv; o < 64; o <<= 1) r += __shfl_xor(r, o, 64);
    return r;
}

// Value index held by lane l (< kNv) after wave_reduce_scatter: bit b of l
// (step b kept the upper half) has weight kNv / 2^(b+1), i.e. the index is
// l bit-reversed over log2(kNv) bits; the map is its own inverse, so value v
// is held by lane scatter_index(v).
__device__ __forceinline__ int scatter_index(int lane) {
    int idx = 0;
#pragma unroll
    for (int b = 0; b < kLogNv; b++)
        if (lane & (1 << b)) idx += (kNv >> 1) >> b;
    return idx;
}

constexpr int kWaves = kBpThreads / 64;

template <int TY, int PX, bool XL>
__global__ __launch_bounds__(kBpThreads) void bp_entropy_adjoint(
    BpArgs a, const float2 *__restrict__ prof, const double *__restrict__ ant,
    const double *__restrict__ r0, const double *__restrict__ f0, const float2 *__restrict__ img,
    const double *__restrict__ ent, double *__restrict__ partial) {
    __shared__ typename BpTile<TY, PX, XL>::Lds L;
    __shared__ float uz[BpTile<TY, PX, XL>::CH];
    // blockIdx.x = split * ntiles + tile; each split covers a pulse range
    const int ntiles = a.tiles_x * a.tiles_y;
    const int split = blockIdx.x / ntiles;
    const int t = swizzle_tile(blockIdx.x - split * ntiles, ntiles);
    const int per_split = (a.Np + a.splits - 1) / a.splits;
    const int nbeg = min(a.Np, split * per_split);
    const int nend = min(a.Np, nbeg + per_split);
    BpTile<TY, PX, XL> T;
    T.init(a, t);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;

    // per-pixel forward image times the entropy weight 2 w_p and times
    // invbin: the profile slope below is per bin, dv/dR = invbin * slope, so
    // z = v' + j k v is formed per bin as (slope + j (k/invbin) v) and the
    // factor invbin rides on these weights
    const float invbin = a.invbin;
    const double E = ent[0], S = ent[1];
    const float invS = (float)(1.0 / S), lnS = (float)log(S), Ef = (float)E;
    float WI[PX], WQ[PX];
#pragma unroll
    for (int j = 0; j < PX; j++) {
        int x, y;
        WI[j] = WQ[j] = 0.f;
        if (T.pixel(j, x, y, a)) {
            const float2 v = img[(size_t)x * a.Ny + y];
            const float m = v.x * v.x + v.y * v.y;
            const float wg = m > 0.f ? -2.f * (__logf(m) - lnS + Ef) * invS : 0.f;
            WI[j] = wg * invbin * v.x;
            WQ[j] = wg * invbin * v.y;
        }
    }
    const float two_pi = 6.283185307179586f;
    const PlainPos pos{ant};

    for (int n0 = nbeg; n0 < nend;) {
        T.stage(a, L, prof, pos, r0, f0, n0, min(BpTile<TY, PX, XL>::CH, nend - n0));
        // uh_z for the chunk (stage() keeps uh_x, uh_y, 1/(2R_c) in pA)
        if (threadIdx.x < L.nc) {
            const int n = n0 + threadIdx.x;
            const double ux = ant[3 * n] - T.pcx, uy = ant[3 * n + 1] - T.pcy, zz = ant[3 * n + 2];
            uz[threadIdx.x] = (float)(zz / sqrt(ux * ux + uy * uy + zz * zz));
        }
        __syncthreads();
        const int nc = L.nc;
        for (int c0 = 0; c0 < nc; c0 += kSub) {
            float acc[kNv];
#pragma unroll
            for (int i = 0; i < kNv; i++) acc[i] = 0.f;
#pragma unroll
            for (int k = 0; k < kSub; k++) {
                const int c = c0 + k;
                if (c < nc) {
                    const float4 A = L.pA[c];
                    const float4 B = L.pB[c];
                    const int4 I = L.pI[c];
                    const float ul = XL ? A.x : A.y;  // uh along the lane axis
                    const float up = XL ? A.y : A.x;  // uh along the pixel axis
                    const float sy = ul * T.qy;
                    const float wmax = B.w;
                    const float kb = two_pi * B.z / invbin;  // rad per bin
                    float sA = 0.f, sB = 0.f, sC = 0.f;
                    if (I.w) {
                        // fast pulse (see the forward kernel): no gate, no
                        // clamp, the phase within the sine unit's range
#pragma unroll
                        for (int j = 0; j < PX; j++) {
                            const float s = fmaf(up, T.qx[j], sy);
                            const float tq = fmaf(-s, s, T.q2[j]);
                            const float g = fmaf(s, A.w, A.z);
                            const float dl = fmaf(tq, g, -s);
                            const float p = fmaf(dl, invbin, B.x);
                            const int i = (int)p;
                            const float4 w = L.win[I.y + i];
                            const float vr = fmaf(p, w.z, w.x);
                            const float vi = fmaf(p, w.w, w.y);
                            const float ph = fmaf(dl, B.z, B.y);
                            const float sn = __builtin_amdgcn_sinf(ph);
                            const float cs = __builtin_amdgcn_cosf(ph);
                            // z / invbin = slope + j (k/invbin) v
                            const float zr = fmaf(-kb, vi, w.z);
                            const float zi = fmaf(kb, vr, w.w);
                            const float er = fmaf(zr, cs, -zi * sn);
                            const float ei = fmaf(zr, sn, zi * cs);
                            const float cpn = fmaf(WI[j], er, WQ[j] * ei);
                            sA += cpn;
                            sB = fmaf(cpn, dl, sB);
                            sC = fmaf(cpn, T.qx[j], sC);
                        }
                    } else {
                        const float glo = (float)(-I.x), ghi = (float)(a.Nfft - 1 - I.x);
#pragma unroll
                        for (int j = 0; j < PX; j++) {
                            const float s = fmaf(up, T.qx[j], sy);
                            const float tq = fmaf(-s, s, T.q2[j]);
                            const float g = fmaf(s, A.w, A.z);
                            const float dl = fmaf(tq, g, -s);
                            const float p = fmaf(dl, invbin, B.x);
                            const int i = (int)__builtin_amdgcn_fmed3f(p, 0.f, wmax);
                            const float4 w = L.win[I.y + i];
                            const float vr = fmaf(p, w.z, w.x);
                            const float vi = fmaf(p, w.w, w.y);
                            // one revolution: v_sin/v_cos take |x| <= 256
                            const float ph = __builtin_amdgcn_fractf(fmaf(dl, B.z, B.y));
                            const float sn = __builtin_amdgcn_sinf(ph);
                            const float cs = __builtin_amdgcn_cosf(ph);
                            const float zr = fmaf(-kb, vi, w.z);
                            const float zi = fmaf(kb, vr, w.w);
                            const float er = fmaf(zr, cs, -zi * sn);
                            const float ei = fmaf(zr, sn, zi * cs);
                            float cpn = fmaf(WI[j], er, WQ[j] * ei);
                            cpn = (p > glo && p < ghi) ? cpn : 0.f;
                            sA += cpn;
                            sB = fmaf(cpn, dl, sB);
                            sC = fmaf(cpn, T.qx[j], sC);
                        }
                    }
                    // sC sums over the pixel axis, sA * qy is the lane axis
                    acc[4 * k + 0] = sA;
                    acc[4 * k + 1] = sB;
                    acc[4 * k + 2] = XL ? sA * T.qy : sC;  // Cx
                    acc[4 * k + 3] = XL ? sC : sA * T.qy;  // Cy
                }
            }
            const float r = wave_reduce_scatter(acc, lane);
            // lane k < kSub gathers pulse c0 + k's four sums of this wave
            const int kk = lane & (kSub - 1);
            float s4[4];
#pragma unroll
            for (int q = 0; q < 4; q++) s4[q] = __shfl(r, scatter_index(4 * kk + q), 64);
            if (lane < kSub && c0 + lane < nc) {
                const int c = c0 + lane;
                const float4 A = L.pA[c];
                const double invR = 2.0 * (double)A.z;
                const double base = (double)s4[0] - (double)s4[1] * invR;
                double *o = partial + (((size_t)(n0 + c) * ntiles + t) * kWaves + wave) * 3;
                o[0] = (double)A.x * base - (double)s4[2] * invR;
                o[1] = (double)A.y * base - (double)s4[3] * invR;
                o[2] = (double)uz[c] * base;
            }
        }
        n0 += nc;
    }
}

// grad[n] = sum over the nshares (tile, wave) shares of partial[n][.].
__global__ __launch_bounds__(64) void adjoint_reduce(const double *__restrict__ partial,
                                                     int nshares, int Np,
                                                     double *__restrict__ grad) {
    const int n = blockIdx.x;
    const int lane = threadIdx.x;
    double g[3] = {0.0, 0.0, 0.0};
    for (int t = lane; t < nshares; t += 64) {
        const double *p = partial + ((size_t)n * nshares + t) * 3;
        g[0] += p[0];
        g[1] += p[1];
        g[2] += p[2];
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) g[q] += __shfl_xor(g[q], o, 64);
    }
    if (lane == 0) {
        grad[3 * n] = g[0];
        grad[3 * n + 1] = g[1];
        grad[3 * n + 2] = g[2];
    }
}

template <int TY, int PX, bool XL = false>
void launch_adj_t(const BpArgs &a, int grid, const float2 *prof, const double *ant,
                  const double *r0, const double *f0, const float2 *img, const double *ent,
                  double *partial, hipStream_t s) {
    hipLaunchKernelGGL((bp_entropy_adjoint<TY, PX, XL>), dim3(grid), dim3(kBpThreads), 0, s, a, prof,
                       ant, r0, f0, img, ent, partial);
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_bp_entropy_grad(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof,
                                    const double *ant, const double *r0, const double *f0,
                                    const void *img, const double *ent, double *grad,
                                    void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!d || !prof || !ant || !r0 || !f0 || !img || !ent || !grad)
        return fail(SARF_ERR_INVALID, "sarf_bp_entropy_grad: null argument");
    if (d->precision != SARF_BP_MIXED)
        return fail(SARF_ERR_UNSUPPORTED, "sarf_bp_entropy_grad: mixed precision only");
    // default: the forward kernel's choice (tiles 6 and 7, 16 pixels per
    // lane, are not instantiated here and fall back to 4)
    int tile = d->tile;
    if (tile <= 0) {
        BpArgs tmp{};
        SARF_TRY(bp_fill_args(ctx, d, 0, &tmp));
        tile = bp_pick_tile(d, tmp.dx, tmp.dy);
        if (tile == 6 || tile == 7) tile = 4;
    }
    BpArgs a{};
    SARF_TRY(bp_fill_args(ctx, d, tile, &a));
    int tx = 0, ty = 0;
    const int tid = bp_tile_shape(tile, &tx, &ty);
    const int ntiles = a.tiles_x * a.tiles_y;
    // Fill the GPU: pulse groups share nothing (partials are per tile and
    // pulse), so split until there are ~16 workgroups per CU (measured best
    // at 512^2 x 1024), >= 32 pulses each.
    int splits = d->pulse_split > 0 ? d->pulse_split : 1;
    if (d->pulse_split <= 0) {
        const int want = 16 * (ctx->cu_count > 0 ? ctx->cu_count : 256);
        while (splits < 64 && ntiles * splits < want && d->Np / (2 * splits) >= 32) splits *= 2;
    }
    a.splits = std::max(1, std::min(splits, d->Np));
    void *buf = nullptr;
    SARF_TRY(scratch_get(ctx, "adjoint.partials",
                         sizeof(double) * 3 * (size_t)ntiles * kWaves * d->Np, &buf));
    double *partial = (double *)buf;
    hipStream_t s = (hipStream_t)stream;
    const float2 *p = (const float2 *)prof;
    const float2 *im = (const float2 *)img;
    const int grid = ntiles * a.splits;
    switch (tid) {
        case 1: launch_adj_t<32, 4>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 2: launch_adj_t<64, 4>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 3: launch_adj_t<64, 8>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 4: launch_adj_t<32, 8>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 5: launch_adj_t<16, 4>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 11: launch_adj_t<32, 4, true>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 12: launch_adj_t<64, 4, true>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 13: launch_adj_t<64, 8, true>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 14: launch_adj_t<32, 8, true>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 15: launch_adj_t<256, 8, true>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        case 16: launch_adj_t<128, 8, true>(a, grid, p, ant, r0, f0, im, ent, partial, s); break;
        default:
            return fail(SARF_ERR_UNSUPPORTED, "sarf_bp_entropy_grad: tile %d not compiled", tid);
    }
    SARF_HIP(hipGetLastError());
    hipLaunchKernelGGL(adjoint_reduce, dim3(d->Np), dim3(64), 0, s, partial, ntiles * kWaves,
                       d->Np, grad);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

// sarf_bp.hip -- backprojection of range profiles onto the z = 0 pixel grid.
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
// Mixed-precision kernel (SARF_BP_MIXED) -- design for gfx950:
//   * A workgroup (256 lanes = 4 wave64) owns a TX x TY pixel tile; each lane
//     owns PX pixels of one image row (same y), lanes run along y so image
//     stores are contiguous (y fastest, column-major Ny x Nx).
//   * Pulses are processed in chunks.  For each (tile, pulse) lanes compute
//     in FP64: u = a_n - p_c (p_c = tile centre), R_c = |u|, dR_c = R_c - R0,
//     the tile's profile window, and the phase of the centre reduced to
//     revolutions, psi_c = frac(2 f0 dR_c / c).  Only these reductions need
//     more than FP32: at R ~ 12.5 km an FP32 range has a 1 mm ulp (0.4 rad
//     at 9.6 GHz).
//   * Per pixel, with q = p - p_c (|q| of a few metres, exact enough in
//     FP32) and the unit vector uh = u/R_c:
//         s = uh.q,  t = |q|^2 - s^2,
//         dR - dR_c = -s + t / (2 (R_c - s)) + O(t^2/R_c^3)
//     which needs no square root; 1/(2(R_c - s)) ~ g0 + g1 s.
//   * The window of profile bins each pulse needs for the tile (found from
//     the tile corners) is staged through LDS as {s_i, s_{i+1} - s_i} pairs,
//     so the lerp is one 16-byte LDS read and two FMAs.
//   * Phase in revolutions: psi = frac(psi_c + (2 f0/c) (dR - dR_c)), then
//     v_sin_f32 / v_cos_f32, which take revolutions.
//   FP32 for interpolation, rotation and accumulation; no MFMA (this is not
//   a contraction).
//
// FP64 kernel (SARF_BP_F64): one lane per pixel, every quantity in double,
// the reference's search rule and FP64 sin/cos: the "all-FP64 build" the
// mixed kernel is measured against at full size.
#include <cmath>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

constexpr int kBpThreads = 256;
constexpr int kBpChunk = 32;    // max pulses staged per chunk
constexpr int kBpCap = 2048;    // LDS window capacity, in bins (16 B each)

struct BpArgs {
    int Nx, Ny, Nfft, Np;
    int tiles_x, tiles_y;
    int accumulate;
    double left, bottom, dx, dy;  // pixel (i,j) = (left + i dx, bottom + j dy)
    double maxWr;
    float invbin;                 // Nfft / maxWr (bins per metre)
};

// XCD-aware tile order: workgroup b runs on XCD b % 8 (round-robin
// dispatch); give each XCD a contiguous run of tiles.  Tiles are numbered
// x fastest, so a run is a band of image rows at similar range and its
// profile windows overlap in that XCD's L2.  Bijective on [0, ntiles).
__device__ __forceinline__ int swizzle_tile(int b, int ntiles) {
    const int full = ntiles & ~7;
    if (b >= full) return b;
    const int per = full >> 3;
    return (b & 7) * per + (b >> 3);
}

template <int TY, int PX>
__global__ __launch_bounds__(kBpThreads) void bp_mixed(BpArgs a, const float2 *__restrict__ prof,
                                                       const double *__restrict__ ant,
                                                       const double *__restrict__ r0,
                                                       const double *__restrict__ f0,
                                                       float2 *__restrict__ img) {
    constexpr int GX = kBpThreads / TY;  // lane groups along x
    constexpr int TX = GX * PX;
    __shared__ float4 win[kBpCap];
    __shared__ float4 pA[kBpChunk];  // uh_x, uh_y, g0, g1
    __shared__ float4 pB[kBpChunk];  // pos_c (window-local), psi_c, kappa, gate flag
    __shared__ int4 pI[kBpChunk];    // start bin, LDS offset (bins), W, interior
    __shared__ int sNc;

    const int t = swizzle_tile(blockIdx.x, a.tiles_x * a.tiles_y);
    const int ty = t / a.tiles_x, tx = t - ty * a.tiles_x;
    const int ix0 = tx * TX, iy0 = ty * TY;
    const double cxi = ix0 + 0.5 * (TX - 1), cyi = iy0 + 0.5 * (TY - 1);
    const double pcx = a.left + cxi * a.dx, pcy = a.bottom + cyi * a.dy;
    const double hx = 0.5 * (TX - 1) * a.dx, hy = 0.5 * (TY - 1) * a.dy;

    const int tid = threadIdx.x;
    const int ly = tid % TY, gx = tid / TY;
    const float qy = (float)((ly - 0.5 * (TY - 1)) * a.dy);
    float qx[PX], q2[PX], accr[PX], acci[PX];
#pragma unroll
    for (int j = 0; j < PX; j++) {
        qx[j] = (float)((gx * PX + j - 0.5 * (TX - 1)) * a.dx);
        q2[j] = qx[j] * qx[j] + qy * qy;
        accr[j] = 0.f;
        acci[j] = 0.f;
    }
    const float invbin = a.invbin;

    for (int n0 = 0; n0 < a.Np;) {
        const int cmax = min(kBpChunk, a.Np - n0);
        __syncthreads();  // previous chunk finished with LDS
        if (tid < cmax) {
            const int n = n0 + tid;
            const double ux = ant[3 * n] - pcx, uy = ant[3 * n + 1] - pcy, uz = ant[3 * n + 2];
            const double Rc = sqrt(ux * ux + uy * uy + uz * uz);
            const double iR = 1.0 / Rc;
            const double dRc = Rc - r0[n];
            const double posg = (dRc / a.maxWr + 0.5) * a.Nfft;
            // |dR - dR_c| over the tile <= |uh_x| hx + |uh_y| hy + |q|^2/(2 R_c)
            const double ext = fabs(ux * iR) * hx + fabs(uy * iR) * hy + (hx * hx + hy * hy) * 0.5 * iR;
            const double span = ext * a.Nfft / a.maxWr;
            const long start = (long)floor(posg - span) - 1;
            // Clamped so a pulse always fits; the host checks the worst case so
            // the clamp never binds for valid geometry.
            const int W = min((int)((long)ceil(posg + span) - start) + 2, kBpCap);
            const double kap = 2.0 * f0[n] / kClight;  // revolutions per metre
            double psi = kap * dRc;
            psi -= rint(psi);
            pA[tid] = make_float4((float)(ux * iR), (float)(uy * iR), (float)(0.5 * iR),
                                  (float)(0.5 * iR * iR));
            pB[tid] = make_float4((float)(posg - (double)start), (float)psi, (float)kap, 0.f);
            const int interior = (start >= 0 && start + W <= a.Nfft) ? 1 : 0;
            pI[tid] = make_int4((int)start, 0, W, interior);
        }
        __syncthreads();
        if (tid == 0) {
            // pack windows back to back while they fit
            int off = 0, c = 0;
            for (; c < cmax; c++) {
                const int W = pI[c].z;
                if (off + W > kBpCap) break;
                pI[c].y = off;
                off += W;
            }
            sNc = c;
        }
        __syncthreads();
        const int nc = sNc;  // >= 1: a single window always fits (host-checked)

        // stage {s_b, s_{b+1} - s_b}: wave w takes pulses w, w+4, ...
        {
            const int wave = tid >> 6, lane = tid & 63;
            for (int c = wave; c < nc; c += kBpThreads / 64) {
                const int4 I = pI[c];
                const float2 *src = prof + (size_t)(n0 + c) * a.Nfft;
                for (int b = lane; b < I.z; b += 64) {
                    const long g = (long)I.x + b;
                    float2 s0 = make_float2(0.f, 0.f), s1 = make_float2(0.f, 0.f);
                    if (g >= 0 && g < a.Nfft) s0 = src[g];
                    if (g + 1 >= 0 && g + 1 < a.Nfft) s1 = src[g + 1];
                    win[I.y + b] = make_float4(s0.x, s0.y, s1.x - s0.x, s1.y - s0.y);
                }
            }
        }
        __syncthreads();

        for (int c = 0; c < nc; c++) {
            const float4 A = pA[c];
            const float4 B = pB[c];
            const int4 I = pI[c];
            const float sy = A.y * qy;
            const int wmax = I.z - 2;
            if (I.w) {
#pragma unroll
                for (int j = 0; j < PX; j++) {
                    const float s = fmaf(A.x, qx[j], sy);
                    const float tq = fmaf(-s, s, q2[j]);
                    const float g = fmaf(s, A.w, A.z);
                    const float dl = fmaf(tq, g, -s);  // dR - dR_c
                    const float pos = fmaf(dl, invbin, B.x);
                    const float fr = __builtin_amdgcn_fractf(pos);
                    const int i = min(max((int)pos, 0), wmax);
                    const float4 w = win[I.y + i];
                    const float vr = fmaf(fr, w.z, w.x);
                    const float vi = fmaf(fr, w.w, w.y);
                    float ph = fmaf(dl, B.z, B.y);
                    ph = __builtin_amdgcn_fractf(ph);
                    const float sn = __builtin_amdgcn_sinf(ph);
                    const float cs = __builtin_amdgcn_cosf(ph);
                    accr[j] = fmaf(vr, cs, fmaf(-vi, sn, accr[j]));
                    acci[j] = fmaf(vr, sn, fmaf(vi, cs, acci[j]));
                }
            } else {
                // swath gate r_0 < dR < r_{N-1}  <=>  0 < pos_global < N - 1
                const float glo = (float)(-I.x);
                const float ghi = (float)(a.Nfft - 1 - I.x);
#pragma unroll
                for (int j = 0; j < PX; j++) {
                    const float s = fmaf(A.x, qx[j], sy);
                    const float tq = fmaf(-s, s, q2[j]);
                    const float g = fmaf(s, A.w, A.z);
                    const float dl = fmaf(tq, g, -s);
                    const float pos = fmaf(dl, invbin, B.x);
                    const float fr = __builtin_amdgcn_fractf(pos);
                    const int i = min(max((int)pos, 0), wmax);
                    const float4 w = win[I.y + i];
                    const bool ok = pos > glo && pos < ghi;
                    const float vr = ok ? fmaf(fr, w.z, w.x) : 0.f;
                    const float vi = ok ? fmaf(fr, w.w, w.y) : 0.f;
                    float ph = fmaf(dl, B.z, B.y);
                    ph = __builtin_amdgcn_fractf(ph);
                    const float sn = __builtin_amdgcn_sinf(ph);
                    const float cs = __builtin_amdgcn_cosf(ph);
                    accr[j] = fmaf(vr, cs, fmaf(-vi, sn, accr[j]));
                    acci[j] = fmaf(vr, sn, fmaf(vi, cs, acci[j]));
                }
            }
        }
        n0 += nc;
    }

    const int y = iy0 + ly;
    if (y < a.Ny) {
#pragma unroll
        for (int j = 0; j < PX; j++) {
            const int x = ix0 + gx * PX + j;
            if (x < a.Nx) {
                float2 *o = img + (size_t)x * a.Ny + y;
                if (a.accumulate) {
                    const float2 old = *o;
                    *o = make_float2(old.x + accr[j], old.y + acci[j]);
                } else {
                    *o = make_float2(accr[j], acci[j]);
                }
            }
        }
    }
}

// All-FP64 reference-precision kernel, one lane per pixel.
__global__ __launch_bounds__(256) void bp_f64(BpArgs a, const double2 *__restrict__ prof,
                                              const double *__restrict__ ant,
                                              const double *__restrict__ r0,
                                              const double *__restrict__ f0,
                                              double2 *__restrict__ img) {
    const long pix = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (long)a.Nx * a.Ny) return;
    const int x = (int)(pix / a.Ny), y = (int)(pix - (long)x * a.Ny);
    const double px = a.left + x * a.dx, py = a.bottom + y * a.dy;
    const int N = a.Nfft;
    const double rmin = -0.5 * a.maxWr, rmax = ((double)(N - 1) / N - 0.5) * a.maxWr;
    double accr = 0.0, acci = 0.0;
    for (int n = 0; n < a.Np; n++) {
        const double ddx = ant[3 * n] - px, ddy = ant[3 * n + 1] - py, ddz = ant[3 * n + 2];
        const double dR = sqrt(ddx * ddx + ddy * ddy + ddz * ddz) - r0[n];
        if (!(dR > rmin && dR < rmax)) continue;
        const double idx = (dR / a.maxWr + 0.5) * N;
        int ri = (int)floor(idx) - 1;
        if (ri < -1) ri = -1;
        // first bin with r > dR (cpuBackProjection.hpp:28-29)
        while (++ri < N && ((double)ri / N - 0.5) * a.maxWr <= dR) {
        }
        if (ri == N || ri == 0) continue;
        const double rl = ((double)(ri - 1) / N - 0.5) * a.maxWr;
        const double rr = ((double)ri / N - 0.5) * a.maxWr;
        const double al = (dR - rl) / (rr - rl);
        const double2 sl = prof[(size_t)n * N + ri - 1], sr = prof[(size_t)n * N + ri];
        const double vr = al * sr.x + (1.0 - al) * sl.x;
        const double vi = al * sr.y + (1.0 - al) * sl.y;
        const double ph = 4.0 * kPi * f0[n] * dR / kClight;
        double sn, cs;
        sincos(ph, &sn, &cs);
        accr += vr * cs - vi * sn;
        acci += vr * sn + vi * cs;
    }
    double2 *o = img + pix;
    if (a.accumulate) {
        *o = make_double2(o->x + accr, o->y + acci);
    } else {
        *o = make_double2(accr, acci);
    }
}

// Tile shapes compiled in: (TY, PX) -> TX = 256/TY * PX.
struct TileShape {
    int id, ty, px, tx;
};
constexpr TileShape kTiles[] = {
    {1, 32, 4, 32},  // 32 x 32
    {2, 64, 4, 16},  // 16 x 64
    {3, 64, 8, 32},  // 32 x 64
    {4, 32, 8, 64},  // 64 x 32
    {5, 16, 4, 64},  // 64 x 16
};

template <int TY, int PX>
void launch_mixed_t(const BpArgs &a, int grid, const float2 *prof, const double *ant,
                    const double *r0, const double *f0, float2 *img, hipStream_t s) {
    hipLaunchKernelGGL((bp_mixed<TY, PX>), dim3(grid), dim3(kBpThreads), 0, s, a, prof, ant, r0,
                       f0, img);
}

}  // namespace

int bp_tile_count() { return (int)(sizeof(kTiles) / sizeof(kTiles[0])); }

// Default tile: the largest tile whose half-extent stays within ~1.65 m in
// each direction (keeps the FP32 local offsets well inside the 1e-4
// rel-L2 budget) and whose worst-case window fits the LDS capacity.
static int pick_tile(const sarf_bp_desc *d, double dx, double dy) {
    const double invbin = d->Nfft / d->maxWr;
    int best = -1;
    long best_px = 0;
    for (const TileShape &t : kTiles) {
        const double hx = 0.5 * (t.tx - 1) * dx, hy = 0.5 * (t.ty - 1) * dy;
        if (hx > 1.65 || hy > 1.65) continue;
        const double wmax = 2.0 * (hx + hy) * invbin + 8;
        if (wmax > kBpCap) continue;
        const long px = (long)t.tx * t.ty;
        if (px > best_px) {
            best_px = px;
            best = t.id;
        }
    }
    return best < 0 ? 1 : best;
}

int launch_backproject(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof, const double *ant,
                       const double *r0, const double *f0, void *img, hipStream_t s) {
    (void)ctx;
    if (d->Nx < 2 || d->Ny < 2 || d->Nfft < 2 || d->Np < 1)
        return fail(SARF_ERR_INVALID, "backproject: bad sizes Nx=%d Ny=%d Nfft=%d Np=%d", d->Nx,
                    d->Ny, d->Nfft, d->Np);
    if (!(d->maxWr > 0) || !(d->Wx > 0) || !(d->Wy > 0) || !std::isfinite(d->x0) ||
        !std::isfinite(d->y0))
        return fail(SARF_ERR_INVALID, "backproject: bad geometry maxWr=%g Wx=%g Wy=%g", d->maxWr,
                    d->Wx, d->Wy);
    BpArgs a{};
    a.Nx = d->Nx;
    a.Ny = d->Ny;
    a.Nfft = d->Nfft;
    a.Np = d->Np;
    a.accumulate = (d->flags & SARF_BP_ACCUMULATE) ? 1 : 0;
    a.dx = d->Wx / (d->Nx - 1);
    a.dy = d->Wy / (d->Ny - 1);
    a.left = d->x0 - d->Wx / 2;
    a.bottom = d->y0 - d->Wy / 2;
    a.maxWr = d->maxWr;
    a.invbin = (float)(d->Nfft / d->maxWr);

    if (d->precision == SARF_BP_F64) {
        const long npix = (long)d->Nx * d->Ny;
        hipLaunchKernelGGL(bp_f64, dim3((npix + 255) / 256), dim3(256), 0, s, a,
                           (const double2 *)prof, ant, r0, f0, (double2 *)img);
        SARF_HIP(hipGetLastError());
        return SARF_OK;
    }
    if (d->precision != SARF_BP_MIXED)
        return fail(SARF_ERR_INVALID, "backproject: unknown precision %d", d->precision);

    const int tile = d->tile > 0 ? d->tile : pick_tile(d, a.dx, a.dy);
    const TileShape *ts = nullptr;
    for (const TileShape &t : kTiles)
        if (t.id == tile) ts = &t;
    if (!ts) return fail(SARF_ERR_INVALID, "backproject: unknown tile id %d", tile);
    {
        const double hx = 0.5 * (ts->tx - 1) * a.dx, hy = 0.5 * (ts->ty - 1) * a.dy;
        const double wmax = 2.0 * (hx + hy) * (d->Nfft / d->maxWr) + 8;
        if (wmax > kBpCap)
            return fail(SARF_ERR_UNSUPPORTED,
                        "backproject: tile %d spans %.0f range bins (> %d); pixel spacing too "
                        "coarse for the range bin width", tile, wmax, kBpCap);
    }
    a.tiles_x = (d->Nx + ts->tx - 1) / ts->tx;
    a.tiles_y = (d->Ny + ts->ty - 1) / ts->ty;
    const int grid = a.tiles_x * a.tiles_y;
    const float2 *p = (const float2 *)prof;
    float2 *o = (float2 *)img;
    switch (ts->id) {
        case 1: launch_mixed_t<32, 4>(a, grid, p, ant, r0, f0, o, s); break;
        case 2: launch_mixed_t<64, 4>(a, grid, p, ant, r0, f0, o, s); break;
        case 3: launch_mixed_t<64, 8>(a, grid, p, ant, r0, f0, o, s); break;
        case 4: launch_mixed_t<32, 8>(a, grid, p, ant, r0, f0, o, s); break;
        case 5: launch_mixed_t<16, 4>(a, grid, p, ant, r0, f0, o, s); break;
    }
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

}  // namespace sarf

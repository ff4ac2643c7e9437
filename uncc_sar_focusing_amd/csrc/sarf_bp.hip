// sarf_bp.hip -- image-forming backprojection kernels and their launcher.
//
// bp_mixed (SARF_BP_MIXED): the tile engine of sarf_bp_core.hpp over plain
// antenna positions, writing the image (or a per-pulse-group partial).
// bp_f64 (SARF_BP_F64): one lane per pixel, every quantity in double, the
// reference's search rule (cpuBackProjection.hpp:28-29) and FP64 sin/cos:
// the "all-FP64 build" the mixed kernel is measured against at full size.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sarf_bp_core.hpp"

namespace sarf {
namespace {

template <int TY, int PX, bool XL>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<TY, PX, XL>())))
void bp_mixed(BpArgs a, const float2 *__restrict__ prof,
                                                       const double *__restrict__ ant,
                                                       const double *__restrict__ r0,
                                                       const double *__restrict__ f0,
                                                       float2 *__restrict__ img) {
    __shared__ typename BpTile<TY, PX, XL>::Lds L;
    __shared__ __align__(16) float fx[BpTile<TY, PX, XL>::kFx];
    // blockIdx.x = split * ntiles + tile: every split of a tile covers a
    // contiguous pulse range and writes its own partial image -- except the
    // head group of a head + tail split, which writes the image itself
    const int ntiles = a.tiles_x * a.tiles_y;
    const int split = blockIdx.x / ntiles;
    const int t = swizzle_tile(blockIdx.x - split * ntiles, ntiles);
    int nbeg, nend;
    split_range(a, split, nbeg, nend);
    int accumulate = a.accumulate;
    if (a.splits > 1 && !(a.tail > 0 && split == 0)) {
        img = a.part + (size_t)(a.tail > 0 ? split - 1 : split) * a.split_stride;
        accumulate = 0;
    }

    BpTile<TY, PX, XL> T;
    T.init(a, t);
    T.run(a, L, fx, prof, PlainPos{ant}, r0, f0, nbeg, nend);
#pragma unroll
    for (int j = 0; j < PX; j++) {
        int x, y;
        if (T.pixel(j, x, y, a)) {
            float2 *o = img + (size_t)x * a.Ny + y;
            if (accumulate) {
                const float2 old = *o;
                *o = make_float2(old.x + T.accr[j], old.y + T.acci[j]);
            } else {
                *o = make_float2(T.accr[j], T.acci[j]);
            }
        }
    }
}

// Sums the per-split partial images in split order (deterministic) into img.
__global__ __launch_bounds__(256) void bp_reduce_splits(const float2 *__restrict__ part,
                                                        long stride, int splits, long n,
                                                        int accumulate, float2 *__restrict__ img) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float2 acc = accumulate ? img[i] : make_float2(0.f, 0.f);
    // partials added in split order; eight loads issued before any is added
    int s = 0;
    for (; s + 8 <= splits; s += 8) {
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = part[(s + u) * stride + i];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            acc.x += v[u].x;
            acc.y += v[u].y;
        }
    }
    for (; s < splits; s++) {
        const float2 v = part[s * stride + i];
        acc.x += v.x;
        acc.y += v.y;
    }
    img[i] = acc;
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

// Tile shapes compiled in: id, extents tx x ty (pixels), lanes TL along y
// (xl = 0) or along x (xl = 1), PX pixels per lane along the other axis.
// Tiles 15 and 16 are long across the line of sight and short along it (for
// a line of sight along y): their profile windows are the narrowest, and with
// 256 or 128 lanes along x every lane coordinate belongs to one wave only, so
// the per-(pulse, lane) terms of the separable range are formed once per
// workgroup instead of once per wave (2048^2 at 0.05 m, 4096 pulses: tile 15
// 6.83-6.89 ms, tile 16 6.93-6.98, tile 13 7.18-7.23).
struct TileShape {
    int id, tx, ty, tl, px, xl;
};
constexpr TileShape kTiles[] = {
    {1, 32, 32, 32, 4, 0},
    {2, 16, 64, 64, 4, 0},
    {3, 32, 64, 64, 8, 0},
    {4, 64, 32, 32, 8, 0},
    {5, 64, 16, 16, 4, 0},
    {6, 64, 64, 64, 16, 0},
    {7, 128, 32, 32, 16, 0},
    {11, 32, 32, 32, 4, 1},
    {12, 64, 16, 64, 4, 1},
    {13, 64, 32, 64, 8, 1},
    {14, 32, 64, 32, 8, 1},
    {15, 256, 8, 256, 8, 1},
    {16, 128, 16, 128, 8, 1},
};

template <int TY, int PX, bool XL = false>
void launch_mixed_t(const BpArgs &a, int grid, const float2 *prof, const double *ant,
                    const double *r0, const double *f0, float2 *img, hipStream_t s) {
    hipLaunchKernelGGL((bp_mixed<TY, PX, XL>), dim3(grid), dim3(kBpThreads), 0, s, a, prof, ant,
                       r0, f0, img);
}

}  // namespace

std::vector<int> bp_tile_ids() {
    std::vector<int> ids;
    for (const TileShape &t : kTiles) ids.push_back(t.id);
    return ids;
}

// The three fastest lanes-along-x tiles and their measured relative costs
// (2048^2 at 0.05 m, 4096 pulses).
constexpr int kFastXl[3] = {15, 16, 13};
constexpr double kFastXlCost[3] = {1.0, 1.013, 1.05};

// Default tile: among the three fastest lanes-along-x tiles (with a hint
// along y), the cheapest by their measured relative speed divided by the
// fraction of the tile grid's pixels that fall inside the image (a
// 290-pixel-wide image in 256-wide tiles computes 512 columns); otherwise
// the first of a measured speed ranking (MI355X, 2048^2 at 0.05 m and 256^2
// at 0.1 m) -- always among the tiles whose line-of-sight-projected
// half extent |los_x| hx + |los_y| hy -- the largest |dR - dR_c| over the
// tile, which sets the FP32 rounding of the local range -- stays within
// 2.6 m (about 3e-5 relative L2 measured) and whose worst-case window fits
// in LDS.  With no LOS hint each axis counts fully and the bound is 3.2 m.
// With a hint, tiles whose lanes run along the cross-range axis rank first
// (their window gathers are LDS broadcasts): lanes along x when
// |los_y| >= |los_x|.  Relative costs at 2048^2, 4096 pulses: tile 15 1.00,
// 16 1.013, 13 1.05.  (Over 64 scenes of 128-1024 pixels a side, one by one,
// tile 13 against 15: 15 is faster on widths just under a multiple of 256,
// slower by up to 20 % on those just above one.)
int bp_pick_tile(const sarf_bp_desc *d, double dx, double dy) {
    static const int kRankY[] = {4, 5, 1, 3, 7, 2, 6};
    static const int kRankX[] = {kFastXl[0], kFastXl[1], kFastXl[2], 4, 5, 1, 3, 14, 12, 11, 7, 2, 6};
    const double invbin = d->Nfft / d->maxWr;
    const bool hint = (d->los_x != 0.0 || d->los_y != 0.0);
    const double lx = hint ? std::fabs(d->los_x) : 1.0, ly = hint ? std::fabs(d->los_y) : 1.0;
    const double bound = hint ? 2.6 : 3.2;
    const bool xlanes = hint && ly >= lx;
    const int *rank = xlanes ? kRankX : kRankY;
    const int nrank = xlanes ? (int)(sizeof(kRankX) / sizeof(int)) : (int)(sizeof(kRankY) / sizeof(int));
    int best = 0;
    double best_cost = 0.0;
    for (int k = 0; k < nrank; k++) {
        const int id = rank[k];
        for (const TileShape &t : kTiles) {
            if (t.id != id) continue;
            const double hx = 0.5 * (t.tx - 1) * dx, hy = 0.5 * (t.ty - 1) * dy;
            if (lx * hx + ly * hy > bound) continue;
            if (2.0 * (hx + hy) * invbin + 8 > kBpCap) continue;
            if (!xlanes || k >= 3) {
                if (best == 0) return id;  // below the measured three: rank order
                continue;
            }
            const double ux = (double)d->Nx / (((d->Nx + t.tx - 1) / t.tx) * (double)t.tx);
            const double uy = (double)d->Ny / (((d->Ny + t.ty - 1) / t.ty) * (double)t.ty);
            const double cost = kFastXlCost[k] / (ux * uy);
            if (best == 0 || cost < best_cost) {
                best = id;
                best_cost = cost;
            }
        }
    }
    return best > 0 ? best : 1;
}

int bp_tile_shape(int id, int *tx, int *ty);

// Whether tile `id` meets the default choice's accuracy bound (LOS-projected
// half extent) and fits its worst-case profile window in LDS.
bool bp_tile_accurate(const sarf_bp_desc *d, int id) {
    int tx = 0, ty = 0;
    if (!bp_tile_shape(id, &tx, &ty) || d->Nx < 2 || d->Ny < 2 || !(d->maxWr > 0)) return false;
    const double dx = d->Wx / (d->Nx - 1), dy = d->Wy / (d->Ny - 1);
    const double invbin = d->Nfft / d->maxWr;
    const bool hint = (d->los_x != 0.0 || d->los_y != 0.0);
    const double lx = hint ? std::fabs(d->los_x) : 1.0, ly = hint ? std::fabs(d->los_y) : 1.0;
    const double bound = hint ? 2.6 : 3.2;
    const double hx = 0.5 * (tx - 1) * dx, hy = 0.5 * (ty - 1) * dy;
    return lx * hx + ly * hy <= bound && 2.0 * (hx + hy) * invbin + 8 <= kBpCap;
}

// Of tiles 15, 16 and 13 (the lanes-along-x tiles of the default pick), the
// accurate one with the lowest measured relative cost over the share of its
// tile grid's pixels inside the image, as bp_pick_tile weighs them; 0 if
// none is accurate.
int bp_pick_xl_tile(const sarf_bp_desc *d) {
    int best = 0;
    double best_cost = 0.0;
    for (int k = 0; k < 3; k++) {
        int tx = 0, ty = 0;
        if (!bp_tile_accurate(d, kFastXl[k]) || !bp_tile_shape(kFastXl[k], &tx, &ty)) continue;
        const double ux = (double)d->Nx / (((d->Nx + tx - 1) / tx) * (double)tx);
        const double uy = (double)d->Ny / (((d->Ny + ty - 1) / ty) * (double)ty);
        const double cost = kFastXlCost[k] / (ux * uy);
        if (best == 0 || cost < best_cost) {
            best = kFastXl[k];
            best_cost = cost;
        }
    }
    return best;
}

// Returns the tile id if known (and its TX, TY), else 0.
int bp_tile_shape(int id, int *tx, int *ty) {
    for (const TileShape &t : kTiles)
        if (t.id == id) {
            *tx = t.tx;
            *ty = t.ty;
            return id;
        }
    return 0;
}

// Waves per SIMD the tile's kernel is built for (= resident workgroups per
// CU: 256-thread workgroups, 4 SIMDs); 1 for an unknown id.
int tile_waves_per_eu(int tile) {
    for (const TileShape &t : kTiles)
        if (t.id == tile) return bp_waves_per_eu_rt(t.tl, t.px, t.xl != 0);
    return 1;
}

// Validates a mixed/F64 descriptor and fills the kernel arguments for tile
// `tile` (0 = default choice).  Splits are left at 1.
int bp_fill_args(sarf_ctx *ctx, const sarf_bp_desc *d, int tile, BpArgs *out) {
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
    a.splits = 1;
    a.tail = 0;
    a.split_stride = 0;
    if (d->precision == SARF_BP_MIXED) {
        if (tile <= 0) tile = bp_pick_tile(d, a.dx, a.dy);
        int tx = 0, ty = 0;
        if (!bp_tile_shape(tile, &tx, &ty))
            return fail(SARF_ERR_INVALID, "backproject: unknown tile id %d", tile);
        const double hx = 0.5 * (tx - 1) * a.dx, hy = 0.5 * (ty - 1) * a.dy;
        const double wmax = 2.0 * (hx + hy) * (d->Nfft / d->maxWr) + 8;
        if (wmax > kBpCap)
            return fail(SARF_ERR_UNSUPPORTED,
                        "backproject: tile %d spans %.0f range bins (> %d); pixel spacing too "
                        "coarse for the range bin width", tile, wmax, kBpCap);
        a.tiles_x = (d->Nx + tx - 1) / tx;
        a.tiles_y = (d->Ny + ty - 1) / ty;
    }
    *out = a;
    return SARF_OK;
}

namespace {

// The geometry a remembered launch shape applies to (spacings as the
// backprojection computes them, W / (N - 1)).
sarf_tune_entry tune_geometry(const sarf_bp_desc *d) {
    sarf_tune_entry e{};
    e.Nx = d->Nx;
    e.Ny = d->Ny;
    e.Np = d->Np;
    e.Nfft = d->Nfft;
    e.dx = d->Wx / (d->Nx - 1);
    e.dy = d->Wy / (d->Ny - 1);
    e.maxWr = d->maxWr;
    e.los_x = d->los_x;
    e.los_y = d->los_y;
    return e;
}

std::array<uint64_t, 8> tune_key(const sarf_tune_entry &e) {
    auto bits = [](double v) {
        uint64_t u;
        std::memcpy(&u, &v, sizeof u);
        return u;
    };
    return {(uint64_t)e.Nx, (uint64_t)e.Ny, (uint64_t)e.Np, (uint64_t)e.Nfft,
            bits(e.dx), bits(e.dy), bits(e.maxWr), bits(e.los_x) * 31 + bits(e.los_y)};
}

std::array<uint64_t, 8> tune_key(const sarf_bp_desc *d) { return tune_key(tune_geometry(d)); }

// Pulse splits.  A tile's pulses are split into groups, each a workgroup of
// its own writing a partial image, summed in order afterwards.  Two rules:
// * head + tail, when the tiles alone fill at least two rounds of resident
//   workgroups: one head group of the first 7/8 of the pulses per tile
//   writes the image directly, and the last 1/8 is spread over splits - 1
//   short groups that the split-major launch order runs last, filling in as
//   CUs free up (default 3 splits).  Measured on MI355X at 2048^2 x 4096
//   pulses, tile 13 (2048 tiles, 1024 resident): 7.24 ms, against 7.32 ms
//   for 8 even groups and 8.04 ms unsplit, with 2 partial images instead of 8
//   (a tail of 1/4, 1/6, 1/12 or 1/16 of the pulses measured 7.27-7.58 ms);
// * even groups otherwise (few tiles: the groups are what fills the GPU),
//   doubling until there are 64 workgroups per CU with at least 16 pulses
//   per group (512^2 x 1024 pulses best at 16-32, 256^2 x 512 at 16).

bool split_head_tail(const sarf_ctx *ctx, int tile, int ntiles, int Np) {
    const long resident = (long)(ctx->cu_count > 0 ? ctx->cu_count : 256) * tile_waves_per_eu(tile);
    return (long)ntiles >= 2 * resident && Np >= 64;
}

int default_splits(const sarf_ctx *ctx, int tile, int ntiles, int Np) {
    if (split_head_tail(ctx, tile, ntiles, Np)) return 3;
    int splits = 1;
    const long want = 64L * (ctx->cu_count > 0 ? ctx->cu_count : 256);
    while (splits < 32 && (long)ntiles * splits < want && Np / (2 * splits) >= 16) splits *= 2;
    return splits;
}

// Tail pulses of a launch with `splits` groups (0: even groups).
int split_tail(const sarf_ctx *ctx, int tile, int ntiles, int Np, int splits) {
    if (splits < 2 || !split_head_tail(ctx, tile, ntiles, Np)) return 0;
    const int tail = std::max(splits - 1, Np / 8);
    return tail < Np ? tail : 0;
}

// Partial images a launch writes.
int split_partials(int splits, int tail) { return splits < 2 ? 0 : tail > 0 ? splits - 1 : splits; }

// Mixed-precision launch with a resolved tile id (> 0) and split count
// (0 = default rule).
// ext_part (with room for ext_bytes) replaces the context's partial-image
// scratch when given.
int launch_mixed_shape(sarf_ctx *ctx, const sarf_bp_desc *d, int tile, int splits,
                       const void *prof, const double *ant, const double *r0, const double *f0,
                       void *img, hipStream_t s, float2 *ext_part = nullptr,
                       size_t ext_bytes = 0) {
    BpArgs a{};
    SARF_TRY(bp_fill_args(ctx, d, tile, &a));
    int tx = 0, ty = 0;
    const int tid = bp_tile_shape(tile, &tx, &ty);
    const int ntiles = a.tiles_x * a.tiles_y;
    if (splits <= 0) splits = default_splits(ctx, tile, ntiles, d->Np);
    splits = std::max(1, std::min(splits, d->Np));
    const long npix = (long)d->Nx * d->Ny;
    float2 *o = (float2 *)img;
    a.splits = splits;
    a.tail = split_tail(ctx, tile, ntiles, d->Np, splits);
    a.split_stride = 0;
    a.part = nullptr;
    const int nparts = split_partials(splits, a.tail);
    if (nparts > 0) {
        void *buf = ext_part;
        const size_t need = sizeof(float2) * (size_t)npix * nparts;
        if (ext_part) {
            if (need > ext_bytes)
                return fail(SARF_ERR_INVALID, "backproject: %d pulse splits of a %dx%d image need "
                            "%zu bytes of partial images, more than the %zu available", splits,
                            d->Nx, d->Ny, need, ext_bytes);
        } else {
            SARF_TRY(scratch_get(ctx, "bp.partials", need, &buf));
        }
        a.part = (float2 *)buf;
        a.split_stride = npix;
    }
    const int grid = ntiles * splits;
    const float2 *p = (const float2 *)prof;
    switch (tid) {
        case 1: launch_mixed_t<32, 4>(a, grid, p, ant, r0, f0, o, s); break;
        case 2: launch_mixed_t<64, 4>(a, grid, p, ant, r0, f0, o, s); break;
        case 3: launch_mixed_t<64, 8>(a, grid, p, ant, r0, f0, o, s); break;
        case 4: launch_mixed_t<32, 8>(a, grid, p, ant, r0, f0, o, s); break;
        case 5: launch_mixed_t<16, 4>(a, grid, p, ant, r0, f0, o, s); break;
        case 6: launch_mixed_t<64, 16>(a, grid, p, ant, r0, f0, o, s); break;
        case 7: launch_mixed_t<32, 16>(a, grid, p, ant, r0, f0, o, s); break;
        case 11: launch_mixed_t<32, 4, true>(a, grid, p, ant, r0, f0, o, s); break;
        case 12: launch_mixed_t<64, 4, true>(a, grid, p, ant, r0, f0, o, s); break;
        case 13: launch_mixed_t<64, 8, true>(a, grid, p, ant, r0, f0, o, s); break;
        case 14: launch_mixed_t<32, 8, true>(a, grid, p, ant, r0, f0, o, s); break;
        case 15: launch_mixed_t<256, 8, true>(a, grid, p, ant, r0, f0, o, s); break;
        case 16: launch_mixed_t<128, 8, true>(a, grid, p, ant, r0, f0, o, s); break;
        default: return fail(SARF_ERR_INVALID, "backproject: unknown tile id %d", tid);
    }
    SARF_HIP(hipGetLastError());
    if (nparts > 0) {
        // even groups: img (+)= sum of all partials; head + tail: the head
        // group wrote (or added into) img, the tail partials are added to it
        const int acc = a.tail > 0 || (d->flags & SARF_BP_ACCUMULATE) ? 1 : 0;
        hipLaunchKernelGGL(bp_reduce_splits, dim3((npix + 255) / 256), dim3(256), 0, s, a.part,
                           (long)npix, nparts, npix, acc, (float2 *)img);
        SARF_HIP(hipGetLastError());
    }
    return SARF_OK;
}

}  // namespace

int launch_reduce_partials(const float2 *part, long stride, int splits, long n, int accumulate,
                           float2 *img, hipStream_t s) {
    hipLaunchKernelGGL(bp_reduce_splits, dim3((n + 255) / 256), dim3(256), 0, s, part, stride,
                       splits, n, accumulate, img);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

// The backprojection of sarf_backproject; with ext_part, partial images go
// to ext_part (ext_bytes) instead of context scratch, and an automatic or
// tuned split count is lowered until they fit.
int launch_backproject_ext(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof,
                           const double *ant, const double *r0, const double *f0, void *img,
                           hipStream_t s, float2 *ext_part, size_t ext_bytes) {
    if (d->precision != SARF_BP_MIXED && d->precision != SARF_BP_F64)
        return fail(SARF_ERR_INVALID, "backproject: unknown precision %d", d->precision);
    if (d->precision == SARF_BP_F64) {
        BpArgs a{};
        SARF_TRY(bp_fill_args(ctx, d, 0, &a));
        const long npix = (long)d->Nx * d->Ny;
        hipLaunchKernelGGL(bp_f64, dim3((npix + 255) / 256), dim3(256), 0, s, a,
                           (const double2 *)prof, ant, r0, f0, (double2 *)img);
        SARF_HIP(hipGetLastError());
        return SARF_OK;
    }
    int tile = d->tile > 0 ? d->tile : 0;
    int splits = d->pulse_split > 0 ? d->pulse_split : 0;
    if (tile == 0 && splits == 0 && d->Nx >= 2 && d->Ny >= 2) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        auto it = ctx->tuned.find(tune_key(d));
        if (it != ctx->tuned.end()) {
            tile = it->second.tile;
            splits = it->second.pulse_split;
        }
    }
    if (tile == 0) {
        if (d->Nx < 2 || d->Ny < 2 || !(d->Wx > 0) || !(d->Wy > 0))
            return fail(SARF_ERR_INVALID, "backproject: bad sizes Nx=%d Ny=%d Wx=%g Wy=%g", d->Nx,
                        d->Ny, d->Wx, d->Wy);
        tile = bp_pick_tile(d, d->Wx / (d->Nx - 1), d->Wy / (d->Ny - 1));
    }
    if (ext_part && d->pulse_split <= 0) {
        int tx = 0, ty = 0;
        if (!bp_tile_shape(tile, &tx, &ty))
            return fail(SARF_ERR_INVALID, "backproject: unknown tile id %d", tile);
        const int ntiles = ((d->Nx + tx - 1) / tx) * ((d->Ny + ty - 1) / ty);
        if (splits <= 0) splits = default_splits(ctx, tile, ntiles, d->Np);
        const size_t img_bytes = sizeof(float2) * (size_t)d->Nx * d->Ny;
        while (splits > 1 &&
               img_bytes * split_partials(splits, split_tail(ctx, tile, ntiles, d->Np, splits)) >
                   ext_bytes)
            splits = split_tail(ctx, tile, ntiles, d->Np, splits) > 0 ? splits - 1 : splits / 2;
    }
    return launch_mixed_shape(ctx, d, tile, splits, prof, ant, r0, f0, img, s, ext_part,
                              ext_bytes);
}

int launch_backproject(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof, const double *ant,
                       const double *r0, const double *f0, void *img, hipStream_t s) {
    return launch_backproject_ext(ctx, d, prof, ant, r0, f0, img, s, nullptr, 0);
}

}  // namespace sarf

using namespace sarf;

namespace {

// Tiles that meet the default choice's accuracy bound (and the LDS window).
std::vector<int> tune_tiles(const sarf_bp_desc *d) {
    std::vector<int> out;
    for (int id : bp_tile_ids())
        if (bp_tile_accurate(d, id)) out.push_back(id);
    return out;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" int sarf_autotune(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof,
                             const double *ant, const double *r0, const double *f0, void *img,
                             int reps, sarf_tune_result *out, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!d || !prof || !ant || !r0 || !f0 || !img || !out)
        return fail(SARF_ERR_INVALID, "sarf_autotune: null argument");
    if (d->precision != SARF_BP_MIXED)
        return fail(SARF_ERR_INVALID, "sarf_autotune: only SARF_BP_MIXED has launch shapes");
    if (d->flags & SARF_BP_ACCUMULATE)
        return fail(SARF_ERR_INVALID, "sarf_autotune: SARF_BP_ACCUMULATE is not allowed");
    if (reps < 1 || reps > 100) return fail(SARF_ERR_INVALID, "sarf_autotune: reps=%d", reps);
    BpArgs chk{};
    SARF_TRY(bp_fill_args(ctx, d, 0, &chk));  // validates sizes and geometry
    hipStream_t s = (hipStream_t)stream;
    sarf_bp_desc dd = *d;
    dd.tile = 0;
    dd.pulse_split = 0;
    EventPair ev;
    SARF_HIP(hipEventCreate(&ev.a));
    SARF_HIP(hipEventCreate(&ev.b));
    auto time_shape = [&](int tile, int splits, float *ms) -> int {
        SARF_TRY(launch_mixed_shape(ctx, &dd, tile, splits, prof, ant, r0, f0, img, s));  // warm
        SARF_HIP(hipEventRecord(ev.a, s));
        for (int r = 0; r < reps; r++)
            SARF_TRY(launch_mixed_shape(ctx, &dd, tile, splits, prof, ant, r0, f0, img, s));
        SARF_HIP(hipEventRecord(ev.b, s));
        SARF_HIP(hipEventSynchronize(ev.b));
        float t = 0.f;
        SARF_HIP(hipEventElapsedTime(&t, ev.a, ev.b));
        *ms = t / reps;
        return SARF_OK;
    };
    const double dx = d->Wx / (d->Nx - 1), dy = d->Wy / (d->Ny - 1);
    const int def_tile = bp_pick_tile(d, dx, dy);
    float def_ms = 0.f;
    SARF_TRY(time_shape(def_tile, 0, &def_ms));
    int best_tile = def_tile, best_split = 0, n = 1;
    float best_ms = def_ms;
    for (int tile : tune_tiles(d)) {
        for (int splits = 1; splits <= 32; splits *= 2) {
            if (splits > 1 && d->Np / splits < 32) break;
            float ms = 0.f;
            SARF_TRY(time_shape(tile, splits, &ms));
            n++;
            // 1 % hysteresis: timing noise alone does not move the pick
            // away from the default or an earlier candidate
            if (ms < 0.99f * best_ms) {
                best_ms = ms;
                best_tile = tile;
                best_split = splits;
            }
        }
    }
    if (best_split == 0) {  // the default won: record its resolved split count
        int tx = 0, ty = 0;
        bp_tile_shape(best_tile, &tx, &ty);
        const int ntiles = ((d->Nx + tx - 1) / tx) * ((d->Ny + ty - 1) / ty);
        best_split = std::max(1, std::min(default_splits(ctx, best_tile, ntiles, d->Np), d->Np));
    }
    // leave the image made with the chosen shape
    SARF_TRY(launch_mixed_shape(ctx, &dd, best_tile, best_split, prof, ant, r0, f0, img, s));
    SARF_HIP(hipStreamSynchronize(s));
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        sarf_tune_entry e = tune_geometry(d);
        e.tile = best_tile;
        e.pulse_split = best_split;
        ctx->tuned[tune_key(e)] = e;
    }
    out->tile = best_tile;
    out->pulse_split = best_split;
    out->ms = best_ms;
    out->default_ms = def_ms;
    out->candidates = n;
    return SARF_OK;
}

extern "C" int sarf_tune_table_get(sarf_ctx *ctx, sarf_tune_entry *out, int cap, int *count) {
    clear_error();
    if (!ctx || !count || cap < 0 || (cap > 0 && !out))
        return fail(SARF_ERR_INVALID, "sarf_tune_table_get: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->mu);
    int i = 0;
    for (const auto &kv : ctx->tuned) {
        if (i < cap) out[i] = kv.second;
        i++;
    }
    *count = i;
    return SARF_OK;
}

extern "C" int sarf_tune_table_set(sarf_ctx *ctx, const sarf_tune_entry *entries, int n) {
    clear_error();
    if (!ctx || n < 0 || (n > 0 && !entries))
        return fail(SARF_ERR_INVALID, "sarf_tune_table_set: bad arguments");
    for (int i = 0; i < n; i++) {
        const sarf_tune_entry &e = entries[i];
        int tx = 0, ty = 0;
        if (e.Nx < 2 || e.Ny < 2 || e.Np < 1 || e.Nfft < 2 || !(e.dx > 0) || !(e.dy > 0) ||
            !(e.maxWr > 0))
            return fail(SARF_ERR_INVALID, "sarf_tune_table_set: entry %d has a bad geometry", i);
        if (!bp_tile_shape(e.tile, &tx, &ty) || e.pulse_split < 1 || e.pulse_split > e.Np)
            return fail(SARF_ERR_INVALID, "sarf_tune_table_set: entry %d: tile %d / pulse_split %d",
                        i, e.tile, e.pulse_split);
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    for (int i = 0; i < n; i++) ctx->tuned[tune_key(entries[i])] = entries[i];
    return SARF_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free backprojection.  Its partial images live in `scratch`, a
// file-scope device array: every GPU the library runs on holds one copy of
// kNoAllocScratchBytes until the process ends.  Host code never passes the
// array's name to a kernel (on the host it only names the symbol); it uses
// the device address hipGetSymbolAddress returned for each device, which
// sarf_ctx_create records through bp_noalloc_init.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
float2 *g_noalloc_ptr[kNoAllocMaxDevices] = {};
}  // namespace

static __device__ __align__(256) float2 scratch[kNoAllocScratchBytes / sizeof(float2)];

namespace sarf {

// Records this device's address of `scratch`; the device must be current.
int bp_noalloc_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        return fail(SARF_ERR_NODEVICE, "device %d: at most %d devices", device,
                    kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return SARF_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SARF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                    hipGetErrorString(e));
    g_noalloc_ptr[device] = (float2 *)p;
    return SARF_OK;
}

}  // namespace sarf

extern "C" int sarf_backproject_noalloc(sarf_ctx *ctx, const sarf_bp_desc *desc,
                                        const void *prof, const double *ant, const double *r0,
                                        const double *f0, void *img, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!desc || !prof || !ant || !r0 || !f0 || !img)
        return fail(SARF_ERR_INVALID, "sarf_backproject_noalloc: null argument");
    float2 *part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        if (ctx->device >= 0 && ctx->device < kNoAllocMaxDevices) part = g_noalloc_ptr[ctx->device];
    }
    if (!part)
        return fail(SARF_ERR_INVALID, "sarf_backproject_noalloc: no scratch address for device %d",
                    ctx->device);
    return launch_backproject_ext(ctx, desc, prof, ant, r0, f0, img, (hipStream_t)stream, part,
                                  kNoAllocScratchBytes);
}

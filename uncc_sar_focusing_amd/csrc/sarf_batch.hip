// sarf_batch.hip -- many candidate images in one persistent launch, scored
// by image entropy without materialising the images.
//
// Replaces the grid-search objective of the reference: kernelWrapper
// (src/autofocus/gridSearchErrorFunctions_dft.cuh:258-336), which rebuilds
// the trajectory of one candidate (:279-291), backprojects it into a device
// new[] image and scores an 8-bit dB 2-D DFT metric, launched per candidate
// by cuGridSearch (src/autofocus/autofocus_gpu.cu:1057-1070; pattern
// autofocus_gpu_optuna.cu:210-308, one kernel per candidate on 1000 streams).
//
// Here: candidate c moves every phase centre along a unit direction,
//   a_n(c) = a_n + (sum_k coef[c,k] basis[k,n]) dir_n,
// e.g. a line-of-sight polynomial error a t^2 + b t^3 + c t^4, or, with one
// direction field per basis, a_n(c) = a_n + sum_k coef[c,k] basis[k,n] dir[k,n],
// which with a_n = 0 and bases 1, i, i^2 along x, y, z is the reference's own
// absolute trajectory model (:279-291).  A persistent
// grid walks (tile, candidate) work items, tile-major, so the workgroups
// resident at one time share a tile and read nearly the same profile windows
// (L2 reuse).  Each item backprojects all pulses into registers with the
// shared tile engine and reduces S = sum |I|^2 and T = sum |I|^2 ln |I|^2 over
// the tile into partials[c][tile] (FP64); a second kernel sums the partials
// in tile order and returns E = ln S - T/S (the entropy of p = |I|^2 / S),
// deterministically.
#include <algorithm>
#include <cmath>

#include "sarf_bp_core.hpp"

namespace sarf {

int bp_pick_tile(const sarf_bp_desc *d, double dx, double dy);
int bp_tile_shape(int id, int *tx, int *ty);
int bp_fill_args(sarf_ctx *ctx, const sarf_bp_desc *d, int tile, BpArgs *a);

namespace {

template <int TY, int PX, bool XL>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(bp_waves_per_eu<TY, PX, XL>())))
void bp_batch_entropy(
    BpArgs a, const float2 *__restrict__ prof, const double *__restrict__ ant,
    const double *__restrict__ r0, const double *__restrict__ f0, const double *__restrict__ dir,
    const double *__restrict__ basis, const double *__restrict__ coef, int nbasis, int ncand,
    int per_basis, double2 *__restrict__ partial) {
    __shared__ typename BpTile<TY, PX, XL>::Lds L;
    __shared__ __align__(16) float fx[BpTile<TY, PX, XL>::kFx];
    __shared__ double2 red[kBpThreads / 64];
    const int ntiles = a.tiles_x * a.tiles_y;
    const long nitems = (long)ntiles * ncand;
    for (long item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int t = (int)(item / ncand);
        const int c = (int)(item - (long)t * ncand);
        BpTile<TY, PX, XL> T;
        T.init(a, t);
        LosPolyPos pos{ant, dir, basis, coef + (size_t)c * nbasis, nbasis, a.Np, per_basis};
        T.run(a, L, fx, prof, pos, r0, f0, 0, a.Np);
        double s = 0.0, tt = 0.0;
#pragma unroll
        for (int j = 0; j < PX; j++) {
            int x, y;
            if (T.pixel(j, x, y, a)) {
                const float m = T.accr[j] * T.accr[j] + T.acci[j] * T.acci[j];
                if (m > 0.f) {
                    s += (double)m;
                    tt += (double)m * (double)__logf(m);
                }
            }
        }
        const double2 r = block_sum2(s, tt, red);
        if (threadIdx.x == 0) partial[(size_t)c * ntiles + t] = r;
    }
}

// E[c] = ln S - T/S from partials[c][0..ntiles), summed in tile order.
__global__ __launch_bounds__(256) void entropy_from_partials(const double2 *__restrict__ partial,
                                                             int ntiles, int ncand,
                                                             double *__restrict__ out) {
    __shared__ double2 red[4];
    const int c = blockIdx.x;
    double s = 0.0, t = 0.0;
    // fixed assignment of tiles to lanes -> deterministic
    for (int i = threadIdx.x; i < ntiles; i += blockDim.x) {
        const double2 v = partial[(size_t)c * ntiles + i];
        s += v.x;
        t += v.y;
    }
    const double2 r = block_sum2(s, t, red);
    if (threadIdx.x == 0) out[c] = r.x > 0.0 ? log(r.x) - r.y / r.x : 0.0;
}

// Entropy of a complex image: per-block partial (S, T), then one block.
template <typename T2>
__global__ __launch_bounds__(256) void image_entropy_partials(const T2 *__restrict__ img, long n,
                                                              double2 *__restrict__ partial) {
    __shared__ double2 red[4];
    double s = 0.0, t = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double m = (double)img[i].x * img[i].x + (double)img[i].y * img[i].y;
        if (m > 0.0) {
            s += m;
            t += m * log(m);
        }
    }
    const double2 r = block_sum2(s, t, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(256) void image_entropy_final(const double2 *__restrict__ partial,
                                                           int nblocks, double *__restrict__ out) {
    __shared__ double2 red[4];
    double s = 0.0, t = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += blockDim.x) {
        s += partial[i].x;
        t += partial[i].y;
    }
    const double2 r = block_sum2(s, t, red);
    if (threadIdx.x == 0) {
        out[0] = r.x > 0.0 ? log(r.x) - r.y / r.x : 0.0;
        out[1] = r.x;
        out[2] = r.y;
    }
}

template <int TY, int PX, bool XL = false>
void launch_batch_t(const BpArgs &a, int grid, const float2 *prof, const double *ant,
                    const double *r0, const double *f0, const double *dir, const double *basis,
                    const double *coef, int nbasis, int ncand, int per_basis, double2 *partial,
                    hipStream_t s) {
    hipLaunchKernelGGL((bp_batch_entropy<TY, PX, XL>), dim3(grid), dim3(kBpThreads), 0, s, a,
                       prof, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial);
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_bp_batch_entropy(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof,
                                     const double *ant, const double *r0, const double *f0,
                                     int ncand, int nbasis, const double *coef,
                                     const double *basis, const double *dir, double *entropy,
                                     void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!d || !prof || !ant || !r0 || !f0 || !entropy || (nbasis > 0 && (!coef || !basis || !dir)))
        return fail(SARF_ERR_INVALID, "sarf_bp_batch_entropy: null argument");
    if (ncand < 1 || nbasis < 0 || nbasis > 16)
        return fail(SARF_ERR_INVALID, "sarf_bp_batch_entropy: ncand=%d nbasis=%d", ncand, nbasis);
    if (d->precision != SARF_BP_MIXED)
        return fail(SARF_ERR_UNSUPPORTED, "sarf_bp_batch_entropy: mixed precision only");
    if (d->flags & ~SARF_BATCH_DIR_PER_BASIS)
        return fail(SARF_ERR_INVALID, "sarf_bp_batch_entropy: flags %d (only "
                    "SARF_BATCH_DIR_PER_BASIS)", d->flags);
    const int per_basis = (d->flags & SARF_BATCH_DIR_PER_BASIS) ? 1 : 0;
    sarf_bp_desc dd = *d;
    dd.flags = 0;
    d = &dd;
    BpArgs a{};
    const int tile = d->tile > 0 ? d->tile : 0;
    SARF_TRY(bp_fill_args(ctx, d, tile, &a));
    int tx = 0, ty = 0;
    const int tid = bp_tile_shape(tile > 0 ? tile : bp_pick_tile(d, a.dx, a.dy), &tx, &ty);
    const int ntiles = a.tiles_x * a.tiles_y;
    const long nitems = (long)ntiles * ncand;
    void *buf = nullptr;
    SARF_TRY(scratch_get(ctx, "batch.partials", sizeof(double2) * (size_t)nitems, &buf));
    double2 *partial = (double2 *)buf;
    // persistent: 4 resident workgroups per CU (LDS-limited), 2 waves of them
    const int cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    const int grid = (int)std::min<long>(nitems, 8L * cus);
    hipStream_t s = (hipStream_t)stream;
    const float2 *p = (const float2 *)prof;
    switch (tid) {
        case 1: launch_batch_t<32, 4>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 2: launch_batch_t<64, 4>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 3: launch_batch_t<64, 8>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 4: launch_batch_t<32, 8>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 5: launch_batch_t<16, 4>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 11: launch_batch_t<32, 4, true>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 12: launch_batch_t<64, 4, true>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 13: launch_batch_t<64, 8, true>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 14: launch_batch_t<32, 8, true>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 15: launch_batch_t<256, 8, true>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        case 16: launch_batch_t<128, 8, true>(a, grid, p, ant, r0, f0, dir, basis, coef, nbasis, ncand, per_basis, partial, s); break;
        default:
            return fail(SARF_ERR_UNSUPPORTED, "sarf_bp_batch_entropy: tile %d not compiled for batches", tid);
    }
    SARF_HIP(hipGetLastError());
    hipLaunchKernelGGL(entropy_from_partials, dim3(ncand), dim3(256), 0, s, partial, ntiles, ncand,
                       entropy);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

extern "C" int sarf_image_entropy(sarf_ctx *ctx, const void *img, long npix, int f64, double *out,
                                  void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!img || !out || npix < 1) return fail(SARF_ERR_INVALID, "sarf_image_entropy: bad argument");
    const int nblocks = (int)std::min<long>((npix + 255) / 256, 1024);
    void *buf = nullptr;
    SARF_TRY(scratch_get(ctx, "entropy.partials", sizeof(double2) * nblocks, &buf));
    hipStream_t s = (hipStream_t)stream;
    if (f64)
        hipLaunchKernelGGL(image_entropy_partials<double2>, dim3(nblocks), dim3(256), 0, s,
                           (const double2 *)img, npix, (double2 *)buf);
    else
        hipLaunchKernelGGL(image_entropy_partials<float2>, dim3(nblocks), dim3(256), 0, s,
                           (const float2 *)img, npix, (double2 *)buf);
    SARF_HIP(hipGetLastError());
    hipLaunchKernelGGL(image_entropy_final, dim3(1), dim3(256), 0, s, (const double2 *)buf, nblocks,
                       out);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

// sarf_output.hip -- 8-bit dB display image of a focused image on the GPU.
//
// Quantisation of the reference's writeBMPFile
// (include/uncc_sar_globals.hpp:66-118), in float with the same operation
// order and no contraction:
//   m      = max over the image of |I| (scaled abs, uncc_complex.hpp:93-102)
//   pixelf = (float)(255.0 / dyn) * (20 * log10(|I| / m) + dyn)
//   byte   = pixelf < 0 ? 0 : (unsigned char)pixelf
// Output rows are top-down as the writer builds them (flipX branch, :85-98):
// row r shows y = Ny-1-r, so out[r*Nx + x] = byte(img[x*Ny + Ny-1-r]).
// The image is read along y (coalesced) and the bytes are transposed through
// LDS so the row-major output is written along x.
#include <algorithm>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

constexpr int kDbThreads = 256;
constexpr int kDbTile = 64;

__device__ __forceinline__ float scaled_abs(float2 v) {
    const float s = fmaxf(fabsf(v.x), fabsf(v.y));
    if (s == 0.0f) return s;
    const float x = __fdiv_rn(v.x, s), y = __fdiv_rn(v.y, s);
    return __fmul_rn(s, __fsqrt_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y))));
}

__global__ __launch_bounds__(kDbThreads) void db8_max_partial(const float2 *__restrict__ img,
                                                              long n, float *__restrict__ part) {
    __shared__ float red[kDbThreads];
    float m = 0.0f;
    for (long i = (long)blockIdx.x * kDbThreads + threadIdx.x; i < n;
         i += (long)gridDim.x * kDbThreads)
        m = fmaxf(m, scaled_abs(img[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = kDbThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kDbThreads) void db8_max_final(float *__restrict__ part, int nb) {
    __shared__ float red[kDbThreads];
    float m = 0.0f;
    for (int i = threadIdx.x; i < nb; i += kDbThreads) m = fmaxf(m, part[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = kDbThreads / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[nb] = red[0];
}

// Workgroup = 64 x 64 pixels (x0.., y0..): 64 lanes along y, 4 rows of x
// per pass; bytes go through LDS and leave along x.
__global__ __launch_bounds__(kDbThreads) void db8_quantize(const float2 *__restrict__ img, int Nx,
                                                           int Ny, const float *__restrict__ mptr,
                                                           float scale, float dyn,
                                                           unsigned char *__restrict__ out) {
    __shared__ unsigned char tile[kDbTile][kDbTile + 4];  // [y_local][x_local]
    const int x0 = blockIdx.x * kDbTile, y0 = blockIdx.y * kDbTile;
    const int ly = threadIdx.x & 63, lx = threadIdx.x >> 6;
    const float m = *mptr;
    const int y = y0 + ly;
    for (int xi = lx; xi < kDbTile; xi += kDbThreads / 64) {
        const int x = x0 + xi;
        unsigned char b = 0;
        if (x < Nx && y < Ny) {
            const float a = scaled_abs(img[(size_t)x * Ny + y]);
            const float l = log10f(__fdiv_rn(a, m));
            const float v = __fmul_rn(scale, __fadd_rn(__fmul_rn(20.0f, l), dyn));
            b = (v < 0.0f || v != v) ? 0 : (unsigned char)fminf(v, 255.0f);
        }
        tile[ly][xi] = b;
    }
    __syncthreads();
    // write: lane along x, rows r = Ny-1-y
    const int wx = threadIdx.x & 63, wy = threadIdx.x >> 6;
    for (int yi = wy; yi < kDbTile; yi += kDbThreads / 64) {
        const int yy = y0 + yi, x = x0 + wx;
        if (yy < Ny && x < Nx) out[(size_t)(Ny - 1 - yy) * Nx + x] = tile[yi][wx];
    }
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_image_db8(sarf_ctx *ctx, const void *img, int Nx, int Ny, float dyn_range_dB,
                              unsigned char *out, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    if (!img || !out) return fail(SARF_ERR_INVALID, "sarf_image_db8: null argument");
    if (Nx < 1 || Ny < 1 || !(dyn_range_dB > 0.0f))
        return fail(SARF_ERR_INVALID, "sarf_image_db8: Nx=%d Ny=%d dyn_range_dB=%g", Nx, Ny,
                    (double)dyn_range_dB);
    hipStream_t s = (hipStream_t)stream;
    const long n = (long)Nx * Ny;
    const int nb = (int)std::min<long>(4L * (ctx->cu_count > 0 ? ctx->cu_count : 256),
                                       (n + kDbThreads - 1) / kDbThreads);
    void *buf = nullptr;
    SARF_TRY(scratch_get(ctx, "db8.max", sizeof(float) * (nb + 1), &buf));
    float *part = (float *)buf;
    hipLaunchKernelGGL(db8_max_partial, dim3(nb), dim3(kDbThreads), 0, s, (const float2 *)img, n,
                       part);
    hipLaunchKernelGGL(db8_max_final, dim3(1), dim3(kDbThreads), 0, s, part, nb);
    const float scale = (float)(255.0 / (double)dyn_range_dB);
    hipLaunchKernelGGL(db8_quantize,
                       dim3((Nx + kDbTile - 1) / kDbTile, (Ny + kDbTile - 1) / kDbTile),
                       dim3(kDbThreads), 0, s, (const float2 *)img, Nx, Ny, part + nb, scale,
                       dyn_range_dB, out);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

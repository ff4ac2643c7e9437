// sarf_output.hip -- 8-bit dB display image of a focused image on the GPU.
//
// Quantisation of the reference's writeBMPFile
// (include/uncc_sar_globals.hpp:66-118), bit for bit:
//   m      = max over the image of |I| (scaled abs, uncc_complex.hpp:93-102)
//   r      = |I| / m
//   pixelf = (float)(255.0 / dyn) * (20 * log10(r) + dyn)   (float, host libm)
//   byte   = pixelf < 0 ? 0 : (unsigned char)pixelf
// The byte is a non-decreasing step function of r, so it is fixed by the
// 255 thresholds t_b = the smallest float r with byte(r) >= b.  They are
// found on the host by bisection over float bit patterns with the same
// float expression and the host's log10 (the function the reference calls),
// and the kernel counts the thresholds at or below r: no device log10 is
// involved, and |I|, m and r are IEEE-exact on both sides.
// Output rows are top-down as the writer builds them (flipX branch, :85-98):
// row r shows y = Ny-1-r, so out[r*Nx + x] = byte(img[x*Ny + Ny-1-r]).
// The image is read along y (coalesced) and the bytes are transposed through
// LDS so the row-major output is written along x.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

constexpr int kDbThreads = 256;
constexpr int kDbTile = 64;

// Correctly rounded float quotient and square root: formed in double and
// rounded to float once.  A double carries 53 bits, more than 2*24 + 2, so
// rounding the (correctly rounded) double result of a float division or
// square root to float gives that float operation's correctly rounded
// result.  (The device's float divide and square root builtins measured one
// ulp off the host's IEEE results on some |I| / max ratios.)
__device__ __forceinline__ float div_rn(float a, float b) {
    return (float)((double)a / (double)b);
}
__device__ __forceinline__ float sqrt_rn(float a) { return (float)sqrt((double)a); }

__device__ __forceinline__ float scaled_abs(float2 v) {
    const float s = fmaxf(fabsf(v.x), fabsf(v.y));
    if (s == 0.0f) return s;
    const float x = div_rn(v.x, s), y = div_rn(v.y, s);
    return __fmul_rn(s, sqrt_rn(__fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y))));
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

// Level thresholds of the quantisation: level[b] (b = 1..255) is the
// smallest r with byte(r) >= b (+inf when no r <= 1 reaches b); level[0] = 0.
struct Db8Levels {
    float t[256];
};

// Workgroup = 64 x 64 pixels (x0.., y0..): 64 lanes along y, 4 rows of x
// per pass; bytes go through LDS and leave along x.
__global__ __launch_bounds__(kDbThreads) void db8_quantize(const float2 *__restrict__ img, int Nx,
                                                           int Ny, const float *__restrict__ mptr,
                                                           Db8Levels lv,
                                                           unsigned char *__restrict__ out) {
    __shared__ unsigned char tile[kDbTile][kDbTile + 4];  // [y_local][x_local]
    __shared__ float level[256];
    level[threadIdx.x] = lv.t[threadIdx.x];
    __syncthreads();
    const int x0 = blockIdx.x * kDbTile, y0 = blockIdx.y * kDbTile;
    const int ly = threadIdx.x & 63, lx = threadIdx.x >> 6;
    const float m = *mptr;
    const int y = y0 + ly;
    for (int xi = lx; xi < kDbTile; xi += kDbThreads / 64) {
        const int x = x0 + xi;
        unsigned char b = 0;
        if (x < Nx && y < Ny) {
            const float r = div_rn(scaled_abs(img[(size_t)x * Ny + y]), m);
            if (r == r) {  // 0/0 (an all-zero image) quantises to 0
                // count of thresholds <= r among level[1..255] (non-decreasing)
                int lo = 0;
#pragma unroll
                for (int step = 128; step > 0; step >>= 1)
                    if (lo + step <= 255 && level[lo + step] <= r) lo += step;
                b = (unsigned char)lo;
            }
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

// byte(r) of the reference's expression, in float with the host libm and no
// contraction (each product and sum rounded to float, as the reference's
// x86-64 build evaluates it).
#pragma clang fp contract(off)
static int db8_byte_host(float r, float dyn) {
    const float scale = (float)(255.0 / (double)dyn);
    volatile float l = std::log10(r);
    volatile float t = 20 * l;
    volatile float u = t + dyn;
    volatile float v = scale * u;
    const float w = v;
    if (w < 0.0f || w != w) return 0;
    return w >= 255.0f ? 255 : (int)(unsigned char)w;
}

static float bits_to_float(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Thresholds by bisection over the bit patterns of r in [0, 1] (positive
// floats are ordered as their bit patterns).
static void db8_levels_host(float dyn, float *t) {
    const uint32_t one = 0x3f800000u;
    t[0] = 0.0f;
    for (int b = 1; b < 256; b++) {
        if (db8_byte_host(1.0f, dyn) < b) {
            t[b] = INFINITY;
            continue;
        }
        uint32_t lo = 0, hi = one;  // byte(lo) < b <= byte(hi) (byte(0) = 0)
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (db8_byte_host(bits_to_float(mid), dyn) >= b) hi = mid;
            else lo = mid;
        }
        t[b] = bits_to_float(hi);
    }
}

}  // namespace
}  // namespace sarf

using namespace sarf;

extern "C" int sarf_image_db8(sarf_ctx *ctx, const void *img, int Nx, int Ny, float dyn_range_dB,
                              unsigned char *out, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
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
    Db8Levels lv;
    db8_levels_host(dyn_range_dB, lv.t);
    hipLaunchKernelGGL(db8_quantize,
                       dim3((Nx + kDbTile - 1) / kDbTile, (Ny + kDbTile - 1) / kDbTile),
                       dim3(kDbThreads), 0, s, (const float2 *)img, Nx, Ny, part + nb, lv, out);
    SARF_HIP(hipGetLastError());
    return SARF_OK;
}

extern "C" int sarf_db8_levels(float dyn_range_dB, float *levels) {
    clear_error();
    if (!levels) return fail(SARF_ERR_INVALID, "sarf_db8_levels: null argument");
    if (!(dyn_range_dB > 0.0f) || !std::isfinite(dyn_range_dB))
        return fail(SARF_ERR_INVALID, "sarf_db8_levels: dyn_range_dB=%g", (double)dyn_range_dB);
    db8_levels_host(dyn_range_dB, levels);
    return SARF_OK;
}

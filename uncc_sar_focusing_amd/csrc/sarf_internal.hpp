// sarf_internal.hpp -- shared host/device declarations of libsarf (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <array>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "sarf.h"

namespace sarf {

constexpr double kClight = 299792458.0;   // uncc_sar_focusing.hpp:34
constexpr double kPi = 3.141592653589793; // uncc_sar_focusing.hpp:36
// Candidates per level of sarf_grid_search.
constexpr long kGridMaxCandidates = 1L << 21;
// Key offset of the PGA Bluestein tables in sarf_ctx::rc_blue.
constexpr long kPgaTableKey = 1L << 40;

// Records a printf-style message as the calling thread's last error and
// returns `code`, so call sites read `return fail(SARF_ERR_INVALID, ...)`.
int fail(int code, const char *fmt, ...);
void clear_error();

#define SARF_HIP(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::sarf::fail(SARF_ERR_HIP, "%s: %s (%s:%d)", #expr,                   \
                                hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)

#define SARF_TRY(expr)                    \
    do {                                  \
        int rc_ = (expr);                 \
        if (rc_ != SARF_OK) return rc_;   \
    } while (0)

// Device scratch that grows on demand and is reused between calls.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
};

// Parallel memcpy between pageable and pinned host memory on a few
// persistent threads (defined with sarf_focus).
class HostCopier;

}  // namespace sarf

struct sarf_ctx {
    int device = 0;
    int cu_count = 0;
    std::mutex mu;
    // DFT twiddle tables e^{+j 2 pi i / N}, keyed by N
    std::map<long, float2 *> tw32;
    std::map<long, double2 *> tw64;
    // Bluestein tables, one allocation per key: range compression of
    // length N in FP32 (key N) or FP64 (key -N); PGA rows of length N
    // (key sarf::kPgaTableKey + N)
    std::map<long, void *> rc_blue;
    // named scratch buffers (partial images, reductions, staging); a buffer
    // that has to grow is retired, not freed, so work enqueued earlier --
    // on any stream, or in a captured graph -- never sees freed memory
    std::map<std::string, sarf::Scratch> scratch;
    std::vector<void *> retired;
    // device state of the lists uploaded by the batched entry points, one
    // allocation per distinct list (entry-point tag + the caller's
    // descriptors), never overwritten: a graph captured with a list keeps
    // valid pointers until sarf_ctx_trim or sarf_ctx_destroy
    std::map<std::vector<unsigned char>, sarf::Scratch> lists;
    // ordering of this context's calls across streams: the end of the last
    // call, recorded on its stream
    hipEvent_t last_ev = nullptr;
    hipStream_t last_stream = nullptr;
    bool last_valid = false;
    // backprojection launch shapes chosen by sarf_autotune, keyed by the
    // image geometry: (tile id, pulse splits)
    std::map<std::array<uint64_t, 8>, sarf_tune_entry> tuned;
    // host-buffer path: pinned staging (grown on demand) and its copy threads
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    sarf::HostCopier *copier = nullptr;
};

namespace sarf {

// Makes ctx->device current.
int use_device(sarf_ctx *ctx);
// Returns a device buffer of at least `bytes` bytes under `name`.
int scratch_get(sarf_ctx *ctx, const char *name, size_t bytes, void **out);
// The device buffer dedicated to one list of a batched entry point: `tag`
// names the entry point, `key` the caller's list (and parameters that shape
// the device state).  A list seen before returns its buffer with *fresh =
// false (nothing to upload, nothing to wait for); a new one is allocated
// (*fresh = true, the caller uploads), which is refused during capture.
int list_buffer(sarf_ctx *ctx, char tag, const std::vector<unsigned char> &key, size_t bytes,
                hipStream_t s, const char *who, void **buf, bool *fresh);

// Orders the calls of one context across streams.  The constructor makes
// stream `s` wait for the end of the context's previous call if that was on
// another stream; the destructor records the end of this call on `s`.  Both
// are skipped while `s` is being captured (a graph orders its own nodes).
struct StreamOrder {
    sarf_ctx *ctx;
    hipStream_t s;
    bool active;
    StreamOrder(sarf_ctx *c, hipStream_t st);
    ~StreamOrder();
    StreamOrder(const StreamOrder &) = delete;
    StreamOrder &operator=(const StreamOrder &) = delete;
};
// Twiddle table for N = 2^log2n (created on first use, synchronously).
int twiddles_f32(sarf_ctx *ctx, int log2n, const float2 **out);
int twiddles_f64(sarf_ctx *ctx, int log2n, const double2 **out);
int twiddles_any_f32(sarf_ctx *ctx, long n, const float2 **out);
int twiddles_any_f64(sarf_ctx *ctx, long n, const double2 **out);

int ilog2_exact(long n);  // log2 if n is a power of two, else -1

// Bluestein chirp and kernel spectrum of a length-N DFT with transform
// length M = 2^L >= 2N - 1, in FP64: c_k = e^{+j pi k^2 / N} (k < N) and
// C = FFT_M(h), h_d = conj(c_d) for d < N, h_{M-d} = conj(c_d) for 0 < d < N.
void bluestein_f64(int N, int L, std::vector<double> &cr, std::vector<double> &ci,
                   std::vector<double> &hr, std::vector<double> &hi);

// Internal launchers shared between translation units.
int launch_range_compress(sarf_ctx *ctx, const sarf_rc_desc *d, const void *ph,
                          const float *taper, const double *f0, const double *r0,
                          void *prof, hipStream_t stream);
int launch_backproject(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof,
                       const double *ant, const double *r0, const double *f0, void *img,
                       hipStream_t stream);
int launch_backproject_ext(sarf_ctx *ctx, const sarf_bp_desc *d, const void *prof,
                           const double *ant, const double *r0, const double *f0, void *img,
                           hipStream_t stream, float2 *ext_part, size_t ext_bytes);
// Records the current device's address of the allocation-free entry's scratch.
int bp_noalloc_init(int device);
// img[i] (+)= sum over s in order of part[s*stride + i], i < n (deterministic).
int launch_reduce_partials(const float2 *part, long stride, int splits, long n, int accumulate,
                           float2 *img, hipStream_t stream);

}  // namespace sarf

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// sarf_api.hip -- context, error reporting and the C entry points of libsarf.
//
// Replaces the reference's GPUMemoryManager / initialize_* / finalize_*
// host plumbing (src/gpuBackProjection/cuda_sar_focusing/cuda_sar_focusing.hpp:
// 93-519) with an explicit context: device buffers are caller-owned, work is
// stream-ordered, errors are returned as codes (never printed), and the device
// is never reset.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "sarf_hostcopy.hpp"
#include "sarf_internal.hpp"

namespace sarf {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

int ilog2_exact(long n) {
    if (n <= 0 || (n & (n - 1)) != 0) return -1;
    int l = 0;
    while ((1L << l) < n) ++l;
    return l;
}

int use_device(sarf_ctx *ctx) {
    if (!ctx) return fail(SARF_ERR_INVALID, "null context");
    SARF_HIP(hipSetDevice(ctx->device));
    return SARF_OK;
}

int scratch_get(sarf_ctx *ctx, const char *name, size_t bytes, void **out) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    Scratch &s = ctx->scratch[name];
    if (s.bytes < bytes) {
        if (s.ptr) {
            // Work enqueued earlier (on any stream, or in a captured graph)
            // may still use the old buffer: keep it until sarf_ctx_trim or
            // sarf_ctx_destroy instead of freeing it here.
            ctx->retired.push_back(s.ptr);
            s.ptr = nullptr;
            s.bytes = 0;
        }
        hipError_t e = hipMalloc(&s.ptr, bytes);
        if (e != hipSuccess) {
            s.ptr = nullptr;
            return fail(SARF_ERR_NOMEM, "scratch '%s' of %zu bytes: %s", name, bytes,
                        hipGetErrorString(e));
        }
        s.bytes = bytes;
    }
    *out = s.ptr;
    return SARF_OK;
}

int list_buffer(sarf_ctx *ctx, char tag, const std::vector<unsigned char> &key, size_t bytes,
                hipStream_t s, const char *who, void **buf, bool *fresh) {
    std::vector<unsigned char> k;
    k.reserve(key.size() + 1);
    k.push_back((unsigned char)tag);
    k.insert(k.end(), key.begin(), key.end());
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->lists.find(k);
    if (it != ctx->lists.end() && it->second.bytes >= bytes) {
        *buf = it->second.ptr;
        *fresh = false;
        return SARF_OK;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SARF_HIP(hipStreamIsCapturing(s, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(SARF_ERR_INVALID, "%s: a new list cannot be uploaded during stream capture; "
                    "make the same call once before capturing", who);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        return fail(SARF_ERR_NOMEM, "%s: list state of %zu bytes: %s", who, bytes,
                    hipGetErrorString(e));
    if (it != ctx->lists.end()) {
        ctx->retired.push_back(it->second.ptr);
        it->second = Scratch{p, bytes};
    } else {
        ctx->lists.emplace(std::move(k), Scratch{p, bytes});
    }
    *buf = p;
    *fresh = true;
    return SARF_OK;
}

static bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone;
}

StreamOrder::StreamOrder(sarf_ctx *c, hipStream_t st) : ctx(c), s(st), active(false) {
    if (!ctx || !ctx->last_ev || capturing(s)) return;
    active = true;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->last_valid && ctx->last_stream != s) (void)hipStreamWaitEvent(s, ctx->last_ev, 0);
}

StreamOrder::~StreamOrder() {
    if (!active) return;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (hipEventRecord(ctx->last_ev, s) == hipSuccess) {
        ctx->last_stream = s;
        ctx->last_valid = true;
    }
}

template <typename T2>
static int make_twiddles(long n, T2 **out) {
    std::vector<T2> h(n);
    for (long i = 0; i < n; i++) {
        // e^{+j 2 pi i / N}, evaluated in double and rounded once
        double a = 2.0 * kPi * (double)i / (double)n;
        h[i].x = (decltype(h[i].x))std::cos(a);
        h[i].y = (decltype(h[i].y))std::sin(a);
    }
    T2 *d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(T2) * n);
    if (e != hipSuccess) return fail(SARF_ERR_NOMEM, "twiddle table: %s", hipGetErrorString(e));
    SARF_HIP(hipMemcpy(d, h.data(), sizeof(T2) * n, hipMemcpyHostToDevice));
    *out = d;
    return SARF_OK;
}

int twiddles_any_f32(sarf_ctx *ctx, long n, const float2 **out) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tw32.find(n);
    if (it == ctx->tw32.end()) {
        float2 *d = nullptr;
        SARF_TRY(make_twiddles(n, &d));
        it = ctx->tw32.emplace(n, d).first;
    }
    *out = it->second;
    return SARF_OK;
}

int twiddles_any_f64(sarf_ctx *ctx, long n, const double2 **out) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->tw64.find(n);
    if (it == ctx->tw64.end()) {
        double2 *d = nullptr;
        SARF_TRY(make_twiddles(n, &d));
        it = ctx->tw64.emplace(n, d).first;
    }
    *out = it->second;
    return SARF_OK;
}

int twiddles_f32(sarf_ctx *ctx, int log2n, const float2 **out) {
    return twiddles_any_f32(ctx, 1L << log2n, out);
}

int twiddles_f64(sarf_ctx *ctx, int log2n, const double2 **out) {
    return twiddles_any_f64(ctx, 1L << log2n, out);
}

// True if the host range [p, p + bytes) is page-locked (hipHostMalloc'd or
// registered), so the GPU can copy from and to it directly.  For pageable
// memory hipPointerGetAttributes fails; its error is cleared here.
static bool host_pinned(const void *p, size_t bytes) {
    for (const char *q : {(const char *)p, (const char *)p + (bytes ? bytes - 1 : 0)}) {
        hipPointerAttribute_t a{};
        if (hipPointerGetAttributes(&a, q) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        if (a.type != hipMemoryTypeHost) return false;
    }
    return true;
}

// Pinned staging of at least `bytes` bytes for the host-buffer path.
int pinned_get(sarf_ctx *ctx, size_t bytes, char **out) {
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) (void)hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
        SARF_HIP(hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault));
        ctx->pinned_bytes = bytes;
    }
    if (!ctx->copier) {
        const unsigned hw = std::thread::hardware_concurrency();
        ctx->copier = new HostCopier((int)std::max(1u, std::min(8u, hw ? hw : 1u)));
    }
    *out = (char *)ctx->pinned;
    return SARF_OK;
}

}  // namespace sarf

using namespace sarf;

extern "C" {

int sarf_version(void) { return SARF_VERSION; }

const char *sarf_last_error(void) { return g_last_error.c_str(); }

int sarf_ctx_create(int device, sarf_ctx **out) {
    clear_error();
    if (!out) return fail(SARF_ERR_INVALID, "sarf_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return fail(SARF_ERR_NODEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return fail(SARF_ERR_NODEVICE, "device %d out of range [0,%d)", device, count);
    hipDeviceProp_t prop;
    SARF_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SARF_ERR_NODEVICE, "device %d is %s; libsarf is built for gfx950 only",
                    device, prop.gcnArchName);
    SARF_HIP(hipSetDevice(device));
    SARF_TRY(bp_noalloc_init(device));
    hipEvent_t ev = nullptr;
    SARF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    sarf_ctx *ctx = new sarf_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    ctx->last_ev = ev;
    *out = ctx;
    return SARF_OK;
}

// Frees what the context keeps only so that earlier work stays valid: the
// retired scratch buffers and the device state of every uploaded list.
static void trim(sarf_ctx *ctx) {
    for (void *p : ctx->retired) (void)hipFree(p);
    ctx->retired.clear();
    for (auto &kv : ctx->lists)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    ctx->lists.clear();
}

int sarf_ctx_trim(sarf_ctx *ctx) {
    clear_error();
    SARF_TRY(use_device(ctx));
    SARF_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lk(ctx->mu);
    trim(ctx);
    return SARF_OK;
}

int sarf_ctx_destroy(sarf_ctx *ctx) {
    clear_error();
    if (!ctx) return SARF_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : ctx->tw32) (void)hipFree(kv.second);
    for (auto &kv : ctx->tw64) (void)hipFree(kv.second);
    for (auto &kv : ctx->rc_blue) (void)hipFree(kv.second);
    trim(ctx);
    if (ctx->last_ev) (void)hipEventDestroy(ctx->last_ev);
    for (auto &kv : ctx->scratch)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    delete ctx->copier;
    delete ctx;
    return SARF_OK;
}

int sarf_range_compress(sarf_ctx *ctx, const sarf_rc_desc *desc, const void *ph,
                        const float *taper, const double *f0, const double *r0, void *prof,
                        void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!desc || !ph || !prof) return fail(SARF_ERR_INVALID, "sarf_range_compress: null argument");
    return launch_range_compress(ctx, desc, ph, taper, f0, r0, prof, (hipStream_t)stream);
}

int sarf_backproject(sarf_ctx *ctx, const sarf_bp_desc *desc, const void *prof,
                     const double *ant, const double *r0, const double *f0, void *img,
                     void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!desc || !prof || !ant || !r0 || !f0 || !img)
        return fail(SARF_ERR_INVALID, "sarf_backproject: null argument");
    return launch_backproject(ctx, desc, prof, ant, r0, f0, img, (hipStream_t)stream);
}

int sarf_focus(sarf_ctx *ctx, const sarf_focus_desc *d, const float *ph, const double *ant,
               const double *r0, const double *f0, float *img) {
    clear_error();
    SARF_TRY(use_device(ctx));
    {
        // the context's scratch may still be in use by its last call on a
        // caller's stream: wait for that call before reusing it
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->last_valid) SARF_HIP(hipEventSynchronize(ctx->last_ev));
    }
    if (!d || !ph || !ant || !r0 || !f0 || !img)
        return fail(SARF_ERR_INVALID, "sarf_focus: null argument");
    if (d->K <= 0 || d->Np <= 0 || d->Nx < 2 || d->Ny < 2 || d->Nfft < d->K)
        return fail(SARF_ERR_INVALID, "sarf_focus: bad sizes K=%d Np=%d Nfft=%d Nx=%d Ny=%d",
                    d->K, d->Np, d->Nfft, d->Nx, d->Ny);
    const bool f64 = d->precision == SARF_BP_F64;
    const size_t csz = f64 ? 16 : 8;
    const size_t ph_elems = (size_t)d->K * d->Np, prof_elems = (size_t)d->Nfft * d->Np;
    const size_t img_elems = (size_t)d->Nx * d->Ny;
    void *dph, *dprof, *dimg, *dgeo;
    SARF_TRY(scratch_get(ctx, "focus.ph", ph_elems * csz, &dph));
    SARF_TRY(scratch_get(ctx, "focus.prof", prof_elems * csz, &dprof));
    SARF_TRY(scratch_get(ctx, "focus.img", img_elems * csz, &dimg));
    SARF_TRY(scratch_get(ctx, "focus.geo", sizeof(double) * 5 * d->Np, &dgeo));
    double *dant = (double *)dgeo, *dr0 = dant + 3 * d->Np, *df0 = dr0 + d->Np;
    SARF_HIP(hipMemcpy(dant, ant, sizeof(double) * 3 * d->Np, hipMemcpyHostToDevice));
    SARF_HIP(hipMemcpy(dr0, r0, sizeof(double) * d->Np, hipMemcpyHostToDevice));
    SARF_HIP(hipMemcpy(df0, f0, sizeof(double) * d->Np, hipMemcpyHostToDevice));
    sarf_bp_desc bp{};
    bp.Nx = d->Nx; bp.Ny = d->Ny; bp.Nfft = d->Nfft; bp.Np = d->Np;
    bp.x0 = d->x0; bp.y0 = d->y0; bp.Wx = d->Wx; bp.Wy = d->Wy; bp.maxWr = d->maxWr;
    bp.precision = d->precision;
    {
        // line-of-sight hint from the mean phase centre (tile choice only)
        double mx = 0, my = 0, mz = 0;
        for (int n = 0; n < d->Np; n++) {
            mx += ant[3 * n];
            my += ant[3 * n + 1];
            mz += ant[3 * n + 2];
        }
        mx = mx / d->Np - d->x0;
        my = my / d->Np - d->y0;
        mz /= d->Np;
        const double nrm = std::sqrt(mx * mx + my * my + mz * mz);
        if (nrm > 0) {
            bp.los_x = mx / nrm;
            bp.los_y = my / nrm;
        }
    }
    if (f64) {
        // validation path: whole transfers, then RC and BP
        std::vector<double> tmp(2 * ph_elems);
        for (size_t i = 0; i < 2 * ph_elems; i++) tmp[i] = ph[i];
        SARF_HIP(hipMemcpy(dph, tmp.data(), ph_elems * csz, hipMemcpyHostToDevice));
        sarf_rc_desc rc{d->K, d->Np, d->Nfft, SARF_RC_F64};
        SARF_TRY(launch_range_compress(ctx, &rc, dph, nullptr, nullptr, nullptr, dprof, nullptr));
        SARF_TRY(launch_backproject(ctx, &bp, dprof, dant, dr0, df0, dimg, nullptr));
        SARF_HIP(hipStreamSynchronize(nullptr));
    } else {
        // Pulse groups: the caller's pageable samples of group g + 1 are
        // copied into pinned staging (several threads) and sent to the GPU
        // while group g is range-compressed and backprojected; the image comes
        // back in chunks, each copied out of pinned staging while the next is
        // in flight.  Backprojection is additive over pulses, so each group
        // accumulates into the image.  Page-locked caller buffers skip the
        // staging: the copies go straight from and to them.
        // The first group's upload is not overlapped with anything, so it is
        // kept short (an eighth of the pulses); the rest split evenly.
        const int Gr = std::max(1, std::min(4, d->Np / 1024));
        const int G = Gr > 1 ? Gr + 1 : 1;
        const int first = G > 1 ? d->Np / 8 : d->Np;
        auto group_start = [&](int g) {
            if (g == 0) return 0;
            if (g >= G) return d->Np;
            return first + (int)((long)(d->Np - first) * (g - 1) / (G - 1));
        };
        size_t gmax = 0;
        for (int g = 0; g < G; g++) {
            const int n0 = group_start(g), n1 = group_start(g + 1);
            gmax = std::max(gmax, sizeof(float2) * (size_t)(n1 - n0) * d->K);
        }
        const size_t slot = (gmax + 255) & ~(size_t)255, img_bytes = img_elems * csz;
        const bool ph_direct = host_pinned(ph, ph_elems * csz);
        const bool img_direct = host_pinned(img, img_bytes);
        char *pin = nullptr;
        const size_t stage = (ph_direct ? 0 : 2 * slot) + (img_direct ? 0 : img_bytes);
        if (stage > 0) SARF_TRY(pinned_get(ctx, stage, &pin));
        char *pin_ph[2] = {nullptr, nullptr}, *pin_img = pin;
        if (!ph_direct) {
            pin_ph[0] = pin;
            pin_ph[1] = pin + slot;
            pin_img = pin + 2 * slot;
        }
        constexpr int kChunks = 4;
        hipStream_t sc = nullptr, s = nullptr;
        SARF_HIP(hipStreamCreateWithFlags(&sc, hipStreamNonBlocking));
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            (void)hipStreamDestroy(sc);
            return fail(SARF_ERR_HIP, "sarf_focus: stream creation failed");
        }
        hipEvent_t copied[2] = {nullptr, nullptr}, dl[kChunks] = {};
        int rc = SARF_OK;
        for (hipEvent_t *e : {&copied[0], &copied[1]})
            if (rc == SARF_OK && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
                rc = fail(SARF_ERR_HIP, "sarf_focus: event creation failed");
        for (int c = 0; c < kChunks && rc == SARF_OK; c++)
            if (hipEventCreateWithFlags(&dl[c], hipEventDisableTiming) != hipSuccess)
                rc = fail(SARF_ERR_HIP, "sarf_focus: event creation failed");
        for (int g = 0; g < G && rc == SARF_OK; g++) {
            const int n0 = group_start(g), n1 = group_start(g + 1);
            const size_t off = (size_t)n0 * d->K, bytes = sizeof(float2) * (size_t)(n1 - n0) * d->K;
            const int k = g & 1;
            const void *src = (const float2 *)ph + off;
            if (!ph_direct) {
                // staging slot k was last sent for group g - 2
                if (g >= 2 && hipEventSynchronize(copied[k]) != hipSuccess) {
                    rc = fail(SARF_ERR_HIP, "sarf_focus: upload of group %d failed", g - 2);
                    break;
                }
                ctx->copier->copy(pin_ph[k], src, bytes);
                src = pin_ph[k];
            }
            if (hipMemcpyAsync((float2 *)dph + off, src, bytes, hipMemcpyHostToDevice, sc) !=
                    hipSuccess ||
                hipEventRecord(copied[k], sc) != hipSuccess ||
                hipStreamWaitEvent(s, copied[k], 0) != hipSuccess) {
                rc = fail(SARF_ERR_HIP, "sarf_focus: upload of pulses %d..%d failed", n0, n1);
                break;
            }
            sarf_rc_desc rcd{d->K, n1 - n0, d->Nfft, 0};
            float2 *gprof = (float2 *)dprof + (size_t)n0 * d->Nfft;
            rc = launch_range_compress(ctx, &rcd, (float2 *)dph + off, nullptr, nullptr, nullptr,
                                       gprof, s);
            if (rc != SARF_OK) break;
            sarf_bp_desc bg = bp;
            bg.Np = n1 - n0;
            bg.flags = g > 0 ? SARF_BP_ACCUMULATE : 0;
            rc = launch_backproject(ctx, &bg, gprof, dant + 3 * (size_t)n0, dr0 + n0, df0 + n0,
                                    dimg, s);
        }
        char *dst_img = img_direct ? (char *)img : pin_img;
        for (int c = 0; c < kChunks && rc == SARF_OK; c++) {
            const size_t b = img_bytes * c / kChunks, e = img_bytes * (c + 1) / kChunks;
            if (hipMemcpyAsync(dst_img + b, (char *)dimg + b, e - b, hipMemcpyDeviceToHost, s) !=
                    hipSuccess ||
                hipEventRecord(dl[c], s) != hipSuccess)
                rc = fail(SARF_ERR_HIP, "sarf_focus: image download failed");
        }
        for (int c = 0; c < kChunks && rc == SARF_OK; c++) {
            const size_t b = img_bytes * c / kChunks, e = img_bytes * (c + 1) / kChunks;
            if (hipEventSynchronize(dl[c]) != hipSuccess) {
                rc = fail(SARF_ERR_HIP, "sarf_focus: %s", hipGetErrorString(hipGetLastError()));
                break;
            }
            if (!img_direct) ctx->copier->copy((char *)img + b, pin_img + b, e - b);
        }
        (void)hipStreamSynchronize(s);
        (void)hipStreamSynchronize(sc);
        for (hipEvent_t e : copied)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : dl)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(s);
        (void)hipStreamDestroy(sc);
        if (rc != SARF_OK) return rc;
    }
    if (f64) {
        std::vector<double> tmp(2 * img_elems);
        SARF_HIP(hipMemcpy(tmp.data(), dimg, img_elems * csz, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < 2 * img_elems; i++) img[i] = (float)tmp[i];
    }
    return SARF_OK;
}

int sarf_focus_mf(sarf_ctx *ctx, const sarf_mf_desc *d, const float *ph, const double *ant,
                  const double *r0, const double *f0, const double *df, float *img) {
    clear_error();
    SARF_TRY(use_device(ctx));
    {
        // the context's scratch may still be in use by its last call on a
        // caller's stream: wait for that call before reusing it
        std::lock_guard<std::mutex> lk(ctx->mu);
        if (ctx->last_valid) SARF_HIP(hipEventSynchronize(ctx->last_ev));
    }
    if (!d || !ph || !ant || !r0 || !f0 || !df || !img)
        return fail(SARF_ERR_INVALID, "sarf_focus_mf: null argument");
    if (d->K < 1 || d->Np < 1 || d->Nx < 2 || d->Ny < 2)
        return fail(SARF_ERR_INVALID, "sarf_focus_mf: bad sizes K=%d Np=%d Nx=%d Ny=%d", d->K,
                    d->Np, d->Nx, d->Ny);
    const size_t ph_bytes = sizeof(float2) * (size_t)d->K * d->Np;
    const size_t img_bytes = sizeof(float2) * (size_t)d->Nx * d->Ny;
    void *dph, *dimg, *dgeo;
    SARF_TRY(scratch_get(ctx, "focus.ph", ph_bytes, &dph));
    SARF_TRY(scratch_get(ctx, "focus.img", img_bytes, &dimg));
    SARF_TRY(scratch_get(ctx, "focus.geo", sizeof(double) * 6 * d->Np, &dgeo));
    double *dant = (double *)dgeo, *dr0 = dant + 3 * d->Np, *df0 = dr0 + d->Np, *ddf = df0 + d->Np;
    SARF_HIP(hipMemcpy(dph, ph, ph_bytes, hipMemcpyHostToDevice));
    SARF_HIP(hipMemcpy(dant, ant, sizeof(double) * 3 * d->Np, hipMemcpyHostToDevice));
    SARF_HIP(hipMemcpy(dr0, r0, sizeof(double) * d->Np, hipMemcpyHostToDevice));
    SARF_HIP(hipMemcpy(df0, f0, sizeof(double) * d->Np, hipMemcpyHostToDevice));
    SARF_HIP(hipMemcpy(ddf, df, sizeof(double) * d->Np, hipMemcpyHostToDevice));
    sarf_mf_desc md = *d;
    if (md.flags & SARF_BP_ACCUMULATE) SARF_HIP(hipMemcpy(dimg, img, img_bytes, hipMemcpyHostToDevice));
    SARF_TRY(sarf_matched_filter(ctx, &md, dph, dant, dr0, df0, ddf, dimg, nullptr));
    SARF_HIP(hipMemcpy(img, dimg, img_bytes, hipMemcpyDeviceToHost));
    return SARF_OK;
}

}  // extern "C"

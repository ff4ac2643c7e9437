/*
 * sarf.h -- C ABI of the MI355X (gfx950) SAR focusing engine (libsarf.so).
 *
 * Plain C: no C++ types cross this boundary, only pointers, sizes and small
 * POD descriptors.  Every entry point returns SARF_OK (0) on success or a
 * negative SARF_ERR_* code; the message of the last failure on the calling
 * thread is returned by sarf_last_error().  No entry point prints, exits or
 * resets the device.
 *
 * Memory: "device" pointers are caller-owned HIP device allocations; work is
 * enqueued on the caller's stream (hipStream_t passed as void*, NULL = the
 * null stream) and is complete only after the caller synchronises that
 * stream.  "host" pointers are ordinary host memory; those entry points
 * synchronise before returning.
 *
 * Layouts (complex = interleaved re,im):
 *   phase history  ph[n*K + k]        K frequency samples per pulse, pulse n
 *   range profile  prof[n*Nfft + m]   bin m <-> range r_m = (m/Nfft - 0.5)*maxWr
 *   antenna        ant[3*n + {0,1,2}] phase-centre x,y,z of pulse n (m), FP64
 *   image          img[x*Ny + y]      Ny x Nx column-major (y fastest), pixel
 *                                     (x0 - Wx/2 + x*Wx/(Nx-1), y0 - Wy/2 + y*Wy/(Ny-1), 0)
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository, charlotte-visionlab/uncc_sar_focusing).
 */
#ifndef SARF_H
#define SARF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SARF_VERSION 101

/* status codes */
#define SARF_OK 0
#define SARF_ERR_INVALID (-1)     /* bad argument or descriptor */
#define SARF_ERR_UNSUPPORTED (-2) /* valid but not supported (e.g. Nfft > 16384) */
#define SARF_ERR_HIP (-3)         /* HIP runtime error */
#define SARF_ERR_NOMEM (-4)       /* device allocation failed */
#define SARF_ERR_NODEVICE (-5)    /* no gfx950 device / bad device index */

/* precision of backprojection */
#define SARF_BP_MIXED 0 /* FP64 per (tile,pulse) geometry, FP32 per pixel (default) */
#define SARF_BP_F64 1   /* all-FP64 per pixel and pulse, FP64 profiles and image */

/* backprojection flags */
#define SARF_BP_ACCUMULATE 1 /* add into img instead of overwriting it */
#define SARF_BATCH_DIR_PER_BASIS 2 /* sarf_bp_batch_entropy: dir holds one field per basis */
#define SARF_SCENES_TILE_MIXED 100 /* sarf_bp_scenes: a tile per scene (see there) */

/* range-compression flags */
#define SARF_RC_TAPER 1 /* multiply sample k by taper[k] (rangeCompress.m:6-12) */
#define SARF_RC_RAMP 2  /* multiply pulse n by exp(-j 4 pi f0[n] R0[n] / c) (rangeCompress.m:17) */
#define SARF_RC_F64 4   /* FP64 in/out (ph: double2, prof: double2) */

typedef struct sarf_ctx sarf_ctx;

/* Library version (SARF_VERSION of the build). */
int sarf_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char *sarf_last_error(void);

/* Create a context bound to HIP device `device` (must be gfx950).
 * Replaces initialize_GPUMATLAB / initialize_CUDAResources
 * (src/gpuBackProjection/cuda_sar_focusing/cuda_sar_focusing.hpp:403-478),
 * without the per-call cudaDeviceReset of finalize_CUDAResources (:482-519).
 * A context may be used by one host thread at a time.  Its calls may go to
 * different streams: a call on a stream other than the previous call's waits
 * (on the device) for that call, since calls share the context's scratch;
 * sarf_focus / sarf_focus_mf wait on the host.  Not while a stream is being
 * captured: a caller that captures work of one context on several streams
 * orders them itself.  Scratch that has to grow is kept, not freed, until
 * sarf_ctx_trim or sarf_ctx_destroy, so work enqueued earlier (or a captured
 * graph) never refers to freed memory.  Tables a transform length needs
 * (twiddles, Bluestein chirps and their spectra) are built on its first use
 * in the context, on the host with a synchronous upload, and kept until
 * sarf_ctx_destroy: capture a call into a HIP graph after one eager call
 * with the same sizes. */
int sarf_ctx_create(int device, sarf_ctx **out);
int sarf_ctx_destroy(sarf_ctx *ctx);

/* Waits for the device, then frees what the context keeps only for work    */
/* enqueued earlier: scratch buffers replaced by larger ones and the device */
/* state of every scene / image list uploaded by sarf_range_compress_scenes,*/
/* sarf_bp_scenes and sarf_pga_scenes.  HIP graphs captured before the call */
/* must not be replayed afterwards; the next batched call uploads again.     */
/* No reference counterpart.                                                 */
int sarf_ctx_trim(sarf_ctx *ctx);

/* ------------------------------------------------------------------------ */
/* Range compression: prof[n] = fftshift(ifft(ph[n] zero-padded to Nfft))   */
/* scaled by 1/Nfft.  Replaces cuifft + cufftNormalize_1DBatch +            */
/* cufftShift_1DBatch (cuda_sar_focusing.hpp:372-399,616-621;               */
/* cuda_BackProjectionKernels.cuh:95-135) and the per-pulse ifftw/fftshift  */
/* of run_bp (src/cpuBackProjection/cpuBackProjection.hpp:152-158), with    */
/* the zero padding of matlab/bpBasic.m:82.                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    int K;     /* frequency samples per pulse, 1 <= K <= Nfft */
    int Np;    /* pulses, >= 1 */
    int Nfft;  /* any length >= K: powers of two up to 2^24 (up to 16384 in one LDS
                  pass, longer through global-memory passes); other lengths up to 2^20
                  by Bluestein's identity (up to 8192 register-resident in one kernel,
                  longer ones and FP64 through the global-memory passes) */
    int flags; /* SARF_RC_* */
} sarf_rc_desc;

int sarf_range_compress(sarf_ctx *ctx, const sarf_rc_desc *desc,
                        const void *ph,       /* device, complex K x Np */
                        const float *taper,   /* device, K, or NULL */
                        const double *f0,     /* device, Np (SARF_RC_RAMP) or NULL */
                        const double *r0,     /* device, Np (SARF_RC_RAMP) or NULL */
                        void *prof,           /* device, complex Nfft x Np */
                        void *stream);

/* Range compression of many independent scenes (FP32), each as one         */
/* sarf_range_compress call: the scenes of one Nfft share one persistent    */
/* launch over all their pulses.  Results equal per-scene calls bit for bit.*/
/* Each distinct scene list is uploaded once (the call waits for the copy)  */
/* into device memory of its own, kept until sarf_ctx_trim or destroy; a    */
/* list seen before only enqueues kernels, so after one call outside a      */
/* stream capture the same call can be captured into a HIP graph, and that  */
/* graph stays valid whatever other lists later calls use (a new list       */
/* during capture returns SARF_ERR_INVALID).  The reference compresses one  */
/* aperture per call (cuda_sar_focusing.hpp:616-621).                       */
typedef struct {
    const void *ph;       /* device float2 K x Np */
    void *prof;           /* device float2 Nfft x Np */
    const float *taper;   /* device, K (SARF_RC_TAPER) or NULL */
    const double *f0;     /* device, Np (SARF_RC_RAMP) or NULL */
    const double *r0;     /* device, Np (SARF_RC_RAMP) or NULL */
    int K, Np;            /* as in sarf_rc_desc */
    int Nfft;             /* power of two, 32 <= Nfft <= 16384 */
    int flags;            /* SARF_RC_TAPER | SARF_RC_RAMP */
} sarf_rc_scene;

int sarf_range_compress_scenes(sarf_ctx *ctx, int nscenes,
                               const sarf_rc_scene *scenes, /* host array */
                               void *stream);

/* ------------------------------------------------------------------------ */
/* Backprojection of range profiles onto the z = 0 pixel grid.              */
/* Replaces backprojection_loop (cuda_BackProjectionKernels.cuh:33-93) and  */
/* computeDifferentialRangeAndPhaseCorrections + interp1                    */
/* (cpuBackProjection.hpp:24-93):                                           */
/*   dR = |a_n - p| - R0_n;  if r_0 < dR < r_{Nfft-1}:                      */
/*   img[p] += lerp(prof[n], dR) * exp(j 4 pi f0_n dR / c)                  */
/* with the CPU reference's interpolation rule (left bin = floor).          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny;   /* image size, >= 2 each */
    int Nfft;     /* profile length (>= 2) */
    int Np;       /* pulses */
    double x0, y0;  /* image centre (m) */
    double Wx, Wy;  /* image extent (m), spacing W/(N-1) */
    double maxWr;   /* unambiguous range c/(2 mean deltaF) = max_Wy_m */
    int precision;  /* SARF_BP_MIXED | SARF_BP_F64 */
    int flags;      /* SARF_BP_ACCUMULATE */
    int tile;       /* 0 = default for the pixel spacing and line of sight; else a
                       tile id: 1..7 (lanes along y) or 11..16 (lanes along x);
                       batches and the gradient take 1..5 and 11..16 */
    int pulse_split;/* 0 = automatic; n >= 1: pulses in n contiguous groups summed
                       in group order -- n even groups, each a partial image, or,
                       when the image's tiles fill two rounds of resident
                       workgroups, one head group of the first 7/8 of the pulses
                       written into img and n - 1 tail groups as partial images */
    double los_x, los_y; /* optional hint: horizontal components of the unit line
                            of sight from the image centre to the aperture centre;
                            (0,0) = unknown.  Only steers the default tile choice. */
} sarf_bp_desc;

int sarf_backproject(sarf_ctx *ctx, const sarf_bp_desc *desc,
                     const void *prof,   /* device, complex Nfft x Np (float2; double2 for F64) */
                     const double *ant,  /* device, Np x 3 */
                     const double *r0,   /* device, Np */
                     const double *f0,   /* device, Np (start frequency of each pulse) */
                     void *img,          /* device, complex Ny x Nx (float2; double2 for F64) */
                     void *stream);

/* Allocation-free sarf_backproject: the same results, but no device memory */
/* is allocated or freed in the call (no hipMalloc/hipFree, no context      */
/* scratch).  Partial images of pulse splits go to a 64 MiB device array    */
/* that the library holds per GPU for the life of the process: an automatic */
/* or tuned split count is lowered until its partial images (Nx*Ny*8 bytes  */
/* each, one per group but the head group) fit, and an explicit             */
/* desc->pulse_split that does not fit returns SARF_ERR_INVALID.            */
/* Every call on one GPU shares that array: calls that overlap on a GPU     */
/* (other streams, other threads, other contexts) race.  No counterpart in  */
/* the reference; it is for callers that cannot allocate -- code under HIP  */
/* graph capture, tight serving loops, harnesses that forbid allocation.    */
int sarf_backproject_noalloc(sarf_ctx *ctx, const sarf_bp_desc *desc,
                             const void *prof, const double *ant, const double *r0,
                             const double *f0, void *img, void *stream);

/* ------------------------------------------------------------------------ */
/* Matched-filter image formation (exact, O(K) per pixel and pulse):        */
/*   img[p] += sum_n sum_k ph[n*K + k] exp(j 4 pi f_{n,k} dR_n(p) / c),     */
/*   f_{n,k} = f0[n] + k df[n],  dR_n(p) = |a_n - p| - R0_n,  z = 0,        */
/* every pixel and every pulse (no swath gate).  Replaces run_mf +          */
/* computeDifferentialRangeAndPhaseCorrectionsMF                            */
/* (src/cpuBackProjection/cpuBackProjection.hpp:165-228) and                */
/* matlab/mfBasic.m:66-89.  run_mf starts at pulse 2 (:218); callers that   */
/* want that pass the pulses from index 2 on.  Range and phase steps per    */
/* pixel and pulse in FP64; the sum over k in FP32 by Horner (K <= 8192).   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny;       /* image size, >= 2 each */
    int K;            /* samples per pulse, 1..8192 */
    int Np;           /* pulses, >= 1 */
    double x0, y0, Wx, Wy;
    int flags;        /* SARF_BP_ACCUMULATE */
    int pulse_split;  /* 0 = automatic; n >= 1 as in sarf_bp_desc */
} sarf_mf_desc;

int sarf_matched_filter(sarf_ctx *ctx, const sarf_mf_desc *desc,
                        const void *ph,     /* device, complex64 K x Np (ph[n*K + k]) */
                        const double *ant,  /* device, Np x 3 */
                        const double *r0,   /* device, Np */
                        const double *f0,   /* device, Np: frequency of sample 0 */
                        const double *df,   /* device, Np: frequency step */
                        void *img,          /* device, complex64 Ny x Nx */
                        void *stream);

/* ------------------------------------------------------------------------ */
/* Launch-shape tuner for sarf_backproject (mixed precision).  Times every   */
/* pixel tile that meets the same accuracy bound as the default choice,     */
/* crossed with pulse splits 1, 2, 4, .. 32 (>= 32 pulses per split), on the */
/* caller's buffers and stream, and waits for the stream.  The fastest pair  */
/* is remembered in the context for this image geometry (Nx, Ny, Np, Nfft,  */
/* spacing, maxWr, LOS hint): later sarf_backproject calls with tile = 0 and */
/* pulse_split = 0 use it.  On return img holds the backprojection made     */
/* with the chosen pair.  desc->flags must not hold SARF_BP_ACCUMULATE;     */
/* desc->tile and desc->pulse_split are ignored.  No reference counterpart: */
/* the reference launches one fixed block shape (src/gpuBackProjection/     */
/* cuda_sar_focusing/cuda_sar_focusing.hpp:632-634).                        */
/* ------------------------------------------------------------------------ */
typedef struct {
    int tile;          /* chosen tile id */
    int pulse_split;   /* chosen split count */
    float ms;          /* its mean time per call (ms) */
    float default_ms;  /* mean time of the untuned default choice (ms) */
    int candidates;    /* (tile, split) pairs timed */
} sarf_tune_result;

int sarf_autotune(sarf_ctx *ctx, const sarf_bp_desc *desc, const void *prof, const double *ant,
                  const double *r0, const double *f0, void *img, int reps,
                  sarf_tune_result *out, void *stream);

/* The launch shapes sarf_autotune remembered, as a table a caller can save */
/* (e.g. as JSON) and load into the context of a later process, so that a   */
/* serving process does not tune again.  An entry holds the geometry it     */
/* applies to -- sizes, spacings W/(N-1), maxWr and the LOS hint, matched   */
/* exactly -- and the chosen tile and pulse split.  No reference            */
/* counterpart (the reference has no tuner).                                */
typedef struct {
    int Nx, Ny, Np, Nfft;
    double dx, dy;        /* Wx/(Nx-1), Wy/(Ny-1) as sarf_bp_desc gives them */
    double maxWr;
    double los_x, los_y;  /* the LOS hint of the tuned descriptor (0, 0 = none) */
    int tile, pulse_split;
} sarf_tune_entry;

/* *count = number of remembered shapes; the first min(cap, *count) are     */
/* copied to out (out may be NULL when cap = 0).                            */
int sarf_tune_table_get(sarf_ctx *ctx, sarf_tune_entry *out, int cap, int *count);
/* Adds or replaces remembered shapes; every entry is validated (known tile, */
/* 1 <= pulse_split <= Np, positive spacings) before any is stored.         */
int sarf_tune_table_set(sarf_ctx *ctx, const sarf_tune_entry *entries, int n);

/* ------------------------------------------------------------------------ */
/* Image entropy E = -sum p ln p, p = |I|^2 / sum |I|^2 (FP64 sums).         */
/* out (device, 3 doubles) = {E, S = sum |I|^2, T = sum |I|^2 ln |I|^2}.     */
/* The focus metric of every autofocus path here; it replaces the           */
/* reference's incompatible metrics (8-bit dB column entropy,               */
/* src/autofocus/gridSearchErrorFunctions.cuh:341-447; DFT energy,          */
/* gridSearchErrorFunctions_dft.cuh:112-203).                               */
/* ------------------------------------------------------------------------ */
int sarf_image_entropy(sarf_ctx *ctx, const void *img, /* device complex, npix */
                       long npix, int f64 /* 0: float2, 1: double2 */,
                       double *out, void *stream);

/* ------------------------------------------------------------------------ */
/* Grid-search objective for many candidate trajectories in one persistent  */
/* launch.  Candidate c moves each phase centre along a unit direction:     */
/*   a_n(c) = a_n + (sum_k coef[c*nbasis + k] * basis[k*Np + n]) * dir_n     */
/* (e.g. a line-of-sight error a t^2 + b t^3 + c t^4), backprojects all     */
/* pulses (mixed precision, bp desc as for sarf_backproject) and writes the  */
/* entropy of candidate c's image to entropy[c]; images are not stored.     */
/* With desc->flags = SARF_BATCH_DIR_PER_BASIS, dir is nbasis x Np x 3 and  */
/*   a_n(c) = a_n + sum_k coef[c*nbasis + k] * basis[k*Np + n] * dir[k][n],  */
/* which with a_n = 0 and bases 1, i, i^2 along x, y, z expresses the       */
/* reference's absolute candidate trajectories (D = 7 and D = 10,           */
/* src/autofocus/gridSearchErrorFunctions_dft.cuh:279-291).                 */
/* Replaces kernelWrapper (src/autofocus/gridSearchErrorFunctions_dft.cuh:  */
/* 258-336) driven by cuGridSearch (src/autofocus/autofocus_gpu.cu:         */
/* 1057-1070).                                                              */
/* ------------------------------------------------------------------------ */
int sarf_bp_batch_entropy(sarf_ctx *ctx, const sarf_bp_desc *desc,
                          const void *prof,     /* device float2 Nfft x Np */
                          const double *ant, const double *r0, const double *f0,
                          int ncand, int nbasis,
                          const double *coef,   /* device, ncand x nbasis */
                          const double *basis,  /* device, nbasis x Np */
                          const double *dir,    /* device, Np x 3 (nbasis x Np x 3 per basis) */
                          double *entropy,      /* device, ncand */
                          void *stream);

/* ------------------------------------------------------------------------ */
/* Multi-resolution grid search: the level loop of the reference's autofocus*/
/* driver (src/autofocus/autofocus_gpu.cu:886-948 for the linear model,     */
/* :1036-1102 for the quadratic one; cuGridSearch's search_by_value_* and   */
/* find_extrema).  Each level evaluates every point of the box              */
/*   p_d = start_d + i_d (end_d - start_d) / (num_d - 1), i_d < num_d        */
/* (an axis with num_d = 1 stays at start_d; first axis slowest) with one    */
/* sarf_bp_batch_entropy launch, takes the first minimum, then shrinks the  */
/* box around it: start'_d = m_d - (end_d - start_d)/4 and, with            */
/* SARF_GRID_SHRINK_REFERENCE, end'_d = m_d + (end_d - start'_d)/4 -- the    */
/* reference's update order, which reads the already-moved start (:938-939)*/
/* -- or, with SARF_GRID_SHRINK_CENTRED, end'_d = m_d + (end_d - start_d)/4. */
/* Models: SARF_GRID_MODEL_COEF -- p is the coefficient vector of           */
/* sarf_bp_batch_entropy's linear model (dims = nbasis; basis and dir as    */
/* there, bp->flags may hold SARF_BATCH_DIR_PER_BASIS); SARF_GRID_MODEL_REF_D7 */
/* / _D10 -- p is the reference's parameter vector of absolute trajectories */
/* (gridSearchErrorFunctions_dft.cuh:279-291: x_i = p1 i p6 + p0, ... for D */
/* = 7; x_i = p2 i^2 + p1 i p9 + p0, ... for D = 10), ant, basis and dir are */
/* ignored.  Synchronous: best (host, dims), *best_entropy, levels[l] (host,*/
/* one per level, may be NULL) and cov (host, dims x dims, may be NULL) are */
/* filled on return.  cov is the covariance estimate of the last level's    */
/* minimiser (the reference prints cuGridSearch's covariance_by_value, a    */
/* library that is not available; this estimate is defined at              */
/* sarf_grid_covariance).                                                   */
/* ------------------------------------------------------------------------ */
#define SARF_GRID_MAX_DIMS 10
#define SARF_GRID_MODEL_COEF 0
#define SARF_GRID_MODEL_REF_D7 7
#define SARF_GRID_MODEL_REF_D10 10
#define SARF_GRID_SHRINK_REFERENCE 0
#define SARF_GRID_SHRINK_CENTRED 1

typedef struct {
    int dims;     /* parameters per candidate, 1..SARF_GRID_MAX_DIMS */
    int model;    /* SARF_GRID_MODEL_* */
    int levels;   /* resolution levels (the reference's multiRes), 1..64 */
    int shrink;   /* SARF_GRID_SHRINK_* */
    double start[SARF_GRID_MAX_DIMS];
    double end[SARF_GRID_MAX_DIMS];
    int num[SARF_GRID_MAX_DIMS];  /* samples per axis, 1..4096; product <= 2^21 */
} sarf_grid_desc;

typedef struct {
    double start[SARF_GRID_MAX_DIMS]; /* the box this level searched */
    double end[SARF_GRID_MAX_DIMS];
    double best[SARF_GRID_MAX_DIMS];  /* its minimiser */
    double entropy;                   /* entropy at the minimiser */
    float gpu_ms;                     /* device time of the level's launch */
    int ncand;                        /* candidates evaluated */
} sarf_grid_level;

int sarf_grid_search(sarf_ctx *ctx, const sarf_bp_desc *bp,
                     const void *prof,     /* device float2 Nfft x Np */
                     const double *ant,    /* device Np x 3 (COEF model) or NULL */
                     const double *r0, const double *f0,
                     const sarf_grid_desc *grid,
                     const double *basis,  /* device (COEF model) or NULL */
                     const double *dir,    /* device (COEF model) or NULL */
                     double *best,         /* host, dims */
                     double *best_entropy, /* host, 1, may be NULL */
                     double *cov,          /* host, dims x dims, may be NULL */
                     sarf_grid_level *levels, /* host, grid->levels, may be NULL */
                     void *stream);

/* Host only: the covariance estimate of sarf_grid_search for objective     */
/* values[c] over the box of `grid` (start/end/num, candidates ordered as   */
/* above) with minimiser index `best`: a Laplace approximation -- a full    */
/* quadratic f = c + g.u + u^T H u / 2 fitted by least squares to the 3^m   */
/* values around the minimiser (window moved inward at the box edge) over   */
/* the m axes with >= 3 samples, u in parameter units, and cov = H^-1 there */
/* when H is positive definite.  Axes of one sample: zero rows and columns; */
/* axes of two samples or a fit that is not positive definite: NaN.  No     */
/* reference counterpart (cuGridSearch's covariance_by_value is absent).    */
int sarf_grid_covariance(const sarf_grid_desc *grid, const double *values, long best,
                         double *cov /* host, dims x dims */);

/* Host only: the reference's trajectory fits that centre its grid          */
/* (bestFit / quadFit, src/autofocus/autofocus_gpu.cu:315-416), over every  */
/* skip-th sample of values (e.g. Ant_x) against the pulse index i:         */
/* least squares with double sums of the float products values[i] * i (and  */
/* * i again), results rounded to float.  coeffs = {intercept, slope} and   */
/* {constant, linear, quadratic}.  Index powers are exact here; the         */
/* reference forms i^4 in int, which wraps from i = 216.                    */
int sarf_fit_linear(const float *values, int n, int skip, float *coeffs /* 2 */);
int sarf_fit_quadratic(const float *values, int n, int skip, float *coeffs /* 3 */);

/* ------------------------------------------------------------------------ */
/* Adjoint backprojection: grad[n] = dE/da_n (FP64, Np x 3), the gradient    */
/* of the image entropy with respect to each antenna phase centre (R0 held   */
/* fixed), given the forward image img = sarf_backproject(...) of the same   */
/* inputs and ent = sarf_image_entropy(img).  Interpolated profile values    */
/* are recomputed, nothing per pixel-pulse is stored.  No counterpart in the */
/* reference; it makes the grid search of src/autofocus/autofocus_gpu.cu     */
/* a gradient search.                                                       */
/* ------------------------------------------------------------------------ */
int sarf_bp_entropy_grad(sarf_ctx *ctx, const sarf_bp_desc *desc,
                         const void *prof,    /* device float2 Nfft x Np */
                         const double *ant, const double *r0, const double *f0,
                         const void *img,     /* device float2 Ny x Nx (forward image) */
                         const double *ent,   /* device {E, S, T} of img */
                         double *grad,        /* device, Np x 3 */
                         void *stream);

/* ------------------------------------------------------------------------ */
/* Backprojection of many independent scenes of uneven size in one          */
/* persistent launch (work items = scene tiles, largest first, drained      */
/* through an atomic counter).  Each scene is described as for              */
/* sarf_backproject, with its own device buffers; images are overwritten.   */
/* tile: 1..5 or 11..16 for every scene, or SARF_SCENES_TILE_MIXED: each   */
/* scene on the one of tiles 15, 16 and 13 that sarf_backproject's default  */
/* weighs cheapest for it (relative speed over the share of its tile grid   */
/* inside the image), all in the same launch; SARF_ERR_INVALID if a scene   */
/* has no accurate tile among the three.  Tiles 15, 16 and 13 run lanes     */
/* along x, the fast orientation for a line of sight along y.  0: a tile    */
/* per scene when every scene has a line-of-sight hint with |los_y| >=      */
/* |los_x| and an accurate tile among the three, else one tile for all,     */
/* chosen for the finest-spaced scene.                                      */
/* Each distinct work list (descriptors and tile) is uploaded once (the     */
/* call waits for the copy) into device memory of its own, kept until       */
/* sarf_ctx_trim or destroy; a list seen before only enqueues kernels, so   */
/* after one call outside a stream capture the same call can be captured   */
/* into a HIP graph, which stays valid whatever lists later calls use (a    */
/* new list during capture returns SARF_ERR_INVALID).                       */
/* The reference focuses one aperture per call (cuda_focus_SAR_image,       */
/* cuda_sar_focusing.hpp:526-723).                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    const void *prof;   /* device float2 Nfft x Np */
    const double *ant;  /* device Np x 3 */
    const double *r0;   /* device Np */
    const double *f0;   /* device Np */
    void *img;          /* device float2 Ny x Nx */
    int Nx, Ny, Nfft, Np;
    double x0, y0, Wx, Wy, maxWr;
    double los_x, los_y; /* optional tile hint, as in sarf_bp_desc */
} sarf_scene;

int sarf_bp_scenes(sarf_ctx *ctx, int nscenes,
                   const sarf_scene *scenes, /* host array of nscenes descriptors */
                   int tile, void *stream);

/* Tile tuner for sarf_bp_scenes: times the batch with every batch tile that*/
/* meets the accuracy bound of every scene (`reps` runs each, after one      */
/* warm run), waits for the stream, and returns the fastest in *best_tile    */
/* (and its mean time in *best_ms, may be NULL).  Images hold the last       */
/* run's result.  No reference counterpart (see sarf_autotune).              */
int sarf_autotune_scenes(sarf_ctx *ctx, int nscenes, const sarf_scene *scenes, int reps,
                         int *best_tile, float *best_ms, void *stream);

/* ------------------------------------------------------------------------ */
/* Phase gradient autofocus of a focused image, in place (matlab/PGA.m:     */
/* 151-232; C++ restatement src/autofocus/cpu/autofocus_cpu.cpp:249-474).   */
/* Rows are range (y), columns azimuth (x); img is img[x*Ny + y].  Runs     */
/* `iters` iterations, or with rms_stop > 0 stops at the first iteration    */
/* whose detrended estimate has rms < rms_stop, without applying that       */
/* estimate -- the stopping rule of the C++ restatement                     */
/* (autofocus_cpu.cpp:432, threshold 0.1 there); everything stays on the    */
/* device, no host synchronisation.  af_ph (device, Nx doubles) receives    */
/* the accumulated detrended phase of the applied iterations, rms (device,  */
/* iters) the rms of each iteration's estimate (NaN for iterations that did */
/* not run when rms_stop > 0).                                               */
/* ------------------------------------------------------------------------ */
typedef struct {
    int Nx, Ny;      /* image size; 2 <= Nx <= 4096 */
    int iters;       /* iterations (the maximum when rms_stop > 0) */
    double db_down;  /* window cut-off below the peak column energy, dB (30) */
    double rms_stop; /* > 0: early-stop threshold (rad); 0: fixed count */
} sarf_pga_desc;

int sarf_pga(sarf_ctx *ctx, const sarf_pga_desc *desc,
             void *img,       /* device float2 Ny x Nx, updated in place */
             double *af_ph,   /* device, Nx */
             double *rms,     /* device, iters */
             void *stream);

/* PGA on many images of different sizes at once: every step of sarf_pga  */
/* is one launch over all images (the row transforms one launch per       */
/* transform length), with results identical to sarf_pga on each image.  */
/* The images' device state (descriptors and work arrays) lives in device  */
/* memory of its own per distinct list (and iters, rms_stop), uploaded once */
/* (the call waits for the copy) and kept until sarf_ctx_trim or destroy,   */
/* so after one call outside a stream capture the same call can be captured*/
/* into a HIP graph that later calls with other lists do not disturb.  No   */
/* reference counterpart: the reference autofocuses one image per call      */
/* (autofocus_cpu.cpp:249-474).                                             */
typedef struct {
    void *img;      /* device float2 Ny x Nx, updated in place */
    double *af_ph;  /* device, Nx */
    double *rms;    /* device, iters */
    int Nx, Ny;     /* 2 <= Nx <= 4096, Ny >= 1 */
} sarf_pga_scene;

int sarf_pga_scenes(sarf_ctx *ctx, int nscenes, const sarf_pga_scene *scenes, int iters,
                    double db_down, double rms_stop /* as in sarf_pga_desc, per image */,
                    void *stream);

/* ------------------------------------------------------------------------ */
/* Whole focusing call on host buffers: H2D, range compression,             */
/* backprojection, D2H.  This is what cuda_focus_SAR_image<float,float>     */
/* (cuda_sar_focusing.hpp:526-723) does; it is the entry the C++ host       */
/* wrappers in include/uncc_sar_focusing.hpp call.  Synchronous.  Pageable  */
/* ph / img go through the context's pinned staging; page-locked ones       */
/* (hipHostMalloc'd or registered) are copied directly.                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    int K, Np, Nfft;
    int Nx, Ny;
    double x0, y0, Wx, Wy, maxWr;
    int precision; /* SARF_BP_MIXED | SARF_BP_F64 (F64: ph/img stay float on the host) */
} sarf_focus_desc;

int sarf_focus(sarf_ctx *ctx, const sarf_focus_desc *desc,
               const float *ph,    /* host, complex K x Np */
               const double *ant,  /* host, Np x 3 */
               const double *r0,   /* host, Np */
               const double *f0,   /* host, Np */
               float *img);        /* host, complex Ny x Nx */

/* Matched-filter image formation on host buffers: H2D, sarf_matched_filter, */
/* D2H.  Synchronous.  This is what focus_SAR_image does when the image      */
/* parameters select MATCHED_FILTER (run_mf, src/cpuBackProjection/          */
/* cpuBackProjection.hpp:204-228, dispatched at :255-257); the C++ wrapper  */
/* in include/uncc_sar_focusing.hpp passes the pulses run_mf uses (index 2  */
/* on, :218).  With SARF_BP_ACCUMULATE in desc->flags img is read and added */
/* to; otherwise it is overwritten.                                         */
int sarf_focus_mf(sarf_ctx *ctx, const sarf_mf_desc *desc,
                  const float *ph,    /* host, complex K x Np */
                  const double *ant,  /* host, Np x 3 */
                  const double *r0,   /* host, Np */
                  const double *f0,   /* host, Np: frequency of sample 0 */
                  const double *df,   /* host, Np: frequency step */
                  float *img);        /* host, complex Ny x Nx */

/* ------------------------------------------------------------------------ */
/* Point-scatterer phase-history synthesis (test/benchmark data source;     */
/* data model of matlab/formImagePointTargetBP.m:63-81):                    */
/*   ph[n,k] += sum_s A_s exp(-j 4 pi (f0_n + k df)(|a_n - p_s| - R0_n)/c)  */
/* Phase argument and its reduction in FP64, rotation in FP32, sums in FP32.*/
/* Adds into ph (caller zeroes it first).                                   */
/* ------------------------------------------------------------------------ */
int sarf_synth_phase_history(sarf_ctx *ctx, int K, int Np, int S,
                             const double *scat,  /* device, S x 3 */
                             const float *amp,    /* device, complex S */
                             const double *ant, const double *r0, const double *f0,
                             double df, void *ph, /* device, float2 K x Np */
                             void *stream);

/* Same data model for many scatterers (distributed clutter): a type-1      */
/* non-uniform FFT per pulse (8-tap exponential-of-semicircle spreading on  */
/* a 2x oversampled grid), O(S + K log K) per pulse instead of O(S K);      */
/* ~1e-6 relative error.  Needs K <= 8192.  Adds into ph.                   */
int sarf_synth_phase_history_nufft(sarf_ctx *ctx, int K, int Np, int S,
                                   const double *scat, const float *amp,
                                   const double *ant, const double *r0, const double *f0,
                                   double df, void *ph, void *stream);

/* ------------------------------------------------------------------------ */
/* 8-bit dB display image: the quantisation of writeBMPFile                 */
/* (include/uncc_sar_globals.hpp:66-118), byte for byte:                    */
/*   pixelf = (float)(255.0/dyn) * (20 log10(|I| / max|I|) + dyn),          */
/*   byte = pixelf < 0 ? 0 : (unsigned char)pixelf,                         */
/* in float with the host's log10, rows top-down, row r showing y = Ny-1-r  */
/* (far range on top, :85-98): out[r*Nx + x].  The kernel compares          */
/* r = |I|/max|I| with the 255 level thresholds of sarf_db8_levels, so no   */
/* device log10 enters the result.                                          */
int sarf_image_db8(sarf_ctx *ctx,
                   const void *img,        /* device, complex64 Ny x Nx (img[x*Ny + y]) */
                   int Nx, int Ny, float dyn_range_dB,
                   unsigned char *out,     /* device, Nx*Ny bytes */
                   void *stream);

/* Host only (no device needed): the thresholds sarf_image_db8 quantises   */
/* with.  levels[b], b = 1..255, is the smallest float r in [0, 1] whose    */
/* byte above is >= b (+inf when no r <= 1 reaches b); levels[0] = 0.  The  */
/* byte of r is the number of levels[1..255] <= r.  Found by bisection over */
/* float bit patterns with the reference's float expression.  No reference  */
/* counterpart (the reference evaluates log10 per pixel, uncc_sar_globals.hpp:*/
/* 99-104).                                                                 */
int sarf_db8_levels(float dyn_range_dB, float *levels /* host, 256 floats */);

#ifdef __cplusplus
}
#endif

#endif /* SARF_H */

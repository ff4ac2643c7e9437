// sarf_gridsearch.hip -- multi-resolution grid-search autofocus driver and
// the trajectory fits that seed it (host code over sarf_bp_batch_entropy).
//
// Reference counterparts: the level loop of src/autofocus/autofocus_gpu.cu
// (linear model :886-948, quadratic model :1036-1102): search a box of
// candidate parameters, take the argmin, shrink the box around it, repeat
// `multiRes` times; bestFit / quadFit (:315-416) give the box centre from the
// recorded antenna track.  The candidates of one level are one persistent
// launch of the batched entropy kernel (sarf_batch.hip); this file only
// enumerates candidates, reads back their entropies and moves the box.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sarf_internal.hpp"

namespace sarf {
namespace {

// Candidate count of a box, or -1 if an axis count is out of range.
long box_count(const sarf_grid_desc *g) {
    long n = 1;
    for (int d = 0; d < g->dims; d++) {
        if (g->num[d] < 1 || g->num[d] > 4096) return -1;
        n *= g->num[d];
        if (n > kGridMaxCandidates) return -1;
    }
    return n;
}

// Grid point of axis d at index i: start + i (end - start) / (num - 1); an
// axis of one sample stays at its start.
double axis_point(const double *start, const double *end, const int *num, int d, int i) {
    if (num[d] <= 1) return start[d];
    return start[d] + (double)i * ((end[d] - start[d]) / (double)(num[d] - 1));
}

// Index vector of candidate c (first axis slowest).
void unflatten(long c, int dims, const int *num, int *idx) {
    for (int d = dims - 1; d >= 0; d--) {
        idx[d] = (int)(c % num[d]);
        c /= num[d];
    }
}

// Number of linear-model coefficients for a model and its parameter count.
int model_nbasis(int model, int dims) {
    if (model == SARF_GRID_MODEL_COEF) return dims;
    if (model == SARF_GRID_MODEL_REF_D7) return 6;
    if (model == SARF_GRID_MODEL_REF_D10) return 9;
    return -1;
}

// Linear-model coefficients of one parameter vector.  For the reference's
// absolute trajectories (gridSearchErrorFunctions_dft.cuh:279-291) the
// positions are sum_k coef_k basis_k[i] along axis k / terms, with bases
// 1, i (D = 7) or 1, i, i^2 (D = 10) per axis, and the velocity scale (p6,
// p9) multiplies the linear term only.
void model_coefs(int model, const double *p, int dims, double *coef) {
    if (model == SARF_GRID_MODEL_COEF) {
        for (int k = 0; k < dims; k++) coef[k] = p[k];
        return;
    }
    const int terms = model == SARF_GRID_MODEL_REF_D7 ? 2 : 3;
    const double scale = p[model == SARF_GRID_MODEL_REF_D7 ? 6 : 9];
    for (int a = 0; a < 3; a++)
        for (int q = 0; q < terms; q++) {
            const int k = a * terms + q;
            coef[k] = p[k] * (q == 1 ? scale : 1.0);
        }
}

// Solves the dense n x n system A x = b in place (partial pivoting).
// Returns false if A is singular to working precision.
bool solve_dense(std::vector<double> &A, std::vector<double> &b, int n) {
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
        if (!(std::fabs(A[piv * n + c]) > 0.0)) return false;
        if (piv != c) {
            for (int k = 0; k < n; k++) std::swap(A[c * n + k], A[piv * n + k]);
            std::swap(b[c], b[piv]);
        }
        for (int r = c + 1; r < n; r++) {
            const double f = A[r * n + c] / A[c * n + c];
            if (f == 0.0) continue;
            for (int k = c; k < n; k++) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; c--) {
        double s = b[c];
        for (int k = c + 1; k < n; k++) s -= A[c * n + k] * b[k];
        b[c] = s / A[c * n + c];
    }
    for (int c = 0; c < n; c++)
        if (!std::isfinite(b[c])) return false;
    return true;
}

}  // namespace

// Covariance of the minimiser from the objective values of one level (a
// Laplace approximation, treating the objective as a negative log
// likelihood): a full quadratic f = c + g.u + u^T H u / 2 is fitted by least
// squares to the values on the 3 x .. x 3 neighbourhood of the argmin (moved
// inward at the box edge) over the axes with at least 3 samples, u in
// parameter units; cov = H^-1 on those axes when H is positive definite.
// Axes with one sample get zero rows and columns; axes with two samples, or
// a fit that is not positive definite, get NaN rows and columns.
int grid_covariance(const sarf_grid_desc *g, const double *start, const double *end,
                    const double *values, long best, double *cov) {
    const int D = g->dims;
    int bidx[SARF_GRID_MAX_DIMS];
    unflatten(best, D, g->num, bidx);
    std::vector<int> free_axes;
    bool nan_axes[SARF_GRID_MAX_DIMS] = {false};
    for (int d = 0; d < D; d++) {
        if (g->num[d] >= 3) free_axes.push_back(d);
        else if (g->num[d] == 2) nan_axes[d] = true;
    }
    for (int i = 0; i < D * D; i++) cov[i] = 0.0;
    const int m = (int)free_axes.size();
    bool ok = m > 0;
    std::vector<double> H((size_t)m * m, 0.0);
    if (ok) {
        // window start per free axis: argmin - 1, clamped into [0, num - 3]
        std::vector<int> w0(m);
        for (int j = 0; j < m; j++) {
            const int d = free_axes[j];
            w0[j] = std::min(std::max(bidx[d] - 1, 0), g->num[d] - 3);
        }
        const int nq = m * (m + 1) / 2;
        const int nu = 1 + m + nq;  // c, g_j, H_jk (j <= k)
        std::vector<double> AtA((size_t)nu * nu, 0.0), Atb(nu, 0.0), row(nu);
        long npts = 1;
        for (int j = 0; j < m; j++) npts *= 3;
        std::vector<double> u(m);
        int idx[SARF_GRID_MAX_DIMS];
        for (long q = 0; q < npts; q++) {
            long r = q;
            for (int d = 0; d < D; d++) idx[d] = bidx[d];
            for (int j = 0; j < m; j++) {
                const int d = free_axes[j];
                idx[d] = w0[j] + (int)(r % 3);
                r /= 3;
                u[j] = axis_point(start, end, g->num, d, idx[d]) -
                       axis_point(start, end, g->num, d, bidx[d]);
            }
            long flat = 0;
            for (int d = 0; d < D; d++) flat = flat * g->num[d] + idx[d];
            const double f = values[flat];
            if (!std::isfinite(f)) {
                ok = false;
                break;
            }
            int k = 0;
            row[k++] = 1.0;
            for (int j = 0; j < m; j++) row[k++] = u[j];
            for (int j = 0; j < m; j++)
                for (int l = j; l < m; l++) row[k++] = (j == l ? 0.5 : 1.0) * u[j] * u[l];
            for (int a = 0; a < nu; a++) {
                Atb[a] += row[a] * f;
                for (int b = 0; b < nu; b++) AtA[(size_t)a * nu + b] += row[a] * row[b];
            }
        }
        if (ok) ok = solve_dense(AtA, Atb, nu);
        if (ok) {
            int k = 1 + m;
            for (int j = 0; j < m; j++)
                for (int l = j; l < m; l++) {
                    H[(size_t)j * m + l] = Atb[k];
                    H[(size_t)l * m + j] = Atb[k];
                    k++;
                }
            // positive definite: Cholesky must succeed
            std::vector<double> L((size_t)m * m, 0.0);
            for (int j = 0; j < m && ok; j++) {
                for (int l = 0; l <= j; l++) {
                    double s = H[(size_t)j * m + l];
                    for (int k2 = 0; k2 < l; k2++) s -= L[(size_t)j * m + k2] * L[(size_t)l * m + k2];
                    if (j == l) {
                        if (!(s > 0.0)) {
                            ok = false;
                            break;
                        }
                        L[(size_t)j * m + j] = std::sqrt(s);
                    } else {
                        L[(size_t)j * m + l] = s / L[(size_t)l * m + l];
                    }
                }
            }
        }
        if (ok) {
            // inverse column by column
            for (int c = 0; c < m && ok; c++) {
                std::vector<double> A(H), e(m, 0.0);
                e[c] = 1.0;
                ok = solve_dense(A, e, m);
                for (int j = 0; j < m && ok; j++)
                    cov[free_axes[j] * D + free_axes[c]] = e[j];
            }
        }
    }
    const double qnan = std::nan("");
    for (int d = 0; d < D; d++) {
        const bool is_free = std::find(free_axes.begin(), free_axes.end(), d) != free_axes.end();
        if (nan_axes[d] || (is_free && !ok))
            for (int e = 0; e < D; e++) {
                cov[d * D + e] = qnan;
                cov[e * D + d] = qnan;
            }
    }
    return SARF_OK;
}

}  // namespace sarf

using namespace sarf;

extern "C" int sarf_grid_search(sarf_ctx *ctx, const sarf_bp_desc *bp, const void *prof,
                                const double *ant, const double *r0, const double *f0,
                                const sarf_grid_desc *g, const double *basis, const double *dir,
                                double *best, double *best_entropy, double *cov,
                                sarf_grid_level *levels, void *stream) {
    clear_error();
    SARF_TRY(use_device(ctx));
    StreamOrder order_(ctx, (hipStream_t)stream);
    if (!bp || !g || !prof || !r0 || !f0 || !best)
        return fail(SARF_ERR_INVALID, "sarf_grid_search: null argument");
    const int D = g->dims;
    if (D < 1 || D > SARF_GRID_MAX_DIMS || g->levels < 1 || g->levels > 64)
        return fail(SARF_ERR_INVALID, "sarf_grid_search: dims=%d levels=%d", D, g->levels);
    if (g->shrink != SARF_GRID_SHRINK_REFERENCE && g->shrink != SARF_GRID_SHRINK_CENTRED)
        return fail(SARF_ERR_INVALID, "sarf_grid_search: shrink %d", g->shrink);
    const int nb = model_nbasis(g->model, D);
    if (nb < 0) return fail(SARF_ERR_INVALID, "sarf_grid_search: model %d", g->model);
    if ((g->model == SARF_GRID_MODEL_REF_D7 && D != 7) ||
        (g->model == SARF_GRID_MODEL_REF_D10 && D != 10))
        return fail(SARF_ERR_INVALID, "sarf_grid_search: model %d needs %d parameters, got %d",
                    g->model, g->model, D);
    if (g->model == SARF_GRID_MODEL_COEF && (!ant || !basis || !dir))
        return fail(SARF_ERR_INVALID, "sarf_grid_search: the coefficient model needs ant, basis "
                    "and dir");
    for (int d = 0; d < D; d++)
        if (!std::isfinite(g->start[d]) || !std::isfinite(g->end[d]) || g->end[d] < g->start[d])
            return fail(SARF_ERR_INVALID, "sarf_grid_search: axis %d box [%g, %g]", d,
                        g->start[d], g->end[d]);
    const long ncand = box_count(g);
    if (ncand < 1)
        return fail(SARF_ERR_INVALID, "sarf_grid_search: axis counts must be 1..4096 with at most "
                    "%ld candidates per level", (long)kGridMaxCandidates);
    const int Np = bp->Np;
    if (Np < 1) return fail(SARF_ERR_INVALID, "sarf_grid_search: Np=%d", Np);
    hipStream_t s = (hipStream_t)stream;

    // device buffers: candidate coefficients, entropies, and for the
    // reference models the zero antenna track, bases and directions
    const size_t model_bytes =
        g->model == SARF_GRID_MODEL_COEF ? 0 : sizeof(double) * ((size_t)Np * 3 + (size_t)nb * Np +
                                                                 (size_t)nb * Np * 3);
    const size_t coef_bytes = sizeof(double) * (size_t)ncand * nb;
    const size_t ent_bytes = sizeof(double) * (size_t)ncand;
    void *buf = nullptr;
    SARF_TRY(scratch_get(ctx, "grid.buf", coef_bytes + ent_bytes + model_bytes, &buf));
    double *dcoef = (double *)buf;
    double *dent = dcoef + (size_t)ncand * nb;
    const double *use_ant = ant, *use_basis = basis, *use_dir = dir;
    sarf_bp_desc bd = *bp;
    if (g->model != SARF_GRID_MODEL_COEF) {
        double *dzero = dent + ncand;
        double *dbasis = dzero + (size_t)Np * 3;
        double *ddir = dbasis + (size_t)nb * Np;
        const int terms = g->model == SARF_GRID_MODEL_REF_D7 ? 2 : 3;
        std::vector<double> h((size_t)Np * 3 + (size_t)nb * Np + (size_t)nb * Np * 3, 0.0);
        double *hb = h.data() + (size_t)Np * 3, *hd = hb + (size_t)nb * Np;
        for (int a = 0; a < 3; a++)
            for (int q = 0; q < terms; q++) {
                const int k = a * terms + q;
                for (int i = 0; i < Np; i++) {
                    hb[(size_t)k * Np + i] = q == 0 ? 1.0 : (q == 1 ? (double)i : (double)i * i);
                    hd[((size_t)k * Np + i) * 3 + a] = 1.0;
                }
            }
        SARF_HIP(hipMemcpyAsync(dzero, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s));
        use_ant = dzero;
        use_basis = dbasis;
        use_dir = ddir;
        bd.flags |= SARF_BATCH_DIR_PER_BASIS;
    } else {
        bd.flags &= SARF_BATCH_DIR_PER_BASIS;
    }
    bd.precision = SARF_BP_MIXED;

    double st[SARF_GRID_MAX_DIMS], en[SARF_GRID_MAX_DIMS];
    for (int d = 0; d < D; d++) {
        st[d] = g->start[d];
        en[d] = g->end[d];
    }
    std::vector<double> hcoef((size_t)ncand * nb), hent(ncand), p(D);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SARF_HIP(hipEventCreate(&e0));
    {
        hipError_t e = hipEventCreate(&e1);
        if (e != hipSuccess) {
            (void)hipEventDestroy(e0);
            return fail(SARF_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    int rc = SARF_OK;
    long bi = 0;
    int idx[SARF_GRID_MAX_DIMS];
    for (int lv = 0; lv < g->levels && rc == SARF_OK; lv++) {
        for (long c = 0; c < ncand; c++) {
            unflatten(c, D, g->num, idx);
            for (int d = 0; d < D; d++) p[d] = axis_point(st, en, g->num, d, idx[d]);
            model_coefs(g->model, p.data(), D, &hcoef[(size_t)c * nb]);
        }
        hipError_t he = hipMemcpyAsync(dcoef, hcoef.data(), coef_bytes, hipMemcpyHostToDevice, s);
        if (he != hipSuccess) {
            rc = fail(SARF_ERR_HIP, "sarf_grid_search: upload: %s", hipGetErrorString(he));
            break;
        }
        (void)hipEventRecord(e0, s);
        rc = sarf_bp_batch_entropy(ctx, &bd, prof, use_ant, r0, f0, (int)ncand, nb, dcoef,
                                   use_basis, use_dir, dent, stream);
        if (rc != SARF_OK) break;
        (void)hipEventRecord(e1, s);
        he = hipMemcpyAsync(hent.data(), dent, ent_bytes, hipMemcpyDeviceToHost, s);
        if (he == hipSuccess) he = hipStreamSynchronize(s);
        if (he != hipSuccess) {
            rc = fail(SARF_ERR_HIP, "sarf_grid_search: %s", hipGetErrorString(he));
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        // first minimum (ties keep the lower index); NaN never wins
        bi = -1;
        for (long c = 0; c < ncand; c++)
            if (std::isfinite(hent[c]) && (bi < 0 || hent[c] < hent[bi])) bi = c;
        if (bi < 0) {
            rc = fail(SARF_ERR_INVALID, "sarf_grid_search: no finite entropy at level %d", lv);
            break;
        }
        unflatten(bi, D, g->num, idx);
        double m[SARF_GRID_MAX_DIMS];
        for (int d = 0; d < D; d++) m[d] = axis_point(st, en, g->num, d, idx[d]);
        if (levels) {
            sarf_grid_level &L = levels[lv];
            std::memset(&L, 0, sizeof L);
            for (int d = 0; d < D; d++) {
                L.start[d] = st[d];
                L.end[d] = en[d];
                L.best[d] = m[d];
            }
            L.entropy = hent[bi];
            L.gpu_ms = ms;
            L.ncand = (int)ncand;
        }
        for (int d = 0; d < D; d++) best[d] = m[d];
        if (best_entropy) *best_entropy = hent[bi];
        if (cov) grid_covariance(g, st, en, hent.data(), bi, cov);
        // shrink around the minimum to half the span; the reference computes
        // the new end from the already-updated start (autofocus_gpu.cu:938-939)
        for (int d = 0; d < D; d++) {
            const double w = en[d] - st[d];
            st[d] = m[d] - w / 4;
            en[d] = g->shrink == SARF_GRID_SHRINK_REFERENCE ? m[d] + (en[d] - st[d]) / 4
                                                            : m[d] + w / 4;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// bestFit (autofocus_gpu.cu:315-371): least-squares line through every
// skip-th value against its index, sums in double with the float product
// values[i] * i of the reference; coeffs = {intercept, slope} as float.
extern "C" int sarf_fit_linear(const float *values, int n, int skip, float *coeffs) {
    clear_error();
    if (!values || !coeffs || n < 2 || skip < 1)
        return fail(SARF_ERR_INVALID, "sarf_fit_linear: n=%d skip=%d", n, skip);
    double N = 0, x1 = 0, x2 = 0, f0 = 0, f1 = 0;
    for (int i = 0; i < n; i += skip) {
        N += 1;
        x1 += i;
        x2 += (double)((long long)i * i);
        f0 += values[i];
        f1 += (float)(values[i] * (float)i);
    }
    const double D = -x1 * x1 + N * x2;
    if (!(D != 0.0)) return fail(SARF_ERR_INVALID, "sarf_fit_linear: fewer than two samples");
    const double a = (N * f1 - f0 * x1) / D;
    const double b = (f0 * x2 - f1 * x1) / D;
    coeffs[0] = (float)b;
    coeffs[1] = (float)a;
    return SARF_OK;
}

// quadFit (autofocus_gpu.cu:373-416): least-squares parabola by Cramer's rule
// on the normal equations; coeffs = {constant, linear, quadratic} as float.
// The index powers are exact (the reference forms i^3 and i^4 in int, which
// wraps from i = 1291 and i = 216).
extern "C" int sarf_fit_quadratic(const float *values, int n, int skip, float *coeffs) {
    clear_error();
    if (!values || !coeffs || n < 3 || skip < 1)
        return fail(SARF_ERR_INVALID, "sarf_fit_quadratic: n=%d skip=%d", n, skip);
    double N = 0, x1 = 0, x2 = 0, x3 = 0, x4 = 0, f0 = 0, f1 = 0, f2 = 0;
    for (int i = 0; i < n; i += skip) {
        const long long ii = i;
        N += 1;
        x1 += i;
        x2 += (double)(ii * ii);
        x3 += (double)(ii * ii * ii);
        x4 += (double)(ii * ii) * (double)(ii * ii);
        f0 += values[i];
        const float vi = values[i] * (float)i;
        f1 += vi;
        f2 += (float)(vi * (float)i);
    }
    const double D = x4 * x1 * x1 - 2 * x1 * x2 * x3 + x2 * x2 * x2 - N * x4 * x2 + N * x3 * x3;
    if (!(D != 0.0)) return fail(SARF_ERR_INVALID, "sarf_fit_quadratic: fewer than three samples");
    const double a = (f2 * x1 * x1 - f1 * x1 * x2 - f0 * x3 * x1 + f0 * x2 * x2 - N * f2 * x2 +
                      N * f1 * x3) / D;
    const double b = (f1 * x2 * x2 - N * f1 * x4 + N * f2 * x3 + f0 * x1 * x4 - f0 * x2 * x3 -
                      f2 * x1 * x2) / D;
    const double c = (f2 * x2 * x2 - f1 * x2 * x3 - f0 * x4 * x2 + f0 * x3 * x3 - f2 * x1 * x3 +
                      f1 * x1 * x4) / D;
    coeffs[0] = (float)c;
    coeffs[1] = (float)b;
    coeffs[2] = (float)a;
    return SARF_OK;
}

// Host only: the covariance estimate of sarf_grid_search for given objective
// values of one level (exported so the estimator can be checked without a
// device).
extern "C" int sarf_grid_covariance(const sarf_grid_desc *g, const double *values, long best,
                                    double *cov) {
    clear_error();
    if (!g || !values || !cov || g->dims < 1 || g->dims > SARF_GRID_MAX_DIMS)
        return fail(SARF_ERR_INVALID, "sarf_grid_covariance: bad argument");
    const long n = box_count(g);
    if (n < 1 || best < 0 || best >= n)
        return fail(SARF_ERR_INVALID, "sarf_grid_covariance: best=%ld of %ld", best, n);
    return grid_covariance(g, g->start, g->end, values, best, cov);
}

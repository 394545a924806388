/*
 * sfh_oracle.c -- CPU FP64 restatement of the StarFormationHistories.jl
 * algorithms on the two data-parallel paths (template-mixture likelihood and
 * smoothed Hess-diagram template deposition).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links it and has no CPU
 * fallback.
 *
 * Every function cites the reference lines it follows (paths relative to the
 * StarFormationHistories.jl v1.3.1 source tree).  Arithmetic is FP64 and is
 * compiled with -ffp-contract=off so that a*b+c is never fused, matching
 * Julia's default (non-@fastmath, non-muladd) semantics for the index math.
 *
 * Conventions
 *   models   : N x T, column-major (each template one contiguous column), the
 *              `stack_models` layout (src/fitting/utilities.jl:12-13).
 *   X        : T x B, column-major (one coefficient vector per walker).
 *   Hess mat : (nx) x (ny), column-major, x (colour) fastest, as the
 *              reference's `zeros(Float64, length(edges[1])-1,
 *              length(edges[2])-1)` (src/StarFormationHistories.jl:584).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Template-mixture likelihood (src/fitting/fitting_base.jl)                 */
/* ------------------------------------------------------------------------ */

/* composite!(composite::AbstractVector, coeffs, models::AbstractMatrix)
 * src/fitting/fitting_base.jl:55-65 -- m = models * coeffs (gemv 'N').
 * Parallel over bin blocks; each bin is summed over templates in order. */
#define ORC_COMPOSITE_BODY(MT)                                                 \
    if (nthreads < 1) nthreads = 1;                                            \
    int64_t blk = 4096;                                                        \
    int64_t nblk = (N + blk - 1) / blk;                                        \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")         \
    for (int64_t b = 0; b < nblk; ++b) {                                       \
        int64_t i0 = b * blk, i1 = i0 + blk < N ? i0 + blk : N;                \
        for (int64_t i = i0; i < i1; ++i) out[i] = 0.0;                        \
        for (int64_t j = 0; j < T; ++j) {                                      \
            const MT* col = models + j * N;                                    \
            double r = coeffs[j];                                              \
            for (int64_t i = i0; i < i1; ++i) out[i] += (double)col[i] * r;    \
        }                                                                      \
    }

ORC_API void orc_composite_f32(const float* models, const double* coeffs,
                               int64_t N, int64_t T, double* out, int nthreads)
{ ORC_COMPOSITE_BODY(float) }

ORC_API void orc_composite_f64(const double* models, const double* coeffs,
                               int64_t N, int64_t T, double* out, int nthreads)
{ ORC_COMPOSITE_BODY(double) }

/* loglikelihood(composite, data) -- src/fitting/fitting_base.jl:84-96.
 *   c_i = max(m_i, eps(T)); term = n_i > 0 ? n_i - c_i - n_i*log(n_i/c_i) : -c_i
 *   returns -typemax(T) (= -Inf for floats) when the sum is exactly zero. */
ORC_API double orc_loglikelihood(const double* composite, const double* data,
                                 int64_t N, double eps)
{
    double result = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        double ci = composite[i] > eps ? composite[i] : eps;
        double ni = data[i];
        result += (ni > 0.0) ? (ni - ci - ni * log(ni / ci)) : -ci;
    }
    return result != 0.0 ? result : -INFINITY;
}

/* ∇loglikelihood(model, composite, data) -- src/fitting/fitting_base.jl:144-160.
 *   sum_i -c_ij * (1 - n_i / max(m_i, eps)) */
ORC_API double orc_grad_single_f64(const double* model, const double* composite,
                                   const double* data, int64_t N, double eps)
{
    double result = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        double ci = composite[i] > eps ? composite[i] : eps;
        result += -model[i] * (1.0 - data[i] / ci);
    }
    return result;
}

/* ∇loglikelihood!(G, composite::Vector, models::Matrix, data::Vector)
 * src/fitting/fitting_base.jl:265-285.  Overwrites composite with
 * R_i = 1 - n_i / max(m_i, eps) (:274-280) then G = -models' * R (:283). */
#define ORC_GRAD_BODY(MT)                                                      \
    if (nthreads < 1) nthreads = 1;                                            \
    for (int64_t i = 0; i < N; ++i) {                                          \
        double ci = composite[i] > eps ? composite[i] : eps;                   \
        composite[i] = 1.0 - data[i] / ci;                                     \
    }                                                                          \
    _Pragma("omp parallel for schedule(static) num_threads(nthreads)")         \
    for (int64_t j = 0; j < T; ++j) {                                          \
        const MT* col = models + j * N;                                        \
        double acc = 0.0;                                                      \
        for (int64_t i = 0; i < N; ++i) acc += (double)col[i] * composite[i];  \
        G[j] = -acc;                                                           \
    }

ORC_API void orc_grad_inplace_f32(double* G, double* composite, const float* models,
                                  const double* data, int64_t N, int64_t T,
                                  double eps, int nthreads)
{ ORC_GRAD_BODY(float) }

ORC_API void orc_grad_inplace_f64(double* G, double* composite, const double* models,
                                  const double* data, int64_t N, int64_t T,
                                  double eps, int nthreads)
{ ORC_GRAD_BODY(double) }

/* fg!(F, G, coeffs, models, data, composite) -- src/fitting/solvers.jl:20-38.
 * want_F / G!=NULL select the branch.  Returns -logL when F is wanted (logL
 * computed before ∇loglikelihood! clobbers composite, :27-29), else 0.
 * G receives the gradient of -logL (G .*= -1, :30).  The two clamps differ in
 * type: loglikelihood uses eps(promote_type(composite, data))
 * (fitting_base.jl:86,90) -> eps; ∇loglikelihood! uses eps(eltype(composite))
 * (:277) -> eps_grad. */
ORC_API double orc_fg_f32(int want_F, double* G, const double* coeffs,
                          const float* models, const double* data,
                          double* composite, int64_t N, int64_t T, double eps,
                          double eps_grad, int nthreads)
{
    orc_composite_f32(models, coeffs, N, T, composite, nthreads);
    double negll = 0.0;
    if (want_F) negll = -orc_loglikelihood(composite, data, N, eps);
    if (G) {
        orc_grad_inplace_f32(G, composite, models, data, N, T, eps_grad, nthreads);
        for (int64_t j = 0; j < T; ++j) G[j] = -G[j];
    }
    return negll;
}

/* B independent fg! evaluations (one per walker column of X), the call pattern
 * of the HMC / ensemble callers (src/fitting/hmc_sample.jl:24-37,
 * src/fitting/mcmc_sample.jl:12-23).  negLL[b] = -logL, G[:,b] = -∇logL. */
ORC_API void orc_fg_batched_f32(const float* models, const double* data,
                                const double* X, int64_t B, int64_t N, int64_t T,
                                double eps, double eps_grad, double* negLL, double* G,
                                int nthreads)
{
    double* comp = (double*)malloc((size_t)N * sizeof(double));
    for (int64_t b = 0; b < B; ++b) {
        double nl = orc_fg_f32(1, G ? G + b * T : NULL, X + b * T, models, data,
                               comp, N, T, eps, eps_grad, nthreads);
        if (negLL) negLL[b] = nl;
    }
    free(comp);
}

/* ------------------------------------------------------------------------ */
/* Template deposition (src/StarFormationHistories.jl)                      */
/* ------------------------------------------------------------------------ */

/* A StepRangeLen of bin edges: first, step, length (number of edges). */
typedef struct { double first, step; int64_t len; } orc_range;

/* histogram_pix(d, edges::AbstractRange) -- :502 */
ORC_API double orc_histogram_pix(double d, orc_range e) { return (d - e.first) / e.step + 1.0; }
/* histogram_data(x, edges::AbstractRange) -- :518 */
ORC_API double orc_histogram_data(double x, orc_range e) { return (x - 1.0) * e.step + e.first; }

/* Julia round(Int, x) (RoundNearest, ties to even): rint in the default
 * rounding mode. */
static int64_t jl_round(double x) { return (int64_t)rint(x); }

/* Gauss-Legendre nodes / weights on (-1,1) -- :303-304 */
static const double GL_X[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
static const double GL_W[3] = {0.5555555555555556, 0.8888888888888888, 0.5555555555555556};

/* gaussian_psf_covariant -- :315-334.  Pixel-integrated covariant Gaussian:
 * exact integral in x, 3-point Gauss-Legendre in y. */
ORC_API double orc_gaussian_psf_covariant(double x, double y, double halfxstep,
                                          double halfystep, double x0, double y0,
                                          double sx, double sy, double cov_mult,
                                          double A, double B)
{
    (void)B;                       /* the reference never adds B here (:333) */
    double dx = x - x0;
    double prefac = A / 2.0 / sqrt(6.283185307179586) / sy;
    double sqrt2 = sqrt(2.0);
    double result = 0.0;
    for (int k = 0; k < 3; ++k) {
        double yv = GL_X[k] * halfystep + y;
        double yw = GL_W[k] * halfystep;
        double Dy = yv - y0;
        double Dx = dx + (Dy * cov_mult);
        double t = Dy / sy;
        result += yw * exp(-(t * t) / 2.0) *
                  (erf((Dx + halfxstep) / sqrt2 / sx) + erf((-Dx + halfxstep) / sqrt2 / sx));
    }
    return result * prefac;
}

/* gaussian_int_general (Float64 method) -- :198-205 */
ORC_API double orc_gaussian_int_general(double dx, double dy, double hx, double hy,
                                        double sx, double sy, double A, double B)
{
    double sqrt2 = sqrt(2.0);
    return A / 4.0 * (erf((dx - hx) / (sqrt2 * sx)) - erf((dx + hx) / (sqrt2 * sx))) *
           (erf((dy - hy) / (sqrt2 * sy)) - erf((dy + hy) / (sqrt2 * sy))) + B;
}

/* addstar!(image::Histogram, obj::RealSpaceKernel, cutout_size) -- :366-408,
 * restricted to output rows [jlo, jhi] (0-based) so that row-parallel
 * callers stay deterministic.  Returns the number of (point, bin) pairs the
 * full footprint evaluates (independent of the row restriction when
 * jlo=0, jhi=ny-1). */
static int64_t addstar_covariant_rows(double* mat, orc_range ex, orc_range ey,
                                      double x0, double y0, double sx, double sy,
                                      double cov_mult, double A,
                                      int64_t jlo, int64_t jhi)
{
    int64_t nx = ex.len - 1, ny = ey.len - 1;
    double cutx = 15.0 * sx, cuty = 10.0 * sy;          /* size(obj), :293 */
    int64_t xp = jl_round(orc_histogram_pix(x0, ex));   /* :378-379 */
    int64_t yp = jl_round(orc_histogram_pix(y0, ey));
    int64_t xo = jl_round(cutx / ex.step / 2.0); if (xo < 1) xo = 1;   /* :381 */
    int64_t yo = jl_round(cuty / ey.step / 2.0); if (yo < 1) yo = 1;   /* :382 */
    int64_t xa = xp - xo > 1 ? xp - xo : 1, xb = xp + xo < nx ? xp + xo : nx;  /* :384 */
    int64_t ya = yp - yo > 1 ? yp - yo : 1, yb = yp + yo < ny ? yp + yo : ny;  /* :385 */
    int64_t lx = xb - xa + 1, ly = yb - ya + 1;
    if (lx <= 1 || ly <= 1) return 0;                   /* :401 length>1 guard */
    /* xriter / yriter: range(histogram_data(first+1/2), histogram_data(last+1/2);
     * length) -- :390-395 */
    double xs = orc_histogram_data((double)xa + 0.5, ex);
    double xe = orc_histogram_data((double)xb + 0.5, ex);
    double ys = orc_histogram_data((double)ya + 0.5, ey);
    double ye = orc_histogram_data((double)yb + 0.5, ey);
    double xst = (xe - xs) / (double)(lx - 1), yst = (ye - ys) / (double)(ly - 1);
    double hx = xst / 2.0, hy = yst / 2.0;
    for (int64_t j = 0; j < ly; ++j) {
        int64_t yind = ya + j;                          /* 1-based */
        if (yind - 1 < jlo || yind - 1 > jhi) continue;
        double yr = ys + (double)j * yst;
        for (int64_t i = 0; i < lx; ++i) {
            int64_t xind = xa + i;
            double xr = xs + (double)i * xst;
            mat[(xind - 1) + nx * (yind - 1)] +=
                orc_gaussian_psf_covariant(xr, yr, hx, hy, x0, y0, sx, sy, cov_mult, A, 0.0);
        }
    }
    return lx * ly;
}

/* addstar!(image::AbstractMatrix, obj::PixelSpaceKernel, cutout_size) -- :348-364
 * with GaussianPSFAsymmetric (:223-265); x0,y0,sx,sy in pixel units. */
static int64_t addstar_asymmetric_rows(double* mat, int64_t nx, int64_t ny,
                                       double x0, double y0, double sx, double sy,
                                       double A, int64_t jlo, int64_t jhi)
{
    int64_t cx = (int64_t)ceil(sx * 10.0), cy = (int64_t)ceil(sy * 10.0);  /* :247 */
    int64_t x = jl_round(x0), y = jl_round(y0);                               /* :350 */
    int64_t xo = cx / 2; if (xo < 1) xo = 1;                                  /* :351 */
    int64_t yo = cy / 2; if (yo < 1) yo = 1;                                  /* :352 */
    int64_t xa = x - xo > 1 ? x - xo : 1, xb = x + xo < nx ? x + xo : nx;     /* :356 */
    int64_t ya = y - yo > 1 ? y - yo : 1, yb = y + yo < ny ? y + yo : ny;     /* :357 */
    int64_t lx = xb - xa + 1, ly = yb - ya + 1;
    if (lx <= 1 || ly <= 1) return 0;                                         /* :359 */
    for (int64_t j = ya; j <= yb; ++j) {
        if (j - 1 < jlo || j - 1 > jhi) continue;
        for (int64_t i = xa; i <= xb; ++i)
            mat[(i - 1) + nx * (j - 1)] +=
                orc_gaussian_int_general((double)i + 0.5 - x0, (double)j + 0.5 - y0,
                                         0.5, 0.5, sx, sy, A, 0.0);           /* :361 */
    }
    return lx * ly;
}

ORC_API int64_t orc_addstar_covariant(double* mat, orc_range ex, orc_range ey,
                                      double x0, double y0, double sx, double sy,
                                      double cov_mult, double A)
{
    return addstar_covariant_rows(mat, ex, ey, x0, y0, sx, sy, cov_mult, A, 0, ey.len - 2);
}

ORC_API int64_t orc_addstar_asymmetric(double* mat, int64_t nx, int64_t ny, double x0,
                                       double y0, double sx, double sy, double A)
{
    return addstar_asymmetric_rows(mat, nx, ny, x0, y0, sx, sy, A, 0, ny - 1);
}

/* bin_cmd_smooth(colors, mags, color_err, mag_err, cov_mult; weights, edges)
 * -- :574-621.  cov_mult == 0 uses GaussianPSFAsymmetric in pixel space
 * (:586-610); otherwise GaussianPSFCovariant in real space (:611-618).
 * cmult (nullable) supplies a per-point covariance multiplier instead of the
 * scalar cov_mult (the reference admits only the scalar, :579).
 * out must hold (ex.len-1)*(ey.len-1) doubles; it is zeroed here (:584).
 * Threads split output rows; every bin still sums its points in point order.
 * Returns the number of (point, bin) pairs evaluated. */
ORC_API int64_t orc_bin_cmd_smooth(const double* colors, const double* mags,
                                   const double* color_err, const double* mag_err,
                                   const double* weights, const double* cmult,
                                   int64_t P, double cov_mult, orc_range ex,
                                   orc_range ey, double* out, int nthreads)
{
    int64_t nx = ex.len - 1, ny = ey.len - 1;
    memset(out, 0, (size_t)(nx * ny) * sizeof(double));
    if (nthreads < 1) nthreads = 1;
    int64_t pairs = 0;
    int64_t rows_per = (ny + nthreads - 1) / nthreads;
    #pragma omp parallel for schedule(static, 1) num_threads(nthreads) reduction(+:pairs)
    for (int t = 0; t < nthreads; ++t) {
        int64_t jlo = (int64_t)t * rows_per, jhi = jlo + rows_per - 1;
        if (jhi > ny - 1) jhi = ny - 1;
        if (jlo > jhi) continue;
        int64_t cnt = 0;
        for (int64_t p = 0; p < P; ++p) {
            double c = cmult ? cmult[p] : cov_mult;
            if (c == 0.0 && !cmult) {
                double xw = ex.step, yw = ey.step;                            /* :598 */
                double x0 = orc_histogram_pix(colors[p], ex);                 /* :601 */
                double y0 = orc_histogram_pix(mags[p], ey);
                double sx = color_err[p] / xw, sy = mag_err[p] / yw;          /* :603-604 */
                cnt += addstar_asymmetric_rows(out, nx, ny, x0, y0, sx, sy, weights[p], jlo, jhi);
            } else {
                cnt += addstar_covariant_rows(out, ex, ey, colors[p], mags[p], color_err[p],
                                              mag_err[p], c, weights[p], jlo, jhi);
            }
        }
        if (t == 0) pairs += cnt;   /* every thread sees the same footprints */
    }
    return pairs;
}

/* Pair count only (the work metric of the deposition path). */
ORC_API int64_t orc_count_pairs(const double* colors, const double* mags,
                                const double* color_err, const double* mag_err,
                                int64_t P, double cov_mult, orc_range ex, orc_range ey)
{
    int64_t nx = ex.len - 1, ny = ey.len - 1, pairs = 0;
    for (int64_t p = 0; p < P; ++p) {
        int64_t lx, ly;
        if (cov_mult == 0.0) {
            double x0 = orc_histogram_pix(colors[p], ex), y0 = orc_histogram_pix(mags[p], ey);
            double sx = color_err[p] / ex.step, sy = mag_err[p] / ey.step;
            int64_t cx = (int64_t)ceil(sx * 10.0), cy = (int64_t)ceil(sy * 10.0);
            int64_t x = jl_round(x0), y = jl_round(y0);
            int64_t xo = cx / 2 < 1 ? 1 : cx / 2, yo = cy / 2 < 1 ? 1 : cy / 2;
            int64_t xa = x - xo > 1 ? x - xo : 1, xb = x + xo < nx ? x + xo : nx;
            int64_t ya = y - yo > 1 ? y - yo : 1, yb = y + yo < ny ? y + yo : ny;
            lx = xb - xa + 1; ly = yb - ya + 1;
        } else {
            int64_t xp = jl_round(orc_histogram_pix(colors[p], ex));
            int64_t yp = jl_round(orc_histogram_pix(mags[p], ey));
            int64_t xo = jl_round(15.0 * color_err[p] / ex.step / 2.0); if (xo < 1) xo = 1;
            int64_t yo = jl_round(10.0 * mag_err[p] / ey.step / 2.0); if (yo < 1) yo = 1;
            int64_t xa = xp - xo > 1 ? xp - xo : 1, xb = xp + xo < nx ? xp + xo : nx;
            int64_t ya = yp - yo > 1 ? yp - yo : 1, yb = yp + yo < ny ? yp + yo : ny;
            lx = xb - xa + 1; ly = yb - ya + 1;
        }
        if (lx > 1 && ly > 1) pairs += lx * ly;
    }
    return pairs;
}

ORC_API int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * sfh.h -- C ABI of libsfh_hip.so, the MI355X (gfx950) implementation of the
 * two data-parallel paths of StarFormationHistories.jl:
 *
 *   P2  template-mixture likelihood: composite!, loglikelihood,
 *       ∇loglikelihood(!), fg! -- plus a batched form over many coefficient
 *       vectors (HMC / ensemble walkers) on the matrix cores;
 *   P1  smoothed Hess-diagram template deposition: bin_cmd_smooth (the kernel
 *       loop of partial_cmd_smooth), single and batched over templates.
 *
 * Plain C: pointers, sizes, status codes.  Every function returns SFH_OK (0)
 * or a negative SFH_ERR_* code; sfh_last_error(ctx) then describes it.  The
 * reference throws ArgumentError from @argcheck on shape mismatches; those
 * cases return SFH_ERR_ARGUMENT here.
 *
 * Pointer conventions
 *   - Functions without a _device suffix take HOST pointers, copy what they
 *     need, and return when the result is in host memory.
 *   - *_device functions take DEVICE pointers on the context's device, enqueue
 *     work on the context's stream and return without synchronising.
 *
 * Layouts (identical to the reference's flattened layout)
 *   models   : nbins x ntemplates, column-major, one template per column
 *              (stack_models, src/fitting/utilities.jl:12-13).  FP32.
 *   X        : ntemplates x B, column-major, one coefficient vector per column.
 *   Hess mat : nx x ny, column-major (colour bin fastest), nx = length of the
 *              colour edge range - 1 (src/StarFormationHistories.jl:584).
 *
 * Threading: calls on one context are serialised by a mutex; distinct
 * contexts are independent (one stream each), mirroring the reference's
 * per-task buffers (src/fitting/hmc_sample.jl:127).
 */
#ifndef SFH_H
#define SFH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFH_OK              0
#define SFH_ERR_ARGUMENT   -1   /* shape / value checks (@argcheck -> ArgumentError) */
#define SFH_ERR_DEVICE     -2   /* HIP runtime or launch failure */
#define SFH_ERR_ALLOC      -3   /* device or pinned allocation failed */
#define SFH_ERR_DOMAIN     -4   /* input outside the reference's domain (e.g. non-finite
                                   kernel centre: Julia's round(Int, NaN) throws InexactError) */

typedef struct sfh_ctx sfh_ctx;
typedef struct sfh_templates sfh_templates;
typedef struct sfh_data sfh_data;

/* Uniform bin edges: `first`, `step`, `length` edges -- the
 * Tuple{AbstractRange,AbstractRange} of `edges` in bin_cmd_smooth /
 * partial_cmd_smooth (src/StarFormationHistories.jl:451-476, :372-374). */
typedef struct {
    double first;
    double step;
    int64_t length;
} sfh_range;

/* Batched precision ladder (sfh_fg_batched*). */
#define SFH_PREC_FP32    0  /* FP32 templates x FP32 coefficients, FP32 MFMA */
#define SFH_PREC_BF16X2  1  /* split-BF16 hi+lo on both operands (3 products), FP32 accumulation */
#define SFH_PREC_F64     2  /* composite on the FP64 matrix cores (IEEE FP64 accumulation),
                               gradient split-BF16 hi+lo: log-likelihood at FP64 accuracy */

/* ---------------------------------------------------------------- context */

/* No reference counterpart (the reference is one Julia process).  `stream` is
 * a hipStream_t to enqueue on, or NULL to create a private stream. */
int sfh_ctx_create(int device, void* stream, sfh_ctx** out);
int sfh_ctx_destroy(sfh_ctx* ctx);
const char* sfh_last_error(const sfh_ctx* ctx);
int sfh_ctx_synchronize(sfh_ctx* ctx);
/* Per-kernel timing with HIP events on the context's stream. */
int sfh_ctx_set_timing(sfh_ctx* ctx, int enable);
/* No reference counterpart.  Capture the context's stream into a HIP graph:
 * every call between begin and end is recorded instead of run (calls that
 * allocate or wait invalidate the capture, so run them once beforehand);
 * sfh_graph_launch replays the graph `times` times on the stream. */
typedef struct sfh_graph sfh_graph;
int sfh_ctx_capture_begin(sfh_ctx* ctx);
int sfh_ctx_capture_end(sfh_ctx* ctx, sfh_graph** out);
int sfh_graph_launch(sfh_graph* graph, int64_t times);
int sfh_graph_destroy(sfh_graph* graph);
int sfh_ctx_kernel_stats(sfh_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches);
int sfh_ctx_reset_stats(sfh_ctx* ctx);
int sfh_version(void);
/* No reference counterpart: a measurement helper for the deposition's
 * roofline.  Runs the row strip's instruction mix (per step one FMA reading a
 * scalar register and one multiply of two vector registers, 8 independent
 * chains) at four waves per SIMD on every CU for `iters` iterations
 * (1 .. 2^24) and returns the VALU lane-instructions per second it reached. */
int sfh_probe_valu_issue(sfh_ctx* ctx, int64_t iters, double* lane_instr_per_s);

/* ------------------------------------------------------------- templates */

/* Upload models (nbins x ntemplates, column-major, FP32) -- the flattened
 * `models::AbstractMatrix` argument of composite!/∇loglikelihood!/fg!
 * (src/fitting/fitting_base.jl:55-65,265-285; src/fitting/solvers.jl:20). */
int sfh_templates_create(sfh_ctx* ctx, const float* models, int64_t nbins,
                         int64_t ntemplates, sfh_templates** out);
/* Same from a device pointer (e.g. templates written by the deposition kernel). */
int sfh_templates_create_device(sfh_ctx* ctx, const float* d_models, int64_t nbins,
                                int64_t ntemplates, sfh_templates** out);
int sfh_templates_destroy(sfh_templates* tpl);
int sfh_templates_shape(const sfh_templates* tpl, int64_t* nbins, int64_t* ntemplates);
/* Device pointer of the resident FP32 models (nbins x ntemplates). */
const float* sfh_templates_device_ptr(const sfh_templates* tpl);

/* Observed Hess diagram n_i (any real counts), FP64. */
int sfh_data_create(sfh_ctx* ctx, const double* data, int64_t nbins, sfh_data** out);
int sfh_data_create_device(sfh_ctx* ctx, const double* d_data, int64_t nbins, sfh_data** out);
int sfh_data_destroy(sfh_data* data);

/* ---------------------------------------------- P2: one coefficient vector */

/* composite!(composite::AbstractVector, coeffs, models::AbstractMatrix)
 * -- src/fitting/fitting_base.jl:55-65.  composite = models * coeffs. */
int sfh_composite(sfh_ctx* ctx, const sfh_templates* tpl, const double* coeffs,
                  double* composite);
int sfh_composite_device(sfh_ctx* ctx, const sfh_templates* tpl, const double* d_coeffs,
                         double* d_composite);

/* loglikelihood(composite, data) -- src/fitting/fitting_base.jl:84-96.
 * `eps` is eps(promote_type(eltype(composite), eltype(data))), the clamp
 * max(m_i, eps); the result is -Inf when the sum is exactly 0 (:95). */
int sfh_loglikelihood(sfh_ctx* ctx, const double* composite, const double* data,
                      int64_t n, double eps, double* out);

/* loglikelihood(coeffs, models::AbstractMatrix, data::AbstractVector)
 * -- src/fitting/fitting_base.jl:116-125. */
int sfh_loglikelihood_coeffs(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                             const double* coeffs, double eps, double* out);

/* ∇loglikelihood(models::AbstractMatrix, composite, data)
 * -- src/fitting/fitting_base.jl:177-182 (composite left unchanged).  `eps`
 * is eps(promote_type(eltype(models), eltype(composite), eltype(data))) (:148). */
int sfh_grad_loglikelihood(sfh_ctx* ctx, const sfh_templates* tpl, const double* composite,
                           const sfh_data* data, double eps, double* G);

/* ∇loglikelihood!(G, composite::AbstractVector, models::AbstractMatrix, data)
 * -- src/fitting/fitting_base.jl:265-285.  Overwrites composite with
 * 1 - n_i/max(m_i, eps) exactly as the reference does (:274-280); `eps` is
 * eps(eltype(composite)) (:277). */
int sfh_grad_loglikelihood_inplace(sfh_ctx* ctx, const sfh_templates* tpl, double* composite,
                                   const sfh_data* data, double eps, double* G);

/* fg!(F, G, coeffs, models, data, composite) -- src/fitting/solvers.jl:20-38.
 * want_F selects the objective; G may be NULL (gradient not wanted).  On
 * return *negLL = -logL (when want_F), G = -∇logL, and composite (if not
 * NULL) holds what the reference leaves there: 1 - n/max(m,eps_grad) when G
 * was requested, else m.  fg! calls loglikelihood(composite, data), whose
 * clamp is eps(promote_type(eltype(composite), eltype(data)))
 * (fitting_base.jl:86) -> `eps`, and ∇loglikelihood!, whose clamp is
 * eps(eltype(composite)) (:277) -> `eps_grad`. */
int sfh_fg(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
           const double* coeffs, int want_F, double eps, double eps_grad, double* negLL,
           double* G, double* composite);
/* Device-resident form: d_negLL (1 double), d_G (ntemplates) and
 * d_composite (nbins) may each be NULL. */
int sfh_fg_device(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                  const double* d_coeffs, int want_F, double eps, double eps_grad,
                  double* d_negLL, double* d_G, double* d_composite);

/* No counterpart in the reference: sfh_fg_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same kernels and results, bit for
 * bit, with the partial sums in a static device buffer of
 * SFH_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) instead of
 * context scratch: no call allocates or frees anything.  Limits: 1..704
 * templates (the single-pass kernel; SFH_ERR_ARGUMENT otherwise) and
 * partials (one LL value + ntemplates gradient values per work item, two
 * items per compute unit) within the buffer (SFH_ERR_ARGUMENT otherwise).
 * Every call on one GPU shares that GPU's buffer: calls that overlap in time
 * on one GPU (several contexts or streams) race. */
#define SFH_NOALLOC_SCRATCH_BYTES (4 * 1024 * 1024)
int sfh_fg_device_noalloc(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                          const double* d_coeffs, int want_F, double eps, double eps_grad,
                          double* d_negLL, double* d_G, double* d_composite);

/* --------------------------------------- P2: many coefficient vectors (new) */

/* B independent fg! evaluations, one per column of X (ntemplates x B), the
 * call pattern of LogDensityProblems.logdensity_and_gradient(::HMCModel)
 * (src/fitting/hmc_sample.jl:24-37) and MCMCModel (src/fitting/mcmc_sample.jl:12-23)
 * over many walkers.  negLL[b] = -logL(X[:,b]); G[:,b] = -∇logL (G may be
 * NULL).  The final log-likelihood reduction is FP64 in every precision.
 * eps / eps_grad: the two clamps of sfh_fg. */
int sfh_fg_batched(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                   const double* X, int64_t B, int precision, double eps, double eps_grad,
                   double* negLL, double* G);
int sfh_fg_batched_device(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                          const double* d_X, int64_t B, int precision, double eps,
                          double eps_grad, double* d_negLL, double* d_G);

/* ------------------------------- P2: many independent fits of uneven size (new) */

/* A group of independent fits (fit f: templates[f], data[f]; any mix of
 * template and bin counts, each fit at most 1024 templates), evaluated in one
 * persistent launch per fg! of every fit and optimised on the device.  The
 * template and data handles stay owned by the caller and must outlive the
 * group; the group keeps its own copy of the templates in a blocked layout
 * (32-bin blocks contiguous over all templates), as much device memory again.  Vectors of all fits are concatenated: fit f's ntemplates values
 * start at offset sum_{g<f} ntemplates(g); its ntemplates^2 matrix at
 * sum_{g<f} ntemplates(g)^2. */
typedef struct sfh_multifit sfh_multifit;

int sfh_multifit_create(sfh_ctx* ctx, int64_t nfits, const sfh_templates* const* templates,
                        const sfh_data* const* data, sfh_multifit** out);
int sfh_multifit_destroy(sfh_multifit* mf);

/* Work decomposition of the grouped fg! launch: about item_elems template
 * elements per work item (whole 32-bin blocks of one fit), workgroups of
 * 64*waves threads (waves in {4, 8, 16}: 4 for fits of up to 1024 templates,
 * 8 up to 704, 16 up to 512; a group whose largest fit exceeds 704 starts at 4)
 * and wg_per_cu persistent workgroups per compute unit. */
int sfh_multifit_get_shape(const sfh_multifit* mf, int64_t* item_elems, int* waves, int* wg_per_cu);
int sfh_multifit_set_shape(sfh_multifit* mf, int64_t item_elems, int waves, int wg_per_cu);
/* Times every candidate shape on this group's fits (reps launches each, at
 * x = ones) and keeps the fastest; returns it (any pointer may be NULL). */
int sfh_multifit_autotune(sfh_multifit* mf, int reps, int64_t* item_elems, int* waves, int* wg_per_cu);

/* fg!(true, G, coeffs_f, models_f, data_f, composite) of src/fitting/solvers.jl:20-38
 * for every fit f in one launch: negLL[f] = -logL_f, G (concatenated, may be
 * NULL) = -∇logL_f.  X: concatenated coefficient vectors.  eps / eps_grad:
 * the two clamps of sfh_fg. */
int sfh_multifit_fg(sfh_multifit* mf, const double* X, double eps, double eps_grad, double* negLL,
                    double* G);
int sfh_multifit_fg_device(sfh_multifit* mf, const double* d_X, double eps, double eps_grad,
                           double* d_negLL, double* d_G);

#define SFH_FIT_LOG 0    /* fit_templates: log(coeffs), MAP (Jacobian -sum log x) then MLE */
#define SFH_FIT_SQRT 1   /* fit_templates_fast: coeffs = theta^2, MLE only */
#define SFH_OPT_BFGS 0   /* dense inverse-Hessian BFGS (Optim.BFGS; inverse Hessian returned) */
#define SFH_OPT_LBFGS 1  /* limited-memory BFGS */

#define SFH_FIT_RUNNING 0
#define SFH_FIT_CONVERGED 1      /* max |gradient| <= g_abstol */
#define SFH_FIT_MAX_ITERATIONS 2
#define SFH_FIT_LINESEARCH_FAILED 3

typedef struct {
    int transform;           /* SFH_FIT_LOG or SFH_FIT_SQRT */
    int method;              /* SFH_OPT_BFGS or SFH_OPT_LBFGS */
    int memory;              /* L-BFGS pairs kept (1..32) */
    int steps_per_graph;     /* optimiser steps per captured HIP graph (1..256) */
    double g_abstol;         /* Optim.Options g_abstol (default 1e-8), on the transformed variables */
    int64_t max_iterations;  /* per stage (Optim.Options iterations, default 1000) */
    double eps;              /* the log-likelihood's clamp, eps(promote_type(models, data));
                                eps(Float64) by default */
    double eps_grad;         /* the gradient's clamp, eps(eltype(composite)) = eps(eltype(models))
                                in fit_templates (solvers.jl:173); eps(Float64) by default */
} sfh_fit_options;

/* Fills *opt with the defaults above (LOG, BFGS, memory 10, 16 steps per graph). */
int sfh_fit_options_default(sfh_fit_options* opt);

/* Per-fit results, concatenated as described above; any pointer may be NULL.
 * Stage 0 is the MAP stage of SFH_FIT_LOG, stage 1 the MLE (the only stage of
 * SFH_FIT_SQRT).  fmin, iterations, evaluations and status hold 2*nfits
 * values, [stage * nfits + f]. */
typedef struct {
    double* mu_map;          /* exp(theta) at the MAP optimum (LOG) */
    double* mu_mle;          /* coefficients at the MLE optimum */
    double* invH_map;        /* BFGS inverse Hessian in the transformed variables, MAP stage */
    double* invH_mle;        /* ... MLE stage */
    double* fmin;            /* objective at the optimum: -logL - sum(theta) (MAP), -logL (MLE) */
    int64_t* iterations;
    int64_t* evaluations;    /* fg! evaluations */
    int32_t* status;         /* SFH_FIT_* */
    int64_t steps;           /* out: optimiser steps run (one grouped fg! each) */
} sfh_fit_results;

/* fit_templates (src/fitting/solvers.jl:172-221) with SFH_FIT_LOG, or
 * fit_templates_fast (:248-275) with SFH_FIT_SQRT, for every fit of the group
 * at once: x0 renormalised as renormalize_x0 does (src/fitting/utilities.jl:104-115;
 * x0 NULL means ones), then quasi-Newton iterations whose line search accepts
 * a step on LineSearches.HagerZhang's Wolfe / approximate-Wolfe conditions
 * (delta 0.1, sigma 0.9, epsilon 1e-6; bracketing with safeguarded cubic
 * steps) from a first trial step of length min(1, |d|)
 * (LineSearches.InitialStatic(1.0, true)), every fit's fg! evaluated in one
 * launch per step, steps captured in HIP graphs and replayed until every fit
 * has stopped. */
int sfh_multifit_fit(sfh_multifit* mf, const double* x0, const sfh_fit_options* opt,
                     sfh_fit_results* res);

/* --------------------------------------------------- P1: template deposition */

/* A function tabulated on a uniform grid: values[k] = f(first + k*step),
 * 0 <= k < length; read back by linear interpolation, clamped at the ends. */
typedef struct {
    double first;
    double step;
    int64_t length;
    const double* values;
} sfh_table;

/* The pair histogram of binary_hess(::RandomBinaryPairs, m_ini, mags, ...)
 * -- src/StarFormationHistories.jl:627-682, for y_index in color_indices.
 * For the npoints-point (downsampled) isochrone, every pair of intervals
 * i < j: magnitudes flux2mag(mag2flux(mag[i]) + mag2flux(mag[j])) in the two
 * colour filters (a = first(color_indices), b = last), weight
 * (m_{i+1}-m_i)(m_{j+1}-m_j)(imf_i+imf_{i+1})(imf_j+imf_{j+1}) * prefac
 * times compl_a(mag_a) * compl_b(mag_b) (prefac = normalize_value /
 * mean_mass / 4 / trapz(m_ini, imf)), measured magnitudes mag + bias(mag),
 * binned (left-closed) at colour a - b and y = a (y_is_a) or b.  The
 * completeness and bias callables arrive tabulated (sfh_table) over the
 * pairs' magnitude range.  out: the FP64 nx x ny histogram (colour fastest),
 * accumulated deterministically. */
int sfh_binary_pairs_hist(sfh_ctx* ctx, int64_t npoints, const double* mini, const double* imf_pdf,
                          const double* mag_a, const double* mag_b, int y_is_a, double prefac,
                          const sfh_table* compl_a, const sfh_table* compl_b, const sfh_table* bias_a,
                          const sfh_table* bias_b, sfh_range xedges, sfh_range yedges, double* out);
/* The same pair histograms for many isochrones in one launch (new API):
 * template t owns points offsets[t] .. offsets[t+1]-1 of the concatenated
 * mini / imf_pdf / mag_a / mag_b and the weight factor prefac[t]; the tables
 * are shared.  out: ntemplates x (nx ny) FP64, template-major. */
int sfh_binary_pairs_hist_batch(sfh_ctx* ctx, int64_t ntemplates, const int64_t* offsets, const double* mini,
                                const double* imf_pdf, const double* mag_a, const double* mag_b, int y_is_a,
                                const double* prefac, const sfh_table* compl_a, const sfh_table* compl_b,
                                const sfh_table* bias_a, const sfh_table* bias_b, sfh_range xedges, sfh_range yedges,
                                double* out);

/* bin_cmd_smooth(colors, mags, color_err, mag_err, cov_mult; weights, edges)
 * -- src/StarFormationHistories.jl:574-621.  cov_mult in {-1,0,1} as the
 * reference checks (:579): 0 deposits GaussianPSFAsymmetric in pixel space
 * (:586-610, :191-265, :348-364); +-1 deposits GaussianPSFCovariant in real
 * space (:611-618, :272-338, :366-408).  out receives the FP64 nx x ny matrix. */
int sfh_bin_cmd_smooth(sfh_ctx* ctx, const double* colors, const double* mags,
                       const double* color_err, const double* mag_err,
                       const double* weights, int64_t npoints, int cov_mult,
                       sfh_range xedges, sfh_range yedges, double* out);

/* Many templates in one launch (no reference counterpart: the reference
 * calls partial_cmd_smooth once per template).  Point arrays are the
 * concatenation of every template's points; d_offsets (ntemplates+1 int64,
 * device) delimits them.  d_cmult, if not NULL, gives a real-valued
 * covariance multiplier per point (the covariant formula with any c);
 * otherwise cov_mult applies to all points.  Writes FP32 templates into
 * d_out, column t at d_out + t*ld_out (ld_out >= nx*ny), the layout
 * sfh_templates_create_device takes.  *pairs_out (may be NULL) receives the
 * number of (point, bin) pairs evaluated (synchronises). */
int sfh_bin_cmd_smooth_batch_device(sfh_ctx* ctx, const double* d_colors, const double* d_mags,
                                    const double* d_color_err, const double* d_mag_err,
                                    const double* d_weights, const double* d_cmult,
                                    const int64_t* d_offsets, int64_t ntemplates,
                                    double cov_mult, sfh_range xedges, sfh_range yedges,
                                    float* d_out, int64_t ld_out, int64_t* pairs_out);

/* ------------------------------------------------- mock observations */

#define SFH_MOCK_MAX_FILTERS 8
#define SFH_MOCK_MAX_IMF_SEGMENTS 8

/* A piecewise power-law IMF: pdf(m) proportional to coefs[k] m^-slopes[k] on
 * [breaks[k], breaks[k+1]), k < nseg (Salpeter1955: one segment; Kroupa2001:
 * two, slopes 1.3 / 2.3, coefficients 1 / 0.5 around 0.5 Msun). */
typedef struct sfh_imf {
    int32_t nseg;
    const double* breaks;   /* nseg + 1, increasing */
    const double* slopes;   /* nseg */
    const double* coefs;    /* nseg */
} sfh_imf;

/* kind 0 NoBinaries, 1 RandomBinaryPairs(fraction), 2 BinaryMassRatio(fraction,
 * Uniform(qmin, qmax)) -- src/simulate.jl:153-259. */
typedef struct sfh_binary_model {
    int32_t kind;
    double fraction, qmin, qmax;
} sfh_binary_model;

/* A device-resident star catalogue: n entries of nstars masses and nfilters
 * magnitudes, with the index of the system (or input star) each came from. */
typedef struct sfh_catalog sfh_catalog;

/* generate_stars_mass (mode 0) / generate_stars_mag (mode 1) --
 * src/simulate.jl:306-339 / :350-404.  The isochrone arrives as the reference
 * prepares it before sampling (sort_ingested :54-63, distance modulus added):
 * mini ascending without duplicates, mags npoints x nfilters row-major.  Mode
 * 0 draws systems until their birth masses reach `limit` and returns those
 * whose primary mass is >= mmin_cut (mass_limits :92-123, host side) with
 * positive flux; mode 1 draws until the flux in limit_filter reaches `limit`
 * (mag2flux of the apparent limit magnitude) and returns systems with positive
 * flux that are brighter than mag_lim in mag_lim_filter (mag_lim = +inf: all).
 * System i draws from Philox4x64-10(counter (i, 0), key (seed, stream)); the
 * running total is kept in 64-bit fixed point (2^62 / limit), so the sample
 * ends at the first system that brings it to the limit whatever the summation
 * order.  *out receives a new catalogue (system order). */
int sfh_generate_stars(sfh_ctx* ctx, int64_t npoints, const double* mini, int32_t nfilters, const double* mags,
                       int32_t mode, double limit, int32_t limit_filter, double mmin_cut, double mag_lim,
                       int32_t mag_lim_filter, const sfh_imf* imf, const sfh_binary_model* binary, uint64_t seed,
                       uint64_t stream, sfh_catalog** out);
/* A catalogue of nfilters magnitudes per star from host memory (row-major). */
int sfh_catalog_from_host(sfh_ctx* ctx, int64_t n, int32_t nfilters, const double* mags, sfh_catalog** out);
int sfh_catalog_shape(const sfh_catalog* c, int64_t* n, int32_t* nstars, int32_t* nfilters, int64_t* sampled);
/* Copies masses (n x nstars), mags (n x nfilters) and source indices (n) to
 * host memory; any pointer may be NULL. */
int sfh_catalog_copy(const sfh_catalog* c, double* masses, double* mags, int64_t* index);
/* Per-filter minimum and maximum of the catalogue's finite magnitudes (lo,
 * hi: nfilters each; +inf / -inf when there are none). */
int sfh_catalog_range(const sfh_catalog* c, double* lo, double* hi);
int sfh_catalog_destroy(sfh_catalog* c);
/* model_cmd(mags, errfuncs, completefuncs, biasfuncs) -- src/simulate.jl:621-674.
 * Star j is kept when a uniform number (Philox block (j, 0), key (seed,
 * stream), word 0) is <= the product of the filters' completeness; kept
 * magnitudes become mag + z err(mag) + bias(mag), z standard normal (Box-
 * Muller).  err, completeness (and bias, may be NULL) are nfilters tables
 * each, interpolated linearly.  *out: the observed catalogue; its indices are
 * the kept input stars. */
int sfh_model_cmd(sfh_ctx* ctx, const sfh_catalog* in, const sfh_table* err, const sfh_table* completeness,
                  const sfh_table* bias, uint64_t seed, uint64_t stream, sfh_catalog** out);
/* bin_cmd(colors, mags; weights, edges) -- src/StarFormationHistories.jl:544-553:
 * the left-closed weighted 2-D histogram (nx x ny, colour fastest), from host
 * arrays (weights may be NULL: unit weights, exact counts) or from the
 * catalogue's colour (filter a - filter b) and y filter. */
int sfh_bin_cmd(sfh_ctx* ctx, int64_t n, const double* colors, const double* mags, const double* weights,
                sfh_range xedges, sfh_range yedges, double* out);
int sfh_bin_cmd_catalog(sfh_ctx* ctx, const sfh_catalog* c, int32_t color_a, int32_t color_b, int32_t y_filter,
                        sfh_range xedges, sfh_range yedges, double* out);

/* ------------------------------ P1: partial_cmd_smooth, prepared on the device */

/* A scalar callable of the template path -- a magnitude -> error, completeness
 * or bias function, or the IMF's mass -> pdf -- in a form the device
 * evaluates (FP64):
 *   SFH_FN_CONST        p[0]
 *   SFH_FN_EXP_PHOTERR  exp_photerr(m, a=p[0], b=p[1], c=p[2], d=p[3]) = a^(b (m - c)) + d
 *                       (src/utilities.jl:218)
 *   SFH_FN_MARTIN2016   Martin2016_complete(m, A=p[0], m50=p[1], rho=p[2]) =
 *                       A / (1 + exp((m - m50) / rho))  (src/utilities.jl:205)
 *   SFH_FN_TABLE        `table`, linearly interpolated and clamped at its ends
 *                       (any other callable, tabulated by the caller)
 *   SFH_FN_POWERLAW     p[0] coefs[k] m^-slopes[k] on [breaks[k], breaks[k+1])
 *                       of `powerlaw`, the last segment closed, 0 outside (an IMF pdf
 *                       with normalisation p[0], e.g. Kroupa2001, Salpeter1955) */
#define SFH_FN_CONST 0
#define SFH_FN_EXP_PHOTERR 1
#define SFH_FN_MARTIN2016 2
#define SFH_FN_TABLE 3
#define SFH_FN_POWERLAW 4
typedef struct sfh_fn {
    int32_t kind;
    double p[4];
    sfh_table table;
    sfh_imf powerlaw;
} sfh_fn;

/* partial_cmd_smooth(m_ini, mags, mag_err_funcs, y_index, color_indices, imf,
 * completeness_funcs, bias_funcs; dmod, normalize_value, mean_mass, edges)
 * -- src/StarFormationHistories.jl:829-893 (binary_model = NoBinaries()) -- for
 * ntemplates isochrones in one call, from host memory: isochrone t owns points
 * offsets[t] .. offsets[t+1]-1 of m_ini and of each filter's magnitude array
 * mags[f] (nfilters host pointers), and has its own dmod, normalize_value and
 * mean_mass.  y_index, color_a, color_b are 0-based filter indices (the
 * reference's y_index and color_indices).  mag_err holds nfilters callables;
 * completeness / bias hold nfilters callables or are NULL (one / zero, the
 * reference's defaults).  The host preparation -- mini_spacing (:110-148),
 * interpolate_mini (:77-81), the callables, calculate_weights (:765-776),
 * midpoints (:417-431) -- runs on the device in FP64, then the kernels are
 * deposited as sfh_bin_cmd_smooth does (FP32 templates).  m_ini must be
 * strictly increasing per isochrone (sort_ingested is the caller's), and each
 * isochrone has 0 or >= 2 points.  correlation: 0 gives the reference's
 * cov_mult; r in (-1, 1) instead gives each covariant kernel the per-point
 * multiplier c = -r / sqrt(1 - r^2) * color_err / mag_err (a colour-magnitude
 * correlation r; no reference counterpart).  Outputs: out (host, nbins x
 * ntemplates column-major, nbins = nx*ny, colour bin fastest) and/or d_out
 * (device, column t at d_out + t*ld_out); *npoints_out / *pairs_out (may be
 * NULL) receive the kernel points and (point, bin) pairs deposited. */
int sfh_partial_cmd_smooth_batch(sfh_ctx* ctx, int64_t ntemplates, const int64_t* offsets, const double* m_ini,
                                 int32_t nfilters, const double* const* mags, int32_t y_index, int32_t color_a,
                                 int32_t color_b, const sfh_fn* mag_err, const sfh_fn* completeness,
                                 const sfh_fn* bias, const sfh_fn* imf, const double* dmod,
                                 const double* normalize_value, const double* mean_mass, double correlation,
                                 sfh_range xedges, sfh_range yedges, float* out, float* d_out, int64_t ld_out,
                                 int64_t* npoints_out, int64_t* pairs_out);

/* ------------------------------------------------ device buffers (no reference counterpart) */

/* Device memory on the context's device for callers without a device array
 * type of their own (a Julia ccall binding): the *_device entry points and
 * the sfh_nn_* primitives take these pointers.  Copies are ordered on the
 * context's stream and return when complete. */
int sfh_buffer_alloc(sfh_ctx* ctx, int64_t bytes, void** d_ptr);
int sfh_buffer_free(sfh_ctx* ctx, void* d_ptr);
int sfh_buffer_upload(sfh_ctx* ctx, void* d_dst, const void* h_src, int64_t bytes);
int sfh_buffer_download(sfh_ctx* ctx, void* h_dst, const void* d_src, int64_t bytes);

/* ------------------------------- learned template surrogate primitives (new) */

/* No reference counterpart: the device operations the template surrogate
 * (a network from (logAge, [M/H], distance modulus, error scale) to a
 * log-template) is trained and evaluated with.  All pointers are device
 * pointers; every call is asynchronous on the context's stream except
 * sfh_nn_mse with a host loss pointer.  dtype 0 = BF16 (stored as uint16),
 * 1 = FP32.  Matrices are row-major. */

/* C (M x N, FP32, ldc) = alpha * A B^T + beta * C, A: M x K (lda), B: N x K
 * (ldb), K a multiple of 16 (zero padded), lda and ldb multiples of 8; on
 * the matrix cores with FP32 accumulation. */
int sfh_nn_gemm_nt(sfh_ctx* ctx, int dtype, const void* A, int64_t lda, const void* B, int64_t ldb, float* C,
                   int64_t ldc, int64_t M, int64_t N, int64_t K, float alpha, float beta);
/* Y (rows x ldy, dtype) = act(X + bias) with zero padding past cols; pre
 * (rows x cols FP32, may be NULL) keeps X + bias.  act 0 = identity, 1 = SiLU. */
int sfh_nn_bias_act(sfh_ctx* ctx, int dtype, const float* X, int64_t ldx, const float* bias, int64_t rows,
                    int64_t cols, int act, float* pre, void* Y, int64_t ldy);
/* A dense layer: pre (M x N FP32, may be NULL) = A W^T + bias (A: M x K, W:
 * N x K, dtype, K a multiple of 16) and Y (M x ldy, dtype) = act(pre) with zero
 * padding past N; YT (N x ldt, dtype, may be NULL) = Y^T with zeros in rows
 * M..ldt-1 (what sfh_nn_transpose of act(pre) gives); z (M x N FP32) is
 * workspace.  sfh_nn_gemm_nt followed by sfh_nn_bias_act, bit for bit, with a
 * split product's slices summed by the activation pass itself. */
int sfh_nn_linear(sfh_ctx* ctx, int dtype, const void* A, int64_t lda, const void* W, int64_t ldw, int64_t M,
                  int64_t N, int64_t K, const float* bias, int act, float* z, float* pre, void* Y, int64_t ldy,
                  void* YT, int64_t ldt);
/* dX = dY * act'(pre) elementwise (dX may alias dY). */
int sfh_nn_act_bwd(sfh_ctx* ctx, const float* dY, const float* pre, int64_t n, int act, float* dX);
/* The same over P pixels of 32 channels (16-byte aligned), with the channel
 * sums dsum[c] = sum_p dX[p, c] (a convolution's bias gradient) from the same
 * pass, added in an order fixed by P. */
int sfh_nn_act_bwd_chansum(sfh_ctx* ctx, const float* dY, const float* pre, int64_t P, int act, float* dX,
                           float* dsum);
/* out (ocols x ldo, dtype) = act(in)^T (in: rows x cols FP32, ldi), zero padded. */
int sfh_nn_transpose(sfh_ctx* ctx, int dtype, const float* in, int64_t ldi, int64_t rows, int64_t cols, void* out,
                     int64_t ldo, int64_t ocols, int act);
/* out[c] = sum_r X[r, c] (+ beta * out[c]), fixed summation order. */
int sfh_nn_colsum(sfh_ctx* ctx, const float* X, int64_t ld, int64_t rows, int64_t cols, float beta, float* out);
/* A dense layer's output gradient d (rows x cols FP32, row stride cols)
 * prepared for its two products in one pass: d' = d * act'(pre), gb[c] =
 * sum_r d'[r, c] (the summation order of sfh_nn_colsum), dl (rows x cols,
 * dtype) = d' and dT (cols x ldt, dtype, 16-byte aligned) = d'^T with zeros
 * for rows <= r < ldt (ldt a multiple of 8): sfh_nn_act_bwd, sfh_nn_colsum,
 * sfh_nn_bias_act and sfh_nn_transpose of the same d in one launch. */
int sfh_nn_dense_bwd_prep(sfh_ctx* ctx, int dtype, const float* d, const float* pre, int64_t rows, int64_t cols,
                          int act, float* gb, void* dl, void* dT, int64_t ldt);
/* Stride-2 4x4 transposed convolution (padding 1), the two gathers around
 * its matrix products: col2im sums the 4 taps landing on each of the
 * B x 2H x 2W x Co outputs from cols (B*H*W rows, column (ky*4+kx)*Co+co) and
 * adds bias; im2col builds those rows from an output gradient. */
int sfh_nn_col2im(sfh_ctx* ctx, const float* cols, int64_t ldc, int64_t B, int64_t H, int64_t W, int64_t Co,
                  const float* bias, float* out);
int sfh_nn_im2col(sfh_ctx* ctx, int dtype, const float* dout, int64_t B, int64_t H, int64_t W, int64_t Co,
                  void* dcols, int64_t ldd);
/* The transpose of im2col, as the weight gradient's operand: dcolsT (16 Co
 * rows of ldo >= B H W, dtype, zero past B H W) with Co in {4, 8, 16, 32}. */
int sfh_nn_im2col_t(sfh_ctx* ctx, int dtype, const float* dout, int64_t B, int64_t H, int64_t W, int64_t Co,
                    void* dcolsT, int64_t ldo);
/* Low-precision copies of weight matrices, one launch for up to 8: for each,
 * dst (row-major, rows x cols) = src (FP32, row-major) and, when dstT is not
 * NULL, dstT (cols x rows) = its transpose, both rounded to dtype. */
typedef struct sfh_nn_wcopy {
    const float* src;
    int64_t rows, cols;
    void* dst;
    void* dstT;
} sfh_nn_wcopy;
int sfh_nn_weight_copies(sfh_ctx* ctx, int dtype, int64_t n, const sfh_nn_wcopy* mats);
/* Gradient of that transposed convolution's input: out (B Hi Wi x Ci, FP32,
 * row stride ldo) = im2col(d) wT^T, d the FP32 output gradient (B x 2Hi x 2Wi x
 * Co), wT (dtype, Ci rows of ldw >= 16 Co, column tap Co + co) -- the product of
 * sfh_nn_im2col and sfh_nn_gemm_nt with the column matrix gathered as it is
 * read instead of written (Co 8, 16 or 32). */
int sfh_nn_conv_t_dinput(sfh_ctx* ctx, int dtype, const float* d, int64_t B, int64_t Hi, int64_t Wi, int64_t Co,
                         const void* wT, int64_t ldw, int64_t Ci, float* out, int64_t ldo);
/* Forward of that transposed convolution with 32 input and 32 output channels:
 * pre (B x 2Hi x 2Wi x 32, FP32) = convolution of in (dtype, B Hi Wi x 32) with
 * w (dtype, 16 x 32 rows of 32: row tap 32 + co) plus bias[co] (may be NULL),
 * and aout (dtype, may be NULL) = SiLU(pre).  The products of sfh_nn_gemm_nt
 * and sfh_nn_col2im (and the SiLU of sfh_nn_bias_act) in one kernel, one output
 * parity class per launch plane, the im2col operand gathered as it is read. */
int sfh_nn_conv_t_fwd(sfh_ctx* ctx, int dtype, const void* in, int64_t B, int64_t Hi, int64_t Wi, int64_t C,
                      const void* w, const float* bias, float* pre, void* aout);
/* Weight gradient of that transposed convolution with 32 input and 32 output
 * channels: dK (16 x 32 rows of 32, row tap 32 + co, FP32) = sum over input
 * pixels p of im2col(d)[p][tap 32 + co] act(pre[p][ci]) (act 0: identity, 1:
 * SiLU), d the FP32 output gradient (B x 2Hi x 2Wi x 32), pre the FP32 layer
 * input before its activation (B Hi Wi x 32; d under 2 GiB); operands rounded
 * to dtype, FP32 accumulation, fixed-order sums over pixel splits. */
int sfh_nn_conv_t_dweight(sfh_ctx* ctx, int dtype, const float* d, const float* pre, int act, int64_t B, int64_t Hi,
                          int64_t Wi, int64_t C, float* dK);
/* The same transposed convolution with one output channel (B x 2H x 2W) and a
 * per-pixel bias map, applied to act(in) (act 0: identity, 1: SiLU, in then
 * holding pre-activations); its backward gives din (the gradient of in:
 * through act'), dsum (Ci: the channel sums of din, the previous layer's bias
 * gradient; needs din), dw (16 x Ci) and dbmap (any may be NULL). */
int sfh_nn_conv1_fwd(sfh_ctx* ctx, const float* in, int64_t B, int64_t H, int64_t W, int64_t Ci, const float* w,
                     const float* bmap, int act, float* out);
int sfh_nn_conv1_bwd(sfh_ctx* ctx, const float* dout, const float* in, int64_t B, int64_t H, int64_t W, int64_t Ci,
                     const float* w, int act, float* din, float* dsum, float* dw, float* dbmap);
/* Mean squared error over the top-left R x S of each of B outputs of H2 x W2
 * against target rows (R x S each): output b against row rows[b] (b when rows
 * is NULL).  dout (may be NULL) gets its gradient.  With loss non-NULL the
 * call waits and returns the loss; loss_acc (device, may be NULL) has the
 * loss added to it on the stream, for asynchronous training loops. */
int sfh_nn_mse(sfh_ctx* ctx, const float* out, const float* target, const int64_t* rows, int64_t B, int64_t H2,
               int64_t W2, int64_t R, int64_t S, float* dout, float* loss, double* loss_acc);
/* Training loss of the surrogate over the R x S crop of B outputs of H2 x W2
 * (log images, target rows chosen by rows as in sfh_nn_mse):
 *   w_mse MSE + mean_b [w_kl KL(p_e || p_s) + w_sum (log sum s - log sum e)^2],
 *   s = exp(scale out + offset), e = exp(scale target + offset), p = s / sum s,
 * the KL being the log-likelihood lost per star by fitting with s instead of
 * e, the last term the error of the template's normalisation.  dout gets the
 * gradient (zero outside the crop); acc (device, 3 doubles, may be NULL) has
 * the MSE, the mean KL and the mean squared log-sum error added.
 * Asynchronous. */
int sfh_nn_template_loss(sfh_ctx* ctx, const float* out, const float* target, const int64_t* rows, int64_t B,
                         int64_t H2, int64_t W2, int64_t R, int64_t S, float scale, float offset, float w_mse,
                         float w_kl, float w_sum, float* dout, double* acc);
/* dst[i, :cols] = src[rows[i], :cols] (FP32). */
int sfh_nn_gather_rows(sfh_ctx* ctx, const float* src, int64_t lds, const int64_t* rows, int64_t n, int64_t cols,
                       float* dst, int64_t ldd);
/* Templates (B rows of ldt, bin = ix + nx iy, as the deposition kernels write
 * them) to training images img[b, ix, iy] = (log(T + eps) - offset) / scale,
 * and network outputs (B x H2 x W2, top-left nx x ny used) back to templates
 * T = max(exp(out scale + offset) - eps, 0). */
int sfh_nn_log_target(sfh_ctx* ctx, const float* T, int64_t ldt, int64_t B, int64_t nx, int64_t ny, float eps,
                      float offset, float scale, float* img);
int sfh_nn_log_template(sfh_ctx* ctx, const float* out, int64_t B, int64_t H2, int64_t W2, int64_t nx, int64_t ny,
                        float eps, float offset, float scale, float* T, int64_t ldt);
/* Per template k of n (rows of E and S, nbins each): out[3k..3k+2] =
 * KL(E/sum E || S'/sum S'), sum E, sum S' with S' = max(S, floor_value), in
 * FP64.  N times the KL is the log-likelihood change of fitting N stars drawn
 * from E with S' instead of E (amplitude free). */
int sfh_nn_template_kl(sfh_ctx* ctx, const float* E, int64_t lde, const float* S, int64_t lds, int64_t n,
                       int64_t nbins, double floor_value, double* out);
/* Step-indexed forms for captured training loops (counter: a device int64):
 * rows_step[i] = rows_all[counter n + i]; counter += inc; and the Adam step
 * with (lr, 1 - b1^t, 1 - b2^t) read from sched[3 counter .. +2]. */
int sfh_nn_select_rows(sfh_ctx* ctx, const int64_t* rows_all, const int64_t* counter, int64_t n, int64_t* rows_step);
int sfh_nn_counter_add(sfh_ctx* ctx, int64_t* counter, int64_t inc);
int sfh_nn_adam_sched(sfh_ctx* ctx, float* p, const float* g, float* m, float* v, int64_t n, const float* sched,
                      const int64_t* counter, float b1, float b2, float eps, float weight_decay);
/* One Adam step (bias-corrected, step >= 1) on FP32 parameters. */
int sfh_nn_adam(sfh_ctx* ctx, float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int64_t step, float weight_decay);

#ifdef __cplusplus
}
#endif

#endif /* SFH_H */

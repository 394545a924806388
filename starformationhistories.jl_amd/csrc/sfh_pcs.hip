// sfh_pcs.hip -- partial_cmd_smooth (src/StarFormationHistories.jl:829-911,
// NoBinaries) for many isochrones at once, entirely on the device: the
// reference's host preparation -- mini_spacing (:110-148), interpolate_mini
// (:77-81), the user callables, calculate_weights (:765-776) and midpoints
// (:417-431) -- then the per-point kernel descriptors and the deposition
// gather of sfh_deposit.hip.
//
//   k_pcs_count   workgroup per isochrone: the resampled point count
//                 1 + sum_i n_i, n_i = ceil(d_i / Δmag) where the CMD distance
//                 d_i of segment i exceeds Δmag, else 1.
//   (host)        prefix sums of the counts -> offsets of the resampled and
//                 kernel points (one synchronisation).
//   k_pcs_points  workgroup per isochrone: a block scan of n_i places every
//                 segment's new masses m_i + j Δm_i / n_i (or m_{i+1}); the
//                 magnitudes of the three filters involved are interpolated
//                 linearly on the isochrone's knots, plus the distance modulus.
//   k_pcs_desc    workgroup per isochrone: the callables at every resampled
//                 point, the weights (trapezoid of the IMF, completeness of the
//                 left point), the midpoint series (the range form when every
//                 step agrees to sqrt(eps), as midpoints(v, true) does) and the
//                 footprint / evaluation constants of each kernel point.
//
// FP64 throughout, in the reference's operation order, with no contracted
// multiply-adds.  Inputs the restatement does not reproduce are refused with
// an error: masses that are not strictly increasing (sort_ingested and
// deduplicate_knots! are the caller's), resampled masses that would repeat
// (unique), and a last resampled mass past the isochrone (Interpolations
// throws BoundsError).
#include "sfh_internal.h"
#include "sfh_psf.h"

#include <algorithm>
#include <cmath>
#include <vector>

// no multiply-add contraction anywhere in this file (the reference's rounding)
#pragma clang fp contract(off)

namespace sfh {
namespace pcs {

constexpr int kWG = 256;
constexpr int64_t kMaxSegmentPoints = 1 << 24;

enum : int { kBadOrder = 1, kBadBounds = 2, kBadFinite = 4, kBadDuplicate = 8, kBadHuge = 16 };

struct Fn {
    int32_t kind;
    double p[4];
    double t_first, t_step;
    int64_t t_len;
    const double* t_vals;
    int32_t nseg;
    const double* breaks;
    const double* slopes;
    const double* coefs;
};

__device__ inline double eval(const Fn& f, double x)
{
    switch (f.kind) {
    case SFH_FN_EXP_PHOTERR:                        // exp_photerr, src/utilities.jl:218
        return pow(f.p[0], f.p[1] * (x - f.p[2])) + f.p[3];
    case SFH_FN_MARTIN2016:                         // Martin2016_complete, src/utilities.jl:205
        return f.p[0] / (1.0 + exp((x - f.p[1]) / f.p[2]));
    case SFH_FN_TABLE: {
        double u = (x - f.t_first) / f.t_step;
        u = fmin(fmax(u, 0.0), (double)(f.t_len - 1));
        int64_t k = (int64_t)u;
        if (k >= f.t_len - 1) k = f.t_len - 2;
        const double r = u - (double)k;
        return f.t_vals[k] + r * (f.t_vals[k + 1] - f.t_vals[k]);
    }
    case SFH_FN_POWERLAW: {
        if (!(x >= f.breaks[0] && x <= f.breaks[f.nseg])) return 0.0;
        int k = 0;
        while (k + 1 < f.nseg && x >= f.breaks[k + 1]) ++k;
        return f.p[0] * (f.coefs[k] * pow(x, -f.slopes[k]));
    }
    default:
        return f.p[0];
    }
}

// Filter roles: 0 = colour filter a (color_indices[1]), 1 = colour filter b,
// 2 = the y filter (the same array as a or b when y is a colour filter).
struct Args {
    const double* m;
    const double* mag[3];
    const int64_t* poff;
    const int64_t* qoff;
    const int64_t* koff;
    const double* dmod;
    const double* norm;
    const double* mean_mass;
    double dmag;
    Fn err[3], compl_[3], bias[3], imf;
    int y_role;            // 0 (y == a), 1 (y == b) or 2 (y not in the colour)
    double corr_c;         // -r / sqrt(1 - r^2) for a correlation r, or 0
    int use_corr;
    double cov_mult;       // -1, 1 (covariant) or 0 (asymmetric)
    psf::Range ex, ey;
    double* mini;          // resampled masses, qoff layout
    double* iso;           // 3 x Qtotal magnitudes (roles a, b, y)
    int64_t qtotal;
    psf::PointDesc* desc;  // koff layout
    unsigned long long* pairs;
    int* bad;
};

__device__ inline double block_sum_d(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// n_i of segment g (global index of its left point): mini_spacing :125-140
__device__ inline int64_t seg_count(const Args& a, int64_t g, int* flags)
{
    const double m0 = a.m[g], m1 = a.m[g + 1];
    if (!(m1 > m0)) *flags |= kBadOrder;
    const double c0 = a.mag[0][g] - a.mag[1][g], c1 = a.mag[0][g + 1] - a.mag[1][g + 1];   // colors :849
    const double dx = c1 - c0, dy = a.mag[2][g + 1] - a.mag[2][g];
    const double d = hypot(dx, dy);
    if (!isfinite(d) || !isfinite(m0) || !isfinite(m1)) {
        *flags |= kBadFinite;
        return 1;
    }
    if (!(d > a.dmag)) return 1;
    const double q = ceil(d / a.dmag);            // round(Int, diffi / Δmag, RoundUp)
    if (q > (double)kMaxSegmentPoints) {
        *flags |= kBadHuge;
        return 1;
    }
    return (int64_t)q;
}

__global__ __launch_bounds__(kWG) void k_pcs_count(Args a, int64_t* __restrict__ counts)
{
    __shared__ double sh[kWG];
    const int64_t t = blockIdx.x;
    const int64_t i0 = a.poff[t], P = a.poff[t + 1] - i0;
    int flags = 0;
    double s = 0.0;                               // exact: integers below 2^53
    for (int64_t i = threadIdx.x; i < P - 1; i += kWG) s += (double)seg_count(a, i0 + i, &flags);
    if (flags) atomicOr(a.bad, flags);
    s = block_sum_d(s, sh);
    if (threadIdx.x == 0) counts[t] = P > 0 ? 1 + (int64_t)s : 0;
}

// Linear interpolation on the knots (interpolate_mini, Gridded(Linear())) of
// every role at mass x lying in [m_k, m_{k+1}] of segment k (searchsorted
// from the right: a mass equal to or past a knot takes the next interval).
__device__ inline void put_point(const Args& a, int64_t i0, int64_t P, int64_t k, double x, int64_t pos, double dm,
                                 int* flags)
{
    int64_t kk = k;
    if (x >= a.m[i0 + k + 1] && k + 1 <= P - 2) kk = k + 1;
    if (x > a.m[i0 + P - 1]) *flags |= kBadBounds;
    const double mlo = a.m[i0 + kk], mhi = a.m[i0 + kk + 1];
    const double tt = (x - mlo) / (mhi - mlo);
    a.mini[pos] = x;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double* y = a.mag[r] + i0;
        a.iso[r * a.qtotal + pos] = (y[kk] * (1.0 - tt) + y[kk + 1] * tt) + dm;   // :854 (+ dmod)
    }
}

__global__ __launch_bounds__(kWG) void k_pcs_points(Args a)
{
    __shared__ int64_t scan[kWG];
    const int64_t t = blockIdx.x;
    const int64_t i0 = a.poff[t], P = a.poff[t + 1] - i0;
    if (P <= 0) return;
    const int64_t base = a.qoff[t];
    const double dm = a.dmod[t];
    int flags = 0;
    if (threadIdx.x == 0) put_point(a, i0, P, 0, a.m[i0], base, dm, &flags);   // new_mini[1] = first(m_ini) :117
    int64_t running = 1;
    for (int64_t c0 = 0; c0 < P - 1; c0 += kWG) {
        const int64_t i = c0 + threadIdx.x;
        const int64_t n = i < P - 1 ? seg_count(a, i0 + i, &flags) : 0;
        // inclusive block scan (Hillis-Steele) of the segment counts
        scan[threadIdx.x] = n;
        __syncthreads();
        for (int off = 1; off < kWG; off <<= 1) {
            const int64_t v = (int)threadIdx.x >= off ? scan[threadIdx.x - off] : 0;
            __syncthreads();
            scan[threadIdx.x] += v;
            __syncthreads();
        }
        const int64_t incl = scan[threadIdx.x], total = scan[kWG - 1];
        __syncthreads();
        if (i < P - 1) {
            const int64_t pos = base + running + incl - n;
            const double m0 = a.m[i0 + i], m1 = a.m[i0 + i + 1];
            if (n >= 2) {                               // interpolated segment :130-135
                const double step = (m1 - m0) / (double)n;
                for (int64_t j = 1; j <= n; ++j) put_point(a, i0, P, i, m0 + step * (double)j, pos + j - 1, dm, &flags);
            } else {                                    // :137
                put_point(a, i0, P, i, m1, pos, dm, &flags);
            }
        }
        running += total;
    }
    if (flags) atomicOr(a.bad, flags);
}

// The four midpoint series of resampled point k (bias-shifted colour and y
// magnitude, colour error, y error: :871-888) and its completeness.
struct PointVals {
    double v[4];
    double compl_;
};

__device__ inline PointVals point_vals(const Args& a, int64_t q)
{
    PointVals o;
    const double ia = a.iso[q], ib = a.iso[a.qtotal + q], iy = a.iso[2 * a.qtotal + q];
    const double ea = eval(a.err[0], ia), eb = eval(a.err[1], ib), ey = eval(a.err[2], iy);   // :855
    double cerr;
    double compl_ = eval(a.compl_[0], ia) * eval(a.compl_[1], ib);
    if (a.y_role == 0) cerr = eb;                   // x_c_idx: the colour filter that is not y :861-862
    else if (a.y_role == 1) cerr = ea;
    else {
        cerr = sqrt(ea * ea + eb * eb);             // :871-872
        compl_ = compl_ * eval(a.compl_[2], iy);    // :874-876
    }
    const double ba = ia + eval(a.bias[0], ia), bb = ib + eval(a.bias[1], ib);   // :884
    const double by = iy + eval(a.bias[2], iy);
    o.v[0] = ba - bb;                               // :885
    o.v[1] = by;
    o.v[2] = cerr;
    o.v[3] = ey;
    o.compl_ = compl_;
    return o;
}

// Julia's isapprox(x, y) with atol = 0, rtol = sqrt(eps(Float64))
__device__ inline bool approx(double x, double y)
{
    return x == y || (isfinite(x) && isfinite(y) && fabs(x - y) <= 1.4901161193847656e-8 * fmax(fabs(x), fabs(y)));
}

__global__ __launch_bounds__(kWG) void k_pcs_desc(Args a)
{
    __shared__ int ranges[4];
    const int64_t t = blockIdx.x;
    const int64_t q0 = a.qoff[t], Q = a.qoff[t + 1] - q0;
    const int64_t K = Q - 1;
    if (K <= 0) return;
    if (threadIdx.x < 4) ranges[threadIdx.x] = 1;
    __syncthreads();
    // midpoints(v, true): the range form when every step is approximately the first (:421-428)
    const PointVals f0 = point_vals(a, q0), f1 = point_vals(a, q0 + 1);
    const PointVals fl = point_vals(a, q0 + K);
    double d0[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) d0[s] = f1.v[s] - f0.v[s];
    int ok = 0xF;
    for (int64_t k = threadIdx.x; k < K; k += kWG) {
        const PointVals u = point_vals(a, q0 + k), w = point_vals(a, q0 + k + 1);
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (!approx(d0[s], w.v[s] - u.v[s])) ok &= ~(1 << s);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
        if (!(ok & (1 << s))) atomicAnd(&ranges[s], 0);
    __syncthreads();
    bool rng[4];
    double start[4], rstep[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        rng[s] = ranges[s] != 0;
        start[s] = f0.v[s] + d0[s] / 2.0;
        const double stop = fl.v[s] - d0[s] / 2.0;
        rstep[s] = K > 1 ? (stop - start[s]) / (double)(K - 1) : 0.0;
    }
    const double nm = a.norm[t], mm = a.mean_mass[t];
    const int64_t kbase = a.koff[t];
    int flags = 0;
    unsigned long long cnt = 0;
    for (int64_t k = threadIdx.x; k < K; k += kWG) {
        const int64_t q = q0 + k;
        const PointVals u = point_vals(a, q), w = point_vals(a, q + 1);
        double mid[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (rng[s]) mid[s] = K == 1 ? start[s] : start[s] + (double)k * rstep[s];
            else mid[s] = u.v[s] + (w.v[s] - u.v[s]) / 2.0;       // v[i] + d[i] / 2 :430
        }
        // calculate_weights :770-773 (completeness of the left point)
        const double m0 = a.mini[q], m1 = a.mini[q + 1];
        if (!(m1 > m0)) flags |= kBadDuplicate;
        double wt = (m1 - m0) * (eval(a.imf, m0) + eval(a.imf, m1)) / 2.0;
        wt = wt * (u.compl_ * nm / mm);
        psf::PointDesc d;
        const bool fin = isfinite(mid[0]) && isfinite(mid[1]) && isfinite(mid[2]) && isfinite(mid[3]) && isfinite(wt);
        if (!fin) {
            d = psf::empty_desc();
            flags |= kBadFinite;
        } else if (a.cov_mult == 0.0) {
            cnt += (unsigned long long)psf::prep_asymmetric(mid[0], mid[1], mid[2], mid[3], wt, a.ex, a.ey, &d);
        } else {
            const double c = a.use_corr ? a.corr_c * mid[2] / mid[3] : a.cov_mult;
            if (!isfinite(c)) {
                d = psf::empty_desc();
                flags |= kBadFinite;
            } else {
                cnt += (unsigned long long)psf::prep_covariant(mid[0], mid[1], mid[2], mid[3], c, wt, a.ex, a.ey, &d);
            }
        }
        a.desc[kbase + k] = d;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(a.pairs, cnt);
    if (flags) atomicOr(a.bad, flags);
}

// ------------------------------------------------------------------ host

// Host staging of every small array the kernels read (tables, power-law
// segments), uploaded in one copy; Fn pointers are patched to the device copy.
struct FnStage {
    std::vector<double> data;
    std::vector<std::pair<const double**, size_t>> fix;   // pointer field, offset into data
    size_t add(const double* h, int64_t n)
    {
        const size_t off = data.size();
        data.insert(data.end(), h, h + n);
        return off;
    }
};

static int stage_fn(sfh_ctx* ctx, const sfh_fn* src, double deflt, Fn* dst, FnStage* st)
{
    *dst = Fn{};
    if (!src) {
        dst->kind = SFH_FN_CONST;
        dst->p[0] = deflt;
        return SFH_OK;
    }
    dst->kind = src->kind;
    for (int k = 0; k < 4; ++k) dst->p[k] = src->p[k];
    switch (src->kind) {
    case SFH_FN_CONST:
    case SFH_FN_EXP_PHOTERR:
        break;
    case SFH_FN_MARTIN2016:
        SFH_ARG_CHECK(ctx, src->p[2] != 0.0, "Martin2016_complete: rho must be non-zero");
        break;
    case SFH_FN_TABLE:
        SFH_ARG_CHECK(ctx, src->table.values && src->table.length >= 2 && src->table.step > 0.0 &&
                               src->table.length < ((int64_t)1 << 26),
                      "sfh_fn table: 2 to 2^26 values and a positive step");
        dst->t_first = src->table.first;
        dst->t_step = src->table.step;
        dst->t_len = src->table.length;
        st->fix.push_back({&dst->t_vals, st->add(src->table.values, src->table.length)});
        break;
    case SFH_FN_POWERLAW: {
        const sfh_imf& pl = src->powerlaw;
        SFH_ARG_CHECK(ctx, pl.nseg >= 1 && pl.nseg <= 16 && pl.breaks && pl.slopes && pl.coefs,
                      "sfh_fn power law: 1..16 segments");
        for (int k = 0; k < pl.nseg; ++k)
            SFH_ARG_CHECK(ctx, pl.breaks[k] < pl.breaks[k + 1], "sfh_fn power law: increasing breaks");
        dst->nseg = pl.nseg;
        st->fix.push_back({&dst->breaks, st->add(pl.breaks, pl.nseg + 1)});
        st->fix.push_back({&dst->slopes, st->add(pl.slopes, pl.nseg)});
        st->fix.push_back({&dst->coefs, st->add(pl.coefs, pl.nseg)});
        break;
    }
    default:
        return set_error(ctx, SFH_ERR_ARGUMENT, "unknown sfh_fn kind %d", src->kind);
    }
    return SFH_OK;
}

static int pcs_error(sfh_ctx* ctx, int b)
{
    if (b & kBadFinite) return set_error(ctx, SFH_ERR_DOMAIN, "non-finite isochrone value or kernel parameter");
    if (b & kBadOrder)
        return set_error(ctx, SFH_ERR_ARGUMENT,
                         "m_ini must be strictly increasing per isochrone (sort_ingested / deduplicate_knots! are "
                         "not applied here)");
    if (b & kBadHuge) return set_error(ctx, SFH_ERR_ARGUMENT, "an isochrone segment needs more than 2^24 points");
    if (b & kBadBounds)
        return set_error(ctx, SFH_ERR_ARGUMENT, "BoundsError: a resampled mass lies past the isochrone's last mass");
    return set_error(ctx, SFH_ERR_ARGUMENT, "resampled masses repeat (mini_spacing's unique() would drop them)");
}

// A named context scratch buffer filled from host memory (n bytes).
static int upload_scratch(sfh_ctx* ctx, const char* name, const void* h, size_t n, void** d)
{
    if (int rc = scratch(ctx, name, n > 0 ? n : 8, d)) return rc;
    if (n) SFH_HIP_TRY(ctx, hipMemcpyAsync(*d, h, n, hipMemcpyHostToDevice, ctx->stream));
    return SFH_OK;
}

}  // namespace pcs
}  // namespace sfh

using namespace sfh;
using namespace sfh::pcs;

extern "C" {

int sfh_partial_cmd_smooth_batch(sfh_ctx* ctx, int64_t ntemplates, const int64_t* offsets, const double* m_ini,
                                 int32_t nfilters, const double* const* mags, int32_t y_index, int32_t color_a,
                                 int32_t color_b, const sfh_fn* mag_err, const sfh_fn* completeness,
                                 const sfh_fn* bias, const sfh_fn* imf, const double* dmod,
                                 const double* normalize_value, const double* mean_mass, double correlation,
                                 sfh_range xedges, sfh_range yedges, float* out, float* d_out, int64_t ld_out,
                                 int64_t* npoints_out, int64_t* pairs_out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_ARG_CHECK(ctx, ntemplates >= 1 && ntemplates < ((int64_t)1 << 31), "1 <= ntemplates < 2^31");
    SFH_ARG_CHECK(ctx, offsets && m_ini && mags && mag_err && imf && dmod && normalize_value && mean_mass,
                  "null pointer");
    SFH_ARG_CHECK(ctx, nfilters >= 1 && nfilters <= 64, "1 <= length(mags) <= 64");
    SFH_ARG_CHECK(ctx, y_index >= 0 && y_index < nfilters && color_a >= 0 && color_a < nfilters && color_b >= 0 &&
                           color_b < nfilters,
                  "y_index / color_indices out of range");
    SFH_ARG_CHECK(ctx, out || d_out, "no output (out and d_out are both NULL)");
    SFH_ARG_CHECK(ctx, std::isfinite(correlation) && correlation > -1.0 && correlation < 1.0,
                  "correlation must lie in (-1, 1) (0: the reference's cov_mult)");
    SFH_ARG_CHECK(ctx, xedges.length >= 2 && yedges.length >= 2 && std::isfinite(xedges.step) &&
                           std::isfinite(yedges.step) && xedges.step > 0.0 && yedges.step > 0.0 &&
                           xedges.length < (1 << 30) && yedges.length < (1 << 30),
                  "edge ranges need >= 2 edges and positive finite steps");
    const int64_t nx = xedges.length - 1, ny = yedges.length - 1, nbins = nx * ny;
    SFH_ARG_CHECK(ctx, nbins > 1, "length(image.weights) > 1 (:367)");
    SFH_ARG_CHECK(ctx, !d_out || ld_out >= nbins, "ld_out < nbins");
    SFH_ARG_CHECK(ctx, offsets[0] == 0, "offsets[0] must be 0");
    for (int64_t t = 0; t < ntemplates; ++t)
        SFH_ARG_CHECK(ctx, offsets[t + 1] >= offsets[t] && offsets[t + 1] - offsets[t] != 1,
                      "offsets non-decreasing, every isochrone with 0 or >= 2 points (interpolation knots) (t=%lld)",
                      (long long)t);
    for (int64_t t = 0; t < ntemplates; ++t)
        SFH_ARG_CHECK(ctx, std::isfinite(dmod[t]) && std::isfinite(normalize_value[t]) && mean_mass[t] != 0.0 &&
                               std::isfinite(mean_mass[t]),
                      "dmod / normalize_value / mean_mass finite, mean_mass non-zero (t=%lld)", (long long)t);
    const int64_t P = offsets[ntemplates];
    const int roles[3] = {color_a, color_b, y_index};
    for (int r = 0; r < 3; ++r) SFH_ARG_CHECK(ctx, mags[roles[r]] || P == 0, "null magnitude array");
    hipStream_t s = ctx->stream;

    Args a{};
    FnStage st;
    for (int r = 0; r < 3; ++r) {
        const int f = roles[r];
        if (int rc = stage_fn(ctx, &mag_err[f], 0.0, &a.err[r], &st)) return rc;
        if (int rc = stage_fn(ctx, completeness ? &completeness[f] : nullptr, 1.0, &a.compl_[r], &st)) return rc;
        if (int rc = stage_fn(ctx, bias ? &bias[f] : nullptr, 0.0, &a.bias[r], &st)) return rc;
    }
    if (int rc = stage_fn(ctx, imf, 0.0, &a.imf, &st)) return rc;
    void* dfn = nullptr;
    if (int rc = upload_scratch(ctx, "pcs_fn", st.data.data(), st.data.size() * sizeof(double), &dfn)) return rc;
    for (auto& fx : st.fix) *fx.first = (const double*)dfn + fx.second;

    // inputs: masses, the filters involved (each once), offsets, per-template scalars
    const size_t pb = (size_t)P * sizeof(double);
    void *dm = nullptr, *dpoff = nullptr, *dsc = nullptr;
    if (int rc = upload_scratch(ctx, "pcs_m", m_ini, pb, &dm)) return rc;
    if (int rc = upload_scratch(ctx, "pcs_poff", offsets, (size_t)(ntemplates + 1) * sizeof(int64_t), &dpoff))
        return rc;
    std::vector<double> scal((size_t)ntemplates * 3);
    std::copy(dmod, dmod + ntemplates, scal.begin());
    std::copy(normalize_value, normalize_value + ntemplates, scal.begin() + ntemplates);
    std::copy(mean_mass, mean_mass + ntemplates, scal.begin() + 2 * ntemplates);
    if (int rc = upload_scratch(ctx, "pcs_scal", scal.data(), scal.size() * sizeof(double), &dsc)) return rc;
    static const char* kMagNames[3] = {"pcs_mag0", "pcs_mag1", "pcs_mag2"};
    void* dmag[3] = {};
    for (int r = 0; r < 3; ++r) {
        int prev = -1;
        for (int q = 0; q < r; ++q)
            if (roles[q] == roles[r]) prev = q;
        if (prev >= 0) dmag[r] = dmag[prev];
        else if (int rc = upload_scratch(ctx, kMagNames[r], mags[roles[r]], pb, &dmag[r])) return rc;
    }
    a.m = (const double*)dm;
    for (int r = 0; r < 3; ++r) a.mag[r] = (const double*)dmag[r];
    a.poff = (const int64_t*)dpoff;
    a.dmod = (const double*)dsc;
    a.norm = (const double*)dsc + ntemplates;
    a.mean_mass = (const double*)dsc + 2 * ntemplates;
    a.dmag = std::min(xedges.step, yedges.step);            // :851
    // y filter in the colour -> covariant kernel with cov_mult -1 (y = a) or
    // 1 (y = b); otherwise the asymmetric kernel (:857-881)
    a.y_role = y_index == color_a ? 0 : y_index == color_b ? 1 : 2;
    a.cov_mult = a.y_role == 0 ? -1.0 : a.y_role == 1 ? 1.0 : 0.0;
    a.use_corr = correlation != 0.0 && a.y_role != 2;
    a.corr_c = -correlation / std::sqrt(1.0 - correlation * correlation);
    a.ex = psf::Range{xedges.first, xedges.step, xedges.length};
    a.ey = psf::Range{yedges.first, yedges.step, yedges.length};
    void *pc = nullptr, *pcnt = nullptr;
    if (int rc = scratch(ctx, "pcs_counts", (size_t)ntemplates * sizeof(int64_t), &pc)) return rc;
    if (int rc = scratch(ctx, "pcs_flags", 64, &pcnt)) return rc;
    SFH_HIP_TRY(ctx, hipMemsetAsync(pcnt, 0, 64, s));
    a.pairs = (unsigned long long*)pcnt;
    a.bad = (int*)((char*)pcnt + 16);

    // 1. resampled point counts
    {
        TimedLaunch tl(ctx, "pcs_count");
        hipLaunchKernelGGL(k_pcs_count, dim3((unsigned)ntemplates), dim3(kWG), 0, s, a, (int64_t*)pc);
    }
    if (int rc = check_launch(ctx, "k_pcs_count")) return rc;
    std::vector<int64_t> counts((size_t)ntemplates);
    int bad = 0;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(counts.data(), pc, counts.size() * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(&bad, a.bad, sizeof(int), hipMemcpyDeviceToHost, s));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (bad) return pcs_error(ctx, bad);
    std::vector<int64_t> qoff((size_t)ntemplates + 1, 0), koff((size_t)ntemplates + 1, 0);
    for (int64_t t = 0; t < ntemplates; ++t) {
        qoff[t + 1] = qoff[t] + counts[t];
        koff[t + 1] = koff[t] + std::max<int64_t>(counts[t] - 1, 0);
    }
    const int64_t Q = qoff[ntemplates], K = koff[ntemplates];
    SFH_ARG_CHECK(ctx, Q < ((int64_t)1 << 34), "too many resampled points (%lld)", (long long)Q);
    void *dq = nullptr, *dk = nullptr, *dmini = nullptr, *diso = nullptr, *ddesc = nullptr;
    if (int rc = upload_scratch(ctx, "pcs_qoff", qoff.data(), qoff.size() * sizeof(int64_t), &dq)) return rc;
    if (int rc = upload_scratch(ctx, "pcs_koff", koff.data(), koff.size() * sizeof(int64_t), &dk)) return rc;
    if (int rc = scratch(ctx, "pcs_mini", (size_t)(Q + 1) * sizeof(double), &dmini)) return rc;
    if (int rc = scratch(ctx, "pcs_iso", (size_t)(Q + 1) * 3 * sizeof(double), &diso)) return rc;
    if (int rc = scratch(ctx, "dep_desc", (size_t)(K + 1) * sizeof(psf::PointDesc), &ddesc)) return rc;
    a.qoff = (const int64_t*)dq;
    a.koff = (const int64_t*)dk;
    a.mini = (double*)dmini;
    a.iso = (double*)diso;
    a.qtotal = Q;
    a.desc = (psf::PointDesc*)ddesc;

    // 2. resampled masses and magnitudes, 3. kernel point descriptors
    {
        TimedLaunch tl(ctx, "pcs_points");
        hipLaunchKernelGGL(k_pcs_points, dim3((unsigned)ntemplates), dim3(kWG), 0, s, a);
    }
    if (int rc = check_launch(ctx, "k_pcs_points")) return rc;
    {
        TimedLaunch tl(ctx, "pcs_desc");
        hipLaunchKernelGGL(k_pcs_desc, dim3((unsigned)ntemplates), dim3(kWG), 0, s, a);
    }
    if (int rc = check_launch(ctx, "k_pcs_desc")) return rc;
    // 4. deposition into the device output (or a staging buffer for `out`)
    float* dst = d_out;
    const int64_t ld = d_out ? ld_out : nbins;
    if (!dst) {
        void* p = nullptr;
        if (int rc = scratch(ctx, "pcs_out", (size_t)nbins * ntemplates * sizeof(float), &p)) return rc;
        dst = (float*)p;
    }
    if (int rc = deposit_from_desc(ctx, ddesc, (const int64_t*)dk, ntemplates, K, a.cov_mult == 0.0, xedges,
                                   yedges, dst, ld))
        return rc;
    unsigned long long pairs = 0;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(&bad, a.bad, sizeof(int), hipMemcpyDeviceToHost, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(&pairs, a.pairs, sizeof(pairs), hipMemcpyDeviceToHost, s));
    if (out) {
        SFH_HIP_TRY(ctx, hipMemcpy2DAsync(out, (size_t)nbins * sizeof(float), dst, (size_t)ld * sizeof(float),
                                          (size_t)nbins * sizeof(float), (size_t)ntemplates, hipMemcpyDeviceToHost,
                                          s));
    }
    SFH_HIP_TRY(ctx, hipStreamSynchronize(s));
    if (bad) return pcs_error(ctx, bad);
    if (npoints_out) *npoints_out = K;
    if (pairs_out) *pairs_out = (int64_t)pairs;
    return SFH_OK;
}

}  // extern "C"

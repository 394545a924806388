// sfh_mock.hip -- the mock-observation path: stellar systems sampled from an
// isochrone and an IMF to a total birth mass or luminosity
// (generate_stars_mass / generate_stars_mag, src/simulate.jl:306-404, with
// sample_system :246-259), observed with Gaussian errors, bias and
// completeness (model_cmd, :621-674) and binned into Hess diagrams (bin_cmd,
// src/StarFormationHistories.jl:544-553).
//
// Randomness is counter based: system (or star) i draws from
// Philox4x64-10(counter = (i, block), key = (seed, stream)), so every system
// is independent of the others and the host restatement draws the same
// numbers.  The reference's sequential "while total < limit" loop becomes:
// per-system contributions (system mass, or flux in the limit filter) in
// 64-bit fixed point (scale 2^62 / limit), per-workgroup integer sums, and
// the first system whose running integer sum reaches 2^62 ends the sample --
// exact in any summation order.  Kept systems are compacted in system order.
#include "sfh_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

struct sfh_catalog {
    sfh_ctx* ctx = nullptr;
    int64_t n = 0, cap = 0;       // entries held / allocated
    int32_t nstars = 0;           // masses per entry (1 or 2; 0 for observed catalogues)
    int32_t nfilters = 0;
    int64_t sampled = 0;          // systems drawn, or input stars of model_cmd
    double* d_masses = nullptr;   // n x nstars
    double* d_mags = nullptr;     // n x nfilters
    int64_t* d_index = nullptr;   // system index, or index of the input star
};

namespace sfh {
namespace mock {

constexpr int kWG = 256;
constexpr int kMaxFilters = SFH_MOCK_MAX_FILTERS;
constexpr int kMaxSeg = SFH_MOCK_MAX_IMF_SEGMENTS;
constexpr long long kTarget = 1LL << 62;

// ---------------------------------------------------------------- Philox
// Philox4x64-10 (Salmon, Moraes, Dror & Shaw, SC'11): ten rounds of two
// 64 x 64 -> 128-bit products with Weyl-sequence key bumps.
struct R4 {
    uint64_t v[4];
};

__device__ inline R4 philox(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1)
{
    constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
    constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
    uint64_t x0 = c0, x1 = c1, x2 = 0, x3 = 0;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t lo0 = M0 * x0, hi0 = __umul64hi(M0, x0);
        const uint64_t lo1 = M1 * x2, hi1 = __umul64hi(M1, x2);
        const uint64_t y0 = hi1 ^ x1 ^ k0, y2 = hi0 ^ x3 ^ k1;
        x0 = y0;
        x1 = lo1;
        x2 = y2;
        x3 = lo0;
        k0 += W0;
        k1 += W1;
    }
    return R4{{x0, x1, x2, x3}};
}

// uniform in [0, 1) from the top 53 bits
__device__ inline double u01(uint64_t r) { return (double)(r >> 11) * 0x1.0p-53; }

// ---------------------------------------------------------------- IMF
// Piecewise power law: pdf(m) = coef_k m^-slope_k on [lo_k, lo_{k+1}); cum_k
// = integral below lo_k (unnormalised), cum_nseg the total.
struct Imf {
    int nseg;
    double lo[kMaxSeg + 1];
    double a[kMaxSeg];       // 1 - slope
    double coef[kMaxSeg];
    double cum[kMaxSeg + 1];
};

__device__ inline double imf_sample(const Imf& f, double u)
{
    const double x = u * f.cum[f.nseg];
    int k = 0;
    while (k + 1 < f.nseg && x >= f.cum[k + 1]) ++k;
    const double y = (x - f.cum[k]) / f.coef[k];
    double m;
    if (f.a[k] == 0.0)
        m = f.lo[k] * exp(y);
    else
        m = pow(pow(f.lo[k], f.a[k]) + y * f.a[k], 1.0 / f.a[k]);
    return fmin(fmax(m, f.lo[k]), f.lo[k + 1]);
}

// ---------------------------------------------------------------- tables
struct Table {
    double first, step;
    int64_t n;
    const double* v;
};

// linear interpolation on a uniform grid, clamped to the ends
__device__ inline double lookup(const Table& t, double x)
{
    double u = (x - t.first) / t.step;
    u = fmin(fmax(u, 0.0), (double)(t.n - 1));
    int64_t k = (int64_t)u;
    if (k >= t.n - 1) k = t.n - 2;
    const double f = u - (double)k;
    return t.v[k] + f * (t.v[k + 1] - t.v[k]);
}

// ---------------------------------------------------------------- sampling
struct GenArgs {
    const double* mini;       // npoints, ascending, no duplicates
    const double* mags;       // npoints x nf (row-major), distance modulus included
    int64_t npoints;
    int nf, nst;              // filters, masses per system
    int mode;                 // 0: to a birth mass (generate_stars_mass), 1: to a flux
    int limf;                 // mode 1: filter of the limit
    int mlf;                  // mode 1: filter of mag_lim
    double mmin_cut;          // mode 0: primary masses below are not returned (mass_limits)
    double mag_lim;           // mode 1
    double scale;             // fixed-point scale of the contributions
    Imf imf;
    int bkind;                // 0 NoBinaries, 1 RandomBinaryPairs, 2 BinaryMassRatio (uniform q)
    double bfrac, qmin, qmax;
    uint64_t key0, key1;
};

__device__ inline double interp_mag(const GenArgs& a, double m, int f)
{
    int64_t lo = 0, hi = a.npoints - 1;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (a.mini[mid] <= m)
            lo = mid;
        else
            hi = mid;
    }
    const double t = (m - a.mini[lo]) / (a.mini[lo + 1] - a.mini[lo]);
    const double y0 = a.mags[lo * a.nf + f], y1 = a.mags[(lo + 1) * a.nf + f];
    return y0 + t * (y1 - y0);
}

// One system per thread: its masses, magnitudes, contribution (fixed point)
// and whether it is returned; per-workgroup sums of the contributions.
__global__ __launch_bounds__(kWG) void k_generate(GenArgs a, int64_t s0, int64_t n, long long* __restrict__ q,
                                                  double* __restrict__ masses, double* __restrict__ mags,
                                                  unsigned char* __restrict__ keep, long long* __restrict__ bsum)
{
#pragma clang fp contract(off)
    __shared__ long long red[kWG];
    const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
    long long qi = 0;
    if (t < n) {
        const int64_t i = s0 + t;
        const R4 r = philox((uint64_t)i, 0, a.key0, a.key1);
        const double ub = u01(r.v[0]), u1 = u01(r.v[1]), u2 = u01(r.v[2]);
        double ms[2] = {0.0, 0.0};
        if (a.bkind == 0) {                                        // sample_system :246
            ms[0] = imf_sample(a.imf, u1);
        } else if (a.bkind == 1) {                                 // :248-259
            ms[0] = imf_sample(a.imf, u1);
            if (ub <= a.bfrac) {
                const double m2 = imf_sample(a.imf, u2);
                if (m2 > ms[0]) {
                    ms[1] = ms[0];
                    ms[0] = m2;
                } else {
                    ms[1] = m2;
                }
            }
        } else {                                                   // :261-277
            const double M = imf_sample(a.imf, u1);
            ms[0] = M;
            if (ub <= a.bfrac) {
                const double qr = a.qmin + (a.qmax - a.qmin) * u2;
                ms[0] = M / (qr + 1.0);
                ms[1] = M - ms[0];
            }
        }
        const double mmin1 = a.mini[0], mmax = a.mini[a.npoints - 1];
        double flux[kMaxFilters];
#pragma unroll
        for (int f = 0; f < kMaxFilters; ++f) flux[f] = 0.0;
        for (int s = 0; s < a.nst; ++s) {
            const double m = ms[s];
            // mass mode: open interval (:327); luminosity mode: closed (:387)
            const bool valid = a.mode == 0 ? (m > mmin1 && m < mmax) : !(m < mmin1 || m > mmax);
            if (!valid) continue;
            for (int f = 0; f < a.nf; ++f) flux[f] += exp10(4.0 * (0.0 - interp_mag(a, m, f)) / 10.0);
        }
        double fsum = 0.0;
        for (int f = 0; f < a.nf; ++f) fsum += flux[f];
        double c;
        bool kp;
        if (a.mode == 0) {
            c = a.nst == 2 ? ms[0] + ms[1] : ms[0];
            kp = !(ms[0] < a.mmin_cut) && fsum > 0.0;
        } else {
            c = flux[a.limf];
            kp = fsum > 0.0 && (isinf(a.mag_lim) || 0.0 - 5.0 * log10(flux[a.mlf]) / 2.0 < a.mag_lim);
        }
        const double qd = c * a.scale;
        qi = qd >= (double)kTarget ? kTarget : llrint(qd);
        q[t] = qi;
        keep[t] = kp ? 1 : 0;
        for (int s = 0; s < a.nst; ++s) masses[t * a.nst + s] = ms[s];
        for (int f = 0; f < a.nf; ++f) mags[t * a.nf + f] = 0.0 - 5.0 * log10(flux[f]) / 2.0;
    }
    // block sum saturating at 2^62: enough to find the crossing block, and
    // no overflow when single systems carry up to the whole limit
    red[threadIdx.x] = qi;
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            const unsigned long long s2 =
                (unsigned long long)red[threadIdx.x] + (unsigned long long)red[threadIdx.x + w];
            red[threadIdx.x] = s2 > (unsigned long long)kTarget ? kTarget : (long long)s2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------- observing
struct ObsArgs {
    int nf;
    int has_bias;
    Table err[kMaxFilters], compl_[kMaxFilters], bias[kMaxFilters];
    uint64_t key0, key1;
};

// Standard normal number f of star j: Box-Muller on uniform pair f / 2
// (pair 0 from words 1, 2 of block 0; pair p > 0 from words 0, 1 of block p),
// cosine for even f, sine for odd f.
__device__ inline double normal_of(const ObsArgs& a, int64_t j, const R4& r0, int f)
{
    const int p = f >> 1;
    uint64_t wa, wb;
    if (p == 0) {
        wa = r0.v[1];
        wb = r0.v[2];
    } else {
        const R4 r = philox((uint64_t)j, (uint64_t)p, a.key0, a.key1);
        wa = r.v[0];
        wb = r.v[1];
    }
    const double rad = sqrt(-2.0 * log(1.0 - u01(wa)));
    double s, c;
    sincospi(2.0 * u01(wb), &s, &c);
    return rad * ((f & 1) ? s : c);
}

__global__ __launch_bounds__(kWG) void k_observe(const double* __restrict__ in, int64_t n, ObsArgs a,
                                                 double* __restrict__ out, unsigned char* __restrict__ keep)
{
#pragma clang fp contract(off)
    const int64_t j = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (j >= n) return;
    const R4 r0 = philox((uint64_t)j, 0, a.key0, a.key1);
    double c = 1.0;
    for (int f = 0; f < a.nf; ++f) c *= lookup(a.compl_[f], in[j * a.nf + f]);       // :629-633
    keep[j] = u01(r0.v[0]) <= c ? 1 : 0;                                               // :635
    for (int f = 0; f < a.nf; ++f) {                                                   // :667-671
        const double m = in[j * a.nf + f];
        const double e = lookup(a.err[f], m), b = a.has_bias ? lookup(a.bias[f], m) : 0.0;
        out[j * a.nf + f] = m + normal_of(a, j, r0, f) * e + b;
    }
}

// ---------------------------------------------------------------- compaction
__global__ __launch_bounds__(kWG) void k_count_keep(const unsigned char* __restrict__ keep, int64_t n,
                                                    long long* __restrict__ counts)
{
    __shared__ int red[kWG];
    const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
    red[threadIdx.x] = (t < n && keep[t]) ? 1 : 0;
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

// kept entry t goes to base + offsets[block] + (kept entries before it in its block)
__global__ __launch_bounds__(kWG) void k_scatter(const unsigned char* __restrict__ keep, int64_t n,
                                                 const long long* __restrict__ offsets,
                                                 const double* __restrict__ masses, int nst,
                                                 const double* __restrict__ mags, int nf, int64_t index0,
                                                 double* __restrict__ omasses, double* __restrict__ omags,
                                                 int64_t* __restrict__ oindex)
{
    __shared__ int wcount[kWG / 64];
    const int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool k = t < n && keep[t];
    const unsigned long long bal = __ballot(k);
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();
    if (!k) return;
    int before = __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; ++w) before += wcount[w];
    const int64_t o = offsets[blockIdx.x] + before;
    for (int s = 0; s < nst; ++s) omasses[o * nst + s] = masses[t * nst + s];
    for (int f = 0; f < nf; ++f) omags[o * nf + f] = mags[t * nf + f];
    oindex[o] = index0 + t;
}

// per-workgroup minimum and maximum of each filter's finite magnitudes
__global__ __launch_bounds__(kWG) void k_mag_range(const double* __restrict__ mags, int64_t n, int nf,
                                                   double* __restrict__ part)
{
    __shared__ double lo_s[kWG], hi_s[kWG];
    for (int f = 0; f < nf; ++f) {
        double lo = INFINITY, hi = -INFINITY;
        for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
            const double v = mags[i * nf + f];
            if (isfinite(v)) {
                lo = fmin(lo, v);
                hi = fmax(hi, v);
            }
        }
        lo_s[threadIdx.x] = lo;
        hi_s[threadIdx.x] = hi;
        __syncthreads();
        for (int w = kWG / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                lo_s[threadIdx.x] = fmin(lo_s[threadIdx.x], lo_s[threadIdx.x + w]);
                hi_s[threadIdx.x] = fmax(hi_s[threadIdx.x], hi_s[threadIdx.x + w]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[((int64_t)blockIdx.x * nf + f) * 2] = lo_s[0];
            part[((int64_t)blockIdx.x * nf + f) * 2 + 1] = hi_s[0];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- binning
// Left-closed bins: bin = floor((v - first) / step), in range when
// 0 <= bin < n.  Weights in 64-bit fixed point (scale) or unit counts.
__global__ __launch_bounds__(kWG) void k_bin(const double* __restrict__ mags, int nf, int ia, int ib, int iy,
                                             const double* __restrict__ xs, const double* __restrict__ ys,
                                             const double* __restrict__ w, int64_t n, double x0, double dx,
                                             int64_t nx, double y0, double dy, int64_t ny, double scale,
                                             unsigned long long* __restrict__ hist)
{
#pragma clang fp contract(off)
    const int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (i >= n) return;
    double x, y;
    if (mags) {
        x = mags[i * nf + ia] - mags[i * nf + ib];
        y = mags[i * nf + iy];
    } else {
        x = xs[i];
        y = ys[i];
    }
    const double fx = floor((x - x0) / dx), fy = floor((y - y0) / dy);
    if (!(fx >= 0.0 && fx < (double)nx && fy >= 0.0 && fy < (double)ny)) return;
    const long long v = w ? llrint(w[i] * scale) : 1LL;
    atomicAdd(hist + (int64_t)fx + nx * (int64_t)fy, (unsigned long long)v);
}

__global__ __launch_bounds__(kWG) void k_hist_to_double(const unsigned long long* __restrict__ h, int64_t n,
                                                        double inv_scale, double* __restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (k < n) out[k] = (double)(long long)h[k] * inv_scale;
}

}  // namespace mock
}  // namespace sfh

using namespace sfh;
using namespace sfh::mock;

namespace {

int catalog_reserve(sfh_catalog* c, int64_t need)
{
    sfh_ctx* ctx = c->ctx;
    if (need <= c->cap) return SFH_OK;
    const int64_t cap = std::max<int64_t>(need, c->cap + c->cap / 2 + 1024);
    double *m = nullptr, *g = nullptr;
    int64_t* ix = nullptr;
    if (c->nstars > 0) SFH_HIP_TRY(ctx, hipMalloc(&m, (size_t)cap * c->nstars * sizeof(double)));
    SFH_HIP_TRY(ctx, hipMalloc(&g, (size_t)cap * std::max(c->nfilters, 1) * sizeof(double)));
    SFH_HIP_TRY(ctx, hipMalloc(&ix, (size_t)cap * sizeof(int64_t)));
    if (c->n > 0) {
        if (m) SFH_HIP_TRY(ctx, hipMemcpyAsync(m, c->d_masses, (size_t)c->n * c->nstars * sizeof(double),
                                               hipMemcpyDeviceToDevice, ctx->stream));
        SFH_HIP_TRY(ctx, hipMemcpyAsync(g, c->d_mags, (size_t)c->n * c->nfilters * sizeof(double),
                                        hipMemcpyDeviceToDevice, ctx->stream));
        SFH_HIP_TRY(ctx, hipMemcpyAsync(ix, c->d_index, (size_t)c->n * sizeof(int64_t), hipMemcpyDeviceToDevice,
                                        ctx->stream));
    }
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    hipFree(c->d_masses);
    hipFree(c->d_mags);
    hipFree(c->d_index);
    c->d_masses = m;
    c->d_mags = g;
    c->d_index = ix;
    c->cap = cap;
    return SFH_OK;
}

// Append the kept entries among the first n of a staged batch to c.
int compact_append(sfh_catalog* c, const unsigned char* keep, int64_t n, const double* masses, const double* mags,
                   int64_t index0)
{
    sfh_ctx* ctx = c->ctx;
    if (n == 0) return SFH_OK;
    const int64_t nb = cdiv(n, kWG);
    void* p = nullptr;
    if (int rc = scratch(ctx, "mock_counts", (size_t)nb * sizeof(long long), &p)) return rc;
    long long* d_counts = (long long*)p;
    hipLaunchKernelGGL(k_count_keep, dim3((unsigned)nb), dim3(kWG), 0, ctx->stream, keep, n, d_counts);
    if (int rc = check_launch(ctx, "k_count_keep")) return rc;
    std::vector<long long> counts(nb);
    SFH_HIP_TRY(ctx, hipMemcpyAsync(counts.data(), d_counts, nb * sizeof(long long), hipMemcpyDeviceToHost,
                                    ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    long long tot = 0;
    for (int64_t b = 0; b < nb; ++b) {
        const long long k = counts[b];
        counts[b] = c->n + tot;
        tot += k;
    }
    if (tot == 0) return SFH_OK;
    if (int rc = catalog_reserve(c, c->n + tot)) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_counts, counts.data(), nb * sizeof(long long), hipMemcpyHostToDevice,
                                    ctx->stream));
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)nb), dim3(kWG), 0, ctx->stream, keep, n,
                       (const long long*)d_counts, masses, c->nstars, mags, c->nfilters, index0, c->d_masses,
                       c->d_mags, c->d_index);
    if (int rc = check_launch(ctx, "k_scatter")) return rc;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    c->n += tot;
    return SFH_OK;
}

int to_table(sfh_ctx* ctx, const sfh_table* t, double* dst, Table& out)
{
    SFH_HIP_TRY(ctx, hipMemcpyAsync(dst, t->values, (size_t)t->length * sizeof(double), hipMemcpyHostToDevice,
                                    ctx->stream));
    out = Table{t->first, t->step, t->length, dst};
    return SFH_OK;
}

bool table_ok(const sfh_table* t) { return t && t->values && t->length >= 2 && t->step > 0.0; }

}  // namespace

extern "C" {

int sfh_generate_stars(sfh_ctx* ctx, int64_t npoints, const double* mini, int32_t nfilters, const double* mags,
                       int32_t mode, double limit, int32_t limit_filter, double mmin_cut, double mag_lim,
                       int32_t mag_lim_filter, const sfh_imf* imf, const sfh_binary_model* binary, uint64_t seed,
                       uint64_t stream, sfh_catalog** out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, out && mini && mags && imf && binary, "sfh_generate_stars: null pointer");
    *out = nullptr;
    SFH_ARG_CHECK(ctx, npoints >= 2 && nfilters >= 1 && nfilters <= kMaxFilters,
                  "sfh_generate_stars: need >= 2 isochrone points and 1..%d filters", kMaxFilters);
    SFH_ARG_CHECK(ctx, mode == 0 || mode == 1, "sfh_generate_stars: mode 0 (mass) or 1 (luminosity)");
    SFH_ARG_CHECK(ctx, limit > 0.0 && std::isfinite(limit), "sfh_generate_stars: limit must be positive and finite");
    SFH_ARG_CHECK(ctx, mode == 0 || (limit_filter >= 0 && limit_filter < nfilters),
                  "sfh_generate_stars: limit filter out of range");
    SFH_ARG_CHECK(ctx, mode == 0 || std::isinf(mag_lim) || (mag_lim_filter >= 0 && mag_lim_filter < nfilters),
                  "sfh_generate_stars: mag_lim filter out of range");
    for (int64_t i = 0; i + 1 < npoints; ++i)
        SFH_ARG_CHECK(ctx, mini[i] < mini[i + 1], "sfh_generate_stars: masses must increase strictly");
    SFH_ARG_CHECK(ctx, imf->nseg >= 1 && imf->nseg <= kMaxSeg && imf->breaks && imf->slopes && imf->coefs,
                  "sfh_generate_stars: the IMF needs 1..%d power-law segments", kMaxSeg);
    SFH_ARG_CHECK(ctx, binary->kind >= 0 && binary->kind <= 2 && binary->fraction >= 0.0 && binary->fraction <= 1.0,
                  "sfh_generate_stars: binary model kind 0..2 with 0 <= fraction <= 1");
    SFH_ARG_CHECK(ctx, binary->kind != 2 || (binary->qmin >= 0.0 && binary->qmax <= 1.0 && binary->qmin <= binary->qmax),
                  "sfh_generate_stars: mass ratios within [0, 1]");
    if (int rc = bind(ctx)) return rc;

    GenArgs a{};
    a.npoints = npoints;
    a.nf = nfilters;
    a.nst = binary->kind == 0 ? 1 : 2;
    a.mode = mode;
    a.limf = limit_filter;
    a.mlf = mag_lim_filter;
    a.mmin_cut = mmin_cut;
    a.mag_lim = mag_lim;
    a.scale = std::ldexp(1.0, 62) / limit;
    a.bkind = binary->kind;
    a.bfrac = binary->fraction;
    a.qmin = binary->qmin;
    a.qmax = binary->qmax;
    a.key0 = seed;
    a.key1 = stream;
    // IMF segments and their integrals
    Imf& f = a.imf;
    f.nseg = imf->nseg;
    f.cum[0] = 0.0;
    double mean_num = 0.0;
    for (int k = 0; k < f.nseg; ++k) {
        const double lo = imf->breaks[k], hi = imf->breaks[k + 1], s = imf->slopes[k], cf = imf->coefs[k];
        SFH_ARG_CHECK(ctx, lo > 0.0 && hi > lo && cf > 0.0 && std::isfinite(s),
                      "sfh_generate_stars: IMF segments need 0 < lo < hi and coef > 0");
        f.lo[k] = lo;
        f.lo[k + 1] = hi;
        f.a[k] = 1.0 - s;
        f.coef[k] = cf;
        const double I = f.a[k] == 0.0 ? cf * std::log(hi / lo)
                                       : cf * (std::pow(hi, f.a[k]) - std::pow(lo, f.a[k])) / f.a[k];
        f.cum[k + 1] = f.cum[k] + I;
        const double a2 = 2.0 - s;
        mean_num += a2 == 0.0 ? cf * std::log(hi / lo) : cf * (std::pow(hi, a2) - std::pow(lo, a2)) / a2;
    }
    const double mean_mass = mean_num / f.cum[f.nseg];

    void* p = nullptr;
    if (int rc = scratch(ctx, "mock_iso", (size_t)npoints * (1 + nfilters) * sizeof(double), &p)) return rc;
    double* d_iso = (double*)p;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_iso, mini, npoints * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_iso + npoints, mags, (size_t)npoints * nfilters * sizeof(double),
                                    hipMemcpyHostToDevice, ctx->stream));
    a.mini = d_iso;
    a.mags = d_iso + npoints;

    // the catalogue takes its reference on the context only once it is complete,
    // so a failure never releases the context while its lock is held
    sfh_catalog* c = new sfh_catalog();
    c->ctx = ctx;
    c->nstars = a.nst;
    c->nfilters = nfilters;
    auto fail = [&](int rc) {
        hipFree(c->d_masses);
        hipFree(c->d_mags);
        hipFree(c->d_index);
        delete c;
        return rc;
    };

    // first batch: the expected number of systems for a birth mass (plus
    // margin); luminosity limits start at 2^20 and size later batches from
    // the mean contribution seen so far
    const double sys_mass = mean_mass * (binary->kind == 1 ? 1.0 + binary->fraction : 1.0);
    int64_t batch = mode == 0 ? (int64_t)(limit / sys_mass * 1.02 + 8.0 * std::sqrt(limit / sys_mass) + 1024.0)
                              : (int64_t)1 << 20;
    const int64_t kMaxBatch = (int64_t)1 << 24;
    long long total = 0;
    int64_t s0 = 0;
    bool done = false;
    while (!done) {
        batch = std::min(std::max<int64_t>(batch, 4096), kMaxBatch);
        const int64_t nb = cdiv(batch, kWG);
        const size_t bytes = (size_t)batch * (sizeof(long long) + (a.nst + nfilters) * sizeof(double) + 1) +
                             (size_t)nb * sizeof(long long) + 256;
        if (int rc = scratch(ctx, "mock_stage", bytes, &p)) return fail(rc);
        long long* d_q = (long long*)p;
        double* d_m = (double*)(d_q + batch);
        double* d_g = d_m + (size_t)batch * a.nst;
        long long* d_bs = (long long*)(d_g + (size_t)batch * nfilters);
        unsigned char* d_keep = (unsigned char*)(d_bs + nb);
        {
            TimedLaunch tl(ctx, "mock_generate");
            hipLaunchKernelGGL(k_generate, dim3((unsigned)nb), dim3(kWG), 0, ctx->stream, a, s0, batch, d_q, d_m,
                               d_g, d_keep, d_bs);
        }
        if (int rc = check_launch(ctx, "k_generate")) return fail(rc);
        std::vector<long long> bs(nb);
        if (hipMemcpyAsync(bs.data(), d_bs, nb * sizeof(long long), hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return fail(set_error(ctx, SFH_ERR_DEVICE, "sfh_generate_stars: copy of the batch sums failed"));
        int64_t kend = batch;
        for (int64_t b = 0; b < nb; ++b) {
            if (total + bs[b] < kTarget) {
                total += bs[b];
                continue;
            }
            // the crossing block: add its systems one by one
            const int64_t t0 = b * kWG, t1 = std::min<int64_t>(batch, t0 + kWG);
            std::vector<long long> qs(t1 - t0);
            if (hipMemcpy(qs.data(), d_q + t0, (t1 - t0) * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(set_error(ctx, SFH_ERR_DEVICE, "sfh_generate_stars: copy of the crossing block failed"));
            for (int64_t t = t0; t < t1; ++t) {
                total += qs[t - t0];
                if (total >= kTarget) {
                    kend = t + 1;
                    break;
                }
            }
            done = true;
            break;
        }
        if (int rc = compact_append(c, d_keep, kend, d_m, d_g, s0)) return fail(rc);
        s0 += kend;
        // a limit no system can reach would keep the reference's loop running
        // for ever: stop when a whole batch added nothing, or past 2^40 systems
        if (!done && total == 0)
            return fail(set_error(ctx, SFH_ERR_DOMAIN, "sfh_generate_stars: %lld systems drawn and none contributes "
                                  "to the limit (no isochrone mass in the IMF range, or no flux in the limit filter)",
                                  (long long)s0));
        if (!done && s0 > ((int64_t)1 << 40))
            return fail(set_error(ctx, SFH_ERR_DOMAIN, "sfh_generate_stars: limit not reached after 2^40 systems"));
        if (!done) {
            // next batch from the mean contribution so far
            const double per = (double)total / (double)s0;
            const double need = per > 0.0 ? ((double)kTarget - (double)total) / per : (double)kMaxBatch;
            batch = (int64_t)(need * 1.02 + 8.0 * std::sqrt(std::max(need, 1.0)) + 1024.0);
        }
    }
    c->sampled = s0;
    resolve_timing(ctx);
    ctx_retain(ctx);
    *out = c;
    return SFH_OK;
}

int sfh_catalog_from_host(sfh_ctx* ctx, int64_t n, int32_t nfilters, const double* mags, sfh_catalog** out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, out && (mags || n == 0) && n >= 0 && nfilters >= 1 && nfilters <= kMaxFilters,
                  "sfh_catalog_from_host: need 1..%d filters", kMaxFilters);
    *out = nullptr;
    if (int rc = bind(ctx)) return rc;
    sfh_catalog* c = new sfh_catalog();
    c->ctx = ctx;
    c->nstars = 0;
    c->nfilters = nfilters;
    int rc = catalog_reserve(c, std::max<int64_t>(n, 1));
    if (!rc && n > 0 &&
        hipMemcpy(c->d_mags, mags, (size_t)n * nfilters * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(ctx, SFH_ERR_DEVICE, "sfh_catalog_from_host: copy failed");
    if (rc) {
        hipFree(c->d_mags);
        hipFree(c->d_index);
        delete c;
        return rc;
    }
    c->n = n;
    c->sampled = n;
    ctx_retain(ctx);
    *out = c;
    return SFH_OK;
}

int sfh_catalog_shape(const sfh_catalog* c, int64_t* n, int32_t* nstars, int32_t* nfilters, int64_t* sampled)
{
    if (!c) return SFH_ERR_ARGUMENT;
    if (n) *n = c->n;
    if (nstars) *nstars = c->nstars;
    if (nfilters) *nfilters = c->nfilters;
    if (sampled) *sampled = c->sampled;
    return SFH_OK;
}

int sfh_catalog_copy(const sfh_catalog* c, double* masses, double* mags, int64_t* index)
{
    if (!c) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = c->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    if (c->n == 0) return SFH_OK;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (masses && c->nstars > 0)
        SFH_HIP_TRY(ctx, hipMemcpy(masses, c->d_masses, (size_t)c->n * c->nstars * sizeof(double),
                                   hipMemcpyDeviceToHost));
    if (mags)
        SFH_HIP_TRY(ctx, hipMemcpy(mags, c->d_mags, (size_t)c->n * c->nfilters * sizeof(double),
                                   hipMemcpyDeviceToHost));
    if (index) SFH_HIP_TRY(ctx, hipMemcpy(index, c->d_index, (size_t)c->n * sizeof(int64_t), hipMemcpyDeviceToHost));
    return SFH_OK;
}

int sfh_catalog_range(const sfh_catalog* c, double* lo, double* hi)
{
    if (!c || !lo || !hi) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = c->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    const int nf = c->nfilters;
    for (int f = 0; f < nf; ++f) {
        lo[f] = INFINITY;
        hi[f] = -INFINITY;
    }
    if (c->n == 0) return SFH_OK;
    const int64_t nb = std::min<int64_t>(cdiv(c->n, kWG), 1024);
    void* p = nullptr;
    if (int rc = scratch(ctx, "mock_range", (size_t)nb * nf * 2 * sizeof(double), &p)) return rc;
    hipLaunchKernelGGL(k_mag_range, dim3((unsigned)nb), dim3(kWG), 0, ctx->stream, (const double*)c->d_mags, c->n,
                       nf, (double*)p);
    if (int rc = check_launch(ctx, "k_mag_range")) return rc;
    std::vector<double> part((size_t)nb * nf * 2);
    SFH_HIP_TRY(ctx, hipMemcpyAsync(part.data(), p, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t b = 0; b < nb; ++b)
        for (int f = 0; f < nf; ++f) {
            lo[f] = std::min(lo[f], part[(b * nf + f) * 2]);
            hi[f] = std::max(hi[f], part[(b * nf + f) * 2 + 1]);
        }
    return SFH_OK;
}

int sfh_catalog_destroy(sfh_catalog* c)
{
    if (!c) return SFH_OK;
    sfh_ctx* ctx = c->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        hipFree(c->d_masses);
        hipFree(c->d_mags);
        hipFree(c->d_index);
        delete c;
    }
    ctx_release(ctx);
    return SFH_OK;
}

int sfh_model_cmd(sfh_ctx* ctx, const sfh_catalog* in, const sfh_table* err, const sfh_table* completeness,
                  const sfh_table* bias, uint64_t seed, uint64_t stream, sfh_catalog** out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, in && err && completeness && out, "sfh_model_cmd: null pointer");
    *out = nullptr;
    SFH_ARG_CHECK(ctx, in->ctx == ctx, "sfh_model_cmd: catalogue from another context");
    const int nf = in->nfilters;
    for (int f = 0; f < nf; ++f) {
        SFH_ARG_CHECK(ctx, table_ok(&err[f]) && table_ok(&completeness[f]) && (!bias || table_ok(&bias[f])),
                      "sfh_model_cmd: tables need >= 2 values on an increasing grid (filter %d)", f);
    }
    if (int rc = bind(ctx)) return rc;
    ObsArgs a{};
    a.nf = nf;
    a.has_bias = bias ? 1 : 0;
    a.key0 = seed;
    a.key1 = stream;
    size_t tab = 0;
    for (int f = 0; f < nf; ++f) tab += err[f].length + completeness[f].length + (bias ? bias[f].length : 0);
    const int64_t n = in->n;
    void* p = nullptr;
    const size_t bytes = tab * sizeof(double) + (size_t)n * nf * sizeof(double) + (size_t)n + 256;
    if (int rc = scratch(ctx, "mock_obs", bytes, &p)) return rc;
    double* cur = (double*)p;
    for (int f = 0; f < nf; ++f) {
        if (int rc = to_table(ctx, &err[f], cur, a.err[f])) return rc;
        cur += err[f].length;
        if (int rc = to_table(ctx, &completeness[f], cur, a.compl_[f])) return rc;
        cur += completeness[f].length;
        if (bias) {
            if (int rc = to_table(ctx, &bias[f], cur, a.bias[f])) return rc;
            cur += bias[f].length;
        }
    }
    double* d_out = cur;
    unsigned char* d_keep = (unsigned char*)(d_out + (size_t)n * nf);
    sfh_catalog* c = new sfh_catalog();
    c->ctx = ctx;
    c->nstars = 0;
    c->nfilters = nf;
    c->sampled = n;
    if (n > 0) {
        {
            TimedLaunch tl(ctx, "mock_observe");
            hipLaunchKernelGGL(k_observe, dim3((unsigned)cdiv(n, kWG)), dim3(kWG), 0, ctx->stream,
                               (const double*)in->d_mags, n, a, d_out, d_keep);
        }
        int rc = check_launch(ctx, "k_observe");
        if (!rc) rc = compact_append(c, d_keep, n, nullptr, d_out, 0);
        if (rc) {
            hipFree(c->d_mags);
            hipFree(c->d_index);
            delete c;
            return rc;
        }
    }
    resolve_timing(ctx);
    ctx_retain(ctx);
    *out = c;
    return SFH_OK;
}

static int bin_common(sfh_ctx* ctx, const double* d_mags, int nf, int ia, int ib, int iy, const double* d_x,
                      const double* d_y, const double* d_w, int64_t n, double wsum_abs, sfh_range xe, sfh_range ye,
                      double* out)
{
    const int64_t nx = xe.length - 1, ny = ye.length - 1, nb = nx * ny;
    const double scale = d_w ? (wsum_abs > 0.0 ? std::ldexp(1.0, 62) / wsum_abs : 1.0) : 1.0;
    void* p = nullptr;
    if (int rc = scratch(ctx, "mock_hist", (size_t)nb * (sizeof(unsigned long long) + sizeof(double)), &p))
        return rc;
    unsigned long long* d_h = (unsigned long long*)p;
    double* d_out = (double*)(d_h + nb);
    SFH_HIP_TRY(ctx, hipMemsetAsync(d_h, 0, (size_t)nb * sizeof(unsigned long long), ctx->stream));
    if (n > 0) {
        TimedLaunch tl(ctx, "mock_bin");
        hipLaunchKernelGGL(k_bin, dim3((unsigned)cdiv(n, kWG)), dim3(kWG), 0, ctx->stream, d_mags, nf, ia, ib, iy,
                           d_x, d_y, d_w, n, xe.first, xe.step, nx, ye.first, ye.step, ny, scale, d_h);
    }
    if (int rc = check_launch(ctx, "k_bin")) return rc;
    hipLaunchKernelGGL(k_hist_to_double, dim3((unsigned)cdiv(nb, kWG)), dim3(kWG), 0, ctx->stream,
                       (const unsigned long long*)d_h, nb, 1.0 / scale, d_out);
    if (int rc = check_launch(ctx, "k_hist_to_double")) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(out, d_out, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    return SFH_OK;
}

int sfh_bin_cmd(sfh_ctx* ctx, int64_t n, const double* colors, const double* mags, const double* weights,
                sfh_range xedges, sfh_range yedges, double* out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, out && n >= 0 && (n == 0 || (colors && mags)), "sfh_bin_cmd: null pointer");
    SFH_ARG_CHECK(ctx, xedges.length >= 2 && yedges.length >= 2 && xedges.step > 0 && yedges.step > 0,
                  "sfh_bin_cmd: edges need >= 2 increasing values");
    if (int rc = bind(ctx)) return rc;
    double wabs = 0.0;
    if (weights)
        for (int64_t i = 0; i < n; ++i) wabs += std::fabs(weights[i]);
    SFH_ARG_CHECK(ctx, std::isfinite(wabs), "sfh_bin_cmd: non-finite weights");
    void* p = nullptr;
    if (int rc = scratch(ctx, "mock_points", (size_t)std::max<int64_t>(n, 1) * 3 * sizeof(double), &p)) return rc;
    double* d_x = (double*)p;
    double* d_y = d_x + n;
    double* d_w = weights ? d_y + n : nullptr;
    if (n > 0) {
        SFH_HIP_TRY(ctx, hipMemcpyAsync(d_x, colors, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        SFH_HIP_TRY(ctx, hipMemcpyAsync(d_y, mags, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        if (weights)
            SFH_HIP_TRY(ctx, hipMemcpyAsync(d_w, weights, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    return bin_common(ctx, nullptr, 0, 0, 0, 0, d_x, d_y, d_w, n, wabs, xedges, yedges, out);
}

int sfh_bin_cmd_catalog(sfh_ctx* ctx, const sfh_catalog* c, int32_t color_a, int32_t color_b, int32_t y_filter,
                        sfh_range xedges, sfh_range yedges, double* out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, c && out && c->ctx == ctx, "sfh_bin_cmd_catalog: null pointer or foreign catalogue");
    SFH_ARG_CHECK(ctx, color_a >= 0 && color_a < c->nfilters && color_b >= 0 && color_b < c->nfilters &&
                       y_filter >= 0 && y_filter < c->nfilters,
                  "sfh_bin_cmd_catalog: filter index out of range");
    SFH_ARG_CHECK(ctx, xedges.length >= 2 && yedges.length >= 2 && xedges.step > 0 && yedges.step > 0,
                  "sfh_bin_cmd_catalog: edges need >= 2 increasing values");
    if (int rc = bind(ctx)) return rc;
    return bin_common(ctx, c->d_mags, c->nfilters, color_a, color_b, y_filter, nullptr, nullptr, nullptr, c->n, 0.0,
                      xedges, yedges, out);
}

}  // extern "C"

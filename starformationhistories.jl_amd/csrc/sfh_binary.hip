// sfh_binary.hip -- the O(P^2) half of binary_hess(::RandomBinaryPairs, ...)
// (src/StarFormationHistories.jl:627-677): every pair of isochrone mass
// intervals (i < j) as an unresolved binary, binned into the Hess diagram.
//
// Workgroup (i, t): primary interval i of template t; its threads walk the
// secondaries j (many templates in one launch, each with its own interval
// range, weight factor and fixed-point scale).
// Per pair (FP64): fluxes of points i and j summed per colour filter,
// magnitude -5 log10(f) / 2 (flux2mag, src/utilities.jl:67), weight
// dM_i dM_j (imf_i + imf_i+1)(imf_j + imf_j+1) prefac times the completeness
// of both filters, bias added (measured = intrinsic + bias), colour and y
// magnitude, left-closed bin.  Completeness and bias are user callables in
// the reference; here they arrive tabulated on a uniform magnitude grid and
// are interpolated linearly.
//
// Accumulation is deterministic: weights are added as 64-bit fixed point
// (scale 2^62 / W, W a bound on the total weight) with integer atomics, after
// the lanes of a wave that hit the same bin (consecutive secondaries mostly
// do) have summed their shares.
#include "sfh_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace sfh {
namespace bin {

constexpr int kWG = 256;

struct Table {
    double first, step;
    int64_t n;
    const double* v;
};

// linear interpolation, clamped to the table's ends
__device__ inline double lookup(const Table& t, double x)
{
    double u = (x - t.first) / t.step;
    u = fmin(fmax(u, 0.0), (double)(t.n - 1));
    int64_t k = (int64_t)u;
    if (k >= t.n - 1) k = t.n - 2;
    const double f = u - (double)k;
    return t.v[k] + f * (t.v[k + 1] - t.v[k]);
}

struct PairArgs {
    const double* flux_a;     // 10^(-0.4 m) of the interval's left point, colour filter a
    const double* flux_b;
    const double* dm;         // mini[i+1] - mini[i]
    const double* imfsum;     // imf[i] + imf[i+1]
    const int64_t* toff;      // template t owns intervals toff[t] .. toff[t+1]-1
    const double* prefac;     // per template
    const double* scale;      // fixed-point scale per template
    int y_is_a;
    Table ca, cb, ba, bb;     // completeness / bias of filters a, b
    double x0, dx, y0, dy;
    int64_t nx, ny;
    unsigned long long* hist; // ntemplates x (nx * ny)
};

__global__ __launch_bounds__(kWG) void k_binary_pairs(PairArgs a)
{
#pragma clang fp contract(off)
    const int64_t t = blockIdx.y;
    const int64_t base = a.toff[t], nint = a.toff[t + 1] - base;
    const int64_t i = blockIdx.x;
    if (i + 1 >= nint) return;
    const int lane = threadIdx.x & 63;
    const double* flux_a = a.flux_a + base;
    const double* flux_b = a.flux_b + base;
    const double* dmv = a.dm + base;
    const double* isv = a.imfsum + base;
    const double prefac = a.prefac[t], scale = a.scale[t];
    unsigned long long* hist = a.hist + t * a.nx * a.ny;
    const double fa_i = flux_a[i], fb_i = flux_b[i];
    const double dmp = dmv[i], mpint = isv[i];
    for (int64_t j0 = i + 1; j0 < nint; j0 += kWG) {
        const int64_t j = j0 + threadIdx.x;
        bool valid = j < nint;
        int64_t bin = -1;
        long long q = 0;
        if (valid) {
            const double ma = -5.0 * log10(fa_i + flux_a[j]) / 2.0;
            const double mb = -5.0 * log10(fb_i + flux_b[j]) / 2.0;
            double w = dmp * dmv[j] * mpint * isv[j] * prefac;            // :661
            w *= lookup(a.ca, ma) * lookup(a.cb, mb);                        // :670-671
            const double pa = ma + lookup(a.ba, ma), pb = mb + lookup(a.bb, mb);   // :677
            const double col = pa - pb, y = a.y_is_a ? pa : pb;              // :678-679
            const double fx = floor((col - a.x0) / a.dx), fy = floor((y - a.y0) / a.dy);
            if (fx >= 0.0 && fx < (double)a.nx && fy >= 0.0 && fy < (double)a.ny) {
                bin = (int64_t)fx + a.nx * (int64_t)fy;
                q = llrint(w * scale);
            }
        }
        valid = bin >= 0;
        // one atomic per distinct bin of the wave
        unsigned long long pending = __ballot(valid);
        while (pending) {
            const int leader = __ffsll((long long)pending) - 1;
            const int64_t lb = __shfl(bin, leader, 64);
            const bool mine = valid && bin == lb;
            long long v = mine ? q : 0;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == leader) atomicAdd(hist + lb, (unsigned long long)v);
            pending &= ~__ballot(mine);
        }
    }
}

// out[t, k] = hist[t, k] (1 / scale[t])
__global__ __launch_bounds__(kWG) void k_fixed_to_double(const unsigned long long* __restrict__ h, int64_t nb,
                                                         int64_t nt, const double* __restrict__ scale,
                                                         double* __restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (k < nb * nt) out[k] = (double)(long long)h[k] * (1.0 / scale[k / nb]);
}

}  // namespace bin
}  // namespace sfh

using namespace sfh;
using namespace sfh::bin;

namespace {

// Pair histograms of nt templates (points offsets[t] .. offsets[t+1]-1) into
// out (nt x nb, FP64, host).
int pairs_hist(sfh_ctx* ctx, int64_t nt, const int64_t* offsets, const double* mini, const double* imf_pdf,
               const double* mag_a, const double* mag_b, int y_is_a, const double* prefac, const sfh_table* const* tabs,
               sfh_range xedges, sfh_range yedges, double* out)
{
    const int64_t nx = xedges.length - 1, ny = yedges.length - 1, nb = nx * ny;
    // host: per-interval fluxes (mag2flux, src/utilities.jl:57) and factors;
    // per template the bound of its total weight and the fixed-point scale
    std::vector<int64_t> toff(nt + 1, 0);
    for (int64_t t = 0; t < nt; ++t) toff[t + 1] = toff[t] + std::max<int64_t>(offsets[t + 1] - offsets[t] - 1, 0);
    const int64_t ni = toff[nt];
    std::vector<double> fa(ni), fb(ni), dm(ni), is(ni), scale(nt, 1.0);
    auto tmax = [](const sfh_table* tb) {
        double m = 0.0;
        for (int64_t k = 0; k < tb->length; ++k) m = std::max(m, std::fabs(tb->values[k]));
        return m;
    };
    const double cmax = tmax(tabs[0]) * tmax(tabs[1]);
    int64_t max_nint = 0;
    for (int64_t t = 0; t < nt; ++t) {
        double asum = 0.0;
        for (int64_t k = 0; k < toff[t + 1] - toff[t]; ++k) {
            const int64_t p = offsets[t] + k, q = toff[t] + k;
            fa[q] = std::pow(10.0, 4.0 * (0.0 - mag_a[p]) / 10.0);
            fb[q] = std::pow(10.0, 4.0 * (0.0 - mag_b[p]) / 10.0);
            dm[q] = mini[p + 1] - mini[p];
            is[q] = imf_pdf[p] + imf_pdf[p + 1];
            asum += std::fabs(dm[q] * is[q]);
        }
        // bound on |sum of the template's pair weights|: prefac cmax_a cmax_b (sum_i |a_i|)^2 / 2
        const double wbound = std::fabs(prefac[t]) * cmax * asum * asum / 2.0;
        SFH_ARG_CHECK(ctx, std::isfinite(wbound), "sfh_binary_pairs_hist: non-finite weights (template %lld)",
                      (long long)t);
        scale[t] = wbound > 0.0 ? std::ldexp(1.0, 62) / wbound : 1.0;
        max_nint = std::max(max_nint, toff[t + 1] - toff[t]);
    }
    if (max_nint < 2 || nb == 0) {
        std::fill(out, out + nt * nb, 0.0);
        return SFH_OK;
    }
    SFH_ARG_CHECK(ctx, nt <= 65535, "sfh_binary_pairs_hist: at most 65535 templates per call");
    size_t tab_elems = 0;
    for (int k = 0; k < 4; ++k) tab_elems += (size_t)tabs[k]->length;
    const size_t dbl = 4 * (size_t)ni + tab_elems + 2 * (size_t)nt + (size_t)(nt + 1) + (size_t)nb * nt;
    void* p = nullptr;
    if (int rc = scratch(ctx, "bin_dbl", dbl * sizeof(double), &p)) return rc;
    void* ph = nullptr;
    if (int rc = scratch(ctx, "bin_hist", (size_t)nb * nt * sizeof(unsigned long long), &ph)) return rc;
    double* d = (double*)p;
    double* d_fa = d;
    double* d_fb = d_fa + ni;
    double* d_dm = d_fb + ni;
    double* d_is = d_dm + ni;
    double* d_tab = d_is + ni;
    double* d_pre = d_tab + tab_elems;
    double* d_scale = d_pre + nt;
    int64_t* d_toff = (int64_t*)(d_scale + nt);
    double* d_out = (double*)(d_toff + nt + 1);
    hipStream_t s = ctx->stream;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_fa, fa.data(), ni * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_fb, fb.data(), ni * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_dm, dm.data(), ni * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_is, is.data(), ni * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_pre, prefac, nt * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_scale, scale.data(), nt * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_toff, toff.data(), (nt + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    Table dt[4];
    double* cur = d_tab;
    for (int k = 0; k < 4; ++k) {
        SFH_HIP_TRY(ctx, hipMemcpyAsync(cur, tabs[k]->values, tabs[k]->length * sizeof(double),
                                        hipMemcpyHostToDevice, s));
        dt[k] = Table{tabs[k]->first, tabs[k]->step, tabs[k]->length, cur};
        cur += tabs[k]->length;
    }
    SFH_HIP_TRY(ctx, hipMemsetAsync(ph, 0, (size_t)nb * nt * sizeof(unsigned long long), s));
    PairArgs a{d_fa, d_fb, d_dm, d_is, d_toff, d_pre, d_scale, y_is_a ? 1 : 0, dt[0], dt[1], dt[2], dt[3],
               xedges.first, xedges.step, yedges.first, yedges.step, nx, ny, (unsigned long long*)ph};
    {
        TimedLaunch tl(ctx, "binary_pairs");
        hipLaunchKernelGGL(k_binary_pairs, dim3((unsigned)(max_nint - 1), (unsigned)nt), dim3(kWG), 0, s, a);
    }
    if (int rc = check_launch(ctx, "k_binary_pairs")) return rc;
    hipLaunchKernelGGL(k_fixed_to_double, dim3((unsigned)cdiv(nb * nt, kWG)), dim3(kWG), 0, s,
                       (const unsigned long long*)ph, nb, nt, (const double*)d_scale, d_out);
    if (int rc = check_launch(ctx, "k_fixed_to_double")) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(out, d_out, (size_t)nb * nt * sizeof(double), hipMemcpyDeviceToHost, s));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(s));
    resolve_timing(ctx);
    return SFH_OK;
}

bool tables_ok(const sfh_table* const* tabs)
{
    for (int k = 0; k < 4; ++k)
        if (!tabs[k] || !tabs[k]->values || tabs[k]->length < 2 || !(tabs[k]->step > 0)) return false;
    return true;
}

}  // namespace

extern "C" {

int sfh_binary_pairs_hist(sfh_ctx* ctx, int64_t npoints, const double* mini, const double* imf_pdf,
                          const double* mag_a, const double* mag_b, int y_is_a, double prefac,
                          const sfh_table* compl_a, const sfh_table* compl_b, const sfh_table* bias_a,
                          const sfh_table* bias_b, sfh_range xedges, sfh_range yedges, double* out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, mini && imf_pdf && mag_a && mag_b && out && npoints >= 0, "sfh_binary_pairs_hist: null pointer");
    SFH_ARG_CHECK(ctx, xedges.length >= 2 && yedges.length >= 2 && xedges.step > 0 && yedges.step > 0,
                  "sfh_binary_pairs_hist: edges need >= 2 increasing values");
    const sfh_table* tabs[4] = {compl_a, compl_b, bias_a, bias_b};
    SFH_ARG_CHECK(ctx, tables_ok(tabs), "sfh_binary_pairs_hist: tables need >= 2 values on an increasing grid");
    if (int rc = bind(ctx)) return rc;
    const int64_t offsets[2] = {0, npoints};
    return pairs_hist(ctx, 1, offsets, mini, imf_pdf, mag_a, mag_b, y_is_a, &prefac, tabs, xedges, yedges, out);
}

int sfh_binary_pairs_hist_batch(sfh_ctx* ctx, int64_t ntemplates, const int64_t* offsets, const double* mini,
                                const double* imf_pdf, const double* mag_a, const double* mag_b, int y_is_a,
                                const double* prefac, const sfh_table* compl_a, const sfh_table* compl_b,
                                const sfh_table* bias_a, const sfh_table* bias_b, sfh_range xedges, sfh_range yedges,
                                double* out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, ntemplates >= 0 && offsets && mini && imf_pdf && mag_a && mag_b && prefac && out,
                  "sfh_binary_pairs_hist_batch: null pointer");
    SFH_ARG_CHECK(ctx, xedges.length >= 2 && yedges.length >= 2 && xedges.step > 0 && yedges.step > 0,
                  "sfh_binary_pairs_hist_batch: edges need >= 2 increasing values");
    const sfh_table* tabs[4] = {compl_a, compl_b, bias_a, bias_b};
    SFH_ARG_CHECK(ctx, tables_ok(tabs), "sfh_binary_pairs_hist_batch: tables need >= 2 values on an increasing grid");
    for (int64_t t = 0; t < ntemplates; ++t)
        SFH_ARG_CHECK(ctx, offsets[t] >= 0 && offsets[t + 1] >= offsets[t],
                      "sfh_binary_pairs_hist_batch: offsets must not decrease");
    if (ntemplates == 0) return SFH_OK;
    if (int rc = bind(ctx)) return rc;
    return pairs_hist(ctx, ntemplates, offsets, mini, imf_pdf, mag_a, mag_b, y_is_a, prefac, tabs, xedges, yedges,
                      out);
}

}  // extern "C"

// sfh_binary.hip -- the O(P^2) half of binary_hess(::RandomBinaryPairs, ...)
// (src/StarFormationHistories.jl:627-677): every pair of isochrone mass
// intervals (i < j) as an unresolved binary, binned into the Hess diagram.
//
// One workgroup per primary interval i; its threads walk the secondaries j.
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
    const double* flux_a;     // 10^(-0.4 m) of points 0..n-2, colour filter a
    const double* flux_b;
    const double* dm;         // mini[i+1] - mini[i]
    const double* imfsum;     // imf[i] + imf[i+1]
    int64_t nint;             // intervals n - 1
    double prefac;
    int y_is_a;
    Table ca, cb, ba, bb;     // completeness / bias of filters a, b
    double x0, dx, y0, dy;
    int64_t nx, ny;
    double scale;             // fixed-point scale
    unsigned long long* hist; // nx * ny
};

__global__ __launch_bounds__(kWG) void k_binary_pairs(PairArgs a)
{
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x;
    const int lane = threadIdx.x & 63;
    const double fa_i = a.flux_a[i], fb_i = a.flux_b[i];
    const double dmp = a.dm[i], mpint = a.imfsum[i];
    for (int64_t j0 = i + 1; j0 < a.nint; j0 += kWG) {
        const int64_t j = j0 + threadIdx.x;
        bool valid = j < a.nint;
        int64_t bin = -1;
        long long q = 0;
        if (valid) {
            const double ma = -5.0 * log10(fa_i + a.flux_a[j]) / 2.0;
            const double mb = -5.0 * log10(fb_i + a.flux_b[j]) / 2.0;
            double w = dmp * a.dm[j] * mpint * a.imfsum[j] * a.prefac;     // :661
            w *= lookup(a.ca, ma) * lookup(a.cb, mb);                        // :670-671
            const double pa = ma + lookup(a.ba, ma), pb = mb + lookup(a.bb, mb);   // :677
            const double col = pa - pb, y = a.y_is_a ? pa : pb;              // :678-679
            const double fx = floor((col - a.x0) / a.dx), fy = floor((y - a.y0) / a.dy);
            if (fx >= 0.0 && fx < (double)a.nx && fy >= 0.0 && fy < (double)a.ny) {
                bin = (int64_t)fx + a.nx * (int64_t)fy;
                q = llrint(w * a.scale);
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
            if (lane == leader) atomicAdd(a.hist + lb, (unsigned long long)v);
            pending &= ~__ballot(mine);
        }
    }
}

__global__ __launch_bounds__(kWG) void k_fixed_to_double(const unsigned long long* __restrict__ h, int64_t n,
                                                         double inv_scale, double* __restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (k < n) out[k] = (double)(long long)h[k] * inv_scale;
}

}  // namespace bin
}  // namespace sfh

using namespace sfh;
using namespace sfh::bin;

extern "C" {

int sfh_binary_pairs_hist(sfh_ctx* ctx, int64_t npoints, const double* mini, const double* imf_pdf,
                          const double* mag_a, const double* mag_b, int y_is_a, double prefac,
                          const sfh_table* compl_a, const sfh_table* compl_b, const sfh_table* bias_a,
                          const sfh_table* bias_b, sfh_range xedges, sfh_range yedges, double* out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, mini && imf_pdf && mag_a && mag_b && out && compl_a && compl_b && bias_a && bias_b,
                  "sfh_binary_pairs_hist: null pointer");
    SFH_ARG_CHECK(ctx, xedges.length >= 2 && yedges.length >= 2 && xedges.step > 0 && yedges.step > 0,
                  "sfh_binary_pairs_hist: edges need >= 2 increasing values");
    const sfh_table* tabs[4] = {compl_a, compl_b, bias_a, bias_b};
    for (const sfh_table* t : tabs)
        SFH_ARG_CHECK(ctx, t->values && t->length >= 2 && t->step > 0,
                      "sfh_binary_pairs_hist: tables need >= 2 values on an increasing grid");
    if (int rc = bind(ctx)) return rc;
    const int64_t nx = xedges.length - 1, ny = yedges.length - 1, nb = nx * ny;
    const int64_t nint = npoints - 1;
    if (nint < 2) {
        std::fill(out, out + nb, 0.0);
        return SFH_OK;
    }
    // host: per-point fluxes (mag2flux, src/utilities.jl:57) and interval factors
    std::vector<double> fa(nint), fb(nint), dm(nint), is(nint);
    double asum = 0.0;
    for (int64_t i = 0; i < nint; ++i) {
        fa[i] = std::pow(10.0, 4.0 * (0.0 - mag_a[i]) / 10.0);
        fb[i] = std::pow(10.0, 4.0 * (0.0 - mag_b[i]) / 10.0);
        dm[i] = mini[i + 1] - mini[i];
        is[i] = imf_pdf[i] + imf_pdf[i + 1];
        asum += std::fabs(dm[i] * is[i]);
    }
    auto tmax = [](const sfh_table* t) {
        double m = 0.0;
        for (int64_t k = 0; k < t->length; ++k) m = std::max(m, std::fabs(t->values[k]));
        return m;
    };
    // bound on |sum of all pair weights|: prefac cmax_a cmax_b (sum_i |a_i|)^2 / 2
    const double wbound = std::fabs(prefac) * tmax(compl_a) * tmax(compl_b) * asum * asum / 2.0;
    if (!(wbound > 0.0) || !std::isfinite(wbound)) {
        std::fill(out, out + nb, 0.0);
        SFH_ARG_CHECK(ctx, wbound == 0.0, "sfh_binary_pairs_hist: non-finite weights");
        return SFH_OK;
    }
    const double scale = std::ldexp(1.0, 62) / wbound;

    // device buffers: 4 interval arrays, 4 tables, the histogram, the FP64 output
    size_t tab_elems = 0;
    for (const sfh_table* t : tabs) tab_elems += (size_t)t->length;
    const size_t dbl = 4 * (size_t)nint + tab_elems + (size_t)nb;
    void* p = nullptr;
    if (int rc = scratch(ctx, "bin_dbl", dbl * sizeof(double), &p)) return rc;
    void* ph = nullptr;
    if (int rc = scratch(ctx, "bin_hist", (size_t)nb * sizeof(unsigned long long), &ph)) return rc;
    double* d = (double*)p;
    double* d_fa = d;
    double* d_fb = d_fa + nint;
    double* d_dm = d_fb + nint;
    double* d_is = d_dm + nint;
    double* d_tab = d_is + nint;
    double* d_out = d_tab + tab_elems;
    hipStream_t s = ctx->stream;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_fa, fa.data(), nint * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_fb, fb.data(), nint * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_dm, dm.data(), nint * sizeof(double), hipMemcpyHostToDevice, s));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_is, is.data(), nint * sizeof(double), hipMemcpyHostToDevice, s));
    Table dt[4];
    double* cur = d_tab;
    for (int k = 0; k < 4; ++k) {
        SFH_HIP_TRY(ctx, hipMemcpyAsync(cur, tabs[k]->values, tabs[k]->length * sizeof(double),
                                        hipMemcpyHostToDevice, s));
        dt[k] = Table{tabs[k]->first, tabs[k]->step, tabs[k]->length, cur};
        cur += tabs[k]->length;
    }
    SFH_HIP_TRY(ctx, hipMemsetAsync(ph, 0, (size_t)nb * sizeof(unsigned long long), s));
    PairArgs a{d_fa, d_fb, d_dm, d_is, nint, prefac, y_is_a ? 1 : 0, dt[0], dt[1], dt[2], dt[3],
               xedges.first, xedges.step, yedges.first, yedges.step, nx, ny, scale, (unsigned long long*)ph};
    {
        TimedLaunch tl(ctx, "binary_pairs");
        hipLaunchKernelGGL(k_binary_pairs, dim3((unsigned)(nint - 1)), dim3(kWG), 0, s, a);
    }
    if (int rc = check_launch(ctx, "k_binary_pairs")) return rc;
    hipLaunchKernelGGL(k_fixed_to_double, dim3((unsigned)cdiv(nb, kWG)), dim3(kWG), 0, s,
                       (const unsigned long long*)ph, nb, 1.0 / scale, d_out);
    if (int rc = check_launch(ctx, "k_fixed_to_double")) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(out, d_out, (size_t)nb * sizeof(double), hipMemcpyDeviceToHost, s));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(s));
    resolve_timing(ctx);
    return SFH_OK;
}

}  // extern "C"

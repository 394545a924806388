// sfh_psf.h -- per-point footprint preparation (FP64, bit-exact index math)
// and per-bin evaluation (FP32) of the two pixel-integrated Gaussian kernels of
// StarFormationHistories.jl, shared by the gfx950 kernels and a host build
// used to check the FP32 arithmetic against the FP64 oracle on the CPU.
//
// Reference: src/StarFormationHistories.jl
//   GaussianPSFCovariant + gaussian_psf_covariant  :272-338
//   addstar!(::Histogram, ::RealSpaceKernel)        :366-408
//   GaussianPSFAsymmetric + gaussian_int_general     :191-265
//   addstar!(::AbstractMatrix, ::PixelSpaceKernel)  :348-364
//   bin_cmd_smooth                                  :574-621
//
// Both kernels reduce to D(u, e) = erf(u + e) - erf(u - e) at a bin centre u
// (in units of sqrt(2) sigma) and half bin width e:
//   covariant  : A hy / (2 sqrt(2 pi) sy) * sum_k w_k exp(-t_k^2/2) D(u_i + g t_k, ex)
//                with t_k = (y_j + xi_k hy - y0)/sy (3-point Gauss-Legendre, :303-304),
//                g = c sy / (sqrt(2) sx)  (the reference's Δx = dx + c Δy, :328);
//   asymmetric : A/4 * D(u_i, ex) * D(v_j, ey)  (:203-204; the two sign flips cancel).
// For e < kSeriesEps, D is evaluated without the erf-difference cancellation:
//   D = (4 e / sqrt(pi)) exp(-u^2) sum_n H_2n(u) e^2n / (2n+1)!
// (mean of exp(-(u+s)^2) over |s| < e; Hermite polynomials H_2n), truncated
// after n = 4 (relative truncation < 1e-9 for |u| < 6).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace sfh {
namespace psf {

constexpr float kSeriesEps = 0.1f;
constexpr double kGLNode = 0.7745966692414834;   // :303
constexpr float kW0 = 0.5555555555555556f;      // :304
constexpr float kW1 = 0.8888888888888888f;

// One prepared kernel: the clipped footprint and its FP32 parameters.
// Footprint indices are 0-based and inclusive; xa > xb marks "nothing to
// deposit" (the reference's length>1 guard, :359 / :401).
struct alignas(16) PointDesc {
    int32_t xa, xb, ya, yb;
    float ax;     // u of the first footprint column centre
    float du;     // u step per column
    float ex;     // half column width in u
    float g;      // covariant: shear c*sy/(sqrt2 sx); asymmetric: unused
    float ay;     // covariant: t of first footprint row centre; asymmetric: v of first row
    float dv;     // t (or v) step per row
    float ey;     // covariant: GL node offset xi*hy/sy; asymmetric: half row width in v
    float pref;   // amplitude prefactor (see header)
};

struct Range {
    double first, step;
    int64_t length;
};

// Julia round(Int, x) (ties to even) with the argument clamped so the cast
// is defined; callers flag non-finite input separately.
__host__ __device__ inline int64_t jl_round(double x)
{
    if (x > 4.0e15) x = 4.0e15;
    if (x < -4.0e15) x = -4.0e15;
    return (int64_t)rint(x);
}

__host__ __device__ inline int64_t jl_ceil(double x)
{
    if (x > 4.0e15) x = 4.0e15;
    if (x < -4.0e15) x = -4.0e15;
    return (int64_t)ceil(x);
}

__host__ __device__ inline int64_t imax(int64_t a, int64_t b) { return a > b ? a : b; }
__host__ __device__ inline int64_t imin(int64_t a, int64_t b) { return a < b ? a : b; }

__host__ __device__ inline PointDesc empty_desc()
{
    PointDesc d{};
    d.xa = 1;
    d.xb = 0;
    d.ya = 1;
    d.yb = 0;
    return d;
}

// Covariant kernel centred at (x0, y0) (real units) -- addstar! :366-408.
// Returns the number of (point, bin) pairs in the footprint (0 if skipped).
// The FP64 expressions keep the reference's operation order; compile with
// fp-contract off so no FMA changes a rounding.
__host__ __device__ inline int64_t prep_covariant(double x0, double y0, double sx, double sy,
                                                  double c, double A, Range ex, Range ey,
                                                  PointDesc* out)
{
#pragma clang fp contract(off)
    const int64_t nx = ex.length - 1, ny = ey.length - 1;
    const int64_t xp = jl_round((x0 - ex.first) / ex.step + 1.0);   // histogram_pix :502, :378-379
    const int64_t yp = jl_round((y0 - ey.first) / ey.step + 1.0);
    const int64_t xo = imax(1, jl_round(15.0 * sx / ex.step / 2.0)); // size(obj) :293, :381
    const int64_t yo = imax(1, jl_round(10.0 * sy / ey.step / 2.0)); // :382
    const int64_t xa = imax(1, xp - xo), xb = imin(nx, xp + xo);    // :384
    const int64_t ya = imax(1, yp - yo), yb = imin(ny, yp + yo);    // :385
    const int64_t lx = xb - xa + 1, ly = yb - ya + 1;
    if (lx <= 1 || ly <= 1) {                                        // :401
        *out = empty_desc();
        return 0;
    }
    // xriter / yriter: range(histogram_data(first+1/2), histogram_data(last+1/2); length) :390-395
    const double xs = ((double)xa + 0.5 - 1.0) * ex.step + ex.first;
    const double xe = ((double)xb + 0.5 - 1.0) * ex.step + ex.first;
    const double ys = ((double)ya + 0.5 - 1.0) * ey.step + ey.first;
    const double ye = ((double)yb + 0.5 - 1.0) * ey.step + ey.first;
    const double xst = (xe - xs) / (double)(lx - 1), yst = (ye - ys) / (double)(ly - 1);
    const double hx = xst / 2.0, hy = yst / 2.0;
    const double s2x = 1.4142135623730951 * sx;
    PointDesc d;
    d.xa = (int32_t)(xa - 1);
    d.xb = (int32_t)(xb - 1);
    d.ya = (int32_t)(ya - 1);
    d.yb = (int32_t)(yb - 1);
    d.ax = (float)((xs - x0) / s2x);
    d.du = (float)(xst / s2x);
    d.ex = (float)(hx / s2x);
    d.g = (float)(c * sy / s2x);
    d.ay = (float)((ys - y0) / sy);
    d.dv = (float)(yst / sy);
    d.ey = (float)(kGLNode * hy / sy);
    double pref = A * hy / (2.0 * 2.5066282746310002 * sy);          // prefac*h_y, :318,:323
    if (d.ex < kSeriesEps) pref *= 4.0 * hx / s2x / 1.7724538509055160;   // series form factor
    d.pref = (float)pref;
    *out = d;
    return lx * ly;
}

// Asymmetric pixel-space kernel -- bin_cmd_smooth :598-609, addstar! :348-364.
// (x0, y0) in real units, converted to fractional pixels like the reference.
__host__ __device__ inline int64_t prep_asymmetric(double col, double mag, double cerr,
                                                   double merr, double A, Range ex, Range ey,
                                                   PointDesc* out)
{
#pragma clang fp contract(off)
    const int64_t nx = ex.length - 1, ny = ey.length - 1;
    const double x0 = (col - ex.first) / ex.step + 1.0;             // :601
    const double y0 = (mag - ey.first) / ey.step + 1.0;             // :602
    const double sx = cerr / ex.step, sy = merr / ey.step;          // :603-604
    const int64_t cx = jl_ceil(sx * 10.0), cy = jl_ceil(sy * 10.0); // size(obj) :247
    const int64_t x = jl_round(x0), y = jl_round(y0);               // :350
    const int64_t xo = imax(1, cx / 2), yo = imax(1, cy / 2);       // :351-352
    const int64_t xa = imax(1, x - xo), xb = imin(nx, x + xo);      // :356
    const int64_t ya = imax(1, y - yo), yb = imin(ny, y + yo);      // :357
    const int64_t lx = xb - xa + 1, ly = yb - ya + 1;
    if (lx <= 1 || ly <= 1) {                                        // :359
        *out = empty_desc();
        return 0;
    }
    const double s2x = 1.4142135623730951 * sx, s2y = 1.4142135623730951 * sy;
    PointDesc d;
    d.xa = (int32_t)(xa - 1);
    d.xb = (int32_t)(xb - 1);
    d.ya = (int32_t)(ya - 1);
    d.yb = (int32_t)(yb - 1);
    d.ax = (float)(((double)xa + 0.5 - x0) / s2x);                  // evaluate at i + 1/2 :361
    d.du = (float)(1.0 / s2x);
    d.ex = (float)(0.5 / s2x);
    d.g = 0.0f;
    d.ay = (float)(((double)ya + 0.5 - y0) / s2y);
    d.dv = (float)(1.0 / s2y);
    d.ey = (float)(0.5 / s2y);
    double pref = A / 4.0;
    if (d.ex < kSeriesEps) pref *= 4.0 * (0.5 / s2x) / 1.7724538509055160;
    if (d.ey < kSeriesEps) pref *= 4.0 * (0.5 / s2y) / 1.7724538509055160;
    d.pref = (float)pref;
    *out = d;
    return lx * ly;
}

// Coefficients of S(w) = s0 + s1 w + s2 w^2 + s3 w^3 + s4 w^4 (w = u^2) for
// the series D = (4e/sqrt(pi)) exp(-w) S(w).
struct Series {
    float s0, s1, s2, s3, s4;
};

__host__ __device__ inline Series series_coeffs(float e)
{
    const float q = e * e;
    const float c1 = q / 6.0f, c2 = q * q / 120.0f, c3 = q * q * q / 5040.0f,
                c4 = q * q * q * q / 362880.0f;
    Series s;
    s.s0 = 1.0f - 2.0f * c1 + 12.0f * c2 - 120.0f * c3 + 1680.0f * c4;
    s.s1 = 4.0f * c1 - 48.0f * c2 + 720.0f * c3 - 13440.0f * c4;
    s.s2 = 16.0f * c2 - 480.0f * c3 + 13440.0f * c4;
    s.s3 = 64.0f * c3 - 3584.0f * c4;
    s.s4 = 256.0f * c4;
    return s;
}

__host__ __device__ inline float fast_exp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return std::exp(x);
#endif
}

// exp(-u^2) * S(u^2)  (the series form of D without its 4e/sqrt(pi) factor)
__host__ __device__ inline float d_series(float u, const Series& s)
{
    float w = u * u;
    w = w < 100.0f ? w : 100.0f;
    const float poly = fmaf(fmaf(fmaf(fmaf(s.s4, w, s.s3), w, s.s2), w, s.s1), w, s.s0);
    return fast_exp(-w) * poly;
}

__host__ __device__ inline float d_direct(float u, float e)
{
    return erff(u + e) - erff(u - e);
}

}  // namespace psf
}  // namespace sfh

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
//   D = (4 e / sqrt(pi)) exp(-u^2) S(u^2),  S(w) = sum_n H_2n(sqrt w) e^2n / (2n+1)!
// (the mean of exp(-(u+s)^2) over |s| < e; Hermite polynomials H_2n), whose
// coefficients in powers of w are kept up to w^3.  Relative to the local
// value at |u| <= 4 the dropped terms are below ~1e-7 when e < 0.1 (the
// terms of order e^2n |H_2n(4)| / (2n+1)! are ~10 e^2, 28 e^4, 30 e^6, 15 e^8).
// For e < 0.026 the factor is kept in the exponent instead (the Gaussian
// form): s0 + s1 w = s0 exp(b w) (1 + O(b^2 w^2)), b = s1 / s0 = 2e^2/3 +
// O(e^4), so D = (4 e / sqrt(pi)) s0 exp(-(1 - b) w), one Gaussian of a
// rescaled coordinate.  ln S(w) = ln s0 + b w - (4/45) e^4 w^2 + ..., so the
// dropped term is (4/45) e^4 w^2 relative to the local value: <= 1.1e-5 at
// |u| = 4 for e < 0.026 (below 8e-8 for e < 0.0077, and always below the
// degree-1 polynomial's own truncation error (2/15) e^4 w^2); relative to
// the kernel's peak it is at most (4/45) e^4 max(w^2 e^-w) = 0.048 e^4,
// below 2.2e-8, under the FP32 rounding of the sums.  On the 700-template
// workload 97% of the pairs take this form.
//
// The series forms run in scaled coordinates a' = a sqrt(log2 e) (the
// Gaussian form: a sqrt((1 - b) log2 e)), so exp(-a^2) = exp2(-a'^2) is one
// hardware exp2, with the coefficients of S rescaled to w' = a'^2.  The direct form shares erf values between adjacent
// columns: the right edge of column c is the left edge of column c + 1.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

namespace sfh {
namespace psf {

constexpr float kSeriesEps = 0.1f;
constexpr float kDeg1Eps = 0.026f;    // the Gaussian form below this half width
constexpr float kDeg2Eps = 0.026f;    // (no degree-2 series: its range is the Gaussian form's)
constexpr double kGLNode = 0.7745966692414834;   // :303
constexpr float kW0 = 0.5555555555555556f;      // :304
constexpr float kW1 = 0.8888888888888888f;
constexpr double kLog2e = 1.4426950408889634;
constexpr double kSqrtLog2e = 1.2011224087864498;
constexpr int kCPT = 16;                        // consecutive columns evaluated per thread

// x evaluation of a point: the erf difference, the Gaussian form, or the
// series of degree 2..3
enum : int32_t { kDirect = 0, kGauss = 1, kSeries2 = 2, kSeries3 = 3 };

// One prepared kernel: the clipped footprint and its FP32 parameters.
// Footprint indices are 0-based and inclusive; xa > xb marks "nothing to
// deposit" (the reference's length>1 guard, :359 / :401).
// Series kinds store ax, du and g in the scaled coordinate a' (see header).
struct alignas(16) PointDesc {
    int32_t xa, xb, ya, yb;
    float ax;     // x coordinate of the first footprint column centre
    float du;     // x step per column
    float ex;     // half column width in u (direct form)
    float g;      // covariant: shear c*sy/(sqrt2 sx); asymmetric: 0
    float ay;     // covariant: t of first footprint row centre; asymmetric: v of first row
    float dv;     // t (or v) step per row
    float ey;     // covariant: GL node offset xi*hy/sy; asymmetric: half row width in v
    float pref;   // amplitude prefactor (see header), with the series form factors
    float s0, s1, s2, s3;   // x series coefficients in w' (series kinds; kGauss: s0 = 1)
    int32_t kind;           // kDirect, kGauss or kSeriesN for the x factor
    float gf, gfi;          // kGauss, covariant: 2^(2 g ey du) and its inverse (the
                            // outer GL nodes' strip multipliers relative to the middle node's)
    int32_t pad0;
    float q[8];             // kGauss: q[j] = 2^-(du (j + 1/2))^2, the strip's column factors
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

// The x-factor form of half width e (in u): its kind, and for the series
// kinds the coefficients rescaled to w' = w log2(e) in d.
__host__ __device__ inline void set_x_form(PointDesc* d, float e)
{
    if (!(e < kSeriesEps)) {
        d->kind = kDirect;
        return;
    }
    d->kind = e < kDeg1Eps ? kGauss : e < kDeg2Eps ? kSeries2 : kSeries3;
    const Series s = series_coeffs(e);
    const double l1 = 1.0 / kLog2e, l2 = l1 * l1, l3 = l2 * l1;
    double scale = kSqrtLog2e;
    if (d->kind == kGauss) {
        // s0 exp(-(1 - b) w), b = s1 / s0: the coordinates carry sqrt(1 - b),
        // the amplitude s0
        const double b = (double)s.s1 / (double)s.s0;
        scale *= sqrt(1.0 - b);
        d->pref = (float)((double)d->pref * (double)s.s0);
        d->s0 = 1.0f;
        d->s1 = d->s2 = d->s3 = 0.0f;
    } else {
        d->s0 = s.s0;
        d->s1 = (float)(s.s1 * l1);
        d->s2 = (float)(s.s2 * l2);
        d->s3 = (float)(s.s3 * l3);
    }
    // coordinates of the series forms in a' = a sqrt(log2 e) (times sqrt(1 - b))
    d->ax = (float)((double)d->ax * scale);
    d->du = (float)((double)d->du * scale);
    d->g = (float)((double)d->g * scale);
    if (d->kind == kGauss) {
        // the column factors of the strip recurrence (gauss_strip) and the
        // outer nodes' multiplier ratios, from the rounded values the kernel uses
        const double du = (double)d->du;
        for (int j = 0; j < 8; ++j) {
            const double sj = du * ((double)j + 0.5);
            d->q[j] = (float)exp2(-sj * sj);
        }
        const double r = 2.0 * (double)d->g * (double)d->ey * du;
        d->gf = (float)exp2(r);
        d->gfi = (float)exp2(-r);
    }
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
    PointDesc d{};
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
    set_x_form(&d, d.ex);
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
    PointDesc d{};
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
    set_x_form(&d, d.ex);
    *out = d;
    return lx * ly;
}

__host__ __device__ inline float fast_exp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __expf(x);
#else
    return std::exp(x);
#endif
}

__host__ __device__ inline float fast_exp2(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_exp2f(x);
#else
    return std::exp2(x);
#endif
}

// exp(-u^2) * S(u^2)  (the full-degree series form of D without its 4e/sqrt(pi)
// factor; the asymmetric kernel's y factor)
__host__ __device__ inline float d_series(float u, const Series& s)
{
    float w = u * u;
    w = w < 100.0f ? w : 100.0f;
    const float poly = fmaf(fmaf(fmaf(fmaf(s.s4, w, s.s3), w, s.s2), w, s.s1), w, s.s0);
    return fast_exp(-w) * poly;
}

__host__ __device__ inline float fast_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

// erfc(x) for x >= 0, branch-free: the Chebyshev-fitted form of Numerical
// Recipes (Press et al., 2nd ed., section 6.2, erfcc), fractional error below
// 1.2e-7 for every x >= 0 in exact arithmetic.
__host__ __device__ inline float erfc_pos(float x)
{
    const float t = fast_rcp(fmaf(0.5f, x, 1.0f));
    float p = 0.17087277f;
    p = fmaf(p, t, -0.82215223f);
    p = fmaf(p, t, 1.48851587f);
    p = fmaf(p, t, -1.13520398f);
    p = fmaf(p, t, 0.27886807f);
    p = fmaf(p, t, -0.18628806f);
    p = fmaf(p, t, 0.09678418f);
    p = fmaf(p, t, 0.37409196f);
    p = fmaf(p, t, 1.00002368f);
    p = fmaf(p, t, -1.26551223f);
    return t * fast_exp(fmaf(-x, x, p));
}

// erf(b1) - erf(b0) of a column from e = erfc(|b|) and the sign of its two
// edges: both >= 0: e0 - e1; both < 0: e1 - e0; across 0: 2 - e1 - e0
__host__ __device__ inline float erf_diff(float e0, bool n0, float e1, bool n1)
{
    return n1 ? e1 - e0 : (n0 ? 2.0f - e1 - e0 : e0 - e1);
}

// erf(u + e) - erf(u - e) through erfc of non-negative arguments (D is even
// in u), so that it keeps its relative accuracy in the tails
__host__ __device__ inline float d_direct(float u, float e)
{
    const float a = fabsf(u);
    const float hi = erfc_pos(a + e);
    return a >= e ? erfc_pos(a - e) - hi : 2.0f - hi - erfc_pos(e - a);
}

// The asymmetric kernel's y factor of a row at coordinate v (half height ey).
__host__ __device__ inline float y_factor(float v, float ey)
{
    return ey < kSeriesEps ? d_series(v, series_coeffs(ey)) : d_direct(v, ey);
}

// Row weights of the covariant kernel: W_k = pref w_k exp(-t_k^2/2) at the GL
// nodes t_k = tj + (k-1) ey, and the x shifts h_k = g t_k.
__host__ __device__ inline void cov_row(const PointDesc& d, int row, float* W, float* h)
{
    const float tj = fmaf((float)(row - d.ya), d.dv, d.ay);
    const float t0 = tj - d.ey, t2 = tj + d.ey;
    const float hl = (float)(-0.5 * kLog2e);
    W[0] = d.pref * kW0 * fast_exp2(hl * t0 * t0);
    W[1] = d.pref * kW1 * fast_exp2(hl * tj * tj);
    W[2] = d.pref * kW0 * fast_exp2(hl * t2 * t2);
    h[0] = d.g * t0;
    h[1] = d.g * tj;
    h[2] = d.g * t2;
}

// The covariant row in the Gaussian form's terms: M_k = pref w_k, the log2 of
// the node's y weight L_k = -(log2 e / 2) t_k^2 (it joins the exponent of the
// strip recurrence, so no exp2 of its own) and the x shifts h_k = g t_k.
__host__ __device__ inline void cov_row_log(const PointDesc& d, int row, float* M, float* L, float* h)
{
    const float tj = fmaf((float)(row - d.ya), d.dv, d.ay);
    const float t0 = tj - d.ey, t2 = tj + d.ey;
    const float hl = (float)(-0.5 * kLog2e);
    M[0] = d.pref * kW0;
    M[1] = d.pref * kW1;
    M[2] = M[0];
    L[0] = hl * t0 * t0;
    L[1] = hl * tj * tj;
    L[2] = hl * t2 * t2;
    h[0] = d.g * t0;
    h[1] = d.g * tj;
    h[2] = d.g * t2;
}

// c du + a for a compile-time column offset c.  On the device the offset is a
// literal of v_fmamk_f32 (D = S0 K + S1, fused): written as a plain fmaf the
// compiler keeps every column's offset in a register of its own (VOP3 takes no
// literal), which costs the kernel sixteen registers.
template <int C>
__host__ __device__ __forceinline__ float col_fma(float du, float a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm("v_fmamk_f32 %0, %1, %2, %3" : "=v"(r) : "v"(du), "i"(__builtin_bit_cast(int, (float)C)), "v"(a));
    return r;
#else
    return fmaf((float)C, du, a);
#endif
}

// part = in ? v : part, as a select: v is computed for every lane first (the
// register barrier keeps the compiler from sinking its computation into a
// branch per column, which serialises the columns' dependency chains)
__host__ __device__ __forceinline__ void masked_set(float& part, float v, bool in)
{
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__ volatile("" : "+v"(v));
#endif
    part = in ? v : part;
}

// f(integral_constant<int, C>) for C = I .. N-1, in order
template <int I, int N, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// The Gaussian form over one row's strip: part[c] += sum_k M[k] 2^(L[k]) E_k(c),
// E_k(c) = 2^-(a_k(c))^2, by a product recurrence from the strip's middle.
// With m the x coordinate midway between columns 7 and 8 and s = c - 7.5,
//   (m + s du)^2 = m^2 + 2 s m du + s^2 du^2,
// so E(8 + j) = P B^j q[j] and E(7 - j) = P B^-(j+1) q[j], P = 2^-(m^2 + m du),
// B = 2^(-2 m du), q[j] = 2^-(du (j + 1/2))^2 (prepared per point).  For the
// three GL nodes (m_k - m_1 = -+g ey) the outer nodes' B are the middle node's
// times 2^(+-2 g ey du) (prepared per point), so a row takes 5 exp2 (3 for the
// asymmetric kernel) instead of one per node and column (an exp2 takes two
// VALU issue slots, an FMA or multiply one), and a column costs one FMA and
// one multiply per node.  The multiplier's rounding grows over at most 8
// steps.  The Gaussian form has du < 0.064 (half bin width e < 0.026), so B
// stays within [2^-2, 2^2] for |m| < 15.  Each column adds the nodes in order
// k = 0, 1, 2.
template <int NODES, bool MASK>
__host__ __device__ __forceinline__ void gauss_strip(const PointDesc& d, const float* M, const float* L,
                                                     const float* h, int col0, float* part)
{
    static_assert(kCPT == 16, "the recurrence runs 8 columns each way from the strip's middle");
    static_assert(NODES == 1 || NODES == 3, "one node (asymmetric) or three GL nodes (covariant)");
    const float base = fmaf((float)(col0 - d.xa), d.du, d.ax);
    const int lo = d.xa - col0, hi = d.xb - col0;
    const float du = d.du;
    constexpr int mid = NODES / 2;
    float pf[NODES], B[NODES], Bi[NODES];
#pragma unroll
    for (int k = 0; k < NODES; ++k) {
        const float m = fmaf(7.5f, du, base + h[k]);
        pf[k] = M[k] * fast_exp2(L[k] - m * (m + du));
        if (k == mid) {
            const float md2 = 2.0f * m * du;
            B[k] = fast_exp2(-md2);
            Bi[k] = fast_exp2(md2);
        }
    }
    if constexpr (NODES == 3) {
        B[0] = B[1] * d.gf;
        Bi[0] = Bi[1] * d.gfi;
        B[2] = B[1] * d.gfi;
        Bi[2] = Bi[1] * d.gf;
    }
#pragma unroll
    for (int k = 0; k < NODES; ++k) {
        float f = pf[k], b = pf[k] * Bi[k];
        const float Bk = B[k], Bik = Bi[k];
        static_for<0, 8>([&](auto jj) {
            constexpr int j = decltype(jj)::value;
            constexpr int cf = 8 + j, cb = 7 - j;
            const float vf = fmaf(f, d.q[j], part[cf]);
            const float vb = fmaf(b, d.q[j], part[cb]);
            if (MASK) {
                masked_set(part[cf], vf, cf >= lo && cf <= hi);
                masked_set(part[cb], vb, cb >= lo && cb <= hi);
            } else {
                part[cf] = vf;
                part[cb] = vb;
            }
            if constexpr (j < 7) {
                f *= Bk;
                b *= Bik;
            }
        });
    }
}

// Adds the contributions of one point to the kCPT columns col0 .. col0+kCPT-1
// of one row: part[c] += sum_k W[k] X(a_k(c)), a_k(c) the x coordinate of
// column col0 + c shifted by h[k] (NODES = 3: covariant GL nodes; NODES = 1:
// asymmetric, W[0] = pref * y factor).  DEG: the x form (kDirect or a series
// degree 2, 3).  MASK: only columns inside [xa, xb] receive anything (without
// it every column does: the footprint spans the whole grid and columns past
// the grid are never written).  Each column adds the nodes in order.
template <int NODES, int DEG, bool MASK>
__host__ __device__ __forceinline__ void add_strip(const PointDesc& d, const float* W, const float* h, int col0,
                                                   float* part)
{
    static_assert(DEG == kDirect || DEG == 2 || DEG == 3, "the Gaussian form is gauss_strip");
    const float base = fmaf((float)(col0 - d.xa), d.du, d.ax);
    const int lo = d.xa - col0, hi = d.xb - col0;
    const float du = d.du;
    if constexpr (DEG == kDirect) {
        // erf at the column edges b(c) = a(c) - ex, c = 0..kCPT, as erfc(|b|) so
        // that D(c) = erf(b(c+1)) - erf(b(c)) keeps its relative accuracy in
        // the tails: both edges >= 0: erfc(b0) - erfc(b1); both < 0:
        // erfc(-b1) - erfc(-b0); across 0: 2 - erfc(b1) - erfc(-b0).  Node by
        // node (each column still adds the nodes in order), so that one
        // node's edge values are live at a time.
#pragma unroll
        for (int k = 0; k < NODES; ++k) {
            const float b0 = base + h[k] - d.ex;
            bool pneg = b0 < 0.0f;
            float prev = erfc_pos(fabsf(b0));
            static_for<0, kCPT>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                const float b1 = col_fma<c + 1>(du, b0);
                const bool neg = b1 < 0.0f;
                const float e1 = erfc_pos(fabsf(b1));
                const float dd = erf_diff(prev, pneg, e1, neg);
                const float v = fmaf(W[k], dd, part[c]);
                if (MASK) masked_set(part[c], v, c >= lo && c <= hi);
                else part[c] = v;
                prev = e1;
                pneg = neg;
#if defined(__HIP_DEVICE_COMPILE__)
                // one column at a time: interleaving the columns' erfc sequences
                // would raise the register count of the whole kernel
                __builtin_amdgcn_sched_barrier(0);
#endif
            });
        }
    } else {
        float a0[NODES], c0[NODES], c1[NODES], c2[NODES], c3[NODES];
#pragma unroll
        for (int k = 0; k < NODES; ++k) {
            a0[k] = base + h[k];
            c0[k] = W[k] * d.s0;
            c1[k] = W[k] * d.s1;
            c2[k] = W[k] * d.s2;
            c3[k] = DEG >= 3 ? W[k] * d.s3 : 0.0f;
        }
        static_for<0, kCPT>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            float v = part[c];
#pragma unroll
            for (int k = 0; k < NODES; ++k) {
                const float a = col_fma<c>(du, a0[k]);
                const float w = a * a;
                float S;
                if constexpr (DEG == 2) S = fmaf(fmaf(c2[k], w, c1[k]), w, c0[k]);
                else S = fmaf(fmaf(fmaf(c3[k], w, c2[k]), w, c1[k]), w, c0[k]);
                v = fmaf(S, fast_exp2(-w), v);
            }
            if (MASK) masked_set(part[c], v, c >= lo && c <= hi);
            else part[c] = v;
#if defined(__HIP_DEVICE_COMPILE__)
            // columns in groups of four: enough independent work per wave, without
            // every column's temporaries live at once
            if constexpr ((c & 3) == 3) __builtin_amdgcn_sched_barrier(0);
#endif
        });
    }
}

// Left edge of footprint column l of a narrow kernel in the erf-difference
// form, for GL node shift h: x0 + l du + h, x0 = ax - ex (the wave-cooperative
// path: one column per lane, the right edge being the next column's left).
__host__ __device__ inline float direct_edge(const PointDesc& d, int l, float h)
{
    return fmaf((float)l, d.du, d.ax - d.ex) + h;
}

// One point's contributions to one row's strip, dispatched on its form.
template <int NODES, bool MASK>
__host__ __device__ __forceinline__ void add_point(const PointDesc& d, const float* W, const float* h, int col0,
                                                   float* part)
{
    switch (d.kind) {
    case kGauss: {
        const float L[NODES] = {};
        gauss_strip<NODES, MASK>(d, W, L, h, col0, part);
        break;
    }
    case kSeries2: add_strip<NODES, 2, MASK>(d, W, h, col0, part); break;
    case kSeries3: add_strip<NODES, 3, MASK>(d, W, h, col0, part); break;
    default: add_strip<NODES, kDirect, MASK>(d, W, h, col0, part); break;
    }
}

}  // namespace psf
}  // namespace sfh

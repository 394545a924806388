// sfh_multifit.hip -- many independent fits of uneven size (each its own
// templates x bins), the fg! of every fit evaluated in one persistent launch,
// and the quasi-Newton iterations of fit_templates / fit_templates_fast
// (src/fitting/solvers.jl:172-275) run on the device for all fits at once.
//
// Grouped fg! (k_multi_fg).  The group keeps a blocked copy of every fit's
// templates, [32-bin block][template][32 bins], so one block of one fit is a
// single contiguous T x 128-byte run.  A work item is a run of blocks of one
// fit.  For a block, each of the 64*NW threads loads one float4 (4 bins) of
// ceil(T/(8 NW)) templates -- 8 threads cover a template's 128 bytes -- and
// keeps them in registers through the whole block (the next block's loads
// in flight meanwhile):
//   composite  m_i  = sum_t C[i,t] x_t      FP64 FMAs, reduced over the 8 NW column slices
//   LL, R      on 32 threads: n - c - n log(n/c), 1 - n/c' with c = max(m, eps),
//              c' = max(m, eps_g)
//   gradient   g_t += sum_i C[i,t] R_i      FP64 FMAs into per-thread registers
// so the templates are read from HBM once per evaluation.  Each item writes
// one LL partial and T gradient partials; the per-fit reduction sums them in
// item order (deterministic).
//
// Optimiser step = three launches over all fits, identical in shape every
// step so a run of steps is captured once in a HIP graph and replayed:
//   k_opt_propose   per fit: search direction when a new iteration starts
//                   (BFGS d = -H g, or the L-BFGS two-loop recursion), the
//                   trial point theta + alpha d, and the coefficients x(theta)
//   k_multi_fg      fg! at every active fit's trial coefficients
//   k_opt_accept    per fit: transformed objective and gradient, the line
//                   search as a state machine (Hager-Zhang acceptance,
//                   bracketing, safeguarded cubic steps), on acceptance the BFGS
//                   inverse-Hessian update (or the L-BFGS pair), convergence
//                   and the MAP -> MLE stage change
// and a one-workgroup count of fits still running ends each graph.
#include "sfh_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace sfh {
namespace mf {

constexpr int kWG = 256;
constexpr int kSB = 32;                  // bins per block: one 128-byte segment of a column
constexpr int kQ = kSB / 4;              // float4 quads per block: thread (slice s, quad q) = 8 s + q
constexpr int kMaxT = 1024;
constexpr int kSingleNW = 8;              // waves per workgroup of the single-fit fg!
constexpr int kSingleMaxK = 11;           // its widest variant (T <= 704) that does not spill

// Pointers read from a FitDesc in memory are generic (flat) to the compiler;
// flat loads also count against the LDS counter, so every LDS access would
// wait for every load in flight.  Loads go through global-space pointers.
typedef float f4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) float gfloat;
typedef const __attribute__((address_space(1))) f4 gfloat4;
typedef const __attribute__((address_space(1))) double gdouble;
constexpr int kMaxMem = 32;

struct FitDesc {
    const float* C;      // ld x T templates, column-major (the caller's)
    const float* Cb;     // templates as fg! reads them: the group's blocked copy
                         // [32-bin block][T][32 bins] (a single fit: C itself)
    const double* n;     // data
    int64_t ld, N, T;
    int64_t xoff;        // offset of the fit's T-vectors
    int64_t hoff;        // offset of its T x T matrix
    int64_t item0, nitems;
    int64_t poff;        // offset of its gradient partials (nitems x T)
};

struct Item {
    int64_t bin0;
    int32_t fit;
    int32_t nblk;        // 32-bin blocks
};

// ------------------------------------------------------------ grouped fg!

// One block of one work item on registers c (already loaded): composite,
// LL and R on 32 threads, gradient accumulation into g.
// OUT: also write the composite m (or R when rout_R) of every bin to rout.
template <int K, int NW, bool OUT = false>
__device__ __forceinline__ void fg_block(const FitDesc& fd, int64_t b0, double eps, double eps_g, int want_G, const double* sx,
                                         double (*sm)[kSB], double* sr, f4 (&c)[K], double nv, double (&g)[K],
                                         double& ll, double* rout = nullptr, int rout_R = 0)
{
    constexpr int SL = 8 * NW;
    const int tid = threadIdx.x, q = tid & (kQ - 1), s = tid >> 3, lane = tid & 63, wave = tid >> 6;
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double xv = sx[s + SL * k];
        m0 = fma((double)c[k].x, xv, m0);
        m1 = fma((double)c[k].y, xv, m1);
        m2 = fma((double)c[k].z, xv, m2);
        m3 = fma((double)c[k].w, xv, m3);
    }
    // sum over the 8 slices of this wave (lane bits 3..5), then over waves in order
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
        m0 += __shfl_xor(m0, off, 64);
        m1 += __shfl_xor(m1, off, 64);
        m2 += __shfl_xor(m2, off, 64);
        m3 += __shfl_xor(m3, off, 64);
    }
    if (lane < kQ) {
        sm[wave][4 * q + 0] = m0;
        sm[wave][4 * q + 1] = m1;
        sm[wave][4 * q + 2] = m2;
        sm[wave][4 * q + 3] = m3;
    }
    __syncthreads();
    if (tid < kSB) {
        double m = sm[0][tid];
#pragma unroll
        for (int w = 1; w < NW; ++w) m += sm[w][tid];
        const int64_t bin = b0 + tid;
        double r = 0.0;
        if (bin < fd.N) {
            const double cl = m > eps ? m : eps;
            const double n = nv;
            const double q = n / cl;
            ll += (n > 0.0) ? (n - cl - n * log(q)) : -cl;
            // the gradient's clamp is eps(eltype(composite)) (fitting_base.jl:277),
            // the LL's eps of the promoted type (:86); they differ only where m
            // is below one of them
            const double cg = m > eps_g ? m : eps_g;
            double qg = q;
            if (cg != cl) qg = n / cg;
            r = 1.0 - qg;
            if constexpr (OUT) {
                if (rout) rout[bin] = rout_R ? r : m;
            }
        }
        sr[tid] = r;
    }
    __syncthreads();
    if (want_G) {
        const double r0 = sr[4 * q], r1 = sr[4 * q + 1], r2 = sr[4 * q + 2], r3 = sr[4 * q + 3];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            // convert again rather than keep the FP64 copies of the block
            // live across the barrier (twice the registers)
            __asm__ volatile("" : "+v"(c[k]));
            g[k] = fma((double)c[k].x, r0, g[k]);
            g[k] = fma((double)c[k].y, r1, g[k]);
            g[k] = fma((double)c[k].z, r2, g[k]);
            g[k] = fma((double)c[k].w, r3, g[k]);
        }
    }
}

// One work item with K columns per thread (K >= ceil(T / (8 NW))).  Every
// load is unconditional -- columns past the last are clamped to it and meet
// x = 0 in the composite; their gradient sums are not written; block indices
// past the item's last are clamped to it -- and two register sets alternate,
// so the loads of the next block are in flight while a block is processed and
// each wait covers only the older set.
// COLMAJOR: the templates are the caller's ld x T column-major matrix (one
// fit read in place: template stride ld, block stride 32) instead of the
// blocked copy (template stride 32, block stride 32 T), and the per-bin
// composite output of fg_block is on.
template <int K, int NW, bool COLMAJOR = false>
__device__ __forceinline__ void fg_item(const FitDesc& fd, const Item& item, int64_t it, double eps, double eps_g, int want_G,
                                        const double* sx, double (*sm)[kSB], double* sr, double* __restrict__ ll_part,
                                        double* __restrict__ g_part, double* rout = nullptr, int rout_R = 0)
{
    constexpr int SL = 8 * NW;    // column slices
    const int tid = threadIdx.x, q = tid & (kQ - 1), s = tid >> 3, wave = tid >> 6;
    const int T = (int)fd.T;
    // element offsets of this thread's columns inside a block (column-major:
    // the host keeps (T - 1) ld + 32 below 2^31)
    int32_t col[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int t = s + SL * k;
        if constexpr (COLMAJOR)
            col[k] = (t < T ? t : T - 1) * (int32_t)fd.ld + 4 * q;
        else
            col[k] = (t < T ? t : T - 1) * kSB + 4 * q;
    }
    const int64_t bstride = COLMAJOR ? (int64_t)kSB : (int64_t)T * kSB;
    gfloat* blk0 = (gfloat*)fd.Cb + (item.bin0 / kSB) * bstride;
    gdouble* nd = (gdouble*)fd.n;
    const int last = item.nblk - 1;
    double g[K];
#pragma unroll
    for (int k = 0; k < K; ++k) g[k] = 0.0;
    double ll = 0.0;    // threads 0..31: their bin of every block

    // a register set: the block's template values and its data counts (every
    // thread loads one count, index clamped into the fit, so the whole set
    // retires together and no later load has to be waited for on its own)
    f4 ca[K], cb[K];
    double na, nb;
    const int64_t nlast = fd.N - 1;
#define SFH_LOAD(R, NV, b)                                                                   \
    do {                                                                                     \
        const int bb_ = (b) < last ? (b) : last;                                             \
        gfloat* p_ = blk0 + (int64_t)bb_ * bstride;                                          \
        _Pragma("unroll") for (int k = 0; k < K; ++k) R[k] = *(gfloat4*)(p_ + col[k]);       \
        const int64_t i_ = item.bin0 + (int64_t)bb_ * kSB + (tid & (kSB - 1));               \
        NV = nd[i_ < nlast ? i_ : nlast];                                                    \
    } while (0)
    SFH_LOAD(ca, na, 0);
    for (int b = 0; b <= last; b += 2) {
        SFH_LOAD(cb, nb, b + 1);
        fg_block<K, NW, COLMAJOR>(fd, item.bin0 + (int64_t)b * kSB, eps, eps_g, want_G, sx, sm, sr, ca, na, g, ll, rout,
                                  rout_R);
        SFH_LOAD(ca, na, b + 2);
        if (b + 1 <= last)
            fg_block<K, NW, COLMAJOR>(fd, item.bin0 + (int64_t)(b + 1) * kSB, eps, eps_g, want_G, sx, sm, sr, cb, nb, g,
                                      ll, rout, rout_R);
    }
#undef SFH_LOAD
    if (wave == 0) {
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) ll += __shfl_xor(ll, off, 64);
        if (tid == 0) ll_part[it] = ll;
    }
    if (want_G) {
        double* gp = g_part + fd.poff + (it - fd.item0) * (int64_t)T;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double v = g[k];
            v += __shfl_xor(v, 1, 64);
            v += __shfl_xor(v, 2, 64);
            v += __shfl_xor(v, 4, 64);
            const int t = s + SL * k;
            if (q == 0 && t < T) gp[t] = v;
        }
    }
}

// Runs the smallest column-count variant K (from 1, 2, 3, 4, 6, 8, 11, 16,
// 22, 32, capped at KMAX) with K >= kmax.
template <int NW, int KMAX, int I = 0>
__device__ __forceinline__ void fg_dispatch(int kmax, const FitDesc& fd, const Item& item, int64_t it, double eps,
                                            double eps_g, int want_G, const double* sx, double (*sm)[kSB], double* sr,
                                            double* __restrict__ ll_part, double* __restrict__ g_part)
{
    constexpr int kBuckets[] = {1, 2, 3, 4, 6, 8, 11, 16, 22, 32};
    constexpr int K = kBuckets[I];
    if constexpr (K >= KMAX) {
        fg_item<KMAX, NW>(fd, item, it, eps, eps_g, want_G, sx, sm, sr, ll_part, g_part);
    } else {
        if (kmax <= K) fg_item<K, NW>(fd, item, it, eps, eps_g, want_G, sx, sm, sr, ll_part, g_part);
        else fg_dispatch<NW, KMAX, I + 1>(kmax, fd, item, it, eps, eps_g, want_G, sx, sm, sr, ll_part, g_part);
    }
}

// Persistent grouped fg!: workgroup w takes items w, w + grid, ...  KMAX
// columns per thread covers the group's largest fit; each item runs the
// smallest column-count variant that covers its own fit.
// Occupancy floor per variant: without it the compiler widens every template
// value to FP64 ahead of the FMAs and doubles the register footprint.
constexpr int fg_waves_per_eu(int kmax) { return kmax <= 4 ? 4 : kmax <= 8 ? 3 : kmax <= 16 ? 2 : 1; }

template <int NW, int KMAX>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(fg_waves_per_eu(KMAX))))
void k_multi_fg(const FitDesc* __restrict__ fits, const Item* __restrict__ items, int64_t nitems,
                const double* __restrict__ X, const int32_t* __restrict__ active, double eps, double eps_g,
                int want_G, double* __restrict__ ll_part, double* __restrict__ g_part)
{
    constexpr int NT = 64 * NW, SL = 8 * NW;
    __shared__ double sx[KMAX * SL];
    __shared__ double sm[NW][kSB];
    __shared__ double sr[kSB];
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const Item item = items[it];
        if (active && !active[item.fit]) continue;
        const FitDesc fd = fits[item.fit];
        const int T = (int)fd.T;
        const int kmax = (T + SL - 1) / SL;
        __syncthreads();    // the previous item is done with sx
        for (int t = threadIdx.x; t < KMAX * SL; t += NT) sx[t] = t < T ? ((gdouble*)X)[fd.xoff + t] : 0.0;
        __syncthreads();
        fg_dispatch<NW, KMAX>(kmax, fd, item, it, eps, eps_g, want_G, sx, sm, sr, ll_part, g_part);
    }
}

// ------------------------------------------------------- block reductions

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the workgroup, identical in every thread (fixed order).
__device__ inline double block_sum(double v, double* red)
{
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

__device__ inline double block_max(double v, double* red)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Per-fit reduction of the item partials: negLL (+Inf for an exact zero sum,
// fitting_base.jl:95) and G.
__global__ __launch_bounds__(kWG) void k_multi_reduce(const FitDesc* __restrict__ fits, const double* __restrict__ ll_part,
                                                      const double* __restrict__ g_part, double* __restrict__ negLL,
                                                      double* __restrict__ G)
{
    const FitDesc fd = fits[blockIdx.x];
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int64_t i = 0; i < fd.nitems; ++i) s += ll_part[fd.item0 + i];
        negLL[blockIdx.x] = (s != 0.0) ? -s : INFINITY;
    }
    if (G) {
        for (int64_t t = threadIdx.x; t < fd.T; t += kWG) {
            double s = 0.0;
            for (int64_t i = 0; i < fd.nitems; ++i) s += g_part[fd.poff + i * fd.T + t];
            G[fd.xoff + t] = s;
        }
    }
}

// ------------------------------------------------------------- optimiser

enum Phase : int32_t { PH_INIT = 0, PH_DIR = 1, PH_TRIAL = 2, PH_DONE = 3 };

// per-fit vectors, each sumT long, stacked in one buffer
enum Vec { V_THETA = 0, V_G, V_D, V_X, V_GLO, V_MUMAP, V_TMP, V_COLSUM, NVEC };

struct OptState {
    int32_t phase, stage;
    int32_t ls_evals, bracketed;
    int32_t nmem, head;
    int32_t status[2];
    int64_t iter, evals;
    int64_t iters[2], evalsv[2];
    double f, dphi0, alpha;
    double a_lo, f_lo, d_lo, a_hi, f_hi, d_hi;
    double fmin[2];
};

struct OptParams {
    int transform, method, memory;
    double g_abstol;
    int64_t max_iterations;
    const FitDesc* fits;
    OptState* st;
    double* V;            // NVEC x sumT
    int64_t sumT;
    double* H;            // BFGS inverse Hessians
    double* Hmap;         // MAP-stage copy (may be null)
    double* S;            // L-BFGS pairs: memory x sumT each
    double* Y;
    double* rho;          // memory x nfits
    int64_t nfits;
    const double* ll_part;
    const double* g_part;
    double* X;            // coefficients handed to k_multi_fg
    int32_t* active;
};

__device__ inline double* vec(const OptParams& p, int v, const FitDesc& fd) { return p.V + v * p.sumT + fd.xoff; }

// Next trial inside a bracket: safeguarded cubic interpolation between its ends.
__device__ inline double zoom_alpha(double a_lo, double f_lo, double d_lo, double a_hi, double f_hi, double d_hi)
{
    const double lo = fmin(a_lo, a_hi), hi = fmax(a_lo, a_hi), w = hi - lo;
    double a = 0.5 * (lo + hi);
    if (isfinite(f_hi) && isfinite(d_hi) && a_lo != a_hi) {
        const double d1 = d_lo + d_hi - 3.0 * (f_lo - f_hi) / (a_lo - a_hi);
        const double disc = d1 * d1 - d_lo * d_hi;
        if (disc >= 0.0) {
            const double d2 = copysign(sqrt(disc), a_hi - a_lo);
            const double den = d_hi - d_lo + 2.0 * d2;
            if (den != 0.0) {
                const double c = a_hi - (a_hi - a_lo) * (d_hi + d2 - d1) / den;
                if (isfinite(c)) a = c;
            }
        }
    }
    if (!(a >= lo + 0.1 * w && a <= hi - 0.1 * w)) a = 0.5 * (lo + hi);
    return a;
}

// x(theta): exp (LOG) or theta^2 (SQRT)
__device__ inline double xform(int transform, double th) { return transform == SFH_FIT_LOG ? exp(th) : th * th; }

// H <- I over the fit's T x T block
__device__ inline void h_identity(double* H, int T)
{
    const int64_t TT = (int64_t)T * T;
    for (int64_t i = threadIdx.x; i < TT; i += kWG) H[i] = (i / T == i % T) ? 1.0 : 0.0;
}

__global__ __launch_bounds__(kWG) void k_opt_init(OptParams p, const double* __restrict__ x0, const double* __restrict__ nsum)
{
    __shared__ double red[4];
    const int f = blockIdx.x;
    const FitDesc fd = p.fits[f];
    const int T = (int)fd.T;
    double* th = vec(p, V_THETA, fd);
    const double* cs = vec(p, V_COLSUM, fd);
    // renormalize_x0: x0 * sum(data) / sum(composite(x0)) (utilities.jl:104-115)
    double part = 0.0;
    for (int t = threadIdx.x; t < T; t += kWG) part += cs[t] * (x0 ? x0[fd.xoff + t] : 1.0);
    const double msum = block_sum(part, red);
    const double scale = msum != 0.0 ? nsum[f] / msum : 1.0;
    for (int t = threadIdx.x; t < T; t += kWG) {
        const double x = (x0 ? x0[fd.xoff + t] : 1.0) * scale;
        th[t] = p.transform == SFH_FIT_LOG ? log(x) : sqrt(x);
    }
    if (p.method == SFH_OPT_BFGS) h_identity(p.H + fd.hoff, T);
    if (threadIdx.x == 0) {
        OptState s{};
        s.phase = PH_INIT;
        s.stage = p.transform == SFH_FIT_LOG ? 0 : 1;
        p.st[f] = s;
    }
}

template <int METHOD>
__global__ __launch_bounds__(kWG) void k_opt_propose(OptParams p)
{
    __shared__ double red[4];
    __shared__ double sg[kMaxT];
    __shared__ double sq[kMaxT];
    __shared__ double salpha[kMaxMem];
    const int f = blockIdx.x;
    const FitDesc fd = p.fits[f];
    const int T = (int)fd.T;
    OptState st = p.st[f];
    if (st.phase == PH_DONE) {
        if (threadIdx.x == 0) p.active[f] = 0;
        return;
    }
    double* th = vec(p, V_THETA, fd);
    double* d = vec(p, V_D, fd);
    double* x = p.X + fd.xoff;

    if (st.phase == PH_DIR) {
        const double* g = vec(p, V_G, fd);
        for (int t = threadIdx.x; t < T; t += kWG) sg[t] = g[t];
        __syncthreads();
        if (METHOD == SFH_OPT_BFGS) {
            // d = -H g (H symmetric: column reads are coalesced)
            const double* H = p.H + fd.hoff;
            for (int i = threadIdx.x; i < T; i += kWG) {
                double acc = 0.0;
                for (int j = 0; j < T; ++j) acc = fma(H[(int64_t)j * T + i], sg[j], acc);
                d[i] = -acc;
            }
        } else {
            // two-loop recursion over the stored pairs, newest first
            for (int t = threadIdx.x; t < T; t += kWG) sq[t] = sg[t];
            const int m = p.memory;
            for (int k = 0; k < st.nmem; ++k) {
                const int slot = (st.head - 1 - k + m) % m;
                const double* S = p.S + (int64_t)slot * p.sumT + fd.xoff;
                const double* Y = p.Y + (int64_t)slot * p.sumT + fd.xoff;
                double part = 0.0;
                for (int t = threadIdx.x; t < T; t += kWG) part += S[t] * sq[t];
                const double a = p.rho[(int64_t)slot * p.nfits + f] * block_sum(part, red);
                if (threadIdx.x == 0) salpha[k] = a;
                for (int t = threadIdx.x; t < T; t += kWG) sq[t] -= a * Y[t];
            }
            double gamma = 1.0;
            if (st.nmem > 0) {
                const int slot = (st.head - 1 + m) % m;
                const double* S = p.S + (int64_t)slot * p.sumT + fd.xoff;
                const double* Y = p.Y + (int64_t)slot * p.sumT + fd.xoff;
                double sy = 0.0, yy = 0.0;
                for (int t = threadIdx.x; t < T; t += kWG) {
                    sy += S[t] * Y[t];
                    yy += Y[t] * Y[t];
                }
                sy = block_sum(sy, red);
                yy = block_sum(yy, red);
                if (yy > 0.0) gamma = sy / yy;
            }
            for (int t = threadIdx.x; t < T; t += kWG) sq[t] *= gamma;
            __syncthreads();
            for (int k = st.nmem - 1; k >= 0; --k) {
                const int slot = (st.head - 1 - k + m) % m;
                const double* S = p.S + (int64_t)slot * p.sumT + fd.xoff;
                const double* Y = p.Y + (int64_t)slot * p.sumT + fd.xoff;
                double part = 0.0;
                for (int t = threadIdx.x; t < T; t += kWG) part += Y[t] * sq[t];
                const double b = p.rho[(int64_t)slot * p.nfits + f] * block_sum(part, red);
                const double a = salpha[k];
                for (int t = threadIdx.x; t < T; t += kWG) sq[t] += S[t] * (a - b);
            }
            for (int t = threadIdx.x; t < T; t += kWG) d[t] = -sq[t];
        }
        __syncthreads();
        double gd = 0.0;
        for (int t = threadIdx.x; t < T; t += kWG) gd += sg[t] * d[t];
        double dphi0 = block_sum(gd, red);
        if (!(dphi0 < 0.0)) {
            // not a descent direction: restart from steepest descent with fresh curvature
            if (METHOD == SFH_OPT_BFGS) h_identity(p.H + fd.hoff, T);
            st.nmem = 0;
            st.head = 0;
            double gg = 0.0;
            for (int t = threadIdx.x; t < T; t += kWG) {
                d[t] = -sg[t];
                gg += sg[t] * sg[t];
            }
            dphi0 = -block_sum(gg, red);
        }
        double dd = 0.0;
        for (int t = threadIdx.x; t < T; t += kWG) dd += d[t] * d[t];
        const double ns = sqrt(block_sum(dd, red));
        // LineSearches.InitialStatic(1.0, scaled = true): a first step of length min(1, |d|)
        st.alpha = ns > 0.0 ? fmin(1.0, ns) / ns : 1.0;
        st.dphi0 = dphi0;
        st.a_lo = 0.0;
        st.f_lo = st.f;
        st.d_lo = dphi0;
        st.a_hi = 0.0;
        st.f_hi = 0.0;
        st.d_hi = 0.0;
        st.bracketed = 0;
        st.ls_evals = 0;
        st.phase = PH_TRIAL;
        double* glo = vec(p, V_GLO, fd);
        for (int t = threadIdx.x; t < T; t += kWG) glo[t] = sg[t];
    }
    __syncthreads();
    // trial coefficients
    const double a = st.phase == PH_INIT ? 0.0 : st.alpha;
    for (int t = threadIdx.x; t < T; t += kWG) {
        const double tt = st.phase == PH_INIT ? th[t] : fma(a, d[t], th[t]);
        x[t] = xform(p.transform, tt);
    }
    if (threadIdx.x == 0) {
        p.active[f] = 1;
        // only the fields this step changes: a whole-struct store would carry
        // the state's arrays through a private-memory copy
        OptState& o = p.st[f];
        o.phase = st.phase;
        o.nmem = st.nmem;
        o.head = st.head;
        o.ls_evals = st.ls_evals;
        o.bracketed = st.bracketed;
        o.alpha = st.alpha;
        o.dphi0 = st.dphi0;
        o.a_lo = st.a_lo;
        o.f_lo = st.f_lo;
        o.d_lo = st.d_lo;
        o.a_hi = st.a_hi;
        o.f_hi = st.f_hi;
        o.d_hi = st.d_hi;
    }
}

template <int METHOD>
__global__ __launch_bounds__(kWG) void k_opt_accept(OptParams p)
{
    __shared__ double red[4];
    __shared__ double ss[kMaxT];
    __shared__ double shy[kMaxT];
    __shared__ double sll;
    const int f = blockIdx.x;
    const FitDesc fd = p.fits[f];
    const int T = (int)fd.T;
    OptState st = p.st[f];
    if (st.phase == PH_DONE) return;
    double* th = vec(p, V_THETA, fd);
    double* g = vec(p, V_G, fd);
    double* d = vec(p, V_D, fd);
    double* glo = vec(p, V_GLO, fd);
    double* gt = vec(p, V_TMP, fd);
    const double* x = p.X + fd.xoff;
    const bool init = st.phase == PH_INIT;
    const double a = init ? 0.0 : st.alpha;

    // objective and gradient in the transformed variables at the trial point
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int64_t i = 0; i < fd.nitems; ++i) s += p.ll_part[fd.item0 + i];
        sll = (s != 0.0) ? -s : INFINITY;
    }
    const bool map = p.transform == SFH_FIT_LOG && st.stage == 0;
    double thsum = 0.0, gd = 0.0;
    for (int t = threadIdx.x; t < T; t += kWG) {
        double G = 0.0;
        for (int64_t i = 0; i < fd.nitems; ++i) G += p.g_part[fd.poff + i * T + t];
        const double tt = init ? th[t] : fma(a, d[t], th[t]);
        double v;
        if (p.transform == SFH_FIT_LOG) v = map ? G * x[t] - 1.0 : G * x[t];   // solvers.jl:185-187, 193-195
        else v = G * 2.0 * tt;                                                // solvers.jl:260-262
        gt[t] = v;
        thsum += tt;
        if (!init) gd += v * d[t];
    }
    thsum = block_sum(thsum, red);
    const double dphi = init ? 0.0 : block_sum(gd, red);
    const double F = sll - (map ? thsum : 0.0);
    st.evals += 1;

    // 0: keep searching, 1: accept the trial, 2: accept the bracket's low end, 3: give up
    int decision = 0;
    if (init) {
        decision = 1;
    } else {
        // Acceptance as in LineSearches.HagerZhang, the line search of
        // fit_templates (Hager & Zhang 2005/2006; delta = 0.1, sigma = 0.9,
        // epsilon = 1e-6): the Wolfe conditions
        //   F <= f0 + delta a phi'(0)  and  phi'(a) >= sigma phi'(0),
        // or the approximate Wolfe conditions
        //   F <= f0 + epsilon |f0|  and  sigma phi'(0) <= phi'(a) <= (2 delta - 1) phi'(0),
        // which rely on the derivative and stay usable where F sits at its
        // rounding floor near the optimum.  The bracket [a_lo, a_hi] keeps
        // phi'(a_lo) < 0 with F(a_lo) <= f0 + epsilon |f0| at its low end.
        const double delta = 0.1, sigma = 0.9, epsk = 1e-6 * fabs(st.f);
        const bool fin = isfinite(F) && isfinite(dphi);
        const bool wolfe = fin && F <= st.f + delta * a * st.dphi0 && dphi >= sigma * st.dphi0;
        const bool approx_wolfe = fin && F <= st.f + epsk && dphi >= sigma * st.dphi0 &&
                                  dphi <= (2.0 * delta - 1.0) * st.dphi0;
        bool lo_moved = false;
        if (wolfe || approx_wolfe) {
            decision = 1;
        } else if (!fin || dphi >= 0.0 || F > st.f + epsk) {
            st.a_hi = a;
            st.f_hi = fin ? F : INFINITY;
            st.d_hi = fin ? dphi : NAN;
            st.bracketed = 1;
        } else {
            st.a_lo = a;
            st.f_lo = F;
            st.d_lo = dphi;
            lo_moved = true;
        }
        if (decision == 0) {
            st.ls_evals += 1;
            const double w = fabs(st.a_hi - st.a_lo);
            const bool exhausted = st.ls_evals >= 30 ||
                                   (st.bracketed && w <= 1e-14 * fmax(1.0, fmax(st.a_hi, st.a_lo)));
            if (exhausted) {
                decision = st.a_lo > 0.0 ? 2 : 3;
            } else {
                st.alpha = st.bracketed ? zoom_alpha(st.a_lo, st.f_lo, st.d_lo, st.a_hi, st.f_hi, st.d_hi)
                                        : fmin(4.0 * a, 1e10);
            }
            if (lo_moved)
                for (int t = threadIdx.x; t < T; t += kWG) glo[t] = gt[t];
        }
    }

    int status = SFH_FIT_RUNNING;
    if (decision == 1 || decision == 2) {
        const double an = decision == 1 ? a : st.a_lo;
        const double Fn = decision == 1 ? F : st.f_lo;
        const double* gn = decision == 1 ? gt : glo;
        // s = an d, y = gn - g; theta += s
        double sy = 0.0, gmax = 0.0;
        for (int t = threadIdx.x; t < T; t += kWG) {
            const double s = init ? 0.0 : an * d[t];
            const double y = gn[t] - g[t];
            ss[t] = s;
            shy[t] = y;
            sy += s * y;
            if (!init) th[t] = fma(an, d[t], th[t]);
            gmax = fmax(gmax, fabs(gn[t]));
        }
        sy = block_sum(sy, red);
        gmax = block_max(gmax, red);
        for (int t = threadIdx.x; t < T; t += kWG) g[t] = gn[t];
        st.f = Fn;
        if (!init) {
            st.iter += 1;
            if (sy > 0.0) {
                if (METHOD == SFH_OPT_BFGS) {
                    // H <- H + (sy + y'Hy)(s s')/sy^2 - (Hy s' + s y'H)/sy   (Hy in shy after this)
                    double* H = p.H + fd.hoff;
                    __syncthreads();
                    double yhy = 0.0;
                    for (int i = threadIdx.x; i < T; i += kWG) {
                        double acc = 0.0;
                        for (int j = 0; j < T; ++j) acc = fma(H[(int64_t)j * T + i], shy[j], acc);
                        yhy += shy[i] * acc;
                        gt[i] = acc;    // scratch: Hy
                    }
                    yhy = block_sum(yhy, red);
                    for (int i = threadIdx.x; i < T; i += kWG) shy[i] = gt[i];
                    __syncthreads();
                    const double c1 = (sy + yhy) / (sy * sy), c2 = 1.0 / sy;
                    const int64_t TT = (int64_t)T * T;
                    for (int64_t e = threadIdx.x; e < TT; e += kWG) {
                        const int i = (int)(e / T), j = (int)(e - (int64_t)i * T);
                        H[e] += c1 * ss[i] * ss[j] - c2 * (shy[i] * ss[j] + ss[i] * shy[j]);
                    }
                } else {
                    const int m = p.memory;
                    const int slot = st.head;
                    double* S = p.S + (int64_t)slot * p.sumT + fd.xoff;
                    double* Y = p.Y + (int64_t)slot * p.sumT + fd.xoff;
                    for (int t = threadIdx.x; t < T; t += kWG) {
                        S[t] = ss[t];
                        Y[t] = shy[t];
                    }
                    if (threadIdx.x == 0) p.rho[(int64_t)slot * p.nfits + f] = 1.0 / sy;
                    st.head = (st.head + 1) % m;
                    st.nmem = st.nmem < m ? st.nmem + 1 : m;
                }
            }
        }
        if (gmax <= p.g_abstol) status = SFH_FIT_CONVERGED;
        else if (st.iter >= p.max_iterations) status = SFH_FIT_MAX_ITERATIONS;
        st.phase = PH_DIR;
    } else if (decision == 3) {
        status = SFH_FIT_LINESEARCH_FAILED;
    }

    if (status != SFH_FIT_RUNNING) {
        const int sg = st.stage;
        if (sg == 0) {
            st.status[0] = status;
            st.iters[0] = st.iter;
            st.evalsv[0] = st.evals;
            st.fmin[0] = st.f;
        } else {
            st.status[1] = status;
            st.iters[1] = st.iter;
            st.evalsv[1] = st.evals;
            st.fmin[1] = st.f;
        }
        __syncthreads();
        if (p.transform == SFH_FIT_LOG && sg == 0) {
            // MAP done: keep exp(theta) and the inverse Hessian, then start the MLE
            // from the MAP optimum (solvers.jl:210); its objective and gradient at
            // the same point follow without a new evaluation: f + sum(theta), g + 1.
            double* mu = vec(p, V_MUMAP, fd);
            double thsum2 = 0.0, gmax = 0.0;
            for (int t = threadIdx.x; t < T; t += kWG) {
                mu[t] = exp(th[t]);
                thsum2 += th[t];
                g[t] += 1.0;
                gmax = fmax(gmax, fabs(g[t]));
            }
            thsum2 = block_sum(thsum2, red);
            gmax = block_max(gmax, red);
            if (METHOD == SFH_OPT_BFGS) {
                double* H = p.H + fd.hoff;
                const int64_t TT = (int64_t)T * T;
                if (p.Hmap)
                    for (int64_t e = threadIdx.x; e < TT; e += kWG) p.Hmap[fd.hoff + e] = H[e];
                __syncthreads();
                h_identity(H, T);
            }
            st.f += thsum2;
            st.stage = 1;
            st.iter = 0;
            st.evals = 0;
            st.nmem = 0;
            st.head = 0;
            st.phase = PH_DIR;
            if (gmax <= p.g_abstol) {
                st.status[1] = SFH_FIT_CONVERGED;
                st.iters[1] = 0;
                st.evalsv[1] = 0;
                st.fmin[1] = st.f;
                st.phase = PH_DONE;
            }
        } else {
            st.phase = PH_DONE;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) p.st[f] = st;
}

__global__ __launch_bounds__(kWG) void k_opt_count(const OptState* __restrict__ st, int64_t nfits, int64_t* __restrict__ out)
{
    __shared__ double red[4];
    double c = 0.0;
    for (int64_t f = threadIdx.x; f < nfits; f += kWG) c += st[f].phase != PH_DONE ? 1.0 : 0.0;
    c = block_sum(c, red);
    if (threadIdx.x == 0) *out = (int64_t)c;
}

// Blocked copy of one fit: Cb[(b*T + t)*32 + i] = C[t*ld + 32 b + i] (bins
// past N are the source's zero padding, 32 b + 31 < ld).
__global__ __launch_bounds__(kWG) void k_block_copy(const float* __restrict__ C, int64_t ld, int64_t T, int64_t nblk,
                                                    float* __restrict__ Cb)
{
    const int64_t n = nblk * T * kQ;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t q = i % kQ, t = (i / kQ) % T, b = i / (kQ * T);
        reinterpret_cast<float4*>(Cb)[i] = *reinterpret_cast<const float4*>(C + t * ld + b * kSB + 4 * q);
    }
}

// Per-template column sums (FP64) for renormalize_x0, and per-fit data sums.
__global__ __launch_bounds__(kWG) void k_colsum(const FitDesc* __restrict__ fits, const int32_t* __restrict__ col_fit,
                                                double* __restrict__ colsum)
{
    __shared__ double red[4];
    const int64_t gcol = blockIdx.x;
    const FitDesc fd = fits[col_fit[gcol]];
    const int64_t t = gcol - fd.xoff;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < fd.N; i += kWG) s += (double)fd.C[t * fd.ld + i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) colsum[gcol] = s;
}

__global__ __launch_bounds__(kWG) void k_datasum(const FitDesc* __restrict__ fits, double* __restrict__ out)
{
    __shared__ double red[4];
    const FitDesc fd = fits[blockIdx.x];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < fd.N; i += kWG) s += fd.n[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ------------------------------------------------------------------- host

// columns-per-thread bound of the kernel variant for a group whose largest
// fit has maxT templates, with NW waves per workgroup (0: no variant)
int kmax_for(int64_t maxT, int nw)
{
    const int need = (int)cdiv(maxT, 8 * nw);
    // The variants that hold their two register sets and gradient sums
    // without spilling: 16-wave workgroups (128 registers per lane) stop at 4
    // columns per thread, 8-wave ones at 11 (T <= 704); 16 columns spill at
    // two waves per SIMD, so fits of 353-1024 templates take the 4-wave,
    // 32-column variant (one wave per SIMD, 512 registers).
    const int cap = nw == 4 ? 32 : nw == 8 ? 11 : 4;
    for (int v : {2, 4, 8, 11, 32})
        if (need <= v && v <= cap) return v;
    return 0;
}

template <int NW, int KMAX>
void launch_fg_t(hipStream_t s, int grid, const FitDesc* fits, const Item* items, int64_t nitems, const double* X,
                 const int32_t* active, double eps, double eps_g, int want_G, double* llp, double* gp)
{
    hipLaunchKernelGGL((k_multi_fg<NW, KMAX>), dim3(grid), dim3(64 * NW), 0, s, fits, items, nitems, X, active, eps,
                       eps_g, want_G, llp, gp);
}

// ------------------------------------------- one fit, templates read in place

// fg! of one template set in one pass over its templates: the grouped
// kernel's work items over the caller's column-major templates (template
// stride ld, block stride 32 -- each thread's float4 is still one 16-byte
// piece of a 128-byte run of one template), item `it` = blocks
// [it nblk / nitems, (it + 1) nblk / nitems) (one item per resident
// workgroup, lengths differing by at most one block), K = the column count T
// needs, the coefficients staged once per workgroup.
template <int NW, int K>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(fg_waves_per_eu(K))))
void k_single_fg(FitDesc fd, int64_t nitems, const double* __restrict__ X, double eps, double eps_g, int want_G,
                 double* __restrict__ ll_part, double* __restrict__ g_part, double* rout, int rout_R)
{
    constexpr int NT = 64 * NW, SL = 8 * NW;
    __shared__ double sx[K * SL];
    __shared__ double sm[NW][kSB];
    __shared__ double sr[kSB];
    for (int t = threadIdx.x; t < K * SL; t += NT) sx[t] = t < fd.T ? ((gdouble*)X)[t] : 0.0;
    __syncthreads();
    const int64_t nblk = (fd.N + kSB - 1) / kSB;
    for (int64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int64_t b0 = it * nblk / nitems, b1 = (it + 1) * nblk / nitems;
        const Item item{b0 * kSB, 0, (int32_t)(b1 - b0)};
        fg_item<K, NW, true>(fd, item, it, eps, eps_g, want_G, sx, sm, sr, ll_part, g_part, rout, rout_R);
    }
}

// Sums of the single-fit partials in a fixed order.  Workgroup w: templates
// 16 w .. 16 w + 15, thread = 16 item slices x 16 templates (each slice sums
// items slice, slice + 16, ... in order; slices then add in order).
// Workgroup 0 also sums the LL partials: LL = -Inf for an exact zero sum
// (fitting_base.jl:95), negLL = -LL.
__global__ __launch_bounds__(kWG) void k_single_reduce(const double* __restrict__ ll_part,
                                                       const double* __restrict__ g_part, int64_t nitems, int64_t T,
                                                       int want_ll, double* __restrict__ negLL,
                                                       double* __restrict__ ll_out, double gsign,
                                                       double* __restrict__ G)
{
    __shared__ double sh[kWG];
    __shared__ double red[4];
    const int tt = threadIdx.x & 15, sl = threadIdx.x >> 4;
    if (G) {
        const int64_t t = (int64_t)blockIdx.x * 16 + tt;
        double v = 0.0;
        if (t < T) {
#pragma unroll 8
            for (int64_t i = sl; i < nitems; i += 16) v += g_part[i * T + t];
        }
        sh[threadIdx.x] = v;
        __syncthreads();
        if (sl == 0 && t < T) {
            double g = sh[tt];
#pragma unroll
            for (int k = 1; k < 16; ++k) g += sh[16 * k + tt];
            G[t] = gsign * g;
        }
    }
    if (want_ll && blockIdx.x == 0) {
        double v = 0.0;
        for (int64_t i = threadIdx.x; i < nitems; i += kWG) v += ll_part[i];
        const double s = block_sum(v, red);
        if (threadIdx.x == 0) {
            const double ll = (s != 0.0) ? s : -INFINITY;
            if (negLL) *negLL = -ll;
            if (ll_out) *ll_out = ll;
        }
    }
}

template <int K>
void launch_single_t(hipStream_t s, int cus, const FitDesc& fd, int64_t nblk, const double* X, double eps,
                     double eps_g, int want_G, double* llp, double* gp, double* rout, int rout_R, int64_t* nitems_out)
{
    // workgroups resident per CU: two of 8 waves where K leaves <= 128 VGPRs
    const int64_t grid = (int64_t)cus * (K <= 4 ? 2 : 1);
    const int64_t nitems = nblk < grid ? nblk : grid;
    *nitems_out = nitems;
    hipLaunchKernelGGL((k_single_fg<kSingleNW, K>), dim3((unsigned)nitems), dim3(64 * kSingleNW), 0, s, fd, nitems,
                       X, eps, eps_g, want_G, llp, gp, rout, rout_R);
}

}  // namespace mf

bool single_pass_covers(const sfh_templates* tpl)
{
    const int64_t ld = cdiv(tpl->nbins, 256) * 256;
    return tpl->ntemplates >= 1 && tpl->ntemplates <= 8 * mf::kSingleNW * mf::kSingleMaxK && tpl->nbins >= 1 &&
           (tpl->ntemplates - 1) * ld + mf::kSB < ((int64_t)1 << 31);
}

// Compute units of each device (cached at the first call on it).
static int g_cus_of[64] = {};

static int device_cus(sfh_ctx* ctx, int* cus)
{
    if (ctx->device < 0 || ctx->device >= 64) return set_error(ctx, SFH_ERR_ARGUMENT, "device index out of range");
    int c = __atomic_load_n(&g_cus_of[ctx->device], __ATOMIC_RELAXED);
    if (!c) {
        SFH_HIP_TRY(ctx, hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, ctx->device));
        __atomic_store_n(&g_cus_of[ctx->device], c, __ATOMIC_RELAXED);
    }
    *cus = c;
    return SFH_OK;
}

// Partials of one single-pass fg!: one LL value per work item, and ntemplates
// gradient values per item when the gradient is wanted; at most 2 items per CU.
static int64_t single_pass_partials(const sfh_templates* tpl, int cus, bool want_G, int64_t* cap)
{
    const int64_t nblk = cdiv(tpl->nbins, mf::kSB);
    *cap = std::min<int64_t>(nblk, 2 * (int64_t)cus);
    return *cap * (1 + (want_G ? tpl->ntemplates : 0));
}

// The launches of a single-pass fg! into caller-provided partial buffers
// (llp: cap doubles, gp: cap x ntemplates doubles or NULL without the
// gradient).  `timed` records the launches in the context's kernel timing.
static int single_pass_launch(sfh_ctx* ctx, const sfh_templates* tpl, int cus, const double* d_data,
                              const double* d_coeffs, double eps, double eps_g, int want_ll, double* d_negLL,
                              double* d_ll, double* d_G, double gsign, double* d_comp_out, int clobber_with_R,
                              double* llp, double* gp, bool timed)
{
    using namespace mf;
    const int64_t T = tpl->ntemplates, N = tpl->nbins;
    FitDesc fd{};
    fd.C = fd.Cb = tpl->d_models;
    fd.n = d_data;
    fd.ld = cdiv(N, 256) * 256;
    fd.N = N;
    fd.T = T;
    // one item (a run of blocks, the next block's loads in flight) per
    // resident workgroup
    const int64_t nblk = cdiv(N, kSB);
    const int want_G = d_G ? 1 : 0;
    const int need = (int)cdiv(T, 8 * kSingleNW);
    int64_t nitems = 0;
    {
        hipEvent_t t0{};
        if (timed) timing_begin(ctx, "single_fg", &t0);
        hipStream_t s = ctx->stream;
        double* l = llp;
        double* g = gp;
        if (need <= 1) launch_single_t<1>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        else if (need <= 2) launch_single_t<2>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        else if (need <= 3) launch_single_t<3>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        else if (need <= 4) launch_single_t<4>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        else if (need <= 6) launch_single_t<6>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        else if (need <= 8) launch_single_t<8>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        else launch_single_t<kSingleMaxK>(s, cus, fd, nblk, d_coeffs, eps, eps_g, want_G, l, g, d_comp_out, clobber_with_R, &nitems);
        if (timed) timing_end(ctx, "single_fg", t0);
    }
    if (int rc = check_launch(ctx, "k_single_fg")) return rc;
    if (want_ll || d_G) {
        hipEvent_t t0{};
        if (timed) timing_begin(ctx, "single_reduce", &t0);
        hipLaunchKernelGGL(k_single_reduce, dim3((unsigned)(d_G ? cdiv(T, 16) : 1)), dim3(kWG), 0, ctx->stream,
                           (const double*)llp, (const double*)gp, nitems, T, want_ll, d_negLL, d_ll, gsign, d_G);
        if (timed) timing_end(ctx, "single_reduce", t0);
    }
    return check_launch(ctx, "k_single_reduce");
}

int single_pass_fg(sfh_ctx* ctx, const sfh_templates* tpl, const double* d_data, const double* d_coeffs, double eps,
                   double eps_g, int want_ll, double* d_negLL, double* d_ll, double* d_G, double gsign,
                   double* d_comp_out, int clobber_with_R)
{
    int cus = 0;
    if (int rc = device_cus(ctx, &cus)) return rc;
    int64_t cap = 0;
    single_pass_partials(tpl, cus, d_G != nullptr, &cap);
    void *llp = nullptr, *gp = nullptr;
    if (int rc = scratch(ctx, "fg1_llpart", (size_t)cap * sizeof(double), &llp)) return rc;
    if (d_G)
        if (int rc = scratch(ctx, "fg1_gpart", (size_t)cap * tpl->ntemplates * sizeof(double), &gp)) return rc;
    return single_pass_launch(ctx, tpl, cus, d_data, d_coeffs, eps, eps_g, want_ll, d_negLL, d_ll, d_G, gsign,
                              d_comp_out, clobber_with_R, (double*)llp, (double*)gp, true);
}

// ------------------------------------------ fg! that allocates nothing

// The partial sums of sfh_fg_device_noalloc live in this fixed device array
// (SFH_NOALLOC_SCRATCH_BYTES; every GPU the library runs on holds its own copy
// until the process ends).  Host code never uses the array's name as an
// address: noalloc_init resolves it per device with hipGetSymbolAddress and
// the entry point passes that pointer to the kernels.  Nothing in it needs
// clearing: every partial the reduction reads is written earlier in the same
// call.  Calls that overlap in time on one device share it and race.
constexpr int64_t kNoAllocDoubles = SFH_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double noalloc_scratch[kNoAllocDoubles];
static double* g_noalloc_ptr[64] = {};
static std::mutex g_noalloc_mu;

int noalloc_init(int device)
{
    if (device < 0 || device >= 64) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return SFH_OK;
    // the caller (sfh_ctx_create) has made `device` current
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)) != hipSuccess || !p) return SFH_ERR_DEVICE;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return SFH_ERR_DEVICE;
    __atomic_store_n(&g_cus_of[device], cus, __ATOMIC_RELAXED);
    g_noalloc_ptr[device] = (double*)p;
    return SFH_OK;
}

}  // namespace sfh

using namespace sfh;
using namespace sfh::mf;

struct sfh_multifit {
    sfh_ctx* ctx = nullptr;
    int64_t nfits = 0, sumT = 0, sumT2 = 0, maxT = 0;
    int cus = 0;
    int nw = 8;                 // waves per grouped-fg! workgroup
    int64_t item_elems = 1 << 18;
    int wg_per_cu = 2;
    std::vector<FitDesc> h_fits;
    FitDesc* d_fits = nullptr;
    float* d_blocked = nullptr;         // blocked template copies of all fits
    Item* d_items = nullptr;
    int64_t nitems = 0, items_cap = 0;
    double* d_llp = nullptr;
    double* d_gp = nullptr;
    int64_t gp_cap = 0;
    double* d_colsum_tmp = nullptr;
    double* d_nsum = nullptr;
    // optimiser buffers
    OptState* d_st = nullptr;
    double* d_V = nullptr;
    double* d_X = nullptr;
    int32_t* d_active = nullptr;
    int64_t* d_count = nullptr;
    double* d_H = nullptr;
    double* d_Hmap = nullptr;
    double* d_S = nullptr;
    double* d_Y = nullptr;
    double* d_rho = nullptr;
    int mem_cap = 0;
};

namespace {

void free_all(sfh_multifit* mf)
{
    void* ptrs[] = {mf->d_fits, mf->d_blocked, mf->d_items, mf->d_llp, mf->d_gp, mf->d_colsum_tmp, mf->d_nsum, mf->d_st, mf->d_V,
                    mf->d_X, mf->d_active, mf->d_count, mf->d_H, mf->d_Hmap, mf->d_S, mf->d_Y, mf->d_rho};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// Work items for the current shape; partial buffers sized to match.
int build_items(sfh_multifit* mf)
{
    sfh_ctx* ctx = mf->ctx;
    std::vector<Item> items;
    int64_t poff = 0;
    for (int64_t f = 0; f < mf->nfits; ++f) {
        FitDesc& fd = mf->h_fits[f];
        const int64_t nblk_total = cdiv(fd.N, kSB);
        const int64_t per = std::max<int64_t>(1, std::min<int64_t>(nblk_total, mf->item_elems / (fd.T * kSB)));
        fd.item0 = (int64_t)items.size();
        for (int64_t b = 0; b < nblk_total; b += per)
            items.push_back(Item{b * kSB, (int32_t)f, (int32_t)std::min<int64_t>(per, nblk_total - b)});
        fd.nitems = (int64_t)items.size() - fd.item0;
        fd.poff = poff;
        poff += fd.nitems * fd.T;
    }
    mf->nitems = (int64_t)items.size();
    if (mf->nitems > mf->items_cap) {
        if (mf->d_items) (void)hipFree(mf->d_items);
        if (mf->d_llp) (void)hipFree(mf->d_llp);
        mf->d_items = nullptr;
        mf->d_llp = nullptr;
        SFH_HIP_TRY(ctx, hipMalloc(&mf->d_items, mf->nitems * sizeof(Item)));
        SFH_HIP_TRY(ctx, hipMalloc(&mf->d_llp, mf->nitems * sizeof(double)));
        mf->items_cap = mf->nitems;
    }
    if (poff > mf->gp_cap) {
        if (mf->d_gp) (void)hipFree(mf->d_gp);
        mf->d_gp = nullptr;
        SFH_HIP_TRY(ctx, hipMalloc(&mf->d_gp, poff * sizeof(double)));
        mf->gp_cap = poff;
    }
    SFH_HIP_TRY(ctx, hipMemcpyAsync(mf->d_items, items.data(), mf->nitems * sizeof(Item), hipMemcpyHostToDevice,
                                    ctx->stream));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(mf->d_fits, mf->h_fits.data(), mf->nfits * sizeof(FitDesc),
                                    hipMemcpyHostToDevice, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

void launch_fg(sfh_multifit* mf, const double* d_X, const int32_t* active, double eps, double eps_g, int want_G)
{
    const int grid = mf->cus * mf->wg_per_cu;
    hipStream_t s = mf->ctx->stream;
    const int km = kmax_for(mf->maxT, mf->nw);
#define SFH_L(NW, KM) launch_fg_t<NW, KM>(s, grid, mf->d_fits, mf->d_items, mf->nitems, d_X, active, eps, eps_g, want_G, mf->d_llp, mf->d_gp)
    if (mf->nw == 4) {
        if (km <= 2) SFH_L(4, 2); else if (km <= 4) SFH_L(4, 4); else if (km <= 8) SFH_L(4, 8);
        else if (km <= 11) SFH_L(4, 11); else SFH_L(4, 32);
    } else if (mf->nw == 8) {
        if (km <= 2) SFH_L(8, 2); else if (km <= 4) SFH_L(8, 4); else if (km <= 8) SFH_L(8, 8);
        else SFH_L(8, 11);
    } else {
        if (km <= 2) SFH_L(16, 2); else SFH_L(16, 4);
    }
#undef SFH_L
}

int fg_device_impl(sfh_multifit* mf, const double* d_X, double eps, double eps_g, double* d_negLL, double* d_G)
{
    sfh_ctx* ctx = mf->ctx;
    {
        TimedLaunch tl(ctx, "multifit_fg");
        launch_fg(mf, d_X, nullptr, eps, eps_g, d_G ? 1 : 0);
    }
    if (int rc = check_launch(ctx, "k_multi_fg")) return rc;
    {
        TimedLaunch tl(ctx, "multifit_reduce");
        hipLaunchKernelGGL(k_multi_reduce, dim3((unsigned)mf->nfits), dim3(kWG), 0, ctx->stream, mf->d_fits, mf->d_llp,
                           mf->d_gp, d_negLL, d_G);
    }
    return check_launch(ctx, "k_multi_reduce");
}

template <typename T>
int ensure(sfh_ctx* ctx, T** p, size_t count)
{
    if (*p) return SFH_OK;
    SFH_HIP_TRY(ctx, hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
    return SFH_OK;
}

}  // namespace

extern "C" {

int sfh_multifit_create(sfh_ctx* ctx, int64_t nfits, const sfh_templates* const* templates,
                        const sfh_data* const* data, sfh_multifit** out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    SFH_ARG_CHECK(ctx, out && templates && data && nfits > 0, "sfh_multifit_create: bad arguments");
    *out = nullptr;
    if (int rc = bind(ctx)) return rc;
    auto* mf = new sfh_multifit();
    mf->ctx = ctx;
    mf->nfits = nfits;
    mf->h_fits.resize(nfits);
    std::vector<int32_t> col_fit;
    std::vector<int64_t> boff;
    int64_t blocked_elems = 0;
    for (int64_t f = 0; f < nfits; ++f) {
        const sfh_templates* t = templates[f];
        const sfh_data* d = data[f];
        if (!t || !d || t->ctx != ctx || d->ctx != ctx || t->nbins != d->nbins || t->ntemplates < 1 ||
            t->ntemplates > kMaxT) {
            delete mf;
            return set_error(ctx, SFH_ERR_ARGUMENT,
                             "sfh_multifit_create: fit %lld needs templates and data of one context with "
                             "axes(models,1) == axes(data,1) and 1..%d templates",
                             (long long)f, kMaxT);
        }
        FitDesc& fd = mf->h_fits[f];
        fd.C = t->d_models;
        fd.n = d->d_data;
        fd.ld = cdiv(t->nbins, 256) * 256;
        fd.N = t->nbins;
        fd.T = t->ntemplates;
        fd.xoff = mf->sumT;
        fd.hoff = mf->sumT2;
        mf->sumT += fd.T;
        mf->sumT2 += fd.T * fd.T;
        mf->maxT = std::max(mf->maxT, fd.T);
        for (int64_t j = 0; j < fd.T; ++j) col_fit.push_back((int32_t)f);
        boff.push_back(blocked_elems);
        blocked_elems += cdiv(fd.N, kSB) * kSB * fd.T;
    }
    // the default 8-wave workgroups hold fits of up to 704 templates; larger
    // ones (up to 1024) take 4-wave workgroups
    if (kmax_for(mf->maxT, mf->nw) == 0) mf->nw = 4;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) {
        delete mf;
        return set_error(ctx, SFH_ERR_DEVICE, "hipGetDeviceProperties failed");
    }
    mf->cus = prop.multiProcessorCount;
    int rc = SFH_OK;
    int32_t* d_colfit = nullptr;
    do {
        if (hipMalloc(&mf->d_blocked, blocked_elems * sizeof(float)) != hipSuccess) {
            rc = set_error(ctx, SFH_ERR_ALLOC, "sfh_multifit_create: blocked template copy (%lld floats) failed",
                           (long long)blocked_elems);
            break;
        }
        for (int64_t f = 0; f < nfits; ++f) {
            FitDesc& fd = mf->h_fits[f];
            fd.Cb = mf->d_blocked + boff[f];
            const int64_t nblk = cdiv(fd.N, kSB);
            const int64_t n = nblk * fd.T * kQ;
            hipLaunchKernelGGL(k_block_copy, dim3((unsigned)std::min<int64_t>(cdiv(n, kWG), 65536)), dim3(kWG), 0,
                               ctx->stream, fd.C, fd.ld, fd.T, nblk, mf->d_blocked + boff[f]);
        }
        if ((rc = check_launch(ctx, "k_block_copy"))) break;
        if (hipMalloc(&mf->d_fits, nfits * sizeof(FitDesc)) != hipSuccess ||
            hipMalloc(&mf->d_V, (size_t)NVEC * mf->sumT * sizeof(double)) != hipSuccess ||
            hipMalloc(&mf->d_nsum, nfits * sizeof(double)) != hipSuccess ||
            hipMalloc(&d_colfit, mf->sumT * sizeof(int32_t)) != hipSuccess) {
            rc = set_error(ctx, SFH_ERR_ALLOC, "sfh_multifit_create: device allocation failed");
            break;
        }
        if ((rc = build_items(mf))) break;
        if (hipMemcpyAsync(d_colfit, col_fit.data(), mf->sumT * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess) {
            rc = set_error(ctx, SFH_ERR_DEVICE, "copy failed");
            break;
        }
        hipLaunchKernelGGL(k_colsum, dim3((unsigned)mf->sumT), dim3(kWG), 0, ctx->stream, mf->d_fits, d_colfit,
                           mf->d_V + V_COLSUM * mf->sumT);
        hipLaunchKernelGGL(k_datasum, dim3((unsigned)nfits), dim3(kWG), 0, ctx->stream, mf->d_fits, mf->d_nsum);
        if ((rc = check_launch(ctx, "k_colsum"))) break;
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = set_error(ctx, SFH_ERR_DEVICE, "sfh_multifit_create: synchronize failed");
            break;
        }
    } while (0);
    if (d_colfit) (void)hipFree(d_colfit);
    if (rc) {
        free_all(mf);
        delete mf;
        return rc;
    }
    ctx_retain(ctx);
    *out = mf;
    return SFH_OK;
}

int sfh_multifit_destroy(sfh_multifit* mf)
{
    if (!mf) return SFH_OK;
    sfh_ctx* ctx = mf->ctx;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        bind(ctx);
        (void)hipStreamSynchronize(ctx->stream);
        free_all(mf);
        delete mf;
    }
    ctx_release(ctx);
    return SFH_OK;
}

int sfh_multifit_get_shape(const sfh_multifit* mf, int64_t* item_elems, int* waves, int* wg_per_cu)
{
    if (!mf) return SFH_ERR_ARGUMENT;
    if (item_elems) *item_elems = mf->item_elems;
    if (waves) *waves = mf->nw;
    if (wg_per_cu) *wg_per_cu = mf->wg_per_cu;
    return SFH_OK;
}

int sfh_multifit_set_shape(sfh_multifit* mf, int64_t item_elems, int waves, int wg_per_cu)
{
    if (!mf) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = mf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    SFH_ARG_CHECK(ctx, item_elems >= kSB && (waves == 4 || waves == 8 || waves == 16) && wg_per_cu >= 1 && wg_per_cu <= 8,
                  "sfh_multifit_set_shape: item_elems >= 32, waves in {4, 8, 16}, 1 <= wg_per_cu <= 8");
    SFH_ARG_CHECK(ctx, kmax_for(mf->maxT, waves) > 0,
                  "sfh_multifit_set_shape: %d waves per workgroup hold at most %d templates per fit (largest: %lld)",
                  waves, waves == 4 ? 1024 : waves == 8 ? 704 : 512, (long long)mf->maxT);
    if (int rc = bind(ctx)) return rc;
    mf->item_elems = item_elems;
    mf->nw = waves;
    mf->wg_per_cu = wg_per_cu;
    return build_items(mf);
}

int sfh_multifit_autotune(sfh_multifit* mf, int reps, int64_t* item_elems, int* waves, int* wg_per_cu)
{
    if (!mf) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = mf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    SFH_ARG_CHECK(ctx, reps >= 1 && reps <= 100, "sfh_multifit_autotune: 1 <= reps <= 100");
    if (int rc = bind(ctx)) return rc;
    void* px = nullptr;
    if (int rc = scratch(ctx, "mf_x", mf->sumT * sizeof(double), &px)) return rc;
    double* d_X = (double*)px;
    std::vector<double> ones(mf->sumT, 1.0);
    SFH_HIP_TRY(ctx, hipMemcpyAsync(d_X, ones.data(), mf->sumT * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipEvent_t e0, e1;
    SFH_HIP_TRY(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(ctx, SFH_ERR_DEVICE, "hipEventCreate failed");
    }
    const double eps = 2.220446049250313e-16;
    const int64_t cand_items[] = {1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20};
    const int cand_nw[] = {4, 8, 16};
    const int cand_wg[] = {1, 2, 3, 4};
    double best = 1e300;
    int64_t best_items = mf->item_elems;
    int best_nw = mf->nw, best_wg = mf->wg_per_cu;
    int rc = SFH_OK;
    for (int64_t ie : cand_items) {
        for (int nw : cand_nw) {
        for (int wg : cand_wg) {
            if (rc) break;
            // no more than 32 waves per compute unit
            if (nw * wg > 32 || kmax_for(mf->maxT, nw) == 0) continue;
            mf->item_elems = ie;
            mf->nw = nw;
            mf->wg_per_cu = wg;
            if ((rc = build_items(mf))) break;
            launch_fg(mf, d_X, nullptr, eps, eps, 1);
            (void)hipEventRecord(e0, ctx->stream);
            for (int r = 0; r < reps; ++r) launch_fg(mf, d_X, nullptr, eps, eps, 1);
            (void)hipEventRecord(e1, ctx->stream);
            if ((rc = check_launch(ctx, "k_multi_fg"))) break;
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = set_error(ctx, SFH_ERR_DEVICE, "sfh_multifit_autotune: synchronize failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                best_items = ie;
                best_nw = nw;
                best_wg = wg;
            }
        }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    mf->item_elems = best_items;
    mf->nw = best_nw;
    mf->wg_per_cu = best_wg;
    if ((rc = build_items(mf))) return rc;
    if (item_elems) *item_elems = best_items;
    if (waves) *waves = best_nw;
    if (wg_per_cu) *wg_per_cu = best_wg;
    return SFH_OK;
}

int sfh_multifit_fg_device(sfh_multifit* mf, const double* d_X, double eps, double eps_grad, double* d_negLL,
                           double* d_G)
{
    if (!mf) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = mf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    SFH_ARG_CHECK(ctx, d_X && d_negLL, "sfh_multifit_fg_device: X and negLL are required");
    if (int rc = bind(ctx)) return rc;
    return fg_device_impl(mf, d_X, eps, eps_grad, d_negLL, d_G);
}

int sfh_multifit_fg(sfh_multifit* mf, const double* X, double eps, double eps_grad, double* negLL, double* G)
{
    if (!mf) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = mf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    SFH_ARG_CHECK(ctx, X && negLL, "sfh_multifit_fg: X and negLL are required");
    if (int rc = bind(ctx)) return rc;
    void *px = nullptr, *pn = nullptr, *pg = nullptr;
    if (int rc = scratch(ctx, "mf_x", mf->sumT * sizeof(double), &px)) return rc;
    if (int rc = scratch(ctx, "mf_nl", mf->nfits * sizeof(double), &pn)) return rc;
    if (G)
        if (int rc = scratch(ctx, "mf_g", mf->sumT * sizeof(double), &pg)) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(px, X, mf->sumT * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (int rc = fg_device_impl(mf, (const double*)px, eps, eps_grad, (double*)pn, (double*)pg)) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(negLL, pn, mf->nfits * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (G) SFH_HIP_TRY(ctx, hipMemcpyAsync(G, pg, mf->sumT * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    return SFH_OK;
}

int sfh_fit_options_default(sfh_fit_options* opt)
{
    if (!opt) return SFH_ERR_ARGUMENT;
    opt->transform = SFH_FIT_LOG;
    opt->method = SFH_OPT_BFGS;
    opt->memory = 10;
    opt->steps_per_graph = 16;
    opt->g_abstol = 1e-8;
    opt->max_iterations = 1000;
    opt->eps = 2.220446049250313e-16;
    opt->eps_grad = 2.220446049250313e-16;
    return SFH_OK;
}

int sfh_multifit_fit(sfh_multifit* mf, const double* x0, const sfh_fit_options* opt, sfh_fit_results* res)
{
    if (!mf) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = mf->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    sfh_fit_options o;
    sfh_fit_options_default(&o);
    if (opt) o = *opt;
    SFH_ARG_CHECK(ctx, o.transform == SFH_FIT_LOG || o.transform == SFH_FIT_SQRT, "sfh_multifit_fit: bad transform");
    SFH_ARG_CHECK(ctx, o.method == SFH_OPT_BFGS || o.method == SFH_OPT_LBFGS, "sfh_multifit_fit: bad method");
    SFH_ARG_CHECK(ctx, o.method != SFH_OPT_LBFGS || (o.memory >= 1 && o.memory <= kMaxMem),
                  "sfh_multifit_fit: 1 <= memory <= %d", kMaxMem);
    SFH_ARG_CHECK(ctx, o.steps_per_graph >= 1 && o.steps_per_graph <= 256, "sfh_multifit_fit: 1 <= steps_per_graph <= 256");
    SFH_ARG_CHECK(ctx, o.max_iterations >= 1 && o.g_abstol >= 0.0 && o.eps > 0.0 && o.eps_grad > 0.0,
                  "sfh_multifit_fit: bad options");
    SFH_ARG_CHECK(ctx, o.method == SFH_OPT_BFGS || (!res || (!res->invH_map && !res->invH_mle)),
                  "sfh_multifit_fit: the inverse Hessian is returned by the BFGS method only");
    if (int rc = bind(ctx)) return rc;
    const bool bfgs = o.method == SFH_OPT_BFGS;
    const bool logt = o.transform == SFH_FIT_LOG;
    if (int rc = ensure(ctx, &mf->d_st, mf->nfits)) return rc;
    if (int rc = ensure(ctx, &mf->d_X, mf->sumT)) return rc;
    if (int rc = ensure(ctx, &mf->d_active, mf->nfits)) return rc;
    if (int rc = ensure(ctx, &mf->d_count, 1)) return rc;
    if (bfgs) {
        if (int rc = ensure(ctx, &mf->d_H, mf->sumT2)) return rc;
        if (logt && res && res->invH_map)
            if (int rc = ensure(ctx, &mf->d_Hmap, mf->sumT2)) return rc;
    } else if (o.memory > mf->mem_cap) {
        for (double** pp : {&mf->d_S, &mf->d_Y, &mf->d_rho}) {
            if (*pp) (void)hipFree(*pp);
            *pp = nullptr;
        }
        if (int rc = ensure(ctx, &mf->d_S, (size_t)o.memory * mf->sumT)) return rc;
        if (int rc = ensure(ctx, &mf->d_Y, (size_t)o.memory * mf->sumT)) return rc;
        if (int rc = ensure(ctx, &mf->d_rho, (size_t)o.memory * mf->nfits)) return rc;
        mf->mem_cap = o.memory;
    }
    void* px0 = nullptr;
    if (x0) {
        if (int rc = scratch(ctx, "mf_x0", mf->sumT * sizeof(double), &px0)) return rc;
        SFH_HIP_TRY(ctx, hipMemcpyAsync(px0, x0, mf->sumT * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    }
    void* ph = nullptr;
    if (int rc = pinned(ctx, sizeof(int64_t), &ph)) return rc;

    OptParams p{};
    p.transform = o.transform;
    p.method = o.method;
    p.memory = o.memory;
    p.g_abstol = o.g_abstol;
    p.max_iterations = o.max_iterations;
    p.fits = mf->d_fits;
    p.st = mf->d_st;
    p.V = mf->d_V;
    p.sumT = mf->sumT;
    p.H = bfgs ? mf->d_H : nullptr;
    p.Hmap = (bfgs && logt && res && res->invH_map) ? mf->d_Hmap : nullptr;
    p.S = bfgs ? nullptr : mf->d_S;
    p.Y = bfgs ? nullptr : mf->d_Y;
    p.rho = bfgs ? nullptr : mf->d_rho;
    p.nfits = mf->nfits;
    p.ll_part = mf->d_llp;
    p.g_part = mf->d_gp;
    p.X = mf->d_X;
    p.active = mf->d_active;

    const unsigned nf = (unsigned)mf->nfits;
    hipLaunchKernelGGL(k_opt_init, dim3(nf), dim3(kWG), 0, ctx->stream, p, (const double*)px0, mf->d_nsum);
    if (int rc = check_launch(ctx, "k_opt_init")) return rc;

    // one graph: steps_per_graph x (propose, grouped fg!, accept), then the count of running fits
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SFH_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < o.steps_per_graph; ++k) {
        if (bfgs) hipLaunchKernelGGL(k_opt_propose<SFH_OPT_BFGS>, dim3(nf), dim3(kWG), 0, ctx->stream, p);
        else hipLaunchKernelGGL(k_opt_propose<SFH_OPT_LBFGS>, dim3(nf), dim3(kWG), 0, ctx->stream, p);
        launch_fg(mf, mf->d_X, mf->d_active, o.eps, o.eps_grad, 1);
        if (bfgs) hipLaunchKernelGGL(k_opt_accept<SFH_OPT_BFGS>, dim3(nf), dim3(kWG), 0, ctx->stream, p);
        else hipLaunchKernelGGL(k_opt_accept<SFH_OPT_LBFGS>, dim3(nf), dim3(kWG), 0, ctx->stream, p);
    }
    hipLaunchKernelGGL(k_opt_count, dim3(1), dim3(kWG), 0, ctx->stream, mf->d_st, mf->nfits, mf->d_count);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    if (le != hipSuccess || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return set_error(ctx, SFH_ERR_DEVICE, "sfh_multifit_fit: graph capture failed: %s",
                         hipGetErrorString(le != hipSuccess ? le : ce));
    }
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_error(ctx, SFH_ERR_DEVICE, "sfh_multifit_fit: graph instantiation failed");
    }
    // at most (stages) x (iterations + 1) x (30 line-search evaluations + 1) steps
    const int64_t max_steps = 2 * (o.max_iterations + 1) * 31;
    int64_t steps = 0;
    int rc = SFH_OK;
    for (;;) {
        {
            TimedLaunch tl(ctx, "multifit_graph");
            if (hipGraphLaunch(exec, ctx->stream) != hipSuccess) {
                rc = set_error(ctx, SFH_ERR_DEVICE, "sfh_multifit_fit: graph launch failed");
                break;
            }
        }
        steps += o.steps_per_graph;
        if (hipMemcpyAsync(ph, mf->d_count, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = set_error(ctx, SFH_ERR_DEVICE, "sfh_multifit_fit: step synchronisation failed");
            break;
        }
        if (*(int64_t*)ph == 0 || steps >= max_steps) break;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    resolve_timing(ctx);
    if (rc) return rc;

    if (res) {
        res->steps = steps;
        std::vector<OptState> st(mf->nfits);
        SFH_HIP_TRY(ctx, hipMemcpy(st.data(), mf->d_st, mf->nfits * sizeof(OptState), hipMemcpyDeviceToHost));
        std::vector<double> th(mf->sumT);
        SFH_HIP_TRY(ctx, hipMemcpy(th.data(), mf->d_V + V_THETA * mf->sumT, mf->sumT * sizeof(double),
                                   hipMemcpyDeviceToHost));
        if (res->mu_mle)
            for (int64_t i = 0; i < mf->sumT; ++i) res->mu_mle[i] = logt ? std::exp(th[i]) : th[i] * th[i];
        if (res->mu_map && logt)
            SFH_HIP_TRY(ctx, hipMemcpy(res->mu_map, mf->d_V + V_MUMAP * mf->sumT, mf->sumT * sizeof(double),
                                       hipMemcpyDeviceToHost));
        if (res->invH_mle && bfgs)
            SFH_HIP_TRY(ctx, hipMemcpy(res->invH_mle, mf->d_H, mf->sumT2 * sizeof(double), hipMemcpyDeviceToHost));
        if (res->invH_map && bfgs && logt)
            SFH_HIP_TRY(ctx, hipMemcpy(res->invH_map, mf->d_Hmap, mf->sumT2 * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t f = 0; f < mf->nfits; ++f) {
            for (int sg = 0; sg < 2; ++sg) {
                const int64_t k = sg * mf->nfits + f;
                if (res->fmin) res->fmin[k] = st[f].fmin[sg];
                if (res->iterations) res->iterations[k] = st[f].iters[sg];
                if (res->evaluations) res->evaluations[k] = st[f].evalsv[sg];
                if (res->status) res->status[k] = st[f].status[sg];
            }
        }
    }
    return SFH_OK;
}

int sfh_fg_device_noalloc(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data, const double* d_coeffs,
                          int want_F, double eps, double eps_grad, double* d_negLL, double* d_G, double* d_composite)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_ARG_CHECK(ctx, tpl && data && d_coeffs, "null pointer");
    SFH_ARG_CHECK(ctx, tpl->nbins == data->nbins, "axes(models,1) == axes(data,1) (%lld vs %lld)",
                  (long long)tpl->nbins, (long long)data->nbins);
    SFH_ARG_CHECK(ctx, !want_F || d_negLL, "d_negLL must be given when want_F");
    SFH_ARG_CHECK(ctx, single_pass_covers(tpl),
                  "sfh_fg_device_noalloc: 1 to %d templates and (ntemplates-1)*ld + 32 < 2^31 (got %lld x %lld)",
                  8 * mf::kSingleNW * mf::kSingleMaxK, (long long)tpl->nbins, (long long)tpl->ntemplates);
    if (ctx->device < 0 || ctx->device >= 64 || !g_noalloc_ptr[ctx->device])
        return set_error(ctx, SFH_ERR_DEVICE, "sfh_fg_device_noalloc: static buffer not initialised on device %d",
                         ctx->device);
    const int cus = __atomic_load_n(&g_cus_of[ctx->device], __ATOMIC_RELAXED);
    int64_t cap = 0;
    const int64_t need = single_pass_partials(tpl, cus, d_G != nullptr, &cap);
    SFH_ARG_CHECK(ctx, need <= kNoAllocDoubles,
                  "sfh_fg_device_noalloc: %lld partial sums exceed the static buffer (%lld doubles)",
                  (long long)need, (long long)kNoAllocDoubles);
    if (!want_F && !d_G && !d_composite) return SFH_OK;
    double* base = g_noalloc_ptr[ctx->device];
    return single_pass_launch(ctx, tpl, cus, data->d_data, d_coeffs, eps, eps_grad, want_F ? 1 : 0, d_negLL,
                              nullptr, d_G, 1.0, d_composite, d_G ? 1 : 0, base, d_G ? base + cap : nullptr,
                              false);
}


}  // extern "C"

// sfh_batched.hip -- B independent fg! evaluations in one pass over the
// templates (the HMC / ensemble call pattern of StarFormationHistories.jl,
// src/fitting/hmc_sample.jl:24-37, src/fitting/mcmc_sample.jl:12-23), on the
// gfx950 matrix cores.
//
//   M   = C X          composite for every walker       (nbins x B),  K = ntemplates
//   LL  = sum_i f(n_i, M_ib) in FP64                    (fitting_base.jl:84-96)
//   R   = 1 - n / max(M, eps)                           (fitting_base.jl:274-280)
//   G   = C^T R        -> gradient of -logL per walker  (fitting_base.jl:283, solvers.jl:30)
//
// Two GEMM launches + two small ones:
//   k_prep_x     coefficient vectors -> MFMA fragment tiles (split into BF16 pieces)
//   k_gemm<LL>   M tile in registers (never written); epilogue: FP64 log-likelihood
//                terms reduced per (bin tile, walker), R written as BF16 (or FP32)
//                fragment tiles for the next GEMM
//   k_gemm<GRAD> G partials over a split of the bins (split-K, fixed order)
//   k_finish     fixed-order reductions -> negLL[b], G[:, b] in FP64
//
// Precision ladder (precision argument):
//   FP32    FP32 templates x FP32 coefficients, v_mfma_f32_32x32x2_f32, FP32 accumulation.
//   BF16X2  templates and coefficients split hi+lo into BF16; products hh+hl+lh
//           with v_mfma_f32_32x32x16_bf16, FP32 accumulation.
//   F64     composite on the FP64 matrix cores (v_mfma_f64_16x16x4f64): FP32
//           templates widened exactly, FP64 coefficients, IEEE FP64 accumulation,
//           so the composite (and the FP64 log-likelihood) matches the FP64
//           oracle to rounding.  More BF16 pieces cannot get there: the BF16
//           MFMA sums its 16 products aligned to the largest one and drops what
//           falls below ~2^-27 of it, a bias no FP64 promotion removes.
// The gradient GEMM uses hi+lo BF16 (3 products; FP32 mode: FP32), FP32
// accumulation folded into FP64 every 32 K-blocks (512 bins; FP32 mode: 8).
//
// Operand layout ("fragment tiles"): for a 32-row x 16-k block, 64 lanes x
// 8 values, lane l holding row l%32 and k = 8*(l/32) + 0..7 (the lane layout of
// the 32x32x16 BF16 MFMA; the 32x32x2 FP32 MFMA walks the same 8 values as 8
// k-steps).  A piece of a block is 1 KiB of BF16 (or 2 KiB of FP32 as two
// 1 KiB halves).  Tiles are stored [kblock][rowblock][piece][lane], so one
// workgroup's operand for one K-block is a contiguous run of memory.  The
// gradient's K (bins) uses the permutation k = {4h..4h+3, 8+4h..8+4h+3}
// (h = l/32), which is exactly how the composite MFMA's accumulator hands the
// bins of a row block to a lane, so R is stored straight from registers.
#include "sfh_internal.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace sfh {
namespace bat {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
// 16-byte fragment as a native vector (arrays of HIP's union-based uint4 are
// not promoted to registers)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWG = 256;   // 4 waves
constexpr int kTR = 2;     // MFMA tiles per wave along rows
constexpr int kTC = 2;     // ... along columns
constexpr int kWR = 2;     // waves along rows
constexpr int kWC = 2;     // waves along columns
constexpr int kRB = kTR * kWR;   // 32-row blocks per workgroup tile (4 -> 128 rows)
constexpr int kCB = kTC * kWC;   // 32-col blocks per workgroup tile (4 -> 128 cols)
constexpr int kBM = 32 * kRB;
constexpr int kBN = 32 * kCB;

enum Mode { FP32 = 0, BF16X2 = 1, F64 = 2, GRAD_BF16 = 3, GRAD_FP32 = 4 };

// pieces per operand
__host__ __device__ constexpr int pieces_a(int m) { return 2; }   // BF16 hi+lo, or two FP32 halves
__host__ __device__ constexpr int pieces_b(int m) { return 2; }
__host__ __device__ constexpr bool is_f32(int m) { return m == FP32 || m == GRAD_FP32; }

// ---------------------------------------------------------------- bf16 split

__device__ inline uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Split v into n BF16 pieces, each capturing the residual of the previous.
template <int N>
__device__ inline void split_f32(float v, uint16_t* out)
{
#pragma unroll
    for (int p = 0; p < N; ++p) {
        const uint16_t h = bf16_rne(v);
        out[p] = h;
        v -= bf16_to_f32(h);
    }
}

template <int N>
__device__ inline void split_f64(double v, uint16_t* out)
{
#pragma unroll
    for (int p = 0; p < N; ++p) {
        const uint16_t h = bf16_rne((float)v);
        out[p] = h;
        v -= (double)bf16_to_f32(h);
    }
}

__device__ inline uint4 pack8(const uint16_t* h)
{
    uint4 r;
    r.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    r.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    r.z = (uint32_t)h[4] | ((uint32_t)h[5] << 16);
    r.w = (uint32_t)h[6] | ((uint32_t)h[7] << 16);
    return r;
}

// bin offset inside a 16-bin K-block for fragment slot j of lane half h
__device__ inline int gperm(int h, int j) { return j < 4 ? 4 * h + j : 8 + 4 * h + (j - 4); }

// ------------------------------------------------- FP64 log-likelihood term
//
// The batched epilogues evaluate one log and one quotient per (bin, walker),
// 2.5e8 of each per 4096-walker launch at 60,000 bins, so they use short FP64
// sequences instead of the library's extra-precise log and IEEE division.

// q = n / c to within one rounding: hardware reciprocal, two Newton steps and
// one residual correction.  Valid for c in [2^-1000, 2^1000]; outside that
// the IEEE division is used.
__device__ inline double quot(double n, double c)
{
    if (!(c >= 0x1p-1000 && c <= 0x1p1000)) return n / c;
    double r = __builtin_amdgcn_rcp(c);
    r = fma(fma(-c, r, 1.0), r, r);
    r = fma(fma(-c, r, 1.0), r, r);
    const double q = n * r;
    return fma(r, fma(-c, q, n), q);
}

// Natural log of a finite x > 0 (any other x: unspecified finite or NaN
// result), relative error a few units in the last place.  x = 2^e m with
// m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s), s = (m-1)/(m+1), |s| <= 0.1716,
// by its Taylor series 2s (1 + z/3 + z^2/5 + ... + z^9/19), z = s^2 (the first
// omitted term is below 2.5e-17 relative); e ln 2 in two parts.
__device__ inline double log_pos(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);          // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752440) {
        m += m;
        e -= 1;
    }
    const double f = m - 1.0;                           // exact
    const double d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    double s = f * r;
    s = fma(r, fma(-d, s, f), s);
    const double z = s * s;
    double p = 1.0 / 19.0;
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double t = s + s;
    const double lm = fma(t * z, p, t);
    const double de = (double)e;
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

// One bin's contribution to logL, n - c - n log(n / c) (-c for n <= 0), and
// R = 1 - n / c, for a composite already clamped to c >= eps
// (src/fitting/fitting_base.jl:84-96, 274-280).
__device__ inline void ll_term(double n, double c, double& ll, float& r)
{
    const double q = quot(n, c);
    ll += (n > 0.0) ? (n - c - n * log_pos(q)) : -c;
    r = (float)(1.0 - q);
}


// ------------------------------------------------------------ layout builders

// Per-template power-of-two scale: e_t = exponent of max_i |C[i,t]| (0 if the
// template is all zero).  The composite GEMM uses C[i,t] * 2^-e_t and
// coefficients * 2^e_t (exact), so every template's leading piece is O(1).
__global__ __launch_bounds__(256) void k_template_exp(const float* __restrict__ C, int64_t ld,
                                                      int64_t nbins, int32_t* __restrict__ ex)
{
    __shared__ float sh[256];
    const int64_t t = blockIdx.x;
    float m = 0.f;
    for (int64_t i = threadIdx.x; i < nbins; i += 256) m = fmaxf(m, fabsf(C[t * ld + i]));
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int e = 0;
        if (sh[0] > 0.f && isfinite(sh[0])) frexpf(sh[0], &e);
        ex[t] = e;
    }
}

// Composite A operand: rows = bins, k = templates.  One thread per
// (kblock, rowblock, lane).  PA BF16 pieces, or FP32 (PA = 2 halves).
template <int M>
__global__ __launch_bounds__(256) void k_build_a1(const float* __restrict__ C, int64_t ld,
                                                  int64_t nbins, int64_t T, const int32_t* __restrict__ ex,
                                                  int64_t nrb, int64_t nkb, uint4* __restrict__ out)
{
    constexpr int PA = pieces_a(M);
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nkb * nrb * 64) return;
    const int lane = (int)(g & 63);
    const int64_t rb = (g >> 6) % nrb, kb = (g >> 6) / nrb;
    const int64_t bin = rb * 32 + (lane & 31);
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t t = kb * 16 + 8 * (lane >> 5) + j;
        v[j] = (t < T && bin < nbins) ? ldexpf(C[t * ld + bin], -ex[t]) : 0.f;
    }
    uint4* o = out + ((kb * nrb + rb) * PA) * 64 + lane;
    if (is_f32(M)) {
        o[0] = make_uint4(__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]), __float_as_uint(v[3]));
        o[64] = make_uint4(__float_as_uint(v[4]), __float_as_uint(v[5]), __float_as_uint(v[6]), __float_as_uint(v[7]));
    } else {
        uint16_t h[PA][8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint16_t s[PA];
            split_f32<PA>(v[j], s);
#pragma unroll
            for (int p = 0; p < PA; ++p) h[p][j] = s[p];
        }
#pragma unroll
        for (int p = 0; p < PA; ++p) o[p * 64] = pack8(h[p]);
    }
}

// Gradient A operand: rows = templates, k = bins (permuted within a K-block).
template <bool F32>
__global__ __launch_bounds__(256) void k_build_a2(const float* __restrict__ C, int64_t ld,
                                                  int64_t nbins, int64_t T, int64_t nrb, int64_t nkb,
                                                  uint4* __restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nkb * nrb * 64) return;
    const int lane = (int)(g & 63);
    const int64_t rb = (g >> 6) % nrb, kb = (g >> 6) / nrb;
    const int64_t t = rb * 32 + (lane & 31);
    const int h = lane >> 5;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t bin = kb * 16 + gperm(h, j);
        v[j] = (t < T && bin < nbins) ? C[t * ld + bin] : 0.f;
    }
    uint4* o = out + ((kb * nrb + rb) * 2) * 64 + lane;
    if (F32) {
        o[0] = make_uint4(__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]), __float_as_uint(v[3]));
        o[64] = make_uint4(__float_as_uint(v[4]), __float_as_uint(v[5]), __float_as_uint(v[6]), __float_as_uint(v[7]));
    } else {
        uint16_t h0[8], h1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint16_t s[2];
            split_f32<2>(v[j], s);
            h0[j] = s[0];
            h1[j] = s[1];
        }
        o[0] = pack8(h0);
        o[64] = pack8(h1);
    }
}

// Coefficients X (T x B, column-major FP64) -> composite B operand tiles.
template <int M>
__global__ __launch_bounds__(256) void k_prep_x(const double* __restrict__ X, int64_t T, int64_t B,
                                                const int32_t* __restrict__ ex, int64_t ncb, int64_t nkb,
                                                uint4* __restrict__ out)
{
    constexpr int PB = pieces_b(M);
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= nkb * ncb * 64) return;
    const int lane = (int)(g & 63);
    const int64_t cb = (g >> 6) % ncb, kb = (g >> 6) / ncb;
    const int64_t w = cb * 32 + (lane & 31);
    double v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t t = kb * 16 + 8 * (lane >> 5) + j;
        v[j] = (t < T && w < B) ? ldexp(X[w * T + t], ex[t]) : 0.0;
    }
    uint4* o = out + ((kb * ncb + cb) * PB) * 64 + lane;
    if (is_f32(M)) {
        float f[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] = (float)v[j];
        o[0] = make_uint4(__float_as_uint(f[0]), __float_as_uint(f[1]), __float_as_uint(f[2]), __float_as_uint(f[3]));
        o[64] = make_uint4(__float_as_uint(f[4]), __float_as_uint(f[5]), __float_as_uint(f[6]), __float_as_uint(f[7]));
    } else {
        uint16_t h[PB][8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint16_t s[PB];
            split_f64<PB>(v[j], s);
#pragma unroll
            for (int p = 0; p < PB; ++p) h[p][j] = s[p];
        }
#pragma unroll
        for (int p = 0; p < PB; ++p) o[p * 64] = pack8(h[p]);
    }
}


// ------------------------------------------------- LDS ring filled by DMA
// gfx950's global_load_lds_dwordx4 writes 16 bytes per lane straight into LDS
// at (wave-uniform base + 16 * lane), so a 1 KiB fragment chunk is one
// instruction per wave and needs no staging registers.  Completion is tracked
// with vmcnt; a bare s_barrier (no fence, which would force vmcnt(0)) then
// publishes the stage to the other waves.
#define SFH_LDS(p) ((__attribute__((address_space(3))) void*)(p))

__device__ inline void dma_chunk(const u32x4* g_chunk, u32x4* lds_chunk, int lane)
{
    __builtin_amdgcn_global_load_lds(g_chunk + lane, SFH_LDS(lds_chunk), 16, 0, 0);
}

// wait until at most n vector-memory operations of this wave are outstanding
__device__ inline void wait_vmcnt(int n)
{
    switch (n) {
    case 0: __builtin_amdgcn_s_waitcnt(0xF70); break;
    case 2: __builtin_amdgcn_s_waitcnt(0xF72); break;
    case 3: __builtin_amdgcn_s_waitcnt(0xF73); break;
    case 4: __builtin_amdgcn_s_waitcnt(0xF74); break;
    case 6: __builtin_amdgcn_s_waitcnt(0xF76); break;
    case 8: __builtin_amdgcn_s_waitcnt(0xF78); break;
    case 9: __builtin_amdgcn_s_waitcnt(0xF79); break;
    case 12: __builtin_amdgcn_s_waitcnt(0xF7C); break;
    default: __builtin_amdgcn_s_waitcnt(0xF70); break;
    }
}

__device__ inline void ring_barrier()
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Barrier for register-staged pipelines: this wave's LDS writes complete
// (lgkmcnt(0)), global loads still in flight stay in flight (no fence, which
// would wait for vmcnt(0)).
__device__ inline void stage_barrier()
{
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// ------------------------------------------------------------------ the GEMM

// hi+lo product table (piece of A, piece of B), smallest first so the FP32
// accumulator adds them before the leading product: (1,0) (0,1) (0,0)
__host__ __device__ constexpr int prod_a(int q) { return q == 0 ? 1 : 0; }
__host__ __device__ constexpr int prod_b(int q) { return q == 1 ? 1 : 0; }
constexpr int kNProd = 3;

template <int Q, int NQ, class F>
__device__ inline void static_for(F&& f)
{
    if constexpr (Q < NQ) {
        f(std::integral_constant<int, Q>{});
        static_for<Q + 1, NQ>(f);
    }
}

// Tile swizzle: consecutive workgroup ids go round-robin over the 8 XCDs, so
// give each XCD runs of ids covering a compact GR x GC patch of tiles; the
// patch's A and B K-blocks are then fetched into that XCD's L2 once and
// reused by its ~32 resident workgroups.
struct Swizzle {
    int nrt, nct, gr, gc, nsr, nsc;
    int xsplit;     // split-K launches: 1-D grid, XCD x runs splits x, x + 8, ... (nsplit % 8 == 0)
};

__device__ inline bool swizzle(const Swizzle& s, int64_t id, int* rt, int* ct)
{
    const int64_t xcd = id & 7, local = id >> 3;
    const int per = s.gr * s.gc;
    const int64_t patch = (local / per) * 8 + xcd;
    const int within = (int)(local % per);
    const int64_t pr = patch / s.nsc, pc = patch % s.nsc;
    if (pr >= s.nsr) return false;
    *rt = (int)(pr * s.gr + within / s.gc);
    *ct = (int)(pc * s.gc + within % s.gc);
    return *rt < s.nrt && *ct < s.nct;
}

struct EpiLL {
    const double* data;      // n_i
    int64_t nbins;
    int64_t B;
    double eps;              // LL clamp: eps(promote_type(composite, data))
    double eps_g;            // R clamp: eps(eltype(composite))
    double eps_hi;           // max(eps, eps_g): below it R is recomputed with eps_g
    double* ll_part;         // [row tile][Bpad]
    uint4* R;                // gradient B operand tiles [kb2][cb][2][64]
    int64_t ncb;             // column blocks (Bpad/32)
    int64_t Bpad;
};

// The LL and R of one (bin, walker) with the reference's two clamps: LL with
// eps of the promoted type (:86), R with eps(eltype(composite)) (:277).  They
// differ only where m < max(eps, eps_g); there R is recomputed (the same
// quotient when the two are equal).
__device__ inline void ll_r_term(const EpiLL& el, double n, double m, double& ll, float& r)
{
    ll_term(n, m > el.eps ? m : el.eps, ll, r);
    if (m < el.eps_hi) r = (float)(1.0 - quot(n, m > el.eps_g ? m : el.eps_g));
}

struct EpiGrad {
    float* gpart;            // [split][Bpad][Tpad2]
    int64_t Bpad, Tpad2;
};

// Workgroup tile (32 kTR WR) x (32 kTC WC): WR x WC waves of 64 x 64.  The
// composite forms use 2 x 2 waves (128 x 128); the split-K gradient uses
// 4 x 2 waves (256 templates x 128 walkers, 512 threads, one workgroup per
// CU): a quarter fewer LDS and L2 bytes per matrix-core product and half the
// re-reads of the R operand, at the same two waves per SIMD.
// Chunks of rowblocks / column blocks past the operand (partial edge tiles)
// are read from the last valid block and their results are not written.
template <int M, bool LL, int WR = kWR, int WC = kWC>
__global__ __launch_bounds__(WR * WC * 64, 1) void k_gemm(const uint4* __restrict__ A_, const uint4* __restrict__ B_,
                                                          int64_t nrb_total, int64_t ncb_total, int64_t kb_per_split,
                                                          int64_t nkb_total, Swizzle sw, EpiLL el, EpiGrad eg)
{
    constexpr int NW = WR * WC;                 // waves
    constexpr int RB = kTR * WR;                // 32-row blocks per tile
    constexpr int CBK = kTC * WC;               // 32-column blocks per tile
    constexpr int BN = 32 * CBK;
    constexpr int PA = pieces_a(M);
    constexpr int PB = pieces_b(M);
    constexpr int CA = RB * PA;                 // A chunks (1 KiB) per K-block
    constexpr int CB = CBK * PB;                // B chunks per K-block
    constexpr int CH = CA + CB;                 // chunks per stage
    constexpr int LPW = CH / NW;                // DMA instructions per wave per stage
    // ring depth: the split-K gradient streams R from HBM and wants a deep ring;
    // the composite GEMMs hit L2 and prefer the occupancy of a shallow one
    constexpr int S = (M == GRAD_BF16 || M == GRAD_FP32) ? 4 : 2;
    constexpr int KP = M == GRAD_BF16 ? 32 : M == GRAD_FP32 ? 8 : 0;
    static_assert(CH % NW == 0, "stage size");

    __shared__ u32x4 ring[S][CH * 64];
    __shared__ double sLL[LL ? WR : 1][LL ? BN : 1];
    const u32x4* A = reinterpret_cast<const u32x4*>(A_);
    const u32x4* Bm = reinterpret_cast<const u32x4*>(B_);

    int rt, ct;
    int64_t split;
    if (sw.xsplit) {
        // consecutive workgroups go round-robin over the 8 XCDs, so workgroup
        // id runs on XCD id % 8: that XCD takes whole splits, its tiles in
        // column-major order, so the workgroups resident together share each
        // R column tile (3 row tiles) and each A row tile (~10 column tiles)
        // through the XCD's L2
        const int64_t id = blockIdx.x, tiles = (int64_t)sw.nrt * sw.nct;
        const int64_t local = id >> 3, tile = local % tiles;
        split = (local / tiles) * 8 + (id & 7);
        ct = (int)(tile / sw.nrt);
        rt = (int)(tile % sw.nrt);
    } else {
        split = blockIdx.y;
        if (!swizzle(sw, blockIdx.x, &rt, &ct)) return;
    }
    const int64_t kb0 = split * kb_per_split;
    int64_t kb1 = kb0 + kb_per_split;
    if (kb1 > nkb_total) kb1 = nkb_total;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / WC, wc = wave % WC;

    f32x16 acc[kTR][kTC];
    double acc64[kTR][kTC][16];
#pragma unroll
    for (int a = 0; a < kTR; ++a)
#pragma unroll
        for (int b = 0; b < kTC; ++b) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                acc[a][b][i] = 0.f;
                acc64[a][b][i] = 0.0;
            }
        }

    // stage kb: chunk c < CA is A rowblock-piece c of this tile, else B
    auto issue = [&](int64_t kb) {
        u32x4* st = ring[(kb - kb0) % S];
#pragma unroll
        for (int i = 0; i < LPW; ++i) {
            const int c = wave * LPW + i;
            const u32x4* g;
            if (c < CA) {
                int64_t rb = (int64_t)rt * RB + c / PA;
                if (rb >= nrb_total) rb = nrb_total - 1;
                g = A + ((kb * nrb_total + rb) * PA + c % PA) * 64;
            } else {
                int64_t cb = (int64_t)ct * CBK + (c - CA) / PB;
                if (cb >= ncb_total) cb = ncb_total - 1;
                g = Bm + ((kb * ncb_total + cb) * PB + (c - CA) % PB) * 64;
            }
            dma_chunk(g, st + c * 64, lane);
        }
    };
    for (int64_t k = kb0; k < kb0 + S - 1 && k < kb1; ++k) issue(k);

    int since = 0;
    for (int64_t kb = kb0; kb < kb1; ++kb) {
        const int64_t later = (kb1 - 1 - kb) < (S - 2) ? (kb1 - 1 - kb) : (S - 2);
        wait_vmcnt((int)later * LPW);
        ring_barrier();
        if (kb + S - 1 < kb1) issue(kb + S - 1);
        const u32x4* st = ring[(kb - kb0) % S];
        const bool more = kb + 1 < kb1;
        // fragments of this wave
        u32x4 fa[kTR][PA], fb[kTC][PB];
#pragma unroll
        for (int a = 0; a < kTR; ++a)
#pragma unroll
            for (int p = 0; p < PA; ++p) fa[a][p] = st[((wr * kTR + a) * PA + p) * 64 + lane];
#pragma unroll
        for (int b = 0; b < kTC; ++b)
#pragma unroll
            for (int p = 0; p < PB; ++p) fb[b][p] = st[(CA + (wc * kTC + b) * PB + p) * 64 + lane];

        if constexpr (is_f32(M)) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
#pragma unroll
                for (int a = 0; a < kTR; ++a)
#pragma unroll
                    for (int b = 0; b < kTC; ++b) {
                        const uint32_t wa = fa[a][s >> 2][s & 3];
                        const uint32_t wb = fb[b][s >> 2][s & 3];
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(__uint_as_float(wa), __uint_as_float(wb),
                                                                        acc[a][b], 0, 0, 0);
                    }
            }
        } else {
            static_for<0, kNProd>([&](auto qc) {
                constexpr int pa_ = prod_a(decltype(qc)::value);
                constexpr int pb_ = prod_b(decltype(qc)::value);
#pragma unroll
                for (int a = 0; a < kTR; ++a)
#pragma unroll
                    for (int b = 0; b < kTC; ++b)
                        acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, fa[a][pa_]), __builtin_bit_cast(bf16x8, fb[b][pb_]),
                            acc[a][b], 0, 0, 0);
            });
        }
        if constexpr (KP > 0) {
            if (++since == KP || !more) {
                since = 0;
#pragma unroll
                for (int a = 0; a < kTR; ++a)
#pragma unroll
                    for (int b = 0; b < kTC; ++b)
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            acc64[a][b][i] += (double)acc[a][b][i];
                            acc[a][b][i] = 0.f;
                        }
            }
        }
    }
    if constexpr (KP == 0) {
#pragma unroll
        for (int a = 0; a < kTR; ++a)
#pragma unroll
            for (int b = 0; b < kTC; ++b)
#pragma unroll
                for (int i = 0; i < 16; ++i) acc64[a][b][i] = (double)acc[a][b][i];
    }

    const int h = lane >> 5, col = lane & 31;
    if constexpr (LL) {
        // rows = bins, cols = walkers
        double part[kTC];
#pragma unroll
        for (int b = 0; b < kTC; ++b) part[b] = 0.0;
#pragma unroll
        for (int a = 0; a < kTR; ++a) {
            const int64_t rbk = (int64_t)rt * RB + wr * kTR + a;   // 32-bin block
#pragma unroll
            for (int b = 0; b < kTC; ++b) {
                const int64_t cbk = (int64_t)ct * CBK + wc * kTC + b;
                const int64_t w = cbk * 32 + col;
                const bool wok = w < el.B;
                float rv[16];
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int64_t bin = rbk * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                    float r = 0.f;
                    if (bin < el.nbins && wok) {
                        const double m = acc64[a][b][i];
                        ll_r_term(el, el.data[bin], m, part[b], r);
                    }
                    rv[i] = r;
                }
                // R fragments: K-blocks 2*rbk (i = 0..7) and 2*rbk+1 (i = 8..15)
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int64_t kb2 = 2 * rbk + half;
                    uint4* o = el.R + ((kb2 * el.ncb + cbk) * 2) * 64 + lane;
                    if constexpr (is_f32(M)) {
                        const int o8 = 8 * half;
                        o[0] = make_uint4(__float_as_uint(rv[o8 + 0]), __float_as_uint(rv[o8 + 1]),
                                          __float_as_uint(rv[o8 + 2]), __float_as_uint(rv[o8 + 3]));
                        o[64] = make_uint4(__float_as_uint(rv[o8 + 4]), __float_as_uint(rv[o8 + 5]),
                                           __float_as_uint(rv[o8 + 6]), __float_as_uint(rv[o8 + 7]));
                    } else {
                        uint16_t h0[8], h1[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            uint16_t s2[2];
                            split_f32<2>(rv[8 * half + j], s2);
                            h0[j] = s2[0];
                            h1[j] = s2[1];
                        }
                        o[0] = pack8(h0);
                        o[64] = pack8(h1);
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kTC; ++b) {
            part[b] += __shfl_xor(part[b], 32, 64);
            if (h == 0) sLL[wr][(wc * kTC + b) * 32 + col] = part[b];
        }
        __syncthreads();
        for (int k = tid; k < BN; k += NW * 64) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < WR; ++r) s += sLL[r][k];
            el.ll_part[(int64_t)rt * el.Bpad + (int64_t)ct * BN + k] = s;
        }
    } else {
        // rows = templates, cols = walkers; write FP32 partials [split][walker][template]
#pragma unroll
        for (int a = 0; a < kTR; ++a) {
            const int64_t rbk = (int64_t)rt * RB + wr * kTR + a;
            if (rbk * 32 >= eg.Tpad2) continue;
#pragma unroll
            for (int b = 0; b < kTC; ++b) {
                const int64_t w = ((int64_t)ct * CBK + wc * kTC + b) * 32 + col;
                if (w >= eg.Bpad) continue;
                float* o = eg.gpart + (split * eg.Bpad + w) * eg.Tpad2 + rbk * 32 + 4 * h;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const float4 v = make_float4((float)acc64[a][b][4 * g + 0], (float)acc64[a][b][4 * g + 1],
                                                 (float)acc64[a][b][4 * g + 2], (float)acc64[a][b][4 * g + 3]);
                    *reinterpret_cast<float4*>(o + 8 * g) = v;
                }
            }
        }
    }
}

// ------------------------------------------------- FP64 matrix-core composite

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// Coefficients X (T x B, column-major FP64) -> [kblock][col tile][16 k][128 walkers] FP64.
// One workgroup per (kblock, col tile): reads run along the templates (16
// consecutive t of a walker, 128 bytes), writes along the walkers, through an
// LDS tile.
__global__ __launch_bounds__(256) void k_prep_x64(const double* __restrict__ X, int64_t T, int64_t B,
                                                  int64_t nct, int64_t nkb, double* __restrict__ out)
{
    __shared__ double tile[16][kBN + 1];
    const int64_t ct = blockIdx.x % nct, kb = blockIdx.x / nct;
#pragma unroll
    for (int j = 0; j < 16 * kBN / 256; ++j) {
        const int idx = threadIdx.x + 256 * j;
        const int c = idx / 16, k = idx % 16;
        const int64_t t = kb * 16 + k, w = ct * kBN + c;
        tile[k][c] = (t < T && w < B) ? X[w * T + t] : 0.0;
    }
    __syncthreads();
    double* o = out + (kb * nct + ct) * 16 * kBN;
#pragma unroll
    for (int j = 0; j < 16 * kBN / 256; ++j) {
        const int idx = threadIdx.x + 256 * j;
        o[idx] = tile[idx / kBN][idx % kBN];
    }
}

// float index of element (row r in [0,128), k in [0,16)) of an FP32 A tile
// stored as fragment tiles [rowblock][half][lane][4]
__device__ inline int a_elem(int r, int k)
{
    const int rb = r >> 5, lanep = (r & 31) + 32 * (k >> 3), j = k & 7;
    return (((rb * 2 + (j >> 2)) * 64 + lanep) << 2) + (j & 3);
}

// Composite on the FP64 matrix cores + the same log-likelihood / R epilogue as
// k_gemm<LL>.  Workgroup tile 128 bins x 128 walkers, 4 waves of 64 x 64, each
// wave 4 x 4 v_mfma_f64_16x16x4f64 tiles (lane l: A[row l%16][k l/16],
// B[k l/16][col l%16], D[row l/16 + 4i][col l%16] in accumulator i).  A (FP32 fragment tiles)
// and X (FP64) K-blocks go global -> registers (two K-blocks ahead, two staging
// sets) -> one of two LDS stages; the R tile is re-laid out through LDS into the
// gradient GEMM's fragment tiles.
__global__ __launch_bounds__(kWG, 1) void k_gemm_f64_ll(const uint4* __restrict__ A_, const double* __restrict__ X,
                                                        int64_t nrb_total, int64_t nct_total, int64_t nkb,
                                                        Swizzle sw, EpiLL el)
{
    constexpr int CA = kRB * 2;                 // 8 KiB: FP32 A tile of one K-block (two halves per rowblock)
    constexpr int CX = 16;                      // 16 KiB: one 128-walker FP64 row per k
    constexpr int STAGE_U4 = (CA + CX) * 64;    // 1536 16-byte words per K-block
    constexpr int NPER = STAGE_U4 / kWG;        // 6 per thread
    constexpr int NPA = CA * 64 / kWG;          // the first 2 of them from the A tile
    static_assert(STAGE_U4 % kWG == 0 && (CA * 64) % kWG == 0, "stage split");
    // two LDS stages (48 KiB) during the K loop; the 128 x 128 FP32 R tile (64 KiB) afterwards
    __shared__ u32x4 lds[4096];
    __shared__ double sLL[kWR][kBN];
    float* sR = reinterpret_cast<float*>(lds);

    int rt, ct;
    if (!swizzle(sw, blockIdx.x, &rt, &ct)) return;
    const u32x4* A = reinterpret_cast<const u32x4*>(A_);
    const u32x4* Xv = reinterpret_cast<const u32x4*>(X);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave / kWC, wc = wave % kWC;
    const int lr = lane & 15, lk = lane >> 4;

    f64x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};

    // element u of a stage: the A tile (contiguous) then the X rows (contiguous)
    const int64_t a_row0 = (int64_t)rt * kRB * 2 * 64;
    const int64_t a_stride = nrb_total * 2 * 64, x_stride = nct_total * CX * 64;
    const u32x4* x0 = Xv + (int64_t)ct * CX * 64;
#define SFH_GLOAD(kb, R)                                                                         \
    do {                                                                                         \
        const u32x4* pa_ = A + (kb) * a_stride + a_row0;                                         \
        const u32x4* px_ = x0 + (kb) * x_stride;                                                 \
        _Pragma("unroll") for (int i_ = 0; i_ < NPER; ++i_)                                      \
            R[i_] = i_ < NPA ? pa_[tid + i_ * kWG] : px_[tid + (i_ - NPA) * kWG];                \
    } while (0)
#define SFH_LSTORE(st, R)                                                                        \
    _Pragma("unroll") for (int i_ = 0; i_ < NPER; ++i_) lds[(st) * STAGE_U4 + tid + i_ * kWG] = R[i_]

    auto compute = [&](int st) {
        const float* cA = reinterpret_cast<const float*>(lds + st * STAGE_U4);
        const double* cX = reinterpret_cast<const double*>(lds + st * STAGE_U4 + CA * 64);
        // fragments of step ks + 1 are read while step ks's products run
        float ga[2][4];
        double fb[2][4];
        auto rd = [&](int ks, int s) {
            const int k = 4 * ks + lk;
#pragma unroll
            for (int a = 0; a < 4; ++a) ga[s][a] = cA[a_elem(wr * 64 + a * 16 + lr, k)];
#pragma unroll
            for (int b = 0; b < 4; ++b) fb[s][b] = cX[k * kBN + wc * 64 + b * 16 + lr];
        };
        rd(0, 0);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int s = ks & 1;
            if (ks < 3) rd(ks + 1, s ^ 1);
            double fa[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) fa[a] = (double)ga[s][a];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[a], fb[s][b], acc[a][b], 0, 0, 0);
        }
    };

    // staging set A holds K-block kb+1, set B holds kb+2 (loads two steps ahead;
    // past the end the last K-block is re-read, keeping the registers defined)
    u32x4 ra[NPER], rb[NPER];
    const int64_t last = nkb - 1;
    SFH_GLOAD(0, ra);
    SFH_LSTORE(0, ra);
    SFH_GLOAD(1 < last ? 1 : last, ra);
    // every load into ra issues before any into rb, as in the loop, so the
    // waits the compiler places at the loop's stores count only ra's loads
    __builtin_amdgcn_sched_barrier(0);
    SFH_GLOAD(2 < last ? 2 : last, rb);
    stage_barrier();
    // no conditional stores inside the loop: a skipped store would leave its
    // staging registers with a load in flight on one path, and the wait the
    // compiler then places before the next load into them drains every load
    __builtin_assume(nkb >= 2);                 // the host pads the template axis to two K-blocks or more
    int64_t kb = 0;
    for (; kb + 1 < nkb; kb += 2) {
        compute(0);
        SFH_LSTORE(1, ra);                      // K-block kb+1
        stage_barrier();
        SFH_GLOAD(kb + 3 < last ? kb + 3 : last, ra);
        compute(1);
        SFH_LSTORE(0, rb);                      // K-block kb+2 (the last one again when kb+2 == nkb)
        stage_barrier();
        SFH_GLOAD(kb + 4 < last ? kb + 4 : last, rb);
    }
    if (kb < nkb) compute(0);                   // odd nkb: stage 0 holds K-block nkb-1
#undef SFH_GLOAD
#undef SFH_LSTORE
    __syncthreads();    // every wave is done with the stages before they become the R tile

    // epilogue: LL terms (FP64) and R into LDS
    double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int cl = wc * 64 + b * 16 + lr;                 // walker within the tile
            const int64_t w = (int64_t)ct * kBN + cl;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int rl = wr * 64 + a * 16 + lk + 4 * i;     // bin within the tile
                const int64_t bin = (int64_t)rt * kBM + rl;
                float r = 0.f;
                if (bin < el.nbins && w < el.B) {
                    const double m = acc[a][b][i];
                    ll_r_term(el, el.data[bin], m, part[b], r);
                }
                sR[rl * kBN + cl] = r;
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        part[b] += __shfl_xor(part[b], 16, 64);
        part[b] += __shfl_xor(part[b], 32, 64);
        if (lk == 0) sLL[wr][wc * 64 + b * 16 + lr] = part[b];
    }
    __syncthreads();
    for (int k = tid; k < kBN; k += kWG)
        el.ll_part[(int64_t)rt * el.Bpad + (int64_t)ct * kBN + k] = sLL[0][k] + sLL[1][k];
    // R fragments: 8 K-blocks x 4 column blocks x 64 lanes per tile
    for (int item = tid; item < 8 * kCB * 64; item += kWG) {
        const int lp = item & 63, cbl = (item >> 6) % kCB, kbl = (item >> 6) / kCB;
        const int cl = cbl * 32 + (lp & 31), hp = lp >> 5;
        uint16_t h0[8], h1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            uint16_t s2[2];
            split_f32<2>(sR[(kbl * 16 + gperm(hp, j)) * kBN + cl], s2);
            h0[j] = s2[0];
            h1[j] = s2[1];
        }
        const int64_t kb2 = (int64_t)rt * 8 + kbl, cb = (int64_t)ct * kCB + cbl;
        uint4* o = el.R + ((kb2 * el.ncb + cb) * 2) * 64 + lp;
        o[0] = pack8(h0);
        o[64] = pack8(h1);
    }
}

// Workgroups [0, nll_blocks): negLL[b] = -(sum_rt ll_part[rt][b]) (-> +Inf
// for an exact 0, fitting_base.jl:95), 16 walkers x 16 row slices per
// workgroup (slice s sums rows s, s + 16, ... in order; the slices then add in
// order).  The rest: G[t + T*b] = sum_split gpart[split][b][t] (the gradient
// of -logL), one element per thread.
constexpr int kFinW = 16, kFinS = 16;
__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ ll_part, int64_t nrt,
                                                const float* __restrict__ gpart, int nsplit, int64_t T,
                                                int64_t B, int64_t Bpad, int64_t Tpad2, int64_t nll_blocks,
                                                double* __restrict__ negLL, double* __restrict__ G)
{
    if ((int64_t)blockIdx.x < nll_blocks) {
        __shared__ double sh[kFinS][kFinW];
        const int wl = threadIdx.x % kFinW, sl = threadIdx.x / kFinW;
        const int64_t b = (int64_t)blockIdx.x * kFinW + wl;
        double s = 0.0;
        if (b < B) {
#pragma unroll 8
            for (int64_t r = sl; r < nrt; r += kFinS) s += ll_part[r * Bpad + b];
        }
        sh[sl][wl] = s;
        __syncthreads();
        if (sl == 0 && b < B) {
            double v = sh[0][wl];
#pragma unroll
            for (int k = 1; k < kFinS; ++k) v += sh[k][wl];
            negLL[b] = (v != 0.0) ? -v : INFINITY;
        }
        return;
    }
    const int64_t g = ((int64_t)blockIdx.x - nll_blocks) * 256 + threadIdx.x;
    if (G && g < T * B) {
        const int64_t b = g / T, t = g - b * T;
        double s = 0.0;
        for (int k = 0; k < nsplit; ++k) s += (double)gpart[((int64_t)k * Bpad + b) * Tpad2 + t];
        G[g] = s;
    }
}

}  // namespace bat

using namespace bat;

// Per-template-set operands for the batched path, built on first use.
struct BatchedOperands {
    int64_t npad = 0, tpad1 = 0, tpad2 = 0;
    int32_t* ex = nullptr;                 // per-template exponents (BF16X2 scaling)
    int32_t* ex0 = nullptr;                // all-zero exponents (FP32 / F64: no scaling)
    uint4* a1[2] = {nullptr, nullptr};     // composite operand: [0] FP32 tiles, [1] BF16 hi+lo tiles
    uint4* a2[2] = {nullptr, nullptr};     // gradient operand:  [0] BF16 hi+lo, [1] FP32
};

static std::mutex g_ops_mu;
static std::map<const sfh_templates*, BatchedOperands> g_ops;

void release_batched(const sfh_templates* t)
{
    std::lock_guard<std::mutex> g(g_ops_mu);
    auto it = g_ops.find(t);
    if (it == g_ops.end()) return;
    BatchedOperands& o = it->second;
    if (o.ex) (void)hipFree(o.ex);
    if (o.ex0) (void)hipFree(o.ex0);
    for (auto* p : o.a1)
        if (p) (void)hipFree(p);
    for (auto* p : o.a2)
        if (p) (void)hipFree(p);
    g_ops.erase(it);
}

static int get_operands(sfh_ctx* ctx, const sfh_templates* tpl, int mode, BatchedOperands** out)
{
    std::lock_guard<std::mutex> g(g_ops_mu);
    BatchedOperands& o = g_ops[tpl];
    const int64_t N = tpl->nbins, T = tpl->ntemplates;
    const int64_t ld = cdiv(N, 256) * 256;
    if (!o.ex) {
        o.npad = cdiv(N, kBM) * kBM;
        o.tpad1 = std::max<int64_t>(cdiv(T, 16), 2) * 16;   // >= 2 K-blocks: the FP64 composite's loop runs in pairs
        o.tpad2 = cdiv(T, kBM) * kBM;
        SFH_HIP_TRY(ctx, hipMalloc(&o.ex, (size_t)T * sizeof(int32_t)));
        SFH_HIP_TRY(ctx, hipMalloc(&o.ex0, (size_t)T * sizeof(int32_t)));
        SFH_HIP_TRY(ctx, hipMemsetAsync(o.ex0, 0, (size_t)T * sizeof(int32_t), ctx->stream));
        hipLaunchKernelGGL(k_template_exp, dim3((unsigned)T), dim3(256), 0, ctx->stream, tpl->d_models, ld, N, o.ex);
        if (int rc = check_launch(ctx, "k_template_exp")) return rc;
    }
    const int64_t nrb1 = o.npad / 32, nkb1 = o.tpad1 / 16;
    const int ai = mode == BF16X2 ? 1 : 0;
    if (!o.a1[ai]) {
        const size_t bytes = (size_t)nkb1 * nrb1 * 2 * 64 * sizeof(uint4);
        if (hipMalloc(&o.a1[ai], bytes) != hipSuccess) {
            o.a1[ai] = nullptr;
            return set_error(ctx, SFH_ERR_ALLOC, "hipMalloc(%zu) for composite operand failed", bytes);
        }
        const int64_t n = nkb1 * nrb1 * 64;
        const dim3 grid((unsigned)cdiv(n, 256));
        if (ai == 0)
            hipLaunchKernelGGL(k_build_a1<FP32>, grid, dim3(256), 0, ctx->stream, tpl->d_models, ld, N, T, o.ex0, nrb1, nkb1, o.a1[0]);
        else
            hipLaunchKernelGGL(k_build_a1<BF16X2>, grid, dim3(256), 0, ctx->stream, tpl->d_models, ld, N, T, o.ex, nrb1, nkb1, o.a1[1]);
        if (int rc = check_launch(ctx, "k_build_a1")) return rc;
    }
    const int gi = mode == FP32 ? 1 : 0;
    if (!o.a2[gi]) {
        const int64_t nrb2 = o.tpad2 / 32, nkb2 = o.npad / 16;
        const size_t bytes = (size_t)nkb2 * nrb2 * 2 * 64 * sizeof(uint4);
        if (hipMalloc(&o.a2[gi], bytes) != hipSuccess) {
            o.a2[gi] = nullptr;
            return set_error(ctx, SFH_ERR_ALLOC, "hipMalloc(%zu) for gradient operand failed", bytes);
        }
        const int64_t n = nkb2 * nrb2 * 64;
        if (gi)
            hipLaunchKernelGGL(k_build_a2<true>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, ctx->stream, tpl->d_models, ld, N, T, nrb2, nkb2, o.a2[1]);
        else
            hipLaunchKernelGGL(k_build_a2<false>, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, ctx->stream, tpl->d_models, ld, N, T, nrb2, nkb2, o.a2[0]);
        if (int rc = check_launch(ctx, "k_build_a2")) return rc;
    }
    *out = &o;
    return SFH_OK;
}

static Swizzle make_swizzle(int nrt, int nct, int gr, int gc)
{
    Swizzle s;
    s.nrt = nrt;
    s.nct = nct;
    s.gr = std::max(1, std::min(gr, nrt));
    s.gc = std::max(1, std::min(gc, nct));
    s.nsr = (int)cdiv(nrt, s.gr);
    s.nsc = (int)cdiv(nct, s.gc);
    s.xsplit = 0;
    return s;
}

static unsigned swizzle_grid(const Swizzle& s)
{
    const int64_t patches = (int64_t)s.nsr * s.nsc;
    return (unsigned)(cdiv(patches, 8) * 8 * s.gr * s.gc);
}

static int fg_batched_impl(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data, const double* d_X,
                           int64_t B, int precision, double eps, double eps_g, double* d_negLL,
                           double* d_G)
{
    SFH_ARG_CHECK(ctx, tpl && data && d_X && d_negLL, "null pointer");
    SFH_ARG_CHECK(ctx, tpl->nbins == data->nbins, "axes(models,1) == axes(data,1)");
    SFH_ARG_CHECK(ctx, B >= 1, "B must be >= 1");
    SFH_ARG_CHECK(ctx, precision == SFH_PREC_FP32 || precision == SFH_PREC_BF16X2 || precision == SFH_PREC_F64,
                  "unknown precision %d", precision);
    const int mode = precision;
    BatchedOperands* ops = nullptr;
    if (int rc = get_operands(ctx, tpl, mode, &ops)) return rc;
    const int64_t N = tpl->nbins, T = tpl->ntemplates;
    const int64_t Bpad = cdiv(B, kBN) * kBN;
    const int64_t nrb1 = ops->npad / 32, nkb1 = ops->tpad1 / 16, ncb = Bpad / 32, nct = Bpad / kBN;
    const int64_t nkb2 = ops->npad / 16, nrt1 = ops->npad / kBM;

    void *px = nullptr, *pr = nullptr, *pll = nullptr, *pg = nullptr;
    const size_t xbytes = mode == F64 ? (size_t)nkb1 * nct * 16 * kBN * sizeof(double)
                                      : (size_t)nkb1 * ncb * 2 * 64 * sizeof(uint4);
    if (int rc = scratch(ctx, "bat_x", xbytes, &px)) return rc;
    if (int rc = scratch(ctx, "bat_r", (size_t)nkb2 * ncb * 2 * 64 * sizeof(uint4), &pr)) return rc;
    if (int rc = scratch(ctx, "bat_ll", (size_t)nrt1 * Bpad * sizeof(double), &pll)) return rc;

    {
        TimedLaunch tl(ctx, "batched_prep_x");
        if (mode == F64) {
            hipLaunchKernelGGL(k_prep_x64, dim3((unsigned)(nkb1 * nct)), dim3(256), 0, ctx->stream, d_X, T, B, nct,
                               nkb1, (double*)px);
        } else {
            const int64_t n = nkb1 * ncb * 64;
            const dim3 grid((unsigned)cdiv(n, 256));
            if (mode == FP32)
                hipLaunchKernelGGL(k_prep_x<FP32>, grid, dim3(256), 0, ctx->stream, d_X, T, B, ops->ex0, ncb, nkb1, (uint4*)px);
            else
                hipLaunchKernelGGL(k_prep_x<BF16X2>, grid, dim3(256), 0, ctx->stream, d_X, T, B, ops->ex, ncb, nkb1, (uint4*)px);
        }
    }
    if (int rc = check_launch(ctx, "k_prep_x")) return rc;

    // composite GEMM + log-likelihood / R epilogue
    {
        const Swizzle sw = make_swizzle((int)nrt1, (int)nct, 4, 8);
        EpiLL el{data->d_data, N, B, eps, eps_g, eps > eps_g ? eps : eps_g, (double*)pll, (uint4*)pr, ncb, Bpad};
        EpiGrad eg{};
        const dim3 grid(swizzle_grid(sw), 1);
        TimedLaunch tl(ctx, "batched_composite_ll");
        if (mode == F64)
            hipLaunchKernelGGL(k_gemm_f64_ll, grid, dim3(kWG), 0, ctx->stream, ops->a1[0], (const double*)px, nrb1,
                               nct, nkb1, sw, el);
        else if (mode == FP32)
            hipLaunchKernelGGL((k_gemm<FP32, true>), grid, dim3(kWG), 0, ctx->stream, ops->a1[0], (const uint4*)px,
                               nrb1, ncb, nkb1, nkb1, sw, el, eg);
        else
            hipLaunchKernelGGL((k_gemm<BF16X2, true>), grid, dim3(kWG), 0, ctx->stream, ops->a1[1], (const uint4*)px,
                               nrb1, ncb, nkb1, nkb1, sw, el, eg);
    }
    if (int rc = check_launch(ctx, "composite GEMM")) return rc;

    int nsplit = 1;
    if (d_G) {
        // 256 x 128 tiles (4 x 2 waves), one 512-thread workgroup per CU; the
        // bins are split so the workgroups fill whole waves of the 256 CUs
        constexpr int GWR = 4, GWC = 2, GRB = kTR * GWR, GCB = kTC * GWC;
        const int64_t nrb2 = ops->tpad2 / 32;
        const int nrt = (int)cdiv(nrb2, GRB), nct2 = (int)cdiv(ncb, GCB);
        const int64_t tiles = (int64_t)nrt * nct2;
        int best = 1;
        double best_cost = 1e300;
        for (int ns = 1; ns <= 64 && ns <= std::max<int64_t>(1, nkb2 / 32); ++ns) {
            const int64_t wg = tiles * ns;
            const double waves = (double)cdiv(wg, 256);
            // time ~ waves of workgroups x K-blocks per workgroup
            const double cost = waves * (double)cdiv(nkb2, ns);
            if (cost < best_cost * 0.999) {
                best_cost = cost;
                best = ns;
            }
        }
        const int64_t kps = cdiv(nkb2, best);
        nsplit = (int)cdiv(nkb2, kps);
        if (int rc = scratch(ctx, "bat_g", (size_t)nsplit * Bpad * ops->tpad2 * sizeof(float), &pg)) return rc;
        Swizzle sw = make_swizzle(nrt, nct2, nrt, std::max(1, 8 / nrt));
        dim3 grid(swizzle_grid(sw), (unsigned)nsplit);
        if (nsplit % 8 == 0) {
            sw.xsplit = 1;
            grid = dim3((unsigned)(tiles * nsplit), 1);
        }
        EpiLL el{};
        EpiGrad eg{(float*)pg, Bpad, ops->tpad2};
        TimedLaunch tl(ctx, "batched_gradient");
        if (mode == FP32)
            hipLaunchKernelGGL((k_gemm<GRAD_FP32, false, GWR, GWC>), grid, dim3(GWR * GWC * 64), 0, ctx->stream,
                               ops->a2[1], (const uint4*)pr, nrb2, ncb, kps, nkb2, sw, el, eg);
        else
            hipLaunchKernelGGL((k_gemm<GRAD_BF16, false, GWR, GWC>), grid, dim3(GWR * GWC * 64), 0, ctx->stream,
                               ops->a2[0], (const uint4*)pr, nrb2, ncb, kps, nkb2, sw, el, eg);
        if (int rc = check_launch(ctx, "gradient GEMM")) return rc;
    }
    {
        const int64_t nll = cdiv(B, kFinW), ng = d_G ? cdiv(T * B, 256) : 0;
        TimedLaunch tl(ctx, "batched_finish");
        hipLaunchKernelGGL(k_finish, dim3((unsigned)(nll + ng)), dim3(256), 0, ctx->stream, (const double*)pll, nrt1,
                           (const float*)pg, nsplit, T, B, Bpad, ops->tpad2, nll, d_negLL, d_G);
    }
    return check_launch(ctx, "k_finish");
}

}  // namespace sfh

using namespace sfh;

extern "C" {

int sfh_fg_batched_device(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data, const double* d_X,
                          int64_t B, int precision, double eps, double eps_grad, double* d_negLL, double* d_G)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    return fg_batched_impl(ctx, tpl, data, d_X, B, precision, eps, eps_grad, d_negLL, d_G);
}

int sfh_fg_batched(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data, const double* X, int64_t B,
                   int precision, double eps, double eps_grad, double* negLL, double* G)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_ARG_CHECK(ctx, tpl && X && negLL, "null pointer");
    SFH_ARG_CHECK(ctx, B >= 1, "B must be >= 1");
    const int64_t T = tpl->ntemplates;
    void *px = nullptr, *pn = nullptr, *pg = nullptr;
    if (int rc = scratch(ctx, "batx_in", (size_t)T * B * sizeof(double), &px)) return rc;
    if (int rc = scratch(ctx, "batx_nll", (size_t)B * sizeof(double), &pn)) return rc;
    if (G)
        if (int rc = scratch(ctx, "batx_g", (size_t)T * B * sizeof(double), &pg)) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(px, X, (size_t)T * B * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (int rc = fg_batched_impl(ctx, tpl, data, (const double*)px, B, precision, eps, eps_grad, (double*)pn, (double*)pg))
        return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(negLL, pn, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (G)
        SFH_HIP_TRY(ctx, hipMemcpyAsync(G, pg, (size_t)T * B * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

}  // extern "C"

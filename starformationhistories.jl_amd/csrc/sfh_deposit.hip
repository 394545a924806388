// sfh_deposit.hip -- smoothed Hess-diagram template deposition on gfx950:
// bin_cmd_smooth (src/StarFormationHistories.jl:574-621), one template or many
// templates per launch.
//
// Two kernels per call:
//   k_dep_prep    one thread per point: the reference's FP64 footprint / index
//                 math (bit-exact, no FMA contraction) and every per-point
//                 constant of the evaluation -> an 80-byte PointDesc.
//   k_dep_gather  bin-owner gather: a thread owns one row and kCPT = 16
//                 consecutive columns of one template, in registers; a
//                 workgroup 256 such strips.  Every wave walks its template's
//                 descriptors in order through scalar loads (the descriptor is
//                 uniform across the wave, so it sits in scalar registers: no
//                 LDS staging, no barriers), skips points whose rows miss its
//                 strips, and adds each point's contribution to its bins, so
//                 there are no atomics and every bin sums its points in point
//                 order (deterministic).  Contributions accumulate in FP32 per
//                 64 points and are flushed into FP64 accumulators.
// The pair work is VALU bound (about 3 + DEG VALU operations and one exp2 per
// GL node and bin for the series forms); HBM traffic is the descriptors in and
// one write of each template out.
#include "sfh_internal.h"
#include "sfh_psf.h"

namespace sfh {

using psf::PointDesc;

constexpr int kDepBlock = 256;
constexpr int kCPT = psf::kCPT;  // columns per thread
constexpr int kChunk = 64;       // points between FP32 -> FP64 flushes

__global__ __launch_bounds__(256) void k_dep_prep(
    const double* __restrict__ colors, const double* __restrict__ mags,
    const double* __restrict__ cerr, const double* __restrict__ merr,
    const double* __restrict__ weights, const double* __restrict__ cmult, double cov_mult,
    int64_t npoints, psf::Range ex, psf::Range ey, int asym, PointDesc* __restrict__ out,
    unsigned long long* __restrict__ pairs, int* __restrict__ bad)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cnt = 0;
    if (p < npoints) {
        const double x = colors[p], y = mags[p], sx = cerr[p], sy = merr[p], A = weights[p];
        const double c = cmult ? cmult[p] : cov_mult;
        PointDesc d;
        if (!(isfinite(x) && isfinite(y) && isfinite(sx) && isfinite(sy) && isfinite(A) && isfinite(c))) {
            d = psf::empty_desc();
            atomicOr(bad, 1);
        } else if (asym) {
            cnt = (unsigned long long)psf::prep_asymmetric(x, y, sx, sy, A, ex, ey, &d);
        } else {
            cnt = (unsigned long long)psf::prep_covariant(x, y, sx, sy, c, A, ex, ey, &d);
        }
        out[p] = d;
    }
    // wave-level sum of pair counts, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(pairs, cnt);
}

// The FP64 accumulators live in LDS (16 per thread, 32 KiB per workgroup),
// touched once per 64 points, so that the FP32 strip and the per-point
// constants fit four waves per SIMD (a fifth, at 95 registers with the
// series' coefficient products formed per column, measured slower).
// Workgroup w of the gather takes (template, row band) order[w] = t * nblk + band
// when an order is given: the (template, band) pairs by decreasing cost, so
// the longest workgroups start first and the launch ends with short ones.
// Without one, template t's band j is (j + t) mod nblk: consecutive workgroups
// go to the 8 XCDs in turn, so without the rotation every template's band j
// would land on XCD j mod 8 and the populated bands of a diagram would load
// some XCDs far more than others.
template <bool kAsym>
__global__ __launch_bounds__(kDepBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_dep_gather(
    const PointDesc* __restrict__ desc, const int64_t* __restrict__ offsets, const int32_t* __restrict__ order,
    int nx, int ny, int ncg, int nblk_per_tpl, float* __restrict__ out32, double* __restrict__ out64,
    int64_t ld_out)
{
    __shared__ double sacc[kCPT][kDepBlock];
    int64_t t;
    int blk;
    if (order) {
        const int id = order[blockIdx.x];
        t = id / nblk_per_tpl;
        blk = id - (int)t * nblk_per_tpl;
    } else {
        t = blockIdx.x / nblk_per_tpl;
        blk = (int)((blockIdx.x - t * nblk_per_tpl + t) % nblk_per_tpl);
    }
    const int64_t p0 = offsets[t], p1 = offsets[t + 1];
    const int slot = blk * kDepBlock + threadIdx.x;
    const int row = slot / ncg;
    const int col0 = (slot - row * ncg) * kCPT;
    const bool valid = row < ny;
    // rows of this wave's 64 strips (wave-uniform)
    const int wslot = __builtin_amdgcn_readfirstlane(blk * kDepBlock + (int)(threadIdx.x & ~63u));
    const int row_lo = wslot / ncg, row_hi = (wslot + 63) / ncg;

    // volatile: the accumulators stay in LDS instead of being promoted to registers
    volatile double* acc = &sacc[0][threadIdx.x];
    float part[kCPT];
#pragma unroll
    for (int c = 0; c < kCPT; ++c) {
        acc[c * kDepBlock] = 0.0;
        part[c] = 0.0f;
    }
    // Each chunk of 64 points: every lane reads one point's footprint (the
    // descriptor's first 16 bytes) and form, a ballot keeps the points whose
    // rows meet this wave's strips, and only those descriptors are loaded --
    // through scalar loads (uniform across the wave), two descriptors per
    // wait.  The hits are taken in two passes: first the common case (the
    // Gaussian form with a footprint spanning the whole grid width), then
    // every other point.  Each pass has one code path, so the strip's registers
    // carry through it without copies; a bin sums the chunk's common-case
    // points, then the others, each in point order (deterministic).
    const PointDesc* __restrict__ td = desc + p0;
    const int np = (int)(p1 - p0);
    const int lane = threadIdx.x & 63;
    const int wave_id = (int)(threadIdx.x >> 6);
    struct Foot {
        int4 f;
        int kind;
    };
    auto footprint = [&](int pc) {
        const int p = pc + lane;
        Foot r;
        if (p < np) {
            r.f = *reinterpret_cast<const int4*>(td + p);
            r.kind = td[p].kind;
        } else {
            r.f = make_int4(1, 0, 1, 0);
            r.kind = psf::kDirect;
        }
        return r;
    };
    auto in_rows = [&](const PointDesc& d) { return valid && row >= d.ya && row <= d.yb; };
    auto add_common = [&](const PointDesc& d) {
        if (!in_rows(d)) return;
        if constexpr (kAsym) {
            float M[1], L[1] = {0.0f}, h[1] = {0.0f};
            M[0] = d.pref * psf::y_factor(fmaf((float)(row - d.ya), d.dv, d.ay), d.ey);
            psf::gauss_strip<1, false>(d, M, L, h, col0, part);
        } else {
            float M[3], L[3], h[3];
            psf::cov_row_log(d, row, M, L, h);
            psf::gauss_strip<3, false>(d, M, L, h, col0, part);
        }
    };
    auto add = [&](const PointDesc& d) {
        if (!in_rows(d) || col0 > d.xb || col0 + kCPT - 1 < d.xa) return;
        if constexpr (kAsym) {
            float W[1], h[1] = {0.0f};
            W[0] = d.pref * psf::y_factor(fmaf((float)(row - d.ya), d.dv, d.ay), d.ey);
            psf::add_point<1, true>(d, W, h, col0, part);
        } else {
            float W[3], h[3];
            psf::cov_row(d, row, W, h);
            psf::add_point<3, true>(d, W, h, col0, part);
        }
    };
    // A point in the erf-difference form (a narrow kernel: its footprint is at
    // most 55 columns wide) with the whole wave cooperating, instead of the
    // strip code in which only the few lanes whose strips meet the footprint
    // would work: lane l takes footprint column xa + l, the point's rows inside
    // this wave's strips one after another; lanes 0 .. nc evaluate erfc at
    // their column's left edge and a column's right edge is the next lane's.
    // Each bin's value (nodes summed in order, in FP32) is added to its
    // owner's FP64 accumulator in LDS: the bins are distinct, so no two lanes
    // write one word, and the owner's own flush comes later in program order.
    auto direct = [&](const PointDesc& d) {
        constexpr int NODES = kAsym ? 1 : 3;
        const int nc = d.xb - d.xa + 1;
        const int r0 = d.ya > row_lo ? d.ya : row_lo, r1 = d.yb < row_hi ? d.yb : row_hi;
        for (int r = r0; r <= r1; ++r) {
            float W[3], h[3] = {0.0f, 0.0f, 0.0f};
            if constexpr (kAsym) W[0] = d.pref * psf::y_factor(fmaf((float)(r - d.ya), d.dv, d.ay), d.ey);
            else psf::cov_row(d, r, W, h);
            float v = 0.0f;
#pragma unroll
            for (int k = 0; k < NODES; ++k) {
                const float b = psf::direct_edge(d, lane, h[k]);
                const float e = psf::erfc_pos(fabsf(b));
                const int n = b < 0.0f ? 1 : 0;
                const float e1 = __shfl_down(e, 1, 64);
                const int n1 = __shfl_down(n, 1, 64);
                v = fmaf(W[k], psf::erf_diff(e, n != 0, e1, n1 != 0), v);
            }
            if (lane < nc) {
                const int col = d.xa + lane;
                const int o = r * ncg + col / kCPT - wslot;
                if (o >= 0 && o < 64) sacc[col % kCPT][wave_id * 64 + o] += (double)v;
            }
        }
    };
    Foot fn = footprint(0);
    for (int pc = 0; pc < np; pc += kChunk) {
        const Foot f = fn;
        fn = footprint(pc + kChunk);                           // next chunk's footprints in flight
        const bool hit = f.f.x <= f.f.y && f.f.z <= row_hi && f.f.w >= row_lo;
        const bool common = f.kind == psf::kGauss && f.f.x == 0 && f.f.y == nx - 1;
        uint64_t hits = __ballot(hit && common);
        while (hits) {
            const int i = __builtin_ctzll(hits);
            hits &= hits - 1;
            const bool two = hits != 0;
            const int j = two ? __builtin_ctzll(hits) : i;
            if (two) hits &= hits - 1;
            const PointDesc d0 = td[pc + i];
            const PointDesc d1 = td[pc + j];
            add_common(d0);
            if (two) add_common(d1);
        }
        hits = __ballot(hit && !common);
        while (hits) {
            const int i = __builtin_ctzll(hits);
            hits &= hits - 1;
            const bool two = hits != 0;
            const int j = two ? __builtin_ctzll(hits) : i;
            if (two) hits &= hits - 1;
            const PointDesc d0 = td[pc + i];
            const PointDesc d1 = td[pc + j];
            // the narrow erf-difference kernels go to the wave-cooperative path
            if (d0.kind == psf::kDirect && d0.xb - d0.xa < 63) direct(d0);
            else add(d0);
            if (two) {
                if (d1.kind == psf::kDirect && d1.xb - d1.xa < 63) direct(d1);
                else add(d1);
            }
        }
#pragma unroll
        for (int c = 0; c < kCPT; ++c) {
            acc[c * kDepBlock] += (double)part[c];
            part[c] = 0.0f;
        }
    }
    if (!valid) return;
    const int64_t base = t * ld_out + (int64_t)row * nx;
#pragma unroll
    for (int c = 0; c < kCPT; ++c) {
        const int col = col0 + c;
        if (col < nx) {
            const double v = acc[c * kDepBlock];
            if (out32) out32[base + col] = (float)v;
            if (out64) out64[base + col] = v;
        }
    }
}

// Relative cost per row strip of each x form (issue slots of the form's strip
// code over the Gaussian form's).
__device__ inline unsigned form_cost(int kind, bool full)
{
    if (kind == psf::kGauss) return full ? 4u : 6u;
    if (kind == psf::kSeries2 || kind == psf::kSeries3) return 12u;
    return 32u;
}

// Estimated cost of every (template, row band) workgroup of the gather: per
// point, the rows x strips of its footprint inside each band it meets, times
// its form's cost.  One workgroup per template, the band sums in LDS (integer
// atomics: the sums do not depend on the order); nblk * 8 bytes of dynamic LDS.
__global__ __launch_bounds__(256) void k_dep_cost(const PointDesc* __restrict__ desc,
                                                  const int64_t* __restrict__ offsets, int nx, int ncg, int nblk,
                                                  unsigned long long* __restrict__ cost)
{
    extern __shared__ unsigned long long band[];
    const int64_t t = blockIdx.x;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) band[b] = 0;
    __syncthreads();
    for (int64_t p = offsets[t] + threadIdx.x; p < offsets[t + 1]; p += blockDim.x) {
        const int4 f = *reinterpret_cast<const int4*>(desc + p);
        if (f.x > f.y) continue;
        const bool full = f.x == 0 && f.y == nx - 1;
        const unsigned w = form_cost(desc[p].kind, full);
        const int nstrips = f.y / kCPT - f.x / kCPT + 1;
        const int b0 = (int)(((int64_t)f.z * ncg) / kDepBlock);
        int b1 = (int)(((int64_t)f.w * ncg + ncg - 1) / kDepBlock);
        if (b1 > nblk - 1) b1 = nblk - 1;
        for (int b = b0; b <= b1; ++b) {
            const int rlo = (int)(((int64_t)b * kDepBlock) / ncg);
            const int rhi = (int)(((int64_t)b * kDepBlock + kDepBlock - 1) / ncg);
            const int r0 = f.z > rlo ? f.z : rlo, r1 = f.w < rhi ? f.w : rhi;
            if (r1 >= r0) atomicAdd(&band[b], (unsigned long long)(r1 - r0 + 1) * nstrips * w);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) cost[t * nblk + b] = band[b];
}

constexpr int kOrderMax = 16384;    // workgroups one ordering launch takes
constexpr int kOrderBuckets = 1024;

// order[] = the n workgroup ids by decreasing cost, to 1/1024 of the largest:
// a counting sort over 1024 cost buckets in one workgroup (the order inside
// a bucket is that of the LDS atomics; it only schedules the gather, whose
// workgroups each write their own band, so the templates do not depend on it).
__global__ __launch_bounds__(1024) void k_dep_order(const unsigned long long* __restrict__ cost, int n,
                                                    int32_t* __restrict__ order)
{
    __shared__ unsigned long long smax[16];
    __shared__ int start[kOrderBuckets];
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    unsigned long long mx = 0;
    for (int i = tid; i < n; i += blockDim.x) mx = cost[i] > mx ? cost[i] : mx;
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if (lane == 0) smax[wv] = mx;
    start[tid] = 0;
    __syncthreads();
    mx = 0;
    for (int w = 0; w < 16; ++w) mx = smax[w] > mx ? smax[w] : mx;
    int shift = 0;
    while ((mx >> shift) >= (unsigned long long)kOrderBuckets) ++shift;
    auto bucket = [&](int i) { return kOrderBuckets - 1 - (int)(cost[i] >> shift); };
    for (int i = tid; i < n; i += blockDim.x) atomicAdd(&start[bucket(i)], 1);
    __syncthreads();
    // exclusive prefix sum of the 1024 counts, one per thread
    const int c = start[tid];
    int incl = c;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wv; ++w) base += wsum[w];
    start[tid] = base + incl - c;
    __syncthreads();
    for (int i = tid; i < n; i += blockDim.x) order[atomicAdd(&start[bucket(i)], 1)] = i;
}

// The gather over descriptors of npoints points (template t owns
// desc[offsets[t] .. offsets[t+1])), with the workgroups in decreasing
// estimated cost when the launch is small enough to order in one pass.
static int launch_gather(sfh_ctx* ctx, const PointDesc* desc, const int64_t* d_offsets, int64_t ntemplates,
                         int64_t npoints, bool asym, int nx, int ny, float* out32, double* out64, int64_t ld_out)
{
    const int ncg = (int)cdiv(nx, kCPT);
    const int nblk = (int)cdiv((int64_t)ncg * ny, kDepBlock);
    const int64_t grid = (int64_t)nblk * ntemplates;
    SFH_ARG_CHECK(ctx, grid < (int64_t)1 << 31, "too many templates for one launch");
    if (grid == 0) return SFH_OK;
    const int32_t* order = nullptr;
    if (ntemplates > 1 && grid <= kOrderMax && npoints > 0) {
        void *pc = nullptr, *po = nullptr;
        if (int rc = scratch(ctx, "dep_cost", (size_t)grid * sizeof(unsigned long long), &pc)) return rc;
        if (int rc = scratch(ctx, "dep_order", (size_t)grid * sizeof(int32_t), &po)) return rc;
        {
            TimedLaunch tl(ctx, "deposit_order");
            hipLaunchKernelGGL(k_dep_cost, dim3((unsigned)ntemplates), dim3(256),
                               (size_t)nblk * sizeof(unsigned long long), ctx->stream, desc, d_offsets, nx, ncg, nblk,
                               (unsigned long long*)pc);
            hipLaunchKernelGGL(k_dep_order, dim3(1), dim3(1024), 0, ctx->stream, (const unsigned long long*)pc,
                               (int)grid, (int32_t*)po);
        }
        if (int rc = check_launch(ctx, "k_dep_order")) return rc;
        order = (const int32_t*)po;
    }
    if (asym) {
        TimedLaunch tl(ctx, "deposit_asym");
        hipLaunchKernelGGL(k_dep_gather<true>, dim3((unsigned)grid), dim3(kDepBlock), 0, ctx->stream, desc, d_offsets,
                           order, nx, ny, ncg, nblk, out32, out64, ld_out);
    } else {
        TimedLaunch tl(ctx, "deposit_cov");
        hipLaunchKernelGGL(k_dep_gather<false>, dim3((unsigned)grid), dim3(kDepBlock), 0, ctx->stream, desc,
                           d_offsets, order, nx, ny, ncg, nblk, out32, out64, ld_out);
    }
    return check_launch(ctx, "k_dep_gather");
}

static int validate_ranges(sfh_ctx* ctx, sfh_range xe, sfh_range ye)
{
    SFH_ARG_CHECK(ctx, xe.length >= 2 && ye.length >= 2, "edge ranges need at least 2 edges");
    SFH_ARG_CHECK(ctx, (xe.length - 1) * (ye.length - 1) > 1, "length(image.weights) > 1 (:367)");
    SFH_ARG_CHECK(ctx, std::isfinite(xe.step) && std::isfinite(ye.step) && xe.step != 0.0 && ye.step != 0.0,
                  "edge steps must be finite and non-zero");
    SFH_ARG_CHECK(ctx, xe.length < (1 << 30) && ye.length < (1 << 30), "grid too large");
    return SFH_OK;
}

// Core launcher on device arrays.  out32 / out64: column t at base + t*ld_out.
static int deposit(sfh_ctx* ctx, const double* colors, const double* mags, const double* cerr,
                   const double* merr, const double* weights, const double* cmult,
                   const int64_t* d_offsets, int64_t ntemplates, int64_t npoints_total,
                   double cov_mult, sfh_range xe, sfh_range ye, float* out32, double* out64,
                   int64_t ld_out, int64_t* pairs_out)
{
    if (int rc = validate_ranges(ctx, xe, ye)) return rc;
    const int nx = (int)(xe.length - 1), ny = (int)(ye.length - 1);
    SFH_ARG_CHECK(ctx, ld_out >= (int64_t)nx * ny, "ld_out < nx*ny");
    const bool asym = (cov_mult == 0.0) && cmult == nullptr;
    void *pdesc = nullptr, *pcnt = nullptr;
    if (int rc = scratch(ctx, "dep_desc", (size_t)(npoints_total + 1) * sizeof(PointDesc), &pdesc))
        return rc;
    if (int rc = scratch(ctx, "dep_counters", 64, &pcnt)) return rc;
    SFH_HIP_TRY(ctx, hipMemsetAsync(pcnt, 0, 64, ctx->stream));
    unsigned long long* d_pairs = (unsigned long long*)pcnt;
    int* d_bad = (int*)((char*)pcnt + 16);
    psf::Range rx{xe.first, xe.step, xe.length}, ry{ye.first, ye.step, ye.length};
    if (npoints_total > 0) {
        TimedLaunch tl(ctx, "deposit_prep");
        hipLaunchKernelGGL(k_dep_prep, dim3((unsigned)cdiv(npoints_total, 256)), dim3(256), 0, ctx->stream,
                           colors, mags, cerr, merr, weights, cmult, cov_mult, npoints_total, rx, ry,
                           asym ? 1 : 0, (PointDesc*)pdesc, d_pairs, d_bad);
    }
    if (int rc = check_launch(ctx, "k_dep_prep")) return rc;
    if (int rc = launch_gather(ctx, (const PointDesc*)pdesc, d_offsets, ntemplates, npoints_total, asym, nx, ny,
                               out32, out64, ld_out))
        return rc;
    // Non-finite inputs: the reference's round(Int, NaN) throws InexactError.
    int bad = 0;
    unsigned long long pairs = 0;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(&pairs, d_pairs, sizeof(pairs), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (bad) return set_error(ctx, SFH_ERR_DOMAIN, "non-finite kernel parameter (InexactError in round(Int, ·))");
    if (pairs_out) *pairs_out = (int64_t)pairs;
    return SFH_OK;
}

// The gather of descriptors already built (by the device preparation of
// partial_cmd_smooth): template t owns desc[koff[t] .. koff[t+1]).
int deposit_from_desc(sfh_ctx* ctx, const void* desc, const int64_t* d_koff, int64_t ntemplates, int64_t npoints,
                      bool asym, sfh_range xe, sfh_range ye, float* out32, int64_t ld_out)
{
    if (int rc = validate_ranges(ctx, xe, ye)) return rc;
    const int nx = (int)(xe.length - 1), ny = (int)(ye.length - 1);
    SFH_ARG_CHECK(ctx, ld_out >= (int64_t)nx * ny, "ld_out < nx*ny");
    return launch_gather(ctx, (const PointDesc*)desc, d_koff, ntemplates, npoints, asym, nx, ny, out32, nullptr,
                         ld_out);
}

}  // namespace sfh

using namespace sfh;

extern "C" {

int sfh_bin_cmd_smooth(sfh_ctx* ctx, const double* colors, const double* mags,
                       const double* color_err, const double* mag_err, const double* weights,
                       int64_t npoints, int cov_mult, sfh_range xedges, sfh_range yedges,
                       double* out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_ARG_CHECK(ctx, cov_mult == -1 || cov_mult == 0 || cov_mult == 1, "cov_mult in (-1, 0, 1) (:579)");
    SFH_ARG_CHECK(ctx, npoints >= 0, "npoints must be >= 0");
    SFH_ARG_CHECK(ctx, out && (npoints == 0 || (colors && mags && color_err && mag_err && weights)),
                  "null pointer");
    if (int rc = validate_ranges(ctx, xedges, yedges)) return rc;
    const int64_t nbins = (xedges.length - 1) * (yedges.length - 1);
    const int64_t np = npoints > 0 ? npoints : 1;
    void* p = nullptr;
    if (int rc = scratch(ctx, "bcs_in", (size_t)np * 5 * sizeof(double) + 2 * sizeof(int64_t), &p)) return rc;
    double* d = (double*)p;
    const double* srcs[5] = {colors, mags, color_err, mag_err, weights};
    for (int k = 0; k < 5 && npoints > 0; ++k)
        SFH_HIP_TRY(ctx, hipMemcpyAsync(d + k * np, srcs[k], (size_t)npoints * sizeof(double),
                                        hipMemcpyHostToDevice, ctx->stream));
    int64_t* doff = (int64_t*)(d + 5 * np);
    int64_t hoff[2] = {0, npoints};
    SFH_HIP_TRY(ctx, hipMemcpyAsync(doff, hoff, sizeof(hoff), hipMemcpyHostToDevice, ctx->stream));
    void* po = nullptr;
    if (int rc = scratch(ctx, "bcs_out", (size_t)nbins * sizeof(double), &po)) return rc;
    int rc = deposit(ctx, d, d + np, d + 2 * np, d + 3 * np, d + 4 * np, nullptr, doff, 1, npoints,
                     (double)cov_mult, xedges, yedges, nullptr, (double*)po, nbins, nullptr);
    if (rc) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(out, po, (size_t)nbins * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

int sfh_bin_cmd_smooth_batch_device(sfh_ctx* ctx, const double* d_colors, const double* d_mags,
                                    const double* d_color_err, const double* d_mag_err,
                                    const double* d_weights, const double* d_cmult,
                                    const int64_t* d_offsets, int64_t ntemplates, double cov_mult,
                                    sfh_range xedges, sfh_range yedges, float* d_out, int64_t ld_out,
                                    int64_t* pairs_out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_ARG_CHECK(ctx, ntemplates >= 1, "ntemplates must be >= 1");
    SFH_ARG_CHECK(ctx, d_offsets && d_out, "null pointer");
    SFH_ARG_CHECK(ctx, std::isfinite(cov_mult), "cov_mult must be finite");
    // The gather kernel trusts the offsets to index the descriptor array, so
    // check them all on the host first (ntemplates+1 integers).
    std::vector<int64_t> hoff((size_t)ntemplates + 1);
    SFH_HIP_TRY(ctx, hipMemcpyAsync(hoff.data(), d_offsets, hoff.size() * sizeof(int64_t),
                                    hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    SFH_ARG_CHECK(ctx, hoff[0] == 0, "offsets[0] must be 0");
    for (int64_t t = 0; t < ntemplates; ++t)
        SFH_ARG_CHECK(ctx, hoff[t + 1] >= hoff[t], "offsets must be non-decreasing (t=%lld)", (long long)t);
    const int64_t last = hoff[ntemplates];
    SFH_ARG_CHECK(ctx, last == 0 || (d_colors && d_mags && d_color_err && d_mag_err && d_weights),
                  "null point array");
    return deposit(ctx, d_colors, d_mags, d_color_err, d_mag_err, d_weights, d_cmult, d_offsets,
                   ntemplates, last, cov_mult, xedges, yedges, d_out, nullptr, ld_out, pairs_out);
}

}  // extern "C"

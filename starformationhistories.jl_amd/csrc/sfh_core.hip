// sfh_core.hip -- context / handle management and the single-coefficient-vector
// likelihood path (composite!, loglikelihood, ∇loglikelihood(!), fg!) of
// StarFormationHistories.jl (src/fitting/fitting_base.jl, src/fitting/solvers.jl)
// for gfx950.
//
// Data layout in HBM
//   models  : FP32, ld x T column-major with ld = nbins rounded up to 256 and
//             the padding rows zero, so every template column starts 1 KiB
//             aligned and float4 loads never straddle a column.
//   data    : FP64 n_i, length nbins.
//   partial : FP64 composite partial sums, nsplit x ld (template split-K).
//
// fg! and loglikelihood from coefficients with T <= 704 run in one pass over
// the templates (k_single_fg + k_single_reduce, sfh_multifit.hip).  The
// kernels here serve composite!, ∇loglikelihood(!) from a given composite and
// fg! of larger template sets (4 launches, each an HBM stream or a tiny
// reduction):
//   k_composite_partial : m_s[i] = sum_{j in split s} C[i,j] r_j      (streams C once)
//   k_composite_finish  : m_i = sum_s m_s[i]; LL term; R_i = 1-n_i/max(m_i,eps)
//   k_grad_partial      : g[j,chunk] = sum_{i in chunk} C[i,j] R_i      (streams C once)
//   k_fg_final          : LL = sum of block partials (fixed order); G = sign * sum_chunk g
// All reductions are in FP64 and in a fixed order (bitwise reproducible).
#include "sfh_internal.h"

#include <cmath>
#include <cstdarg>
#include <cstring>

namespace sfh {

int set_error(sfh_ctx* ctx, int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx) ctx->last_error = buf;
    return code;
}

int bind(sfh_ctx* ctx)
{
    SFH_HIP_TRY(ctx, hipSetDevice(ctx->device));
    return SFH_OK;
}

int scratch(sfh_ctx* ctx, const char* name, size_t bytes, void** out)
{
    DevBuf& b = ctx->scratch[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            SFH_HIP_TRY(ctx, hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes < 256 ? 256 : bytes;
        if (hipMalloc(&b.ptr, want) != hipSuccess) {
            b.ptr = nullptr;
            return set_error(ctx, SFH_ERR_ALLOC, "hipMalloc(%zu) for scratch '%s' failed", want, name);
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return SFH_OK;
}

int pinned(sfh_ctx* ctx, size_t bytes, void** out)
{
    if (ctx->pinned_bytes < bytes) {
        if (ctx->pinned) {
            SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            SFH_HIP_TRY(ctx, hipHostFree(ctx->pinned));
            ctx->pinned = nullptr;
            ctx->pinned_bytes = 0;
        }
        if (hipHostMalloc(&ctx->pinned, bytes, hipHostMallocDefault) != hipSuccess) {
            ctx->pinned = nullptr;
            return set_error(ctx, SFH_ERR_ALLOC, "hipHostMalloc(%zu) failed", bytes);
        }
        ctx->pinned_bytes = bytes;
    }
    *out = ctx->pinned;
    return SFH_OK;
}

static hipEvent_t take_event(sfh_ctx* ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timing_begin(sfh_ctx* ctx, const char* name, hipEvent_t* start_out)
{
    (void)name;
    *start_out = nullptr;
    if (!ctx->timing) return;
    hipEvent_t e = take_event(ctx);
    if (e && hipEventRecord(e, ctx->stream) == hipSuccess) *start_out = e;
}

void timing_end(sfh_ctx* ctx, const char* name, hipEvent_t start)
{
    if (!ctx->timing || !start) return;
    hipEvent_t e = take_event(ctx);
    if (!e) return;
    if (hipEventRecord(e, ctx->stream) != hipSuccess) return;
    ctx->pending.push_back(PendingTiming{name, start, e});
    if (ctx->pending.size() > 4096) resolve_timing(ctx);
}

void resolve_timing(sfh_ctx* ctx)
{
    for (auto& p : ctx->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.stop) == hipSuccess &&
            hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            KernelStat& s = ctx->stats[p.name];
            s.total_ms += ms;
            s.launches += 1;
        }
        ctx->event_pool.push_back(p.start);
        ctx->event_pool.push_back(p.stop);
    }
    ctx->pending.clear();
}

int check_launch(sfh_ctx* ctx, const char* name)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(ctx, SFH_ERR_DEVICE, "launch of %s failed: %s", name, hipGetErrorString(e));
    return SFH_OK;
}

// ------------------------------------------------------------------ kernels

constexpr int kBlock = 256;
constexpr int kCompBins = 1024;   // bins per composite workgroup (4 per thread)
constexpr int kGradChunk = 4096;  // bins per gradient workgroup
constexpr int kGradTplPerWave = 4;

static inline int64_t ld_of(int64_t nbins) { return cdiv(nbins, 256) * 256; }

// m_s[i] = sum_{j in [j0,j1)} C[i + ld*j] * r_j in FP64.  Each thread owns 4
// consecutive bins (one float4 per template); the template loop is unrolled
// so each lane keeps 8 independent 16-byte loads in flight.
__global__ __launch_bounds__(kBlock) void k_composite_partial(
    const float* __restrict__ C, const double* __restrict__ r, int64_t ld, int64_t T,
    int64_t tsplit, double* __restrict__ part)
{
    const int64_t bin0 = ((int64_t)blockIdx.x * kCompBins) + (int64_t)threadIdx.x * 4;
    const int64_t s = blockIdx.y;
    const int64_t j0 = s * tsplit;
    int64_t j1 = j0 + tsplit;
    if (j1 > T) j1 = T;
    if (bin0 >= ld) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const float* p = C + bin0;
    int64_t j = j0;
    for (; j + 8 <= j1; j += 8) {
        float4 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const float4*>(p + (j + u) * ld);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double c = r[j + u];
            a0 = fma((double)v[u].x, c, a0);
            a1 = fma((double)v[u].y, c, a1);
            a2 = fma((double)v[u].z, c, a2);
            a3 = fma((double)v[u].w, c, a3);
        }
    }
    for (; j < j1; ++j) {
        const float4 v = *reinterpret_cast<const float4*>(p + j * ld);
        const double c = r[j];
        a0 = fma((double)v.x, c, a0);
        a1 = fma((double)v.y, c, a1);
        a2 = fma((double)v.z, c, a2);
        a3 = fma((double)v.w, c, a3);
    }
    double* o = part + s * ld + bin0;
    reinterpret_cast<double2*>(o)[0] = make_double2(a0, a1);
    reinterpret_cast<double2*>(o)[1] = make_double2(a2, a3);
}

// Fixed-order block sum of one double per thread; result valid in thread 0.
__device__ inline double block_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    return sh[0];
}

// Per bin: m = sum over splits (fixed order).  Then, for bins < nbins:
//   LL term (fitting_base.jl:90-92) accumulated into one FP64 partial per block;
//   R = 1 - n/max(m,eps_g) (fitting_base.jl:277-279) written to R (zero for
//   the padding rows so the gradient pass reads finite values); eps clamps the
//   LL with eps of the promoted type (:86), eps_g the gradient with
//   eps(eltype(composite)) (:277);
//   composite_out gets m (or R when clobber_with_R, the ∇loglikelihood! contract).
__global__ __launch_bounds__(kBlock) void k_composite_finish(
    const double* __restrict__ part, int64_t ld, int nsplit, int64_t nbins,
    const double* __restrict__ data, double eps, double eps_g, int want_ll, double* __restrict__ ll_part,
    double* __restrict__ R, double* __restrict__ composite_out, int clobber_with_R)
{
    __shared__ double sh[kBlock];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double term = 0.0;
    if (i < ld) {
        double m = 0.0;
        for (int s = 0; s < nsplit; ++s) m += part[(int64_t)s * ld + i];
        if (i < nbins) {
            const double c = m > eps ? m : eps;
            const double cg = m > eps_g ? m : eps_g;
            const double n = data ? data[i] : 0.0;
            if (want_ll) term = (n > 0.0) ? (n - c - n * log(n / c)) : -c;
            const double rr = 1.0 - n / cg;
            if (R) R[i] = rr;
            if (composite_out) composite_out[i] = clobber_with_R ? rr : m;
        } else if (R) {
            R[i] = 0.0;
        }
    }
    if (want_ll) {
        const double s = block_sum(term, sh);
        if (threadIdx.x == 0) ll_part[blockIdx.x] = s;
    }
}

// g[j*nch + chunk] = sum_{i in chunk} C[i,j] R_i.  The R chunk sits in LDS
// (FP64, 32 KiB); each wave streams kGradTplPerWave template columns with
// 16-byte loads and reduces each dot product across its 64 lanes.
__global__ __launch_bounds__(kBlock) void k_grad_partial(
    const float* __restrict__ C, const double* __restrict__ R, int64_t ld, int64_t T,
    double* __restrict__ gpart, int64_t nch)
{
    __shared__ double sR[kGradChunk];
    const int64_t chunk = blockIdx.x;
    const int64_t base = chunk * kGradChunk;
    int64_t len = ld - base;
    if (len > kGradChunk) len = kGradChunk;
    for (int k = threadIdx.x; k < kGradChunk; k += kBlock) sR[k] = (k < len) ? R[base + k] : 0.0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t jbase = ((int64_t)blockIdx.y * 4 + wave) * kGradTplPerWave;
    for (int t = 0; t < kGradTplPerWave; ++t) {
        const int64_t j = jbase + t;
        if (j >= T) break;
        const float* col = C + j * ld + base;
        double acc = 0.0;
#pragma unroll 4
        for (int v = 0; v < kGradChunk / 256; ++v) {
            const int k = (v * 64 + lane) * 4;
            if (k < len) {
                const float4 c = *reinterpret_cast<const float4*>(col + k);
                acc = fma((double)c.x, sR[k + 0], acc);
                acc = fma((double)c.y, sR[k + 1], acc);
                acc = fma((double)c.z, sR[k + 2], acc);
                acc = fma((double)c.w, sR[k + 3], acc);
            }
        }
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) gpart[j * nch + chunk] = acc;
    }
}

// LL = sum of block partials in a fixed order (-Inf if exactly 0,
// fitting_base.jl:95); negLL = -LL.  G_j = gsign * sum_chunk g[j, chunk].
__global__ __launch_bounds__(kBlock) void k_fg_final(
    const double* __restrict__ ll_part, int64_t nll, double* __restrict__ negLL,
    double* __restrict__ ll_out, const double* __restrict__ gpart, int64_t nch, int64_t T,
    double gsign, double* __restrict__ G)
{
    __shared__ double sh[kBlock];
    if (nll > 0) {
        double v = 0.0;
        for (int64_t b = threadIdx.x; b < nll; b += kBlock) v += ll_part[b];
        const double s = block_sum(v, sh);
        if (threadIdx.x == 0) {
            const double ll = (s != 0.0) ? s : -INFINITY;
            if (negLL) *negLL = -ll;
            if (ll_out) *ll_out = ll;
        }
    }
    if (G) {
        for (int64_t j = threadIdx.x; j < T; j += kBlock) {
            double g = 0.0;
            for (int64_t c = 0; c < nch; ++c) g += gpart[j * nch + c];
            G[j] = gsign * g;
        }
    }
}

__global__ void k_scale_copy(const float* __restrict__ src, int64_t nbins, int64_t T,
                             int64_t ld, float* __restrict__ dst)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ld * T) return;
    const int64_t j = idx / ld, i = idx - j * ld;
    dst[idx] = (i < nbins) ? src[j * nbins + i] : 0.0f;
}

// ---------------------------------------------------------- launch helpers

static int choose_split(int64_t ld, int64_t T)
{
    // Aim for >= ~4096 waves in flight on 256 CUs; keep >= 8 templates per split.
    const int64_t waves_per_split = cdiv(ld, kCompBins) * (kBlock / 64);
    int64_t s = cdiv(4096, waves_per_split);
    const int64_t smax = cdiv(T, 8);
    if (s > smax) s = smax;
    if (s < 1) s = 1;
    if (s > 64) s = 64;
    return (int)s;
}

int launch_composite(sfh_ctx* ctx, const sfh_templates* tpl, const double* d_coeffs,
                     double* d_partials, int nsplit)
{
    const int64_t ld = ld_of(tpl->nbins);
    const int64_t tsplit = cdiv(tpl->ntemplates, nsplit);
    dim3 grid((unsigned)cdiv(ld, kCompBins), (unsigned)nsplit);
    {
        TimedLaunch t(ctx, "composite_partial");
        hipLaunchKernelGGL(k_composite_partial, grid, dim3(kBlock), 0, ctx->stream,
                           tpl->d_models, d_coeffs, ld, tpl->ntemplates, tsplit, d_partials);
    }
    return check_launch(ctx, "k_composite_partial");
}

// The full single-vector evaluation on device pointers.
//   mode bits: want composite_out (m or R), want LL, want G.
struct FgArgs {
    const double* d_coeffs = nullptr;   // NULL: d_comp_in supplies m directly
    const double* d_comp_in = nullptr;  // composite given by the caller (loglikelihood(composite, data))
    const double* d_data = nullptr;
    int64_t nbins = 0;
    double eps = 0.0;                   // LL clamp: eps(promote_type(composite, data))
    double eps_g = 0.0;                 // gradient clamp: eps(eltype(composite))
    int want_ll = 0;
    double* d_negLL = nullptr;
    double* d_ll = nullptr;
    double* d_G = nullptr;              // NULL: no gradient
    double gsign = 1.0;
    double* d_comp_out = nullptr;
    int clobber_with_R = 0;
};

static int run_fg(sfh_ctx* ctx, const sfh_templates* tpl, const FgArgs& a)
{
    const int64_t nbins = a.nbins;
    const int64_t ld = ld_of(nbins);
    const int64_t T = tpl ? tpl->ntemplates : 0;
    // fg! from coefficients: one pass over the templates where the shape allows
    if (a.d_coeffs && a.d_data && (a.want_ll || a.d_G) && single_pass_covers(tpl))
        return single_pass_fg(ctx, tpl, a.d_data, a.d_coeffs, a.eps, a.eps_g, a.want_ll, a.d_negLL, a.d_ll, a.d_G, a.gsign,
                              a.d_comp_out, a.clobber_with_R);
    int nsplit = 1;
    const double* part = a.d_comp_in;
    if (a.d_coeffs) {
        nsplit = choose_split(ld, T);
        void* p = nullptr;
        int rc = scratch(ctx, "fg_partials", (size_t)nsplit * ld * sizeof(double), &p);
        if (rc) return rc;
        rc = launch_composite(ctx, tpl, a.d_coeffs, (double*)p, nsplit);
        if (rc) return rc;
        part = (const double*)p;
    }
    const int64_t nfin = cdiv(ld, kBlock);
    void* llp = nullptr;
    void* rbuf = nullptr;
    int rc = scratch(ctx, "fg_llpart", (size_t)nfin * sizeof(double), &llp);
    if (rc) return rc;
    if (a.d_G) {
        rc = scratch(ctx, "fg_R", (size_t)ld * sizeof(double), &rbuf);
        if (rc) return rc;
    }
    {
        TimedLaunch t(ctx, "composite_finish");
        hipLaunchKernelGGL(k_composite_finish, dim3((unsigned)nfin), dim3(kBlock), 0, ctx->stream,
                           part, ld, nsplit, nbins, a.d_data, a.eps, a.eps_g, a.want_ll, (double*)llp,
                           (double*)rbuf, a.d_comp_out, a.clobber_with_R);
    }
    rc = check_launch(ctx, "k_composite_finish");
    if (rc) return rc;
    const int64_t nch = cdiv(ld, kGradChunk);
    void* gp = nullptr;
    if (a.d_G) {
        rc = scratch(ctx, "fg_gpart", (size_t)nch * T * sizeof(double), &gp);
        if (rc) return rc;
        dim3 grid((unsigned)nch, (unsigned)cdiv(T, 4 * kGradTplPerWave));
        {
            TimedLaunch t(ctx, "grad_partial");
            hipLaunchKernelGGL(k_grad_partial, grid, dim3(kBlock), 0, ctx->stream, tpl->d_models,
                               (const double*)rbuf, ld, T, (double*)gp, nch);
        }
        rc = check_launch(ctx, "k_grad_partial");
        if (rc) return rc;
    }
    if (a.want_ll || a.d_G) {
        TimedLaunch t(ctx, "fg_final");
        hipLaunchKernelGGL(k_fg_final, dim3(1), dim3(kBlock), 0, ctx->stream, (const double*)llp,
                           a.want_ll ? nfin : (int64_t)0, a.d_negLL, a.d_ll, (const double*)gp, nch,
                           T, a.gsign, a.d_G);
    }
    return check_launch(ctx, "k_fg_final");
}

}  // namespace sfh

using namespace sfh;

// ======================================================================= C ABI

extern "C" {

int sfh_version(void) { return 10000; }

// Errors raised before a context exists (sfh_last_error(NULL) reports them).
static thread_local std::string g_create_error;

static int create_fail(int code, const char* what, hipError_t e)
{
    char buf[512];
    snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
    g_create_error = buf;
    return code;
}

int sfh_ctx_create(int device, void* stream, sfh_ctx** out)
{
    if (!out) return SFH_ERR_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess) return create_fail(SFH_ERR_DEVICE, "hipGetDeviceCount", e);
    if (ndev <= 0) {
        g_create_error = "no HIP device visible";
        return SFH_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        g_create_error = "device index out of range";
        return SFH_ERR_ARGUMENT;
    }
    sfh_ctx* c = new sfh_ctx();
    c->device = device;
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete c;
        return create_fail(SFH_ERR_DEVICE, "hipSetDevice", e);
    }
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return create_fail(SFH_ERR_DEVICE, "hipStreamCreateWithFlags", e);
        }
        c->own_stream = true;
    }
    if (int rc = noalloc_init(device)) {
        if (c->own_stream) hipStreamDestroy(c->stream);
        delete c;
        g_create_error = "noalloc_init: hipGetSymbolAddress of the static fg! buffer failed";
        return rc;
    }
    *out = c;
    return SFH_OK;
}

static void ctx_teardown(sfh_ctx* ctx)
{
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        resolve_timing(ctx);
        for (auto& kv : ctx->scratch)
            if (kv.second.ptr) hipFree(kv.second.ptr);
        for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
        if (ctx->pinned) hipHostFree(ctx->pinned);
        if (ctx->own_stream) hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

}  // extern "C"

namespace sfh {

void ctx_retain(sfh_ctx* ctx)
{
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ++ctx->children;
}

void ctx_release(sfh_ctx* ctx)
{
    bool last = false;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        last = --ctx->children == 0 && ctx->closed;
    }
    if (last) ctx_teardown(ctx);
}

// The deposition strip's instruction mix as independent streams, four waves
// per SIMD: per step, each of 8 chains takes one FMA that reads a scalar
// register (the strip's accumulate with a column factor) and one multiply of
// two vector registers (the strip's recurrence), 64 VALU instructions per
// iteration and lane.  Measures the VALU issue rate such code reaches.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_valu_issue_probe(float s, int iters,
                                                                                                   float* out)
{
    float acc[8], f[8], g[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        acc[i] = 0.0f;
        f[i] = 1.0f + 1e-3f * (float)i + 1e-7f * (float)threadIdx.x;
        g[i] = 1.0f - 1e-9f * (float)(threadIdx.x + i);
    }
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                acc[i] = fmaf(f[i], s, acc[i]);
                f[i] *= g[i];
            }
    }
    float t = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i) t += acc[i] + f[i];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = t;
}

}  // namespace sfh

extern "C" {

// Handles may outlive the call that closes their context (e.g. garbage
// collection order at interpreter exit): the teardown then waits for the
// last of them.
int sfh_ctx_destroy(sfh_ctx* ctx)
{
    if (!ctx) return SFH_OK;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        if (ctx->closed) return SFH_OK;
        ctx->closed = true;
        if (ctx->children > 0) return SFH_OK;
    }
    ctx_teardown(ctx);
    return SFH_OK;
}

const char* sfh_last_error(const sfh_ctx* ctx) { return ctx ? ctx->last_error.c_str() : g_create_error.c_str(); }

int sfh_ctx_synchronize(sfh_ctx* ctx)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    return SFH_OK;
}

int sfh_ctx_capture_begin(sfh_ctx* ctx)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    SFH_ARG_CHECK(ctx, !ctx->timing, "sfh_ctx_capture_begin: switch launch timing off first");
    SFH_HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    return SFH_OK;
}

int sfh_ctx_capture_end(sfh_ctx* ctx, sfh_graph** out)
{
    if (!ctx || !out) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    *out = nullptr;
    if (int rc = bind(ctx)) return rc;
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
    if (ce != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return set_error(ctx, SFH_ERR_DEVICE, "sfh_ctx_capture_end: capture failed (%s); an allocation or "
                         "synchronisation inside the captured calls invalidates it", hipGetErrorString(ce));
    }
    hipGraphExec_t exec = nullptr;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_error(ctx, SFH_ERR_DEVICE, "sfh_ctx_capture_end: graph instantiation failed");
    }
    sfh_graph* gr = new sfh_graph{ctx, graph, exec};
    ctx_retain(ctx);
    *out = gr;
    return SFH_OK;
}

int sfh_graph_launch(sfh_graph* gr, int64_t times)
{
    if (!gr || times < 0) return SFH_ERR_ARGUMENT;
    sfh_ctx* ctx = gr->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    for (int64_t k = 0; k < times; ++k) SFH_HIP_TRY(ctx, hipGraphLaunch(gr->exec, ctx->stream));
    return SFH_OK;
}

int sfh_graph_destroy(sfh_graph* gr)
{
    if (!gr) return SFH_OK;
    sfh_ctx* ctx = gr->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        (void)hipGraphExecDestroy(gr->exec);
        (void)hipGraphDestroy(gr->graph);
        delete gr;
    }
    ctx_release(ctx);
    return SFH_OK;
}

int sfh_ctx_set_timing(sfh_ctx* ctx, int enable)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->timing = enable != 0;
    return SFH_OK;
}

int sfh_ctx_kernel_stats(sfh_ctx* ctx, const char* kernel, double* total_ms, int64_t* launches)
{
    if (!ctx || !kernel) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    resolve_timing(ctx);
    auto it = ctx->stats.find(kernel);
    if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.total_ms;
    if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.launches;
    return SFH_OK;
}

int sfh_probe_valu_issue(sfh_ctx* ctx, int64_t iters, double* lane_instr_per_s)
{
    if (!ctx || !lane_instr_per_s || iters < 1 || iters > (1 << 24)) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    int cus = 0;
    SFH_HIP_TRY(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int blocks = 4 * cus;     // 4 workgroups of 4 waves per CU: 4 waves per SIMD
    void* out = nullptr;
    if (int rc = scratch(ctx, "valu_probe", (size_t)blocks * 256 * sizeof(float), &out)) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SFH_HIP_TRY(ctx, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_error(ctx, SFH_ERR_DEVICE, "hipEventCreate failed");
    }
    hipLaunchKernelGGL(k_valu_issue_probe, dim3(blocks), dim3(256), 0, ctx->stream, 0.5f, 16, (float*)out);
    (void)hipEventRecord(e0, ctx->stream);
    hipLaunchKernelGGL(k_valu_issue_probe, dim3(blocks), dim3(256), 0, ctx->stream, 0.5f, (int)iters, (float*)out);
    (void)hipEventRecord(e1, ctx->stream);
    int rc = check_launch(ctx, "k_valu_issue_probe");
    float ms = 0.0f;
    if (rc == SFH_OK) {
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.0f))
            rc = set_error(ctx, SFH_ERR_DEVICE, "probe timing failed");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *lane_instr_per_s = (double)blocks * 256.0 * 64.0 * (double)iters / ((double)ms * 1e-3);
    return SFH_OK;
}

int sfh_ctx_reset_stats(sfh_ctx* ctx)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (int rc = bind(ctx)) return rc;
    resolve_timing(ctx);
    ctx->stats.clear();
    return SFH_OK;
}

static int templates_make(sfh_ctx* ctx, const float* src, bool src_device, int64_t nbins,
                          int64_t ntemplates, sfh_templates** out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, out && src, "null pointer");
    SFH_ARG_CHECK(ctx, nbins > 0 && ntemplates > 0, "nbins and ntemplates must be positive");
    if (int rc = bind(ctx)) return rc;
    const int64_t ld = ld_of(nbins);
    sfh_templates* t = new sfh_templates();
    t->ctx = ctx;
    t->nbins = nbins;
    t->ntemplates = ntemplates;
    if (hipMalloc(&t->d_models, (size_t)ld * ntemplates * sizeof(float)) != hipSuccess) {
        delete t;
        return set_error(ctx, SFH_ERR_ALLOC, "hipMalloc for %lld x %lld templates failed",
                         (long long)ld, (long long)ntemplates);
    }
    int rc = SFH_OK;
    if (src_device) {
        const int64_t n = ld * ntemplates;
        hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, ctx->stream,
                           src, nbins, ntemplates, ld, t->d_models);
        rc = check_launch(ctx, "k_scale_copy");
        if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = set_error(ctx, SFH_ERR_DEVICE, "template copy failed");
    } else {
        hipError_t e = hipMemsetAsync(t->d_models, 0, (size_t)ld * ntemplates * sizeof(float), ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpy2DAsync(t->d_models, ld * sizeof(float), src, nbins * sizeof(float),
                                 nbins * sizeof(float), ntemplates, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = set_error(ctx, SFH_ERR_DEVICE, "template upload: %s", hipGetErrorString(e));
    }
    if (rc) {
        hipFree(t->d_models);
        delete t;
        return rc;
    }
    ctx_retain(ctx);
    *out = t;
    return SFH_OK;
}

int sfh_templates_create(sfh_ctx* ctx, const float* models, int64_t nbins, int64_t ntemplates,
                         sfh_templates** out)
{
    return templates_make(ctx, models, false, nbins, ntemplates, out);
}

int sfh_templates_create_device(sfh_ctx* ctx, const float* d_models, int64_t nbins,
                                int64_t ntemplates, sfh_templates** out)
{
    return templates_make(ctx, d_models, true, nbins, ntemplates, out);
}

int sfh_templates_destroy(sfh_templates* t)
{
    if (!t) return SFH_OK;
    sfh_ctx* ctx = t->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        release_batched(t);
        hipFree(t->d_models);
        delete t;
    }
    ctx_release(ctx);
    return SFH_OK;
}

int sfh_templates_shape(const sfh_templates* t, int64_t* nbins, int64_t* ntemplates)
{
    if (!t) return SFH_ERR_ARGUMENT;
    if (nbins) *nbins = t->nbins;
    if (ntemplates) *ntemplates = t->ntemplates;
    return SFH_OK;
}

const float* sfh_templates_device_ptr(const sfh_templates* t) { return t ? t->d_models : nullptr; }

static int data_make(sfh_ctx* ctx, const double* src, bool dev, int64_t nbins, sfh_data** out)
{
    if (!ctx) return SFH_ERR_ARGUMENT;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    SFH_ARG_CHECK(ctx, out && src, "null pointer");
    SFH_ARG_CHECK(ctx, nbins > 0, "nbins must be positive");
    if (int rc = bind(ctx)) return rc;
    sfh_data* d = new sfh_data();
    d->ctx = ctx;
    d->nbins = nbins;
    if (hipMalloc(&d->d_data, (size_t)nbins * sizeof(double)) != hipSuccess) {
        delete d;
        return set_error(ctx, SFH_ERR_ALLOC, "hipMalloc for data failed");
    }
    hipError_t e = hipMemcpyAsync(d->d_data, src, (size_t)nbins * sizeof(double),
                                  dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        hipFree(d->d_data);
        delete d;
        return set_error(ctx, SFH_ERR_DEVICE, "data upload: %s", hipGetErrorString(e));
    }
    ctx_retain(ctx);
    *out = d;
    return SFH_OK;
}

int sfh_data_create(sfh_ctx* ctx, const double* data, int64_t nbins, sfh_data** out)
{
    return data_make(ctx, data, false, nbins, out);
}

int sfh_data_create_device(sfh_ctx* ctx, const double* d_data, int64_t nbins, sfh_data** out)
{
    return data_make(ctx, d_data, true, nbins, out);
}

int sfh_data_destroy(sfh_data* d)
{
    if (!d) return SFH_OK;
    sfh_ctx* ctx = d->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        hipFree(d->d_data);
        delete d;
    }
    ctx_release(ctx);
    return SFH_OK;
}

// ---------------------------------------------------------------- P2 entry

#define SFH_ENTER(ctx)                                        \
    if (!(ctx)) return SFH_ERR_ARGUMENT;                      \
    std::lock_guard<std::recursive_mutex> _guard((ctx)->mu);  \
    if (int _rc = bind(ctx)) return _rc;

// Upload `n` doubles from host into a named scratch buffer.
static int upload(sfh_ctx* ctx, const char* name, const double* h, int64_t n, double** d)
{
    void* p = nullptr;
    if (int rc = scratch(ctx, name, (size_t)(n > 0 ? n : 1) * sizeof(double), &p)) return rc;
    SFH_HIP_TRY(ctx, hipMemcpyAsync(p, h, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    *d = (double*)p;
    return SFH_OK;
}

static int download(sfh_ctx* ctx, double* h, const double* d, int64_t n)
{
    SFH_HIP_TRY(ctx, hipMemcpyAsync(h, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    return SFH_OK;
}

int sfh_composite_device(sfh_ctx* ctx, const sfh_templates* tpl, const double* d_coeffs,
                         double* d_composite)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, tpl && d_coeffs && d_composite, "null pointer");
    FgArgs a;
    a.d_coeffs = d_coeffs;
    a.nbins = tpl->nbins;
    a.d_data = nullptr;            // no LL, no R: data is not read
    a.d_comp_out = d_composite;
    a.eps = 0.0;
    a.eps_g = 0.0;
    return run_fg(ctx, tpl, a);
}

int sfh_composite(sfh_ctx* ctx, const sfh_templates* tpl, const double* coeffs, double* composite)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, tpl && coeffs && composite, "null pointer");
    double *dc = nullptr, *dm = nullptr;
    if (int rc = upload(ctx, "h_coeffs", coeffs, tpl->ntemplates, &dc)) return rc;
    void* p = nullptr;
    if (int rc = scratch(ctx, "h_composite", (size_t)tpl->nbins * sizeof(double), &p)) return rc;
    dm = (double*)p;
    if (int rc = sfh_composite_device(ctx, tpl, dc, dm)) return rc;
    if (int rc = download(ctx, composite, dm, tpl->nbins)) return rc;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

int sfh_loglikelihood(sfh_ctx* ctx, const double* composite, const double* data, int64_t n,
                      double eps, double* out)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, composite && data && out, "null pointer");
    SFH_ARG_CHECK(ctx, n > 0, "axes(composite) == axes(data) must be non-empty");
    const int64_t ld = ld_of(n);
    void* p = nullptr;
    if (int rc = scratch(ctx, "ll_comp", (size_t)ld * sizeof(double), &p)) return rc;
    SFH_HIP_TRY(ctx, hipMemsetAsync(p, 0, (size_t)ld * sizeof(double), ctx->stream));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(p, composite, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double* dd = nullptr;
    if (int rc = upload(ctx, "ll_data", data, n, &dd)) return rc;
    void* r = nullptr;
    if (int rc = scratch(ctx, "ll_out", sizeof(double), &r)) return rc;
    FgArgs a;
    a.d_comp_in = (const double*)p;
    a.d_data = dd;
    a.nbins = n;
    a.eps = eps;
    a.eps_g = eps;
    a.want_ll = 1;
    a.d_ll = (double*)r;
    if (int rc = run_fg(ctx, nullptr, a)) return rc;
    if (int rc = download(ctx, out, (const double*)r, 1)) return rc;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

int sfh_fg_device(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                  const double* d_coeffs, int want_F, double eps, double eps_grad, double* d_negLL,
                  double* d_G, double* d_composite)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, tpl && data && d_coeffs, "null pointer");
    SFH_ARG_CHECK(ctx, tpl->nbins == data->nbins,
                  "axes(models,1) == axes(data,1) (%lld vs %lld)", (long long)tpl->nbins,
                  (long long)data->nbins);
    FgArgs a;
    a.d_coeffs = d_coeffs;
    a.d_data = data->d_data;
    a.nbins = tpl->nbins;
    a.eps = eps;
    a.eps_g = eps_grad;
    a.want_ll = want_F ? 1 : 0;
    a.d_negLL = d_negLL;
    a.d_G = d_G;
    a.gsign = 1.0;                 // fg! returns the gradient of -logL (solvers.jl:30)
    a.d_comp_out = d_composite;
    a.clobber_with_R = d_G ? 1 : 0;
    return run_fg(ctx, tpl, a);
}

int sfh_fg(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data, const double* coeffs,
           int want_F, double eps, double eps_grad, double* negLL, double* G, double* composite)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, tpl && data && coeffs, "null pointer");
    SFH_ARG_CHECK(ctx, !want_F || negLL, "negLL must be given when want_F");
    const int64_t T = tpl->ntemplates;
    // The context's pinned host staging holds [negLL, G (T), coefficients (T)]:
    // the coefficients go up from it in one DMA, and the final reduction
    // writes negLL and G into it directly, so no copy comes back (the pageable
    // copies this replaces cost ~30 us per call).  The call holds the context
    // lock until the stream is synchronised, so the staging is not shared.
    void* ph = nullptr;
    if (int rc = pinned(ctx, (size_t)(1 + 2 * T) * sizeof(double), &ph)) return rc;
    double* hout = (double*)ph;
    double* hin = hout + 1 + T;
    if (T > 0) std::memcpy(hin, coeffs, (size_t)T * sizeof(double));
    void* dout = nullptr;
    SFH_HIP_TRY(ctx, hipHostGetDevicePointer(&dout, ph, 0));
    void *pc = nullptr, *pm = nullptr;
    if (int rc = scratch(ctx, "h_coeffs", (size_t)(T > 0 ? T : 1) * sizeof(double), &pc)) return rc;
    if (T > 0)
        SFH_HIP_TRY(ctx, hipMemcpyAsync(pc, hin, (size_t)T * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    if (composite) {
        if (int rc = scratch(ctx, "h_composite", (size_t)tpl->nbins * sizeof(double), &pm)) return rc;
    }
    if (int rc = sfh_fg_device(ctx, tpl, data, (const double*)pc, want_F, eps, eps_grad,
                               want_F ? (double*)dout : nullptr, G ? (double*)dout + 1 : nullptr, (double*)pm))
        return rc;
    if (composite)
        if (int rc = download(ctx, composite, (const double*)pm, tpl->nbins)) return rc;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (want_F) *negLL = hout[0];
    if (G && T > 0) std::memcpy(G, hout + 1, (size_t)T * sizeof(double));
    return SFH_OK;
}

int sfh_loglikelihood_coeffs(sfh_ctx* ctx, const sfh_templates* tpl, const sfh_data* data,
                             const double* coeffs, double eps, double* out)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, out, "null pointer");
    double negll = 0.0;
    if (int rc = sfh_fg(ctx, tpl, data, coeffs, 1, eps, eps, &negll, nullptr, nullptr)) return rc;
    *out = -negll;
    return SFH_OK;
}

static int grad_common(sfh_ctx* ctx, const sfh_templates* tpl, double* composite,
                       const sfh_data* data, double eps, double* G, bool clobber)
{
    SFH_ARG_CHECK(ctx, tpl && data && composite && G, "null pointer");
    SFH_ARG_CHECK(ctx, tpl->nbins == data->nbins,
                  "axes(models,1) == axes(data,1) == axes(composite,1)");
    const int64_t n = tpl->nbins, ld = ld_of(n);
    void* p = nullptr;
    if (int rc = scratch(ctx, "gr_comp", (size_t)ld * sizeof(double), &p)) return rc;
    SFH_HIP_TRY(ctx, hipMemsetAsync(p, 0, (size_t)ld * sizeof(double), ctx->stream));
    SFH_HIP_TRY(ctx, hipMemcpyAsync(p, composite, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    void *pg = nullptr, *pr = nullptr;
    if (int rc = scratch(ctx, "h_G", (size_t)tpl->ntemplates * sizeof(double), &pg)) return rc;
    if (clobber)
        if (int rc = scratch(ctx, "h_composite", (size_t)n * sizeof(double), &pr)) return rc;
    FgArgs a;
    a.d_comp_in = (const double*)p;
    a.d_data = data->d_data;
    a.nbins = n;
    a.eps = eps;
    a.eps_g = eps;
    a.want_ll = 0;
    a.d_G = (double*)pg;
    a.gsign = -1.0;                // ∇loglikelihood: G = -models' * R (fitting_base.jl:283)
    a.d_comp_out = (double*)pr;
    a.clobber_with_R = 1;
    if (int rc = run_fg(ctx, tpl, a)) return rc;
    if (int rc = download(ctx, G, (const double*)pg, tpl->ntemplates)) return rc;
    if (clobber)
        if (int rc = download(ctx, composite, (const double*)pr, n)) return rc;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

int sfh_grad_loglikelihood(sfh_ctx* ctx, const sfh_templates* tpl, const double* composite,
                           const sfh_data* data, double eps, double* G)
{
    SFH_ENTER(ctx);
    return grad_common(ctx, tpl, const_cast<double*>(composite), data, eps, G, false);
}

int sfh_grad_loglikelihood_inplace(sfh_ctx* ctx, const sfh_templates* tpl, double* composite,
                                   const sfh_data* data, double eps, double* G)
{
    SFH_ENTER(ctx);
    return grad_common(ctx, tpl, composite, data, eps, G, true);
}


int sfh_buffer_alloc(sfh_ctx* ctx, int64_t bytes, void** d_ptr)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, d_ptr && bytes >= 0, "sfh_buffer_alloc: bad arguments");
    *d_ptr = nullptr;
    void* p = nullptr;
    if (hipMalloc(&p, (size_t)(bytes > 0 ? bytes : 1)) != hipSuccess)
        return set_error(ctx, SFH_ERR_ALLOC, "hipMalloc(%lld) failed", (long long)bytes);
    *d_ptr = p;
    return SFH_OK;
}

int sfh_buffer_free(sfh_ctx* ctx, void* d_ptr)
{
    SFH_ENTER(ctx);
    if (!d_ptr) return SFH_OK;
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    SFH_HIP_TRY(ctx, hipFree(d_ptr));
    return SFH_OK;
}

int sfh_buffer_upload(sfh_ctx* ctx, void* d_dst, const void* h_src, int64_t bytes)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, bytes >= 0 && (bytes == 0 || (d_dst && h_src)), "sfh_buffer_upload: bad arguments");
    if (bytes) SFH_HIP_TRY(ctx, hipMemcpyAsync(d_dst, h_src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

int sfh_buffer_download(sfh_ctx* ctx, void* h_dst, const void* d_src, int64_t bytes)
{
    SFH_ENTER(ctx);
    SFH_ARG_CHECK(ctx, bytes >= 0 && (bytes == 0 || (d_src && h_dst)), "sfh_buffer_download: bad arguments");
    if (bytes) SFH_HIP_TRY(ctx, hipMemcpyAsync(h_dst, d_src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return SFH_OK;
}

}  // extern "C"

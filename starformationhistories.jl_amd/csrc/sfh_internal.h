// sfh_internal.h -- context, handles, error plumbing and launch timing shared
// by the HIP translation units of libsfh_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sfh.h"

namespace sfh {

// A grow-only device buffer owned by a context (no shrinking: the resident
// working set of a fit is reused call after call).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct KernelStat {
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
};

}  // namespace sfh

struct sfh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::recursive_mutex mu;
    std::string last_error;
    bool timing = false;
    std::map<std::string, sfh::KernelStat> stats;
    std::vector<sfh::PendingTiming> pending;
    std::vector<hipEvent_t> event_pool;
    // named scratch buffers
    std::map<std::string, sfh::DevBuf> scratch;
    // small pinned host staging buffer
    void* pinned = nullptr;
    size_t pinned_bytes = 0;
    // handles (templates, data, fit groups) created on this context and not
    // yet destroyed; sfh_ctx_destroy defers the teardown until they are gone
    int64_t children = 0;
    bool closed = false;
};

struct sfh_templates {
    sfh_ctx* ctx = nullptr;
    int64_t nbins = 0, ntemplates = 0;
    float* d_models = nullptr;          // ld x ntemplates, column-major, ld = nbins rounded up to 256
};

struct sfh_graph {
    sfh_ctx* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
struct sfh_data {
    sfh_ctx* ctx = nullptr;
    int64_t nbins = 0;
    double* d_data = nullptr;
};

namespace sfh {

int set_error(sfh_ctx* ctx, int code, const char* fmt, ...);

#define SFH_HIP_TRY(ctx, expr)                                                     \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess)                                                      \
            return ::sfh::set_error((ctx), SFH_ERR_DEVICE, "%s failed: %s (%s:%d)", \
                                    #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
    } while (0)

#define SFH_ARG_CHECK(ctx, cond, ...)                                              \
    do {                                                                           \
        if (!(cond)) return ::sfh::set_error((ctx), SFH_ERR_ARGUMENT, __VA_ARGS__); \
    } while (0)

// Make the context's device current for the calling thread.
int bind(sfh_ctx* ctx);

// Grow-only named scratch buffer.
int scratch(sfh_ctx* ctx, const char* name, size_t bytes, void** out);
// Pinned host staging buffer of at least `bytes`.
int pinned(sfh_ctx* ctx, size_t bytes, void** out);

// Launch timing: call before / after each kernel launch.
void timing_begin(sfh_ctx* ctx, const char* name, hipEvent_t* start_out);
void timing_end(sfh_ctx* ctx, const char* name, hipEvent_t start);
void resolve_timing(sfh_ctx* ctx);

// Handle bookkeeping: a handle created on ctx calls ctx_retain; its destroy
// calls ctx_release last, outside any lock on ctx (it may free ctx).
void ctx_retain(sfh_ctx* ctx);
void ctx_release(sfh_ctx* ctx);

// Check the last launch error.
int check_launch(sfh_ctx* ctx, const char* name);

struct TimedLaunch {
    sfh_ctx* ctx;
    const char* name;
    hipEvent_t start{};
    TimedLaunch(sfh_ctx* c, const char* n) : ctx(c), name(n) { timing_begin(ctx, name, &start); }
    ~TimedLaunch() { timing_end(ctx, name, start); }
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Frees the batched-path operand layouts built for a template set.
void release_batched(const sfh_templates* t);

// fg! of one template set in one pass over its column-major templates (the
// grouped fg! kernel of sfh_multifit.hip with one fit).  single_pass_covers
// says whether the shape fits it (1..1024 templates, 32-bit offsets).
bool single_pass_covers(const sfh_templates* tpl);
int single_pass_fg(sfh_ctx* ctx, const sfh_templates* tpl, const double* d_data, const double* d_coeffs, double eps,
                   double eps_g, int want_ll, double* d_negLL, double* d_ll, double* d_G, double gsign, double* d_comp_out,
                   int clobber_with_R);

// Resolves, once per device, the address of the static buffer that
// sfh_fg_device_noalloc uses (that device current).  Called by sfh_ctx_create.
int noalloc_init(int device);

// Deposition gather over descriptors already built on the device (template t
// owns desc[koff[t] .. koff[t+1]), koff on the device); FP32 out, column t at
// out32 + t*ld_out.
int deposit_from_desc(sfh_ctx* ctx, const void* desc, const int64_t* d_koff, int64_t ntemplates, int64_t npoints,
                      bool asym, sfh_range xe, sfh_range ye, float* out32, int64_t ld_out);

// Launchers implemented per translation unit.
int launch_composite(sfh_ctx* ctx, const sfh_templates* tpl, const double* d_coeffs,
                     double* d_partials, int nsplit);

}  // namespace sfh

// psf_emu.cpp -- host build of the deposition kernel's arithmetic (the
// __host__ __device__ functions of sfh_psf.h) arranged exactly like
// k_dep_gather: FP32 per-point contributions summed per 64-point chunk -- the
// chunk's common-case points (Gaussian form, footprint spanning the grid
// width) first, then the others, each in point order -- then flushed into
// FP64; narrow erf-difference kernels add each bin straight into FP64.  Lets the CPU test-suite check the FP32 formulation and the footprint
// index math against the FP64 oracle without a GPU.
#include "../../starformationhistories.jl_amd/csrc/sfh_psf.h"

#include <cstring>
#include <vector>

using namespace sfh::psf;

extern "C" int64_t emu_bin_cmd_smooth(const double* colors, const double* mags, const double* cerr,
                                      const double* merr, const double* weights, const double* cmult,
                                      int64_t P, double cov_mult, Range ex, Range ey, double* out)
{
    const int nx = (int)(ex.length - 1), ny = (int)(ey.length - 1);
    const bool asym = cov_mult == 0.0 && cmult == nullptr;
    std::vector<PointDesc> desc((size_t)P);
    int64_t pairs = 0;
    for (int64_t p = 0; p < P; ++p) {
        const double c = cmult ? cmult[p] : cov_mult;
        pairs += asym ? prep_asymmetric(colors[p], mags[p], cerr[p], merr[p], weights[p], ex, ey, &desc[p])
                      : prep_covariant(colors[p], mags[p], cerr[p], merr[p], c, weights[p], ex, ey, &desc[p]);
    }
    // the kernel's strips: a row's columns in runs of kCPT from column 0
    const int ncg = (nx + kCPT - 1) / kCPT;
    std::vector<double> acc((size_t)ncg * kCPT * ny, 0.0);
    std::vector<float> part((size_t)ncg * kCPT * ny, 0.0f);
    const size_t ld = (size_t)ncg * kCPT;
    for (int64_t pc = 0; pc < P; pc += 64) {
        const int64_t pe = pc + 64 < P ? pc + 64 : P;
        for (int pass = 0; pass < 2; ++pass) {
            for (int64_t q = pc; q < pe; ++q) {
                const PointDesc& d = desc[q];
                if (d.xa > d.xb) continue;
                const bool common = d.kind == kGauss && d.xa == 0 && d.xb == nx - 1;
                if (common != (pass == 0)) continue;
                // narrow erf-difference kernels: the kernel's wave-cooperative
                // path, one column at a time, each bin straight into FP64
                const bool coop = d.kind == kDirect && d.xb - d.xa < 63;
                for (int row = d.ya; row <= d.yb; ++row) {
                    float W[3], L[3] = {0.0f, 0.0f, 0.0f}, h[3] = {0.0f, 0.0f, 0.0f};
                    if (asym) W[0] = d.pref * y_factor(fmaf((float)(row - d.ya), d.dv, d.ay), d.ey);
                    else if (common) cov_row_log(d, row, W, L, h);
                    else cov_row(d, row, W, h);
                    if (coop) {
                        const int nodes = asym ? 1 : 3;
                        for (int l = 0; l <= d.xb - d.xa; ++l) {
                            float v = 0.0f;
                            for (int k = 0; k < nodes; ++k) {
                                const float b0 = direct_edge(d, l, h[k]), b1 = direct_edge(d, l + 1, h[k]);
                                v = fmaf(W[k], erf_diff(erfc_pos(fabsf(b0)), b0 < 0.0f, erfc_pos(fabsf(b1)), b1 < 0.0f),
                                         v);
                            }
                            acc[(size_t)row * ld + d.xa + l] += (double)v;
                        }
                        continue;
                    }
                    for (int cg = 0; cg < ncg; ++cg) {
                        const int col0 = cg * kCPT;
                        float* pp = &part[(size_t)row * ld + col0];
                        if (common) {
                            if (asym) gauss_strip<1, false>(d, W, L, h, col0, pp);
                            else gauss_strip<3, false>(d, W, L, h, col0, pp);
                            continue;
                        }
                        if (col0 > d.xb || col0 + kCPT - 1 < d.xa) continue;
                        if (asym) add_point<1, true>(d, W, h, col0, pp);
                        else add_point<3, true>(d, W, h, col0, pp);
                    }
                }
            }
        }
        for (size_t k = 0; k < acc.size(); ++k) {
            acc[k] += (double)part[k];
            part[k] = 0.0f;
        }
    }
    for (int row = 0; row < ny; ++row)
        for (int col = 0; col < nx; ++col) out[(size_t)row * nx + col] = acc[(size_t)row * ld + col];
    return pairs;
}

// sfh_nn.hip -- device primitives of the learned template surrogate: a
// network from (logAge, [M/H], distance modulus, error scale) to a
// log-template, trained on templates from the deposition kernel.
//
// Every matrix product is C = alpha A B^T (+ beta C) with A (M x K) and B
// (N x K) row-major and K contiguous ("NT"), on the matrix cores:
// v_mfma_f32_32x32x16_bf16 for BF16 operands, v_mfma_f32_32x32x2f32 for FP32
// operands (used to check gradients), FP32 accumulation either way.  Callers
// keep K a multiple of 16 with zero padding.  The stride-2, 4x4 transposed
// convolutions of the decoder become such products plus gathers:
//   forward   cols[p, (co,ky,kx)] = in[p, :] . W[(co,ky,kx), :]    (GEMM)
//             out[b,Y,X,co] = sum of the 4 taps (ky,kx) that land on (Y,X) (col2im)
//   backward  dcols[p, (co,ky,kx)] = dout[b, 2y-1+ky, 2x-1+kx, co]    (im2col)
//             din = dcols W,  dW = dcols^T in                          (GEMMs)
// Activations are channels-last (pixel-major, channel fastest).
#include "sfh_internal.h"

#include <cmath>

namespace sfh {
namespace nn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kWG = 256;

__device__ inline uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);   // inf / nan unchanged
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float sigmoid(float x) { return 1.f / (1.f + __expf(-x)); }

template <typename T> __device__ inline T from_f32(float f);
template <> __device__ inline float from_f32<float>(float f) { return f; }
template <> __device__ inline uint16_t from_f32<uint16_t>(float f) { return to_bf16(f); }

// ------------------------------------------------------------------- GEMM
//
// Workgroup tile 128 x 128, four waves of 64 x 64, each 2 x 2 MFMA tiles of
// 32 x 32.  Lane l of a 32 x 32 tile holds row (or column) l % 32 and the 8
// k-values 8 (l / 32) .. +7 of the 16-wide K step; the accumulator element i
// is row 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.

template <typename T>
struct Frag;
template <>
struct Frag<uint16_t> {
    u32x4 v;    // 8 BF16
};
template <>
struct Frag<float> {
    f4 v[2];    // 8 FP32
};

template <typename T>
__device__ inline void load_frag(const T* __restrict__ p, bool ok, Frag<T>& f);
template <>
__device__ inline void load_frag<uint16_t>(const uint16_t* __restrict__ p, bool ok, Frag<uint16_t>& f)
{
    f.v = ok ? *reinterpret_cast<const u32x4*>(p) : u32x4{0u, 0u, 0u, 0u};
}
template <>
__device__ inline void load_frag<float>(const float* __restrict__ p, bool ok, Frag<float>& f)
{
    f.v[0] = ok ? *reinterpret_cast<const f4*>(p) : f4{0.f, 0.f, 0.f, 0.f};
    f.v[1] = ok ? *reinterpret_cast<const f4*>(p + 4) : f4{0.f, 0.f, 0.f, 0.f};
}

__device__ inline f32x16 mma(const Frag<uint16_t>& a, const Frag<uint16_t>& b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a.v), __builtin_bit_cast(bf16x8, b.v),
                                                   c, 0, 0, 0);
}
__device__ inline f32x16 mma(const Frag<float>& a, const Frag<float>& b, f32x16 c)
{
#pragma unroll
    for (int s = 0; s < 8; ++s) c = __builtin_amdgcn_mfma_f32_32x32x2f32(a.v[s >> 2][s & 3], b.v[s >> 2][s & 3], c, 0, 0, 0);
    return c;
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_gemm_nt(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                 int64_t ldb, float* __restrict__ C, int64_t ldc, int64_t M,
                                                 int64_t N, int64_t K, float alpha, float beta)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t m0 = (int64_t)blockIdx.y * 128 + (wave >> 1) * 64;
    const int64_t n0 = (int64_t)blockIdx.x * 128 + (wave & 1) * 64;
    const int r = lane & 31, h = lane >> 5;
    f32x16 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
    const T* pa[2];
    const T* pb[2];
    bool oka[2], okb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int64_t row = m0 + a * 32 + r;
        oka[a] = row < M;
        pa[a] = A + (oka[a] ? row : 0) * lda + 8 * h;
        const int64_t col = n0 + a * 32 + r;
        okb[a] = col < N;
        pb[a] = B + (okb[a] ? col : 0) * ldb + 8 * h;
    }
    Frag<T> fa[2], fb[2], na[2], nb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        load_frag(pa[a], oka[a] && K > 0, fa[a]);
        load_frag(pb[a], okb[a] && K > 0, fb[a]);
    }
    for (int64_t k = 0; k < K; k += 16) {
        const bool more = k + 16 < K;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            load_frag(pa[a] + (more ? k + 16 : 0), oka[a] && more, na[a]);
            load_frag(pb[a] + (more ? k + 16 : 0), okb[a] && more, nb[a]);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = mma(fa[a], fb[b], acc[a][b]);
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            fa[a] = na[a];
            fb[a] = nb[a];
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int64_t row = m0 + a * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
                const int64_t col = n0 + b * 32 + r;
                if (row < M && col < N) {
                    float* c = C + row * ldc + col;
                    *c = beta == 0.f ? alpha * acc[a][b][i] : alpha * acc[a][b][i] + beta * *c;
                }
            }
}

// ------------------------------------------------------------ elementwise

// Y = act(X + bias[col]) over rows x cols (row-major, ldx); pre (rows x cols,
// may be null) keeps X + bias; Y (ldy >= cols) gets zeros in cols..ldy-1.
template <typename T>
__global__ __launch_bounds__(kWG) void k_bias_act(const float* __restrict__ X, int64_t ldx,
                                                  const float* __restrict__ bias, int64_t rows, int64_t cols,
                                                  int act, float* __restrict__ pre, T* __restrict__ Y, int64_t ldy)
{
    const int64_t n = rows * ldy;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t rr = i / ldy, c = i - rr * ldy;
        float y = 0.f;
        if (c < cols) {
            const float x = X[rr * ldx + c] + (bias ? bias[c] : 0.f);
            if (pre) pre[rr * cols + c] = x;
            y = act == 1 ? x * sigmoid(x) : x;
        }
        Y[i] = from_f32<T>(y);
    }
}

// dX = dY * act'(pre), elementwise over n (in place allowed)
__global__ __launch_bounds__(kWG) void k_act_bwd(const float* dY, const float* __restrict__ pre, int64_t n, int act,
                                                 float* dX)
{
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        float d = dY[i];
        if (act == 1) {
            const float x = pre[i], s = sigmoid(x);
            d *= s * (1.f + x * (1.f - s));
        }
        dX[i] = d;
    }
}

// out[c * ldo + r] = act(in[r * ldi + c]) for r < rows, c < cols; zeros for
// rows <= r < ldo and, with ocols > cols, for the rows cols..ocols-1 of out.
template <typename T>
__global__ __launch_bounds__(kWG) void k_transpose(const float* __restrict__ in, int64_t ldi, int64_t rows,
                                                   int64_t cols, T* __restrict__ out, int64_t ldo, int64_t ocols,
                                                   int act)
{
    __shared__ float tile[32][33];
    const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int j = ty; j < 32; j += 8) {
        const int64_t r = r0 + j, c = c0 + tx;
        float v = (r < rows && c < cols) ? in[r * ldi + c] : 0.f;
        if (act == 1) v = v * sigmoid(v);
        tile[j][tx] = v;
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t c = c0 + j, r = r0 + tx;
        if (c < ocols && r < ldo) out[c * ldo + r] = from_f32<T>(tile[tx][j]);
    }
}

// column sums of a rows x cols matrix: per-chunk partials then a fixed-order sum
__global__ __launch_bounds__(kWG) void k_colsum_part(const float* __restrict__ X, int64_t ld, int64_t rows,
                                                     int64_t cols, int64_t chunk, float* __restrict__ part)
{
    const int64_t c = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (c >= cols) return;
    const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = r0 + chunk < rows ? r0 + chunk : rows;
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += X[r * ld + c];
    part[(int64_t)blockIdx.y * cols + c] = s;
}

__global__ __launch_bounds__(kWG) void k_colsum_final(const float* __restrict__ part, int64_t nparts, int64_t cols,
                                                      float beta, float* __restrict__ out)
{
    const int64_t c = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int64_t p = 0; p < nparts; ++p) s += part[p * cols + c];
    out[c] = beta == 0.f ? s : s + beta * out[c];
}

// ---------------------------------------------------------- convolutions

// Transposed convolution, stride 2, 4x4 kernel, padding 1: input H x W,
// output 2H x 2W; output row Y gets input rows y = (Y + 1 - ky) / 2 for the
// two ky of parity (Y + 1) % 2.  cols: [B*H*W][ldc], column co*16 + ky*4 + kx.
__global__ __launch_bounds__(kWG) void k_col2im(const float* __restrict__ cols, int64_t ldc, int64_t B, int64_t H,
                                                int64_t W, int64_t Co, const float* __restrict__ bias,
                                                float* __restrict__ out)
{
    const int64_t H2 = 2 * H, W2 = 2 * W, n = B * H2 * W2 * Co;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t co = i % Co, X = (i / Co) % W2, Y = (i / (Co * W2)) % H2, b = i / (Co * W2 * H2);
        float s = bias ? bias[co] : 0.f;
        const int ky0 = (int)((Y + 1) & 1), kx0 = (int)((X + 1) & 1);
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const int ky = ky0 + 2 * dy;
            const int64_t y = (Y + 1 - ky) / 2;
            if (y < 0 || y >= H) continue;
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int kx = kx0 + 2 * dx;
                const int64_t x = (X + 1 - kx) / 2;
                if (x < 0 || x >= W) continue;
                s += cols[((b * H + y) * W + x) * ldc + co * 16 + ky * 4 + kx];
            }
        }
        out[i] = s;
    }
}

// dcols[(b,y,x)][co*16+ky*4+kx] = dout[b][2y-1+ky][2x-1+kx][co] (0 outside),
// row length ldd (>= 16 Co, zero padded)
template <typename T>
__global__ __launch_bounds__(kWG) void k_im2col(const float* __restrict__ dout, int64_t B, int64_t H, int64_t W,
                                                int64_t Co, T* __restrict__ dcols, int64_t ldd)
{
    const int64_t H2 = 2 * H, W2 = 2 * W, n = B * H * W * ldd;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t j = i % ldd, p = i / ldd;
        float v = 0.f;
        if (j < 16 * Co) {
            const int64_t co = j / 16, ky = (j / 4) % 4, kx = j % 4;
            const int64_t x = p % W, y = (p / W) % H, b = p / (W * H);
            const int64_t Y = 2 * y - 1 + ky, X = 2 * x - 1 + kx;
            if (Y >= 0 && Y < H2 && X >= 0 && X < W2) v = dout[((b * H2 + Y) * W2 + X) * Co + co];
        }
        dcols[i] = from_f32<T>(v);
    }
}

// Single-channel output layer (Co = 1) of the same transposed convolution on
// the vector units: out[b,Y,X] = bmap[Y,X] + sum_taps sum_ci in[b,y,x,ci] w[tap,ci]
__global__ __launch_bounds__(kWG) void k_conv1_fwd(const float* __restrict__ in, int64_t B, int64_t H, int64_t W,
                                                   int64_t Ci, const float* __restrict__ w,
                                                   const float* __restrict__ bmap, float* __restrict__ out)
{
    const int64_t H2 = 2 * H, W2 = 2 * W, n = B * H2 * W2;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t X = i % W2, Y = (i / W2) % H2, b = i / (W2 * H2);
        float s = bmap[Y * W2 + X];
        const int ky0 = (int)((Y + 1) & 1), kx0 = (int)((X + 1) & 1);
        for (int dy = 0; dy < 2; ++dy) {
            const int ky = ky0 + 2 * dy;
            const int64_t y = (Y + 1 - ky) / 2;
            if (y < 0 || y >= H) continue;
            for (int dx = 0; dx < 2; ++dx) {
                const int kx = kx0 + 2 * dx;
                const int64_t x = (X + 1 - kx) / 2;
                if (x < 0 || x >= W) continue;
                const float* pin = in + ((b * H + y) * W + x) * Ci;
                const float* pw = w + (ky * 4 + kx) * Ci;
                for (int64_t c = 0; c < Ci; ++c) s += pin[c] * pw[c];
            }
        }
        out[i] = s;
    }
}

// din[b,y,x,ci] = sum_taps dout[b,2y-1+ky,2x-1+kx] w[tap,ci]
__global__ __launch_bounds__(kWG) void k_conv1_bwd_in(const float* __restrict__ dout, int64_t B, int64_t H,
                                                      int64_t W, int64_t Ci, const float* __restrict__ w,
                                                      float* __restrict__ din)
{
    const int64_t H2 = 2 * H, W2 = 2 * W, n = B * H * W * Ci;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t c = i % Ci, p = i / Ci, x = p % W, y = (p / W) % H, b = p / (W * H);
        float s = 0.f;
        for (int ky = 0; ky < 4; ++ky) {
            const int64_t Y = 2 * y - 1 + ky;
            if (Y < 0 || Y >= H2) continue;
            for (int kx = 0; kx < 4; ++kx) {
                const int64_t X = 2 * x - 1 + kx;
                if (X < 0 || X >= W2) continue;
                s += dout[(b * H2 + Y) * W2 + X] * w[(ky * 4 + kx) * Ci + c];
            }
        }
        din[i] = s;
    }
}

// dw[tap,ci] partials: workgroup g sums pixels g*chunk .. over its range;
// thread t handles (tap, ci) = (t / Ci, t % Ci), 16 Ci <= 512 threads worth
__global__ __launch_bounds__(kWG) void k_conv1_bwd_w(const float* __restrict__ dout, const float* __restrict__ in,
                                                     int64_t B, int64_t H, int64_t W, int64_t Ci, int64_t chunk,
                                                     float* __restrict__ part)
{
    const int64_t H2 = 2 * H, W2 = 2 * W, P = B * H * W;
    const int64_t p0 = (int64_t)blockIdx.x * chunk, p1 = p0 + chunk < P ? p0 + chunk : P;
    for (int64_t t = threadIdx.x; t < 16 * Ci; t += kWG) {
        const int64_t tap = t / Ci, c = t % Ci, ky = tap / 4, kx = tap % 4;
        float s = 0.f;
        for (int64_t p = p0; p < p1; ++p) {
            const int64_t x = p % W, y = (p / W) % H, b = p / (W * H);
            const int64_t Y = 2 * y - 1 + ky, X = 2 * x - 1 + kx;
            if (Y < 0 || Y >= H2 || X < 0 || X >= W2) continue;
            s += in[p * Ci + c] * dout[(b * H2 + Y) * W2 + X];
        }
        part[(int64_t)blockIdx.x * 16 * Ci + t] = s;
    }
}

// Masked mean squared error over the top-left R x S of each B x H2 x W2
// output: loss partials per workgroup and dout = 2 (out - target) / (B R S)
// inside the crop, 0 outside.
__global__ __launch_bounds__(kWG) void k_mse(const float* __restrict__ out, const float* __restrict__ target,
                                             int64_t B, int64_t H2, int64_t W2, int64_t R, int64_t S,
                                             float* __restrict__ dout, float* __restrict__ part)
{
    __shared__ float red[kWG];
    const int64_t n = B * H2 * W2;
    const float scale = 2.f / (float)(B * R * S);
    float acc = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t X = i % W2, Y = (i / W2) % H2, b = i / (W2 * H2);
        float d = 0.f;
        if (Y < R && X < S) {
            d = out[i] - target[(b * R + Y) * S + X];
            acc += d * d;
        }
        if (dout) dout[i] = scale * d;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Adam step on FP32 master parameters (bias-corrected moments)
__global__ __launch_bounds__(kWG) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2, float wd)
{
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

inline unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(cdiv(n, kWG), 8192); }

}  // namespace nn
}  // namespace sfh

using namespace sfh;
using namespace sfh::nn;

#define SFH_NN_BEGIN(ctx)                                  \
    if (!ctx) return SFH_ERR_ARGUMENT;                     \
    std::lock_guard<std::recursive_mutex> _g(ctx->mu);     \
    if (int _rc = bind(ctx)) return _rc;

extern "C" {

int sfh_nn_gemm_nt(sfh_ctx* ctx, int dtype, const void* A, int64_t lda, const void* B, int64_t ldb, float* C,
                   int64_t ldc, int64_t M, int64_t N, int64_t K, float alpha, float beta)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, (dtype == 0 || dtype == 1) && K % 16 == 0 && lda >= K && ldb >= K && ldc >= N && M >= 0 && N >= 0,
                  "sfh_nn_gemm_nt: dtype 0 (BF16) or 1 (FP32), K a multiple of 16, lda, ldb >= K, ldc >= N");
    SFH_ARG_CHECK(ctx, lda % 8 == 0 && ldb % 8 == 0, "sfh_nn_gemm_nt: lda, ldb multiples of 8");
    if (M == 0 || N == 0) return SFH_OK;
    dim3 grid((unsigned)cdiv(N, 128), (unsigned)cdiv(M, 128));
    TimedLaunch tl(ctx, "nn_gemm");
    if (dtype == 0)
        hipLaunchKernelGGL(k_gemm_nt<uint16_t>, grid, dim3(kWG), 0, ctx->stream, (const uint16_t*)A, lda,
                           (const uint16_t*)B, ldb, C, ldc, M, N, K, alpha, beta);
    else
        hipLaunchKernelGGL(k_gemm_nt<float>, grid, dim3(kWG), 0, ctx->stream, (const float*)A, lda, (const float*)B,
                           ldb, C, ldc, M, N, K, alpha, beta);
    return check_launch(ctx, "k_gemm_nt");
}

int sfh_nn_bias_act(sfh_ctx* ctx, int dtype, const float* X, int64_t ldx, const float* bias, int64_t rows,
                    int64_t cols, int act, float* pre, void* Y, int64_t ldy)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, (dtype == 0 || dtype == 1) && ldy >= cols && ldx >= cols && (act == 0 || act == 1),
                  "sfh_nn_bias_act: bad arguments");
    const int64_t n = rows * ldy;
    if (dtype == 0)
        hipLaunchKernelGGL(k_bias_act<uint16_t>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, X, ldx, bias, rows,
                           cols, act, pre, (uint16_t*)Y, ldy);
    else
        hipLaunchKernelGGL(k_bias_act<float>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, X, ldx, bias, rows, cols,
                           act, pre, (float*)Y, ldy);
    return check_launch(ctx, "k_bias_act");
}

int sfh_nn_act_bwd(sfh_ctx* ctx, const float* dY, const float* pre, int64_t n, int act, float* dX)
{
    SFH_NN_BEGIN(ctx);
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, dY, pre, n, act, dX);
    return check_launch(ctx, "k_act_bwd");
}

int sfh_nn_transpose(sfh_ctx* ctx, int dtype, const float* in, int64_t ldi, int64_t rows, int64_t cols, void* out,
                     int64_t ldo, int64_t ocols, int act)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, ldo >= rows && ocols >= cols && ldi >= cols && (act == 0 || act == 1),
                  "sfh_nn_transpose: bad arguments");
    dim3 grid((unsigned)cdiv(ldo, 32), (unsigned)cdiv(ocols, 32));
    if (dtype == 0)
        hipLaunchKernelGGL(k_transpose<uint16_t>, grid, dim3(kWG), 0, ctx->stream, in, ldi, rows, cols, (uint16_t*)out,
                           ldo, ocols, act);
    else
        hipLaunchKernelGGL(k_transpose<float>, grid, dim3(kWG), 0, ctx->stream, in, ldi, rows, cols, (float*)out, ldo,
                           ocols, act);
    return check_launch(ctx, "k_transpose");
}

int sfh_nn_colsum(sfh_ctx* ctx, const float* X, int64_t ld, int64_t rows, int64_t cols, float beta, float* out)
{
    SFH_NN_BEGIN(ctx);
    const int64_t chunk = 256, nparts = std::max<int64_t>(1, cdiv(rows, chunk));
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_colsum", (size_t)nparts * cols * sizeof(float), &p)) return rc;
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)cdiv(cols, kWG), (unsigned)nparts), dim3(kWG), 0, ctx->stream, X,
                       ld, rows, cols, chunk, (float*)p);
    hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(cols, kWG)), dim3(kWG), 0, ctx->stream, (const float*)p,
                       nparts, cols, beta, out);
    return check_launch(ctx, "k_colsum");
}

int sfh_nn_col2im(sfh_ctx* ctx, const float* cols, int64_t ldc, int64_t B, int64_t H, int64_t W, int64_t Co,
                  const float* bias, float* out)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, ldc >= 16 * Co, "sfh_nn_col2im: ldc >= 16 Co");
    const int64_t n = B * 4 * H * W * Co;
    hipLaunchKernelGGL(k_col2im, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, cols, ldc, B, H, W, Co, bias, out);
    return check_launch(ctx, "k_col2im");
}

int sfh_nn_im2col(sfh_ctx* ctx, int dtype, const float* dout, int64_t B, int64_t H, int64_t W, int64_t Co,
                  void* dcols, int64_t ldd)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, ldd >= 16 * Co, "sfh_nn_im2col: ldd >= 16 Co");
    const int64_t n = B * H * W * ldd;
    if (dtype == 0)
        hipLaunchKernelGGL(k_im2col<uint16_t>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, dout, B, H, W, Co,
                           (uint16_t*)dcols, ldd);
    else
        hipLaunchKernelGGL(k_im2col<float>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, dout, B, H, W, Co,
                           (float*)dcols, ldd);
    return check_launch(ctx, "k_im2col");
}

int sfh_nn_conv1_fwd(sfh_ctx* ctx, const float* in, int64_t B, int64_t H, int64_t W, int64_t Ci, const float* w,
                     const float* bmap, float* out)
{
    SFH_NN_BEGIN(ctx);
    const int64_t n = B * 4 * H * W;
    hipLaunchKernelGGL(k_conv1_fwd, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, in, B, H, W, Ci, w, bmap, out);
    return check_launch(ctx, "k_conv1_fwd");
}

int sfh_nn_conv1_bwd(sfh_ctx* ctx, const float* dout, const float* in, int64_t B, int64_t H, int64_t W, int64_t Ci,
                     const float* w, float* din, float* dw, float* dbmap)
{
    SFH_NN_BEGIN(ctx);
    const int64_t P = B * H * W;
    if (din)
        hipLaunchKernelGGL(k_conv1_bwd_in, dim3(grid_for(P * Ci)), dim3(kWG), 0, ctx->stream, dout, B, H, W, Ci, w,
                           din);
    if (dw) {
        const int64_t chunk = 1024, nparts = cdiv(P, chunk);
        void* p = nullptr;
        if (int rc = scratch(ctx, "nn_conv1", (size_t)nparts * 16 * Ci * sizeof(float), &p)) return rc;
        hipLaunchKernelGGL(k_conv1_bwd_w, dim3((unsigned)nparts), dim3(kWG), 0, ctx->stream, dout, in, B, H, W, Ci,
                           chunk, (float*)p);
        hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(16 * Ci, kWG)), dim3(kWG), 0, ctx->stream,
                           (const float*)p, nparts, 16 * Ci, 0.f, dw);
    }
    if (dbmap) {
        // dbmap[Y,X] = sum_b dout[b,Y,X]: column sums of a B x (4 H W) matrix
        const int64_t cols = 4 * H * W;
        void* p = nullptr;
        if (int rc = scratch(ctx, "nn_colsum", (size_t)cols * sizeof(float), &p)) return rc;
        hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)cdiv(cols, kWG), 1), dim3(kWG), 0, ctx->stream, dout, cols,
                           B, cols, B, (float*)p);
        hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(cols, kWG)), dim3(kWG), 0, ctx->stream,
                           (const float*)p, (int64_t)1, cols, 0.f, dbmap);
    }
    return check_launch(ctx, "k_conv1_bwd");
}

int sfh_nn_mse(sfh_ctx* ctx, const float* out, const float* target, int64_t B, int64_t H2, int64_t W2, int64_t R,
               int64_t S, float* dout, float* loss)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, R <= H2 && S <= W2 && loss, "sfh_nn_mse: crop within the output, loss required");
    const int nb = 512;
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_mse", nb * sizeof(float), &p)) return rc;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(kWG), 0, ctx->stream, out, target, B, H2, W2, R, S, dout, (float*)p);
    float part[nb];
    SFH_HIP_TRY(ctx, hipMemcpyAsync(part, p, nb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[i];
    *loss = (float)(s / (double)(B * R * S));
    return SFH_OK;
}

int sfh_nn_adam(sfh_ctx* ctx, float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                float eps, int64_t step, float weight_decay)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, step >= 1, "sfh_nn_adam: step >= 1");
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2,
                       weight_decay);
    return check_launch(ctx, "k_adam");
}

}  // extern "C"

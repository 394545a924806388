// sfh_nn.hip -- device primitives of the learned template surrogate: a
// network from (logAge, [M/H], distance modulus, error scale) to a
// log-template, trained on templates from the deposition kernel.
//
// Every matrix product runs on the matrix cores: v_mfma_f32_32x32x16_bf16 for
// BF16 operands, v_mfma_f32_32x32x2f32 for FP32 operands (used to check
// gradients), FP32 accumulation either way.  Plain products are C = alpha A B^T
// (+ beta C) with A (M x K) and B (N x K) row-major and K contiguous ("NT"),
// K a multiple of 16 with zero padding.  The stride-2, 4x4 transposed
// convolutions of the decoder are, in matrix form,
//   forward   cols[p, (ky,kx,co)] = in[p, :] . W[(ky,kx,co), :]    (GEMM)
//             out[b,Y,X,co] = sum of the 4 taps (ky,kx) that land on (Y,X) (col2im)
//   backward  dcols[p, (ky,kx,co)] = dout[b, 2y-1+ky, 2x-1+kx, co]    (im2col)
//             din = dcols W,  dW = dcols^T in                          (GEMMs)
// and that composition is kept as entry points; the training path runs
// each line as one product whose operands are gathered as they are loaded
// (k_conv_fwd by output parity class, k_conv_dinput, k_conv_dweight), so no
// column matrix is written.  The single-channel output layer works on the
// pixel grid with 16 x 16 x 4 FP32 MFMA tap sums (k_conv1_*).  Activations are
// channels-last (pixel-major, channel fastest) and weight rows tap-major
// ((ky*4 + kx)*Co + co), so the gathers move runs of Co consecutive floats.
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

// Unmasked loads for operands whose out-of-range rows are clamped to a valid
// row: their products land only in accumulator entries that are never stored.
__device__ inline void load_frag_raw(const uint16_t* __restrict__ p, Frag<uint16_t>& f)
{
    f.v = *reinterpret_cast<const u32x4*>(p);
}
__device__ inline void load_frag_raw(const float* __restrict__ p, Frag<float>& f)
{
    f.v[0] = *reinterpret_cast<const f4*>(p);
    f.v[1] = *reinterpret_cast<const f4*>(p + 4);
}

// Branch-free masked loads: a buffer resource over a tensor of nbytes
// (< 2^31) returns zeros for an offset at or past nbytes without touching
// memory, so a gather whose lanes fall outside the grid needs no branch (a
// branch around a load makes the wait that follows it cover every load in
// flight).  kNoRow is such an offset.
constexpr uint32_t kNoRow = 0x80000000u;
__device__ inline __amdgpu_buffer_rsrc_t tensor_rsrc(const void* p, uint32_t nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)nbytes, 0x00020000);
}
__device__ inline u32x4 buf_load16(__amdgpu_buffer_rsrc_t rs, uint32_t off)
{
    return __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
}
__device__ inline void load_frag_buf(__amdgpu_buffer_rsrc_t rs, uint32_t off, Frag<uint16_t>& f)
{
    f.v = buf_load16(rs, off);
}
__device__ inline void load_frag_buf(__amdgpu_buffer_rsrc_t rs, uint32_t off, Frag<float>& f)
{
    f.v[0] = __builtin_bit_cast(f4, buf_load16(rs, off));
    f.v[1] = __builtin_bit_cast(f4, buf_load16(rs, off + 16));
}

__device__ inline void load_frag_lds(const uint16_t* p, Frag<uint16_t>& f) { f.v = *reinterpret_cast<const u32x4*>(p); }
__device__ inline void load_frag_lds(const float* p, Frag<float>& f)
{
    f.v[0] = *reinterpret_cast<const f4*>(p);
    f.v[1] = *reinterpret_cast<const f4*>(p + 4);
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

// Workgroup tile (WM * TM * 32) x (WN * TN * 32): WM x WN waves, each TM x TN
// MFMA tiles of 32 x 32.  Wide outputs use 2 x 2 waves of 2 x 2 tiles (128 x
// 128); outputs at most 32 wide (a layer's channels, a kernel's taps) use 4 x 1
// waves of 2 x 1 tiles (256 x 32), which wastes no matrix-core work on padding.
// Split K: workgroup z covers k in [z kchunk, min(K, (z + 1) kchunk)); with
// more than one slice each writes its raw sums to C + z M N (ldc = N), summed
// later in order with alpha and beta.
template <typename T, int WM, int WN, int TM, int TN>
__global__ __launch_bounds__(kWG) void k_gemm_nt(const T* __restrict__ A, int64_t lda, const T* __restrict__ B,
                                                 int64_t ldb, float* __restrict__ C, int64_t ldc, int64_t M,
                                                 int64_t N, int64_t K, float alpha, float beta, int64_t kchunk)
{
    static_assert(WM * WN * 64 == kWG, "four waves");
    {
        const int64_t kb = (int64_t)blockIdx.z * kchunk;
        A += kb;
        B += kb;
        K = K - kb < kchunk ? K - kb : kchunk;
        if (gridDim.z > 1) C += (int64_t)blockIdx.z * M * N;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t m0 = (int64_t)blockIdx.y * (WM * TM * 32) + (wave / WN) * (TM * 32);
    const int64_t n0 = (int64_t)blockIdx.x * (WN * TN * 32) + (wave % WN) * (TN * 32);
    const int r = lane & 31, h = lane >> 5;
    f32x16 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
#pragma unroll
            for (int i = 0; i < 16; ++i) acc[a][b][i] = 0.f;
    const T* pa[TM];
    const T* pb[TN];
    bool oka[TM], okb[TN];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
        const int64_t row = m0 + a * 32 + r;
        oka[a] = row < M;
        pa[a] = A + (oka[a] ? row : 0) * lda + 8 * h;
    }
#pragma unroll
    for (int b = 0; b < TN; ++b) {
        const int64_t col = n0 + b * 32 + r;
        okb[b] = col < N;
        pb[b] = B + (okb[b] ? col : 0) * ldb + 8 * h;
    }
    // ring of D K-steps in flight: slot s holds step k0 + 16 s and is refilled
    // with step k0 + 16 (s + D) once its products are issued.  The loads are
    // unmasked (no branch around them, so they stay in flight across
    // iterations): rows and columns past M and N read the clamped row 0 into
    // accumulator entries that are never stored, and a step past K (reading
    // step 0 again) is never multiplied.  The main loop runs whole groups of D
    // steps without branches; the last K / 16 mod D steps follow it.
    constexpr int D = sizeof(T) == 2 ? 4 : 2;
    Frag<T> fa[D][TM], fb[D][TN];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        const int64_t ks = 16 * s < K ? 16 * s : 0;
#pragma unroll
        for (int a = 0; a < TM; ++a) load_frag_raw(pa[a] + ks, fa[s][a]);
#pragma unroll
        for (int b = 0; b < TN; ++b) load_frag_raw(pb[b] + ks, fb[s][b]);
        __builtin_amdgcn_sched_barrier(0);
    }
    const int64_t nsteps = K / 16, kfull = nsteps / D * (16 * D);
    for (int64_t k0 = 0; k0 < kfull; k0 += 16 * D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b) acc[a][b] = mma(fa[s][a], fb[s][b], acc[a][b]);
            const int64_t kn = k0 + 16 * (s + D) < K ? k0 + 16 * (s + D) : 0;
#pragma unroll
            for (int a = 0; a < TM; ++a) load_frag_raw(pa[a] + kn, fa[s][a]);
#pragma unroll
            for (int b = 0; b < TN; ++b) load_frag_raw(pb[b] + kn, fb[s][b]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    const int rem = (int)(nsteps - kfull / 16);
#pragma unroll
    for (int s = 0; s < D - 1; ++s)
        if (s < rem) {
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b) acc[a][b] = mma(fa[s][a], fb[s][b], acc[a][b]);
        }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b)
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
// YT (may be null) also gets Y's transpose: YT[c ldt + r] for c < cols, with
// zeros for rows <= r < ldt.
template <typename T>
__device__ inline void store_transposed(T* __restrict__ YT, int64_t ldt, int64_t rows, int64_t rr, int64_t c, T y)
{
    YT[c * ldt + rr] = y;
    if (rr == rows - 1)
        for (int64_t r2 = rows; r2 < ldt; ++r2) YT[c * ldt + r2] = from_f32<T>(0.f);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_bias_act(const float* __restrict__ X, int64_t ldx,
                                                  const float* __restrict__ bias, int64_t rows, int64_t cols,
                                                  int act, float* __restrict__ pre, T* __restrict__ Y, int64_t ldy,
                                                  T* __restrict__ YT, int64_t ldt)
{
    const int64_t n = rows * ldy;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t rr = i / ldy, c = i - rr * ldy;
        float y = 0.f;
        if (c < cols) {
            const float x = X[rr * ldx + c] + (bias ? bias[c] : 0.f);
            if (pre) pre[rr * cols + c] = x;
            y = act == 1 ? x * sigmoid(x) : x;
            if (YT) store_transposed(YT, ldt, rows, rr, c, from_f32<T>(y));
        }
        Y[i] = from_f32<T>(y);
    }
}

// Y = act(x) with x = (the sum of a split product's nparts slices) + bias[c]
// over rows x cols (slice p of element e at part[p rows cols + e]); the
// slices are added as k_colsum_final adds them (four interleaved groups,
// ((s0 + s1) + s2) + s3), so the result equals that sum followed by
// k_bias_act bit for bit.  pre (may be null) keeps x; Y (ldy >= cols) gets
// zeros in cols..ldy-1.
template <typename T>
__global__ __launch_bounds__(kWG) void k_splitk_bias_act(const float* __restrict__ part, int64_t nparts, int64_t rows,
                                                         int64_t cols, const float* __restrict__ bias, int act,
                                                         float* __restrict__ pre, T* __restrict__ Y, int64_t ldy,
                                                         T* __restrict__ YT, int64_t ldt)
{
    const int64_t n = rows * ldy, rc = rows * cols;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t rr = i / ldy, c = i - rr * ldy;
        float y = 0.f;
        if (c < cols) {
            const int64_t e = rr * cols + c;
            float g[4] = {0.f, 0.f, 0.f, 0.f};
            for (int64_t p = 0; p < nparts; ++p) g[p & 3] += part[p * rc + e];
            const float x = 1.f * (((g[0] + g[1]) + g[2]) + g[3]) + (bias ? bias[c] : 0.f);
            if (pre) pre[e] = x;
            y = act == 1 ? x * sigmoid(x) : x;
            if (YT) store_transposed(YT, ldt, rows, rr, c, from_f32<T>(y));
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

// Low-precision copies of up to kMaxWCopies weight matrices in one launch:
// per 32 x 32 tile of src (rows x cols, row-major FP32), dst gets the tile in
// place (row stride cols) and dstT (may be null) its transpose (row stride
// rows), both rounded to T.  Workgroup w takes tile w - first[m] of the
// matrix m whose tile range holds it.
constexpr int kMaxWCopies = 8;
struct WCopies {
    int n;
    int64_t first[kMaxWCopies + 1];       // tile prefix sums
    const float* src[kMaxWCopies];
    int64_t rows[kMaxWCopies], cols[kMaxWCopies];
    void* dst[kMaxWCopies];
    void* dstT[kMaxWCopies];
};

template <typename T>
__global__ __launch_bounds__(kWG) void k_weight_copies(WCopies wc)
{
    __shared__ float tile[32][33];
    const int64_t g = blockIdx.x;
    int m = 0;
    while (m + 1 < wc.n && g >= wc.first[m + 1]) ++m;
    const int64_t rows = wc.rows[m], cols = wc.cols[m];
    const int64_t tcols = (cols + 31) / 32, t = g - wc.first[m];
    const int64_t r0 = (t / tcols) * 32, c0 = (t % tcols) * 32;
    const float* src = wc.src[m];
    T* dst = (T*)wc.dst[m];
    T* dstT = (T*)wc.dstT[m];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    // the four loads issued together from clamped (in-range) positions; the
    // positions past the matrix are neither stored nor transposed
    float v[4];
    const int64_t cc = c0 + tx < cols ? c0 + tx : cols - 1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int64_t r = r0 + ty + 8 * k;
        v[k] = src[(r < rows ? r : rows - 1) * cols + cc];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = ty + 8 * k;
        const int64_t r = r0 + j, c = c0 + tx;
        const bool ok = r < rows && c < cols;
        if (ok) dst[r * cols + c] = from_f32<T>(v[k]);
        tile[j][tx] = ok ? v[k] : 0.f;
    }
    if (!dstT) return;
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int64_t c = c0 + j, r = r0 + tx;
        if (c < cols && r < rows) dstT[c * rows + r] = from_f32<T>(tile[tx][j]);
    }
}

// column sums of a rows x cols matrix: per-chunk partials then a fixed-order sum
// Column sums in two deterministic passes.  Workgroup (x, y): columns
// 64 x .. 64 x + 63, rows of chunk y; thread (g, c) = (t / 64, t % 64) adds
// rows g, g + 4, ..., then the four row groups are added in order.
__global__ __launch_bounds__(kWG) void k_colsum_part(const float* __restrict__ X, int64_t ld, int64_t rows,
                                                     int64_t cols, int64_t chunk, float* __restrict__ part)
{
    __shared__ float red[4][64];
    const int g = threadIdx.x >> 6, cl = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * 64 + cl;
    const int64_t r0 = (int64_t)blockIdx.y * chunk, r1 = r0 + chunk < rows ? r0 + chunk : rows;
    float s = 0.f;
    if (c < cols) {
#pragma unroll 8
        for (int64_t r = r0 + g; r < r1; r += 4) s += X[r * ld + c];
    }
    red[g][cl] = s;
    __syncthreads();
    if (g == 0 && c < cols) part[(int64_t)blockIdx.y * cols + c] = ((red[0][cl] + red[1][cl]) + red[2][cl]) + red[3][cl];
}
// out[(c / N) ldo + c % N] = alpha (sum of the partials of column c) + beta out[...];
// partial p of column c at part[p ldp + c]
__global__ __launch_bounds__(kWG) void k_colsum_final(const float* __restrict__ part, int64_t nparts, int64_t ldp,
                                                      int64_t cols, float alpha, float beta, float* __restrict__ out,
                                                      int64_t N, int64_t ldo)
{
    __shared__ float red[4][64];
    const int g = threadIdx.x >> 6, cl = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * 64 + cl;
    float s = 0.f;
    if (c < cols) {
#pragma unroll 8
        for (int64_t p = g; p < nparts; p += 4) s += part[p * ldp + c];
    }
    red[g][cl] = s;
    __syncthreads();
    if (g == 0 && c < cols) {
        s = alpha * (((red[0][cl] + red[1][cl]) + red[2][cl]) + red[3][cl]);
        float* o = out + (c / N) * ldo + c % N;
        *o = beta == 0.f ? s : s + beta * *o;
    }
}

// ---------------------------------------------------------- convolutions

// Transposed convolution, stride 2, 4x4 kernel, padding 1: input H x W,
// output 2H x 2W; output row Y gets input rows y = (Y + 1 - ky) / 2 for the
// two ky of parity (Y + 1) % 2.  cols: [B*H*W][ldc], column (ky*4 + kx)*Co + co.
__global__ __launch_bounds__(kWG) void k_col2im(const float* __restrict__ cols, int64_t ldc, int64_t B, int64_t H,
                                                int64_t W, int64_t Co, const float* __restrict__ bias,
                                                const float* __restrict__ bmap, float* __restrict__ out)
{
    const int64_t H2 = 2 * H, W2 = 2 * W, n = B * H2 * W2 * Co;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t co = i % Co, X = (i / Co) % W2, Y = (i / (Co * W2)) % H2, b = i / (Co * W2 * H2);
        float s = bias ? bias[co] : 0.f;
        if (bmap) s += bmap[i - b * H2 * W2 * Co];
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
                s += cols[((b * H + y) * W + x) * ldc + (ky * 4 + kx) * Co + co];
            }
        }
        out[i] = s;
    }
}

// dcols[(b,y,x)][(ky*4+kx)*Co + co] = dout[b][2y-1+ky][2x-1+kx][co] (0 outside),
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
            const int64_t co = j % Co, tap = j / Co, ky = tap / 4, kx = tap % 4;
            const int64_t x = p % W, y = (p / W) % H, b = p / (W * H);
            const int64_t Y = 2 * y - 1 + ky, X = 2 * x - 1 + kx;
            if (Y >= 0 && Y < H2 && X >= 0 && X < W2) v = dout[((b * H2 + Y) * W2 + X) * Co + co];
        }
        dcols[i] = from_f32<T>(v);
    }
}

// ------------------------------------------ single-channel output layer, 32 in

// The last transposed convolution has 32 input channels and one output
// channel.  As matrix products its operands would be a P x 16 column matrix
// and FP32 transposes of the 32-channel input; these three kernels work on the
// pixel grid directly, each reading the 32-channel tensor once.
constexpr int kC1Ci = 32;
constexpr int kC1TY = 8, kC1TX = 32;                       // input tile of the forward kernel
constexpr int kC1HY = kC1TY + 2, kC1HX = kC1TX + 2;        // with a one-pixel halo

__device__ inline f4 act4(f4 v, int act)
{
    if (act == 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = v[j] * sigmoid(v[j]);
    }
    return v;
}

// out[b][Y][X] = bmap[Y][X] + sum over the (y, x, tap) that reach (Y, X) of
// act(in[b][y][x][:]) . w[tap][:].  A workgroup walks input tiles (one image's
// kC1TY x kC1TX pixels) t = blockIdx.x, + gridDim.x, ...: the tile's pixels
// (+ a one-pixel halo, zeros outside the image) are staged activated in LDS,
// their 16 tap values computed into LDS (thread = tap x pixel slice, its tap's
// 32 weights in registers), then every output pixel of the tile's 2x-scaled
// footprint adds its four; the next tile's inputs are loaded into registers
// while the current one is computed.
__global__ __launch_bounds__(kWG) void k_conv1_fwd(const float* __restrict__ in, int B, int H, int W,
                                                   const float* __restrict__ w, const float* __restrict__ bmap,
                                                   int act, int64_t ntiles, float* __restrict__ out)
{
    constexpr int NP = kC1HY * kC1HX, NQ = kC1Ci / 4;
    constexpr int NPER = (NP * NQ + kWG - 1) / kWG;
    __shared__ f4 xs[NP][NQ + 1];    // + 1: rows 16 B apart in the banks for the MFMA operand reads
    __shared__ float tv[NP][17];
    const int tilesx = (W + kC1TX - 1) / kC1TX, tilesy = (H + kC1TY - 1) / kC1TY;
    const int H2 = 2 * H, W2 = 2 * W;
    f4 raw[NPER];
    // the tile's pixels (+ halo) through a buffer resource over its image:
    // pixels outside read zeros with no branch around the loads, so these
    // prefetches stay in flight while the current tile is computed
    const uint32_t img_bytes = (uint32_t)(H * W * kC1Ci * 4);
    auto fetch = [&](int64_t t) {
        const int tx = (int)(t % tilesx), ty = (int)((t / tilesx) % tilesy);
        const int64_t b = t / ((int64_t)tilesx * tilesy);
        const int y0 = ty * kC1TY, x0 = tx * kC1TX;
        const __amdgpu_buffer_rsrc_t rs = tensor_rsrc(in + b * H * (int64_t)W * kC1Ci, img_bytes);
#pragma unroll
        for (int j = 0; j < NPER; ++j) {
            const int i = threadIdx.x + kWG * j;
            const int p = i / NQ, q = i % NQ;
            const int yy = y0 - 1 + p / kC1HX, xx = x0 - 1 + p % kC1HX;
            const bool ok = i < NP * NQ && yy >= 0 && yy < H && xx >= 0 && xx < W;
            raw[j] = __builtin_bit_cast(f4, buf_load16(rs, ok ? (uint32_t)(((yy * W + xx) * kC1Ci + 4 * q) * 4) : kNoRow));
        }
    };
    // the output pixel map's values for the tile's outputs, read before the
    // next tile's prefetch is issued so that waiting for them does not wait
    // for the prefetch (loads retire in order)
    constexpr int NOUT = 4 * kC1TY * kC1TX / kWG;
    static_assert(NOUT * kWG == 4 * kC1TY * kC1TX, "whole output passes");
    const __amdgpu_buffer_rsrc_t rb = tensor_rsrc(bmap ? bmap : in, bmap ? (uint32_t)(4 * H * W * 4) : 0u);
    // tap values tv[p][tap] = sum_c act(x[p][c]) w[tap][c] as 16 x 16 x 4 FP32
    // MFMA products: rows = 16 pixels, columns = the 16 taps, k-step ks takes
    // channel 8 g + ks from lane group g = lane / 16, so a lane reads its 8
    // channels of a pixel as two 16-byte LDS loads and keeps its tap's 8
    // weights in registers.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mt = lane & 15, mg = lane >> 4;
    float wk[8];
    {
        const f4 w0 = *reinterpret_cast<const f4*>(w + mt * kC1Ci + 8 * mg);
        const f4 w1 = *reinterpret_cast<const f4*>(w + mt * kC1Ci + 8 * mg + 4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            wk[j] = w0[j];
            wk[4 + j] = w1[j];
        }
    }
    if ((int64_t)blockIdx.x < ntiles) fetch(blockIdx.x);
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int tx = (int)(t % tilesx), ty = (int)((t / tilesx) % tilesy);
        const int64_t b = t / ((int64_t)tilesx * tilesy);
        const int y0 = ty * kC1TY, x0 = tx * kC1TX;
        __syncthreads();                           // the previous tile's LDS reads are done
#pragma unroll
        for (int j = 0; j < NPER; ++j) {
            const int i = threadIdx.x + kWG * j;
            if (i < NP * NQ) xs[i / NQ][i % NQ] = act4(raw[j], act);   // act(0) = 0 keeps the halo zero
        }
        __syncthreads();
        float bm[NOUT];
#pragma unroll
        for (int k = 0; k < NOUT; ++k) {
            const int o = threadIdx.x + kWG * k;
            const int Y = 2 * y0 + o / (2 * kC1TX), X = 2 * x0 + o % (2 * kC1TX);
            bm[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                  rb, Y < H2 && X < W2 ? (uint32_t)((Y * W2 + X) * 4) : kNoRow, 0, 0));
        }
        if (t + gridDim.x < ntiles) fetch(t + gridDim.x);
        for (int rb = wave; rb < (NP + 15) / 16; rb += kWG / 64) {
            const int p = rb * 16 + mt, pr = p < NP ? p : NP - 1;    // rows past NP: results unused
            const f4 x0 = xs[pr][2 * mg], x1 = xs[pr][2 * mg + 1];
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < 4; ++j) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x0[j], wk[j], acc, 0, 0, 0);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x1[j], wk[4 + j], acc, 0, 0, 0);
            // lane holds column (tap) mt of rows 4 mg .. 4 mg + 3
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int pd = rb * 16 + 4 * mg + i;
                if (pd < NP) tv[pd][mt] = acc[i];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NOUT; ++k) {
            const int o = threadIdx.x + kWG * k;
            const int Y = 2 * y0 + o / (2 * kC1TX), X = 2 * x0 + o % (2 * kC1TX);
            if (Y >= H2 || X >= W2) continue;
            float s = bm[k];    // zero without a map (an empty buffer reads zeros)
            const int ky0 = (Y + 1) & 1, kx0 = (X + 1) & 1;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                const int ky = ky0 + 2 * dy, y = (Y + 1 - ky) / 2;
                if (y < 0 || y >= H) continue;
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                    const int kx = kx0 + 2 * dx, x = (X + 1 - kx) / 2;
                    if (x < 0 || x >= W) continue;
                    s += tv[(y - y0 + 1) * kC1HX + (x - x0 + 1)][ky * 4 + kx];
                }
            }
            out[(b * H2 + Y) * (int64_t)W2 + X] = s;
        }
    }
}

// The 16 output-gradient values a pixel's taps reach: d[ky*4 + kx] =
// dout[b][2y - 1 + ky][2x - 1 + kx] (0 outside the output grid).
// (32-bit offsets inside one image: the host keeps 2H x 2W below 2^31.)
__device__ inline void conv1_window(const float* __restrict__ dout, int64_t b, int y, int x, int H2, int W2,
                                    float (&d)[16])
{
    const float* img = dout + b * H2 * (int64_t)W2;
    const int base = (2 * y - 1) * W2 + (2 * x - 1);
    bool xok[4];
#pragma unroll
    for (int kx = 0; kx < 4; ++kx) xok[kx] = 2 * x - 1 + kx >= 0 && 2 * x - 1 + kx < W2;
#pragma unroll
    for (int ky = 0; ky < 4; ++ky) {
        const bool yok = 2 * y - 1 + ky >= 0 && 2 * y - 1 + ky < H2;
#pragma unroll
        for (int kx = 0; kx < 4; ++kx) d[ky * 4 + kx] = (yok && xok[kx]) ? img[base + ky * W2 + kx] : 0.f;
    }
}

// n / d for n, d < 2^31 through an FP64 reciprocal (inv = 1 / d) and one
// exact correction step.
__device__ inline int64_t udiv31(int64_t n, int64_t d, double inv)
{
    int64_t q = (int64_t)((double)n * inv);
    const int64_t r = n - q * d;
    if (r < 0) --q;
    else if (r >= d) ++q;
    return q;
}

// Per-workgroup channel sums of 32-channel pixels: thread t holds f4 sums of
// channel quad t % 8 (a grid-stride loop whose stride is a multiple of 8
// quads keeps each thread on one quad); the workgroup's 32 pixel slots are
// added through LDS in slot order into part[wg][c].
__device__ inline void chan_part32(f4 s, f4* red, float* __restrict__ part)
{
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 32) {
        const int q = threadIdx.x >> 2, j = threadIdx.x & 3;
        float t = 0.f;
        for (int slot = 0; slot < kWG / 8; ++slot) t += red[slot * 8 + q][j];
        part[(int64_t)blockIdx.x * 32 + threadIdx.x] = t;
    }
}

// din[p][c] = sum_tap window(p)[tap] w[tap][c], times act'(in[p][c]) with
// act = SiLU (din is then the gradient of the pre-activation in), as 16 x 16 x 4
// FP32 MFMA products: blocks of 16 pixels (rows) by the 32 channels (two
// column blocks), k-step ks taking taps 4 ks .. 4 ks + 3 (lane l: pixel l % 16,
// tap 4 ks + l / 16, i.e. kernel row ks and column l / 16; its weights held in
// registers); the window through a buffer resource over dout (dout_bytes <
// 2^31), so off-grid taps and pixels past P read zeros with no branch.  Waves
// walk the blocks grid-stride.  With part non-null, also the workgroup's
// channel sums of din: per lane over its rows, then the four lane groups with
// shuffles and the waves through LDS, in a fixed order, into part[wg][c].
__global__ __launch_bounds__(kWG) void k_conv1_din(const float* __restrict__ dout, uint32_t dout_bytes,
                                                   const float* __restrict__ in, int B, int H, int W, double inv_w,
                                                   double inv_h, const float* __restrict__ w, int act,
                                                   float* __restrict__ din, float* __restrict__ part)
{
    constexpr int NWV = kWG / 64;
    __shared__ float red[NWV][kC1Ci];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mt = lane & 15, mk = lane >> 4;
    const int64_t P = (int64_t)B * H * W, nblk = (P + 15) / 16;
    const int H2 = 2 * H, W2 = 2 * W;
    const __amdgpu_buffer_rsrc_t rs = tensor_rsrc(dout, dout_bytes);
    float wb0[4], wb1[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        wb0[ks] = w[(4 * ks + mk) * kC1Ci + mt];
        wb1[ks] = w[(4 * ks + mk) * kC1Ci + 16 + mt];
    }
    float cs0 = 0.f, cs1 = 0.f;
    for (int64_t bi = (int64_t)blockIdx.x * NWV + wave; bi < nblk; bi += (int64_t)gridDim.x * NWV) {
        const int64_t pa = bi * 16 + mt, pq = pa < P ? pa : P - 1;
        const int64_t rr = udiv31(pq, W, inv_w), b = udiv31(rr, H, inv_h);
        const int x = (int)(pq - rr * W), y = (int)(rr - b * H);
        const int X = 2 * x - 1 + mk;
        float a[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int Y = 2 * y - 1 + ks;
            const bool ok = pa < P && Y >= 0 && Y < H2 && X >= 0 && X < W2;
            a[ks] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                  rs, ok ? (uint32_t)(((b * H2 + Y) * W2 + X) * 4) : kNoRow, 0, 0));
        }
        // activations of the rows this lane's results belong to: pixels 4 mk + r
        float x0[4], x1[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t pd = bi * 16 + 4 * mk + r, pc = pd < P ? pd : P - 1;
            x0[r] = act == 1 ? in[pc * kC1Ci + mt] : 0.f;
            x1[r] = act == 1 ? in[pc * kC1Ci + 16 + mt] : 0.f;
        }
        f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], wb0[ks], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], wb1[ks], acc1, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t pd = bi * 16 + 4 * mk + r;
            float v0 = acc0[r], v1 = acc1[r];
            if (act == 1) {
                const float s0 = sigmoid(x0[r]), s1 = sigmoid(x1[r]);
                v0 *= s0 * (1.f + x0[r] * (1.f - s0));
                v1 *= s1 * (1.f + x1[r] * (1.f - s1));
            }
            if (pd < P) {
                din[pd * kC1Ci + mt] = v0;
                din[pd * kC1Ci + 16 + mt] = v1;
                cs0 += v0;
                cs1 += v1;
            }
        }
    }
    if (part) {
        cs0 += __shfl_xor(cs0, 16, 64);
        cs0 += __shfl_xor(cs0, 32, 64);
        cs1 += __shfl_xor(cs1, 16, 64);
        cs1 += __shfl_xor(cs1, 32, 64);
        if (mk == 0) {
            red[wave][mt] = cs0;
            red[wave][16 + mt] = cs1;
        }
        __syncthreads();
        if (threadIdx.x < kC1Ci) {
            float t = red[0][threadIdx.x];
#pragma unroll
            for (int k = 1; k < NWV; ++k) t += red[k][threadIdx.x];
            part[(int64_t)blockIdx.x * kC1Ci + threadIdx.x] = t;
        }
    }
}

// dX = dY act'(pre) over P pixels of 32 channels (dX may alias dY), with the
// workgroup's channel sums of dX (chan_part32); thread = (pixel, 4 channels).
__global__ __launch_bounds__(kWG) void k_act_bwd_chan32(const float* dY, const float* __restrict__ pre, int64_t P,
                                                        int act, float* dX, float* __restrict__ part)
{
    __shared__ f4 red[kWG];
    f4 sum = {0.f, 0.f, 0.f, 0.f};
    for (int64_t g = (int64_t)blockIdx.x * kWG + threadIdx.x; g < P * 8; g += (int64_t)gridDim.x * kWG) {
        f4 d = reinterpret_cast<const f4*>(dY)[g];
        if (act == 1) {
            const f4 x = reinterpret_cast<const f4*>(pre)[g];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float sg = sigmoid(x[j]);
                d[j] *= sg * (1.f + x[j] * (1.f - sg));
            }
        }
        reinterpret_cast<f4*>(dX)[g] = d;
        sum += d;
    }
    chan_part32(sum, red, part);
}

// out[c] = sum over nparts (at most kChanMaxParts) of part[w][c] (32
// channels): 32 groups of 32 threads take parts w = g, g + 32, ..., all
// loaded before the first is added, then the groups are added in order.
constexpr int kChanFinWG = 1024, kChanMaxParts = 2048;
__global__ __launch_bounds__(kChanFinWG) void k_chan_final32(const float* __restrict__ part, int64_t nparts,
                                                             float* __restrict__ out)
{
    __shared__ float red[32][33];
    const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
    constexpr int kPer = kChanMaxParts / 32;
    float v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t w = g + 32 * k;
        v[k] = w < nparts ? part[w * 32 + c] : 0.f;
    }
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < kPer; ++k) t += v[k];
    red[g][c] = t;
    __syncthreads();
    if (threadIdx.x < 32) {
        float u = 0.f;
        for (int k = 0; k < 32; ++k) u += red[k][threadIdx.x];
        out[threadIdx.x] = u;
    }
}

// Partial weight gradients: part[wg][tap*32 + c] = sum over the workgroup's
// pixels of window(p)[tap] act(in[p][c]), as 16 x 16 x 4 FP32 MFMA products
// over groups of 4 pixels: lane l takes tap (A rows) and channel (B columns)
// l % 16 and pixel l / 16 of a group, two products per group for channels
// 0-15 and 16-31; wave w takes the groups w, w + NWV, ... of the workgroup's
// pixel range, U groups per step with the next step's loads in flight.  The
// gradient window comes through a buffer resource over dout (dout_bytes <
// 2^31): taps off the output grid and pixels past the range read zeros with
// no branch; their activations are read from a clamped pixel and multiply
// those zeros.  The waves' 16 x 32 tiles are added through LDS in wave order.
constexpr int kC1DwWG = 512;
__global__ __launch_bounds__(kC1DwWG) void k_conv1_dw(const float* __restrict__ dout, uint32_t dout_bytes,
                                                      const float* __restrict__ in, int B, int H, int W,
                                                      double inv_w, double inv_h, int64_t per_wg, int act,
                                                      float* __restrict__ part)
{
    constexpr int NWV = kC1DwWG / 64, U = 8, STEP = 4 * NWV;    // pixels between a lane's groups
    __shared__ float red[NWV][16 * kC1Ci];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mt = lane & 15, mk = lane >> 4;
    const int ky = mt >> 2, kx = mt & 3;
    const int64_t P = (int64_t)B * H * W;
    const int64_t p0 = (int64_t)blockIdx.x * per_wg;
    const int64_t p1 = p0 + per_wg < P ? p0 + per_wg : P;
    const int H2 = 2 * H, W2 = 2 * W;
    const __amdgpu_buffer_rsrc_t rs = tensor_rsrc(dout, dout_bytes);
    // the next group to load: pixel p and its coordinates, advanced by STEP
    int64_t p = p0 + 4 * wave + mk;
    int64_t rr = udiv31(p < P ? p : P - 1, W, inv_w), b = udiv31(rr, H, inv_h);
    int x = (int)((p < P ? p : P - 1) - rr * W), y = (int)(rr - b * H);
    auto load_group = [&](float& a, float& v0, float& v1) {
        const int Y = 2 * y - 1 + ky, X = 2 * x - 1 + kx;
        const bool ok = p < p1 && Y >= 0 && Y < H2 && X >= 0 && X < W2;
        a = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                          rs, ok ? (uint32_t)(((b * H2 + Y) * W2 + X) * 4) : kNoRow, 0, 0));
        const int64_t pc = p < p1 ? p : p1 - 1;
        v0 = in[pc * kC1Ci + mt];
        v1 = in[pc * kC1Ci + 16 + mt];
        p += STEP;
        x += STEP;
        while (x >= W) {
            x -= W;
            if (++y == H) {
                y = 0;
                ++b;
            }
        }
    };
    const int64_t ngroups = (p1 - p0 + 3) / 4;                  // groups of 4 pixels in the range
    const int64_t nmine = ngroups > wave ? (ngroups - wave + NWV - 1) / NWV : 0;
    f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    float ca[U], c0[U], c1[U];
#pragma unroll
    for (int u = 0; u < U; ++u) load_group(ca[u], c0[u], c1[u]);
    for (int64_t g = 0; g < nmine; g += U) {
        float na[U], n0[U], n1[U];
        const bool more = g + U < nmine;
        if (more) {
#pragma unroll
            for (int u = 0; u < U; ++u) load_group(na[u], n0[u], n1[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // groups past this wave's last read zeros through the window (p >= p1)
            const float s0 = act == 1 ? c0[u] * sigmoid(c0[u]) : c0[u];
            const float s1 = act == 1 ? c1[u] * sigmoid(c1[u]) : c1[u];
            acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[u], s0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[u], s1, acc1, 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ca[u] = na[u];
                c0[u] = n0[u];
                c1[u] = n1[u];
            }
        }
    }
    // lane holds taps 4 mk .. 4 mk + 3 of channels mt and 16 + mt
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        red[wave][(4 * mk + r) * kC1Ci + mt] = acc0[r];
        red[wave][(4 * mk + r) * kC1Ci + 16 + mt] = acc1[r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 16 * kC1Ci; i += kC1DwWG) {
        float s = red[0][i];
#pragma unroll
        for (int k = 1; k < NWV; ++k) s += red[k][i];
        part[(int64_t)blockIdx.x * 16 * kC1Ci + i] = s;
    }
}

// ------------------------------------- transposed-convolution input gradient

__device__ inline void to_frag(f4 v0, f4 v1, Frag<float>& f)
{
    f.v[0] = v0;
    f.v[1] = v1;
}
__device__ inline void to_frag(f4 v0, f4 v1, Frag<uint16_t>& f)
{
    f.v = u32x4{(uint32_t)to_bf16(v0[0]) | ((uint32_t)to_bf16(v0[1]) << 16),
                (uint32_t)to_bf16(v0[2]) | ((uint32_t)to_bf16(v0[3]) << 16),
                (uint32_t)to_bf16(v1[0]) | ((uint32_t)to_bf16(v1[1]) << 16),
                (uint32_t)to_bf16(v1[2]) | ((uint32_t)to_bf16(v1[3]) << 16)};
}

// ------------------------------------- dense-layer gradient preparation

__device__ inline void store8(uint16_t* p, f4 v0, f4 v1)
{
    Frag<uint16_t> f;
    to_frag(v0, v1, f);
    *reinterpret_cast<u32x4*>(p) = f.v;
}
__device__ inline void store8(float* p, f4 v0, f4 v1)
{
    reinterpret_cast<f4*>(p)[0] = v0;
    reinterpret_cast<f4*>(p)[1] = v1;
}

// What a small batch needs of one dense layer's output gradient d (rows x
// cols FP32, row stride cols) before its two products, in one pass with one
// thread per column: d' = d act'(pre), the bias gradient gb[c] = sum_r d'[r][c]
// (four interleaved row sums added as ((s0 + s1) + s2) + s3, the order of
// k_colsum_final), the copy dl[r cols + c] = d'[r][c] and the transpose
// dT[c ldt + r] (zeros for rows <= r < ldt; ldt a multiple of 8, 8 rows per
// store), both rounded to T.  Rows go in blocks of 32 whose loads are all
// issued before the first is used (rows past the last read the last row and
// are replaced by zeros); one-wave workgroups spread the columns over the CUs.
constexpr int kPrepWG = 64;
template <typename T>
__global__ __launch_bounds__(kPrepWG) void k_dense_bwd_prep(const float* __restrict__ d,
                                                            const float* __restrict__ pre, int64_t rows, int64_t cols,
                                                            int act, float* __restrict__ gb, T* __restrict__ dl,
                                                            T* __restrict__ dT, int64_t ldt)
{
    const int64_t c = (int64_t)blockIdx.x * kPrepWG + threadIdx.x;
    if (c >= cols) return;
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    for (int64_t r0 = 0; r0 < ldt; r0 += 32) {
        float x[32], p[32];
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int64_t r = r0 + j < rows ? r0 + j : rows - 1;
            x[j] = rows > 0 ? d[r * cols + c] : 0.f;
            p[j] = rows > 0 && act == 1 ? pre[r * cols + c] : 0.f;
        }
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int64_t r = r0 + j;
            float v = 0.f;
            if (r < rows) {
                v = x[j];
                if (act == 1) {
                    const float sg = sigmoid(p[j]);
                    v *= sg * (1.f + p[j] * (1.f - sg));
                }
                s[j & 3] += v;                // r0 is a multiple of 32, so r % 4 == j % 4
                dl[r * cols + c] = from_f32<T>(v);
            }
            x[j] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (r0 + 8 * q < ldt)
                store8(dT + c * ldt + r0 + 8 * q, f4{x[8 * q], x[8 * q + 1], x[8 * q + 2], x[8 * q + 3]},
                       f4{x[8 * q + 4], x[8 * q + 5], x[8 * q + 6], x[8 * q + 7]});
    }
    gb[c] = ((s[0] + s[1]) + s[2]) + s[3];
}

// Gradient of a stride-2, 4x4, padding-1 transposed convolution's input:
// out[p][ci] = sum_{tap, co} d[b][2y - 1 + ky][2x - 1 + kx][co] wT[ci][tap Co + co]
// for input pixel p = (b, y, x), tap = ky 4 + kx.  The product of the im2col
// rows of d (K = 16 Co) with wT, with 256 x 32 workgroup tiles (4 waves x two
// 32-row MFMA tiles) as k_gemm_nt<.., 4, 1, 2, 1> and a register ring of three
// K-steps, except that the A operand is gathered from the FP32 gradient d as it is
// loaded (a lane's 8 consecutive k are 8 consecutive channels of one output
// pixel, zero outside the grid) and rounded to T, so the column matrix is
// never written.  Rows run along grid.x, 32-column tiles along grid.y; with
// S > 1 slices of the taps (K / S each) along grid.z, slice z writing its
// partial products to out + z P ldo, summed afterwards in slice order.
template <typename T, int Co, int S>
__global__ __launch_bounds__(kWG) void k_conv_dinput(const float* __restrict__ d, uint32_t d_bytes, int B, int Hi,
                                                     int Wi, double inv_w, double inv_h, const T* __restrict__ wT,
                                                     int64_t ldw, int64_t Ci, float* __restrict__ out, int64_t ldo)
{
    constexpr int TM = 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = (int64_t)B * Hi * Wi;
    const int64_t m0 = (int64_t)blockIdx.x * 256 + wave * (TM * 32);
    const int64_t n0 = (int64_t)blockIdx.y * 32;
    const int r = lane & 31, h = lane >> 5;
    const int H2 = 2 * Hi, W2 = 2 * Wi;
    constexpr int K = 16 * Co / S, NS = K / 16;     // this slice's K and K-steps
    static_assert(NS >= 1 && K * S == 16 * Co, "whole K-steps per slice");
    const int kz = (int)blockIdx.z * K;
    if (S > 1) out += (int64_t)blockIdx.z * P * ldo;
    const __amdgpu_buffer_rsrc_t rs = tensor_rsrc(d, d_bytes);
    uint32_t img[TM];    // element offset of the lane's image
    int y2[TM], x2[TM];
    bool oka[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
        const int64_t p = m0 + a * 32 + r;
        oka[a] = p < P;
        const int64_t pp = oka[a] ? p : 0;
        const int64_t rr = udiv31(pp, Wi, inv_w), b = udiv31(rr, Hi, inv_h);
        img[a] = (uint32_t)(b * H2 * (int64_t)W2 * Co);
        x2[a] = 2 * (int)(pp - rr * Wi) - 1;
        y2[a] = 2 * (int)(rr - b * Hi) - 1;
    }
    const int64_t col = n0 + r;
    const bool okb = col < Ci;
    // columns past Ci read row 0 of wT into outputs that are never stored
    const T* pb = wT + (okb ? col : 0) * ldw + 8 * h;
    // the A operand's 8 FP32 values, rounded to T when the product reads
    // them; outside the grid the buffer's range check returns zeros
    struct Raw {
        f4 lo, hi;
    };
    auto load_a = [&](int a, int k0, Raw& v) {
        const int tap = k0 / Co, co0 = k0 - tap * Co;
        const int Y = y2[a] + (tap >> 2), X = x2[a] + (tap & 3);
        const bool ok = oka[a] && Y >= 0 && Y < H2 && X >= 0 && X < W2;
        const uint32_t off = ok ? (img[a] + (uint32_t)((Y * W2 + X) * Co + co0)) * 4u : kNoRow;
        v.lo = __builtin_bit_cast(f4, buf_load16(rs, off));
        v.hi = __builtin_bit_cast(f4, buf_load16(rs, off + 16));
    };
    f32x16 acc[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[a][i] = 0.f;
    // a ring of three fragment sets: the loads of step s + 2 are issued
    // before the products of step s
    Raw fa[3][TM];
    Frag<T> fb[3];
#pragma unroll
    for (int s = 0; s < 2 && s < NS; ++s) {
#pragma unroll
        for (int a = 0; a < TM; ++a) load_a(a, kz + 16 * s + 8 * h, fa[s][a]);
        load_frag_raw(pb + kz + 16 * s, fb[s]);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (s + 2 < NS) {
#pragma unroll
            for (int a = 0; a < TM; ++a) load_a(a, kz + 16 * (s + 2) + 8 * h, fa[(s + 2) % 3][a]);
            load_frag_raw(pb + kz + 16 * (s + 2), fb[(s + 2) % 3]);
        }
#pragma unroll
        for (int a = 0; a < TM; ++a) {
            Frag<T> f;
            to_frag(fa[s % 3][a].lo, fa[s % 3][a].hi, f);
            acc[a] = mma(f, fb[s % 3], acc[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int64_t row = m0 + a * 32 + 8 * (i >> 2) + 4 * h + (i & 3);
            if (row < P && okb) out[row * ldo + col] = acc[a][i];
        }
}

// out[0..n) = v (a kernel, so that the call can sit in a captured graph)
__global__ __launch_bounds__(kWG) void k_fill(float* __restrict__ out, int64_t n, float v)
{
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) out[i] = v;
}

// ----------------------------------------- transposed-convolution forward

// Forward of a stride-2, 4x4, padding-1 transposed convolution with 32 input
// and 32 output channels, by output parity class (py, px) = blockIdx.z: the
// outputs (b, 2y + py, 2x + px) for every input pixel (b, y, x) receive the
// four taps ky = ky0(py) + {0, 2}, kx = kx0(px) + {0, 2} from input pixels
// y + (py + 1 - ky) / 2, x + (px + 1 - kx) / 2 (zero outside), so each class is
// a product with M = input pixels, N = 32, K = 4 taps x 32 input channels: the
// A rows gathered from the input as they are loaded, the B rows from the
// tap-major weights w[tap 32 + co][ci].  Epilogue: pre = acc + bias[co]
// (FP32) and, when aout is given, aout = SiLU(pre) rounded to T.  256 x 32
// workgroup tiles as k_conv_dinput.
template <typename T>
__global__ __launch_bounds__(kWG) void k_conv_fwd(const T* __restrict__ in, uint32_t in_bytes, int B, int Hi, int Wi,
                                                  double inv_w, double inv_h, const T* __restrict__ w,
                                                  const float* __restrict__ bias, float* __restrict__ pre,
                                                  T* __restrict__ aout)
{
    constexpr int TM = 2, C = 32, K = 4 * C;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int py = blockIdx.z >> 1, px = blockIdx.z & 1;
    const int ky0 = (py + 1) & 1, kx0 = (px + 1) & 1;
    const int64_t P = (int64_t)B * Hi * Wi;
    const int64_t m0 = (int64_t)blockIdx.x * 256 + wave * (TM * 32);
    const int r = lane & 31, h = lane >> 5;
    const int H2 = 2 * Hi, W2 = 2 * Wi;
    // output offset of each of the workgroup's 256 rows (-1 past P), one
    // row per thread, so that the epilogue divides nothing
    __shared__ int64_t qrow[256];
    {
        const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
        int64_t q = -1;
        if (p < P) {
            const int64_t rr = udiv31(p, Wi, inv_w), b = udiv31(rr, Hi, inv_h);
            const int x = (int)(p - rr * Wi), y = (int)(rr - b * Hi);
            q = ((b * H2 + 2 * y + py) * (int64_t)W2 + 2 * x + px) * C;
        }
        qrow[threadIdx.x] = q;
    }
    const __amdgpu_buffer_rsrc_t rs = tensor_rsrc(in, in_bytes);
    uint32_t img[TM];    // element offset of the lane's image
    int yy[TM], xx[TM];
    bool oka[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
        const int64_t p = m0 + a * 32 + r;
        oka[a] = p < P;
        const int64_t pp = oka[a] ? p : 0;
        const int64_t rr = udiv31(pp, Wi, inv_w), b = udiv31(rr, Hi, inv_h);
        img[a] = (uint32_t)(b * Hi * (int64_t)Wi * C);
        xx[a] = (int)(pp - rr * Wi);
        yy[a] = (int)(rr - b * Hi);
    }
    // k = j 32 + ci with j = 2 dy + dx: tap (ky0 + 2 dy, kx0 + 2 dx); taps
    // outside the grid read zeros through the buffer's range check
    auto load_a = [&](int a, int k0, Frag<T>& f) {
        const int j = k0 >> 5, ci0 = k0 & 31;
        const int ky = ky0 + 2 * (j >> 1), kx = kx0 + 2 * (j & 1);
        const int y = yy[a] + (py + 1 - ky) / 2, x = xx[a] + (px + 1 - kx) / 2;
        const bool ok = oka[a] && y >= 0 && y < Hi && x >= 0 && x < Wi;
        const uint32_t e = img[a] + (uint32_t)((y * Wi + x) * C + ci0);
        load_frag_buf(rs, ok ? e * (uint32_t)sizeof(T) : kNoRow, f);
    };
    auto load_b = [&](int k0, Frag<T>& f) {
        const int j = k0 >> 5, ci0 = k0 & 31;
        const int tap = (ky0 + 2 * (j >> 1)) * 4 + kx0 + 2 * (j & 1);
        load_frag(w + (int64_t)(tap * C + r) * C + ci0, true, f);
    };
    f32x16 acc[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[a][i] = 0.f;
    // a ring of three fragment sets: the loads of step s + 2 are issued
    // before the products of step s
    constexpr int NS = K / 16;
    Frag<T> fa[3][TM], fb[3];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
        for (int a = 0; a < TM; ++a) load_a(a, 16 * s + 8 * h, fa[s][a]);
        load_b(16 * s + 8 * h, fb[s]);
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (s + 2 < NS) {
#pragma unroll
            for (int a = 0; a < TM; ++a) load_a(a, 16 * (s + 2) + 8 * h, fa[(s + 2) % 3][a]);
            load_b(16 * (s + 2) + 8 * h, fb[(s + 2) % 3]);
        }
#pragma unroll
        for (int a = 0; a < TM; ++a) acc[a] = mma(fa[s % 3][a], fb[s % 3], acc[a]);
    }
    const float bv = bias ? bias[r] : 0.f;
    __syncthreads();    // qrow
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int64_t qb = qrow[wave * (TM * 32) + a * 32 + 8 * (i >> 2) + 4 * h + (i & 3)];
            if (qb < 0) continue;
            const int64_t q = qb + r;
            const float v = acc[a][i] + bv;
            pre[q] = v;
            if (aout) aout[q] = from_f32<T>(v * sigmoid(v));
        }
}

// ------------------------------------ transposed-convolution weight gradient

// Partial weight gradients of a stride-2, 4x4, padding-1 transposed
// convolution with 32 input and 32 output channels:
//   part[s][tap 32 + co][ci] = sum over pixels p of split s of
//                              d[b][2y - 1 + ky][2x - 1 + kx][co] act(pre[p][ci])
// (tap = ky 4 + kx).  Workgroup = (ky, split); its 4 waves are kx = 0..3 and
// each owns one 32 x 32 tile.  Per chunk of 64 pixels the activated inputs and
// the four taps' gradient windows are staged in LDS channel-major ([channel]
// [pixel], rows padded so a lane's 16-byte fragment reads are conflict-free),
// rounded to T; then every wave runs four 32x32x16 MFMAs while the next
// chunk's loads are in flight in registers.  Operands are read
// from the tensors directly: no transposed column matrix, no transposed input.
constexpr int kDwC = 32, kDwChunk = 64;
template <typename T>
__global__ __launch_bounds__(kWG) void k_conv_dweight(const float* __restrict__ d, uint32_t d_bytes,
                                                      const float* __restrict__ pre, int act, int B, int Hi, int Wi,
                                                      double inv_w, double inv_h, int64_t per,
                                                      float* __restrict__ part)
{
    constexpr int PADL = 16 / sizeof(T);            // 16 bytes of padding per row
    constexpr int LD = kDwChunk + PADL;
    __shared__ __align__(16) T sA[kDwC][LD];       // act(pre)[ci][pixel]
    __shared__ __align__(16) T sD[4][kDwC][LD];    // d window of tap (ky, kx)[co][pixel]
    const int ky = blockIdx.x;
    const int64_t P = (int64_t)B * Hi * Wi;
    const int64_t p0 = (int64_t)blockIdx.y * per;
    const int64_t p1 = p0 + per < P ? p0 + per : P;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 31, h = lane >> 5;
    const int H2 = 2 * Hi, W2 = 2 * Wi;
    // staging role of this thread in every chunk: channel quad q of pixel
    // slots pq and pq + 32, for the activations and the four taps' windows
    const int q = threadIdx.x & 7, pq = threadIdx.x >> 3;
    // the next chunk's raw values, loaded without branches (activations from
    // a clamped pixel, masked when staged; gradient windows through a buffer
    // resource over d, zeros off the grid) and activated only when staged, so
    // nothing waits for them before the current chunk's products
    const __amdgpu_buffer_rsrc_t rs = tensor_rsrc(d, d_bytes);
    f4 va[2], vd[2][4];
    bool vok[2];
    auto fetch = [&](int64_t c0) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int64_t p = c0 + pq + 32 * s;
            const bool ok = p < p1;
            const int64_t pp = ok ? p : p0;
            const int64_t rr = udiv31(pp, Wi, inv_w), b = udiv31(rr, Hi, inv_h);
            const int x2 = 2 * (int)(pp - rr * Wi) - 1, y2 = 2 * (int)(rr - b * Hi) - 1;
            vok[s] = ok;
            va[s] = *reinterpret_cast<const f4*>(pre + pp * kDwC + 4 * q);
            const int Y = y2 + ky;
            const bool yok = ok && Y >= 0 && Y < H2;
#pragma unroll
            for (int kx = 0; kx < 4; ++kx) {
                const int X = x2 + kx;
                const uint32_t off = (uint32_t)((((b * H2 + Y) * W2 + X) * kDwC + 4 * q) * 4);
                vd[s][kx] = __builtin_bit_cast(f4, buf_load16(rs, yok && X >= 0 && X < W2 ? off : kNoRow));
            }
        }
    };
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    if (p0 < p1) fetch(p0);
    for (int64_t c0 = p0; c0 < p1; c0 += kDwChunk) {
        __syncthreads();                           // the previous chunk's fragments are read
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int pl = pq + 32 * s;
            const f4 av = vok[s] ? act4(va[s], act) : f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sA[4 * q + j][pl] = from_f32<T>(av[j]);
#pragma unroll
                for (int kx = 0; kx < 4; ++kx) sD[kx][4 * q + j][pl] = from_f32<T>(vd[s][kx][j]);
            }
        }
        __syncthreads();
        if (c0 + kDwChunk < p1) fetch(c0 + kDwChunk);     // in flight during this chunk's products
#pragma unroll
        for (int ks = 0; ks < kDwChunk / 16; ++ks) {
            Frag<T> fa, fb;
            load_frag_lds(&sD[wave][r][ks * 16 + 8 * h], fa);
            load_frag_lds(&sA[r][ks * 16 + 8 * h], fb);
            acc = mma(fa, fb, acc);
        }
    }
    float* o = part + (int64_t)blockIdx.y * (16 * kDwC * kDwC) + (int64_t)(ky * 4 + wave) * kDwC * kDwC;
#pragma unroll
    for (int i = 0; i < 16; ++i) o[(8 * (i >> 2) + 4 * h + (i & 3)) * kDwC + r] = acc[i];
}

// Four consecutive channels per thread (Co % 4 == 0, ldc / ldd % 4 == 0),
// 32-bit index arithmetic (P * ldd < 2^31).
template <typename T>
__device__ inline void store4(T* p, f4 v);
template <>
__device__ inline void store4<float>(float* p, f4 v) { *reinterpret_cast<f4*>(p) = v; }
template <>
__device__ inline void store4<uint16_t>(uint16_t* p, f4 v)
{
    const uint32_t a = (uint32_t)to_bf16(v[0]) | ((uint32_t)to_bf16(v[1]) << 16);
    const uint32_t b = (uint32_t)to_bf16(v[2]) | ((uint32_t)to_bf16(v[3]) << 16);
    *reinterpret_cast<uint2*>(p) = make_uint2(a, b);
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_im2col4(const float* __restrict__ dout, int B, int H, int W, int Co,
                                                 T* __restrict__ dcols, int ldd)
{
    const int H2 = 2 * H, W2 = 2 * W, q4 = ldd / 4, n = B * H * W * q4;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const int j = (i % q4) * 4, p = i / q4;
        f4 v = {0.f, 0.f, 0.f, 0.f};
        if (j < 16 * Co) {
            const int co = j % Co, tap = j / Co, ky = tap >> 2, kx = tap & 3;
            const int x = p % W, y = (p / W) % H, b = p / (W * H);
            const int Y = 2 * y - 1 + ky, X = 2 * x - 1 + kx;
            if (Y >= 0 && Y < H2 && X >= 0 && X < W2)
                v = *reinterpret_cast<const f4*>(dout + ((int64_t)(b * H2 + Y) * W2 + X) * Co + co);
        }
        store4<T>(dcols + (int64_t)p * ldd + j, v);
    }
}

__global__ __launch_bounds__(kWG) void k_col2im4(const float* __restrict__ cols, int ldc, int B, int H, int W,
                                                 int Co, const float* __restrict__ bias, float* __restrict__ out)
{
    const int H2 = 2 * H, W2 = 2 * W, c4 = Co / 4, n = B * H2 * W2 * c4;
    for (int i = blockIdx.x * kWG + threadIdx.x; i < n; i += gridDim.x * kWG) {
        const int co = (i % c4) * 4, X = (i / c4) % W2, Y = (i / (c4 * W2)) % H2, b = i / (c4 * W2 * H2);
        f4 s = bias ? *reinterpret_cast<const f4*>(bias + co) : f4{0.f, 0.f, 0.f, 0.f};
        const int ky0 = (Y + 1) & 1, kx0 = (X + 1) & 1;
#pragma unroll
        for (int dy = 0; dy < 2; ++dy) {
            const int ky = ky0 + 2 * dy, y = (Y + 1 - ky) / 2;
            if (y < 0 || y >= H) continue;
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
                const int kx = kx0 + 2 * dx, x = (X + 1 - kx) / 2;
                if (x < 0 || x >= W) continue;
                s += *reinterpret_cast<const f4*>(cols + (int64_t)((b * H + y) * W + x) * ldc + (ky * 4 + kx) * Co +
                                                  co);
            }
        }
        *reinterpret_cast<f4*>(out + ((int64_t)(b * H2 + Y) * W2 + X) * Co + co) = s;
    }
}

// The weight-gradient operand directly: dcolsT[j][p] = dcols[p][j] for
// j < 16 Co, p < ldo (zero for p >= P); Co in {4, 8, 16, 32}.  Workgroup
// (x, y) gathers TP pixels x the 256 / Co taps of block y into LDS (float4
// reads of 4 consecutive channels), then writes each of its rows' TP pixels
// contiguously (8 threads per row).
template <typename T, int TP>
__global__ __launch_bounds__(kWG) void k_im2col_t(const float* __restrict__ dout, int B, int H, int W, int Co,
                                                  T* __restrict__ dcolsT, int ldo)
{
    __shared__ T tile[256][TP + 2];
    const int H2 = 2 * H, W2 = 2 * W, P = B * H * W, c4 = Co / 4;
    const int tpb = 256 / Co, tap0 = blockIdx.y * tpb, ntap = tpb < 16 - tap0 ? tpb : 16 - tap0;
    const int p0 = blockIdx.x * TP;
    const int units = TP * ntap * c4;
    for (int u = threadIdx.x; u < units; u += kWG) {
        const int co = (u % c4) * 4, tl = (u / c4) % ntap, pl = u / (c4 * ntap);
        const int tap = tap0 + tl, p = p0 + pl;
        f4 v = {0.f, 0.f, 0.f, 0.f};
        if (p < P) {
            const int x = p % W, y = (p / W) % H, b = p / (W * H);
            const int Y = 2 * y - 1 + (tap >> 2), X = 2 * x - 1 + (tap & 3);
            if (Y >= 0 && Y < H2 && X >= 0 && X < W2)
                v = *reinterpret_cast<const f4*>(dout + ((int64_t)(b * H2 + Y) * W2 + X) * Co + co);
        }
        const int jl = tl * Co + co;
#pragma unroll
        for (int e = 0; e < 4; ++e) tile[jl + e][pl] = from_f32<T>(v[e]);
    }
    __syncthreads();
    constexpr int per = TP / 8;
    const int rows = ntap * Co;
    for (int r = threadIdx.x / 8; r < rows; r += kWG / 8) {
        const int c0 = (threadIdx.x % 8) * per;
        T* dst = dcolsT + (int64_t)(tap0 * Co + r) * ldo + p0 + c0;
#pragma unroll
        for (int e = 0; e < per; ++e)
            if (p0 + c0 + e < ldo) dst[e] = tile[r][c0 + e];
    }
}

// Masked mean squared error over the top-left R x S of each B x H2 x W2
// output: loss partials per workgroup and dout = 2 (out - target) / (B R S)
// inside the crop, 0 outside.  Output b is compared with target row rows[b]
// (b itself when rows is NULL).
__global__ __launch_bounds__(kWG) void k_mse(const float* __restrict__ out, const float* __restrict__ target,
                                             const int64_t* __restrict__ rows, int64_t B, int64_t H2, int64_t W2,
                                             int64_t R, int64_t S, float* __restrict__ dout,
                                             float* __restrict__ part)
{
    __shared__ float red[kWG];
    const int64_t n = B * H2 * W2;
    const float scale = 2.f / (float)(B * R * S);
    float acc = 0.f;
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const int64_t X = i % W2, Y = (i / W2) % H2, b = i / (W2 * H2);
        float d = 0.f;
        if (Y < R && X < S) {
            const int64_t tb = rows ? rows[b] : b;
            d = out[i] - target[(tb * R + Y) * S + X];
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

// Template loss: w_mse MSE(out, target) + mean_b [w_kl KL(p_e || p_s) +
// w_sum (log Ss - log Se)^2] over the R x S crop, with s = exp(sc out + off),
// e = exp(sc target + off) (outputs clamped at 15 inside the exponential),
// Ss = sum s, Se = sum e and p = s / Ss.  Pass 1 (grid kTLChunks x B): per
// chunk of an image, the sums of s, e, e (t - y) and (y - t)^2.  Pass 2:
// dout = w_mse 2 (y - t) / (B R S) + sc [w_kl (p_s - p_e) +
// 2 w_sum (log Ss - log Se) p_s] / B inside the crop, 0 outside.  Pass 3 adds
// the three terms (unweighted) to acc[0..2].
constexpr int kTLChunks = 16;

__global__ __launch_bounds__(kWG) void k_tl_sums(const float* __restrict__ out, const float* __restrict__ target,
                                                 const int64_t* __restrict__ rows, int64_t H2, int64_t W2, int64_t R,
                                                 int64_t S, float sc, float off, double* __restrict__ part)
{
    __shared__ double red[4][kWG];
    const int64_t b = blockIdx.y, n = R * S;
    const int64_t tb = rows ? rows[b] : b;
    const int64_t c0 = n * blockIdx.x / kTLChunks, c1 = n * (blockIdx.x + 1) / kTLChunks;
    double ss = 0.0, se = 0.0, a = 0.0, q = 0.0;
    // (Y, X) of element i, stepped with i (by kWG / S rows and kWG % S columns)
    // instead of divided per element; unrolled so that several iterations'
    // loads are in flight together
    int64_t Y = (c0 + threadIdx.x) / S, X = (c0 + threadIdx.x) % S;
    const int64_t dy = kWG / S, dx = kWG % S;
#pragma unroll 4
    for (int64_t i = c0 + threadIdx.x; i < c1; i += kWG) {
        const float y = out[(b * H2 + Y) * W2 + X], t = target[tb * n + i];
        X += dx;
        Y += dy;
        if (X >= S) {
            X -= S;
            ++Y;
        }
        const float sv = expf(fminf(y, 15.f) * sc + off), ev = expf(t * sc + off);
        ss += sv;
        se += ev;
        a += (double)ev * (double)(t - y);
        q += (double)(y - t) * (double)(y - t);
    }
    red[0][threadIdx.x] = ss;
    red[1][threadIdx.x] = se;
    red[2][threadIdx.x] = a;
    red[3][threadIdx.x] = q;
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(b * kTLChunks + blockIdx.x) * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__device__ inline void tl_totals(const double* __restrict__ part, int64_t b, double& ss, double& se)
{
    ss = 0.0;
    se = 0.0;
    for (int c = 0; c < kTLChunks; ++c) {
        ss += part[(b * kTLChunks + c) * 4];
        se += part[(b * kTLChunks + c) * 4 + 1];
    }
}

__global__ __launch_bounds__(kWG) void k_tl_grad(const float* __restrict__ out, const float* __restrict__ target,
                                                 const int64_t* __restrict__ rows, int64_t B, int64_t H2, int64_t W2,
                                                 int64_t R, int64_t S, float sc, float off, float w_mse, float w_kl,
                                                 float w_sum, const double* __restrict__ part,
                                                 float* __restrict__ dout)
{
    const int64_t b = blockIdx.y, n2 = H2 * W2;
    const int64_t tb = rows ? rows[b] : b;
    double ssd, sed;
    tl_totals(part, b, ssd, sed);
    const float inv_ss = (float)(1.0 / ssd), inv_se = (float)(1.0 / sed);
    const float gm = w_mse * 2.f / (float)(B * R * S), gk = w_kl * sc / (float)B;
    const float gs = (float)(2.0 * w_sum * (log(ssd) - log(sed))) * sc / (float)B;
    // (Y, X) of element i, stepped with i instead of divided per element
    const int64_t i0 = (int64_t)blockIdx.x * kWG + threadIdx.x, step = (int64_t)gridDim.x * kWG;
    const int64_t dy = step / W2, dx = step % W2;
    int64_t Y = i0 / W2, X = i0 % W2;
#pragma unroll 4
    for (int64_t i = i0; i < n2; i += step) {
        float d = 0.f;
        if (Y < R && X < S) {
            const float y = out[b * n2 + i], t = target[tb * R * S + Y * S + X];
            const float ps = expf(fminf(y, 15.f) * sc + off) * inv_ss, pe = expf(t * sc + off) * inv_se;
            d = gm * (y - t) + gk * (ps - pe) + gs * ps;
        }
        dout[b * n2 + i] = d;
        X += dx;
        Y += dy;
        if (X >= W2) {
            X -= W2;
            ++Y;
        }
    }
}

// acc[0] += MSE, acc[1] += mean over images of KL = sc A / Se + log Ss - log Se,
// acc[2] += mean of (log Ss - log Se)^2
__global__ __launch_bounds__(64) void k_tl_accumulate(const double* __restrict__ part, int64_t B, int64_t R,
                                                      int64_t S, float sc, double* __restrict__ acc)
{
    // thread t handles images t, t + 64, ...; the per-thread sums are added in order
    __shared__ double red[3][64];
    double q = 0.0, kl = 0.0, sq = 0.0;
    for (int64_t b = threadIdx.x; b < B; b += 64) {
        double ss, se, a = 0.0;
        tl_totals(part, b, ss, se);
        for (int c = 0; c < kTLChunks; ++c) {
            a += part[(b * kTLChunks + c) * 4 + 2];
            q += part[(b * kTLChunks + c) * 4 + 3];
        }
        const double ls = log(ss) - log(se);
        kl += (double)sc * a / se + ls;
        sq += ls * ls;
    }
    red[0][threadIdx.x] = q;
    red[1][threadIdx.x] = kl;
    red[2][threadIdx.x] = sq;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int i = 1; i < 64; ++i) {
        q += red[0][i];
        kl += red[1][i];
        sq += red[2][i];
    }
    acc[0] += q / (double)(B * R * S);
    acc[1] += kl / (double)B;
    acc[2] += sq / (double)B;
}

// acc[0] += (sum of the nb partials, in a fixed order) / denom
__global__ __launch_bounds__(64) void k_mse_accumulate(const float* __restrict__ part, int nb, double denom,
                                                       double* __restrict__ acc)
{
    __shared__ double red[64];
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += 64) s += (double)part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int i = 1; i < 64; ++i) s += red[i];
    acc[0] += s / denom;
}

// dst[i, c] = src[rows[i], c] for c < cols
__global__ __launch_bounds__(kWG) void k_gather_rows(const float* __restrict__ src, int64_t lds,
                                                     const int64_t* __restrict__ rows, int64_t n, int64_t cols,
                                                     float* __restrict__ dst, int64_t ldd)
{
    for (int64_t t = (int64_t)blockIdx.x * kWG + threadIdx.x; t < n * cols; t += (int64_t)gridDim.x * kWG) {
        const int64_t i = t / cols, c = t % cols;
        dst[i * ldd + c] = src[rows[i] * lds + c];
    }
}

// Template layout (bin = ix + nx iy, one template per row of ldt) <-> image
// layout (image b: nx rows of ny, row ix holds the magnitude bins of colour
// bin ix).  Both directions are 32 x 32 tile transposes through LDS, so reads
// and writes stay coalesced.
constexpr int kTile = 32;

// img[b, ix, iy] = (log(T[b, ix + nx iy] + eps) - offset) / scale
__global__ __launch_bounds__(kWG) void k_log_target(const float* __restrict__ T, int64_t ldt, int64_t nx, int64_t ny,
                                                    float eps, float offset, float inv_scale, float* __restrict__ img)
{
    __shared__ float tile[kTile][kTile + 1];
    const int64_t b = blockIdx.z;
    const int64_t ix0 = (int64_t)blockIdx.x * kTile, iy0 = (int64_t)blockIdx.y * kTile;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;     // 32 x 8
    const float* Tb = T + b * ldt;
    for (int r = ty; r < kTile; r += kWG / kTile) {
        const int64_t ix = ix0 + tx, iy = iy0 + r;
        if (ix < nx && iy < ny) tile[r][tx] = (logf(Tb[ix + nx * iy] + eps) - offset) * inv_scale;
    }
    __syncthreads();
    float* ib = img + b * nx * ny;
    for (int r = ty; r < kTile; r += kWG / kTile) {
        const int64_t ix = ix0 + r, iy = iy0 + tx;
        if (ix < nx && iy < ny) ib[ix * ny + iy] = tile[tx][r];
    }
}

// T[b, ix + nx iy] = max(exp(out[b, ix, iy] scale + offset) - eps, 0) from the
// top-left nx x ny of each H2 x W2 network output
__global__ __launch_bounds__(kWG) void k_log_template(const float* __restrict__ out, int64_t H2, int64_t W2,
                                                      int64_t nx, int64_t ny, float eps, float offset, float scale,
                                                      float* __restrict__ T, int64_t ldt)
{
    __shared__ float tile[kTile][kTile + 1];
    const int64_t b = blockIdx.z;
    const int64_t ix0 = (int64_t)blockIdx.x * kTile, iy0 = (int64_t)blockIdx.y * kTile;
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const float* ob = out + b * H2 * W2;
    for (int r = ty; r < kTile; r += kWG / kTile) {
        const int64_t ix = ix0 + r, iy = iy0 + tx;
        if (ix < nx && iy < ny) tile[r][tx] = fmaxf(expf(ob[ix * W2 + iy] * scale + offset) - eps, 0.f);
    }
    __syncthreads();
    float* Tb = T + b * ldt;
    for (int r = ty; r < kTile; r += kWG / kTile) {
        const int64_t ix = ix0 + tx, iy = iy0 + r;
        if (ix < nx && iy < ny) Tb[ix + nx * iy] = tile[tx][r];
    }
}

// Per template k (one workgroup): with s' = max(S, floor), the sums
// Se = sum E, Ss = sum s', A = sum_{E > 0} E log(E / s') in FP64 and
// KL(E / Se || s' / Ss) = A / Se - log Se + log Ss.  out[3k..3k+2] = KL, Se, Ss.
__global__ __launch_bounds__(kWG) void k_template_kl(const float* __restrict__ E, int64_t lde,
                                                     const float* __restrict__ S, int64_t lds, int64_t nbins,
                                                     double floor_, double* __restrict__ out)
{
    __shared__ double red[3][kWG];
    const int64_t k = blockIdx.x;
    const float* e = E + k * lde;
    const float* s = S + k * lds;
    double se = 0.0, ss = 0.0, a = 0.0;
    for (int64_t i = threadIdx.x; i < nbins; i += kWG) {
        const double ei = e[i], si = fmax((double)s[i], floor_);
        se += ei;
        ss += si;
        if (ei > 0.0) a += ei * log(ei / si);
    }
    red[0][threadIdx.x] = se;
    red[1][threadIdx.x] = ss;
    red[2][threadIdx.x] = a;
    __syncthreads();
    for (int w = kWG / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        se = red[0][0];
        ss = red[1][0];
        a = red[2][0];
        out[3 * k] = se > 0.0 ? a / se - log(se) + log(ss) : 0.0;
        out[3 * k + 1] = se;
        out[3 * k + 2] = ss;
    }
}

__global__ __launch_bounds__(64) void k_select_rows(const int64_t* __restrict__ rows_all,
                                                    const int64_t* __restrict__ counter, int64_t n,
                                                    int64_t* __restrict__ rows_step)
{
    const int64_t base = counter[0] * n;
    for (int64_t i = threadIdx.x; i < n; i += 64) rows_step[i] = rows_all[base + i];
}

__global__ __launch_bounds__(64) void k_counter_add(int64_t* counter, int64_t inc)
{
    if (threadIdx.x == 0) counter[0] += inc;
}

// Adam with (lr, c1, c2) from a per-step schedule
__global__ __launch_bounds__(kWG) void k_adam_sched(float* __restrict__ p, const float* __restrict__ g,
                                                    float* __restrict__ m, float* __restrict__ v, int64_t n,
                                                    const float* __restrict__ sched,
                                                    const int64_t* __restrict__ counter, float b1, float b2,
                                                    float eps, float wd)
{
    const int64_t c = counter[0];
    const float lr = sched[3 * c], c1 = sched[3 * c + 1], c2 = sched[3 * c + 2];
    for (int64_t i = (int64_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (int64_t)gridDim.x * kWG) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
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
// workgroups of the channel-sum kernels over P pixels of 32 channels: a fixed
// function of P, so the partials are always added in the same order
inline int64_t chan_grid(int64_t P)
{
    return std::max<int64_t>(1, std::min<int64_t>(cdiv(P * 8, kWG), kChanMaxParts));
}

}  // namespace nn
}  // namespace sfh

using namespace sfh;
using namespace sfh::nn;

#define SFH_NN_BEGIN(ctx)                                  \
    if (!ctx) return SFH_ERR_ARGUMENT;                     \
    std::lock_guard<std::recursive_mutex> _g(ctx->mu);     \
    if (int _rc = bind(ctx)) return _rc;

namespace {

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int colsum(sfh_ctx* ctx, const float* X, int64_t ld, int64_t rows, int64_t cols, float alpha, float beta, float* out,
           int64_t N, int64_t ldo, const char* name);

// C = alpha A B^T + beta C.  Products with few output tiles and a long K (the
// weight gradients, K = pixels: K >= 1024, slices of at least 256; a small
// batch's dense layers, M <= 64: K >= 256, slices of at least 64) are split
// along K so that the launch fills the chip: about 1024 workgroups, summed in
// order.  A workgroup's K-steps run one after another, so a few tiles with a
// long K would leave the launch bound by the latency of its loads.
int gemm(sfh_ctx* ctx, int dtype, const void* A, int64_t lda, const void* B, int64_t ldb, float* C, int64_t ldc,
         int64_t M, int64_t N, int64_t K, float alpha, float beta, float** slices = nullptr,
         int64_t* nslices = nullptr)
{
    // with slices: a split product's raw slices (splits x M N) are returned
    // unsummed for the caller to add (alpha 1, beta 0 then), else C is written
    if (nslices) *nslices = 1;
    if (M == 0 || N == 0) return SFH_OK;
    const bool narrow = N <= 32;
    const int64_t tm = narrow ? 256 : 128, tn = narrow ? 32 : 128;
    const int64_t tiles = cdiv(N, tn) * cdiv(M, tm);
    int64_t splits = 1, kchunk = std::max<int64_t>(K, 16);
    const int64_t min_slice = M <= 64 ? 64 : 256;
    if (tiles < 512 && K >= 4 * min_slice) {
        splits = std::min<int64_t>(std::max<int64_t>(1, 1024 / tiles), K / min_slice);
        kchunk = cdiv(cdiv(K, splits), 16) * 16;
        splits = cdiv(K, kchunk);
    }
    float* out = C;
    int64_t ldo = ldc;
    if (splits > 1) {
        void* p = nullptr;
        if (int rc = scratch(ctx, "nn_splitk", (size_t)splits * M * N * sizeof(float), &p)) return rc;
        out = (float*)p;
        ldo = N;
    }
    dim3 grid((unsigned)cdiv(N, tn), (unsigned)cdiv(M, tm), (unsigned)splits);
    // slices hold raw sums; alpha and beta are applied when they are added
    const float ka = splits > 1 ? 1.f : alpha, kb = splits > 1 ? 0.f : beta;
    {
        TimedLaunch tl(ctx, "nn_gemm");
        if (dtype == 0) {
            const uint16_t *a = (const uint16_t*)A, *b = (const uint16_t*)B;
            if (narrow)
                hipLaunchKernelGGL((k_gemm_nt<uint16_t, 4, 1, 2, 1>), grid, dim3(kWG), 0, ctx->stream, a, lda, b, ldb,
                                   out, ldo, M, N, K, ka, kb, kchunk);
            else
                hipLaunchKernelGGL((k_gemm_nt<uint16_t, 2, 2, 2, 2>), grid, dim3(kWG), 0, ctx->stream, a, lda, b, ldb,
                                   out, ldo, M, N, K, ka, kb, kchunk);
        } else {
            const float *a = (const float*)A, *b = (const float*)B;
            if (narrow)
                hipLaunchKernelGGL((k_gemm_nt<float, 4, 1, 2, 1>), grid, dim3(kWG), 0, ctx->stream, a, lda, b, ldb,
                                   out, ldo, M, N, K, ka, kb, kchunk);
            else
                hipLaunchKernelGGL((k_gemm_nt<float, 2, 2, 2, 2>), grid, dim3(kWG), 0, ctx->stream, a, lda, b, ldb,
                                   out, ldo, M, N, K, ka, kb, kchunk);
        }
    }
    if (int rc = check_launch(ctx, "k_gemm_nt")) return rc;
    // the slices are the rows of a splits x (M N) matrix
    if (splits > 1 && slices) {
        *slices = out;
        *nslices = splits;
        return SFH_OK;
    }
    if (splits > 1) return colsum(ctx, out, M * N, splits, M * N, alpha, beta, C, N, ldc, "nn_splitk_part");
    return SFH_OK;
}

// Column sums of X (rows x cols, ld) into out[(c / N) ldo + c % N] with alpha
// and beta; N = ldo = cols for a plain vector.
int colsum(sfh_ctx* ctx, const float* X, int64_t ld, int64_t rows, int64_t cols, float alpha, float beta, float* out,
           int64_t N, int64_t ldo, const char* name)
{
    if (cols == 0) return SFH_OK;
    if (rows <= 256) {
        // few rows (bias gradients of a batch, short split-K sums): one pass
        hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(cols, 64)), dim3(kWG), 0, ctx->stream, X, rows, ld,
                           cols, alpha, beta, out, N, ldo);
        return check_launch(ctx, "k_colsum");
    }
    const int64_t nparts = std::min<int64_t>(std::max<int64_t>(1, cdiv(rows, 64)), 1024);
    const int64_t chunk = std::max<int64_t>(1, cdiv(rows, nparts));
    void* p = nullptr;
    if (int rc = scratch(ctx, name, (size_t)nparts * cols * sizeof(float), &p)) return rc;
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)cdiv(cols, 64), (unsigned)nparts), dim3(kWG), 0, ctx->stream, X,
                       ld, rows, cols, chunk, (float*)p);
    hipLaunchKernelGGL(k_colsum_final, dim3((unsigned)cdiv(cols, 64)), dim3(kWG), 0, ctx->stream, (const float*)p,
                       nparts, cols, cols, alpha, beta, out, N, ldo);
    return check_launch(ctx, "k_colsum");
}

}  // namespace

// The direct output-layer kernels: 32 input channels, 16-byte aligned
// tensors, 32-bit pixel coordinates.
static bool conv1_direct(int64_t B, int64_t H, int64_t W, int64_t Ci, const float* in, const float* w,
                         const float* din = nullptr)
{
    const auto al = [](const float* q) { return !q || ((uintptr_t)q & 15) == 0; };
    // one input image (the forward's buffer loads: 32-bit byte offsets) and the
    // whole output gradient (the weight gradient's) under 2 GiB
    return Ci == kC1Ci && al(in) && al(w) && al(din) && B >= 1 && H * W * kC1Ci * 4 < ((int64_t)1 << 31) &&
           B * H * W * 16 < ((int64_t)1 << 31) &&
           B * H * W * kC1Ci < ((int64_t)1 << 40) && B * cdiv(H, kC1TY) * cdiv(W, kC1TX) < ((int64_t)1 << 31);
}

extern "C" {

int sfh_nn_gemm_nt(sfh_ctx* ctx, int dtype, const void* A, int64_t lda, const void* B, int64_t ldb, float* C,
                   int64_t ldc, int64_t M, int64_t N, int64_t K, float alpha, float beta)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, (dtype == 0 || dtype == 1) && K % 16 == 0 && lda >= K && ldb >= K && ldc >= N && M >= 0 && N >= 0,
                  "sfh_nn_gemm_nt: dtype 0 (BF16) or 1 (FP32), K a multiple of 16, lda, ldb >= K, ldc >= N");
    SFH_ARG_CHECK(ctx, lda % 8 == 0 && ldb % 8 == 0, "sfh_nn_gemm_nt: lda, ldb multiples of 8");
    return gemm(ctx, dtype, A, lda, B, ldb, C, ldc, M, N, K, alpha, beta);
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
                           cols, act, pre, (uint16_t*)Y, ldy, (uint16_t*)nullptr, (int64_t)0);
    else
        hipLaunchKernelGGL(k_bias_act<float>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, X, ldx, bias, rows, cols,
                           act, pre, (float*)Y, ldy, (float*)nullptr, (int64_t)0);
    return check_launch(ctx, "k_bias_act");
}

int sfh_nn_linear(sfh_ctx* ctx, int dtype, const void* A, int64_t lda, const void* W, int64_t ldw, int64_t M,
                  int64_t N, int64_t K, const float* bias, int act, float* z, float* pre, void* Y, int64_t ldy,
                  void* YT, int64_t ldt)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, A && W && z && Y && (dtype == 0 || dtype == 1) && (act == 0 || act == 1) && M >= 0 && N >= 0 &&
                           K >= 0 && K % 16 == 0 && lda >= K && ldw >= K && lda % 8 == 0 && ldw % 8 == 0 &&
                           ldy >= N && (!YT || ldt >= M),
                  "sfh_nn_linear: dtype 0 or 1, act 0 or 1, K a multiple of 16, lda and ldw >= K (multiples of 8), "
                  "ldy >= N, ldt >= M");
    if (M == 0) return SFH_OK;
    float* slices = nullptr;
    int64_t ns = 1;
    if (int rc = gemm(ctx, dtype, A, lda, W, ldw, z, N, M, N, K, 1.f, 0.f, &slices, &ns)) return rc;
    const int64_t n = M * ldy;
    if (ns > 1) {
        if (dtype == 0)
            hipLaunchKernelGGL(k_splitk_bias_act<uint16_t>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream,
                               (const float*)slices, ns, M, N, bias, act, pre, (uint16_t*)Y, ldy, (uint16_t*)YT, ldt);
        else
            hipLaunchKernelGGL(k_splitk_bias_act<float>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream,
                               (const float*)slices, ns, M, N, bias, act, pre, (float*)Y, ldy, (float*)YT, ldt);
        return check_launch(ctx, "k_splitk_bias_act");
    }
    if (dtype == 0)
        hipLaunchKernelGGL(k_bias_act<uint16_t>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, (const float*)z, N, bias,
                           M, N, act, pre, (uint16_t*)Y, ldy, (uint16_t*)YT, ldt);
    else
        hipLaunchKernelGGL(k_bias_act<float>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, (const float*)z, N, bias, M,
                           N, act, pre, (float*)Y, ldy, (float*)YT, ldt);
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

int sfh_nn_act_bwd_chansum(sfh_ctx* ctx, const float* dY, const float* pre, int64_t P, int act, float* dX,
                           float* dsum)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, dY && dX && dsum && P >= 0 && (act == 0 || (act == 1 && pre)) && aligned16(dY) &&
                           aligned16(dX) && (!pre || aligned16(pre)),
                  "sfh_nn_act_bwd_chansum: bad arguments");
    const int64_t nwg = chan_grid(P);
    void* cp = nullptr;
    if (int rc = scratch(ctx, "nn_chan_part", (size_t)nwg * 32 * sizeof(float), &cp)) return rc;
    hipLaunchKernelGGL(k_act_bwd_chan32, dim3((unsigned)nwg), dim3(kWG), 0, ctx->stream, dY, pre, P, act, dX,
                       (float*)cp);
    if (int rc = check_launch(ctx, "k_act_bwd_chan32")) return rc;
    hipLaunchKernelGGL(k_chan_final32, dim3(1), dim3(kChanFinWG), 0, ctx->stream, (const float*)cp, nwg, dsum);
    return check_launch(ctx, "k_chan_final32");
}

int sfh_nn_dense_bwd_prep(sfh_ctx* ctx, int dtype, const float* d, const float* pre, int64_t rows, int64_t cols,
                          int act, float* gb, void* dl, void* dT, int64_t ldt)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, d && gb && dl && dT && (act == 0 || (act == 1 && pre)) && rows >= 0 && cols >= 0 &&
                           ldt >= rows && ldt % 8 == 0 && ((uintptr_t)dT & 15u) == 0,
                  "sfh_nn_dense_bwd_prep: bad arguments");
    if (cols == 0) return SFH_OK;
    const dim3 grid((unsigned)cdiv(cols, kPrepWG));
    if (dtype == 0)
        hipLaunchKernelGGL(k_dense_bwd_prep<uint16_t>, grid, dim3(kPrepWG), 0, ctx->stream, d, pre, rows, cols, act,
                           gb, (uint16_t*)dl, (uint16_t*)dT, ldt);
    else
        hipLaunchKernelGGL(k_dense_bwd_prep<float>, grid, dim3(kPrepWG), 0, ctx->stream, d, pre, rows, cols, act,
                           gb, (float*)dl, (float*)dT, ldt);
    return check_launch(ctx, "k_dense_bwd_prep");
}

int sfh_nn_weight_copies(sfh_ctx* ctx, int dtype, int64_t n, const sfh_nn_wcopy* mats)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, (dtype == 0 || dtype == 1) && n >= 0 && n <= kMaxWCopies && (n == 0 || mats),
                  "sfh_nn_weight_copies: dtype 0 or 1, at most %d matrices", kMaxWCopies);
    WCopies wc{};
    wc.n = (int)n;
    int64_t tiles = 0;
    for (int64_t m = 0; m < n; ++m) {
        SFH_ARG_CHECK(ctx, mats[m].src && mats[m].dst && mats[m].rows >= 1 && mats[m].cols >= 1,
                      "sfh_nn_weight_copies: matrix %lld needs src, dst and a positive shape", (long long)m);
        wc.first[m] = tiles;
        wc.src[m] = mats[m].src;
        wc.rows[m] = mats[m].rows;
        wc.cols[m] = mats[m].cols;
        wc.dst[m] = mats[m].dst;
        wc.dstT[m] = mats[m].dstT;
        tiles += cdiv(mats[m].rows, 32) * cdiv(mats[m].cols, 32);
    }
    wc.first[n] = tiles;
    if (tiles == 0) return SFH_OK;
    SFH_ARG_CHECK(ctx, tiles < ((int64_t)1 << 31), "sfh_nn_weight_copies: too many tiles");
    if (dtype == 0)
        hipLaunchKernelGGL(k_weight_copies<uint16_t>, dim3((unsigned)tiles), dim3(kWG), 0, ctx->stream, wc);
    else
        hipLaunchKernelGGL(k_weight_copies<float>, dim3((unsigned)tiles), dim3(kWG), 0, ctx->stream, wc);
    return check_launch(ctx, "k_weight_copies");
}

int sfh_nn_colsum(sfh_ctx* ctx, const float* X, int64_t ld, int64_t rows, int64_t cols, float beta, float* out)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, X && out && rows >= 0 && cols >= 0 && ld >= cols, "sfh_nn_colsum: bad arguments");
    return colsum(ctx, X, ld, rows, cols, 1.f, beta, out, cols, cols, "nn_colsum");
}

int sfh_nn_col2im(sfh_ctx* ctx, const float* cols, int64_t ldc, int64_t B, int64_t H, int64_t W, int64_t Co,
                  const float* bias, float* out)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, ldc >= 16 * Co, "sfh_nn_col2im: ldc >= 16 Co");
    const int64_t n = B * 4 * H * W * Co;
    if (Co % 4 == 0 && ldc % 4 == 0 && B * H * W * ldc < ((int64_t)1 << 31) && n < ((int64_t)1 << 31) &&
        aligned16(cols) && aligned16(out) && (!bias || aligned16(bias))) {
        hipLaunchKernelGGL(k_col2im4, dim3(grid_for(n / 4)), dim3(kWG), 0, ctx->stream, cols, (int)ldc, (int)B,
                           (int)H, (int)W, (int)Co, bias, out);
        return check_launch(ctx, "k_col2im4");
    }
    hipLaunchKernelGGL(k_col2im, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, cols, ldc, B, H, W, Co, bias,
                       (const float*)nullptr, out);
    return check_launch(ctx, "k_col2im");
}

int sfh_nn_im2col(sfh_ctx* ctx, int dtype, const float* dout, int64_t B, int64_t H, int64_t W, int64_t Co,
                  void* dcols, int64_t ldd)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, ldd >= 16 * Co, "sfh_nn_im2col: ldd >= 16 Co");
    const int64_t n = B * H * W * ldd;
    if (Co % 4 == 0 && ldd % 4 == 0 && n < ((int64_t)1 << 31) && aligned16(dout) && aligned16(dcols)) {
        const int64_t n4 = n / 4;
        if (dtype == 0)
            hipLaunchKernelGGL(k_im2col4<uint16_t>, dim3(grid_for(n4)), dim3(kWG), 0, ctx->stream, dout, (int)B,
                               (int)H, (int)W, (int)Co, (uint16_t*)dcols, (int)ldd);
        else
            hipLaunchKernelGGL(k_im2col4<float>, dim3(grid_for(n4)), dim3(kWG), 0, ctx->stream, dout, (int)B, (int)H,
                               (int)W, (int)Co, (float*)dcols, (int)ldd);
        return check_launch(ctx, "k_im2col4");
    }
    if (dtype == 0)
        hipLaunchKernelGGL(k_im2col<uint16_t>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, dout, B, H, W, Co,
                           (uint16_t*)dcols, ldd);
    else
        hipLaunchKernelGGL(k_im2col<float>, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, dout, B, H, W, Co,
                           (float*)dcols, ldd);
    return check_launch(ctx, "k_im2col");
}

int sfh_nn_im2col_t(sfh_ctx* ctx, int dtype, const float* dout, int64_t B, int64_t H, int64_t W, int64_t Co,
                    void* dcolsT, int64_t ldo)
{
    SFH_NN_BEGIN(ctx);
    const int64_t P = B * H * W;
    SFH_ARG_CHECK(ctx, dout && dcolsT && (Co == 4 || Co == 8 || Co == 16 || Co == 32) && ldo >= P &&
                       16 * Co * ldo < ((int64_t)1 << 31) && aligned16(dout),
                  "sfh_nn_im2col_t: Co in {4, 8, 16, 32}, ldo >= B H W");
    const unsigned gy = (unsigned)cdiv(16, 256 / Co);
    if (dtype == 0)
        hipLaunchKernelGGL((k_im2col_t<uint16_t, 64>), dim3((unsigned)cdiv(ldo, 64), gy), dim3(kWG), 0, ctx->stream,
                           dout, (int)B, (int)H, (int)W, (int)Co, (uint16_t*)dcolsT, (int)ldo);
    else
        hipLaunchKernelGGL((k_im2col_t<float, 32>), dim3((unsigned)cdiv(ldo, 32), gy), dim3(kWG), 0, ctx->stream,
                           dout, (int)B, (int)H, (int)W, (int)Co, (float*)dcolsT, (int)ldo);
    return check_launch(ctx, "k_im2col_t");
}

// The single-channel output layer as the other layers: a matrix product to
// the 16 taps of each input pixel, then the col2im gather (with the per-pixel
// bias map); FP32 operands on the matrix cores.
// act(in) for the matrix-product form of the output layer (act = identity: in).
static int conv1_activated(sfh_ctx* ctx, const float* in, int64_t P, int64_t Ci, int act, const float** out)
{
    *out = in;
    if (act == 0) return SFH_OK;
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_conv1_act", (size_t)P * Ci * sizeof(float), &p)) return rc;
    hipLaunchKernelGGL(k_bias_act<float>, dim3(grid_for(P * Ci)), dim3(kWG), 0, ctx->stream, in, Ci,
                       (const float*)nullptr, P, Ci, act, (float*)nullptr, (float*)p, Ci, (float*)nullptr, (int64_t)0);
    *out = (const float*)p;
    return check_launch(ctx, "k_bias_act");
}

int sfh_nn_conv1_fwd(sfh_ctx* ctx, const float* in, int64_t B, int64_t H, int64_t W, int64_t Ci, const float* w,
                     const float* bmap, int act, float* out)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, in && w && out && B >= 0 && H >= 1 && W >= 1 && Ci >= 16 && Ci % 16 == 0 &&
                           (act == 0 || act == 1),
                  "sfh_nn_conv1_fwd: Ci a positive multiple of 16, act 0 (identity) or 1 (SiLU)");
    const int64_t P = B * H * W;
    if (P == 0) return SFH_OK;
    if (conv1_direct(B, H, W, Ci, in, w)) {
        // two workgroups per CU (their LDS), each walking tiles
        const int64_t ntiles = B * cdiv(H, kC1TY) * cdiv(W, kC1TX);
        int cus = 0;
        SFH_HIP_TRY(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        const int64_t grid = std::min<int64_t>(ntiles, 2 * (int64_t)std::max(cus, 1));
        hipLaunchKernelGGL(k_conv1_fwd, dim3((unsigned)grid), dim3(kWG), 0, ctx->stream, in, (int)B, (int)H, (int)W,
                           w, bmap, act, ntiles, out);
        return check_launch(ctx, "k_conv1_fwd");
    }
    if (int rc = conv1_activated(ctx, in, P, Ci, act, &in)) return rc;
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_conv1_cols", (size_t)P * 16 * sizeof(float), &p)) return rc;
    float* cols = (float*)p;
    if (int rc = gemm(ctx, 1, in, Ci, w, Ci, cols, 16, P, 16, Ci, 1.f, 0.f)) return rc;
    const int64_t n = P * 4;
    hipLaunchKernelGGL(k_col2im, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, (const float*)cols, (int64_t)16, B, H,
                       W, (int64_t)1, (const float*)nullptr, bmap, out);
    return check_launch(ctx, "k_col2im");
}

// Backward (matrix-product form): dcols = im2col(dout) (P x 16); din = dcols w
// (P x Ci), times act'(in); dw = dcols^T act(in) (16 x Ci, K = pixels, split
// along K); dbmap = sum over b of dout.
int sfh_nn_conv1_bwd(sfh_ctx* ctx, const float* dout, const float* in, int64_t B, int64_t H, int64_t W, int64_t Ci,
                     const float* w, int act, float* din, float* dsum, float* dw, float* dbmap)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, dout && B >= 0 && H >= 1 && W >= 1 && Ci >= 16 && Ci % 16 == 0 && (!din || w) &&
                           (!dw || in) && (act == 0 || act == 1) && (act == 0 || !din || in) && (!dsum || din),
                  "sfh_nn_conv1_bwd: Ci a positive multiple of 16, w with din, in with dw (and with din when act "
                  "is 1), act 0 or 1");
    const int64_t P = B * H * W, Pp = cdiv(P, 16) * 16;
    if (P == 0) return SFH_OK;
    void* p = nullptr;
    if ((din || dw) && conv1_direct(B, H, W, Ci, (dw || act) ? in : nullptr, din ? w : nullptr, din)) {
        if (din) {
            const int64_t nwg = chan_grid(P);
            void* cp = nullptr;
            if (dsum)
                if (int rc = scratch(ctx, "nn_chan_part", (size_t)nwg * 32 * sizeof(float), &cp)) return rc;
            hipLaunchKernelGGL(k_conv1_din, dim3((unsigned)nwg), dim3(kWG), 0, ctx->stream, dout, (uint32_t)(P * 16),
                               in, (int)B, (int)H, (int)W, 1.0 / (double)W, 1.0 / (double)H, w, act, din,
                               (float*)cp);
            if (int rc = check_launch(ctx, "k_conv1_din")) return rc;
            if (dsum) {
                hipLaunchKernelGGL(k_chan_final32, dim3(1), dim3(kChanFinWG), 0, ctx->stream, (const float*)cp, nwg,
                                   dsum);
                if (int rc = check_launch(ctx, "k_chan_final32")) return rc;
            }
        }
        if (dw) {
            // one partial per workgroup (at most 256: a single-pass column sum)
            int64_t nwg = std::min<int64_t>(256, cdiv(P, kC1DwWG / 8));
            const int64_t per = cdiv(P, nwg);
            nwg = cdiv(P, per);
            if (int rc = scratch(ctx, "nn_conv1_dw", (size_t)nwg * 16 * kC1Ci * sizeof(float), &p)) return rc;
            hipLaunchKernelGGL(k_conv1_dw, dim3((unsigned)nwg), dim3(kC1DwWG), 0, ctx->stream, dout,
                               (uint32_t)(P * 16), in, (int)B, (int)H, (int)W, 1.0 / (double)W, 1.0 / (double)H, per,
                               act, (float*)p);
            if (int rc = check_launch(ctx, "k_conv1_dw")) return rc;
            if (int rc = colsum(ctx, (const float*)p, 16 * kC1Ci, nwg, 16 * kC1Ci, 1.f, 0.f, dw, 16 * kC1Ci,
                                16 * kC1Ci, "nn_conv1_dw_sum"))
                return rc;
        }
    } else if (din || dw) {
        const float* pre = in;
        if (dw)
            if (int rc = conv1_activated(ctx, pre, P, Ci, act, &in)) return rc;
        if (int rc = scratch(ctx, "nn_conv1_cols", (size_t)P * 16 * sizeof(float), &p)) return rc;
        float* dcols = (float*)p;
        hipLaunchKernelGGL(k_im2col<float>, dim3(grid_for(P * 16)), dim3(kWG), 0, ctx->stream, dout, B, H, W,
                           (int64_t)1, dcols, (int64_t)16);
        if (int rc = check_launch(ctx, "k_im2col")) return rc;
        if (din) {
            // din = dcols (P x 16) times w (16 x Ci): B operand w^T (Ci x 16)
            void* q = nullptr;
            if (int rc = scratch(ctx, "nn_conv1_wT", (size_t)16 * Ci * sizeof(float), &q)) return rc;
            hipLaunchKernelGGL(k_transpose<float>, dim3(1, (unsigned)cdiv(Ci, 32)), dim3(kWG), 0, ctx->stream, w, Ci,
                               (int64_t)16, Ci, (float*)q, (int64_t)16, Ci, 0);
            if (int rc = gemm(ctx, 1, dcols, 16, q, 16, din, Ci, P, Ci, 16, 1.f, 0.f)) return rc;
            if (act) {
                hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(P * Ci)), dim3(kWG), 0, ctx->stream, (const float*)din, pre,
                                   P * Ci, act, din);
                if (int rc = check_launch(ctx, "k_act_bwd")) return rc;
            }
            if (dsum)
                if (int rc = colsum(ctx, din, Ci, P, Ci, 1.f, 0.f, dsum, Ci, Ci, "nn_conv1_dsum")) return rc;
        }
        if (dw) {
            // dw = dcols^T (16 x P) times in (P x Ci): operands dcols^T and in^T, K = Pp (zero padded)
            void* q = nullptr;
            if (int rc = scratch(ctx, "nn_conv1_T", (size_t)(16 + Ci) * Pp * sizeof(float), &q)) return rc;
            float* colsT = (float*)q;
            float* inT = colsT + 16 * Pp;
            hipLaunchKernelGGL(k_transpose<float>, dim3((unsigned)cdiv(Pp, 32), 1), dim3(kWG), 0, ctx->stream,
                               (const float*)dcols, (int64_t)16, P, (int64_t)16, colsT, Pp, (int64_t)16, 0);
            hipLaunchKernelGGL(k_transpose<float>, dim3((unsigned)cdiv(Pp, 32), (unsigned)cdiv(Ci, 32)), dim3(kWG), 0,
                               ctx->stream, in, Ci, P, Ci, inT, Pp, Ci, 0);
            if (int rc = check_launch(ctx, "k_transpose")) return rc;
            if (int rc = gemm(ctx, 1, colsT, Pp, inT, Pp, dw, Ci, 16, Ci, Pp, 1.f, 0.f)) return rc;
        }
    }
    if (dbmap)
        if (int rc = colsum(ctx, dout, 4 * H * W, B, 4 * H * W, 1.f, 0.f, dbmap, 4 * H * W, 4 * H * W,
                            "nn_colsum_map"))
            return rc;
    return SFH_OK;
}

int sfh_nn_conv_t_dinput(sfh_ctx* ctx, int dtype, const float* d, int64_t B, int64_t Hi, int64_t Wi, int64_t Co,
                         const void* wT, int64_t ldw, int64_t Ci, float* out, int64_t ldo)
{
    SFH_NN_BEGIN(ctx);
    const int64_t P = B * Hi * Wi;
    SFH_ARG_CHECK(ctx, d && wT && out && (dtype == 0 || dtype == 1) && B >= 0 && Hi >= 1 && Wi >= 1 &&
                           (Co == 8 || Co == 16 || Co == 32) && Ci >= 1 && ldw >= 16 * Co && ldw % 8 == 0 && ldo >= Ci &&
                           aligned16(d) && aligned16(wT) && P < ((int64_t)1 << 31) &&
                           16 * Hi * Wi * Co < ((int64_t)1 << 31),
                  "sfh_nn_conv_t_dinput: Co 8, 16 or 32, ldw >= 16 Co (a multiple of 8), ldo >= Ci, "
                  "16-byte aligned d and wT, B Hi Wi < 2^31, one gradient image under 2 GiB");
    if (P == 0) return SFH_OK;
    const double iw = 1.0 / (double)Wi, ih = 1.0 / (double)Hi;
    // batch chunks whose gradient images span less than 2^31 bytes (the gathers'
    // 32-bit buffer offsets); images are independent, so chunking is exact
    const int64_t img_bytes = 16 * Hi * Wi * Co, nb = std::max<int64_t>(1, (((int64_t)1 << 31) - 1) / img_bytes);
    for (int64_t b0 = 0; b0 < B; b0 += nb) {
        const int64_t bn = std::min(nb, B - b0), Pc = bn * Hi * Wi;
        const float* dc = d + b0 * (img_bytes / 4);
        float* oc = out + b0 * Hi * Wi * ldo;
        const uint32_t nbytes = (uint32_t)(bn * img_bytes);
        // fewer workgroups than CUs: four slices of the taps, each a quarter of
        // the K-steps one workgroup runs in sequence, summed in slice order
        const int64_t wgs = cdiv(Pc, 256) * cdiv(Ci, 32);
        const int S = wgs < 256 ? 4 : 1;
        float* dst = oc;
        int64_t ldd = ldo;
        if (S > 1) {
            void* p = nullptr;
            if (int rc = scratch(ctx, "nn_dinput_slices", (size_t)S * Pc * Ci * sizeof(float), &p)) return rc;
            dst = (float*)p;
            ldd = Ci;
        }
        const dim3 grid((unsigned)cdiv(Pc, 256), (unsigned)cdiv(Ci, 32), (unsigned)S);
#define SFH_DIN(TT, CC)                                                                                                \
    do {                                                                                                               \
        if (S > 1)                                                                                                     \
            hipLaunchKernelGGL((k_conv_dinput<TT, CC, 4>), grid, dim3(kWG), 0, ctx->stream, dc, nbytes, (int)bn,        \
                               (int)Hi, (int)Wi, iw, ih, (const TT*)wT, ldw, Ci, dst, ldd);                             \
        else                                                                                                           \
            hipLaunchKernelGGL((k_conv_dinput<TT, CC, 1>), grid, dim3(kWG), 0, ctx->stream, dc, nbytes, (int)bn,        \
                               (int)Hi, (int)Wi, iw, ih, (const TT*)wT, ldw, Ci, dst, ldd);                             \
    } while (0)
        if (dtype == 0) {
            if (Co == 8) SFH_DIN(uint16_t, 8); else if (Co == 16) SFH_DIN(uint16_t, 16); else SFH_DIN(uint16_t, 32);
        } else {
            if (Co == 8) SFH_DIN(float, 8); else if (Co == 16) SFH_DIN(float, 16); else SFH_DIN(float, 32);
        }
#undef SFH_DIN
        if (int rc = check_launch(ctx, "k_conv_dinput")) return rc;
        if (S > 1)
            if (int rc = colsum(ctx, dst, Pc * Ci, S, Pc * Ci, 1.f, 0.f, oc, Ci, ldo, "nn_dinput_sum")) return rc;
    }
    return SFH_OK;
}

int sfh_nn_conv_t_fwd(sfh_ctx* ctx, int dtype, const void* in, int64_t B, int64_t Hi, int64_t Wi, int64_t C,
                      const void* w, const float* bias, float* pre, void* aout)
{
    SFH_NN_BEGIN(ctx);
    const int64_t P = B * Hi * Wi;
    SFH_ARG_CHECK(ctx, in && w && pre && (dtype == 0 || dtype == 1) && C == 32 && B >= 0 && Hi >= 1 && Wi >= 1 &&
                           aligned16(in) && aligned16(w) && P < ((int64_t)1 << 31) &&
                           4 * Hi * Wi * C < ((int64_t)1 << 31),
                  "sfh_nn_conv_t_fwd: 32 channels, 16-byte aligned in and w, B Hi Wi < 2^31, one input image "
                  "under 2 GiB");
    if (P == 0) return SFH_OK;
    // batch chunks whose input images span less than 2^31 bytes (the gathers'
    // 32-bit buffer offsets); images are independent, so chunking is exact
    const int64_t esz = dtype == 0 ? 2 : 4, img = Hi * Wi * C;
    const int64_t nb = std::max<int64_t>(1, (((int64_t)1 << 31) - 1) / (img * esz));
    for (int64_t b0 = 0; b0 < B; b0 += nb) {
        const int64_t bn = std::min(nb, B - b0), Pc = bn * Hi * Wi;
        const uint32_t nbytes = (uint32_t)(bn * img * esz);
        const int64_t oo = b0 * 4 * img;    // output image offset (2Hi x 2Wi x C)
        const dim3 grid((unsigned)cdiv(Pc, 256), 1, 4);
        if (dtype == 0)
            hipLaunchKernelGGL(k_conv_fwd<uint16_t>, grid, dim3(kWG), 0, ctx->stream, (const uint16_t*)in + b0 * img,
                               nbytes, (int)bn, (int)Hi, (int)Wi, 1.0 / (double)Wi, 1.0 / (double)Hi,
                               (const uint16_t*)w, bias, pre + oo, aout ? (uint16_t*)aout + oo : nullptr);
        else
            hipLaunchKernelGGL(k_conv_fwd<float>, grid, dim3(kWG), 0, ctx->stream, (const float*)in + b0 * img, nbytes,
                               (int)bn, (int)Hi, (int)Wi, 1.0 / (double)Wi, 1.0 / (double)Hi, (const float*)w, bias,
                               pre + oo, aout ? (float*)aout + oo : nullptr);
        if (int rc = check_launch(ctx, "k_conv_fwd")) return rc;
    }
    return SFH_OK;
}

int sfh_nn_conv_t_dweight(sfh_ctx* ctx, int dtype, const float* d, const float* pre, int act, int64_t B, int64_t Hi,
                          int64_t Wi, int64_t C, float* dK)
{
    SFH_NN_BEGIN(ctx);
    const int64_t P = B * Hi * Wi;
    SFH_ARG_CHECK(ctx, d && pre && dK && (dtype == 0 || dtype == 1) && (act == 0 || act == 1) && C == kDwC &&
                           B >= 0 && Hi >= 1 && Wi >= 1 && aligned16(d) && aligned16(pre) &&
                           P < ((int64_t)1 << 31) && 16 * P * C < ((int64_t)1 << 31),
                  "sfh_nn_conv_t_dweight: 32 channels, act 0 or 1, 16-byte aligned d and pre, the gradient d "
                  "(B x 2Hi x 2Wi x 32 FP32) under 2 GiB");
    constexpr int64_t NOUT = 16 * kDwC * kDwC;
    if (P == 0) {
        hipLaunchKernelGGL(k_fill, dim3(grid_for(NOUT)), dim3(kWG), 0, ctx->stream, dK, NOUT, 0.f);
        return check_launch(ctx, "k_fill");
    }
    // pixel splits: one partial tile set per split, at most 256 (a single-pass column sum)
    int64_t nsplit = std::min<int64_t>(256, cdiv(P, kDwChunk));
    const int64_t per = cdiv(cdiv(P, nsplit), kDwChunk) * kDwChunk;
    nsplit = cdiv(P, per);
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_conv_dw", (size_t)nsplit * NOUT * sizeof(float), &p)) return rc;
    const dim3 grid(4, (unsigned)nsplit);
    if (dtype == 0)
        hipLaunchKernelGGL(k_conv_dweight<uint16_t>, grid, dim3(kWG), 0, ctx->stream, d, (uint32_t)(16 * P * C), pre,
                           act, (int)B, (int)Hi, (int)Wi, 1.0 / (double)Wi, 1.0 / (double)Hi, per, (float*)p);
    else
        hipLaunchKernelGGL(k_conv_dweight<float>, grid, dim3(kWG), 0, ctx->stream, d, (uint32_t)(16 * P * C), pre, act,
                           (int)B, (int)Hi, (int)Wi, 1.0 / (double)Wi, 1.0 / (double)Hi, per, (float*)p);
    if (int rc = check_launch(ctx, "k_conv_dweight")) return rc;
    return colsum(ctx, (const float*)p, NOUT, nsplit, NOUT, 1.f, 0.f, dK, NOUT, NOUT, "nn_conv_dw_sum");
}

int sfh_nn_mse(sfh_ctx* ctx, const float* out, const float* target, const int64_t* rows, int64_t B, int64_t H2,
               int64_t W2, int64_t R, int64_t S, float* dout, float* loss, double* loss_acc)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, R <= H2 && S <= W2 && B >= 1, "sfh_nn_mse: crop within the output, B >= 1");
    const int nb = 512;
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_mse", nb * sizeof(float), &p)) return rc;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(kWG), 0, ctx->stream, out, target, rows, B, H2, W2, R, S, dout,
                       (float*)p);
    if (int rc = check_launch(ctx, "k_mse")) return rc;
    const double denom = (double)(B * R * S);
    if (loss_acc) {
        hipLaunchKernelGGL(k_mse_accumulate, dim3(1), dim3(64), 0, ctx->stream, (const float*)p, nb, denom, loss_acc);
        if (int rc = check_launch(ctx, "k_mse_accumulate")) return rc;
    }
    if (!loss) return SFH_OK;
    float part[nb];
    SFH_HIP_TRY(ctx, hipMemcpyAsync(part, p, nb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    SFH_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    double s = 0.0;
    for (int i = 0; i < nb; ++i) s += part[i];
    *loss = (float)(s / denom);
    return SFH_OK;
}

int sfh_nn_template_loss(sfh_ctx* ctx, const float* out, const float* target, const int64_t* rows, int64_t B,
                         int64_t H2, int64_t W2, int64_t R, int64_t S, float scale, float offset, float w_mse,
                         float w_kl, float w_sum, float* dout, double* acc)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, out && target && dout && B >= 1 && B <= 65535 && R >= 1 && S >= 1 && R <= H2 && S <= W2 &&
                       scale > 0.f,
                  "sfh_nn_template_loss: bad arguments");
    void* p = nullptr;
    if (int rc = scratch(ctx, "nn_tl", (size_t)B * kTLChunks * 4 * sizeof(double), &p)) return rc;
    double* part = (double*)p;
    hipLaunchKernelGGL(k_tl_sums, dim3(kTLChunks, (unsigned)B), dim3(kWG), 0, ctx->stream, out, target, rows, H2, W2,
                       R, S, scale, offset, part);
    hipLaunchKernelGGL(k_tl_grad, dim3((unsigned)std::min<int64_t>(cdiv(H2 * W2, kWG), 64), (unsigned)B), dim3(kWG),
                       0, ctx->stream, out, target, rows, B, H2, W2, R, S, scale, offset, w_mse, w_kl, w_sum,
                       (const double*)part, dout);
    if (acc)
        hipLaunchKernelGGL(k_tl_accumulate, dim3(1), dim3(64), 0, ctx->stream, (const double*)part, B, R, S, scale,
                           acc);
    return check_launch(ctx, "k_tl");
}

int sfh_nn_gather_rows(sfh_ctx* ctx, const float* src, int64_t lds, const int64_t* rows, int64_t n, int64_t cols,
                       float* dst, int64_t ldd)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, src && rows && dst && n >= 0 && cols >= 0 && lds >= cols && ldd >= cols,
                  "sfh_nn_gather_rows: bad arguments");
    if (n == 0 || cols == 0) return SFH_OK;
    hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(n * cols)), dim3(kWG), 0, ctx->stream, src, lds, rows, n, cols,
                       dst, ldd);
    return check_launch(ctx, "k_gather_rows");
}

int sfh_nn_log_target(sfh_ctx* ctx, const float* T, int64_t ldt, int64_t B, int64_t nx, int64_t ny, float eps,
                      float offset, float scale, float* img)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, T && img && B >= 0 && nx >= 1 && ny >= 1 && ldt >= nx * ny && eps > 0.f && scale > 0.f &&
                       B <= 65535,
                  "sfh_nn_log_target: bad arguments");
    if (B == 0) return SFH_OK;
    hipLaunchKernelGGL(k_log_target, dim3((unsigned)cdiv(nx, kTile), (unsigned)cdiv(ny, kTile), (unsigned)B),
                       dim3(kWG), 0, ctx->stream, T, ldt, nx, ny, eps, offset, 1.f / scale, img);
    return check_launch(ctx, "k_log_target");
}

int sfh_nn_log_template(sfh_ctx* ctx, const float* out, int64_t B, int64_t H2, int64_t W2, int64_t nx, int64_t ny,
                        float eps, float offset, float scale, float* T, int64_t ldt)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, out && T && B >= 0 && nx >= 1 && ny >= 1 && nx <= H2 && ny <= W2 && ldt >= nx * ny &&
                       B <= 65535,
                  "sfh_nn_log_template: bad arguments");
    if (B == 0) return SFH_OK;
    hipLaunchKernelGGL(k_log_template, dim3((unsigned)cdiv(nx, kTile), (unsigned)cdiv(ny, kTile), (unsigned)B),
                       dim3(kWG), 0, ctx->stream, out, H2, W2, nx, ny, eps, offset, scale, T, ldt);
    return check_launch(ctx, "k_log_template");
}

int sfh_nn_template_kl(sfh_ctx* ctx, const float* E, int64_t lde, const float* S, int64_t lds, int64_t n,
                       int64_t nbins, double floor_value, double* out)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, E && S && out && n >= 0 && nbins >= 1 && lde >= nbins && lds >= nbins && floor_value > 0.0 &&
                       n <= 2147483647,
                  "sfh_nn_template_kl: bad arguments");
    if (n == 0) return SFH_OK;
    hipLaunchKernelGGL(k_template_kl, dim3((unsigned)n), dim3(kWG), 0, ctx->stream, E, lde, S, lds, nbins,
                       floor_value, out);
    return check_launch(ctx, "k_template_kl");
}

int sfh_nn_select_rows(sfh_ctx* ctx, const int64_t* rows_all, const int64_t* counter, int64_t n, int64_t* rows_step)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, rows_all && counter && rows_step && n >= 1, "sfh_nn_select_rows: bad arguments");
    hipLaunchKernelGGL(k_select_rows, dim3(1), dim3(64), 0, ctx->stream, rows_all, counter, n, rows_step);
    return check_launch(ctx, "k_select_rows");
}

int sfh_nn_counter_add(sfh_ctx* ctx, int64_t* counter, int64_t inc)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, counter, "sfh_nn_counter_add: null counter");
    hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(64), 0, ctx->stream, counter, inc);
    return check_launch(ctx, "k_counter_add");
}

int sfh_nn_adam_sched(sfh_ctx* ctx, float* p, const float* g, float* m, float* v, int64_t n, const float* sched,
                      const int64_t* counter, float b1, float b2, float eps, float weight_decay)
{
    SFH_NN_BEGIN(ctx);
    SFH_ARG_CHECK(ctx, p && g && m && v && sched && counter, "sfh_nn_adam_sched: null pointer");
    hipLaunchKernelGGL(k_adam_sched, dim3(grid_for(n)), dim3(kWG), 0, ctx->stream, p, g, m, v, n, sched, counter, b1,
                       b2, eps, weight_decay);
    return check_launch(ctx, "k_adam_sched");
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

// BF16 implicit-GEMM 3D convolution for gfx950 MFMA, FP32 accumulation: forward with fused
// bias + GELU, input gradient (dgrad) and weight gradient (wgrad).  No reference
// counterpart: the reference has no neural surrogate (its design loops are
// torchfdtd/design.py and torchfdtd/design_problem.py); these kernels serve the learned
// screening model of the inverse-design loop.
//
// Layouts: activations channels-last [n][d][h][w][c] (bf16); forward weights [co][kp]
// with k = tap*ci + c, tap = (kd*ks + kh)*ks + kw, zero-padded to kp (a multiple of 32);
// dgrad weights [ci][ks^3*co] (k = tap*co + c); wgrad reads dY transposed [co][m].
//
// One workgroup (4 waves) computes a 64 x 64 tile of C = A B over K in steps of 32: the
// A tile (64 rows x 32 k) and B tile (64 columns x 32 k) are gathered from global memory
// (the implicit im2col) into LDS with k fastest, and each wave multiplies its 32 x 32
// quarter with 2 x 2 v_mfma_f32_16x16x32_bf16.  blockIdx.z splits K for the weight
// gradient, whose reduction runs over every output position.
#include <hip/hip_runtime.h>

#include "../../include/tfdtd.h"
#include "common.h"

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;
typedef unsigned short bf16_bits;

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;   // LDS rows padded by 16 bytes

// Round to nearest even (finite inputs), as torch's float -> bfloat16 conversion.
__device__ __forceinline__ bf16_bits float_to_bf16(float f) {
    const unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_bits)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    return (bf16_bits)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

struct Conv {
    int n, di, hi, wi, ci;      // input
    int dout, hout, wout, co;   // output
    int ks, stride, pad;
    int kp;                     // forward K (ks^3 * ci) padded to a multiple of 32
};

struct Gemm {
    int op;                     // 0 forward, 1 dgrad, 2 wgrad
    Conv c;
    int M, N, K;                // C is M x N, reduction K (padded)
    int k_per_split;
    const bf16_bits* a;         // forward: x; dgrad: dY [m_out][co]; wgrad: dY^T [co][m_out]
    const bf16_bits* b;         // forward: W [co][kp]; dgrad: Wd [ci][ks^3 co]; wgrad: x
    const float* bias;          // forward only (may be null)
    int activation;             // forward: 1 GELU, 0 identity
    float* out;                 // forward: pre-activation [M][N]; dgrad: dX [M][N]; wgrad: partial [split][M][N]
    bf16_bits* act;             // forward: bf16 activation [M][N]
};

// Input position of output position m and tap (kd, kh, kw); false when it lies in the padding.
__device__ __forceinline__ bool input_of(const Conv& c, long long m, int tap, long long* cell) {
    const int wo = (int)(m % c.wout);
    const int ho = (int)((m / c.wout) % c.hout);
    const int dd = (int)((m / ((long long)c.wout * c.hout)) % c.dout);
    const long long nb = m / ((long long)c.wout * c.hout * c.dout);
    const int kw = tap % c.ks, kh = (tap / c.ks) % c.ks, kd = tap / (c.ks * c.ks);
    const int z = dd * c.stride - c.pad + kd, y = ho * c.stride - c.pad + kh, x = wo * c.stride - c.pad + kw;
    if (z < 0 || z >= c.di || y < 0 || y >= c.hi || x < 0 || x >= c.wi) return false;
    *cell = ((nb * c.di + z) * c.hi + y) * c.wi + x;
    return true;
}

// Output position that input position m_in reaches through tap; false when none.
__device__ __forceinline__ bool output_of(const Conv& c, long long m, int tap, long long* cell) {
    const int x = (int)(m % c.wi);
    const int y = (int)((m / c.wi) % c.hi);
    const int z = (int)((m / ((long long)c.wi * c.hi)) % c.di);
    const long long nb = m / ((long long)c.wi * c.hi * c.di);
    const int kw = tap % c.ks, kh = (tap / c.ks) % c.ks, kd = tap / (c.ks * c.ks);
    const int tz = z + c.pad - kd, ty = y + c.pad - kh, tx = x + c.pad - kw;
    if (tz < 0 || ty < 0 || tx < 0 || tz % c.stride || ty % c.stride || tx % c.stride) return false;
    const int dd = tz / c.stride, ho = ty / c.stride, wo = tx / c.stride;
    if (dd >= c.dout || ho >= c.hout || wo >= c.wout) return false;
    *cell = ((nb * c.dout + dd) * c.hout + ho) * c.wout + wo;
    return true;
}

__device__ __forceinline__ void load8(bf16_bits* dst, const bf16_bits* src) {
    *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
}

__device__ __forceinline__ void zero8(bf16_bits* dst) { *reinterpret_cast<uint4*>(dst) = make_uint4(0, 0, 0, 0); }

// Element (m, k) of the forward im2col matrix of the input x.
__device__ __forceinline__ bf16_bits im2col(const Conv& c, const bf16_bits* x, long long m, int k) {
    if (k >= c.ks * c.ks * c.ks * c.ci) return 0;
    long long cell;
    if (!input_of(c, m, k / c.ci, &cell)) return 0;
    return x[cell * c.ci + k % c.ci];
}

// A tile: sA[row][k] for rows row0.., k0..k0+31.  Thread t fills row t/4, k chunk (t%4)*8.
__device__ __forceinline__ void fill_a(const Gemm& g, bf16_bits* sA, long long row0, int k0) {
    const int t = threadIdx.x, r = t >> 2, q = (t & 3) * 8;
    bf16_bits* dst = sA + r * LDK + q;
    const long long row = row0 + r;
    const int k = k0 + q;
    const Conv& c = g.c;
    if (row >= g.M || k >= g.K) { zero8(dst); return; }
    if (g.op == 0) {
        if (c.ci % 8 == 0) {
            long long cell;
            if (k < c.ks * c.ks * c.ks * c.ci && input_of(c, row, k / c.ci, &cell)) load8(dst, g.a + cell * c.ci + k % c.ci);
            else zero8(dst);
        } else {
            for (int j = 0; j < 8; ++j) dst[j] = im2col(c, g.a, row, k + j);
        }
    } else if (g.op == 1) {
        // dY gathered at the output position this input position reaches by tap k / co.
        long long cell;
        if (output_of(c, row, k / c.co, &cell)) load8(dst, g.a + cell * c.co + k % c.co);
        else zero8(dst);
    } else {
        // dY^T [co][m]: 8 consecutive output positions of channel `row`.
        const long long m_total = (long long)c.n * c.dout * c.hout * c.wout;
        if (k + 8 <= m_total && (m_total % 8) == 0) load8(dst, g.a + row * m_total + k);
        else for (int j = 0; j < 8; ++j) dst[j] = k + j < m_total ? g.a[row * m_total + k + j] : 0;
    }
}

// B tile: sB[col][k] for columns col0.., k0..k0+31.
__device__ __forceinline__ void fill_b(const Gemm& g, bf16_bits* sB, int col0, int k0) {
    const int t = threadIdx.x;
    const Conv& c = g.c;
    if (g.op != 2) {
        const int r = t >> 2, q = (t & 3) * 8;
        bf16_bits* dst = sB + r * LDK + q;
        const int col = col0 + r, k = k0 + q;
        if (col >= g.N || k >= g.K) { zero8(dst); return; }
        load8(dst, g.b + (long long)col * g.K + k);          // weights padded to K columns
    } else {
        // Weight gradient: B(m, k) = im2col(m, k) for 32 reductions m and 64 columns k.  Thread t
        // takes reduction row t/8 and columns (t%8)*8 .. +7, loading along k and storing transposed.
        const int mr = t >> 3, q = (t & 7) * 8;
        const long long m = (long long)k0 + mr;
        const long long m_total = (long long)c.n * c.dout * c.hout * c.wout;
        const int kk = col0 + q;
        const int kmax = c.ks * c.ks * c.ks * c.ci;
        bf16_bits v[8];
        if (m >= m_total) {
            for (int j = 0; j < 8; ++j) v[j] = 0;
        } else if (c.ci % 8 == 0 && kk + 8 <= kmax) {
            long long cell;
            if (input_of(c, m, kk / c.ci, &cell)) *reinterpret_cast<uint4*>(v) = *reinterpret_cast<const uint4*>(g.b + cell * c.ci + kk % c.ci);
            else for (int j = 0; j < 8; ++j) v[j] = 0;
        } else {
            for (int j = 0; j < 8; ++j) v[j] = im2col(c, g.b, m, kk + j);
        }
        for (int j = 0; j < 8; ++j) sB[(q + j) * LDK + mr] = v[j];
    }
}

__global__ __launch_bounds__(256) void gemm_kernel(const Gemm g) {
    __shared__ __attribute__((aligned(16))) bf16_bits sA[BM * LDK];
    __shared__ __attribute__((aligned(16))) bf16_bits sB[BN * LDK];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long row0 = (long long)blockIdx.x * BM;
    const int col0 = blockIdx.y * BN;
    const int kb = blockIdx.z * g.k_per_split, ke = min(g.K, kb + g.k_per_split);
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    f4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += BK) {
        __syncthreads();
        fill_a(g, sA, row0, k0);
        fill_b(g, sB, col0, k0);
        __syncthreads();
        // Lane l supplies row (l % 16) and k (l / 16) * 8 .. +7 of each 16 x 32 operand.
        const int r = lane & 15, kq = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(sA + (wr + 16 * i + r) * LDK + kq);
        for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(sB + (wc + 16 * j + r) * LDK + kq);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // D layout: lane l holds rows (l / 16) * 4 + e, column l % 16 of each 16 x 16 tile.
    const int cn = lane & 15, rq = (lane >> 4) * 4;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int e = 0; e < 4; ++e) {
                const long long row = row0 + wr + 16 * i + rq + e;
                const int col = col0 + wc + 16 * j + cn;
                if (row >= g.M || col >= g.N) continue;
                float v = acc[i][j][e];
                if (g.op == 2) {
                    g.out[((long long)blockIdx.z * g.M + row) * g.N + col] = v;
                } else if (g.op == 0) {
                    if (g.bias) v += g.bias[col];
                    g.out[row * g.N + col] = v;
                    g.act[row * g.N + col] = float_to_bf16(g.activation ? gelu(v) : v);
                } else {
                    g.out[row * g.N + col] = v;
                }
            }
}

// Sum of the K splits of the weight gradient, in split order.
__global__ __launch_bounds__(256) void split_sum_kernel(const float* partial, int splits, long long count, float* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += partial[z * count + i];
    out[i] = s;
}

// dPre = dY * GELU'(pre) (or dY), written in FP32 and BF16.
__global__ __launch_bounds__(256) void gelu_backward_kernel(const float* dy, const float* pre, long long count, int activation,
                                                           float* dpre, bf16_bits* dpre16) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float v = activation ? dy[i] * gelu_grad(pre[i]) : dy[i];
    dpre[i] = v;
    dpre16[i] = float_to_bf16(v);
}

int check_conv(const tfdtd_conv& d) {
    if (d.n < 1 || d.di < 1 || d.hi < 1 || d.wi < 1 || d.ci < 1 || d.co < 1)
        return tfdtd_fail(TFDTD_EINVAL, "conv: sizes must be positive");
    if ((d.ks != 1 && d.ks != 3) || d.stride < 1 || d.stride > 2 || d.pad < 0 || d.pad >= d.ks)
        return tfdtd_fail(TFDTD_EINVAL, "conv: kernel 1 or 3, stride 1 or 2, 0 <= pad < kernel");
    const int k = d.ks * d.ks * d.ks * d.ci;
    if (d.kp < k || d.kp % 32) return tfdtd_fail(TFDTD_EINVAL, "conv: kp must be ks^3*ci rounded up to a multiple of 32");
    return TFDTD_OK;
}

Conv make_conv(const tfdtd_conv& d) {
    Conv c;
    c.n = d.n; c.di = d.di; c.hi = d.hi; c.wi = d.wi; c.ci = d.ci; c.co = d.co;
    c.ks = d.ks; c.stride = d.stride; c.pad = d.pad; c.kp = d.kp;
    c.dout = (d.di + 2 * d.pad - d.ks) / d.stride + 1;
    c.hout = (d.hi + 2 * d.pad - d.ks) / d.stride + 1;
    c.wout = (d.wi + 2 * d.pad - d.ks) / d.stride + 1;
    return c;
}

}  // namespace

extern "C" {

int tfdtd_conv_output(const tfdtd_conv* d, int32_t* dout, int32_t* hout, int32_t* wout) {
    if (!d) return tfdtd_fail(TFDTD_EINVAL, "conv: null descriptor");
    const int status = check_conv(*d);
    if (status) return status;
    const Conv c = make_conv(*d);
    if (c.dout < 1 || c.hout < 1 || c.wout < 1) return tfdtd_fail(TFDTD_EINVAL, "conv: empty output");
    *dout = c.dout; *hout = c.hout; *wout = c.wout;
    return TFDTD_OK;
}

int tfdtd_conv_forward(const tfdtd_conv* d, const void* x, const void* weight, const float* bias, int32_t activation,
                       float* pre, void* act, void* stream) {
    if (!d || !x || !weight || !pre || !act) return tfdtd_fail(TFDTD_EINVAL, "conv forward: null pointer");
    int status = check_conv(*d);
    if (status) return status;
    Gemm g{};
    g.op = 0;
    g.c = make_conv(*d);
    g.M = (int)((long long)d->n * g.c.dout * g.c.hout * g.c.wout);
    g.N = d->co;
    g.K = d->kp;
    g.k_per_split = g.K;
    g.a = static_cast<const bf16_bits*>(x);
    g.b = static_cast<const bf16_bits*>(weight);
    g.bias = bias;
    g.activation = activation;
    g.out = pre;
    g.act = static_cast<bf16_bits*>(act);
    if ((long long)d->n * g.c.dout * g.c.hout * g.c.wout > 0x7fffffffLL) return tfdtd_fail(TFDTD_EINVAL, "conv: too many positions");
    hipLaunchKernelGGL(gemm_kernel, dim3((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, 1), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g);
    return tfdtd_check_launch("conv forward");
}

int tfdtd_conv_dgrad(const tfdtd_conv* d, const void* dy, const void* weight_t, float* dx, void* stream) {
    if (!d || !dy || !weight_t || !dx) return tfdtd_fail(TFDTD_EINVAL, "conv dgrad: null pointer");
    int status = check_conv(*d);
    if (status) return status;
    if (d->co % 8) return tfdtd_fail(TFDTD_EINVAL, "conv dgrad: output channels must be a multiple of 8");
    Gemm g{};
    g.op = 1;
    g.c = make_conv(*d);
    g.M = (int)((long long)d->n * d->di * d->hi * d->wi);
    g.N = d->ci;
    g.K = d->ks * d->ks * d->ks * d->co;
    g.k_per_split = g.K;
    g.a = static_cast<const bf16_bits*>(dy);
    g.b = static_cast<const bf16_bits*>(weight_t);
    g.out = dx;
    hipLaunchKernelGGL(gemm_kernel, dim3((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, 1), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g);
    return tfdtd_check_launch("conv dgrad");
}

int tfdtd_conv_wgrad(const tfdtd_conv* d, const void* dy_t, const void* x, int32_t splits, float* partial, float* dw,
                     void* stream) {
    if (!d || !dy_t || !x || !partial || !dw || splits < 1) return tfdtd_fail(TFDTD_EINVAL, "conv wgrad: bad arguments");
    int status = check_conv(*d);
    if (status) return status;
    Gemm g{};
    g.op = 2;
    g.c = make_conv(*d);
    const long long m_total = (long long)d->n * g.c.dout * g.c.hout * g.c.wout;
    g.M = d->co;
    g.N = d->kp;
    g.K = (int)((m_total + BK - 1) / BK * BK);
    g.k_per_split = (int)((g.K / BK + splits - 1) / splits * BK);
    g.a = static_cast<const bf16_bits*>(dy_t);
    g.b = static_cast<const bf16_bits*>(x);
    g.out = partial;
    const int used = (g.K + g.k_per_split - 1) / g.k_per_split;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(gemm_kernel, dim3((g.M + BM - 1) / BM, (g.N + BN - 1) / BN, used), dim3(256), 0, s, g);
    status = tfdtd_check_launch("conv wgrad");
    if (status) return status;
    const long long count = (long long)g.M * g.N;
    hipLaunchKernelGGL(split_sum_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, partial, used, count, dw);
    return tfdtd_check_launch("conv wgrad sum");
}

int tfdtd_gelu_backward(const float* dy, const float* pre, int64_t count, int32_t activation, float* dpre, void* dpre16,
                        void* stream) {
    if (!dy || (activation && !pre) || !dpre || !dpre16 || count < 0) return tfdtd_fail(TFDTD_EINVAL, "gelu backward: bad arguments");
    if (count == 0) return TFDTD_OK;
    hipLaunchKernelGGL(gelu_backward_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       dy, pre, (long long)count, activation, dpre, static_cast<bf16_bits*>(dpre16));
    return tfdtd_check_launch("gelu backward");
}

}  // extern "C"

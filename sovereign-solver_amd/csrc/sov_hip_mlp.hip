// The learned warm start's network: a fully connected MLP with GELU hidden
// layers, BF16 weights and activations and FP32 accumulation on the MFMA
// matrix cores (v_mfma_f32_32x32x16_bf16), trained on the GPU with Adam.
//
// No reference counterpart: the network predicts a dual warm start and a
// primal weight for PDHG on a parametric LP family.
//
// Layout.  Parameters are one flat FP32 array (the Adam master copy): for
// layer l = 1..L, W_l (w[l] x w[l-1], row-major) then b_l (w[l]).  A BF16
// copy of the same array feeds the matrix cores.  Activations are row-major
// [batch x features]: BF16 inputs of each product, FP32 pre-activations kept
// for the backward pass.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sov_hip_internal.hpp"

struct sov_hip_mlp {
    int L = 0;
    std::vector<int> w;                 // widths[0..L]
    int max_batch = 0;
    int max_w = 0;
    std::vector<int64_t> woff, boff;    // offsets of W_l, b_l (l = 1..L; index l)
    int64_t nparams = 0;
    float* P = nullptr;                 // FP32 master parameters
    float* G = nullptr;                 // gradients
    float* M1 = nullptr;                // Adam moments
    float* M2 = nullptr;
    uint16_t* Pb = nullptr;             // BF16 copy of P
    uint16_t* xb = nullptr;             // BF16 input (max_batch x w[0])
    std::vector<float*> pre;            // l = 1..L: FP32 pre-activations (max_batch x w[l])
    std::vector<uint16_t*> act;         // l = 1..L-1: BF16 GELU outputs
    float* dpre_f[2] = {nullptr, nullptr};      // FP32 d(pre) ping-pong (max_batch x max_w)
    uint16_t* dpre_b[2] = {nullptr, nullptr};   // BF16 copies
    double* red = nullptr;              // loss partials
    int last_batch = 0;                 // batch of the last forward pass
    int device = 0;
    hipStream_t stream = nullptr;       // the creating context's stream (destroy waits for it only)
};

namespace sovhip {

// ---------------------------------------------------------------------------
// BF16 and GELU
// ---------------------------------------------------------------------------

// float -> bf16, round to nearest even; NaN stays a quiet NaN.
__host__ __device__ __forceinline__ uint16_t to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

__host__ __device__ __forceinline__ float from_bf16(uint16_t h) {
    return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16);
}

// GELU(x) = x Phi(x), Phi the standard normal CDF; GELU'(x) = Phi(x) + x phi(x).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// ---------------------------------------------------------------------------
// GEMM on the matrix cores
// ---------------------------------------------------------------------------

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// C = op(A) op(B), op(A) M x K, op(B) K x N, BF16 in, FP32 accumulate.
//   op(A)(i, k) = ta ? A[k lda + i] : A[i lda + k]
//   op(B)(k, j) = tb ? B[j ldb + k] : B[k ldb + j]
// Epilogue, per element (i, j) of the FP32 result r:
//   r += bias[j]                      (bias != null)
//   r *= gelu'(gate[i ldg + j])       (gate != null)
//   C[i ldc + j] = r                  (C != null)
//   Cb[i ldcb + j] = bf16(act ? gelu(r) : r)   (Cb != null)
struct GemmArgs {
    int M, N, K;
    const uint16_t* A;
    int64_t lda;
    int ta;
    const uint16_t* B;
    int64_t ldb;
    int tb;
    float* C;
    int64_t ldc;
    uint16_t* Cb;
    int64_t ldcb;
    const float* bias;
    const float* gate;
    int64_t ldg;
    int act;
};

constexpr int kTile = 64;      // block tile (4 waves of 32 x 32)
constexpr int kTileK = 32;     // k per LDS stage: two 16-deep MFMA steps
constexpr int kLdsRow = 40;    // LDS row stride in elements (80 bytes: 16-byte aligned, banks staggered)

// Stage op(A) rows [row0, row0 + 64) x k [k0, k0 + 32) into sa[i][k] and
// op(B) columns [col0, col0 + 64) x the same k into sb[j][k]: k contiguous
// for both, whatever the storage order, zero outside the matrices.  Each of
// the 256 threads moves 8 elements of each; consecutive threads read
// consecutive addresses of the stored operand.
__device__ __forceinline__ void stage(const GemmArgs& g, int row0, int col0, int k0, uint16_t* sa, uint16_t* sb) {
    const int t = threadIdx.x;
    if (!g.ta) {   // A stored M x K: 8 consecutive k of one row
        const int i = t / 4, kv = (t % 4) * 8, gi = row0 + i;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gk = k0 + kv + q;
            sa[i * kLdsRow + kv + q] = (gi < g.M && gk < g.K) ? g.A[(int64_t)gi * g.lda + gk] : uint16_t(0);
        }
    } else {       // A stored K x M: 8 consecutive rows of one k
        const int k = t / 8, iv = (t % 8) * 8, gk = k0 + k;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gi = row0 + iv + q;
            sa[(iv + q) * kLdsRow + k] = (gi < g.M && gk < g.K) ? g.A[(int64_t)gk * g.lda + gi] : uint16_t(0);
        }
    }
    if (g.tb) {    // B stored N x K: 8 consecutive k of one column
        const int j = t / 4, kv = (t % 4) * 8, gj = col0 + j;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gk = k0 + kv + q;
            sb[j * kLdsRow + kv + q] = (gj < g.N && gk < g.K) ? g.B[(int64_t)gj * g.ldb + gk] : uint16_t(0);
        }
    } else {       // B stored K x N: 8 consecutive columns of one k
        const int k = t / 8, jv = (t % 8) * 8, gk = k0 + k;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int gj = col0 + jv + q;
            sb[(jv + q) * kLdsRow + k] = (gj < g.N && gk < g.K) ? g.B[(int64_t)gk * g.ldb + gj] : uint16_t(0);
        }
    }
}

// Each wave computes a 32 x 32 tile with v_mfma_f32_32x32x16_bf16: lane l
// supplies op(A) row (l % 32) and op(B) column (l % 32) at the 8 k values
// 8 (l / 32) .. +7 of each 16-deep step (one 16-byte LDS read each), and
// holds C rows 8 i + 4 (l / 32) + r (i, r in 0..3) of column l % 32 in
// acc[4 i + r].
__global__ void __launch_bounds__(kBlock) k_gemm(GemmArgs g) {
    __shared__ __align__(16) uint16_t sa[kTile * kLdsRow];
    __shared__ __align__(16) uint16_t sb[kTile * kLdsRow];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int brow = blockIdx.y * kTile, bcol = blockIdx.x * kTile;
    const int wr = (wave / 2) * 32, wc = (wave % 2) * 32;
    const int li = lane % 32, kq = (lane / 32) * 8;
    f32x16 acc = {};
    for (int k0 = 0; k0 < g.K; k0 += kTileK) {
        stage(g, brow, bcol, k0, sa, sb);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTileK; ks += 16) {
            const uint4 av = *reinterpret_cast<const uint4*>(sa + (wr + li) * kLdsRow + ks + kq);
            const uint4 bv = *reinterpret_cast<const uint4*>(sb + (wc + li) * kLdsRow + ks + kq);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                          __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int j = bcol + wc + li;
    if (j >= g.N) return;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const int i = brow + wr + 8 * (t / 4) + (lane / 32) * 4 + (t % 4);
        if (i >= g.M) continue;
        float r = acc[t];
        if (g.bias) r += g.bias[j];
        if (g.gate) r *= gelu_grad(g.gate[(int64_t)i * g.ldg + j]);
        if (g.C) g.C[(int64_t)i * g.ldc + j] = r;
        if (g.Cb) g.Cb[(int64_t)i * g.ldcb + j] = to_bf16(g.act ? gelu(r) : r);
    }
}

static int gemm(sov_hip_ctx* ctx, const GemmArgs& g) {
    if (g.M <= 0 || g.N <= 0) return SOV_OK;
    dim3 grid((g.N + kTile - 1) / kTile, (g.M + kTile - 1) / kTile);
    k_gemm<<<grid, kBlock, 0, ctx->stream>>>(g);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

// ---------------------------------------------------------------------------
// elementwise kernels
// ---------------------------------------------------------------------------

__global__ void k_f32_to_bf16(int64_t n, const float* __restrict__ s, uint16_t* __restrict__ d) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = to_bf16(s[i]);
}

// Column sums of an M x N FP32 matrix (bias gradients): thread per column.
__global__ void k_col_sums(int M, int N, const float* __restrict__ a, float* __restrict__ out) {
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < N; j += (int64_t)gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int i = 0; i < M; ++i) s += a[(int64_t)i * N + j];
        out[j] = s;
    }
}

// Mean squared error over a batch: L = sum((y - t)^2 w) / (M N), w = weight[j]
// (all ones when null); dy = dL/dy in FP32; per-block partial sums of the
// loss into partial[].
__global__ void k_mse(int M, int N, const float* __restrict__ y, const float* __restrict__ t,
                      const float* __restrict__ weight, float* __restrict__ dy, double* __restrict__ partial) {
    double acc = 0.0;
    const int64_t total = (int64_t)M * N;
    const double scale = 1.0 / (double)total;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % N);
        const double wj = weight ? (double)weight[j] : 1.0;
        const double d = (double)y[e] - (double)t[e];
        acc += wj * d * d;
        dy[e] = (float)(2.0 * wj * d * scale);
    }
    double sm[1] = {acc * scale}, mx[1] = {0.0};
    block_reduce_write<1, 0>(sm, mx, partial + blockIdx.x);
}

// One Adam step (Kingma & Ba) on the FP32 master parameters, with bias
// correction; refreshes the BF16 copy.
__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ pb, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.0f - b1) * gi;
        const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        const float w = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
        p[i] = w;
        pb[i] = to_bf16(w);
    }
}

static void free_mlp(sov_hip_mlp* h) {
    for (void* p : {(void*)h->P, (void*)h->G, (void*)h->M1, (void*)h->M2, (void*)h->Pb, (void*)h->xb,
                    (void*)h->dpre_f[0], (void*)h->dpre_f[1], (void*)h->dpre_b[0], (void*)h->dpre_b[1],
                    (void*)h->red})
        (void)hipFree(p);
    for (float* p : h->pre) (void)hipFree(p);
    for (uint16_t* p : h->act) (void)hipFree(p);
    delete h;
}

constexpr int kLossBlocks = 256;

}  // namespace sovhip

using namespace sovhip;

extern "C" {

int sov_hip_mlp_create(sov_hip_ctx* ctx, int32_t layers, const int32_t* widths, int32_t max_batch,
                       sov_hip_mlp** out) {
    if (ctx == nullptr || out == nullptr || widths == nullptr || layers < 1 || layers > 16 || max_batch < 1)
        return fail(SOV_INVALID_ARGUMENT, "mlp_create args");
    *out = nullptr;
    for (int l = 0; l <= layers; ++l)
        if (widths[l] < 1 || widths[l] > 65536) return fail(SOV_INVALID_ARGUMENT, "mlp_create: bad width");
    auto* h = new (std::nothrow) sov_hip_mlp();
    if (h == nullptr) return fail(SOV_OUT_OF_MEMORY, "mlp");
    h->L = layers;
    h->w.assign(widths, widths + layers + 1);
    h->max_batch = max_batch;
    h->max_w = *std::max_element(h->w.begin(), h->w.end());
    h->device = ctx->device;
    h->stream = ctx->stream;
    h->woff.assign(layers + 1, 0);
    h->boff.assign(layers + 1, 0);
    int64_t off = 0;
    for (int l = 1; l <= layers; ++l) {
        h->woff[l] = off;
        off += (int64_t)h->w[l] * h->w[l - 1];
        h->boff[l] = off;
        off += h->w[l];
    }
    h->nparams = off;
    h->pre.assign(layers + 1, nullptr);
    h->act.assign(layers + 1, nullptr);
    const int64_t mb = max_batch;
    hipError_t e = hipMalloc(&h->P, sizeof(float) * off);
    if (e == hipSuccess) e = hipMalloc(&h->G, sizeof(float) * off);
    if (e == hipSuccess) e = hipMalloc(&h->M1, sizeof(float) * off);
    if (e == hipSuccess) e = hipMalloc(&h->M2, sizeof(float) * off);
    if (e == hipSuccess) e = hipMalloc(&h->Pb, sizeof(uint16_t) * off);
    if (e == hipSuccess) e = hipMalloc(&h->xb, sizeof(uint16_t) * mb * h->w[0]);
    for (int l = 1; l <= layers && e == hipSuccess; ++l) {
        e = hipMalloc(&h->pre[l], sizeof(float) * mb * h->w[l]);
        if (e == hipSuccess && l < layers) e = hipMalloc(&h->act[l], sizeof(uint16_t) * mb * h->w[l]);
    }
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipMalloc(&h->dpre_f[k], sizeof(float) * mb * h->max_w);
        if (e == hipSuccess) e = hipMalloc(&h->dpre_b[k], sizeof(uint16_t) * mb * h->max_w);
    }
    if (e == hipSuccess) e = hipMalloc(&h->red, sizeof(double) * (kLossBlocks + 1));
    if (e == hipSuccess) e = hipMemsetAsync(h->M1, 0, sizeof(float) * off, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->M2, 0, sizeof(float) * off, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->P, 0, sizeof(float) * off, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->Pb, 0, sizeof(uint16_t) * off, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        free_mlp(h);
        return hip_fail(e, "mlp_create");
    }
    *out = h;
    return SOV_OK;
}

void sov_hip_mlp_destroy(sov_hip_mlp* h) {
    if (h == nullptr) return;
    int prev = -1;
    if (hipGetDevice(&prev) == hipSuccess && prev != h->device) (void)hipSetDevice(h->device);
    else prev = -1;
    // Only the creating context's stream can still be using these buffers.
    if (h->stream != nullptr) (void)hipStreamSynchronize(h->stream);
    free_mlp(h);
    if (prev >= 0) (void)hipSetDevice(prev);
}

int64_t sov_hip_mlp_num_params(const sov_hip_mlp* h) { return h ? h->nparams : -1; }

int sov_hip_mlp_set_params(sov_hip_ctx* ctx, sov_hip_mlp* h, const float* params) {
    if (ctx == nullptr || h == nullptr || params == nullptr) return fail(SOV_INVALID_ARGUMENT, "mlp_set_params args");
    SOV_HIP_TRY(hipMemcpyAsync(h->P, params, sizeof(float) * h->nparams, hipMemcpyHostToDevice, ctx->stream));
    k_f32_to_bf16<<<launch_blocks(h->nparams), kBlock, 0, ctx->stream>>>(h->nparams, h->P, h->Pb);
    SOV_HIP_TRY(check_launch());
    SOV_HIP_TRY(hipMemsetAsync(h->M1, 0, sizeof(float) * h->nparams, ctx->stream));
    SOV_HIP_TRY(hipMemsetAsync(h->M2, 0, sizeof(float) * h->nparams, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_mlp_get_params(sov_hip_ctx* ctx, const sov_hip_mlp* h, float* params, float* grads) {
    if (ctx == nullptr || h == nullptr) return fail(SOV_INVALID_ARGUMENT, "mlp_get_params args");
    if (params) SOV_HIP_TRY(hipMemcpyAsync(params, h->P, sizeof(float) * h->nparams, hipMemcpyDeviceToHost, ctx->stream));
    if (grads) SOV_HIP_TRY(hipMemcpyAsync(grads, h->G, sizeof(float) * h->nparams, hipMemcpyDeviceToHost, ctx->stream));
    SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SOV_OK;
}

int sov_hip_mlp_forward(sov_hip_ctx* ctx, sov_hip_mlp* h, int32_t batch, const float* d_x, float* d_y) {
    if (ctx == nullptr || h == nullptr || d_x == nullptr || batch < 1 || batch > h->max_batch)
        return fail(SOV_INVALID_ARGUMENT, "mlp_forward args");
    k_f32_to_bf16<<<launch_blocks((int64_t)batch * h->w[0]), kBlock, 0, ctx->stream>>>((int64_t)batch * h->w[0], d_x,
                                                                                      h->xb);
    SOV_HIP_TRY(check_launch());
    for (int l = 1; l <= h->L; ++l) {
        GemmArgs g{};
        g.M = batch;
        g.N = h->w[l];
        g.K = h->w[l - 1];
        g.A = l == 1 ? h->xb : h->act[l - 1];
        g.lda = h->w[l - 1];
        g.ta = 0;
        g.B = h->Pb + h->woff[l];   // op(B)(k, j) = W_l[j][k]
        g.ldb = h->w[l - 1];
        g.tb = 1;
        g.C = h->pre[l];
        g.ldc = h->w[l];
        g.bias = h->P + h->boff[l];
        if (l < h->L) {
            g.Cb = h->act[l];
            g.ldcb = h->w[l];
            g.act = 1;
        }
        const int rc = gemm(ctx, g);
        if (rc != SOV_OK) return rc;
    }
    if (d_y)
        SOV_HIP_TRY(hipMemcpyAsync(d_y, h->pre[h->L], sizeof(float) * batch * h->w[h->L], hipMemcpyDeviceToDevice,
                                   ctx->stream));
    h->last_batch = batch;
    return SOV_OK;
}

int sov_hip_mlp_backward(sov_hip_ctx* ctx, sov_hip_mlp* h, const float* d_dy) {
    if (ctx == nullptr || h == nullptr || d_dy == nullptr || h->last_batch < 1)
        return fail(SOV_INVALID_ARGUMENT, "mlp_backward args (run a forward pass first)");
    const int B = h->last_batch;
    int cur = 0;
    SOV_HIP_TRY(hipMemcpyAsync(h->dpre_f[cur], d_dy, sizeof(float) * B * h->w[h->L], hipMemcpyDeviceToDevice,
                               ctx->stream));
    k_f32_to_bf16<<<launch_blocks((int64_t)B * h->w[h->L]), kBlock, 0, ctx->stream>>>((int64_t)B * h->w[h->L], d_dy,
                                                                                      h->dpre_b[cur]);
    SOV_HIP_TRY(check_launch());
    for (int l = h->L; l >= 1; --l) {
        const uint16_t* in = l == 1 ? h->xb : h->act[l - 1];
        // dW_l = dpre_l^T in_l
        GemmArgs gw{};
        gw.M = h->w[l];
        gw.N = h->w[l - 1];
        gw.K = B;
        gw.A = h->dpre_b[cur];
        gw.lda = h->w[l];
        gw.ta = 1;
        gw.B = in;
        gw.ldb = h->w[l - 1];
        gw.tb = 0;
        gw.C = h->G + h->woff[l];
        gw.ldc = h->w[l - 1];
        int rc = gemm(ctx, gw);
        if (rc != SOV_OK) return rc;
        k_col_sums<<<launch_blocks(h->w[l]), kBlock, 0, ctx->stream>>>(B, h->w[l], h->dpre_f[cur], h->G + h->boff[l]);
        SOV_HIP_TRY(check_launch());
        if (l == 1) break;
        // dpre_{l-1} = (dpre_l W_l) * gelu'(pre_{l-1})
        GemmArgs gd{};
        gd.M = B;
        gd.N = h->w[l - 1];
        gd.K = h->w[l];
        gd.A = h->dpre_b[cur];
        gd.lda = h->w[l];
        gd.ta = 0;
        gd.B = h->Pb + h->woff[l];   // op(B)(k, j) = W_l[k][j]
        gd.ldb = h->w[l - 1];
        gd.tb = 0;
        gd.gate = h->pre[l - 1];
        gd.ldg = h->w[l - 1];
        gd.C = h->dpre_f[1 - cur];
        gd.ldc = h->w[l - 1];
        gd.Cb = h->dpre_b[1 - cur];
        gd.ldcb = h->w[l - 1];
        rc = gemm(ctx, gd);
        if (rc != SOV_OK) return rc;
        cur = 1 - cur;
    }
    return SOV_OK;
}

int sov_hip_mlp_adam(sov_hip_ctx* ctx, sov_hip_mlp* h, float lr, float beta1, float beta2, float eps, int32_t step) {
    if (ctx == nullptr || h == nullptr || step < 1) return fail(SOV_INVALID_ARGUMENT, "mlp_adam args");
    const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
    k_adam<<<launch_blocks(h->nparams), kBlock, 0, ctx->stream>>>(h->nparams, h->P, h->G, h->M1, h->M2, h->Pb, lr,
                                                                  beta1, beta2, eps, c1, c2);
    SOV_HIP_TRY(check_launch());
    return SOV_OK;
}

int sov_hip_mlp_mse(sov_hip_ctx* ctx, sov_hip_mlp* h, int32_t batch, const float* d_y, const float* d_target,
                    const float* d_weight, float* d_dy, double* loss) {
    if (ctx == nullptr || h == nullptr || d_y == nullptr || d_target == nullptr || d_dy == nullptr || batch < 1)
        return fail(SOV_INVALID_ARGUMENT, "mlp_mse args");
    const int N = h->w[h->L];
    k_mse<<<kLossBlocks, kBlock, 0, ctx->stream>>>(batch, N, d_y, d_target, d_weight, d_dy, h->red);
    SOV_HIP_TRY(check_launch());
    if (loss != nullptr) {
        std::vector<double> part(kLossBlocks);
        SOV_HIP_TRY(hipMemcpyAsync(part.data(), h->red, sizeof(double) * kLossBlocks, hipMemcpyDeviceToHost,
                                   ctx->stream));
        SOV_HIP_TRY(hipStreamSynchronize(ctx->stream));
        double s = 0.0;
        for (double v : part) s += v;
        *loss = s;
    }
    return SOV_OK;
}

}  // extern "C"

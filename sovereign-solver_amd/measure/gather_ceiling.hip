// Random-gather ceiling of the GPU: gathers/s from vectors of 64 KB .. 32 MB
// with uniformly random indices streamed from memory (the access pattern of a
// CSR product on a random sparse matrix), plus a streaming-copy bandwidth.
// The step kernels of the PDHG engine are bounded by this gather rate, not by
// HBM bandwidth.  Usage: bin/gather_ceiling (prints one line per case).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <random>
#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

__global__ void k_copy(int64_t n, const float4* __restrict__ a, float4* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) b[i] = a[i];
}
// each thread: sums x[idx[k]] * v[k] over a contiguous chunk of CH entries (like SpMV with G=1 rows of CH)
template <int CH>
__global__ void k_gather(int64_t nrows, const int* __restrict__ idx, const float* __restrict__ val,
                         const float* __restrict__ x, float* __restrict__ y) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < CH; ++k) { int64_t p = r * CH + k; acc += val[p] * x[idx[p]]; }
    y[r] = acc;
}
// wave-cooperative: G lanes per row, row length L, interleaved
template <int G>
__global__ void k_gather_g(int64_t nrows, int L, const int* __restrict__ idx, const float* __restrict__ val,
                           const float* __restrict__ x, float* __restrict__ y) {
    int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / G;
    int lane = threadIdx.x % G;
    float acc = 0.f;
    if (r < nrows) for (int k = lane; k < L; k += G) { int64_t p = r * L + k; acc += val[p] * x[idx[p]]; }
    for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (r < nrows && lane == 0) y[r] = acc;
}

int main() {
    const int64_t nnz = 21 * 1024 * 1024;  // ~22M entries
    std::vector<int> hidx(nnz);
    float *dval, *dx, *dy; int* didx;
    CK(hipMalloc(&dval, nnz * 4)); CK(hipMalloc(&didx, nnz * 4));
    CK(hipMalloc(&dx, 64ll << 20)); CK(hipMalloc(&dy, nnz * 4));
    CK(hipMemset(dval, 0, nnz * 4)); CK(hipMemset(dx, 0, 64ll << 20));
    hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
    // stream bandwidth
    {
        int64_t n4 = (256ll << 20) / 16; float4 *a, *b; CK(hipMalloc(&a, n4 * 16)); CK(hipMalloc(&b, n4 * 16));
        CK(hipMemset(a, 0, n4 * 16));
        float best = 1e9;
        for (int rep = 0; rep < 5; ++rep) {
            CK(hipEventRecord(e0)); k_copy<<<8192, 256>>>(n4, a, b); CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
            float ms; CK(hipEventElapsedTime(&ms, e0, e1)); if (rep) best = std::min(best, ms);
        }
        printf("copy 256MB: %.1f us -> %.0f GB/s (read+write)\n", best * 1e3, 2.0 * n4 * 16 / best / 1e6);
        CK(hipFree(a)); CK(hipFree(b));
    }
    std::mt19937_64 rng(1);
    for (int64_t xs : {64ll << 10, 256ll << 10, 1ll << 20, 2ll << 20, 4ll << 20, 8ll << 20, 32ll << 20}) {
        const int64_t xn = xs / 4;
        std::uniform_int_distribution<int64_t> d(0, xn - 1);
        for (int64_t k = 0; k < nnz; ++k) hidx[k] = (int)d(rng);
        CK(hipMemcpy(didx, hidx.data(), nnz * 4, hipMemcpyHostToDevice));
        auto run = [&](auto launch, const char* name) {
            float best = 1e9;
            for (int rep = 0; rep < 6; ++rep) {
                CK(hipEventRecord(e0)); launch(); CK(hipGetLastError()); CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
                float ms; CK(hipEventElapsedTime(&ms, e0, e1)); if (rep) best = std::min(best, ms);
            }
            printf("x=%6lld KB %-14s %8.1f us  %6.1f Ggathers/s  stream %6.0f GB/s\n", (long long)(xs >> 10), name,
                   best * 1e3, nnz / best / 1e6, nnz * 8.0 / best / 1e6);
        };
        run([&] { k_gather<2><<<(nnz / 2 + 255) / 256, 256>>>(nnz / 2, didx, dval, dx, dy); }, "G1 len2");
        run([&] { k_gather<8><<<(nnz / 8 + 255) / 256, 256>>>(nnz / 8, didx, dval, dx, dy); }, "G1 len8");
        run([&] { k_gather_g<8><<<(nnz / 40 * 8 + 255) / 256, 256>>>(nnz / 40, 40, didx, dval, dx, dy); }, "G8 len40");
        run([&] { k_gather_g<32><<<(nnz / 200 * 32 + 255) / 256, 256>>>(nnz / 200, 200, didx, dval, dx, dy); }, "G32 len200");
        run([&] { k_gather_g<64><<<(nnz / 200 * 64 + 255) / 256, 256>>>(nnz / 200, 200, didx, dval, dx, dy); }, "G64 len200");
    }
    return 0;
}
